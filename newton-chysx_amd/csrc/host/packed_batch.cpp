#include "packed_batch.h"

#include <algorithm>
#include <limits>
#include <stdexcept>
#include <string>

namespace chysx {
namespace engine {

namespace {

unsigned encode(unsigned kind, int a, int b, std::size_t id) {
    if (id > CHX_CONTRIB_MAX_ELEMENT) throw std::runtime_error("chysx::engine: element index exceeds the 26-bit gather-map field");
    return (kind << 30) | (unsigned(a) << 28) | (unsigned(b) << 26) | static_cast<unsigned>(id);
}

template <class T> void upload(DeviceArray<T>& a, std::size_t& bytes) {
    a.copy_to_device();
    bytes += a.gpu_size() * sizeof(T);
}
template <class T> void alloc(DeviceArray<T>& a, std::size_t n, std::size_t& bytes) {
    a.ensure_device(n);
    bytes += n * sizeof(T);
}

chx_instance make_instance(const InstanceSource& s) {
    chx_instance I{};
    I.pos = s.pos;
    I.vel = s.vel;
    I.inv_mass = s.inv_mass;
    I.n_rows = s.n;
    I.dt = s.dt;
    I.damping = s.damping;
    I.gx = s.gx;
    I.gy = s.gy;
    I.gz = s.gz;
    I.max_iter = s.max_iter;
    I.tol = s.tol;
    return I;
}

}  // namespace

void PackedBatch::pack_elements(const std::vector<InstanceSource>& src) {
    auto& P = pins_.host();
    auto& T = tris_.host();
    auto& D = dihs_.host();
    P.clear();
    T.clear();
    D.clear();
    for (std::size_t i = 0; i < src.size(); ++i) {
        const InstanceSource& s = src[i];
        const int inst = static_cast<int>(i);
        if (s.pins) {
            const float k = s.pins->stiffness();
            for (int c = 0; c < s.pins->size(); ++c) {
                const auto& t = s.pins->targets()[c];
                P.push_back(chx_pin{s.pins->indices()[c], inst, t.x, t.y, t.z, k});
            }
        }
        for (const constraint::TriangleStretchConstraint* ts : {s.stretch, s.shear}) {
            if (!ts) continue;
            const float k = ts->stiffness();
            for (int t = 0; t < ts->size(); ++t) {
                const auto& v = ts->indices()[t];
                const auto& dm = ts->Dm_inv()[t].data;
                // k_eff = areas[t] * k (triangle_stretch_constraint.cu:193).
                T.push_back(chx_tri{v.x, v.y, v.z, inst, dm[0], dm[1], dm[2], dm[3], ts->areas()[t] * k});
            }
        }
        if (s.bending) {
            const float k = s.bending->stiffness();
            for (int e = 0; e < s.bending->size(); ++e) {
                const auto& v = s.bending->indices()[e];
                D.push_back(chx_dih{v.x, v.y, v.z, v.w, inst, s.bending->rest_angles()[e], k, 0});
            }
        }
    }
}

void PackedBatch::build(const std::vector<InstanceSource>& src, int lanes_per_row, float* external_x) {
    if (lanes_per_row != 4 && lanes_per_row != 8 && lanes_per_row != 16)
        throw std::invalid_argument("chysx::engine: lanes_per_row must be 4, 8 or 16");
    if (external_x && src.size() != 1) throw std::invalid_argument("chysx::engine: external warm start needs one instance");
    device_bytes_ = 0;
    const int I = static_cast<int>(src.size());
    long long N = 0, NNZ = 0;
    int tiles = 0;
    for (const auto& s : src) {
        if (!s.pattern || s.pattern->n != s.n) throw std::invalid_argument("chysx::engine: instance pattern missing or sized wrong");
        N += s.n;
        NNZ += s.pattern->nnz();
        tiles += (s.n + kTileRows - 1) / kTileRows;
    }
    if (N >= std::numeric_limits<int>::max() || NNZ >= std::numeric_limits<int>::max())
        throw std::runtime_error("chysx::engine: batch too large for 32-bit row / block indices");

    // ---- rows, pattern, tiles ------------------------------------------------
    inst_host_.assign(I, chx_instance{});
    nnz_begin_.assign(I, 0);
    auto& ro = row_offsets_.host();
    auto& ci = col_indices_.host();
    auto& ti = tiles_.host();
    auto& ri = row_inst_.host();
    ro.assign(static_cast<std::size_t>(N) + 1, 0);
    ci.assign(static_cast<std::size_t>(NNZ), 0);
    ti.clear();
    ri.clear();
    if (I > 1) ri.resize(static_cast<std::size_t>(N));
    int rb = 0, nb = 0, tb = 0;
    for (int i = 0; i < I; ++i) {
        const InstanceSource& s = src[i];
        chx_instance inst = make_instance(s);
        inst.row_begin = rb;
        inst.tile_begin = tb;
        for (int r0 = 0; r0 < s.n; r0 += kTileRows) {
            ti.insert(ti.end(), {i, rb + r0, rb + std::min(s.n, r0 + kTileRows), 0});
            ++tb;
        }
        inst.tile_end = tb;
        inst_host_[i] = inst;
        nnz_begin_[i] = nb;
        const auto& p = *s.pattern;
        for (int r = 0; r < s.n; ++r) ro[rb + r] = nb + p.row_offsets[r];
        for (int k = 0; k < p.nnz(); ++k) ci[nb + k] = rb + p.col_indices[k];
        if (I > 1) std::fill(ri.begin() + rb, ri.begin() + rb + s.n, i);
        rb += s.n;
        nb += p.nnz();
    }
    ro[N] = static_cast<int>(NNZ);

    // ---- elements -------------------------------------------------------------
    pack_elements(src);
    const std::size_t nP = pins_.cpu_size(), nT = tris_.cpu_size(), nD = dihs_.cpu_size();

    // ---- gather map: per-slot contribution lists -------------------------------
    // Global slot ids: diagonal of packed row g -> g; off-diagonal packed
    // block k -> N + k.  Entries are appended in element order (pins,
    // stretch, shear, bending; element index; a; b), the order of the
    // reference's launches, so each slot sums its terms in that order.
    const std::size_t S = static_cast<std::size_t>(N + NNZ);
    std::vector<unsigned long long> count(S + 1, 0);
    auto walk = [&](auto&& visit) {
        std::size_t pin_base = 0, tri_base = 0, dih_base = 0;
        for (int i = 0; i < I; ++i) {
            const InstanceSource& s = src[i];
            const long long rbi = inst_host_[i].row_begin, nbi = nnz_begin_[i];
            auto gslot = [&](int slot) -> std::size_t {
                return slot < 0 ? static_cast<std::size_t>(rbi + (-slot - 1)) : static_cast<std::size_t>(N + nbi + slot);
            };
            if (s.pins) {
                const auto& sl = s.pins->hessian_slots();
                if (static_cast<int>(sl.size()) != s.pins->size()) throw std::runtime_error("chysx::engine: pin slots not bound");
                for (int c = 0; c < s.pins->size(); ++c) visit(gslot(sl[c]), CHX_CONTRIB_PIN, 0, 0, pin_base + c);
                pin_base += s.pins->size();
            }
            for (const constraint::TriangleStretchConstraint* ts : {s.stretch, s.shear}) {
                if (!ts) continue;
                const auto& sl = ts->hessian_slots();
                if (sl.size() != 9u * ts->size()) throw std::runtime_error("chysx::engine: triangle slots not bound");
                for (int t = 0; t < ts->size(); ++t)
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) visit(gslot(sl[9 * t + 3 * a + b]), CHX_CONTRIB_TRI, a, b, tri_base + t);
                tri_base += ts->size();
            }
            if (s.bending) {
                const auto& sl = s.bending->hessian_slots();
                if (sl.size() != 16u * s.bending->size()) throw std::runtime_error("chysx::engine: dihedral slots not bound");
                for (int e = 0; e < s.bending->size(); ++e)
                    for (int a = 0; a < 4; ++a)
                        for (int b = 0; b < 4; ++b) visit(gslot(sl[16 * e + 4 * a + b]), CHX_CONTRIB_DIH, a, b, dih_base + e);
                dih_base += s.bending->size();
            }
        }
    };
    walk([&](std::size_t g, unsigned, int, int, std::size_t) { ++count[g + 1]; });
    for (std::size_t k = 0; k < S; ++k) count[k + 1] += count[k];
    if (count[S] > std::numeric_limits<unsigned>::max())
        throw std::runtime_error("chysx::engine: gather map exceeds 2^32 entries");
    auto& off = contrib_offsets_.host();
    off.resize(S + 1);
    for (std::size_t k = 0; k <= S; ++k) off[k] = static_cast<unsigned>(count[k]);
    auto& ent = contribs_.host();
    ent.assign(count[S], 0u);
    std::vector<unsigned> fill(off.begin(), off.end() - 1);
    walk([&](std::size_t g, unsigned kind, int a, int b, std::size_t id) { ent[fill[g]++] = encode(kind, a, b, id); });
    std::vector<unsigned long long>().swap(count);

    // ---- device arrays ----------------------------------------------------------
    upload(row_offsets_, device_bytes_);
    upload(col_indices_, device_bytes_);
    upload(contrib_offsets_, device_bytes_);
    upload(contribs_, device_bytes_);
    if (I > 1) {
        upload(tiles_, device_bytes_);
        upload(row_inst_, device_bytes_);
        inst_.host() = inst_host_;
        upload(inst_, device_bytes_);
    } else {
        tiles_.release();
        row_inst_.release();
        inst_.release();
    }
    upload(pins_, device_bytes_);
    upload(tris_, device_bytes_);
    upload(dihs_, device_bytes_);
    const std::size_t n = static_cast<std::size_t>(N);
    alloc(pin_rec_, nP * CHX_PIN_REC, device_bytes_);
    alloc(tri_rec_, nT * CHX_TRI_REC, device_bytes_);
    alloc(dih_rec_, nD * CHX_DIH_REC, device_bytes_);
    alloc(diag_, 9 * n, device_bytes_);
    alloc(values_, 9 * static_cast<std::size_t>(NNZ), device_bytes_);
    alloc(rhs_, 3 * n, device_bytes_);
    if (!external_x) {
        alloc(x_, 3 * n, device_bytes_);
        x_.zero_device();
    } else {
        x_.release();
    }
    alloc(r_, 3 * n, device_bytes_);
    alloc(z_, 3 * n, device_bytes_);
    alloc(p0_, 3 * n, device_bytes_);
    alloc(p1_, 3 * n, device_bytes_);
    alloc(q_, 3 * n, device_bytes_);
    alloc(minv_, 9 * n, device_bytes_);
    alloc(partials_, static_cast<std::size_t>(CHX_TILE_PARTIALS) * std::max(tiles, 1), device_bytes_);
    alloc(inst_iters_, std::max(I, 1), device_bytes_);
    alloc(inst_done_, std::max(I, 1), device_bytes_);
    alloc(inst_rr_, std::max(I, 1), device_bytes_);
    alloc(inst_bb_, std::max(I, 1), device_bytes_);
    alloc(sync_, CHX_SYNC_WORDS, device_bytes_);
    sync_.zero_device();
    inst_iters_.zero_device();
    inst_rr_.zero_device();
    inst_bb_.zero_device();
    check(chx_device_synchronize(), "engine", "batch upload");

    // ---- descriptor ---------------------------------------------------------------
    chx_cloth_batch& B = desc_;
    B = chx_cloth_batch{};
    B.n_instances = I;
    B.n_rows = static_cast<int>(N);
    B.n_tiles = tiles;
    B.nnz = static_cast<int>(NNZ);
    B.n_pins = static_cast<int>(nP);
    B.n_tris = static_cast<int>(nT);
    B.n_dihs = static_cast<int>(nD);
    B.lanes_per_row = lanes_per_row;
    B.grid_blocks = 0;
    if (I == 1) B.single = inst_host_[0];
    B.inst = I > 1 ? inst_.gpu_data() : nullptr;
    B.row_inst = I > 1 ? row_inst_.gpu_data() : nullptr;
    B.tiles = I > 1 ? tiles_.gpu_data() : nullptr;
    B.pins = pins_.gpu_data();
    B.tris = tris_.gpu_data();
    B.dihs = dihs_.gpu_data();
    B.pin_rec = pin_rec_.gpu_data();
    B.tri_rec = tri_rec_.gpu_data();
    B.dih_rec = dih_rec_.gpu_data();
    B.row_offsets = row_offsets_.gpu_data();
    B.col_indices = col_indices_.gpu_data();
    B.diag = diag_.gpu_data();
    B.values = values_.gpu_data();
    B.contrib_offsets = contrib_offsets_.gpu_data();
    B.contribs = contribs_.gpu_data();
    B.rhs = rhs_.gpu_data();
    B.x = external_x ? external_x : x_.gpu_data();
    B.r = r_.gpu_data();
    B.z = z_.gpu_data();
    B.p0 = p0_.gpu_data();
    B.p1 = p1_.gpu_data();
    B.q = q_.gpu_data();
    B.minv = minv_.gpu_data();
    B.partials = partials_.gpu_data();
    B.inst_iters = inst_iters_.gpu_data();
    B.inst_rr = inst_rr_.gpu_data();
    B.inst_bb = inst_bb_.gpu_data();
    B.inst_done = inst_done_.gpu_data();
    B.sync = sync_.gpu_data();
    inst_dirty_ = false;
}

void PackedBatch::update_elements(const std::vector<InstanceSource>& src) {
    const std::size_t nP = pins_.cpu_size(), nT = tris_.cpu_size(), nD = dihs_.cpu_size();
    pack_elements(src);
    if (pins_.cpu_size() != nP || tris_.cpu_size() != nT || dihs_.cpu_size() != nD)
        throw std::runtime_error("chysx::engine: element counts changed; rebuild the batch");
    pins_.copy_to_device();
    tris_.copy_to_device();
    dihs_.copy_to_device();
    desc_.pins = pins_.gpu_data();
    desc_.tris = tris_.gpu_data();
    desc_.dihs = dihs_.gpu_data();
}

void PackedBatch::set_instance(int i, const InstanceSource& s) {
    chx_instance I = make_instance(s);
    chx_instance& cur = inst_host_.at(i);
    I.row_begin = cur.row_begin;
    I.n_rows = cur.n_rows;
    I.tile_begin = cur.tile_begin;
    I.tile_end = cur.tile_end;
    const bool same = I.pos == cur.pos && I.vel == cur.vel && I.inv_mass == cur.inv_mass && I.dt == cur.dt &&
                      I.damping == cur.damping && I.gx == cur.gx && I.gy == cur.gy && I.gz == cur.gz &&
                      I.max_iter == cur.max_iter && I.tol == cur.tol;
    cur = I;
    if (desc_.n_instances == 1) desc_.single = I;
    else if (!same) inst_dirty_ = true;
}

void PackedBatch::flush_instances(std::uintptr_t stream) {
    if (!inst_dirty_ || desc_.n_instances <= 1) return;
    check(chx_stream_synchronize(stream), "engine", "instance table update");
    inst_.host() = inst_host_;
    inst_.copy_to_device(stream);
    desc_.inst = inst_.gpu_data();
    inst_dirty_ = false;
}

void PackedBatch::step(std::uintptr_t stream) {
    flush_instances(stream);
    check(chx_cloth_step(&desc_, stream), "cloth::ClothSimulator", "step");
}
void PackedBatch::assemble(std::uintptr_t stream) {
    flush_instances(stream);
    check(chx_cloth_assemble(&desc_, stream), "cloth::ClothSimulator", "assemble");
}
void PackedBatch::solve(bool finalize, std::uintptr_t stream) {
    flush_instances(stream);
    check(chx_cloth_solve(&desc_, finalize ? 1 : 0, stream), "solver::PCGSolver", "solve");
}

void PackedBatch::read_stats(std::vector<int>& iters, std::vector<double>& rr, std::vector<double>& bb,
                             std::uintptr_t stream) {
    check_status(stream);
    const int I = desc_.n_instances;
    iters.assign(I, 0);
    rr.assign(I, 0.0);
    bb.assign(I, 0.0);
    if (I == 0) return;
    check(chx_memcpy_d2h(iters.data(), desc_.inst_iters, sizeof(int) * I, stream), "engine", "stats copy");
    check(chx_memcpy_d2h(rr.data(), desc_.inst_rr, sizeof(double) * I, stream), "engine", "stats copy");
    check(chx_memcpy_d2h(bb.data(), desc_.inst_bb, sizeof(double) * I, stream), "engine", "stats copy");
}

void PackedBatch::check_status(std::uintptr_t stream) {
    if (desc_.sync) check(chx_batch_check(&desc_, stream), "solver::PCGSolver", "solve");
}

void PackedBatch::copy_diag(float* out, int rb, int n) const {
    check(chx_memcpy_d2h(out, desc_.diag + 9L * rb, sizeof(float) * 9 * n, 0), "engine", "diag copy");
}
void PackedBatch::copy_values(float* out, int nb, int nnz) const {
    if (nnz > 0) check(chx_memcpy_d2h(out, desc_.values + 9L * nb, sizeof(float) * 9 * nnz, 0), "engine", "values copy");
}
void PackedBatch::copy_rhs(float* out, int rb, int n) const {
    check(chx_memcpy_d2h(out, desc_.rhs + 3L * rb, sizeof(float) * 3 * n, 0), "engine", "rhs copy");
}
void PackedBatch::copy_x(float* out, int rb, int n) const {
    check(chx_memcpy_d2h(out, desc_.x + 3L * rb, sizeof(float) * 3 * n, 0), "engine", "dx copy");
}
void PackedBatch::upload_x(const float* in, int rb, int n) {
    check(chx_memcpy_h2d(desc_.x + 3L * rb, in, sizeof(float) * 3 * n, 0), "engine", "dx upload");
}

}  // namespace engine
}  // namespace chysx
