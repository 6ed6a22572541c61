#include "cloth_simulator.h"

#include <vector>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace chysx {
namespace cloth {

// Move-only, as the reference (cloth_simulator.h:72-75).
static_assert(std::is_move_constructible<ClothSimulator>::value && std::is_move_assignable<ClothSimulator>::value &&
                  !std::is_copy_constructible<ClothSimulator>::value,
              "ClothSimulator is move-only");

void ClothSimulator::set_external_buffers(std::uintptr_t pos_ptr, std::uintptr_t vel_ptr, int particle_count,
                                          std::uintptr_t inv_mass_ptr, std::uintptr_t force_ptr) noexcept {
    const auto n = static_cast<std::size_t>(particle_count < 0 ? 0 : particle_count);
    buffers_.pos = DeviceSpan<math::Vec3f>::from_raw(pos_ptr, n);
    buffers_.vel = DeviceSpan<math::Vec3f>::from_raw(vel_ptr, n);
    buffers_.inv_mass = DeviceSpan<float>::from_raw(inv_mass_ptr, n);
    buffers_.ext_force = DeviceSpan<math::Vec3f>::from_raw(force_ptr, n);
}

// cloth_simulator.cu:277-287
void ClothSimulator::set_pins(const int* host_indices, const math::Vec3f* host_targets, int n, float stiffness) {
    pins_.set_stiffness(stiffness);
    pins_.set_pins(host_indices, host_targets, n);
    topology_dirty_ = true;
}

void ClothSimulator::clear_pins() noexcept {
    try {
        pins_.set_pins(nullptr, nullptr, 0);
    } catch (...) {
    }
    topology_dirty_ = true;
}

// cloth_simulator.cu:294-313: the targets go to the device at once, on the
// caller's stream, through a page-locked staging buffer (no wait unless the
// host is two updates ahead of the stream; the null stream waits).  The pin
// set and the pattern are unchanged, so a captured step graph reads the new
// targets on its next replay.
void ClothSimulator::update_pin_targets(const math::Vec3f* host_targets, int n_pins, std::uintptr_t cuda_stream) {
    if (host_targets == nullptr || n_pins <= 0) return;
    if (pins_.size() != n_pins)
        throw std::invalid_argument(
            "ClothSimulator::update_pin_targets: n_pins does not match the currently installed pin set; call "
            "set_pins(...) to change the pin set instead.");
    std::memcpy(pins_.targets().data(), host_targets, sizeof(math::Vec3f) * n_pins);
    // Not built yet (or rebuilt on the next step): the build uploads them.
    if (!topology_dirty_ && H_num_block_rows_ > 0)
        inst_.upload_pin_targets(reinterpret_cast<const float*>(host_targets), n_pins, cuda_stream);
}

// cloth_simulator.cu:315-377 (edge list; the self-collision topology the
// reference also builds here is not part of this step).
void ClothSimulator::set_mesh(const math::Vec3i* host_triangles, int n_triangles) {
    if (n_triangles < 0) throw std::invalid_argument("ClothSimulator::set_mesh: negative triangle count");
    buffers_.tris.assign(host_triangles, host_triangles + n_triangles);
    buffers_.edges = topology::build_edges(buffers_.tris.data(), n_triangles);
    topology_dirty_ = true;
}

void ClothSimulator::build_springs_from_current_positions(float stiffness, std::uintptr_t /*cuda_stream*/) {
    const int n_edges = static_cast<int>(buffers_.edges.size());
    springs_.set_stiffness(stiffness);
    if (n_edges == 0) {
        springs_.set_springs(nullptr, 0);
        topology_dirty_ = true;
        return;
    }
    if (buffers_.pos.data() == nullptr)
        throw std::runtime_error(
            "ClothSimulator::build_springs_from_current_positions: external positions must be set first via "
            "set_external_buffers().");
    springs_.set_springs(buffers_.edges.data(), n_edges);
    topology_dirty_ = true;
}

void ClothSimulator::build_fem_stretch_from_current_positions(float stiffness, std::uintptr_t cuda_stream) {
    const int n_tris = static_cast<int>(buffers_.tris.size());
    fem_stretch_.set_stiffness(stiffness);
    if (n_tris == 0) {
        fem_stretch_.set_triangles_from_positions(nullptr, 0, buffers_.pos, cuda_stream, 0.0f);
        topology_dirty_ = true;
        return;
    }
    if (buffers_.pos.data() == nullptr)
        throw std::runtime_error(
            "ClothSimulator::build_fem_stretch_from_current_positions: external positions must be set first via "
            "set_external_buffers().");
    fem_stretch_.set_triangles_from_positions(buffers_.tris.data(), n_tris, buffers_.pos, cuda_stream, 0.0f);
    topology_dirty_ = true;
}

void ClothSimulator::build_fem_shear_from_current_positions(float stiffness, std::uintptr_t cuda_stream) {
    constexpr float kQuarterPi = 0.7853981633974483f;  // cloth_simulator.cu:428
    const int n_tris = static_cast<int>(buffers_.tris.size());
    fem_shear_.set_stiffness(stiffness);
    if (n_tris == 0) {
        fem_shear_.set_triangles_from_positions(nullptr, 0, buffers_.pos, cuda_stream, kQuarterPi);
        topology_dirty_ = true;
        return;
    }
    if (buffers_.pos.data() == nullptr)
        throw std::runtime_error(
            "ClothSimulator::build_fem_shear_from_current_positions: external positions must be set first via "
            "set_external_buffers().");
    fem_shear_.set_triangles_from_positions(buffers_.tris.data(), n_tris, buffers_.pos, cuda_stream, kQuarterPi);
    topology_dirty_ = true;
}

void ClothSimulator::build_fem_stvk_from_current_positions(float mu, float lambda, float thickness,
                                                           std::uintptr_t cuda_stream) {
    fem_stvk_.set_material(mu, lambda, thickness);
    const int n_tris = static_cast<int>(buffers_.tris.size());
    if (n_tris == 0) {
        fem_stvk_.set_triangles_from_positions(nullptr, 0, buffers_.pos, cuda_stream);
        topology_dirty_ = true;
        return;
    }
    if (buffers_.pos.data() == nullptr)
        throw std::runtime_error(
            "ClothSimulator::build_fem_stvk_from_current_positions: external positions must be set first via "
            "set_external_buffers().");
    fem_stvk_.set_triangles_from_positions(buffers_.tris.data(), n_tris, buffers_.pos, cuda_stream);
    topology_dirty_ = true;
}

// cloth_simulator.cu:451-514
void ClothSimulator::build_bending_from_current_positions(float stiffness, std::uintptr_t cuda_stream) {
    bending_.set_stiffness(stiffness);
    const int n_tris = static_cast<int>(buffers_.tris.size());
    if (n_tris == 0) {
        bending_.set_dihedrals_from_positions(nullptr, 0, buffers_.pos, cuda_stream);
        topology_dirty_ = true;
        return;
    }
    if (buffers_.pos.data() == nullptr)
        throw std::runtime_error(
            "ClothSimulator::build_bending_from_current_positions: external positions must be set first via "
            "set_external_buffers().");
    const std::vector<math::Vec4i> dih = topology::detect_dihedrals(buffers_.tris.data(), n_tris);
    bending_.set_dihedrals_from_positions(dih.empty() ? nullptr : dih.data(), static_cast<int>(dih.size()),
                                          buffers_.pos, cuda_stream);
    topology_dirty_ = true;
}

// cloth_simulator.cu:516-575
void ClothSimulator::redistribute_mass_area_weighted(float surface_density, std::uintptr_t inv_mass_ptr,
                                                     int particle_count, std::uintptr_t cuda_stream) {
    if (particle_count <= 0) return;
    if (inv_mass_ptr == 0)
        throw std::invalid_argument("ClothSimulator::redistribute_mass_area_weighted: inv_mass_ptr must be non-null");
    if (surface_density <= 0.0f)
        throw std::invalid_argument("ClothSimulator::redistribute_mass_area_weighted: surface_density must be positive");
    const int n_tri = static_cast<int>(buffers_.tris.size());
    if (n_tri == 0)
        throw std::runtime_error(
            "ClothSimulator::redistribute_mass_area_weighted: call set_mesh(...) before redistributing mass.");
    if (buffers_.pos.data() == nullptr || static_cast<int>(buffers_.pos.size()) < particle_count)
        throw std::runtime_error(
            "ClothSimulator::redistribute_mass_area_weighted: external positions must be set first via "
            "set_external_buffers().");
    std::vector<int> off, list;
    topology::vertex_triangles(buffers_.tris.data(), n_tri, particle_count, off, list);
    DeviceArray<math::Vec3i> d_tris;
    DeviceArray<int> d_off, d_list;
    d_tris.host() = buffers_.tris;
    d_off.host() = off;
    d_list.host() = list;
    d_tris.copy_to_device(cuda_stream);
    d_off.copy_to_device(cuda_stream);
    d_list.copy_to_device(cuda_stream);
    check(chx_mass_area_weighted(reinterpret_cast<const int*>(d_tris.gpu_data()),
                                 reinterpret_cast<const float*>(buffers_.pos.data()), d_off.gpu_data(),
                                 d_list.gpu_data(), particle_count, surface_density / 3.0f,
                                 reinterpret_cast<float*>(inv_mass_ptr), cuda_stream),
          "cloth", "area-weighted mass kernel");
    // The temporaries are freed on return; wait for the kernel that reads them.
    check(chx_stream_synchronize(cuda_stream), "cloth", "area-weighted mass");
}

void ClothSimulator::set_lanes_per_row(int lanes) {
    if (lanes != 4 && lanes != 8 && lanes != 16)
        throw std::invalid_argument("ClothSimulator::set_lanes_per_row: lanes must be 4, 8 or 16");
    lanes_per_row_ = lanes;
    runner_.set_lanes_per_row(lanes);
}

void ClothSimulator::set_bf16_blocks(bool on) {
    if (on != bf16_) topology_dirty_ = true;  // the block copies are (de)allocated with the instance
    bf16_ = on;
    runner_.set_bf16(on);
}

// cloth_simulator.cu:795-812: dx (the PCG warm start) is zeroed only when
// its size changes.
void ClothSimulator::resize_work_buffers(int n, std::uintptr_t stream) {
    if (dx_.gpu_size() != static_cast<std::size_t>(n)) {
        dx_.ensure_device(n);
        dx_.zero_device(stream);
        check(stream ? chx_stream_synchronize(stream) : chx_device_synchronize(), "cloth", "dx zero fill");
    }
}

// cloth_simulator.cu:698-793
void ClothSimulator::prepare_topology() {
    const int N = buffers_.particle_count();
    if (N <= 0) throw std::runtime_error("ClothSimulator::prepare_topology: no particles bound");
    std::vector<int> rows, cols;
    topology::append_triangle_pairs(fem_stretch_.indices(), rows, cols);
    topology::append_triangle_pairs(fem_shear_.indices(), rows, cols);
    topology::append_triangle_pairs(fem_stvk_.indices(), rows, cols);
    topology::append_dihedral_pairs(bending_.indices(), rows, cols);
    pattern_ = topology::build_pattern(N, rows.data(), cols.data(), static_cast<int>(rows.size()));
    pins_.bind_hessian_layout(pattern_);
    fem_stretch_.bind_hessian_layout(pattern_);
    fem_shear_.bind_hessian_layout(pattern_);
    fem_stvk_.bind_hessian_layout(pattern_);
    bending_.bind_hessian_layout(pattern_);
}

void ClothSimulator::ensure_hessian_topology(std::uintptr_t stream) {
    const int N = buffers_.particle_count();
    if (N <= 0) return;
    prepare_topology();
    inst_.build(source(), stream);
    runner_.set_lanes_per_row(lanes_per_row_);
    runner_.set_bf16(bf16_);
    runner_.bind_single(inst_.descriptor(source()), stream);
    packed_k_[0] = pins_.stiffness();
    packed_k_[1] = fem_stretch_.stiffness();
    packed_k_[2] = fem_shear_.stiffness();
    packed_k_[3] = bending_.stiffness();
    packed_stvk_[0] = fem_stvk_.mu();
    packed_stvk_[1] = fem_stvk_.lambda();
    packed_stvk_[2] = fem_stvk_.thickness();
    H_num_block_rows_ = N;
    topology_dirty_ = false;
    shapes_dirty_ = false;  // build() uploaded the shape list
}

engine::InstanceSource ClothSimulator::source() const {
    engine::InstanceSource s;
    s.n = buffers_.particle_count();
    s.pos = reinterpret_cast<float*>(buffers_.pos.data());
    s.vel = reinterpret_cast<float*>(buffers_.vel.data());
    s.inv_mass = buffers_.inv_mass.data();
    s.pattern = &pattern_;
    s.x = reinterpret_cast<float*>(dx_.gpu_data());
    s.pins = &pins_;
    s.stretch = &fem_stretch_;
    s.shear = &fem_shear_;
    s.stvk = &fem_stvk_;
    s.bending = &bending_;
    s.damping = material_.damping;
    s.gx = material_.gx;
    s.gy = material_.gy;
    s.gz = material_.gz;
    s.max_iter = pcg_max_iterations_;
    s.tol = pcg_tolerance_;
    s.bf16_blocks = bf16_;
    s.planes = &planes_;
    s.boxes = &boxes_;
    s.contact_thickness = contact_thickness_;
    s.contact_stiffness = contact_stiffness_;
    s.contact_friction = contact_friction_;
    s.contact_friction_eps = contact_friction_eps_;
    return s;
}

void ClothSimulator::add_static_plane(const collision::PlaneShape& p) {
    planes_.push_back(chx_plane{p.n.x, p.n.y, p.n.z, p.d});
    shapes_dirty_ = true;
}

void ClothSimulator::add_static_box(const collision::BoxShape& b) {
    chx_box c{};
    c.cx = b.center.x; c.cy = b.center.y; c.cz = b.center.z;
    c.hx = b.half_ext.x; c.hy = b.half_ext.y; c.hz = b.half_ext.z;
    const math::Vec3f* ax[3] = {&b.ex, &b.ey, &b.ez};
    float* out[3] = {c.ex, c.ey, c.ez};
    for (int a = 0; a < 3; ++a) {
        out[a][0] = ax[a]->x;
        out[a][1] = ax[a]->y;
        out[a][2] = ax[a]->z;
    }
    boxes_.push_back(c);
    shapes_dirty_ = true;
}

void ClothSimulator::clear_static_shapes() {
    planes_.clear();
    boxes_.clear();
    shapes_dirty_ = true;
}

// Stiffness is read from the constraint sets at every launch in the
// reference; re-pack the element arrays when one changed since packing.
void ClothSimulator::check_elements_current(std::uintptr_t stream) {
    const bool k_changed = packed_k_[0] != pins_.stiffness() || packed_k_[1] != fem_stretch_.stiffness() ||
                           packed_k_[2] != fem_shear_.stiffness() || packed_k_[3] != bending_.stiffness() ||
                           packed_stvk_[0] != fem_stvk_.mu() || packed_stvk_[1] != fem_stvk_.lambda() ||
                           packed_stvk_[2] != fem_stvk_.thickness();
    if (k_changed) {
        inst_.update_elements(source(), stream);
        packed_k_[0] = pins_.stiffness();
        packed_k_[1] = fem_stretch_.stiffness();
        packed_k_[2] = fem_shear_.stiffness();
        packed_k_[3] = bending_.stiffness();
        packed_stvk_[0] = fem_stvk_.mu();
        packed_stvk_[1] = fem_stvk_.lambda();
        packed_stvk_[2] = fem_stvk_.thickness();
    }
    if (shapes_dirty_) {
        inst_.update_shapes(source(), stream);
        shapes_dirty_ = false;
    }
}

// cloth_simulator.cu:814-1261 (PCG branch).
void ClothSimulator::step(float dt, std::uintptr_t cuda_stream) {
    const int n = buffers_.particle_count();
    if (n <= 0) return;
    if (solver_type_ == SolverType::VBD)
        throw std::runtime_error("chysx::cloth::ClothSimulator::step: the VBD solver is not implemented by this library");
    if (buffers_.pos.data() == nullptr || buffers_.vel.data() == nullptr)
        throw std::runtime_error(
            "chysx::cloth::ClothSimulator::step: pos / vel spans must be set via set_external_buffers() before "
            "stepping.");
    if (!(dt > 0.0f)) throw std::invalid_argument("chysx::cloth::ClothSimulator::step: dt must be positive");
    ProfileRange range("chysx::cloth::step");  // cloth_simulator.cu:815
    resize_work_buffers(n, cuda_stream);
    if (topology_dirty_ || H_num_block_rows_ != n) {
        ProfileRange rebuild("step::topology_rebuild");  // cloth_simulator.cu:1119
        ensure_hessian_topology(cuda_stream);
    } else {
        check_elements_current(cuda_stream);
    }
    engine::InstanceSource s = source();
    s.dt = dt;
    runner_.update_single(inst_.descriptor(s));
    runner_.step(cuda_stream);
}

chx_instance ClothSimulator::prepare_instance(float dt, std::uintptr_t cuda_stream) {
    const int n = buffers_.particle_count();
    if (n <= 0) throw std::runtime_error("ClothSimulator::prepare_instance: no particles bound");
    if (buffers_.pos.data() == nullptr || buffers_.vel.data() == nullptr)
        throw std::runtime_error("ClothSimulator::prepare_instance: pos / vel must be set via set_external_buffers()");
    if (!(dt > 0.0f)) throw std::invalid_argument("ClothSimulator::prepare_instance: dt must be positive");
    resize_work_buffers(n, cuda_stream);
    if (topology_dirty_ || H_num_block_rows_ != n) ensure_hessian_topology(cuda_stream);
    else check_elements_current(cuda_stream);
    engine::InstanceSource s = source();
    s.dt = dt;
    return inst_.descriptor(s);
}

int ClothSimulator::last_pcg_iterations(std::uintptr_t cuda_stream) {
    if (H_num_block_rows_ == 0) return 0;
    std::vector<int> it;
    std::vector<double> rr, bb;
    runner_.read_stats(it, rr, bb, cuda_stream);
    return it[0];
}

double ClothSimulator::last_relative_residual(std::uintptr_t cuda_stream) {
    if (H_num_block_rows_ == 0) return 0.0;
    std::vector<int> it;
    std::vector<double> rr, bb;
    runner_.read_stats(it, rr, bb, cuda_stream);
    return bb[0] > 0.0 ? std::sqrt(rr[0] / bb[0]) : std::sqrt(rr[0]);
}

namespace {
// Device blocks are grouped paired planes (per group of CHX_BLOCK_GROUP
// blocks: components (c0,c1) .. (c6,c7) as 64 pairs each, then c8 as 64
// words); the reference's debug taps return row-major 3x3 blocks.
inline std::size_t grouped(int c, std::size_t i) {
    const std::size_t G = CHX_BLOCK_GROUP, g = (i / G) * 9 * G, l = i % G;
    const std::size_t cc = static_cast<std::size_t>(c);
    return cc < 8 ? g + (cc / 2) * 2 * G + 2 * l + cc % 2 : g + 8 * G + l;
}
inline std::size_t group_floats(std::size_t count) {
    return 9 * ((count + CHX_BLOCK_GROUP - 1) / CHX_BLOCK_GROUP * CHX_BLOCK_GROUP);
}
template <class T>
void copy_planar_blocks(T* out, const T* dev, int count, const char* what) {
    std::vector<T> planes(group_floats(static_cast<std::size_t>(count)));
    check(chx_memcpy_d2h(planes.data(), dev, sizeof(T) * planes.size(), 0), "cloth", what);
    for (int i = 0; i < count; ++i)
        for (int c = 0; c < 9; ++c)
            out[static_cast<std::size_t>(9) * i + c] = planes[grouped(c, static_cast<std::size_t>(i))];
}
// Symmetric blocks (the diagonal) keep their upper triangle: groups of 64
// blocks as three planes of pairs (xx,xy), (xz,yy), (yz,zz).
template <class T>
void copy_sym_blocks(T* out, const T* dev, int count, const char* what) {
    const std::size_t G = CHX_BLOCK_GROUP;
    const std::size_t groups = (static_cast<std::size_t>(count) + G - 1) / G;
    std::vector<T> planes(6 * G * groups);
    check(chx_memcpy_d2h(planes.data(), dev, sizeof(T) * planes.size(), 0), "cloth", what);
    static const int pos[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int i = 0; i < count; ++i) {
        const std::size_t ii = static_cast<std::size_t>(i);
        for (int q = 0; q < 6; ++q) {
            const std::size_t qq = static_cast<std::size_t>(q);
            const T v = planes[(ii / G) * 6 * G + (qq / 2) * 2 * G + 2 * (ii % G) + qq % 2];
            out[9 * ii + 3 * pos[q][0] + pos[q][1]] = v;
            out[9 * ii + 3 * pos[q][1] + pos[q][0]] = v;
        }
    }
}
// Off-diagonal blocks live once per symmetric pair (upper blocks over
// n_upper value slots); out[k] = the block of CSR index k, row-major: the
// stored upper block, or its partner's transpose for a lower one.
template <class T>
void copy_sliced_blocks(T* out, const T* dev, int n_upper, const std::vector<int>& slots, int nnz, const char* what) {
    std::vector<T> planes(group_floats(static_cast<std::size_t>(n_upper)));
    check(chx_memcpy_d2h(planes.data(), dev, sizeof(T) * planes.size(), 0), "cloth", what);
    for (int k = 0; k < nnz; ++k) {
        const bool lower = slots[k] < 0;
        const std::size_t slot = static_cast<std::size_t>(lower ? -1 - slots[k] : slots[k]);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                const int c = lower ? 3 * b + a : 3 * a + b;
                out[static_cast<std::size_t>(9) * k + 3 * a + b] = planes[grouped(c, slot)];
            }
    }
}
}  // namespace

void ClothSimulator::debug_copy_hessian_diag(float* out) const {
    const int n = H_num_block_rows_;
    if (n == 0) return;
    if (out == nullptr) throw std::invalid_argument("ClothSimulator::debug_copy_hessian_diag: out must be non-null");
    check(chx_device_synchronize(), "cloth", "device synchronize(debug_copy_hessian_diag)");
    copy_sym_blocks(out, inst_.diag(), n, "diag copy");
}

void ClothSimulator::debug_copy_bf16_matrix(float* out_diag_c, uint16_t* out_values16) const {
    const int n = H_num_block_rows_;
    if (n == 0) return;
    if (!bf16_ || inst_.diag_c() == nullptr)
        throw std::logic_error("ClothSimulator::debug_copy_bf16_matrix: BF16 mode is off or not built yet");
    check(chx_device_synchronize(), "cloth", "device synchronize(debug_copy_bf16_matrix)");
    copy_planar_blocks(out_diag_c, inst_.diag_c(), n, "diag_c copy");
    if (pattern_.nnz() > 0)
        copy_sliced_blocks(out_values16, inst_.values16(), inst_.n_upper(), inst_.block_slots(), pattern_.nnz(),
                           "values16 copy");
}

void ClothSimulator::debug_copy_hessian_csr(int* out_row_offsets, int* out_col_indices, float* out_values) const {
    const int n = H_num_block_rows_;
    if (n == 0) return;
    if (out_row_offsets == nullptr)
        throw std::invalid_argument("ClothSimulator::debug_copy_hessian_csr: out_row_offsets null");
    check(chx_device_synchronize(), "cloth", "device synchronize(debug_copy_hessian_csr)");
    std::memcpy(out_row_offsets, pattern_.row_offsets.data(), sizeof(int) * (n + 1));
    const int nnz = pattern_.nnz();
    if (nnz > 0) {
        if (out_col_indices == nullptr || out_values == nullptr)
            throw std::invalid_argument(
                "ClothSimulator::debug_copy_hessian_csr: col_indices/values null but nnz_off > 0");
        std::memcpy(out_col_indices, pattern_.col_indices.data(), sizeof(int) * nnz);
        copy_sliced_blocks(out_values, inst_.values(), inst_.n_upper(), inst_.block_slots(), nnz, "values copy");
    }
}

void ClothSimulator::debug_copy_last_rhs(float* out) const {
    const int n = H_num_block_rows_;
    if (n == 0) return;
    if (out == nullptr) throw std::invalid_argument("ClothSimulator::debug_copy_last_rhs: out must be non-null");
    check(chx_device_synchronize(), "cloth", "device synchronize(debug_copy_last_rhs)");
    check(chx_memcpy_d2h(out, inst_.rhs(), sizeof(float) * 3 * n, 0), "cloth", "rhs copy");
}

void ClothSimulator::debug_copy_last_dx(float* out) const {
    const int n = H_num_block_rows_;
    if (n == 0) return;
    if (out == nullptr) throw std::invalid_argument("ClothSimulator::debug_copy_last_dx: out must be non-null");
    check(chx_device_synchronize(), "cloth", "device synchronize(debug_copy_last_dx)");
    check(chx_memcpy_d2h(out, dx_.gpu_data(), sizeof(float) * 3 * n, 0), "cloth", "dx copy");
}

void ClothSimulator::debug_set_warm_start(const float* in) {
    const int n = buffers_.particle_count();
    if (n <= 0) return;
    if (in == nullptr) throw std::invalid_argument("ClothSimulator::debug_set_warm_start: null input");
    resize_work_buffers(n);
    check(chx_memcpy_h2d(dx_.gpu_data(), in, sizeof(float) * 3 * n, 0), "cloth", "dx upload");
}

double ClothSimulator::compute_energy(int kind_mask) {
    const int n = buffers_.particle_count();
    if (n <= 0) return 0.0;
    resize_work_buffers(n);
    if (topology_dirty_ || H_num_block_rows_ != n) ensure_hessian_topology();
    else check_elements_current();
    runner_.update_single(inst_.descriptor(source()));
    return runner_.energy(kind_mask, 0);
}

}  // namespace cloth
}  // namespace chysx
