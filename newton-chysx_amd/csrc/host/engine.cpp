#include "engine.h"

#include <algorithm>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace chysx {
namespace engine {

namespace {

// Host-side gather-map entry over global element indices: bits 31..30
// kind, 29..28 a, 27..26 b, 25..0 the element index.  build_units turns it
// into the device's 16-bit entry over unit-local positions.
constexpr unsigned kMaxElement = (1u << 26) - 1u;
unsigned encode(unsigned kind, int a, int b, std::size_t id) {
    if (id > kMaxElement)
        throw std::runtime_error("chysx::engine: element index exceeds the 26-bit gather-map field");
    return (kind << 30) | (unsigned(a) << 28) | (unsigned(b) << 26) | static_cast<unsigned>(id);
}
constexpr long pad8(long v) { return (v + 7) & ~7L; }
// LDS record floats of an element of kind k (CHX_CONTRIB_*).
constexpr long asm_rec_floats(int k) {
    return k == int(CHX_CONTRIB_PIN) ? CHX_ASM_PIN_REC
                                     : (k == int(CHX_CONTRIB_TRI) ? CHX_ASM_TRI_REC
                                                                  : (k == int(CHX_CONTRIB_DIH) ? CHX_ASM_DIH_REC
                                                                                               : CHX_ASM_STVK_REC));
}

template <class T> void upload(DeviceArray<T>& a, std::size_t& bytes, std::uintptr_t stream) {
    a.copy_to_device(stream);
    bytes += a.gpu_size() * sizeof(T);
}
template <class T> void alloc(DeviceArray<T>& a, std::size_t n, std::size_t& bytes) {
    a.ensure_device(n);
    bytes += n * sizeof(T);
}

int item_count(int n) { return (n + CHX_ITEM_SIZE - 1) / CHX_ITEM_SIZE; }

// Waits for `stream` only, or for the whole device on the null stream.
void sync_stream(std::uintptr_t stream, const char* what) {
    if (stream == 0) check(chx_device_synchronize(), "engine", what);
    else check(chx_stream_synchronize(stream), "engine", what);
}

}  // namespace

// ---------------------------------------------------------------------------
// InstanceData
// ---------------------------------------------------------------------------

void InstanceData::pack_elements(const InstanceSource& s) {
    auto& P = pins_.host();
    auto& PT = pin_targets_.host();
    auto& T = tris_.host();
    auto& D = dihs_.host();
    auto& V = stvks_.host();
    P.clear();
    PT.clear();
    T.clear();
    D.clear();
    V.clear();
    if (s.pins) {
        const float k = s.pins->stiffness();
        for (int c = 0; c < s.pins->size(); ++c) {
            const auto& t = s.pins->targets()[c];
            P.push_back(chx_pin{s.pins->indices()[c], k});
            PT.insert(PT.end(), {t.x, t.y, t.z});
        }
    }
    for (const constraint::TriangleStretchConstraint* ts : {s.stretch, s.shear}) {
        if (!ts) continue;
        const float k = ts->stiffness();
        for (int t = 0; t < ts->size(); ++t) {
            const auto& v = ts->indices()[t];
            const auto& dm = ts->Dm_inv()[t].data;
            // k_eff = areas[t] * k (triangle_stretch_constraint.cu:193).
            T.push_back(chx_tri{v.x, v.y, v.z, 0, dm[0], dm[1], dm[2], dm[3], ts->areas()[t] * k});
        }
    }
    if (s.stvk) {
        const float mu = s.stvk->mu(), lam = s.stvk->lambda(), h = s.stvk->thickness();
        for (int t = 0; t < s.stvk->size(); ++t) {
            const auto& v = s.stvk->indices()[t];
            const auto& dm = s.stvk->Dm_inv()[t].data;
            const float ah = s.stvk->areas()[t] * h;
            V.push_back(chx_stvk{v.x, v.y, v.z, 0, dm[0], dm[1], dm[2], dm[3], mu * ah, lam * ah, 0, 0});
        }
    }
    if (s.bending) {
        const float k = s.bending->stiffness();
        for (int e = 0; e < s.bending->size(); ++e) {
            const auto& v = s.bending->indices()[e];
            D.push_back(chx_dih{v.x, v.y, v.z, v.w, s.bending->rest_angles()[e], k, 0, 0});
        }
    }
}

// Assembly units (chx_unit): the rows are cut into units whose LDS image
// fits CHX_ASM_LDS_BYTES -- the records of every element with a vertex in
// the unit's rows, the contribution entries of the unit's slots and the
// slots' offsets -- with at most CHX_CONTRIB_MAX_LOCAL + 1 elements of a
// kind: up to four whole slices, or the fewest runs of near-equal rows of
// one slice when the slice alone does not fit (the greedy split, full runs
// and a short last one, made the assembly of the 512-cloth batch 5 % slower).  Every unit gets the list of its elements (ascending per
// kind) and the contribution lists of its slots (diagonal blocks of its
// rows, then its upper value slots slice by slice, block index jj, then
// row) as 16-bit entries whose element field is the element's position in
// the unit's list of that kind; the entries and the offsets are padded to
// 8-entry (16-byte) runs so the kernel stages them with 16-byte loads.
// goff / gent: the per-slot lists over the slot order [diagonal blocks
// 0..n) + [value slots n..n+n_upper) with global element indices.
void InstanceData::build_units(const std::vector<unsigned>& goff, const std::vector<unsigned>& gent) {
    const int n = n_;
    constexpr int kKinds = 4;  // CHX_CONTRIB_PIN, _TRI, _DIH, _STVK
    static_assert(CHX_CONTRIB_PIN == 0 && CHX_CONTRIB_TRI == 1 && CHX_CONTRIB_DIH == 2 && CHX_CONTRIB_STVK == 3,
                  "kind codes index the per-kind tables");
    const std::size_t counts[kKinds] = {pins_.cpu_size(), tris_.cpu_size(), dihs_.cpu_size(), stvks_.cpu_size()};
    const int SR = CHX_SLICE_ROWS;
    const auto& uo = upper_offsets_.host();
    // slots and contribution entries each row brings to its unit
    std::vector<long> row_slots(static_cast<std::size_t>(n)), row_words(static_cast<std::size_t>(n));
    for (int r = 0; r < n; ++r) {
        const int sl = r / SR, width = (uo[sl + 1] - uo[sl]) / SR;
        long w = goff[static_cast<std::size_t>(r) + 1] - goff[static_cast<std::size_t>(r)];
        for (int jj = 0; jj < width; ++jj) {
            const std::size_t g = static_cast<std::size_t>(n) + static_cast<std::size_t>(uo[sl] + jj * SR + (r - sl * SR));
            w += goff[g + 1] - goff[g];
        }
        row_slots[static_cast<std::size_t>(r)] = 1 + width;
        row_words[static_cast<std::size_t>(r)] = w;
    }
    // vertex -> incident elements, per kind (CSR)
    std::vector<int> voff[kKinds], vel[kKinds];
    auto incidence = [&](int kind, int per, auto vertex_of) {
        auto& off = voff[kind];
        auto& el = vel[kind];
        off.assign(static_cast<std::size_t>(n) + 1, 0);
        for (std::size_t e = 0; e < counts[kind]; ++e)
            for (int j = 0; j < per; ++j) ++off[static_cast<std::size_t>(vertex_of(e, j)) + 1];
        for (int v = 0; v < n; ++v) off[v + 1] += off[v];
        el.assign(static_cast<std::size_t>(off[n]), 0);
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (std::size_t e = 0; e < counts[kind]; ++e)
            for (int j = 0; j < per; ++j) el[fill[vertex_of(e, j)]++] = static_cast<int>(e);
    };
    incidence(CHX_CONTRIB_PIN, 1, [&](std::size_t e, int) { return pins_.host()[e].v; });
    incidence(CHX_CONTRIB_TRI, 3, [&](std::size_t e, int j) {
        const chx_tri& t = tris_.host()[e];
        return j == 0 ? t.v0 : (j == 1 ? t.v1 : t.v2);
    });
    incidence(CHX_CONTRIB_DIH, 4, [&](std::size_t e, int j) {
        const chx_dih& d = dihs_.host()[e];
        return j == 0 ? d.v0 : (j == 1 ? d.v1 : (j == 2 ? d.v2 : d.v3));
    });
    incidence(CHX_CONTRIB_STVK, 3, [&](std::size_t e, int j) {
        const chx_stvk& t = stvks_.host()[e];
        return j == 0 ? t.v0 : (j == 1 ? t.v1 : t.v2);
    });
    // stamp[k][e]: the unit that holds element e; mark[k][e]: trial token
    std::vector<int> stamp[kKinds], mark[kKinds];
    for (int k = 0; k < kKinds; ++k) {
        stamp[k].assign(counts[k], -1);
        mark[k].assign(counts[k], -1);
    }
    int token = 0;
    // What rows [ra, rb) add to unit uid: new elements per kind, slots, entries.
    struct Need {
        long elems[kKinds] = {0, 0, 0, 0};
        long slots = 0, words = 0;
        void add(const Need& o) {
            for (int k = 0; k < kKinds; ++k) elems[k] += o.elems[k];
            slots += o.slots;
            words += o.words;
        }
        bool fits() const {
            long floats = 0;
            for (int k = 0; k < kKinds; ++k) {
                if (elems[k] > static_cast<long>(CHX_CONTRIB_MAX_LOCAL) + 1) return false;
                floats += elems[k] * asm_rec_floats(k);
            }
            return ((4 * floats + 15) & ~15L) + 2 * pad8(words) + 2 * pad8(slots + 1) <= CHX_ASM_LDS_BYTES;
        }
    };
    auto added = [&](int ra, int rb, int uid) {
        ++token;
        Need d;
        for (int r = ra; r < rb; ++r) {
            d.slots += row_slots[static_cast<std::size_t>(r)];
            d.words += row_words[static_cast<std::size_t>(r)];
            for (int k = 0; k < kKinds; ++k)
                for (int q = voff[k][r]; q < voff[k][r + 1]; ++q) {
                    const int e = vel[k][q];
                    if (stamp[k][e] != uid && mark[k][e] != token) {
                        mark[k][e] = token;
                        ++d.elems[k];
                    }
                }
        }
        return d;
    };
    auto commit = [&](int ra, int rb, int uid) {
        for (int r = ra; r < rb; ++r)
            for (int k = 0; k < kKinds; ++k)
                for (int q = voff[k][r]; q < voff[k][r + 1]; ++q) stamp[k][vel[k][q]] = uid;
    };
    const int slices = (n + SR - 1) / SR;
    std::vector<std::pair<int, int>> ranges;
    int u_r0 = -1, u_slices = 0;
    Need u_need;
    for (int sl = 0; sl < slices; ++sl) {
        const int sa = sl * SR, sb = std::min(n, sa + SR);
        const int uid = static_cast<int>(ranges.size());
        if (u_r0 >= 0) {
            Need t = u_need;
            t.add(added(sa, sb, uid));
            if (u_slices < 4 && t.fits()) {
                commit(sa, sb, uid);
                u_need = t;
                ++u_slices;
                continue;
            }
            ranges.emplace_back(u_r0, sa);
            u_r0 = -1;
        }
        const int nid = static_cast<int>(ranges.size());
        const Need whole = added(sa, sb, nid);
        if (whole.fits()) {
            commit(sa, sb, nid);
            u_r0 = sa;
            u_need = whole;
            u_slices = 1;
            continue;
        }
        // the slice alone does not fit: split it into the fewest runs of
        // near-equal rows that fit (uid -2 is no unit's: added() then counts
        // every element of the rows)
        const int len = sb - sa;
        int parts = 2;
        for (;; ++parts) {
            if (parts > len) break;
            bool ok = true;
            for (int i = 0; i < parts && ok; ++i)
                ok = added(sa + i * len / parts, sa + (i + 1) * len / parts, -2).fits();
            if (ok) break;
        }
        if (parts > len) {
            for (int r = sa; r < sb; ++r)
                if (!added(r, r + 1, -2).fits())
                    throw std::invalid_argument(
                        "chysx::engine: the elements of one vertex exceed an assembly unit's LDS image "
                        "(CHX_ASM_LDS_BYTES)");
            parts = len;
        }
        for (int i = 0; i < parts; ++i) {
            const int a = sa + i * len / parts, b = sa + (i + 1) * len / parts;
            commit(a, b, static_cast<int>(ranges.size()));
            ranges.emplace_back(a, b);
        }
    }
    if (u_r0 >= 0) ranges.emplace_back(u_r0, n);

    auto& units = units_.host();
    auto& uel = unit_elems_.host();
    auto& off = contrib_offsets_.host();
    auto& ent = contribs_.host();
    units.clear();
    uel.clear();
    off.clear();
    ent.clear();
    ent.reserve(gent.size() + 8 * ranges.size());
    std::vector<int> local[kKinds];
    for (int k = 0; k < kKinds; ++k) local[k].assign(counts[k], -1);
    std::vector<int> list[kKinds];
    for (std::size_t u = 0; u < ranges.size(); ++u) {
        const int r0 = ranges[u].first, r1 = ranges[u].second;
        chx_unit U{};
        U.row_begin = r0;
        U.row_end = r1;
        U.elem_begin = static_cast<int>(uel.size());
        U.offs_begin = static_cast<int>(off.size());
        U.word_begin = static_cast<int>(ent.size());
        ++token;
        for (int k = 0; k < kKinds; ++k) list[k].clear();
        for (int r = r0; r < r1; ++r)
            for (int k = 0; k < kKinds; ++k)
                for (int q = voff[k][r]; q < voff[k][r + 1]; ++q) {
                    const int e = vel[k][q];
                    if (mark[k][e] != token) {
                        mark[k][e] = token;
                        list[k].push_back(e);
                    }
                }
        Need need;
        const int order[kKinds] = {CHX_CONTRIB_PIN, CHX_CONTRIB_TRI, CHX_CONTRIB_STVK, CHX_CONTRIB_DIH};
        for (int k : order) {
            std::sort(list[k].begin(), list[k].end());
            for (std::size_t i = 0; i < list[k].size(); ++i) local[k][list[k][i]] = static_cast<int>(i);
            uel.insert(uel.end(), list[k].begin(), list[k].end());
            need.elems[k] = static_cast<long>(list[k].size());
        }
        U.n_pins = static_cast<int>(list[CHX_CONTRIB_PIN].size());
        U.n_tris = static_cast<int>(list[CHX_CONTRIB_TRI].size());
        U.n_stvks = static_cast<int>(list[CHX_CONTRIB_STVK].size());
        U.n_dihs = static_cast<int>(list[CHX_CONTRIB_DIH].size());
        const std::size_t wbase = ent.size();
        auto emit = [&](std::size_t g) {
            off.push_back(static_cast<uint16_t>(ent.size() - wbase));
            for (unsigned c = goff[g]; c < goff[g + 1]; ++c) {
                const unsigned w = gent[c];
                const unsigned kind = w >> 30;
                const int e = static_cast<int>(w & kMaxElement);
                const int l = local[kind][static_cast<std::size_t>(e)];
                if (l < 0 || stamp[kind][static_cast<std::size_t>(e)] < 0)
                    throw std::logic_error("chysx::engine: contribution outside its assembly unit");
                ent.push_back(static_cast<uint16_t>((kind << 14) | (((w >> 28) & 3u) << 12) |
                                                    (((w >> 26) & 3u) << 10) | static_cast<unsigned>(l)));
            }
        };
        for (int r = r0; r < r1; ++r) emit(static_cast<std::size_t>(r));
        for (int sl = r0 / SR; sl <= (r1 - 1) / SR; ++sl) {
            const int a = std::max(r0, sl * SR), b = std::min(r1, (sl + 1) * SR);
            const int width = (uo[sl + 1] - uo[sl]) / SR;
            for (int jj = 0; jj < width; ++jj)
                for (int r = a; r < b; ++r)
                    emit(static_cast<std::size_t>(n) + static_cast<std::size_t>(uo[sl] + jj * SR + (r - sl * SR)));
        }
        U.n_slots = static_cast<int>(off.size()) - U.offs_begin;
        need.slots = U.n_slots;
        need.words = static_cast<long>(ent.size() - wbase);
        if (!need.fits()) throw std::logic_error("chysx::engine: assembly unit over its LDS capacity");
        off.push_back(static_cast<uint16_t>(ent.size() - wbase));
        while (off.size() % 8 != 0) off.push_back(off.back());
        while ((ent.size() - wbase) % 8 != 0) ent.push_back(0);
        U.n_words = static_cast<int>(ent.size() - wbase);
        for (int k = 0; k < kKinds; ++k)
            for (int e : list[k]) local[k][e] = -1;
        units.push_back(U);
    }
}

void InstanceData::build(const InstanceSource& s, std::uintptr_t stream) {
    if (!s.pattern || s.pattern->n != s.n) throw std::invalid_argument("chysx::engine: pattern missing or sized wrong");
    bytes_ = 0;
    n_ = s.n;
    nnz_ = s.pattern->nnz();
    row_offsets_.host() = s.pattern->row_offsets;
    col_indices_.host() = s.pattern->col_indices;
    pack_elements(s);

    // Symmetric sliced layout (slice s = rows 64s..64s+63).  Entries: the
    // j-th off-diagonal column of row 64s+l at entry slot
    // slice_offsets[s] + 64 j + l, encoded (col << 6) | jj with jj the
    // block's index among its owner row's upper blocks (the owner is the
    // smaller index: the row for col > row, the column for col < row).
    // Values: upper block jj of row 64s+l at value slot
    // upper_offsets[s] + 64 jj + l; lower blocks are their partners'
    // transposes.
    {
        const auto& ro = row_offsets_.host();
        const auto& ci = col_indices_.host();
        const int S_ = CHX_SLICE_ROWS;
        if (n_ >= (1 << 26)) throw std::invalid_argument("chysx::engine: a cloth instance holds fewer than 2^26 vertices");
        const int slices = (n_ + S_ - 1) / S_;
        // first upper block of each row (columns ascending: lower ones first)
        std::vector<int> up0(static_cast<std::size_t>(n_));
        for (int r = 0; r < n_; ++r) {
            int k = ro[r];
            while (k < ro[r + 1] && ci[k] < r) ++k;
            up0[r] = k;
            if (ro[r + 1] - k > 64)
                throw std::invalid_argument("chysx::engine: a vertex couples to more than 64 higher-numbered vertices");
        }
        auto& so = slice_offsets_.host();
        auto& uo = upper_offsets_.host();
        so.assign(static_cast<std::size_t>(slices) + 1, 0);
        uo.assign(static_cast<std::size_t>(slices) + 1, 0);
        std::size_t total = 0, utotal = 0;
        for (int sl = 0; sl < slices; ++sl) {
            int width = 0, uwidth = 1;
            for (int r = sl * S_; r < std::min(n_, (sl + 1) * S_); ++r) {
                width = std::max(width, ro[r + 1] - ro[r]);
                uwidth = std::max(uwidth, ro[r + 1] - up0[r]);
            }
            total += static_cast<std::size_t>(S_) * width;
            utotal += static_cast<std::size_t>(S_) * uwidth;
            if (total > static_cast<std::size_t>(std::numeric_limits<int>::max()) ||
                utotal > static_cast<std::size_t>(std::numeric_limits<int>::max()))
                throw std::invalid_argument("chysx::engine: sliced block layout exceeds 2^31 slots");
            so[sl + 1] = static_cast<int>(total);
            uo[sl + 1] = static_cast<int>(utotal);
        }
        npad_ = static_cast<int>(total);
        nup_ = static_cast<int>(utotal);
        auto vslot = [&](int owner, int jj) { return uo[owner / S_] + jj * S_ + owner % S_; };
        // jj of the block (owner, col) among owner's upper blocks
        auto upper_index = [&](int owner, int col) {
            const auto b = ci.begin() + up0[owner], e = ci.begin() + ro[owner + 1];
            return static_cast<int>(std::lower_bound(b, e, col) - b);
        };
        auto& se = slice_entries_.host();
        auto& bs = block_slots_.host();
        se.assign(std::max<std::size_t>(total, 1), 0);
        bs.assign(std::max<std::size_t>(static_cast<std::size_t>(nnz_), 1), 0);
        for (int sl = 0; sl < slices; ++sl) {
            const int width = (so[sl + 1] - so[sl]) / S_;
            for (int l = 0; l < S_; ++l) {
                const int r = sl * S_ + l;
                const int len = r < n_ ? ro[r + 1] - ro[r] : 0;
                for (int j = 0; j < width; ++j) {
                    const int slot = so[sl] + j * S_ + l;
                    if (j < len) {
                        const int k = ro[r] + j, c = ci[k];
                        const bool lower = c < r;
                        const int jj = lower ? upper_index(c, r) : k - up0[r];
                        se[slot] = (c << 6) | jj;
                        bs[k] = lower ? -1 - vslot(c, jj) : vslot(r, jj);
                    } else {
                        se[slot] = std::min(r, n_ - 1) << 6;
                    }
                }
            }
        }
        // 16-bit form ((col - row + 2048) << 4 | jj, row = the slot's row),
        // used when every entry fits: grids up to ~1000 vertices wide.
        auto& s16 = slice_entries16_.host();
        s16.assign(std::max<std::size_t>(total, 1), 0);
        entries16_ = true;
        for (int sl = 0; sl < slices && entries16_; ++sl) {
            const int width = (so[sl + 1] - so[sl]) / S_;
            for (int l = 0; l < S_ && entries16_; ++l) {
                const int r = sl * S_ + l;
                for (int j = 0; j < width; ++j) {
                    const int slot = so[sl] + j * S_ + l;
                    const int c = se[slot] >> 6, jj = se[slot] & 63, off = c - r;
                    if (off < -2048 || off > 2047 || jj > 15) {
                        entries16_ = false;
                        break;
                    }
                    s16[slot] = static_cast<uint16_t>(((off + 2048) << 4) | jj);
                }
            }
        }
        if (!entries16_) s16.clear();
    }


    // Gather map.  Slots: diagonal block of row r -> r, upper off-diagonal
    // block of CSR index k -> n + its value slot (block_slots[k]), so the
    // gather writes value slots in order; lower blocks (their partners'
    // transposes) are not assembled and get no entries.  Entries are
    // appended in element order (pins, stretch, shear, StVK, bending;
    // element; a; b) — the order of the reference's launches, StVK (no
    // reference counterpart) after the BW98 triangles — so every slot sums
    // its terms in that order.
    const std::size_t S = static_cast<std::size_t>(n_) + static_cast<std::size_t>(nup_);
    constexpr std::size_t kNoSlot = ~std::size_t(0);
    const auto& bslots = block_slots_.host();
    auto slot_of = [&](int slot) -> std::size_t {
        if (slot < 0) return static_cast<std::size_t>(-slot - 1);
        const int v = bslots[static_cast<std::size_t>(slot)];
        return v < 0 ? kNoSlot : static_cast<std::size_t>(n_) + static_cast<std::size_t>(v);
    };
    auto walk = [&](auto&& visit) {
        if (s.pins) {
            const auto& sl = s.pins->hessian_slots();
            if (static_cast<int>(sl.size()) != s.pins->size()) throw std::runtime_error("chysx::engine: pin slots not bound");
            for (int c = 0; c < s.pins->size(); ++c) visit(slot_of(sl[c]), CHX_CONTRIB_PIN, 0, 0, std::size_t(c));
        }
        std::size_t tri_base = 0;
        for (const constraint::TriangleStretchConstraint* ts : {s.stretch, s.shear}) {
            if (!ts) continue;
            const auto& sl = ts->hessian_slots();
            if (sl.size() != 9u * ts->size()) throw std::runtime_error("chysx::engine: triangle slots not bound");
            for (int t = 0; t < ts->size(); ++t)
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) visit(slot_of(sl[9 * t + 3 * a + b]), CHX_CONTRIB_TRI, a, b, tri_base + t);
            tri_base += ts->size();
        }
        if (s.stvk) {
            const auto& sl = s.stvk->hessian_slots();
            if (sl.size() != 9u * s.stvk->size()) throw std::runtime_error("chysx::engine: StVK slots not bound");
            for (int t = 0; t < s.stvk->size(); ++t)
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        visit(slot_of(sl[9 * t + 3 * a + b]), CHX_CONTRIB_STVK, a, b, std::size_t(t));
        }
        if (s.bending) {
            const auto& sl = s.bending->hessian_slots();
            if (sl.size() != 16u * s.bending->size()) throw std::runtime_error("chysx::engine: dihedral slots not bound");
            for (int e = 0; e < s.bending->size(); ++e)
                for (int a = 0; a < 4; ++a)
                    for (int b = 0; b < 4; ++b)
                        visit(slot_of(sl[16 * e + 4 * a + b]), CHX_CONTRIB_DIH, a, b, std::size_t(e));
        }
    };
    std::vector<unsigned long long> count(S + 1, 0);
    walk([&](std::size_t g, unsigned, int, int, std::size_t) {
        if (g != kNoSlot) ++count[g + 1];
    });
    for (std::size_t k = 0; k < S; ++k) count[k + 1] += count[k];
    if (count[S] > std::numeric_limits<unsigned>::max())
        throw std::runtime_error("chysx::engine: gather map exceeds 2^32 entries");
    // per-slot lists with global element indices (slot order as above)
    std::vector<unsigned> goff(S + 1);
    for (std::size_t k = 0; k <= S; ++k) goff[k] = static_cast<unsigned>(count[k]);
    std::vector<unsigned> gent(count[S], 0u);
    {
        std::vector<unsigned> fill(goff.begin(), goff.end() - 1);
        walk([&](std::size_t g, unsigned kind, int a, int b, std::size_t id) {
            if (g != kNoSlot) gent[fill[g]++] = encode(kind, a, b, id);
        });
    }
    build_units(goff, gent);
    const auto& ent = contribs_.host();

    // The kernels address every per-instance array with 32-bit byte offsets.
    const auto fits = [](std::size_t count, std::size_t bytes_each) {
        return count * bytes_each < (std::size_t(1) << 32);
    };
    if (!fits(static_cast<std::size_t>(n_) + CHX_BLOCK_GROUP, 36) || !fits(static_cast<std::size_t>(nup_), 36) ||
        !fits(static_cast<std::size_t>(npad_), 4) ||
        !fits(tris_.cpu_size(), sizeof(chx_tri)) || !fits(dihs_.cpu_size(), sizeof(chx_dih)) ||
        !fits(pins_.cpu_size(), sizeof(chx_pin)) || !fits(ent.size(), sizeof(uint16_t)) ||
        !fits(stvks_.cpu_size(), sizeof(chx_stvk)) || !fits(unit_elems_.cpu_size(), sizeof(int)))
        throw std::invalid_argument(
            "chysx::engine: a per-instance array would reach 4 GiB (at most 119,304,647 vertices or off-diagonal "
            "blocks per cloth); split the cloth into several instances");
    upload(row_offsets_, bytes_, stream);
    upload(col_indices_, bytes_, stream);
    upload(slice_offsets_, bytes_, stream);
    upload(slice_entries_, bytes_, stream);
    if (entries16_) upload(slice_entries16_, bytes_, stream);
    else slice_entries16_.release();
    upload(upper_offsets_, bytes_, stream);
    upload(block_slots_, bytes_, stream);
    upload(contrib_offsets_, bytes_, stream);
    upload(contribs_, bytes_, stream);
    upload(units_, bytes_, stream);
    upload(unit_elems_, bytes_, stream);
    upload(pins_, bytes_, stream);
    upload(pin_targets_, bytes_, stream);
    upload(tris_, bytes_, stream);
    upload(dihs_, bytes_, stream);
    upload(stvks_, bytes_, stream);
    const std::size_t n = static_cast<std::size_t>(n_);
    // grouped planar block arrays hold whole groups of CHX_BLOCK_GROUP blocks
    const std::size_t ng = (n + CHX_BLOCK_GROUP - 1) / CHX_BLOCK_GROUP * CHX_BLOCK_GROUP;
    const std::size_t nug = std::max<std::size_t>(static_cast<std::size_t>(nup_), CHX_BLOCK_GROUP);
    alloc(diag_, 9 * ng, bytes_);
    alloc(values_, 9 * nug, bytes_);
    alloc(rhs_, 3 * n, bytes_);
    alloc(r_, 3 * n, bytes_);
    alloc(z_, 3 * n, bytes_);
    alloc(p0_, 3 * n, bytes_);
    alloc(p1_, 3 * n, bytes_);
    alloc(q_, 3 * n, bytes_);
    alloc(q1_, 3 * n, bytes_);
    alloc(minv_, 9 * ng, bytes_);
    // Column window of every stage of CHX_STAGE_ROWS rows (its own rows
    // included): the rows of the source vector its SpMV reads.
    {
        const auto& ro = row_offsets_.host();
        const auto& ci = col_indices_.host();
        auto& w = windows_.host();
        const int stages = (n_ + CHX_STAGE_ROWS - 1) / CHX_STAGE_ROWS;
        w.assign(2 * static_cast<std::size_t>(stages), 0);
        for (int st = 0; st < stages; ++st) {
            const int r0 = st * CHX_STAGE_ROWS, r1 = std::min(n_, r0 + CHX_STAGE_ROWS);
            int lo = r0, hi = r1 - 1;
            for (int k = ro[r0]; k < ro[r1]; ++k) {
                lo = std::min(lo, ci[k]);
                hi = std::max(hi, ci[k]);
            }
            w[2 * st] = lo;
            w[2 * st + 1] = hi;
        }
        upload(windows_, bytes_, stream);
    }
    update_shapes(s, stream);
    if (s.bf16_blocks) {
        alloc(diag_c_, 9 * ng, bytes_);
        alloc(values16_, 9 * nug, bytes_);
    } else {
        diag_c_.release();
        values16_.release();
    }
    built_ = true;
}

void InstanceData::update_shapes(const InstanceSource& s, std::uintptr_t stream) {
    planes_.host() = s.planes ? *s.planes : std::vector<chx_plane>{};
    boxes_.host() = s.boxes ? *s.boxes : std::vector<chx_box>{};
    if (planes_.cpu_size() > 0) planes_.copy_to_device(stream);
    if (boxes_.cpu_size() > 0) boxes_.copy_to_device(stream);
}

void InstanceData::update_elements(const InstanceSource& s, std::uintptr_t stream) {
    const std::size_t nP = pins_.cpu_size(), nT = tris_.cpu_size(), nD = dihs_.cpu_size(), nV = stvks_.cpu_size();
    pack_elements(s);
    if (pins_.cpu_size() != nP || tris_.cpu_size() != nT || dihs_.cpu_size() != nD || stvks_.cpu_size() != nV)
        throw std::runtime_error("chysx::engine: element counts changed; rebuild the instance");
    pins_.copy_to_device(stream);
    pin_targets_.copy_to_device(stream);
    tris_.copy_to_device(stream);
    dihs_.copy_to_device(stream);
    stvks_.copy_to_device(stream);
}

void InstanceData::upload_pin_targets(const float* host_xyz, int n, std::uintptr_t stream) {
    if (n <= 0) return;
    if (static_cast<std::size_t>(n) * 3 != pin_targets_.cpu_size() || pin_targets_.gpu_size() != pin_targets_.cpu_size())
        throw std::invalid_argument("chysx::engine: pin target count differs from the built pin set");
    std::memcpy(pin_targets_.cpu_data(), host_xyz, sizeof(float) * 3 * static_cast<std::size_t>(n));
    stager_.upload(pin_targets_.gpu_data(), host_xyz, sizeof(float) * 3 * static_cast<std::size_t>(n), stream);
}

chx_instance InstanceData::descriptor(const InstanceSource& s) const {
    chx_instance I;
    std::memset(&I, 0, sizeof(I));  // padding too: descriptors are compared bytewise
    I.pos = s.pos;
    I.vel = s.vel;
    I.inv_mass = s.inv_mass;
    I.row_offsets = row_offsets_.gpu_data();
    I.col_indices = col_indices_.gpu_data();
    I.diag = diag_.gpu_data();
    I.values = values_.gpu_data();
    I.pins = pins_.gpu_data();
    I.pin_targets = pin_targets_.gpu_data();
    I.tris = tris_.gpu_data();
    I.dihs = dihs_.gpu_data();
    I.units = units_.gpu_data();
    I.unit_elems = unit_elems_.gpu_data();
    I.n_units = static_cast<int>(units_.cpu_size());
    I.stvks = stvks_.gpu_data();
    I.contrib_offsets = contrib_offsets_.gpu_data();
    I.contribs = contribs_.gpu_data();
    I.rhs = rhs_.gpu_data();
    I.x = s.x;
    I.r = r_.gpu_data();
    I.z = z_.gpu_data();
    I.p0 = p0_.gpu_data();
    I.p1 = p1_.gpu_data();
    I.q = q_.gpu_data();
    I.q1 = q1_.gpu_data();
    I.minv = minv_.gpu_data();
    I.diag_c = s.bf16_blocks ? diag_c_.gpu_data() : nullptr;
    I.values16 = s.bf16_blocks ? values16_.gpu_data() : nullptr;
    I.windows = windows_.gpu_data();
    I.slice_offsets = slice_offsets_.gpu_data();
    I.slice_entries = slice_entries_.gpu_data();
    I.slice_entries16 = entries16_ ? slice_entries16_.gpu_data() : nullptr;
    I.upper_offsets = upper_offsets_.gpu_data();
    I.block_slots = block_slots_.gpu_data();
    I.n_padded = npad_;
    I.n_upper = nup_;
    I.planes = planes_.cpu_size() > 0 ? planes_.gpu_data() : nullptr;
    I.boxes = boxes_.cpu_size() > 0 ? boxes_.gpu_data() : nullptr;
    I.n_planes = static_cast<int>(planes_.cpu_size());
    I.n_boxes = static_cast<int>(boxes_.cpu_size());
    I.contact_thickness = s.contact_thickness;
    I.contact_stiffness = s.contact_stiffness;
    I.contact_friction = s.contact_friction;
    I.contact_friction_eps = s.contact_friction_eps;
    I.n_rows = n_;
    I.nnz = nnz_;
    I.n_pins = static_cast<int>(pins_.cpu_size());
    I.n_tris = static_cast<int>(tris_.cpu_size());
    I.n_dihs = static_cast<int>(dihs_.cpu_size());
    I.n_stvks = static_cast<int>(stvks_.cpu_size());
    I.tile_begin = 0;
    I.tile_end = item_count(n_);
    I.dt = s.dt;
    I.damping = s.damping;
    I.gx = s.gx;
    I.gy = s.gy;
    I.gz = s.gz;
    I.max_iter = s.max_iter;
    I.tol = s.tol;
    return I;
}

// ---------------------------------------------------------------------------
// BatchRunner
// ---------------------------------------------------------------------------

void BatchRunner::set_lanes_per_row(int L) {
    if (L != 4 && L != 8 && L != 16) throw std::invalid_argument("chysx::engine: lanes_per_row must be 4, 8 or 16");
    lanes_ = L;
    desc_.lanes_per_row = L;
}

void BatchRunner::size_outputs(int n_instances, int n_tiles, std::uintptr_t stream) {
    const std::size_t I = static_cast<std::size_t>(std::max(n_instances, 1));
    if (inst_iters_.gpu_size() < I) {
        inst_iters_.ensure_device(I);
        inst_done_.ensure_device(I);
        inst_rr_.ensure_device(I);
        inst_bb_.ensure_device(I);
        inst_rho_.ensure_device(I);
        inst_iters_.zero_device(stream);
        inst_rr_.zero_device(stream);
        inst_bb_.zero_device(stream);
        inst_rho_.zero_device(stream);
    }
    if (status_host_ == nullptr) {
        void* p = nullptr;
        check(chx_host_malloc(&p, 64), "engine", "status word");
        status_host_ = static_cast<unsigned*>(p);
        *status_host_ = 0u;
    }
    const std::size_t P = static_cast<std::size_t>(CHX_TILE_PARTIALS) * std::max(n_tiles, 1);
    if (partials_.gpu_size() < P) partials_.ensure_device(P);
    if (sync_.gpu_size() == 0) {
        sync_.ensure_device(CHX_SYNC_WORDS);
        sync_.zero_device(stream);
    }
    desc_.partials = partials_.gpu_data();
    desc_.inst_iters = inst_iters_.gpu_data();
    desc_.inst_rr = inst_rr_.gpu_data();
    desc_.inst_bb = inst_bb_.gpu_data();
    desc_.inst_done = inst_done_.gpu_data();
    desc_.inst_rho = inst_rho_.gpu_data();
    desc_.sync = sync_.gpu_data();
    desc_.status_host = status_host_;
}

void BatchRunner::release() noexcept {
    // a replay may still be running on the caller's stream
    if (graph_ != nullptr) (void)chx_device_synchronize();
    drop_graph();
    if (status_host_ != nullptr) {
        // a queued solve may still report into the word
        (void)chx_device_synchronize();
        (void)chx_host_free(status_host_);
        status_host_ = nullptr;
    }
}

BatchRunner& BatchRunner::operator=(BatchRunner&& o) noexcept {
    if (this == &o) return *this;
    release();
    desc_ = o.desc_;
    graph_ = o.graph_;
    graph_desc_ = o.graph_desc_;
    graph_captures_ = o.graph_captures_;
    lanes_ = o.lanes_;
    grid_ = o.grid_;
    tile_rows_ = o.tile_rows_;
    bf16_ = o.bf16_;
    table_host_ = std::move(o.table_host_);
    table_ = std::move(o.table_);
    tiles_ = std::move(o.tiles_);
    elem_items_ = std::move(o.elem_items_);
    unit_items_ = std::move(o.unit_items_);
    partials_ = std::move(o.partials_);
    inst_rr_ = std::move(o.inst_rr_);
    inst_bb_ = std::move(o.inst_bb_);
    inst_rho_ = std::move(o.inst_rho_);
    inst_iters_ = std::move(o.inst_iters_);
    inst_done_ = std::move(o.inst_done_);
    sync_ = std::move(o.sync_);
    trace_ = std::move(o.trace_);
    status_host_ = o.status_host_;
    o.graph_ = nullptr;
    o.status_host_ = nullptr;
    o.desc_ = chx_cloth_batch{};
    return *this;
}

void BatchRunner::set_trace(int entries) {
    if (entries <= 0) {
        desc_.trace = nullptr;
        desc_.trace_entries = 0;
        return;
    }
    check(chx_device_synchronize(), "engine", "trace buffer");
    trace_.ensure_device(2 * static_cast<std::size_t>(entries));
    desc_.trace = trace_.gpu_data();
    desc_.trace_entries = entries;
}

std::vector<unsigned long long> BatchRunner::read_trace(std::uintptr_t stream) {
    std::vector<unsigned long long> out;
    if (desc_.trace == nullptr) return out;
    out.resize(2 * static_cast<std::size_t>(desc_.trace_entries));
    check(chx_memcpy_d2h(out.data(), desc_.trace, out.size() * sizeof(unsigned long long), stream), "engine",
          "trace read");
    check(chx_stream_synchronize(stream), "engine", "trace read");
    return out;
}

void BatchRunner::set_tile_rows(int rows) {
    if (rows < CHX_STAGE_ROWS || rows % CHX_STAGE_ROWS != 0 || rows > 64 * CHX_STAGE_ROWS)
        throw std::invalid_argument("chysx::engine: tile rows must be a multiple of CHX_STAGE_ROWS, at most 64 stages");
    if (rows == tile_rows_) return;
    tile_rows_ = rows;
    if (!bound()) return;
    // Queued launches may still read the old list.
    check(chx_device_synchronize(), "engine", "tile list rebuild");
    if (desc_.inst == nullptr) {
        std::vector<chx_instance> one{desc_.single};
        build_tiles(one);
        desc_.single = one[0];
        desc_.tiles = nullptr;  // implied by tile_rows
    } else {
        build_tiles(table_host_);
        table_.host() = table_host_;
        table_.copy_to_device();
        desc_.inst = table_.gpu_data();
    }
    size_outputs(desc_.n_instances, desc_.n_tiles);
    check(chx_device_synchronize(), "engine", "tile list rebuild");
}

void BatchRunner::build_tiles(std::vector<chx_instance>& ds, std::uintptr_t stream) {
    auto& ti = tiles_.host();
    ti.clear();
    for (std::size_t i = 0; i < ds.size(); ++i) {
        chx_instance& I = ds[i];
        I.tile_begin = static_cast<int>(ti.size() / 4);
        for (int r0 = 0; r0 < I.n_rows; r0 += tile_rows_)
            ti.insert(ti.end(), {static_cast<int>(i), r0, std::min(I.n_rows, r0 + tile_rows_), 0});
        I.tile_end = static_cast<int>(ti.size() / 4);
    }
    if (ti.size() / 4 > static_cast<std::size_t>(std::numeric_limits<int>::max()))
        throw std::runtime_error("chysx::engine: too many row tiles");
    tiles_.copy_to_device(stream);
    desc_.n_tiles = static_cast<int>(ti.size() / 4);
    desc_.tiles = tiles_.gpu_data();
    desc_.tile_rows = tile_rows_;
}

// 1 when every instance with rows carries 16-bit slice entries.
static int entries16_all(const std::vector<chx_instance>& ds) {
    for (const chx_instance& d : ds)
        if (d.n_rows > 0 && d.slice_entries16 == nullptr) return 0;
    return 1;
}

void BatchRunner::bind_single(const chx_instance& d, std::uintptr_t stream) {
    std::vector<chx_instance> one{d};
    build_tiles(one, stream);
    desc_.tiles = nullptr;  // implied by tile_rows: no tile record loads
    desc_.n_instances = 1;
    desc_.n_elem_items = 0;
    desc_.n_unit_items = 0;
    desc_.lanes_per_row = lanes_;
    desc_.grid_blocks = grid_;
    desc_.bf16_blocks = bf16_ ? 1 : 0;
    desc_.entries16 = entries16_all(one);
    desc_.single = one[0];
    desc_.inst = nullptr;
    desc_.elem_items = nullptr;
    desc_.unit_items = nullptr;
    size_outputs(1, desc_.n_tiles, stream);
    sync_stream(stream, "bind");
}

void BatchRunner::bind_many(const std::vector<chx_instance>& ds, std::uintptr_t stream) {
    if (ds.empty()) throw std::invalid_argument("chysx::engine: empty batch");
    if (ds.size() == 1) {
        // one instance runs as a single-instance batch (descriptor in the
        // kernel arguments, implied work lists): the same kernels and bits
        bind_single(ds[0], stream);
        table_host_ = {desc_.single};
        return;
    }
    table_host_ = ds;
    build_tiles(table_host_, stream);
    auto& ei = elem_items_.host();
    auto& si = unit_items_.host();
    ei.clear();
    si.clear();
    for (std::size_t i = 0; i < table_host_.size(); ++i) {
        const chx_instance& I = table_host_[i];
        const int inst = static_cast<int>(i);
        const int counts[4] = {I.n_pins, I.n_tris, I.n_dihs, I.n_stvks};
        for (int kind = 0; kind < 4; ++kind)
            for (int b = 0; b < counts[kind]; b += CHX_ITEM_SIZE)
                ei.insert(ei.end(), {inst, kind, b, std::min(counts[kind], b + CHX_ITEM_SIZE)});
        if (I.n_rows > 0)
            for (int u = 0; u < I.n_units; ++u) si.insert(si.end(), {inst, u, 0, 0});
    }
    elem_items_.copy_to_device(stream);
    unit_items_.copy_to_device(stream);
    table_.host() = table_host_;
    table_.copy_to_device(stream);
    desc_.n_instances = static_cast<int>(ds.size());
    desc_.n_elem_items = static_cast<int>(ei.size() / 4);
    desc_.n_unit_items = static_cast<int>(si.size() / 4);
    desc_.lanes_per_row = lanes_;
    desc_.grid_blocks = grid_;
    desc_.bf16_blocks = bf16_ ? 1 : 0;
    desc_.entries16 = entries16_all(table_host_);
    desc_.single = chx_instance{};
    desc_.inst = table_.gpu_data();
    desc_.elem_items = elem_items_.gpu_data();
    desc_.unit_items = unit_items_.gpu_data();
    size_outputs(desc_.n_instances, desc_.n_tiles, stream);
    sync_stream(stream, "bind");
}

void BatchRunner::update_many(const std::vector<chx_instance>& ds, std::uintptr_t stream) {
    if (ds.size() != table_host_.size()) throw std::invalid_argument("chysx::engine: instance count changed");
    if (desc_.inst == nullptr) {
        update_single(ds[0]);
        table_host_[0] = desc_.single;
        return;
    }
    bool changed = false;
    for (std::size_t i = 0; i < ds.size(); ++i) {
        chx_instance d = ds[i];
        d.tile_begin = table_host_[i].tile_begin;
        d.tile_end = table_host_[i].tile_end;
        if (std::memcmp(&d, &table_host_[i], sizeof(chx_instance)) != 0) {
            table_host_[i] = d;
            changed = true;
        }
    }
    desc_.entries16 = entries16_all(table_host_);
    if (changed) {
        // The previous table may still be read by queued launches.
        check(chx_stream_synchronize(stream), "engine", "instance table update");
        table_.host() = table_host_;
        table_.copy_to_device(stream);
        desc_.inst = table_.gpu_data();
    }
}

void BatchRunner::step(std::uintptr_t stream) {
    check_host_status(stream);
    check(chx_cloth_step(&desc_, stream), "cloth::ClothSimulator", "step");
}

void BatchRunner::step_graph(std::uintptr_t stream) {
    if (stream == 0) throw std::invalid_argument("chysx::cloth::ClothBatch: graph replay needs a non-null stream");
    check_host_status(stream);
    if (graph_ == nullptr || std::memcmp(&graph_desc_, &desc_, sizeof(chx_cloth_batch)) != 0) {
        // A replaced graph may still run on the stream: wait before dropping it.
        check(chx_stream_synchronize(stream), "cloth::ClothBatch", "graph capture");
        drop_graph();
        check(chx_cloth_step_graph(&desc_, stream, &graph_), "cloth::ClothBatch", "graph capture");
        std::memcpy(&graph_desc_, &desc_, sizeof(chx_cloth_batch));
        ++graph_captures_;
    }
    check(chx_graph_launch(graph_, stream), "cloth::ClothBatch", "graph replay");
}

void BatchRunner::drop_graph() noexcept {
    if (graph_ != nullptr) (void)chx_graph_destroy(graph_);
    graph_ = nullptr;
}
void BatchRunner::assemble(std::uintptr_t stream) {
    check(chx_cloth_assemble(&desc_, stream), "cloth::ClothSimulator", "assemble");
}
void BatchRunner::solve(bool finalize, std::uintptr_t stream) {
    check_host_status(stream);
    check(chx_cloth_solve(&desc_, finalize ? 1 : 0, stream), "solver::PCGSolver", "solve");
}
double BatchRunner::energy(int kind_mask, std::uintptr_t stream) {
    double e = 0.0;
    check(chx_batch_energy(&desc_, kind_mask, &e, stream), "cloth", "energy");
    return e;
}

void BatchRunner::read_stats(std::vector<int>& iters, std::vector<double>& rr, std::vector<double>& bb,
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

void BatchRunner::check_status(std::uintptr_t stream) {
    if (desc_.sync) check(chx_batch_check(&desc_, stream), "solver::PCGSolver", "solve");
}

}  // namespace engine
}  // namespace chysx
