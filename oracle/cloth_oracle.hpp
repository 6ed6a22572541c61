// CPU oracle for the implicit cloth step — TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU, the algorithm of ChysX's implicit-Euler
// cloth step so that the HIP implementation can be checked against it.  It
// is never linked into, called by, or shipped with the product library:
// only tests/, the smoke check and the CPU-baseline leg of bench.py load it
// (through oracle_capi.cpp / liboracle_cpu.so).
//
// Every function cites the reference file:line it follows (paths relative to
// the reference repository root).  The restatement is templated on the
// scalar type: `float` reproduces the reference's FP32 semantics, `double`
// is the full-FP64 accuracy oracle.
//
// Deliberate, documented differences from the reference:
//   * float atomics (ChysX/src/sparse/block_csr_atomic.cuh:33-45,
//     triangle_stretch_constraint.cu:206-216, ...) become a sequential
//     scatter in element order (pins, stretch, shear, bending), which is one
//     of the orders the reference's atomics may produce;
//   * PCG inner products accumulate in FP64 over fixed 4096-row chunks
//     (deterministic for any thread count); the reference accumulates in
//     FP32 with atomics (ChysX/src/solver/pcg_solver.cu:75-90);
//   * an optional convergence test ||r||_2 <= tol * ||b||_2 (FP64) is added;
//     tol <= 0 reproduces the reference's fixed-iteration loop
//     (ChysX/src/solver/pcg_solver.cu:176-214).
//
// Parity pinning of each part is listed in tests/test_oracle_*.py.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace chx_oracle {

// ---------------------------------------------------------------------------
// Small vector / matrix helpers (row-major 3x3 and 2x2, as
// ChysX/src/math/matrix.cuh:13-16,130-133 store them).
// ---------------------------------------------------------------------------
template <class T> struct V3 {
    T x{}, y{}, z{};
    V3() = default;
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
};
template <class T> inline V3<T> operator+(const V3<T>& a, const V3<T>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(const V3<T>& a, const V3<T>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(const V3<T>& a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline T dot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(const V3<T>& a, const V3<T>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T length(const V3<T>& a) { return std::sqrt(dot(a, a)); }

template <class T> struct M2 { T m[4]{}; };  // row-major [r*2+c]
template <class T> struct M3 { T m[9]{}; };  // row-major [r*3+c]

template <class T> inline M3<T> outer(const V3<T>& a, const V3<T>& b) {
    M3<T> r;
    r.m[0] = a.x * b.x; r.m[1] = a.x * b.y; r.m[2] = a.x * b.z;
    r.m[3] = a.y * b.x; r.m[4] = a.y * b.y; r.m[5] = a.y * b.z;
    r.m[6] = a.z * b.x; r.m[7] = a.z * b.y; r.m[8] = a.z * b.z;
    return r;
}
template <class T> inline M3<T> identity3() { M3<T> r; r.m[0] = r.m[4] = r.m[8] = T(1); return r; }
template <class T> inline M3<T> transpose(const M3<T>& a) {
    M3<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i * 3 + j] = a.m[j * 3 + i];
    return r;
}
template <class T> inline M3<T> operator+(M3<T> a, const M3<T>& b) { for (int i = 0; i < 9; ++i) a.m[i] += b.m[i]; return a; }
template <class T> inline M3<T> operator-(M3<T> a, const M3<T>& b) { for (int i = 0; i < 9; ++i) a.m[i] -= b.m[i]; return a; }
template <class T> inline M3<T> operator*(M3<T> a, T s) { for (int i = 0; i < 9; ++i) a.m[i] *= s; return a; }
template <class T> inline V3<T> operator*(const M3<T>& a, const V3<T>& v) {
    return {a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z,
            a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
// Adjugate / determinant inverse, no singular guard:
// ChysX/src/math/matrix.cuh:226-241.
template <class T> inline M3<T> inverse(const M3<T>& a) {
    const T* m = a.m;
    M3<T> r;
    r.m[0] = m[4] * m[8] - m[5] * m[7];
    r.m[1] = -m[1] * m[8] + m[2] * m[7];
    r.m[2] = m[1] * m[5] - m[2] * m[4];
    r.m[3] = -m[3] * m[8] + m[5] * m[6];
    r.m[4] = m[0] * m[8] - m[2] * m[6];
    r.m[5] = -m[0] * m[5] + m[2] * m[3];
    r.m[6] = m[3] * m[7] - m[4] * m[6];
    r.m[7] = -m[0] * m[7] + m[1] * m[6];
    r.m[8] = m[0] * m[4] - m[1] * m[3];
    const T det = m[0] * r.m[0] + m[1] * r.m[3] + m[2] * r.m[6];
    return r * (T(1) / det);
}

// ---------------------------------------------------------------------------
// Topology (integer work; bit-exact parity targets)
// ---------------------------------------------------------------------------

// Unique undirected edges, sorted: ChysX/src/cloth/cloth_simulator.cu:330-349.
inline std::vector<std::array<int, 2>> build_edges(const int* tris, int n_tris) {
    std::vector<std::pair<int, int>> tmp;
    tmp.reserve(static_cast<size_t>(n_tris) * 3);
    for (int t = 0; t < n_tris; ++t) {
        const int a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        tmp.emplace_back(std::min(a, b), std::max(a, b));
        tmp.emplace_back(std::min(b, c), std::max(b, c));
        tmp.emplace_back(std::min(c, a), std::max(c, a));
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    std::vector<std::array<int, 2>> out(tmp.size());
    for (size_t i = 0; i < tmp.size(); ++i) out[i] = {tmp[i].first, tmp[i].second};
    return out;
}

// Dihedral detection over directed edges, emitted in the iteration order of
// a std::unordered_map<int64,int> reserved to 3T and filled in triangle
// order: ChysX/src/cloth/cloth_simulator.cu:481-507.  With libstdc++ this
// reproduces the reference's order on Linux exactly.
inline std::vector<std::array<int, 4>> detect_dihedrals(const int* tris, int n_tris) {
    auto pack = [](int from, int to) -> std::int64_t {
        return (static_cast<std::int64_t>(from) << 32) |
               static_cast<std::int64_t>(static_cast<std::uint32_t>(to));
    };
    std::unordered_map<std::int64_t, int> third;
    third.reserve(static_cast<size_t>(3 * n_tris));
    for (int t = 0; t < n_tris; ++t) {
        const int v0 = tris[3 * t], v1 = tris[3 * t + 1], v2 = tris[3 * t + 2];
        third[pack(v0, v1)] = v2;
        third[pack(v1, v2)] = v0;
        third[pack(v2, v0)] = v1;
    }
    std::vector<std::array<int, 4>> out;
    out.reserve(static_cast<size_t>(2 * n_tris));
    for (const auto& kv : third) {
        const int a = static_cast<int>(kv.first >> 32);
        const int b = static_cast<int>(kv.first & 0xffffffff);
        if (a >= b) continue;
        auto rev = third.find(pack(b, a));
        if (rev == third.end()) continue;
        out.push_back({a, b, kv.second, rev->second});
    }
    return out;
}

// Off-diagonal block-CSR pattern: drop i==j, sort, unique.
// ChysX/src/sparse/block_csr.cu:91-147.
struct Pattern {
    int n = 0;
    std::vector<int> row_offsets;  // n+1
    std::vector<int> col_indices;  // nnz
};

inline Pattern build_pattern(int n, const std::vector<int>& rows, const std::vector<int>& cols) {
    Pattern p;
    p.n = n;
    std::vector<std::pair<int, int>> tmp;
    tmp.reserve(rows.size());
    for (size_t k = 0; k < rows.size(); ++k) {
        const int r = rows[k], c = cols[k];
        if (r < 0 || r >= n || c < 0 || c >= n)
            throw std::out_of_range("oracle build_pattern: pair out of range");
        if (r == c) continue;
        tmp.emplace_back(r, c);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    p.row_offsets.assign(static_cast<size_t>(n) + 1, 0);
    p.col_indices.resize(tmp.size());
    for (size_t i = 0; i < tmp.size(); ++i) {
        p.row_offsets[tmp[i].first + 1] += 1;
        p.col_indices[i] = tmp[i].second;
    }
    for (int i = 0; i < n; ++i) p.row_offsets[i + 1] += p.row_offsets[i];
    return p;
}

// Slot encoding: diag -> -r-1, off-diag -> index in values (binary search).
// ChysX/src/sparse/block_csr.cu:168-204.
inline int resolve_slot(const Pattern& p, int r, int c) {
    if (r < 0 || r >= p.n || c < 0 || c >= p.n)
        throw std::out_of_range("oracle resolve_slots: pair out of range");
    if (r == c) return -r - 1;
    const auto first = p.col_indices.begin() + p.row_offsets[r];
    const auto last = p.col_indices.begin() + p.row_offsets[r + 1];
    const auto it = std::lower_bound(first, last, c);
    if (it == last || *it != c)
        throw std::out_of_range("oracle resolve_slots: off-diagonal (row, col) not in topology");
    return static_cast<int>(it - p.col_indices.begin());
}

// Per-element slot table slot[c*K*K + a*K + b]: ConstraintN<K>::
// bind_hessian_layout, ChysX/src/constraint/core/constraint_n.h:115-153.
template <int K>
inline std::vector<int> slot_table(const Pattern& p, const std::vector<std::array<int, K>>& elems) {
    std::vector<int> s(elems.size() * K * K);
    for (size_t c = 0; c < elems.size(); ++c)
        for (int a = 0; a < K; ++a)
            for (int b = 0; b < K; ++b)
                s[c * K * K + a * K + b] = resolve_slot(p, elems[c][a], elems[c][b]);
    return s;
}

// ---------------------------------------------------------------------------
// Element math
// ---------------------------------------------------------------------------

// Rest shape of one triangle: ChysX/src/constraint/cloth/
// triangle_stretch_constraint.cu:50-110 (kMinNorm = 1e-6 at :39).
template <class T>
inline void triangle_rest_shape(const V3<T>& a, const V3<T>& b, const V3<T>& c,
                                T cos_t, T sin_t, M2<T>& Dm_inv, T& area) {
    const T kMin = T(1.0e-6f);
    const V3<T> ab = b - a, ac = c - a;
    const T lab = length(ab), lac = length(ac);
    const V3<T> ab_hat = (lab > kMin) ? ab * (T(1) / lab) : V3<T>(T(1), T(0), T(0));
    const T u_c = dot(ab_hat, ac);
    const T v_c2 = lac * lac - u_c * u_c;
    const T v_c = (v_c2 > T(0)) ? std::sqrt(v_c2) : T(0);
    area = T(0.5) * lab * v_c;
    const T det = lab * v_c;
    M2<T> raw;
    if (det > kMin) {
        // Mat2 inverse: ChysX/src/math/matrix.cuh:112-117 with Dm=[[lab,u_c],[0,v_c]].
        const T d = lab * v_c - u_c * T(0);
        const T inv_det = T(1) / d;
        raw.m[0] = v_c * inv_det;
        raw.m[1] = -u_c * inv_det;
        raw.m[2] = -T(0) * inv_det;
        raw.m[3] = lab * inv_det;
    } else {
        raw.m[0] = T(1); raw.m[1] = T(0); raw.m[2] = T(0); raw.m[3] = T(1);
    }
    // Dm_inv = raw * R^{-1}, R^{-1} = [[c, s], [-s, c]] (:100-109).
    const T R[4] = {cos_t, sin_t, -sin_t, cos_t};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            Dm_inv.m[i * 2 + j] = raw.m[i * 2 + 0] * R[0 * 2 + j] + raw.m[i * 2 + 1] * R[1 * 2 + j];
}

// BW98 stretch energy (unclamped norms): triangle_stretch_constraint.cu:116-147.
template <class T>
inline T triangle_energy(const V3<T>& x0, const V3<T>& x1, const V3<T>& x2, const M2<T>& D, T area, T k) {
    const V3<T> a = x1 - x0, b = x2 - x0;
    const V3<T> wu = a * D.m[0] + b * D.m[2];
    const V3<T> wv = a * D.m[1] + b * D.m[3];
    const T cu = length(wu) - T(1), cv = length(wv) - T(1);
    return T(0.5) * (area * k) * (cu * cu + cv * cv);
}

// Gradient of one triangle: triangle_stretch_constraint.cu:153-217.
template <class T>
inline void triangle_gradient(const V3<T>& x0, const V3<T>& x1, const V3<T>& x2, const M2<T>& D,
                              T area, T k, V3<T> g[3]) {
    const T kMin = T(1.0e-6f);
    const V3<T> a = x1 - x0, b = x2 - x0;
    const V3<T> wu = a * D.m[0] + b * D.m[2];
    const V3<T> wv = a * D.m[1] + b * D.m[3];
    T lu = length(wu), lv = length(wv);
    if (lu < kMin) lu = kMin;
    if (lv < kMin) lv = kMin;
    const T cu = lu - T(1), cv = lv - T(1);
    const V3<T> wu_hat = wu * (T(1) / lu), wv_hat = wv * (T(1) / lv);
    const T dwu[3] = {-D.m[0] - D.m[2], D.m[0], D.m[2]};
    const T dwv[3] = {-D.m[1] - D.m[3], D.m[1], D.m[3]};
    const T k_eff = area * k;
    for (int i = 0; i < 3; ++i)
        g[i] = (wu_hat * (cu * dwu[i]) + wv_hat * (cv * dwv[i])) * k_eff;
}

// PSD-filtered Hessian blocks H[a*3+b]: triangle_stretch_constraint.cu:231-311.
template <class T>
inline void triangle_hessian(const V3<T>& x0, const V3<T>& x1, const V3<T>& x2, const M2<T>& D,
                             T area, T k, M3<T> H[9]) {
    const T kMin = T(1.0e-6f);
    const V3<T> a = x1 - x0, b = x2 - x0;
    const V3<T> wu = a * D.m[0] + b * D.m[2];
    const V3<T> wv = a * D.m[1] + b * D.m[3];
    T lu = length(wu), lv = length(wv);
    if (lu < kMin) lu = kMin;
    if (lv < kMin) lv = kMin;
    const T inv_lu = T(1) / lu, inv_lv = T(1) / lv;
    const T cu = lu - T(1), cv = lv - T(1);
    const V3<T> wu_hat = wu * inv_lu, wv_hat = wv * inv_lv;
    const T dwu[3] = {-D.m[0] - D.m[2], D.m[0], D.m[2]};
    const T dwv[3] = {-D.m[1] - D.m[3], D.m[1], D.m[3]};
    const T k_eff = area * k;
    const M3<T> I3 = identity3<T>();
    const M3<T> Pu = I3 - outer(wu_hat, wu_hat);
    const M3<T> Pv = I3 - outer(wv_hat, wv_hat);
    V3<T> dcu[3], dcv[3];
    for (int i = 0; i < 3; ++i) { dcu[i] = wu_hat * dwu[i]; dcv[i] = wv_hat * dwv[i]; }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            M3<T> h = (outer(dcu[i], dcu[j]) + outer(dcv[i], dcv[j])) * k_eff;
            if (cu > T(0)) h = h + Pu * (k_eff * cu * inv_lu * dwu[i] * dwu[j]);
            if (cv > T(0)) h = h + Pv * (k_eff * cv * inv_lv * dwv[i] * dwv[j]);
            H[i * 3 + j] = h;
        }
}

// ---------------------------------------------------------------------------
// StVK membrane.  No reference counterpart: the reference keeps
// ClothMaterial::lame_mu / lame_lambda as unused placeholders
// (ChysX/src/cloth/cloth_material.h:93-116), so this part is pinned by
// closed forms, finite differences and invariants only.
//   F = [wu wv] (3x2) from the stretch rest shape (theta = 0) of
//   triangle_rest_shape, G = (F^T F - I)/2,
//   E = muA ||G||_F^2 + lamA/2 (tr G)^2, muA = mu*area*h, lamA = lam*area*h.
//   dE/dF = F S, S = 2 muA G + lamA tr(G) I (2x2, symmetric).
//   Second variation: 2 muA ||dG||^2 + lamA (tr dG)^2 + S : (dF^T dF).  The
//   first two terms are PSD for muA > 0, lamA >= 0; the projected Hessian
//   clamps S to its non-negative eigen-part S+ (exact when S is PSD).
// ---------------------------------------------------------------------------
template <class T>
inline void stvk_deformation(const V3<T>& x0, const V3<T>& x1, const V3<T>& x2, const M2<T>& D, V3<T>& fu,
                             V3<T>& fv) {
    const V3<T> a = x1 - x0, b = x2 - x0;
    fu = a * D.m[0] + b * D.m[2];
    fv = a * D.m[1] + b * D.m[3];
}

// Stress S = {Suu, Svv, Suv} and Green strain {Guu, Gvv, Guv}.
template <class T>
inline void stvk_stress(const V3<T>& fu, const V3<T>& fv, T muA, T lamA, T S[3], T G[3]) {
    G[0] = T(0.5) * (dot(fu, fu) - T(1));
    G[1] = T(0.5) * (dot(fv, fv) - T(1));
    G[2] = T(0.5) * dot(fu, fv);
    const T tr = G[0] + G[1];
    S[0] = T(2) * muA * G[0] + lamA * tr;
    S[1] = T(2) * muA * G[1] + lamA * tr;
    S[2] = T(2) * muA * G[2];
}

template <class T>
inline T stvk_energy(const V3<T>& x0, const V3<T>& x1, const V3<T>& x2, const M2<T>& D, T muA, T lamA) {
    V3<T> fu, fv;
    stvk_deformation(x0, x1, x2, D, fu, fv);
    T S[3], G[3];
    stvk_stress(fu, fv, muA, lamA, S, G);
    const T tr = G[0] + G[1];
    return muA * (G[0] * G[0] + G[1] * G[1] + T(2) * G[2] * G[2]) + T(0.5) * lamA * tr * tr;
}

template <class T>
inline void stvk_gradient(const V3<T>& x0, const V3<T>& x1, const V3<T>& x2, const M2<T>& D, T muA, T lamA,
                          V3<T> g[3]) {
    V3<T> fu, fv;
    stvk_deformation(x0, x1, x2, D, fu, fv);
    T S[3], G[3];
    stvk_stress(fu, fv, muA, lamA, S, G);
    const V3<T> Pu = fu * S[0] + fv * S[2], Pv = fu * S[2] + fv * S[1];
    const T dwu[3] = {-D.m[0] - D.m[2], D.m[0], D.m[2]};
    const T dwv[3] = {-D.m[1] - D.m[3], D.m[1], D.m[3]};
    for (int i = 0; i < 3; ++i) g[i] = Pu * dwu[i] + Pv * dwv[i];
}

// Non-negative eigen-part of the symmetric 2x2 {a, c; c, b}.
template <class T>
inline void psd2(T a, T b, T c, T out[3]) {
    const T m = T(0.5) * (a + b), d = T(0.5) * (a - b);
    const T r = std::sqrt(d * d + c * c);
    const T l1 = m + r, l2 = m - r;  // l1 >= l2
    if (l2 >= T(0)) { out[0] = a; out[1] = b; out[2] = c; return; }
    if (l1 <= T(0)) { out[0] = out[1] = out[2] = T(0); return; }
    // Only l1 > 0: l1 v v^T, v the unit eigenvector of l1.
    T vx, vy;
    if (d >= T(0)) { vx = d + r; vy = c; } else { vx = c; vy = r - d; }
    const T nn = vx * vx + vy * vy;
    if (nn <= T(0)) { out[0] = l1; out[1] = T(0); out[2] = T(0); return; }
    out[0] = l1 * vx * vx / nn;
    out[1] = l1 * vy * vy / nn;
    out[2] = l1 * vx * vy / nn;
}

// Hessian blocks H[a*3+b] (projected unless `exact`).
template <class T>
inline void stvk_hessian(const V3<T>& x0, const V3<T>& x1, const V3<T>& x2, const M2<T>& D, T muA, T lamA,
                         M3<T> H[9], bool exact = false) {
    V3<T> fu, fv;
    stvk_deformation(x0, x1, x2, D, fu, fv);
    T S[3], G[3];
    stvk_stress(fu, fv, muA, lamA, S, G);
    T Sp[3] = {S[0], S[1], S[2]};
    if (!exact) psd2(S[0], S[1], S[2], Sp);
    const M3<T> I3 = identity3<T>();
    const M3<T> Huu = outer(fu, fu) * (T(2) * muA + lamA) + outer(fv, fv) * muA + I3 * Sp[0];
    const M3<T> Hvv = outer(fv, fv) * (T(2) * muA + lamA) + outer(fu, fu) * muA + I3 * Sp[1];
    const M3<T> Huv = outer(fv, fu) * muA + outer(fu, fv) * lamA + I3 * Sp[2];
    const M3<T> Hvu = transpose(Huv);
    const T dwu[3] = {-D.m[0] - D.m[2], D.m[0], D.m[2]};
    const T dwv[3] = {-D.m[1] - D.m[3], D.m[1], D.m[3]};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            H[i * 3 + j] = Huu * (dwu[i] * dwu[j]) + Huv * (dwu[i] * dwv[j]) + Hvu * (dwv[i] * dwu[j]) +
                           Hvv * (dwv[i] * dwv[j]);
}

// Dihedral frame (Bridson shape vectors): ChysX/src/constraint/cloth/
// bending_constraint.cu:48-126 (kMinNorm = 1e-12 at :34).
template <class T> struct DihedralFrame {
    T theta{};
    V3<T> a[4];
    bool valid = false;
};

template <class T>
inline DihedralFrame<T> dihedral_frame(const V3<T>& v0, const V3<T>& v1, const V3<T>& v2, const V3<T>& v3) {
    const T kMin = T(1.0e-12f);
    DihedralFrame<T> F;
    const V3<T> e10 = v1 - v0, e20 = v2 - v0, e30 = v3 - v0;
    const T l = length(e10);
    if (l < kMin) return F;
    const T inv_l = T(1) / l;
    const V3<T> e10_hat = e10 * inv_l;
    const V3<T> n1_raw = cross(e20, e10), n2_raw = cross(e10, e30);
    const T n1_len = length(n1_raw), n2_len = length(n2_raw);
    if (n1_len < kMin || n2_len < kMin) return F;
    const V3<T> n1 = n1_raw * (T(1) / n1_len), n2 = n2_raw * (T(1) / n2_len);
    T s = dot(cross(n1, n2), e10_hat);
    if (s > T(1)) s = T(1);
    if (s < T(-1)) s = T(-1);
    const T c = dot(n1, n2);
    F.theta = std::atan2(s, c);
    const T inv_l2 = inv_l * inv_l;
    const T om1 = dot(e10, e20) * inv_l2, om2 = dot(e10, e30) * inv_l2;
    const V3<T> p1 = e20 - e10 * om1, p2 = e30 - e10 * om2;
    const T h1 = length(p1), h2 = length(p2);
    if (h1 < kMin || h2 < kMin) return F;
    const T ih1 = T(1) / h1, ih2 = T(1) / h2;
    const T t1[4] = {(om1 - T(1)) * ih1, -om1 * ih1, T(1) * ih1, T(0)};
    const T t2[4] = {(om2 - T(1)) * ih2, -om2 * ih2, T(0), T(1) * ih2};
    for (int i = 0; i < 4; ++i) F.a[i] = n1 * t1[i] + n2 * t2[i];
    F.valid = true;
    return F;
}

// ---------------------------------------------------------------------------
// The simulator: ChysX/src/cloth/cloth_simulator.{h,cu}, PCG branch.
// ---------------------------------------------------------------------------
template <class T> struct TriSet {
    std::vector<std::array<int, 3>> idx;
    std::vector<M2<T>> Dm_inv;
    std::vector<T> area;
    T k = T(0);
    std::vector<int> slots;
};

template <class T> struct StvkSet {
    std::vector<std::array<int, 3>> idx;
    std::vector<M2<T>> Dm_inv;
    std::vector<T> muA, lamA;
    std::vector<int> slots;
};

// ---------------------------------------------------------------------------
// Static plane / box contact: penalty E = k/2 max(0, h - d)^2 on the deepest
// primitive per particle, Gauss-Newton diagonal block, optional IPC-style
// lagged Coulomb friction (ChysX/src/collision/static_contact.h:13-60,
// static_contact.cu:41-366, friction.cuh:23-34).
// ---------------------------------------------------------------------------
template <class T> struct PlaneShape { V3<T> n; T d = T(0); };  // dot(n, x) + d = 0
template <class T> struct BoxShape { V3<T> c, h, ex, ey, ez; };   // centre, half extents, frame columns

// static_contact.cu:41-48
template <class T> inline T plane_signed_dist(const V3<T>& p, const PlaneShape<T>& s, V3<T>& n) {
    n = s.n;
    return dot(s.n, p) + s.d;
}

// static_contact.cu:50-96: oriented-box signed distance (positive outside,
// the face distance inside) and its outward normal.
template <class T> inline T box_signed_dist(const V3<T>& p, const BoxShape<T>& b, V3<T>& n) {
    const V3<T> rel = p - b.c;
    const T qx = dot(rel, b.ex), qy = dot(rel, b.ey), qz = dot(rel, b.ez);
    const T dx = std::fabs(qx) - b.h.x, dy = std::fabs(qy) - b.h.y, dz = std::fabs(qz) - b.h.z;
    if (dx <= T(0) && dy <= T(0) && dz <= T(0)) {
        if (dx >= dy && dx >= dz) {
            n = (qx >= T(0)) ? b.ex : b.ex * T(-1);
            return dx;
        }
        if (dy >= dz) {
            n = (qy >= T(0)) ? b.ey : b.ey * T(-1);
            return dy;
        }
        n = (qz >= T(0)) ? b.ez : b.ez * T(-1);
        return dz;
    }
    const T cx = dx > T(0) ? dx : T(0), cy = dy > T(0) ? dy : T(0), cz = dz > T(0) ? dz : T(0);
    const T dist = std::sqrt(cx * cx + cy * cy + cz * cz);
    const T inv = T(1) / std::max(dist, T(1.0e-30f));
    const T sx = qx > T(0) ? cx : -cx, sy = qy > T(0) ? cy : -cy, sz = qz > T(0) ? cz : -cz;
    n = b.ex * (sx * inv) + b.ey * (sy * inv) + b.ez * (sz * inv);
    return dist;
}

// friction.cuh:23-34: 1/u past eps, linear ramp 2/eps -> 1/eps below.
template <class T> inline T ipc_f1_over_x(T u, T eps) {
    if (u > eps) return T(1) / u;
    return (-u / eps + T(2)) / eps;
}

template <class T> struct StaticContact {
    std::vector<PlaneShape<T>> planes;
    std::vector<BoxShape<T>> boxes;
    T thickness = T(0), stiffness = T(0), friction = T(0), eps = T(1.0e-4f);
    bool active() const { return stiffness > T(0) && (!planes.empty() || !boxes.empty()); }
    // static_contact.cu:106-160: deepest penetration (planes, then boxes;
    // a later primitive wins only when strictly deeper) and the lagged
    // tangential slip u_t = (v dt) - n (n . v dt).
    void detect(const V3<T>& x, const V3<T>& v, T dt, V3<T>& n, T& depth, V3<T>& ut, T& unorm) const {
        depth = T(0);
        n = V3<T>();
        for (const auto& pl : planes) {
            V3<T> nn;
            const T d = thickness - plane_signed_dist(x, pl, nn);
            if (d > depth) { depth = d; n = nn; }
        }
        for (const auto& b : boxes) {
            V3<T> nn;
            const T d = thickness - box_signed_dist(x, b, nn);
            if (d > depth) { depth = d; n = nn; }
        }
        ut = V3<T>();
        unorm = T(0);
        if (depth > T(0) && dt > T(0)) {
            const V3<T> dxv = v * dt;
            ut = dxv - n * dot(n, dxv);
            unorm = std::sqrt(ut.x * ut.x + ut.y * ut.y + ut.z * ut.z);
        }
    }
};

// ---------------------------------------------------------------------------
// Contact SpMV sidecar (ChysX/src/collision/contact_spmv.h:12-52,
// contact_spmv.cu:46-268): contact c couples particles id[0..3] with
// weights w[0..3], normal n, depth; H_c = k (n n^T) + a_c (I - n n^T) with
// a_c = mu k depth f1(|u_t|) when friction is on.  The full operator is
// C[i][j] = sum_c w_i w_j H_c over the contact's slots; bake adds the
// diagonal terms (i == j slots) to the diagonal blocks, apply adds the
// off-diagonal ones (i != j slots) to y.  FP64 sums in contact order.
// ---------------------------------------------------------------------------
struct ContactOp {
    const int* pairs = nullptr;     // [m][4]
    const double* weights = nullptr;  // [m][8]: w0..w3, nx, ny, nz, depth
    const double* slips = nullptr;  // [m][4] or null
    int count = 0;
    double k = 0.0, mu = 0.0, eps = 1.0e-4;
    void block(int c, double H[9]) const {
        const double* w = weights + 8 * c;
        const double nx = w[4], ny = w[5], nz = w[6];
        double a = 0.0;
        if (mu > 0.0 && slips) a = mu * (k * w[7]) * ipc_f1_over_x(slips[4 * c + 3], eps);
        const double nn[3] = {nx, ny, nz};
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) H[3 * r + q] = (k - a) * nn[r] * nn[q] + (r == q ? a : 0.0);
    }
    // diag[i] += alpha sum w_i^2 H_c
    void bake(double* diag, double alpha) const {
        for (int c = 0; c < count; ++c) {
            double H[9];
            block(c, H);
            for (int s = 0; s < 4; ++s) {
                const double wi = weights[8 * c + s];
                if (wi == 0.0) continue;
                for (int e = 0; e < 9; ++e) diag[9 * pairs[4 * c + s] + e] += alpha * wi * wi * H[e];
            }
        }
    }
    // y[i] += alpha sum_{s != t} w_s w_t H_c x[id_t]
    void apply(const double* x, double* y, double alpha) const {
        for (int c = 0; c < count; ++c) {
            double H[9];
            block(c, H);
            for (int s = 0; s < 4; ++s) {
                const double ws = weights[8 * c + s];
                if (ws == 0.0) continue;
                double t[3] = {0.0, 0.0, 0.0};
                for (int u = 0; u < 4; ++u) {
                    if (u == s) continue;
                    const double wu = weights[8 * c + u];
                    for (int r = 0; r < 3; ++r) t[r] += wu * x[3 * pairs[4 * c + u] + r];
                }
                const int i = pairs[4 * c + s];
                for (int r = 0; r < 3; ++r)
                    y[3 * i + r] += alpha * ws * (H[3 * r] * t[0] + H[3 * r + 1] * t[1] + H[3 * r + 2] * t[2]);
            }
        }
    }
};

template <class T> struct Sim {
    int n = 0;
    std::vector<V3<T>> pos, vel;
    std::vector<T> inv_mass;
    bool has_inv_mass = false;
    T damping = T(0), gx = T(0), gy = T(0), gz = T(-9.81f);

    std::vector<std::array<int, 3>> tris;
    std::vector<std::array<int, 2>> edges;
    TriSet<T> stretch, shear;
    StvkSet<T> stvk;
    std::vector<std::array<int, 4>> dih;
    std::vector<T> rest_angle;
    T k_bend = T(0);
    std::vector<int> dih_slots;
    std::vector<int> pin_idx;
    std::vector<V3<T>> pin_target;
    T k_pin = T(1.0e6f);
    std::vector<int> pin_slots;
    StaticContact<T> contact;

    bool topology_dirty = true;
    int topo_n = 0;
    Pattern pat;
    std::vector<M3<T>> diag, values;
    std::vector<V3<T>> rhs, dx;
    int max_iter = 50;
    double tol = 0.0;
    int last_iters = 0;
    double last_rr = 0.0, last_bb = 0.0;
    double last_rho = 0.0;  // <r, z> of the final iterate (pcg_solver.cu:319-323)

    // ---- setup (ChysX/src/cloth/cloth_simulator.cu:248-514) -------------
    void set_state(int count, const double* p, const double* v, const double* w) {
        n = count;
        pos.resize(n); vel.resize(n);
        for (int i = 0; i < n; ++i) {
            pos[i] = V3<T>(T(p[3 * i]), T(p[3 * i + 1]), T(p[3 * i + 2]));
            vel[i] = V3<T>(T(v[3 * i]), T(v[3 * i + 1]), T(v[3 * i + 2]));
        }
        has_inv_mass = (w != nullptr);
        inv_mass.assign(n, T(0));
        if (w) for (int i = 0; i < n; ++i) inv_mass[i] = T(w[i]);
    }

    void set_mesh(const int* t, int count) {
        if (count < 0) throw std::invalid_argument("set_mesh: negative triangle count");
        tris.resize(count);
        for (int i = 0; i < count; ++i) tris[i] = {t[3 * i], t[3 * i + 1], t[3 * i + 2]};
        edges = build_edges(t, count);
        topology_dirty = true;
    }

    void check_indices(const int* v, int count) const {
        for (int i = 0; i < count; ++i)
            if (v[i] < 0 || v[i] >= n) throw std::out_of_range("oracle: element vertex index out of range");
    }

    void build_tri_set(TriSet<T>& s, T k, float rot) {
        for (const auto& t : tris) check_indices(t.data(), 3);
        s.k = k;
        s.idx = tris;
        s.Dm_inv.assign(tris.size(), M2<T>{});
        s.area.assign(tris.size(), T(0));
        // cos/sin evaluated on the host in float (triangle_stretch_constraint.cu:341-342).
        const T c = T(std::cos(rot)), sn = T(std::sin(rot));
        for (size_t i = 0; i < tris.size(); ++i)
            triangle_rest_shape(pos[tris[i][0]], pos[tris[i][1]], pos[tris[i][2]], c, sn, s.Dm_inv[i], s.area[i]);
        topology_dirty = true;
    }
    void build_stretch(T k) { build_tri_set(stretch, k, 0.0f); }
    // kQuarterPi: ChysX/src/cloth/cloth_simulator.cu:428.
    void build_shear(T k) { build_tri_set(shear, k, 0.7853981633974483f); }

    // StVK membrane with Lame parameters (mu, lam) and thickness h, rest
    // shape from the current positions (theta = 0).
    void build_stvk(T mu, T lam, T h) {
        for (const auto& t : tris) check_indices(t.data(), 3);
        stvk.idx = tris;
        stvk.Dm_inv.assign(tris.size(), M2<T>{});
        stvk.muA.assign(tris.size(), T(0));
        stvk.lamA.assign(tris.size(), T(0));
        for (size_t i = 0; i < tris.size(); ++i) {
            T area = T(0);
            triangle_rest_shape(pos[tris[i][0]], pos[tris[i][1]], pos[tris[i][2]], T(1), T(0), stvk.Dm_inv[i], area);
            stvk.muA[i] = mu * area * h;
            stvk.lamA[i] = lam * area * h;
        }
        topology_dirty = true;
    }

    void build_bending(T k) {
        k_bend = k;
        std::vector<int> flat(tris.size() * 3);
        for (size_t i = 0; i < tris.size(); ++i)
            for (int j = 0; j < 3; ++j) flat[3 * i + j] = tris[i][j];
        dih = detect_dihedrals(flat.data(), static_cast<int>(tris.size()));
        for (const auto& d : dih) check_indices(d.data(), 4);
        rest_angle.assign(dih.size(), T(0));
        // bending_constraint.cu:132-148.
        for (size_t e = 0; e < dih.size(); ++e) {
            const auto F = dihedral_frame(pos[dih[e][0]], pos[dih[e][1]], pos[dih[e][2]], pos[dih[e][3]]);
            rest_angle[e] = F.valid ? F.theta : T(0);
        }
        topology_dirty = true;
    }

    void set_pins(const int* idx, const double* targets, int count, T k) {
        k_pin = k;
        pin_idx.assign(idx, idx + count);
        pin_target.resize(count);
        for (int i = 0; i < count; ++i)
            pin_target[i] = V3<T>(T(targets[3 * i]), T(targets[3 * i + 1]), T(targets[3 * i + 2]));
        topology_dirty = true;
    }

    // update_pin_targets (cloth_simulator.cu:294-313): same pin set, new
    // targets; the pattern and the warm start are untouched.
    void update_pin_targets(const double* targets, int count) {
        if (count != static_cast<int>(pin_idx.size()))
            throw std::invalid_argument("update_pin_targets: n_pins does not match the installed pin set");
        for (int i = 0; i < count; ++i)
            pin_target[i] = V3<T>(T(targets[3 * i]), T(targets[3 * i + 1]), T(targets[3 * i + 2]));
    }

    // Area-weighted lumped mass: cloth_simulator.cu:516-575, 151-190.
    void redistribute_mass(T rho) {
        for (const auto& t : tris) check_indices(t.data(), 3);
        std::vector<T> m(n, T(0));
        const T d3 = rho / T(3);
        for (const auto& t : tris) {
            const V3<T> e1 = pos[t[1]] - pos[t[0]], e2 = pos[t[2]] - pos[t[0]];
            const T area = T(0.5) * length(cross(e1, e2));
            const T dm = d3 * area;
            m[t[0]] += dm; m[t[1]] += dm; m[t[2]] += dm;
        }
        has_inv_mass = true;
        inv_mass.resize(n);
        for (int i = 0; i < n; ++i) inv_mass[i] = (m[i] > T(1.0e-12f)) ? T(1) / m[i] : T(0);
    }

    // ensure_hessian_topology: cloth_simulator.cu:698-793.
    void ensure_topology() {
        std::vector<int> rows, cols;
        auto tri_pairs = [&](const TriSet<T>& s) {
            for (const auto& t : s.idx)
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        if (a != b && t[a] != t[b]) { rows.push_back(t[a]); cols.push_back(t[b]); }
        };
        tri_pairs(stretch);
        tri_pairs(shear);
        for (const auto& t : stvk.idx)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    if (a != b && t[a] != t[b]) { rows.push_back(t[a]); cols.push_back(t[b]); }
        for (const auto& d : dih)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    if (a != b && d[a] != d[b]) { rows.push_back(d[a]); cols.push_back(d[b]); }
        pat = build_pattern(n, rows, cols);
        pin_slots.resize(pin_idx.size());
        for (size_t c = 0; c < pin_idx.size(); ++c) pin_slots[c] = resolve_slot(pat, pin_idx[c], pin_idx[c]);
        stretch.slots = slot_table<3>(pat, stretch.idx);
        shear.slots = slot_table<3>(pat, shear.idx);
        stvk.slots = slot_table<3>(pat, stvk.idx);
        dih_slots = slot_table<4>(pat, dih);
        topo_n = n;
        topology_dirty = false;
    }

    // ---- assembly (sequential scatter in reference launch order) --------
    void add_block(int slot, const M3<T>& h) {
        M3<T>& dst = (slot < 0) ? diag[-slot - 1] : values[slot];
        for (int q = 0; q < 9; ++q) dst.m[q] += h.m[q];
    }

    // Gradient of E at x (pins, stretch, shear, bending):
    // cloth_simulator.cu:914-924; pin_constraint.cu:61-75.
    void gradient(const std::vector<V3<T>>& x, std::vector<V3<T>>& g) const {
        g.assign(n, V3<T>());
        for (size_t c = 0; c < pin_idx.size(); ++c) {
            const int i = pin_idx[c];
            g[i] = g[i] + (x[i] - pin_target[c]) * k_pin;
        }
        auto tri_grad = [&](const TriSet<T>& s) {
            for (size_t t = 0; t < s.idx.size(); ++t) {
                V3<T> ge[3];
                triangle_gradient(x[s.idx[t][0]], x[s.idx[t][1]], x[s.idx[t][2]], s.Dm_inv[t], s.area[t], s.k, ge);
                for (int a = 0; a < 3; ++a) g[s.idx[t][a]] = g[s.idx[t][a]] + ge[a];
            }
        };
        tri_grad(stretch);
        tri_grad(shear);
        for (size_t t = 0; t < stvk.idx.size(); ++t) {
            V3<T> ge[3];
            stvk_gradient(x[stvk.idx[t][0]], x[stvk.idx[t][1]], x[stvk.idx[t][2]], stvk.Dm_inv[t], stvk.muA[t],
                          stvk.lamA[t], ge);
            for (int a = 0; a < 3; ++a) g[stvk.idx[t][a]] = g[stvk.idx[t][a]] + ge[a];
        }
        // bending_constraint.cu:180-206.
        for (size_t e = 0; e < dih.size(); ++e) {
            const auto F = dihedral_frame(x[dih[e][0]], x[dih[e][1]], x[dih[e][2]], x[dih[e][3]]);
            if (!F.valid) continue;
            const T coeff = k_bend * (F.theta - rest_angle[e]);
            for (int i = 0; i < 4; ++i) g[dih[e][i]] = g[dih[e][i]] + F.a[i] * coeff;
        }
    }

    double energy(const std::vector<V3<T>>& x) const {
        double E = 0.0;
        for (size_t c = 0; c < pin_idx.size(); ++c) {
            const V3<T> d = x[pin_idx[c]] - pin_target[c];
            E += double(T(0.5) * k_pin * dot(d, d));
        }
        for (const TriSet<T>* s : {&stretch, &shear})
            for (size_t t = 0; t < s->idx.size(); ++t)
                E += double(triangle_energy(x[s->idx[t][0]], x[s->idx[t][1]], x[s->idx[t][2]], s->Dm_inv[t], s->area[t], s->k));
        for (size_t t = 0; t < stvk.idx.size(); ++t)
            E += double(stvk_energy(x[stvk.idx[t][0]], x[stvk.idx[t][1]], x[stvk.idx[t][2]], stvk.Dm_inv[t],
                                    stvk.muA[t], stvk.lamA[t]));
        for (size_t e = 0; e < dih.size(); ++e) {
            const auto F = dihedral_frame(x[dih[e][0]], x[dih[e][1]], x[dih[e][2]], x[dih[e][3]]);
            if (!F.valid) continue;
            const T d = F.theta - rest_angle[e];
            E += double(T(0.5) * k_bend * d * d);
        }
        return E;
    }

    // H_E(x) into (diag, values): cloth_simulator.cu:1129-1146.
    void hessian(const std::vector<V3<T>>& x) {
        diag.assign(n, M3<T>{});
        values.assign(pat.col_indices.size(), M3<T>{});
        for (size_t c = 0; c < pin_idx.size(); ++c) {  // pin_constraint.cu:79-92
            M3<T> h;
            h.m[0] = h.m[4] = h.m[8] = k_pin;
            add_block(pin_slots[c], h);
        }
        for (const TriSet<T>* s : {&stretch, &shear})
            for (size_t t = 0; t < s->idx.size(); ++t) {
                M3<T> H[9];
                triangle_hessian(x[s->idx[t][0]], x[s->idx[t][1]], x[s->idx[t][2]], s->Dm_inv[t], s->area[t], s->k, H);
                for (int q = 0; q < 9; ++q) add_block(s->slots[9 * t + q], H[q]);
            }
        for (size_t t = 0; t < stvk.idx.size(); ++t) {
            M3<T> H[9];
            stvk_hessian(x[stvk.idx[t][0]], x[stvk.idx[t][1]], x[stvk.idx[t][2]], stvk.Dm_inv[t], stvk.muA[t],
                         stvk.lamA[t], H);
            for (int q = 0; q < 9; ++q) add_block(stvk.slots[9 * t + q], H[q]);
        }
        for (size_t e = 0; e < dih.size(); ++e) {  // bending_constraint.cu:217-245
            const auto F = dihedral_frame(x[dih[e][0]], x[dih[e][1]], x[dih[e][2]], x[dih[e][3]]);
            if (!F.valid) continue;
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) add_block(dih_slots[16 * e + 4 * i + j], outer(F.a[i], F.a[j]) * k_bend);
        }
    }

    // ---- PCG ------------------------------------------------------------
    // y = A x (diag first, then off-diagonals in CSR order): block_csr.cu:54-83.
    void spmv(const std::vector<V3<T>>& xin, std::vector<V3<T>>& y) const {
        y.resize(n);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            V3<T> acc = diag[i] * xin[i];
            for (int k = pat.row_offsets[i]; k < pat.row_offsets[i + 1]; ++k)
                acc = acc + values[k] * xin[pat.col_indices[k]];
            y[i] = acc;
        }
    }

    // FP64 inner product over fixed 4096-row chunks (thread-count independent).
    static double dot64(const std::vector<V3<T>>& a, const std::vector<V3<T>>& b) {
        const int n_ = static_cast<int>(a.size());
        const int chunk = 4096;
        const int nc = (n_ + chunk - 1) / chunk;
        std::vector<double> part(nc, 0.0);
#pragma omp parallel for schedule(static)
        for (int c = 0; c < nc; ++c) {
            double s = 0.0;
            const int e = std::min(n_, (c + 1) * chunk);
            for (int i = c * chunk; i < e; ++i)
                s += double(a[i].x) * double(b[i].x) + double(a[i].y) * double(b[i].y) + double(a[i].z) * double(b[i].z);
            part[c] = s;
        }
        double s = 0.0;
        for (double v : part) s += v;
        return s;
    }

    // Scalar divide guard |den| > 1e-37: ChysX/src/solver/pcg_solver.cu:122-127.
    static T guarded_div(double a, double b) {
        return (b > 1e-37 || b < -1e-37) ? T(a / b) : T(0);
    }

    // Contact sidecar of the standalone solve (pcg_solver.h:111-131): when
    // set, every product is (A + C_off) x (pcg_solver.cu:159-160, 180-181).
    const ContactOp* sidecar = nullptr;
    void product(const std::vector<V3<T>>& src, std::vector<V3<T>>& dst) {
        spmv(src, dst);
        if (!sidecar) return;
        std::vector<double> xs(3 * size_t(n)), ys(3 * size_t(n), 0.0);
        for (int i = 0; i < n; ++i) {
            xs[3 * i] = double(src[i].x);
            xs[3 * i + 1] = double(src[i].y);
            xs[3 * i + 2] = double(src[i].z);
        }
        sidecar->apply(xs.data(), ys.data(), 1.0);
        for (int i = 0; i < n; ++i)
            dst[i] = V3<T>(dst[i].x + T(ys[3 * i]), dst[i].y + T(ys[3 * i + 1]), dst[i].z + T(ys[3 * i + 2]));
    }

    // Block-Jacobi PCG, warm-started from x: ChysX/src/solver/pcg_solver.cu:134-215.
    int pcg(const std::vector<V3<T>>& b, std::vector<V3<T>>& x) {
        std::vector<M3<T>> Minv(n);
        std::vector<V3<T>> r(n), z(n), p(n), q(n), t(n);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) Minv[i] = inverse(diag[i]);
        product(x, t);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            r[i] = b[i] - t[i];
            z[i] = Minv[i] * r[i];
            p[i] = z[i];
        }
        double rho = dot64(r, z);
        const double bb = dot64(b, b);
        double rr = dot64(r, r);
        const bool use_tol = tol > 0.0;
        const double thresh = tol * tol * bb;
        int it = 0;
        for (; it < max_iter; ++it) {
            if (use_tol && rr <= thresh) break;
            product(p, q);
            const double sigma = dot64(p, q);
            const T alpha = guarded_div(rho, sigma);
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) {
                x[i] = x[i] + p[i] * alpha;
                r[i] = r[i] - q[i] * alpha;
                z[i] = Minv[i] * r[i];
            }
            const double rho_new = dot64(r, z);
            rr = dot64(r, r);
            const T beta = guarded_div(rho_new, rho);
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) p[i] = z[i] + p[i] * beta;
            rho = rho_new;
        }
        last_iters = it;
        last_rr = rr;
        last_bb = bb;
        last_rho = rho;
        return it;
    }

    // ---- step: ClothSimulator::step, cloth_simulator.cu:814-1261 --------
    int step(T dt) {
        if (n <= 0) return 0;
        if (!(dt > T(0))) throw std::invalid_argument("step: dt must be positive");
        if (static_cast<int>(dx.size()) != n) dx.assign(n, V3<T>());  // :802-811
        // predictor (:886-900, :66-106)
        std::vector<V3<T>> x_n = pos, x_t(n);
        std::vector<T> mass(n);
        const V3<T> g(gx, gy, gz);
        for (int i = 0; i < n; ++i) {
            x_t[i] = V3<T>(pos[i].x + dt * vel[i].x + dt * dt * g.x,
                           pos[i].y + dt * vel[i].y + dt * dt * g.y,
                           pos[i].z + dt * vel[i].z + dt * dt * g.z);
            if (!has_inv_mass) mass[i] = T(1);
            else mass[i] = (inv_mass[i] > T(1.0e-12f)) ? T(1) / inv_mass[i] : T(1.0e8f);
        }
        // gradient at x_n and rhs (:914-1061, :127-142)
        std::vector<V3<T>> grad;
        gradient(x_n, grad);
        // static contact at x_n after the elastic gradients (:997-1008;
        // static_contact.cu:175-216): g += -k depth n, and with friction
        // g += -alpha u_t (the reference's sign: the force side gets +alpha u_t)
        const bool contact_on = contact.active();
        std::vector<V3<T>> c_n, c_ut;
        std::vector<T> c_depth, c_unorm;
        if (contact_on) {
            c_n.resize(n); c_ut.resize(n); c_depth.resize(n); c_unorm.resize(n);
            for (int i = 0; i < n; ++i) {
                contact.detect(x_n[i], vel[i], dt, c_n[i], c_depth[i], c_ut[i], c_unorm[i]);
                if (c_depth[i] <= T(0)) continue;
                const T kd = -contact.stiffness * c_depth[i];
                grad[i].x += kd * c_n[i].x;
                grad[i].y += kd * c_n[i].y;
                grad[i].z += kd * c_n[i].z;
                if (contact.friction > T(0)) {
                    const T fn = contact.stiffness * c_depth[i];
                    const T alpha = contact.friction * fn * ipc_f1_over_x(c_unorm[i], contact.eps);
                    grad[i].x += -alpha * c_ut[i].x;
                    grad[i].y += -alpha * c_ut[i].y;
                    grad[i].z += -alpha * c_ut[i].z;
                }
            }
        }
        const T inv_dt2 = T(1) / (dt * dt);
        rhs.resize(n);
        for (int i = 0; i < n; ++i) {
            const T m = mass[i] * inv_dt2;
            rhs[i] = V3<T>(m * (x_t[i].x - x_n[i].x) - grad[i].x,
                           m * (x_t[i].y - x_n[i].y) - grad[i].y,
                           m * (x_t[i].z - x_n[i].z) - grad[i].z);
        }
        // Coulomb-cone projection of the assembled rhs (:1073-1078;
        // static_contact.cu:299-351)
        if (contact_on && contact.friction > T(0)) {
            for (int i = 0; i < n; ++i) {
                if (c_depth[i] <= T(0)) continue;
                const V3<T> cn = c_n[i];
                const T push = contact.stiffness * c_depth[i];
                const V3<T> F0(rhs[i].x - push * cn.x, rhs[i].y - push * cn.y, rhs[i].z - push * cn.z);
                const T f0n = F0.x * cn.x + F0.y * cn.y + F0.z * cn.z;
                const T fnm = std::max(T(0), -f0n);
                if (fnm <= T(0)) continue;
                const V3<T> F0t(F0.x - f0n * cn.x, F0.y - f0n * cn.y, F0.z - f0n * cn.z);
                const T ft2 = F0t.x * F0t.x + F0t.y * F0t.y + F0t.z * F0t.z;
                if (ft2 <= T(0)) continue;
                const T ft = std::sqrt(ft2);
                const T cone = contact.friction * fnm;
                if (ft <= cone) {
                    rhs[i].x -= F0t.x; rhs[i].y -= F0t.y; rhs[i].z -= F0t.z;
                } else {
                    const T shrink = cone / ft;
                    rhs[i].x -= shrink * F0t.x; rhs[i].y -= shrink * F0t.y; rhs[i].z -= shrink * F0t.z;
                }
            }
        }
        if (topology_dirty || topo_n != n) ensure_topology();
        // Hessian at x_n + inertia diagonal (:1129-1146, :199-210)
        hessian(x_n);
        for (int i = 0; i < n; ++i) {
            const T d = mass[i] * inv_dt2;
            diag[i].m[0] += d; diag[i].m[4] += d; diag[i].m[8] += d;
        }
        // static contact diagonal after the inertia (:1189-1193;
        // static_contact.cu:236-296): k n n^T + alpha (I - n n^T)
        if (contact_on) {
            for (int i = 0; i < n; ++i) {
                if (c_depth[i] <= T(0)) continue;
                const T k = contact.stiffness, nx = c_n[i].x, ny = c_n[i].y, nz = c_n[i].z;
                T a00 = k * nx * nx, a01 = k * nx * ny, a02 = k * nx * nz;
                T a11 = k * ny * ny, a12 = k * ny * nz, a22 = k * nz * nz;
                if (contact.friction > T(0)) {
                    const T alpha = contact.friction * (k * c_depth[i]) * ipc_f1_over_x(c_unorm[i], contact.eps);
                    a00 += alpha * (T(1) - nx * nx);
                    a01 += alpha * (-nx * ny);
                    a02 += alpha * (-nx * nz);
                    a11 += alpha * (T(1) - ny * ny);
                    a12 += alpha * (-ny * nz);
                    a22 += alpha * (T(1) - nz * nz);
                }
                M3<T>& A = diag[i];
                A.m[0] += a00; A.m[1] += a01; A.m[2] += a02;
                A.m[3] += a01; A.m[4] += a11; A.m[5] += a12;
                A.m[6] += a02; A.m[7] += a12; A.m[8] += a22;
            }
        }
        const int it = pcg(rhs, dx);
        // finalize (:222-244)
        const T inv_dt = T(1) / dt;
        const T decay = (damping > T(0)) ? std::exp(-damping * dt) : T(1);
        for (int i = 0; i < n; ++i) {
            pos[i] = x_n[i] + dx[i];
            V3<T> v = dx[i] * inv_dt;
            if (damping > T(0)) v = v * decay;
            vel[i] = v;
        }
        return it;
    }
};

}  // namespace chx_oracle
