// Per-element cloth math on the device: BW98 stretch/shear triangles,
// dihedral bending and penalty pins.  The formulas follow
// ChysX/src/constraint/cloth/{triangle_stretch,bending,pin}_constraint.cu;
// the Hessian is stored in a factored form so that one record per element
// serves every 3x3 block of that element:
//
//   triangle  H_ab = (dwu_a dwu_b) Mu + (dwv_a dwv_b) Mv,   g_a = dwu_a Gu + dwv_a Gv
//             Mu   = kA wu^ wu^T + s_u (I - wu^ wu^T),   s_u = [cu > 0] kA cu / |wu|
//             Gu   = kA cu wu^
//   dihedral  H_ij = k a_i a_j^T,   g_i = k (theta - theta0) a_i
//   pin       H    = k I,           g   = k (x - t)
//
// which is algebraically the reference's per-block expression
// (triangle_stretch_constraint.cu:231-311, bending_constraint.cu:217-245).
#pragma once

#include "chx_device.h"

namespace chx {

constexpr float kTriMinNorm = 1.0e-6f;   // triangle_stretch_constraint.cu:39
constexpr float kDihMinNorm = 1.0e-12f;  // bending_constraint.cu:34

// Rest shape: triangle_stretch_constraint.cu:50-110.
__device__ __forceinline__ void tri_rest_shape(f3 a, f3 b, f3 c, float cos_t, float sin_t, float* dm, float& area) {
    const f3 ab = sub(b, a), ac = sub(c, a);
    const float lab = len(ab), lac = len(ac);
    const f3 ab_hat = (lab > kTriMinNorm) ? mul(ab, 1.0f / lab) : mk(1.0f, 0.0f, 0.0f);
    const float u_c = dot(ab_hat, ac);
    const float v_c2 = lac * lac - u_c * u_c;
    const float v_c = (v_c2 > 0.0f) ? sqrtf(v_c2) : 0.0f;
    area = 0.5f * lab * v_c;
    const float det = lab * v_c;
    float r0, r1, r2, r3;
    if (det > kTriMinNorm) {
        const float inv_det = 1.0f / (lab * v_c - u_c * 0.0f);
        r0 = v_c * inv_det;
        r1 = -u_c * inv_det;
        r2 = -0.0f * inv_det;
        r3 = lab * inv_det;
    } else {
        r0 = 1.0f; r1 = 0.0f; r2 = 0.0f; r3 = 1.0f;
    }
    // Dm_inv = raw * [[c, s], [-s, c]]
    dm[0] = r0 * cos_t + r1 * (-sin_t);
    dm[1] = r0 * sin_t + r1 * cos_t;
    dm[2] = r2 * cos_t + r3 * (-sin_t);
    dm[3] = r2 * sin_t + r3 * cos_t;
}

// Triangle record (CHX_TRI_REC floats):
// [0..2] Gu, [3..5] Gv, [6..11] Mu (xx xy xz yy yz zz), [12..17] Mv,
// [18..20] dwu, [21..23] dwv.
__device__ __forceinline__ void tri_record(f3 x0, f3 x1, f3 x2, float d00, float d01, float d10, float d11, float kA,
                                           float* rec) {
    const f3 a = sub(x1, x0), b = sub(x2, x0);
    const f3 wu = add(mul(a, d00), mul(b, d10));
    const f3 wv = add(mul(a, d01), mul(b, d11));
    float lu = len(wu), lv = len(wv);
    if (lu < kTriMinNorm) lu = kTriMinNorm;
    if (lv < kTriMinNorm) lv = kTriMinNorm;
    const float inv_lu = 1.0f / lu, inv_lv = 1.0f / lv;
    const float cu = lu - 1.0f, cv = lv - 1.0f;
    const f3 hu = mul(wu, inv_lu), hv = mul(wv, inv_lv);
    const f3 Gu = mul(hu, kA * cu), Gv = mul(hv, kA * cv);
    const float su = (cu > 0.0f) ? kA * cu * inv_lu : 0.0f;
    const float sv = (cv > 0.0f) ? kA * cv * inv_lv : 0.0f;
    const float tu = kA - su, tv = kA - sv;
    rec[0] = Gu.x; rec[1] = Gu.y; rec[2] = Gu.z;
    rec[3] = Gv.x; rec[4] = Gv.y; rec[5] = Gv.z;
    rec[6] = tu * hu.x * hu.x + su; rec[7] = tu * hu.x * hu.y; rec[8] = tu * hu.x * hu.z;
    rec[9] = tu * hu.y * hu.y + su; rec[10] = tu * hu.y * hu.z; rec[11] = tu * hu.z * hu.z + su;
    rec[12] = tv * hv.x * hv.x + sv; rec[13] = tv * hv.x * hv.y; rec[14] = tv * hv.x * hv.z;
    rec[15] = tv * hv.y * hv.y + sv; rec[16] = tv * hv.y * hv.z; rec[17] = tv * hv.z * hv.z + sv;
    rec[18] = -d00 - d10; rec[19] = d00; rec[20] = d10;
    rec[21] = -d01 - d11; rec[22] = d01; rec[23] = d11;
}

// BW98 energy (norms unclamped): triangle_stretch_constraint.cu:116-147.
__device__ __forceinline__ double tri_energy(f3 x0, f3 x1, f3 x2, float d00, float d01, float d10, float d11, float kA) {
    const f3 a = sub(x1, x0), b = sub(x2, x0);
    const f3 wu = add(mul(a, d00), mul(b, d10));
    const f3 wv = add(mul(a, d01), mul(b, d11));
    const float cu = len(wu) - 1.0f, cv = len(wv) - 1.0f;
    return double(0.5f * kA * (cu * cu + cv * cv));
}

// ---------------------------------------------------------------------------
// StVK membrane (no reference counterpart).  F = [fu fv] from Dm^-1,
// G = (F^T F - I)/2, E = muA ||G||^2 + lamA/2 (tr G)^2, S = 2 muA G +
// lamA tr(G) I, dE/dF = F S.  Hessian: the Gauss-Newton part
// 2 muA ||dG||^2 + lamA (tr dG)^2 (PSD) plus the stress part S+ : dF^T dF
// with S clamped to its non-negative eigen-part S+ (exact when S is PSD):
//   Huu = (2 muA + lamA) fu fu^T + muA fv fv^T + S+uu I
//   Hvv = (2 muA + lamA) fv fv^T + muA fu fu^T + S+vv I
//   Huv = muA fv fu^T + lamA fu fv^T + S+uv I,   Hvu = Huv^T
// and H_ab = dwu_a dwu_b Huu + dwu_a dwv_b Huv + dwv_a dwu_b Hvu + dwv_a dwv_b Hvv.
// Record (CHX_STVK_REC floats): [0..2] Gu = (F S)_u, [3..5] Gv, [6..11] Huu
// (xx xy xz yy yz zz), [12..17] Hvv, [18..26] Huv row-major, [27..29] dwu,
// [30..32] dwv, [33..35] 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stvk_psd2(float a, float b, float c, float& pa, float& pb, float& pc) {
    const float m = 0.5f * (a + b), d = 0.5f * (a - b);
    const float r = sqrtf(d * d + c * c);
    const float l1 = m + r, l2 = m - r;
    if (l2 >= 0.0f) { pa = a; pb = b; pc = c; return; }
    if (l1 <= 0.0f) { pa = pb = pc = 0.0f; return; }
    float vx, vy;
    if (d >= 0.0f) { vx = d + r; vy = c; } else { vx = c; vy = r - d; }
    const float nn = vx * vx + vy * vy;
    if (nn <= 0.0f) { pa = l1; pb = 0.0f; pc = 0.0f; return; }
    const float s = l1 / nn;
    pa = s * vx * vx;
    pb = s * vy * vy;
    pc = s * vx * vy;
}

__device__ __forceinline__ void stvk_record(f3 x0, f3 x1, f3 x2, float d00, float d01, float d10, float d11,
                                            float muA, float lamA, float* rec) {
    const f3 a = sub(x1, x0), b = sub(x2, x0);
    const f3 fu = add(mul(a, d00), mul(b, d10));
    const f3 fv = add(mul(a, d01), mul(b, d11));
    const float Guu = 0.5f * (dot(fu, fu) - 1.0f), Gvv = 0.5f * (dot(fv, fv) - 1.0f), Guv = 0.5f * dot(fu, fv);
    const float tr = Guu + Gvv;
    const float Suu = 2.0f * muA * Guu + lamA * tr, Svv = 2.0f * muA * Gvv + lamA * tr, Suv = 2.0f * muA * Guv;
    const f3 Pu = add(mul(fu, Suu), mul(fv, Suv)), Pv = add(mul(fu, Suv), mul(fv, Svv));
    float Puu, Pvv, Puv;
    stvk_psd2(Suu, Svv, Suv, Puu, Pvv, Puv);
    const float c1 = 2.0f * muA + lamA;
    rec[0] = Pu.x; rec[1] = Pu.y; rec[2] = Pu.z;
    rec[3] = Pv.x; rec[4] = Pv.y; rec[5] = Pv.z;
    // Huu
    rec[6] = c1 * fu.x * fu.x + muA * fv.x * fv.x + Puu;
    rec[7] = c1 * fu.x * fu.y + muA * fv.x * fv.y;
    rec[8] = c1 * fu.x * fu.z + muA * fv.x * fv.z;
    rec[9] = c1 * fu.y * fu.y + muA * fv.y * fv.y + Puu;
    rec[10] = c1 * fu.y * fu.z + muA * fv.y * fv.z;
    rec[11] = c1 * fu.z * fu.z + muA * fv.z * fv.z + Puu;
    // Hvv
    rec[12] = c1 * fv.x * fv.x + muA * fu.x * fu.x + Pvv;
    rec[13] = c1 * fv.x * fv.y + muA * fu.x * fu.y;
    rec[14] = c1 * fv.x * fv.z + muA * fu.x * fu.z;
    rec[15] = c1 * fv.y * fv.y + muA * fu.y * fu.y + Pvv;
    rec[16] = c1 * fv.y * fv.z + muA * fu.y * fu.z;
    rec[17] = c1 * fv.z * fv.z + muA * fu.z * fu.z + Pvv;
    // Huv[r][c] = muA fv_r fu_c + lamA fu_r fv_c + S+uv delta_rc
    const float fur[3] = {fu.x, fu.y, fu.z}, fvr[3] = {fv.x, fv.y, fv.z};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) rec[18 + 3 * r + c] = muA * fvr[r] * fur[c] + lamA * fur[r] * fvr[c] + (r == c ? Puv : 0.0f);
    rec[27] = -d00 - d10; rec[28] = d00; rec[29] = d10;
    rec[30] = -d01 - d11; rec[31] = d01; rec[32] = d11;
    rec[33] = 0.0f; rec[34] = 0.0f; rec[35] = 0.0f;
}

__device__ __forceinline__ double stvk_energy(f3 x0, f3 x1, f3 x2, float d00, float d01, float d10, float d11,
                                              float muA, float lamA) {
    const f3 a = sub(x1, x0), b = sub(x2, x0);
    const f3 fu = add(mul(a, d00), mul(b, d10));
    const f3 fv = add(mul(a, d01), mul(b, d11));
    const float Guu = 0.5f * (dot(fu, fu) - 1.0f), Gvv = 0.5f * (dot(fv, fv) - 1.0f), Guv = 0.5f * dot(fu, fv);
    const float tr = Guu + Gvv;
    return double(muA * (Guu * Guu + Gvv * Gvv + 2.0f * Guv * Guv) + 0.5f * lamA * tr * tr);
}

// Dihedral frame: bending_constraint.cu:54-126.
struct DihFrame {
    float theta;
    f3 a[4];
    bool valid;
};

__device__ __forceinline__ DihFrame dihedral_frame(f3 v0, f3 v1, f3 v2, f3 v3) {
    DihFrame F;
    F.valid = false;
    F.theta = 0.0f;
    const f3 e10 = sub(v1, v0), e20 = sub(v2, v0), e30 = sub(v3, v0);
    const float l = len(e10);
    if (l < kDihMinNorm) return F;
    const float inv_l = 1.0f / l;
    const f3 e10_hat = mul(e10, inv_l);
    const f3 n1r = cross(e20, e10), n2r = cross(e10, e30);
    const float n1l = len(n1r), n2l = len(n2r);
    if (n1l < kDihMinNorm || n2l < kDihMinNorm) return F;
    const f3 n1 = mul(n1r, 1.0f / n1l), n2 = mul(n2r, 1.0f / n2l);
    float s = dot(cross(n1, n2), e10_hat);
    if (s > 1.0f) s = 1.0f;
    if (s < -1.0f) s = -1.0f;
    const float c = dot(n1, n2);
    F.theta = atan2f(s, c);
    const float inv_l2 = inv_l * inv_l;
    const float om1 = dot(e10, e20) * inv_l2, om2 = dot(e10, e30) * inv_l2;
    const f3 p1 = sub(e20, mul(e10, om1)), p2 = sub(e30, mul(e10, om2));
    const float h1 = len(p1), h2 = len(p2);
    if (h1 < kDihMinNorm || h2 < kDihMinNorm) return F;
    const float ih1 = 1.0f / h1, ih2 = 1.0f / h2;
    const float t1[4] = {(om1 - 1.0f) * ih1, -om1 * ih1, ih1, 0.0f};
    const float t2[4] = {(om2 - 1.0f) * ih2, -om2 * ih2, 0.0f, ih2};
#pragma unroll
    for (int i = 0; i < 4; ++i) F.a[i] = add(mul(n1, t1[i]), mul(n2, t2[i]));
    F.valid = true;
    return F;
}

// ---- static plane / box contact ------------------------------------------
// Oriented-box signed distance (positive outside, the nearest face's gap
// inside) and outward unit normal (ChysX/src/collision/static_contact.cu:50-96).
__device__ __forceinline__ float box_signed_dist(f3 p, f3 c, f3 h, f3 ex, f3 ey, f3 ez, f3& n) {
    const f3 rel = sub(p, c);
    const float qx = dot(rel, ex), qy = dot(rel, ey), qz = dot(rel, ez);
    const float dx = fabsf(qx) - h.x, dy = fabsf(qy) - h.y, dz = fabsf(qz) - h.z;
    if (dx <= 0.0f && dy <= 0.0f && dz <= 0.0f) {
        if (dx >= dy && dx >= dz) {
            n = (qx >= 0.0f) ? ex : mul(ex, -1.0f);
            return dx;
        }
        if (dy >= dz) {
            n = (qy >= 0.0f) ? ey : mul(ey, -1.0f);
            return dy;
        }
        n = (qz >= 0.0f) ? ez : mul(ez, -1.0f);
        return dz;
    }
    const float cx = dx > 0.0f ? dx : 0.0f, cy = dy > 0.0f ? dy : 0.0f, cz = dz > 0.0f ? dz : 0.0f;
    const float dist = sqrtf(cx * cx + cy * cy + cz * cz);
    const float inv = 1.0f / fmaxf(dist, 1.0e-30f);
    const float sx = qx > 0.0f ? cx : -cx, sy = qy > 0.0f ? cy : -cy, sz = qz > 0.0f ? cz : -cz;
    n = add(add(mul(ex, sx * inv), mul(ey, sy * inv)), mul(ez, sz * inv));
    return dist;
}

// Slip ramp of the lagged Coulomb friction: 1/u past eps, linear from 2/eps
// at rest to 1/eps at eps (ChysX/src/collision/friction.cuh:23-34).
__device__ __forceinline__ float ipc_f1_over_x(float u, float eps) {
    if (u > eps) return 1.0f / u;
    return (-u / eps + 2.0f) / eps;
}

}  // namespace chx
