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

}  // namespace chx
