// C ABI: the implicit cloth substep on gfx950.
//
//   assembly       one workgroup per assembly unit (a row range of one
//                  instance): the unit's element force / Hessian records are
//                  computed into LDS from x_n, then one thread per block the
//                  unit owns sums that block's contributions in the fixed
//                  order of its contribution list, so the assembled matrix is
//                  bitwise reproducible with no float atomics.  Diagonal
//                  blocks also form the right-hand side
//                  b = m/dt^2 (x~ - x_n) - grad E(x_n) and the inertia block
//                  (cloth_simulator.cu:66-142,199-210).
//   PCG            one persistent cooperative kernel per solve: block-Jacobi
//                  PCG, FP64 per-tile partial sums reduced in a fixed order,
//                  one grid barrier per iteration (FP32 cloth instances) or
//                  two (BF16 mode, caller-provided row-major systems),
//                  per-instance convergence, then the finalize
//                  x = x_n + dx, v = dx/dt (cloth_simulator.cu:222-244).
//
// Every instance owns its arrays (local indices); the batch gives work lists
// of row tiles (CHX_TILE_ROWS rows of one instance), element items (energy)
// and assembly units.  Each row tile's partial sums live in
// batch->partials[tile * CHX_TILE_PARTIALS + field] (fields: P_* for the
// two-barrier form, Q_* for the single-barrier form).
#include <algorithm>
#include <mutex>
#include <type_traits>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "chx_contact.h"
#include "chx_device.h"
#include "chx_elements.h"
#include "chx_internal.h"

namespace chx {
namespace {

constexpr int kBlock = 256;
static_assert(CHX_ITEM_SIZE == kBlock, "one thread per work-item entry");

// Tile partials: rho = <r,z> and ||r||^2 per parity, sigma = <p,q>, ||b||^2,
// and per parity the refinement state of that residual (BF16 mode): 2 *
// (refinements so far) + 1 if it is the true residual.  Every tile of an
// instance holds the same state.
enum { P_RHO0 = 0, P_RHO1 = 1, P_RR0 = 2, P_RR1 = 3, P_SIG = 4, P_BB = 5, P_TRUE0 = 6, P_TRUE1 = 7 };

// Divide guard |den| > 1e-37 (ChysX/src/solver/pcg_solver.cu:122-127), with
// the FP64 inner products rounded to FP32 once.
__device__ __forceinline__ float guarded_div(double a, double b) {
    return (b > 1e-37 || b < -1e-37) ? float(a / b) : 0.0f;
}

// Resets the persistent solver's barrier / control words before a PCG launch
// on the same stream.  The status word and the abort report next to it are
// kept (all == false): a timed-out solve stays reported until
// chx_batch_check reads and clears it.
__device__ __forceinline__ void reset_sync(unsigned* sync, bool all) {
    for (int k = threadIdx.x; k < SYNC_WORDS; k += blockDim.x)
        if (all || k < SYNC_STATUS || k > SYNC_WAITING) sync[k] = 0u;
}

// ---------------------------------------------------------------------------
// Assembly: one kernel, one workgroup per assembly unit (a row range of one
// instance, chx_unit).
//
//   phase 1  every element with a vertex in the unit's rows -- the unit's
//            element list, ascending per kind -- gathers x_n from the
//            instance's AoS positions and writes its force / Hessian record
//            into LDS as 16-byte quads, quad-major (quad q of local element
//            e at q * count + e), so each record is read from LDS instead
//            of from memory once per block it touches.  The unit's
//            contribution entries and slot offsets are staged into LDS
//            alongside with 16-byte loads;
//   phase 2  one work item per block the unit owns (the diagonal blocks of
//            its rows on kDiagLanes lanes each, then its upper value slots
//            on one) sums the block's contributions in the fixed order of
//            its contribution list, all operands in LDS, so the assembled
//            matrix is bitwise
//            reproducible with no float atomics.  Diagonal blocks also form
//            the right-hand side b = m/dt^2 (x~ - x_n) - grad E(x_n) and the
//            inertia block (cloth_simulator.cu:66-142,199-210).
//
// Records hold the Hessian in factored form (chx_elements.h): all blocks of
// an element come from one record.  The element formulas are the
// reference's (triangle_stretch_constraint.cu:153-311,
// bending_constraint.cu:180-245, pin_constraint.cu:61-92).
// ---------------------------------------------------------------------------
// LDS records of a unit, kind by kind, quad-major: 16-byte quad q of
// local element e of a kind with n elements at quad q * n + e, so a wave's
// lanes writing quad q of consecutive elements hit consecutive quads (every
// bank once) and lanes reading quad q of scattered elements spread over
// all banks (an element-major stride of 4, 6 or 8 quads would keep them on
// a quarter, half or eighth of the banks).
// Triangle (24): Dm^-1 (0..3), Mu, Mv upper triangles (4..15), Gu, Gv
// (16..21).  StVK (32): Dm^-1 (0..3), Gu, Gv (4..9), Huu, Hvv upper
// triangles (10..21), Huv row-major (22..30).  Dihedral (16): quad j holds
// a_j and, in its last float, k (theta - theta0) (j = 0), k (j = 1) or 0,
// so a block (a, b) reads quads a and b whole.  Pin (4): k (x - t), k.
static_assert(CHX_ASM_TRI_REC == 24 && CHX_ASM_STVK_REC == 32 && CHX_ASM_DIH_REC == 16 && CHX_ASM_PIN_REC == 4,
              "record layouts");
static_assert(CHX_ASM_LDS_BYTES % 16 == 0, "16-byte staging");
// Floats tri_record / stvk_record write (chx_elements.h): the LDS records
// keep their first 18 / 27 and Dm^-1 in place of the shape derivatives.
constexpr int kTriRec = 24, kStvkRec = 36;
// 16-byte vectors of a unit's entries and offsets each thread holds in
// registers across phase 1 (the rest is copied after it).
constexpr int kStageRegs = 2;
// Lanes per diagonal block in the assembly's block phase (a power of two
// dividing the wavefront, so a block's lanes share a wavefront).
// Four lanes per diagonal list (~24 entries on a triangle mesh with
// bending) against one per upper list (at most ~10): 0.92 ms per assembly
// of the 512-cloth batch, against 0.99 with two lanes, 1.04 with eight and
// 1.27 with one; two lanes per upper list as well: 1.01.
constexpr int kDiagLanes = 4;
static_assert((kDiagLanes & (kDiagLanes - 1)) == 0 && kWave % kDiagLanes == 0 && kBlock % kDiagLanes == 0,
              "diagonal lane groups stay inside a wavefront");
typedef unsigned v4u __attribute__((ext_vector_type(4)));

// Finish of one assembled slot: a diagonal block (row s) with the
// predictor, right-hand side, inertia and static contact, or an upper value
// slot s (planar FP32 block, and its BF16 copy in BF16 mode).
__device__ __forceinline__ void finish_slot(const Inst& I, bool is_diag, unsigned s, float (&h)[9], f3 grad) {
    if (is_diag) {
        const unsigned row = s;
        const f3 p = ld3g(I.pos, row), v = ld3g(I.vel, row);
        float m = 1.0f;  // mass_from_inv_mass_kernel, cloth_simulator.cu:66-77
        if (I.inv_mass != nullptr) {
            const float w = ldx(I.inv_mass, row);
            m = (w > 1.0e-12f) ? (1.0f / w) : 1.0e8f;
        }
        const float dt = I.dt;
        const float inv_dt2 = 1.0f / (dt * dt);
        // x~ = x + dt v + dt^2 g (compute_x_tilde_kernel, :83-106)
        const f3 xt = mk(p.x + dt * v.x + dt * dt * I.gx, p.y + dt * v.y + dt * dt * I.gy,
                         p.z + dt * v.z + dt * dt * I.gz);
        const float mi = m * inv_dt2;
        // Static plane / box contact at x_n, after the elastic gradients
        // (cloth_simulator.cu:997-1008; static_contact.cu:106-216): the
        // deepest primitive's penalty gradient -k depth n and, with
        // friction, -alpha u_t (the reference's sign) join the gradient.
        f3 cn = mk(0.f, 0.f, 0.f), ut = mk(0.f, 0.f, 0.f);
        float depth = 0.0f, unorm = 0.0f;
        const bool contact = I.contact_stiffness > 0.0f && (I.n_planes > 0 || I.n_boxes > 0);
        if (contact) {
            const float th = I.contact_thickness;
            for (int q = 0; q < I.n_planes; ++q) {
                const gp<const float> pl = rec_at(gp<const float>(I.planes), unsigned(q), 4u);
                const f3 nn = mk(pl[0], pl[1], pl[2]);
                const float d = th - (dot(nn, p) + pl[3]);
                if (d > depth) {
                    depth = d;
                    cn = nn;
                }
            }
            for (int q = 0; q < I.n_boxes; ++q) {
                const gp<const float> b = rec_at(gp<const float>(I.boxes), unsigned(q), 16u);
                f3 nn;
                const float d = th - box_signed_dist(p, mk(b[0], b[1], b[2]), mk(b[3], b[4], b[5]),
                                                     mk(b[6], b[7], b[8]), mk(b[9], b[10], b[11]),
                                                     mk(b[12], b[13], b[14]), nn);
                if (d > depth) {
                    depth = d;
                    cn = nn;
                }
            }
            if (depth > 0.0f) {
                const f3 dxv = mul(v, dt);
                ut = sub(dxv, mul(cn, dot(cn, dxv)));
                unorm = sqrtf(ut.x * ut.x + ut.y * ut.y + ut.z * ut.z);
                const float kd = -I.contact_stiffness * depth;
                grad = mk(grad.x + kd * cn.x, grad.y + kd * cn.y, grad.z + kd * cn.z);
                if (I.contact_friction > 0.0f) {
                    const float fn = I.contact_stiffness * depth;
                    const float alpha = I.contact_friction * fn * ipc_f1_over_x(unorm, I.contact_friction_eps);
                    grad = mk(grad.x + -alpha * ut.x, grad.y + -alpha * ut.y, grad.z + -alpha * ut.z);
                }
            }
        }
        f3 b = mk(mi * (xt.x - p.x) - grad.x, mi * (xt.y - p.y) - grad.y, mi * (xt.z - p.z) - grad.z);
        if (depth > 0.0f && I.contact_friction > 0.0f) {
            // Coulomb-cone projection of the assembled force
            // (cloth_simulator.cu:1073-1078; static_contact.cu:299-351)
            const float push = I.contact_stiffness * depth;
            const f3 F0 = mk(b.x - push * cn.x, b.y - push * cn.y, b.z - push * cn.z);
            const float f0n = F0.x * cn.x + F0.y * cn.y + F0.z * cn.z;
            const float fnm = fmaxf(0.0f, -f0n);
            if (fnm > 0.0f) {
                const f3 F0t = mk(F0.x - f0n * cn.x, F0.y - f0n * cn.y, F0.z - f0n * cn.z);
                const float ft2 = F0t.x * F0t.x + F0t.y * F0t.y + F0t.z * F0t.z;
                if (ft2 > 0.0f) {
                    const float ft = sqrtf(ft2);
                    const float cone = I.contact_friction * fnm;
                    if (ft <= cone) {
                        b = sub(b, F0t);
                    } else {
                        const float shrink = cone / ft;
                        b = mk(b.x - shrink * F0t.x, b.y - shrink * F0t.y, b.z - shrink * F0t.z);
                    }
                }
            }
        }
        st3g(I.rhs, row, b);
        h[0] += mi;
        h[4] += mi;
        h[8] += mi;
        if (depth > 0.0f) {
            // contact diagonal after the inertia (cloth_simulator.cu:1189-1193;
            // static_contact.cu:236-296): k n n^T + alpha (I - n n^T)
            const float k = I.contact_stiffness, nx = cn.x, ny = cn.y, nz = cn.z;
            float a00 = k * nx * nx, a01 = k * nx * ny, a02 = k * nx * nz;
            float a11 = k * ny * ny, a12 = k * ny * nz, a22 = k * nz * nz;
            if (I.contact_friction > 0.0f) {
                const float fn = k * depth;
                const float alpha = I.contact_friction * fn * ipc_f1_over_x(unorm, I.contact_friction_eps);
                a00 += alpha * (1.0f - nx * nx);
                a01 += alpha * (-nx * ny);
                a02 += alpha * (-nx * nz);
                a11 += alpha * (1.0f - ny * ny);
                a12 += alpha * (-ny * nz);
                a22 += alpha * (1.0f - nz * nz);
            }
            h[0] += a00; h[1] += a01; h[2] += a02;
            h[3] += a01; h[4] += a11; h[5] += a12;
            h[6] += a02; h[7] += a12; h[8] += a22;
        }
        // Symmetric grouped planar blocks (bp6): the upper triangle of the
        // block row's diagonal block.
#pragma unroll
        for (unsigned q = 0; q < 6; q += 2) st_pair(I.diag, bp6(q, row), h[sym6_pos(q)], h[sym6_pos(q + 1)]);
    } else {
        const unsigned slot = s;
#pragma unroll
        for (unsigned k = 0; k < 8; k += 2) st_pair(I.values, bp(k, slot), h[k], h[k + 1]);
        stx(I.values, bp(8u, slot), h[8]);
        if (I.values16 != nullptr) {
#pragma unroll
            for (unsigned k = 0; k < 8; k += 2)
                st_pair(I.values16, bp(k, slot), f32_to_bf16(h[k]), f32_to_bf16(h[k + 1]));
            stx(I.values16, bp(8u, slot), f32_to_bf16(h[8]));
        }
    }
}

struct AsmLds {
    // records (floats), then the entries, then the slot offsets (uint16)
    __align__(16) unsigned char arena[CHX_ASM_LDS_BYTES];
    int sl_first[4];   // slices of the unit: first value slot, lane range, width
    int sl_l0[4];
    int sl_cnt[4];
    int sl_w[4];
    int sl_end[4];     // cumulative upper slots through this slice
};

__global__ __launch_bounds__(kBlock) void assemble_kernel(chx_cloth_batch B) {
    if (blockIdx.x == 0) reset_sync(B.sync, false);
    __shared__ AsmLds L;
    int inst, u;
    get_unit_item(B, blockIdx.x, inst, u);
    const Inst I = inst_at(B, inst);
    const gp<const int> ud = rec_at((gp<const int>)I.units, unsigned(u), unsigned(sizeof(chx_unit) / sizeof(int)));
    const int r0 = ud[0], r1 = ud[1], e0 = ud[2];
    const int np = ud[3], nt = ud[4], nv = ud[5], nd = ud[6];
    const int n_slots = ud[7], offs0 = ud[8], word0 = ud[9], n_words = ud[10];
    const int tid = int(threadIdx.x);
    // LDS image: records, then entries at byte wofs, then offsets
    const int rec_floats = CHX_ASM_PIN_REC * np + CHX_ASM_TRI_REC * nt + CHX_ASM_STVK_REC * nv + CHX_ASM_DIH_REC * nd;
    const int wofs = (4 * rec_floats + 15) & ~15;
    const int n_wvec = n_words >> 3, n_ovec = (n_slots + 8) >> 3;
    v4u* const Wv = reinterpret_cast<v4u*>(L.arena + wofs);
    v4u* const Ov = Wv + n_wvec;
    const uint16_t* const Wl = reinterpret_cast<const uint16_t*>(Wv);
    const uint16_t* const Ol = reinterpret_cast<const uint16_t*>(Ov);
    // staging loads of entries + offsets (vector i < n_wvec: entries, then
    // offsets), issued before the element work and stored after it
    const gp<const v4u> wsrc = (gp<const v4u>)rec_at(I.contribs, unsigned(word0), 1u);
    const gp<const v4u> osrc = (gp<const v4u>)rec_at(I.contrib_offsets, unsigned(offs0), 1u);
    const int n_vec = n_wvec + n_ovec;
    auto stage_src = [&](int i) -> v4u { return i < n_wvec ? wsrc[i] : osrc[i - n_wvec]; };
    v4u held[kStageRegs];
#pragma unroll
    for (int k = 0; k < kStageRegs; ++k) {
        const int i = tid + k * kBlock;
        if (i < n_vec) held[k] = stage_src(i);
    }
    // the unit's slices (whole slices, up to four, or part of one)
    const int s_lo = r0 / int(CHX_SLICE_ROWS), s_hi = (r1 - 1) / int(CHX_SLICE_ROWS);
    const int ns = s_hi - s_lo + 1;
    if (tid == 0) {
        int acc = 0;
        for (int q = 0; q < ns; ++q) {
            const int sl = s_lo + q;
            const int a = max(r0, sl * int(CHX_SLICE_ROWS)), b = min(r1, (sl + 1) * int(CHX_SLICE_ROWS));
            const int uo0 = ldx(I.upper_offsets, unsigned(sl)), uo1 = ldx(I.upper_offsets, unsigned(sl + 1));
            L.sl_first[q] = uo0;
            L.sl_l0[q] = a - sl * int(CHX_SLICE_ROWS);
            L.sl_cnt[q] = b - a;
            L.sl_w[q] = (uo1 - uo0) / int(CHX_SLICE_ROWS);
            acc += (b - a) * L.sl_w[q];
            L.sl_end[q] = acc;
        }
    }
    // ---- phase 1: records of the unit's elements into LDS ----
    float4* const Rp = reinterpret_cast<float4*>(L.arena);
    float4* const Rt = Rp + (CHX_ASM_PIN_REC / 4) * np;
    float4* const Rv = Rt + (CHX_ASM_TRI_REC / 4) * nt;
    float4* const Rd = Rv + (CHX_ASM_STVK_REC / 4) * nv;
    auto st4 = [](float4* p, float a, float b, float c, float d) { *p = make_float4(a, b, c, d); };
    // float f of a record (quad f / 4 at f / 4 * n + l)
    auto rf = [](const float4* R, int n, unsigned l, unsigned f) {
        return reinterpret_cast<const float*>(R + (f >> 2) * unsigned(n) + l)[f & 3u];
    };
    const gp<const float> pos = I.pos;
    // pins and StVK triangles (a few per cloth, or StVK cloths only): one
    // element per thread
    for (int l = tid; l < np; l += kBlock) {
        const unsigned id = unsigned(ldx(I.unit_elems, unsigned(e0 + l)));
        const chx_pin pn = ldg(I.pins, id);
        const f3 g = mul(sub(ld3g(pos, unsigned(pn.v)), ld3g(I.pin_targets, id)), pn.k);  // pin_constraint.cu:61-75
        st4(Rp + l, g.x, g.y, g.z, pn.k);
    }
    for (int l = tid; l < nv; l += kBlock) {
        const unsigned id = unsigned(ldx(I.unit_elems, unsigned(e0 + np + nt + l)));
        const chx_stvk t = ldg(I.stvks, id);
        float rec[kStvkRec];
        stvk_record(ld3g(pos, unsigned(t.v0)), ld3g(pos, unsigned(t.v1)), ld3g(pos, unsigned(t.v2)), t.dm00, t.dm01,
                    t.dm10, t.dm11, t.muA, t.lamA, rec);
        st4(Rv + l, t.dm00, t.dm01, t.dm10, t.dm11);
#pragma unroll
        for (int q = 0; q < 7; ++q)
            st4(Rv + (q + 1) * nv + l, rec[4 * q], rec[4 * q + 1], rec[4 * q + 2], q < 6 ? rec[4 * q + 3] : 0.0f);
    }
    // BW98 triangles and dihedrals
    auto tri_out = [&](int l, const chx_tri& t, f3 x0, f3 x1, f3 x2) {
        float rec[kTriRec];
        tri_record(x0, x1, x2, t.dm00, t.dm01, t.dm10, t.dm11, t.kA, rec);
        float4* o = Rt + l;
        st4(o, t.dm00, t.dm01, t.dm10, t.dm11);
        st4(o + nt, rec[6], rec[7], rec[8], rec[9]);
        st4(o + 2 * nt, rec[10], rec[11], rec[12], rec[13]);
        st4(o + 3 * nt, rec[14], rec[15], rec[16], rec[17]);
        st4(o + 4 * nt, rec[0], rec[1], rec[2], rec[3]);
        st4(o + 5 * nt, rec[4], rec[5], 0.0f, 0.0f);
    };
    auto dih_out = [&](int l, const chx_dih& d, f3 x0, f3 x1, f3 x2, f3 x3) {
        const DihFrame F = dihedral_frame(x0, x1, x2, x3);
        // An invalid frame contributes nothing (bending_constraint.cu:190-191,227-228).
        float a[12];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const f3 v = F.valid ? F.a[j] : mk(0.0f, 0.0f, 0.0f);
            a[3 * j] = v.x;
            a[3 * j + 1] = v.y;
            a[3 * j + 2] = v.z;
        }
        float4* o = Rd + l;
        st4(o, a[0], a[1], a[2], F.valid ? d.k * (F.theta - d.rest) : 0.0f);
        st4(o + nd, a[3], a[4], a[5], F.valid ? d.k : 0.0f);
        st4(o + 2 * nd, a[6], a[7], a[8], 0.0f);
        st4(o + 3 * nd, a[9], a[10], a[11], 0.0f);
    };
    if (nt > 0 && nd > 0) {
        // one of each per thread and round, the two dependent load chains
        // (list entry -> element -> positions) issued side by side; lanes
        // past a kind's count redo its last element and store the same bits
        // to its slot, so no branch separates (or sinks) the two chains
        for (int l = tid; l < max(nt, nd); l += kBlock) {
            const int lt = min(l, nt - 1), ld = min(l, nd - 1);
            const unsigned idt = unsigned(ldx(I.unit_elems, unsigned(e0 + np + lt)));
            const unsigned idd = unsigned(ldx(I.unit_elems, unsigned(e0 + np + nt + nv + ld)));
            const chx_tri t = ldg(I.tris, idt);
            const chx_dih d = ldg(I.dihs, idd);
            const f3 t0 = ld3g(pos, unsigned(t.v0)), t1 = ld3g(pos, unsigned(t.v1)), t2 = ld3g(pos, unsigned(t.v2));
            const f3 d0 = ld3g(pos, unsigned(d.v0)), d1 = ld3g(pos, unsigned(d.v1));
            const f3 d2 = ld3g(pos, unsigned(d.v2)), d3 = ld3g(pos, unsigned(d.v3));
            tri_out(lt, t, t0, t1, t2);
            dih_out(ld, d, d0, d1, d2, d3);
        }
    } else {
        for (int l = tid; l < nt; l += kBlock) {
            const chx_tri t = ldg(I.tris, unsigned(ldx(I.unit_elems, unsigned(e0 + np + l))));
            tri_out(l, t, ld3g(pos, unsigned(t.v0)), ld3g(pos, unsigned(t.v1)), ld3g(pos, unsigned(t.v2)));
        }
        for (int l = tid; l < nd; l += kBlock) {
            const chx_dih d = ldg(I.dihs, unsigned(ldx(I.unit_elems, unsigned(e0 + np + nt + nv + l))));
            dih_out(l, d, ld3g(pos, unsigned(d.v0)), ld3g(pos, unsigned(d.v1)), ld3g(pos, unsigned(d.v2)),
                    ld3g(pos, unsigned(d.v3)));
        }
    }
    // entries and offsets into LDS: the held vectors, then the rest
#pragma unroll
    for (int k = 0; k < kStageRegs; ++k) {
        const int i = tid + k * kBlock;
        if (i < n_vec) Wv[i] = held[k];
    }
    for (int i = tid + kStageRegs * kBlock; i < n_vec; i += kBlock) Wv[i] = stage_src(i);
    __syncthreads();
    // ---- phase 2: one work item per owned block, a diagonal block's list
    // (pins, triangles and dihedrals of its vertex: the longest lists) split
    // over kDiagLanes adjacent lanes -- lane q sums entries c0 + q,
    // c0 + q + kDiagLanes, ... -- and combined by a fixed xor butterfly ----
    const int R = r1 - r0;
    const int n_items = kDiagLanes * R + (n_slots - R);
    for (int it = tid; it < n_items; it += kBlock) {
        const bool is_diag = it < kDiagLanes * R;
        const int t = is_diag ? it / kDiagLanes : it - (kDiagLanes - 1) * R;  // the unit's slot
        const unsigned lane_q = is_diag ? unsigned(it % kDiagLanes) : 0u;
        const unsigned stride = is_diag ? unsigned(kDiagLanes) : 1u;
        unsigned s;  // diagonal: row; upper: value slot
        if (is_diag) {
            s = unsigned(r0 + t);
        } else {
            int q = 0, v = t - R;
            while (v >= L.sl_end[q]) ++q;
            if (q > 0) v -= L.sl_end[q - 1];
            const int jj = v / L.sl_cnt[q], l = L.sl_l0[q] + v % L.sl_cnt[q];
            s = unsigned(L.sl_first[q] + jj * int(CHX_SLICE_ROWS) + l);
        }
        const unsigned c0 = Ol[t] + lane_q, c1 = Ol[t + 1];
        float h[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        f3 grad = mk(0.f, 0.f, 0.f);
        for (unsigned c = c0; c < c1; c += stride) {
            const unsigned w = Wl[c];
            const unsigned kind = w >> 14, a = (w >> 12) & 3u, b = (w >> 10) & 3u, l = w & CHX_CONTRIB_MAX_LOCAL;
            if (kind == CHX_CONTRIB_PIN) {
                const float4 r = Rp[l];
                h[0] += r.w;
                h[4] += r.w;
                h[8] += r.w;
                grad = add(grad, mk(r.x, r.y, r.z));
            } else if (kind == CHX_CONTRIB_TRI) {
                const float4* r = Rt + l;
                const float4 dm = r[0], q1 = r[nt], q2 = r[2 * nt], q3 = r[3 * nt];
                // Mu = (q1.x q1.y q1.z q1.w q2.x q2.y), Mv = (q2.z q2.w q3.x q3.y q3.z q3.w)
                const float dwu[3] = {-dm.x - dm.z, dm.x, dm.z}, dwv[3] = {-dm.y - dm.w, dm.y, dm.w};
                const float su = dwu[a] * dwu[b];
                const float sv = dwv[a] * dwv[b];
                const float m0 = su * q1.x + sv * q2.z, m1 = su * q1.y + sv * q2.w, m2 = su * q1.z + sv * q3.x;
                const float m3 = su * q1.w + sv * q3.y, m4 = su * q2.x + sv * q3.z, m5 = su * q2.y + sv * q3.w;
                h[0] += m0; h[1] += m1; h[2] += m2;
                h[3] += m1; h[4] += m3; h[5] += m4;
                h[6] += m2; h[7] += m4; h[8] += m5;
                if (is_diag && a == b) {
                    const float4 g0 = r[4 * nt], g1 = r[5 * nt];  // Gu = g0.xyz, Gv = (g0.w, g1.x, g1.y)
                    const float ua = dwu[a], va = dwv[a];
                    grad = add(grad, mk(ua * g0.x + va * g0.w, ua * g0.y + va * g1.x, ua * g0.z + va * g1.y));
                }
            } else if (kind == CHX_CONTRIB_STVK) {
                auto r = [&](int f) { return rf(Rv, nv, l, unsigned(4 + f)); };
                const float4 dm = Rv[l];
                const float dwu[3] = {-dm.x - dm.z, dm.x, dm.z}, dwv[3] = {-dm.y - dm.w, dm.y, dm.w};
                const float wua = dwu[a], wva = dwv[a], wub = dwu[b], wvb = dwv[b];
                const float su = wua * wub, sv = wva * wvb, suv = wua * wvb, svu = wva * wub;
                // Huu / Hvv symmetric (xx xy xz yy yz zz), Huv full; Hvu = Huv^T.
                const float uu[9] = {r(6), r(7), r(8), r(7), r(9), r(10), r(8), r(10), r(11)};
                const float vv[9] = {r(12), r(13), r(14), r(13), r(15), r(16), r(14), r(16), r(17)};
#pragma unroll
                for (int q = 0; q < 3; ++q)
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        h[3 * q + c] += su * uu[3 * q + c] + sv * vv[3 * q + c] + suv * r(18 + 3 * q + c) +
                                        svu * r(18 + 3 * c + q);
                if (is_diag && a == b)
                    grad = add(grad, mk(wua * r(0) + wva * r(3), wua * r(1) + wva * r(4), wua * r(2) + wva * r(5)));
            } else {
                const float4 qa = Rd[a * unsigned(nd) + l], qb = Rd[b * unsigned(nd) + l];
                const float k = rf(Rd, nd, l, 7u);
                const f3 ai = mk(qa.x, qa.y, qa.z);
                const f3 aj = mk(qb.x, qb.y, qb.z);
                h[0] += (ai.x * aj.x) * k; h[1] += (ai.x * aj.y) * k; h[2] += (ai.x * aj.z) * k;
                h[3] += (ai.y * aj.x) * k; h[4] += (ai.y * aj.y) * k; h[5] += (ai.y * aj.z) * k;
                h[6] += (ai.z * aj.x) * k; h[7] += (ai.z * aj.y) * k; h[8] += (ai.z * aj.z) * k;
                if (is_diag && a == b) grad = add(grad, mul(ai, rf(Rd, nd, l, 3u)));
            }
        }
        if (is_diag) {
            // the diagonal group's lanes are adjacent and all active
#pragma unroll
            for (int m = 1; m < kDiagLanes; m <<= 1) {
#pragma unroll
                for (int k = 0; k < 9; ++k) h[k] += __shfl_xor(h[k], m);
                grad = mk(grad.x + __shfl_xor(grad.x, m), grad.y + __shfl_xor(grad.y, m),
                          grad.z + __shfl_xor(grad.z, m));
            }
            if (lane_q != 0u) continue;
        }
        finish_slot(I, is_diag, s, h, grad);
    }
}

// BF16 mode: the diagonal of the inner matrix absorbs the rounding of its
// row's off-diagonal blocks, D_c = D + sum_j (V_ij - bf16(V_ij)), so the
// BF16 matrix maps block-constant vectors exactly as the FP32 one does
// (A_c (1 (x) c) = A (1 (x) c)).  The stiffness rows sum to zero, so without
// it the rounding of the off-diagonal blocks perturbs the smooth modes --
// the slowest ones for CG -- by O(eps * stiffness) against eigenvalues of
// O(stiffness / n^2); with it the perturbation acts on differences
// x_j - x_i only.  One workgroup per row tile, one thread per block row.
__global__ __launch_bounds__(kBlock) void compensate_kernel(chx_cloth_batch B) {
    int inst, rb, re;
    get_tile(B, blockIdx.x, inst, rb, re);
    const Inst I = inst_at(B, inst);
    for (unsigned row = unsigned(rb) + threadIdx.x; row < unsigned(re); row += blockDim.x) {
        const unsigned k0 = ldx(I.row_offsets, row), k1 = ldx(I.row_offsets, row + 1u);
        float d[9];
        ld6p(gp<const float>(I.diag), row, d);
        for (unsigned k = k0; k < k1; ++k) {
            // lower blocks: the partner's rounding, transposed
            const int bs = ldx(I.block_slots, k);
            const bool tr = bs < 0;
            const unsigned slot = unsigned(tr ? -1 - bs : bs);
#pragma unroll
            for (unsigned c = 0; c < 9; ++c) {
                const unsigned cs = tr ? (c % 3u) * 3u + c / 3u : c;
                const float v = ldx(gp<const float>(I.values), bp(cs, slot));
                d[c] += v - bf16_to_f32(f32_to_bf16(v));
            }
        }
#pragma unroll
        for (unsigned c = 0; c < 9; ++c) stx(I.diag_c, bp(c, row), d[c]);
    }
}

__global__ void reset_sync_kernel(unsigned* sync, int all) { reset_sync(sync, all != 0); }

// ---------------------------------------------------------------------------
// SpMV row loop: L lanes per block row, 64/L rows per wave and row group; R
// row groups per trip (R * 64 / L rows per wave), whose loads are all in
// flight together.  ro_at(r) / ci_at(k) give row offsets and column
// indices, src(j) the source vector entry j; fin(row, Ax, x_row) runs on the
// first lane of each row after the segmented reduction (x_row = src(row)),
// rows in ascending order per lane.
//
// Each lane takes U = 16/L of its row's off-diagonal blocks per pass (a pass
// covers 16 blocks of a row), loading them at clamped indices without
// branches so that all their loads are in flight together; the products are
// masked and added in ascending block order.  The diagonal block is applied
// by every lane of the row (same addresses, no branch) and kept on lane 0.
// Per lane the sum is diag + off[sub] + off[sub+L] + ..., then a segmented
// butterfly over the L lanes.
// ---------------------------------------------------------------------------
template <int L, int R = 1, class Mat, class RoAt, class CiAt, class Src, class Fin>
__device__ __forceinline__ void spmv_core(Mat mat, int rb, int re, RoAt ro_at, CiAt ci_at, Src src, Fin fin) {
    constexpr int RPW = kWave / L;
    constexpr int U = (16 / L) > 1 ? (16 / L) : 1;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int g = lane / L, sub_ = lane % L;
    const int step = nw * RPW;  // rows between a wave's row groups
    for (int base = rb + wid * RPW; base < re; base += step * R) {
        unsigned rowc[R], beg[R], end[R];
        bool valid[R];
        f3 xr[R], acc[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int row = base + q * step + g;
            valid[q] = row < re;
            rowc[q] = unsigned(valid[q] ? row : rb);
            beg[q] = ro_at(rowc[q]);
            end[q] = valid[q] ? ro_at(rowc[q] + 1u) : beg[q];
            xr[q] = src(rowc[q]);
            const f3 dprod = mat.diag(rowc[q], xr[q]);
            acc[q] = (valid[q] && sub_ == 0) ? dprod : mk(0.f, 0.f, 0.f);
        }
        for (unsigned p = 0;; p += unsigned(U * L)) {
            unsigned k0[R];
            bool live[R];
            bool any = false;
            unsigned kv = 0;  // an index inside some live row, for clamping
#pragma unroll
            for (int q = R - 1; q >= 0; --q) {
                k0[q] = beg[q] + unsigned(sub_) + p;
                live[q] = k0[q] < end[q];
                if (live[q]) {
                    any = true;
                    kv = k0[q];
                }
            }
            if (!any) break;
            f3 part[R][U];
#pragma unroll
            for (int q = 0; q < R; ++q)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const unsigned k = k0[q] + unsigned(u * L);
                    const unsigned kc = (k < end[q]) ? k : kv;
                    part[q][u] = mat.off(kc, src(ci_at(kc)));
                }
#pragma unroll
            for (int q = 0; q < R; ++q)
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (k0[q] + unsigned(u * L) < end[q]) acc[q] = add(acc[q], part[q][u]);
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
            acc[q].x = seg_sum<L>(acc[q].x);
            acc[q].y = seg_sum<L>(acc[q].y);
            acc[q].z = seg_sum<L>(acc[q].z);
        }
#pragma unroll
        for (int q = 0; q < R; ++q)
            if (valid[q] && sub_ == 0) fin(rowc[q], acc[q], xr[q]);
    }
}

// ---------------------------------------------------------------------------
// Sliced symmetric SpMV: one lane per block row, one wave per slice of 64
// rows (rb a multiple of 64).  slot_at(s) gives slice s's first entry slot,
// ent_at(slot) an entry (col << 6) | jj, uoff_at(s) slice s's first value
// slot, len_at(row) the row's block count, src(j) the source entry j;
// fin(row, Ax, x_row) on every valid row's lane.  Entry j of a row is the
// upper block jj of the row (col > row) or the transpose of upper block jj
// of row col (col < row).  A wave walks the slice's width in batches of KS
// entries whose loads are all in flight together (entries past the width
// clamp to the batch's first, padding entries load and are masked); the
// j-th entries of a wave are adjacent words, and so are, on a structured
// mesh, the value slots they name.  Per row the sum is diag + off[0] +
// off[1] + ... in column order.
// ---------------------------------------------------------------------------
// Compile-time tuning constants of the persistent PCG (LDS and register
// budget per workgroup): entries per batch of the sliced SpMV, staged
// column-index and source-row capacities, compacted-list instance limit.
#ifndef CHX_PCG_SLICE_BATCH
#define CHX_PCG_SLICE_BATCH 4
#endif
#ifndef CHX_PCG_CI_CAP
#define CHX_PCG_CI_CAP 4096
#endif
#ifndef CHX_PCG_WIN_CAP
#define CHX_PCG_WIN_CAP 2048
#endif
#ifndef CHX_PCG_COMPACT_MAX
#define CHX_PCG_COMPACT_MAX 1024
#endif
constexpr int kSliceBatch = CHX_PCG_SLICE_BATCH;
constexpr unsigned kSlice = CHX_SLICE_ROWS;
static_assert(CHX_SLICE_ROWS == kWave, "one lane per row of a slice");
static_assert(CHX_STAGE_ROWS % CHX_SLICE_ROWS == 0, "stages hold whole slices");

// No per-row prefetch: fin(row, Ax, x_row).
struct NoPre {
    __device__ int operator()(unsigned) const { return 0; }
};

template <bool E16, class Mat, class SlotAt, class EntAt, class UoffAt, class LenAt, class Src, class Fin,
          class Pre = NoPre>
__device__ __forceinline__ void spmv_sliced(Mat mat, int rb, int re, SlotAt slot_at, EntAt ent_at, UoffAt uoff_at,
                                            LenAt len_at, Src src, Fin fin, Pre pre = Pre{}) {
    constexpr int KS = kSliceBatch;
    const unsigned lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    for (int s0 = rb + wid * kWave; s0 < re; s0 += nw * kWave) {
        const unsigned sl = unsigned(s0) / kSlice;
        const unsigned base = slot_at(sl);
        const unsigned width = (slot_at(sl + 1u) - base) / kSlice;
        const unsigned ubase = uoff_at(sl) + lane;
        const unsigned row = unsigned(s0) + lane;
        const bool valid = row < unsigned(re);
        const unsigned rowc = valid ? row : unsigned(s0);
        const unsigned len = valid ? len_at(rowc) : 0u;
        const auto pv = pre(rowc);  // in flight with the row's block loads
        const f3 xr = src(rowc);
        f3 acc = mat.diag(rowc, xr);
        for (unsigned j0 = 0; j0 < width; j0 += KS) {
            f3 part[KS];
#pragma unroll
            for (int u = 0; u < KS; ++u) {
                const unsigned j = j0 + unsigned(u);
                const unsigned e = unsigned(ent_at(base + (j < width ? j : j0) * kSlice + lane));
                // 16-bit entries hold col - row + 2048 (row = this lane's row
                // of the slice, valid or not) and jj < 16.
                const unsigned col = E16 ? unsigned(s0) + lane + (e >> 4) - 2048u : e >> 6;
                const unsigned jj = E16 ? (e & 15u) : (e & 63u);
                const bool lower = col < rowc;
                const unsigned vs = lower ? uoff_at(col / kSlice) + (col & (kSlice - 1u)) + jj * kSlice
                                          : ubase + jj * kSlice;
                part[u] = mat.off_t(vs, src(col), lower);
            }
#pragma unroll
            for (int u = 0; u < KS; ++u)
                if (j0 + unsigned(u) < len) acc = add(acc, part[u]);
        }
        if (valid) {
            if constexpr (std::is_same_v<Pre, NoPre>) fin(rowc, acc, xr);
            else fin(rowc, acc, xr, pv);
        }
    }
}

// ---------------------------------------------------------------------------
// LDS of the persistent PCG kernel.  A stage is CHX_STAGE_ROWS rows of a
// tile: before its SpMV the workgroup loads the stage's row offsets, column
// indices and source-vector window (every row its blocks reference, from
// chx_instance::windows) into LDS with independent coalesced loads, so the
// SpMV's dependent chain row offset -> column -> gathered row runs in LDS and
// only the block values come from HBM.  The source rows live in a ring of
// kWinCap rows (row j at j mod kWinCap) that persists across the stages of
// a tile: consecutive stages' windows overlap by the matrix's bandwidth on
// each side, and a stage loads (and, in the single-barrier form, computes)
// only the rows its predecessor did not hold.  Stages beyond the LDS
// capacity (or without a window) run the same SpMV from global memory.
// ---------------------------------------------------------------------------
constexpr int kStage = CHX_STAGE_ROWS;
static_assert(kStage == kBlock, "one thread per stage row for the row offsets");
constexpr unsigned kCiCap = CHX_PCG_CI_CAP;   // column indices per staged stage
constexpr unsigned kWinCap = CHX_PCG_WIN_CAP;  // source rows held (ring capacity)
static_assert((kWinCap & (kWinCap - 1u)) == 0u, "the ring indexes rows modulo a power of two");
constexpr unsigned kWinMask = kWinCap - 1u;
constexpr int kChunk = 16;          // tiles whose instance scalars are summed together
static_assert(kChunk * 16 == kBlock, "16 lanes per tile in chunk_sums");
constexpr int kCompactMax = CHX_PCG_COMPACT_MAX;  // instances up to which the active tiles are compacted
#ifndef CHX_PCG_SUM_DEPTH
#define CHX_PCG_SUM_DEPTH 8
#endif
constexpr int kSumDepth = CHX_PCG_SUM_DEPTH;  // tiles per lane whose partials load together (instance sums)
static_assert(kCompactMax <= 4 * kBlock, "build_active scans 4 instances per thread");

// Shared by both SpMV forms.
struct PcgShared {
    double red[24];
    double sc[kChunk][8];
    // Active tile list (batches of up to kCompactMax instances): instance i
    // owns tiles [tb[i], tb[i] + tc[i]); pre[i] = active tiles before it.
    int tb[kCompactMax];
    int tc[kCompactMax];
    int pre[kCompactMax];
    int wsum[kBlock / kWave];
    int nact;
    // the instance descriptors of the chunk's tiles (instance-table batches):
    // one load round per chunk instead of one per tile
    chx_instance cdesc[kChunk];
    int ctile[kChunk];  // the tiles of the chunk at hand, -1 past the end
    int cinst[kChunk];  // their instances and row ranges (tile records)
    int crb[kChunk];
    int cre[kChunk];
};

// Stage arrays of the LDS-staged SpMV.
constexpr unsigned kWinSlices = kWinCap / CHX_SLICE_ROWS + 2;  // slices a window touches
struct PcgStageLds {
    int ro[kStage + 1];
    int ci[kCiCap];
    float4 win[kWinCap];  // source rows, one 16-byte slot each (one LDS access per row)
    int uo[kWinSlices];  // sliced layout: first value slot of the window's slices
};

// ST: LDS-staged SpMV (stage arrays, 64 KB per workgroup) or the global
// form (gathers through L1/L2, ~13 KB, more workgroups per CU).
template <bool ST>
struct PcgLds : PcgShared {
    PcgStageLds st;
};
template <>
struct PcgLds<false> : PcgShared {};

__device__ __forceinline__ f3 lds3(const float4* w, unsigned i) {
    const float4 v = w[i];
    return {v.x, v.y, v.z};
}

// Loads stage [sb, se) of instance I into S (true), or leaves it to the
// global path (false: no window, or over capacity).  srcg(j) is the source
// vector entry j.  The ring holds rows [rlo, rhi] on entry (rlo > rhi:
// none) and the stage's window [lo, rhi] on return; only the window's rows
// outside the old range are loaded.  Uniform across the workgroup.
// SL: the column indices come from the sliced layout (slots of the stage's
// slices) instead of the CSR order.
template <int EL, class SrcG>
__device__ __forceinline__ bool stage_load(const Inst& I, int sb, int se, SrcG srcg, PcgStageLds& S,
                                           unsigned& beg0, unsigned& lo, unsigned& rlo, unsigned& rhi) {
    constexpr bool SL = EL != 0, E16 = EL == 2;
    if (I.windows == nullptr) return false;
    const unsigned rb0 = unsigned(ldx(I.row_offsets, unsigned(sb)));
    const unsigned b0 = SL ? unsigned(ldx(I.slice_offsets, unsigned(sb) / kSlice)) : rb0;
    const unsigned nnz_s =
        (SL ? unsigned(ldx(I.slice_offsets, (unsigned(se) + kSlice - 1u) / kSlice))
            : unsigned(ldx(I.row_offsets, unsigned(se)))) - b0;
    const gp<const int> cols = SL ? I.slice_entries : I.col_indices;
    const unsigned st = unsigned(sb) / unsigned(kStage);
    const unsigned wlo = unsigned(ldx(I.windows, 2u * st)), whi = unsigned(ldx(I.windows, 2u * st + 1u));
    const unsigned W = whi - wlo + 1u;
    if (nnz_s > kCiCap || W > kWinCap) return false;
    // window rows not in the ring: [wlo, a_end) below it, [b_beg, whi] above
    const bool keep = rlo <= rhi && rlo <= whi && wlo <= rhi;
    const unsigned a_end = keep ? min(whi + 1u, max(rlo, wlo)) : whi + 1u;
    const unsigned b_beg = keep ? min(whi + 1u, max(wlo, rhi + 1u)) : whi + 1u;
    const unsigned nA = a_end - wlo;
    const unsigned nNew = nA + (whi + 1u - b_beg);
    const unsigned R = unsigned(se - sb);
    const unsigned tid = threadIdx.x;
    __syncthreads();  // the previous stage's readers are done
    // 16-bit entries move as 32-bit pairs (a stage's entries are whole
    // slices: an even count from an even index), 4096 per pass
    constexpr unsigned kEntPass = E16 ? 4096u : 2048u;
    const unsigned iters = max(max((nnz_s + kEntPass - 1u) / kEntPass, (nNew + 1023u) / 1024u), 1u);
    const unsigned nw16 = nnz_s / 2u;
    const gp<const unsigned> ent32 = (gp<const unsigned>)(gp<const uint16_t>(I.slice_entries16) + b0);
    for (unsigned it = 0; it < iters; ++it) {
        int cv[8];
        f3 wv[4];
        if (nnz_s > 0) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const unsigned k = it * 2048u + unsigned(u) * 256u + tid;
                if constexpr (E16) cv[u] = int(ldx(ent32, k < nw16 ? k : 0u));
                else cv[u] = ldx(cols, b0 + (k < nnz_s ? k : 0u));
            }
        }
        auto new_row = [&](unsigned w) { return w < nA ? wlo + w : b_beg + (w - nA); };
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const unsigned w = it * 1024u + unsigned(u) * 256u + tid;
            wv[u] = srcg(new_row(w < nNew ? w : 0u));
        }
        if (it == 0 && tid < R) S.ro[tid] = ldx(I.row_offsets, unsigned(sb) + tid) - int(rb0);
        if (nnz_s > 0) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const unsigned k = it * 2048u + unsigned(u) * 256u + tid;
                if constexpr (E16) {
                    if (k < nw16) reinterpret_cast<unsigned*>(S.ci)[k] = unsigned(cv[u]);
                } else {
                    if (k < nnz_s) S.ci[k] = cv[u];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const unsigned w = it * 1024u + unsigned(u) * 256u + tid;
            if (w < nNew) {
                const unsigned slot = new_row(w) & kWinMask;
                S.win[slot] = make_float4(wv[u].x, wv[u].y, wv[u].z, 0.0f);
            }
        }
    }
    if (tid == 0) S.ro[R] = ldx(I.row_offsets, unsigned(se)) - int(rb0);
    if (SL) {
        const unsigned s_lo = wlo / kSlice, s_n = whi / kSlice - s_lo + 1u;
        if (tid < s_n) S.uo[tid] = ldx(I.upper_offsets, s_lo + tid);
    }
    __syncthreads();
    beg0 = b0;
    lo = wlo;
    rlo = wlo;
    rhi = whi;
    return true;
}

// SpMV over rows [rb, re) of instance I (a tile): staged, one stage at a
// time, or straight from global memory; SL: sliced layout (cloth
// instances), else CSR with L lanes per row (caller-provided row-major
// blocks).
// pre(row) (sliced layout only): the row's loads for fin, issued before
// its block loads; fin then takes it as a fourth argument.
template <int L, bool ST, int EL, class Mat, class SrcG, class Fin, class Pre = NoPre>
__device__ __forceinline__ void tile_spmv(const Inst& I, Mat mat, int rb, int re, PcgLds<ST>& S, SrcG srcg,
                                          Fin fin, Pre pre = Pre{}) {
    constexpr bool SL = EL != 0, E16 = EL == 2;
    auto global_rows = [&](int b, int e) {
        if constexpr (SL) {
            spmv_sliced<E16>(
                mat, b, e, [&](unsigned sl) { return unsigned(ldx(I.slice_offsets, sl)); },
                [&](unsigned k) {
                    if constexpr (E16) return unsigned(ldx(I.slice_entries16, k));
                    else return unsigned(ldx(I.slice_entries, k));
                },
                [&](unsigned sl) { return unsigned(ldx(I.upper_offsets, sl)); },
                [&](unsigned r) { return unsigned(ldx(I.row_offsets, r + 1u) - ldx(I.row_offsets, r)); }, srcg,
                fin, pre);
        } else {
            spmv_core<L>(
                mat, b, e, [&](unsigned r) { return unsigned(ldx(I.row_offsets, r)); },
                [&](unsigned k) { return unsigned(ldx(I.col_indices, k)); }, srcg, fin);
        }
    };
    if constexpr (!ST) {
        global_rows(rb, re);
    } else {
        PcgStageLds& T = S.st;
        unsigned rlo = 1u, rhi = 0u;  // source rows the ring holds (none yet)
        for (int sb = rb; sb < re; sb += kStage) {
            const int se = min(re, sb + kStage);
            unsigned beg0 = 0, lo = 0;
            if (stage_load<EL>(I, sb, se, srcg, T, beg0, lo, rlo, rhi)) {
                if constexpr (SL) {
                    const unsigned s_lo = lo / kSlice;
                    spmv_sliced<E16>(
                        mat, sb, se, [&](unsigned sl) { return unsigned(ldx(I.slice_offsets, sl)); },
                        [&](unsigned k) {
                            if constexpr (E16) return unsigned(reinterpret_cast<const uint16_t*>(T.ci)[k - beg0]);
                            else return unsigned(T.ci[k - beg0]);
                        },
                        [&](unsigned sl) { return unsigned(T.uo[sl - s_lo]); },
                        [&](unsigned r) { return unsigned(T.ro[r - unsigned(sb) + 1u] - T.ro[r - unsigned(sb)]); },
                        [&](unsigned j) { return lds3(T.win, j & kWinMask); }, fin, pre);
                } else {
                    spmv_core<L>(
                        mat, sb, se, [&](unsigned r) { return unsigned(T.ro[r - unsigned(sb)]) + beg0; },
                        [&](unsigned k) { return unsigned(T.ci[k - beg0]); },
                        [&](unsigned j) { return lds3(T.win, j & kWinMask); }, fin);
                }
            } else {
                global_rows(sb, se);
            }
        }
    }
}

// Scalars of the instances of the chunk's tiles t = S.ctile[m] (m < kChunk):
// S.sc[m][0..3] = sums over the instance's tiles of partial fields f0..f3
// (and the tile mean of f4 into S.sc[m][5] when F4, a per-instance value
// every tile holds), S.sc[m][4] = 1.0 when the instance has
// finished its solve.  16 lanes per tile sum lane-strided over the
// instance's tiles, then a 16-lane butterfly, so an instance's sums are the
// same bits in every workgroup.
template <bool F4 = false>
__device__ __forceinline__ void chunk_sums(const chx_cloth_batch& B, gp<const double> part, int f0, int f1, int f2,
                                           int f3_, PcgShared& S, int f4 = 0) {
    __syncthreads();
    const int m = int(threadIdx.x) >> 4, l = int(threadIdx.x) & 15;
    const int t = S.ctile[m];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
    double done = 0.0, ntl = 1.0;
    if (t >= 0) {
        const int inst = S.cinst[m];
        int tb, te;
        if (B.inst != nullptr && B.n_instances <= kCompactMax) {  // tile ranges in LDS
            tb = S.tb[inst];
            te = tb + S.tc[inst];
        } else {
            const Inst I = inst_at(B, inst);
            tb = I.tile_begin;
            te = I.tile_end;
        }
        // Four of the lane's tiles per pass: their loads are issued together,
        // then added in ascending tile order (the same sums as one at a time).
        for (int u0 = tb + l; u0 < te; u0 += 64) {
            double v[4][5];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int u = u0 + 16 * q;
                if (u < te) {
                    const gp<const double> p = rec_at(part, unsigned(u), unsigned(CHX_TILE_PARTIALS));
                    v[q][0] = p[f0];
                    v[q][1] = p[f1];
                    v[q][2] = p[f2];
                    v[q][3] = p[f3_];
                    v[q][4] = F4 ? p[f4] : 0.0;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (u0 + 16 * q < te) {
                    s0 += v[q][0];
                    s1 += v[q][1];
                    s2 += v[q][2];
                    s3 += v[q][3];
                    if (F4) s4 += v[q][4];
                }
            }
        }
        if (l == 0) done = __hip_atomic_load(&B.inst_done[inst], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ? 1.0 : 0.0;
        ntl = double(te - tb);
    }
    s0 = seg_sum<16>(s0);
    s1 = seg_sum<16>(s1);
    s2 = seg_sum<16>(s2);
    s3 = seg_sum<16>(s3);
    if (F4) s4 = seg_sum<16>(s4);
    if (l == 0) {
        S.sc[m][0] = s0;
        S.sc[m][1] = s1;
        S.sc[m][2] = s2;
        S.sc[m][3] = s3;
        S.sc[m][4] = done;
        S.sc[m][5] = F4 ? s4 / ntl : 0.0;
    }
    __syncthreads();
}

// ---- work distribution ----------------------------------------------------
// A phase's tiles form a list of T entries: every tile (static), or the tiles
// of the instances still solving (compacted, in instance order).  With G a
// multiple of 8 the list is cut into 8 contiguous segments and segment x
// goes to the workgroups g = x (mod 8), which the dispatcher places on one
// XCD, so neighbouring tiles (overlapping source windows) share an L2.
__device__ __forceinline__ int list_entry(int T, int i) {
    const int G = int(gridDim.x), g = int(blockIdx.x);
    int a;
    if ((G & 7) == 0) {
        const int gx = G >> 3, seg = (T + 7) >> 3;
        const int pos = (g >> 3) + i * gx;
        if (pos >= seg) return -1;
        a = (g & 7) * seg + pos;
    } else {
        a = g + i * G;
    }
    return a < T ? a : -1;
}

// Tile of list entry a of the compacted list: the instance j with
// pre[j] <= a < pre[j] + (active tiles of j), found by binary search.
__device__ __forceinline__ int compact_tile(const PcgShared& S, int n, int a) {
    int lo = 0, hi = n;  // pre[lo] <= a, answer in [lo, hi)
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (S.pre[mid] <= a) lo = mid;
        else hi = mid;
    }
    return S.tb[lo] + (a - S.pre[lo]);
}

// Static per-launch table of the instances' tile ranges (compacted mode).
__device__ __forceinline__ void load_tile_ranges(const chx_cloth_batch& B, PcgShared& S) {
    const gp<const int> tab = glb(reinterpret_cast<const int*>(B.inst));
    constexpr unsigned words = unsigned(sizeof(chx_instance) / 4);
    constexpr unsigned ob = unsigned(offsetof(chx_instance, tile_begin) / 4);
    constexpr unsigned oe = unsigned(offsetof(chx_instance, tile_end) / 4);
    for (int i = threadIdx.x; i < B.n_instances; i += blockDim.x) {
        const gp<const int> d = rec_at(tab, unsigned(i), words);
        const int b = d[ob], e = d[oe];
        S.tb[i] = b;
        S.tc[i] = e - b;
    }
}

// Builds pre[] / nact from the done flags: instance i counts tc[i] tiles
// while it is solving, 0 once finished.  The flags must not change while
// any workgroup builds (every workgroup has to derive the same list).
// phase >= 0 (single-barrier solve): done flags hold the phase an instance
// stopped in, plus one; an instance stays on the list of the phase it stops
// in, so workgroups that read its flag before or after it is set in this
// phase derive the same list.
__device__ __forceinline__ void build_active(const chx_cloth_batch& B, PcgShared& S, int phase = -1) {
    const int n = B.n_instances;
    const int base = int(threadIdx.x) * 4;
    int c[4];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = base + q;
        c[q] = 0;
        if (i < n) {
            const int d = __hip_atomic_load(&B.inst_done[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (d == 0 || (phase >= 0 && d > phase)) c[q] = S.tc[i];
        }
        sum += c[q];
    }
    // inclusive scan of `sum` over the workgroup: wave scan, then wave totals
    const int lane = int(threadIdx.x) & (kWave - 1), wave = int(threadIdx.x) / kWave;
    int incl = sum;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int v = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += v;
    }
    __syncthreads();  // the previous list's readers are done
    if (lane == kWave - 1) S.wsum[wave] = incl;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; ++w) before += S.wsum[w];
    int run = before + incl - sum;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = base + q;
        if (i < n) S.pre[i] = run;
        run += c[q];
    }
    if (threadIdx.x == blockDim.x - 1) S.nact = before + incl;
    __syncthreads();
}

// Fills S.ctile with the tiles of this workgroup's chunk c of the phase
// list; returns false when the workgroup has no tile left (uniform).
__device__ __forceinline__ bool next_chunk(const chx_cloth_batch& B, PcgShared& S, bool compact, int c) {
    __syncthreads();  // the previous chunk's readers are done
    const int T = compact ? S.nact : B.n_tiles;
    if (threadIdx.x < kChunk) {
        const int a = list_entry(T, c * kChunk + int(threadIdx.x));
        const int t = (a < 0) ? -1 : (compact ? compact_tile(S, B.n_instances, a) : a);
        S.ctile[threadIdx.x] = t;
        // the chunk's tile records, loaded once for the phase's tile loop
        if (t >= 0) {
            int inst, rb, re;
            get_tile(B, t, inst, rb, re);
            S.cinst[threadIdx.x] = inst;
            S.crb[threadIdx.x] = rb;
            S.cre[threadIdx.x] = re;
        }
    }
    __syncthreads();
    if (B.inst != nullptr && S.ctile[0] >= 0) {
        constexpr int kWords = int(sizeof(chx_instance) / sizeof(unsigned long long));
        static_assert(sizeof(chx_instance) % sizeof(unsigned long long) == 0, "descriptor copy in 8-byte words");
        for (int k = int(threadIdx.x); k < kChunk * kWords; k += int(blockDim.x)) {
            const int m = k / kWords, w = k - m * kWords;
            if (S.ctile[m] >= 0) {
                const gp<const unsigned long long> src = (gp<const unsigned long long>)rec_at(
                    glb((const chx_instance*)B.inst), unsigned(S.cinst[m]), 1u);
                reinterpret_cast<unsigned long long*>(&S.cdesc[m])[w] = src[w];
            }
        }
        __syncthreads();
    }
    return S.ctile[0] >= 0;
}

// Descriptor of chunk tile m's instance: the LDS copy next_chunk made, or
// the single instance's.
__device__ __forceinline__ Inst chunk_inst(const chx_cloth_batch& B, const PcgShared& S, int m) {
    return B.inst != nullptr ? make_inst(&S.cdesc[m]) : make_inst(&B.single);
}

// Tile record m of the chunk at hand (uniform values in scalar registers).
__device__ __forceinline__ void chunk_tile(const PcgShared& S, int m, int& inst, int& rb, int& re) {
    inst = __builtin_amdgcn_readfirstlane(S.cinst[m]);
    rb = __builtin_amdgcn_readfirstlane(S.crb[m]);
    re = __builtin_amdgcn_readfirstlane(S.cre[m]);
}

// What an instance does in iteration slot k.  `true_res`: the residual at
// hand is b - A x with the FP32 matrix (always in FP32 mode; in BF16 mode
// after the set-up and after a refinement slot); `refines`: refinement slots
// spent so far.  A BF16 instance whose recursive residual meets the
// tolerance spends the slot on the true residual instead of stopping (outer
// iterative refinement), up to kMaxRefine times; after that the recursive
// residual decides, as in FP32 mode.  (With x stored in FP32 the true
// residual of a stiff system can sit above the tolerance for good; the
// recursive one keeps falling.)  Finishing is final: once an instance
// stops, its tiles no longer refresh the parity-indexed partials, so the
// test is never re-run on them (`finished`).
constexpr int kMaxRefine = 1;
enum Slot { STOP, ITERATE, REFINE };
__device__ __forceinline__ Slot instance_slot(const Inst& I, int k, double rr, double bb, bool true_res, int refines,
                                              bool finished) {
    if (finished || k >= I.max_iter) return STOP;
    if (I.tol > 0.0 && rr <= I.tol * I.tol * bb) return (true_res || refines >= kMaxRefine) ? STOP : REFINE;
    return ITERATE;
}

// ---------------------------------------------------------------------------
// Persistent block-Jacobi PCG (ChysX/src/solver/pcg_solver.cu:134-215):
//   r = b - A x, z = M^-1 r, p = z, rho = <r,z>
//   loop: q = A p; alpha = rho/<p,q>; x += alpha p; r -= alpha q;
//         z = M^-1 r; rho' = <r,z>; beta = rho'/rho; p = z + beta p
// The p update of iteration k is folded into the SpMV of iteration k: the
// SpMV gathers z_j + beta p_{k-1,j} (double-buffered p), which removes one
// grid barrier and one pass over p per iteration.
// ---------------------------------------------------------------------------
// Setup / refinement step of one tile: r = b - A32 x (true residual with
// the FP32 matrix), M^-1 = inverse(D) (setup only), z = M^-1 r, and p = z
// when pdst is given (setup); tile partials <r,z>, <r,r> (and <b,b> at
// setup).  M^-1 is stored grouped planar (bp).
// fin(row, A x, x_row) of a tile SpMV, with the row's contact cross-slot
// terms added to A x when a row-major solve carries a contact sidecar
// (pcg_solver.cu:159-160, 180-181: every A x of the iteration is (A + C) x).
// src(j) is the SpMV's source vector read from global memory.
template <bool SL, class Src, class Fin>
__device__ __forceinline__ auto plus_contact(const chx_cloth_batch& B, const Inst& I, Src src, Fin fin) {
    return [&B, &I, src, fin](unsigned row, f3 Ax, f3 xr) {
        if constexpr (!SL) {
            if (B.contact_rows != nullptr)
                Ax = add(Ax, contact_row(B.contact, B.contact_rows, B.contact_entries, row, unsigned(I.n_rows), 1.0f,
                                         src));
        }
        fin(row, Ax, xr);
    };
}

template <int L, bool ST, int EL, class MatR>
__device__ __forceinline__ void residual_tile(const chx_cloth_batch& B, const Inst& I, MatR mat, int rb, int re,
                                              bool setup, gp<float> pdst, PcgLds<ST>& S, double (&acc)[3]) {
    constexpr bool SL = EL != 0;
    const gp<const float> xv = I.x;
    auto xs = [&](unsigned j) { return ld3g(xv, j); };
    tile_spmv<L, ST, EL>(I, mat, rb, re, S, xs, plus_contact<SL>(B, I, xs, [&](unsigned row, f3 Ax, f3) {
        const f3 b = ld3g(I.rhs, row);
        const f3 r = sub(b, Ax);
        float mi[9];
        if (setup) {
            float d9[9];
            mat.diag_block(row, d9);
            inverse9(d9, mi);
            if constexpr (SL) {  // symmetric: the upper triangle
#pragma unroll
                for (unsigned q = 0; q < 6; q += 2)
                    st_pair(I.minv, bp6(q, row), mi[sym6_pos(q)], mi[sym6_pos(q + 1)]);
            } else {
#pragma unroll
                for (unsigned k = 0; k < 9; ++k) stx(I.minv, bp(k, row), mi[k]);
            }
        } else {
            if constexpr (SL) {
                ld6p(gp<const float>(I.minv), row, mi);
            } else {
#pragma unroll
                for (unsigned k = 0; k < 9; ++k) mi[k] = ldx(gp<const float>(I.minv), bp(k, row));
            }
        }
        const f3 z = mv9(mi, r);
        st3g(I.r, row, r);
        st3g(I.z, row, z);
        if (pdst != nullptr) st3g(pdst, row, z);
        acc[0] += dot64(r, z);
        acc[1] += dot64(r, r);
        acc[2] += dot64(b, b);
    }));
}

// Matrix variants of the PCG kernel: the view of the inner SpMV (A) and of
// the exact residual / preconditioner set-up (R).
//   V = 0  row-major 3x3 blocks (standalone chx_pcg_solve_bsr3)
//   V = 1  grouped planar FP32 blocks (cloth instances)
//   V = 2  grouped planar BF16 copies for the inner SpMV, FP32 for refinement
template <int V>
struct PcgMats;
template <>
struct PcgMats<0> {
    static __device__ __forceinline__ Mat32G a(const Inst& I) { return Mat32G{I.diag, I.values}; }
    static __device__ __forceinline__ Mat32G r(const Inst& I) { return a(I); }
};
template <>
struct PcgMats<1> {
    static __device__ __forceinline__ MatP32 a(const Inst& I) {
        return MatP32{I.diag, I.values};
    }
    static __device__ __forceinline__ MatP32 r(const Inst& I) { return a(I); }
};
template <>
struct PcgMats<2> {
    static __device__ __forceinline__ MatP16 a(const Inst& I) {
        return MatP16{I.diag_c, I.values16};
    }
    static __device__ __forceinline__ MatP32 r(const Inst& I) {
        return MatP32{I.diag, I.values};
    }
};

// V = 3 / 4: V = 1 / 2 with 16-bit slice entries (same matrix views).
template <>
struct PcgMats<3> : PcgMats<1> {};
template <>
struct PcgMats<4> : PcgMats<2> {};

// Optional solver trace (workgroup 0, thread 0): real-time stamp and
// active-instance count at barrier e.
__device__ __forceinline__ void trace_point(const chx_cloth_batch& B, int e, unsigned active) {
    if (B.trace != nullptr && blockIdx.x == 0 && threadIdx.x == 0 && e < B.trace_entries) {
        const gp<unsigned long long> t = glb(B.trace);
        t[2 * e] = __builtin_amdgcn_s_memrealtime();
        t[2 * e + 1] = active;
    }
}

// Finalize of every tile (cloth_simulator.cu:222-244): pos = x_n + dx,
// vel = dx / dt, times exp(-damping dt) when damping > 0.
__device__ __forceinline__ void finalize_tiles(const chx_cloth_batch& B, int G) {
    for (int t = blockIdx.x; t < B.n_tiles; t += G) {
        int inst, rb, re;
        get_tile(B, t, inst, rb, re);
        const Inst I = inst_at(B, inst);
        const float inv_dt = 1.0f / I.dt;
        const float decay = (I.damping > 0.0f) ? expf(-I.damping * I.dt) : 1.0f;
        for (unsigned row = unsigned(rb) + threadIdx.x; row < unsigned(re); row += blockDim.x) {
            const f3 d = ld3g(I.x, row);
            const f3 p = ld3g(I.pos, row);
            st3g(I.pos, row, add(p, d));
            f3 v = mul(d, inv_dt);
            if (I.damping > 0.0f) v = mul(v, decay);
            st3g(I.vel, row, v);
        }
    }
}

template <int L, int V, bool ST>
__global__ __launch_bounds__(kBlock) void pcg_kernel(chx_cloth_batch B, int finalize) {
    constexpr bool BF16 = V == 2 || V == 4;
    constexpr bool SL = V != 0;                       // cloth instances: sliced blocks
    constexpr int EL = V == 0 ? 0 : (V >= 3 ? 2 : 1);  // entry layout: CSR, sliced 32-bit, sliced 16-bit
    __shared__ PcgLds<ST> S;
    const int G = int(gridDim.x);
    unsigned* sync = B.sync;
    const Barrier bar = make_barrier(B);
    const gp<double> part = glb(B.partials);
    // BF16 mode: the inner iterations are CG on the BF16 matrix (compensated
    // FP32 diagonal) and test the recursively updated residual; when it meets
    // the tolerance, the next slot recomputes the true residual with the FP32
    // matrix.  The solve stops only on a true residual; otherwise CG restarts
    // from it (p = z): outer iterative refinement against the FP32 residual.

    // ---- setup: r = b - A x, M^-1, z = M^-1 r, p = z; rho, ||r||^2, ||b||^2 ----
    if (blockIdx.x == 0 && threadIdx.x == 0) st_relaxed(&sync[SYNC_ACTIVE], unsigned(B.n_instances));
    if (blockIdx.x == 0 && B.trace != nullptr)
        for (int e = threadIdx.x; e < 2 * B.trace_entries; e += blockDim.x) glb(B.trace)[e] = 0ull;
    for (int t = blockIdx.x; t < B.n_tiles; t += G) {
        int inst, rb, re;
        get_tile(B, t, inst, rb, re);
        const Inst I = inst_at(B, inst);
        if (threadIdx.x == 0 && t == I.tile_begin)
            __hip_atomic_store(&B.inst_done[inst], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        double acc[3] = {0.0, 0.0, 0.0};
        residual_tile<L, ST, EL>(B, I, PcgMats<V>::r(I), rb, re, true, I.p0, S, acc);
        block_sum(acc, S.red);
        if (threadIdx.x == 0) {
            const gp<double> p = rec_at(part, unsigned(t), unsigned(CHX_TILE_PARTIALS));
            p[P_RHO0] = acc[0];
            p[P_RR0] = acc[1];
            p[P_BB] = acc[2];
            p[P_TRUE0] = 1.0;
        }
    }
    // Compacted work lists for batches of up to kCompactMax instances: built
    // at the start of phase B, when no done flag changes (instances stop in
    // phase A only), and reused by the next phase A.  Phase A of iteration 0
    // runs the static list (every instance is solving).
    const bool compactable = B.inst != nullptr && B.n_instances <= kCompactMax;
    if (compactable) load_tile_ranges(B, S);
    bool compact = false;
    if (!grid_sync(bar, G)) return;
    trace_point(B, 0, unsigned(B.n_instances));

    for (int k = 0;; ++k) {
        const int cur = k & 1;
        const bool first = (k == 0);  // p_0 = z_0, written by the set-up
        // ---- phase A: p_k = z + beta p_{k-1}; q = A p_k; sigma = <p_k, q> ----
        // A refining instance rewrites its parity cur^1 partials during this
        // phase; its own tiles read them here but do not use them (no beta).
        for (int c = 0; next_chunk(B, S, compact, c); ++c) {
            chunk_sums<BF16>(B, part, P_RHO0 + cur, P_RHO0 + (cur ^ 1), P_RR0 + cur, P_BB, S, P_TRUE0 + cur);
            for (int m = 0; m < kChunk; ++m) {
                const int t = S.ctile[m];
                if (t < 0) break;
                int inst, rb, re;
                chunk_tile(S, m, inst, rb, re);
                const Inst I = chunk_inst(B, S, m);
                const double rho = S.sc[m][0], rho_prev = S.sc[m][1], rr = S.sc[m][2], bb = S.sc[m][3];
                const bool finished = S.sc[m][4] != 0.0;
                const int rstate = BF16 ? int(S.sc[m][5] + 0.5) : 1;
                const bool true_res = (rstate & 1) != 0;
                const Slot slot = instance_slot(I, k, rr, bb, true_res, rstate >> 1, finished);
                if (slot == STOP) {
                    if (threadIdx.x == 0 && t == I.tile_begin && !finished) {
                        __hip_atomic_store(&B.inst_done[inst], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        B.inst_iters[inst] = k;
                        B.inst_rr[inst] = rr;
                        B.inst_bb[inst] = bb;
                        if (B.inst_rho != nullptr) B.inst_rho[inst] = rho;
                        __hip_atomic_fetch_sub(&sync[SYNC_ACTIVE], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    continue;
                }
                const gp<float> pprev = cur ? I.p0 : I.p1;  // p_{k-1}
                const gp<float> pk = cur ? I.p1 : I.p0;     // p_k
                if (BF16 && slot == REFINE) {
                    // True residual r = b - A32 x, z = M^-1 r and their sums
                    // become the parity cur^1 partials, marked true.
                    double acc[3] = {0.0, 0.0, 0.0};
                    residual_tile<L, ST, EL>(B, I, PcgMats<V>::r(I), rb, re, false, gp<float>(nullptr), S, acc);
                    block_sum(acc, S.red);
                    if (threadIdx.x == 0) {
                        const gp<double> p = rec_at(part, unsigned(t), unsigned(CHX_TILE_PARTIALS));
                        p[P_RHO0 + (cur ^ 1)] = acc[0];
                        p[P_RR0 + (cur ^ 1)] = acc[1];
                        p[P_TRUE0 + (cur ^ 1)] = double(rstate + 3);  // one more refinement, true
                    }
                    continue;
                }
                const gp<const float> z = I.z;
                const gp<float> q = I.q;
                double acc[1] = {0.0};
                auto fin = [&](unsigned row, f3 Ap, f3 pr) {
                    if (!first) st3g(pk, row, pr);
                    st3g(q, row, Ap);
                    acc[0] += dot64(pr, Ap);
                };
                if (first) {
                    auto pj = [&](unsigned jj) { return ld3g(pk, jj); };
                    tile_spmv<L, ST, EL>(I, PcgMats<V>::a(I), rb, re, S, pj, plus_contact<SL>(B, I, pj, fin));
                } else if (BF16 && true_res) {
                    // restart after a refinement slot: p_k = z_k
                    tile_spmv<L, ST, EL>(I, PcgMats<V>::a(I), rb, re, S, [&](unsigned jj) { return ld3g(z, jj); }, fin);
                } else {
                    const float beta = guarded_div(rho, rho_prev);
                    auto pj = [&](unsigned jj) {
                        const f3 zj = ld3g(z, jj), pp = ld3g(pprev, jj);
                        return mk(fmaf(beta, pp.x, zj.x), fmaf(beta, pp.y, zj.y), fmaf(beta, pp.z, zj.z));
                    };
                    tile_spmv<L, ST, EL>(I, PcgMats<V>::a(I), rb, re, S, pj, plus_contact<SL>(B, I, pj, fin));
                }
                block_sum(acc, S.red);
                if (threadIdx.x == 0) rec_at(part, unsigned(t), unsigned(CHX_TILE_PARTIALS))[P_SIG] = acc[0];
            }
        }
        unsigned active = 0;
        if (!grid_sync(bar, G, cur, &active)) return;
        trace_point(B, 2 * k + 1, active);
        if (active == 0u) break;

        // ---- phase B: alpha; x += alpha p; r -= alpha q; z = M^-1 r; rho', rr' --
        const int nxt = cur ^ 1;
        if (compactable) {
            build_active(B, S);
            compact = true;
        }
        for (int c = 0; next_chunk(B, S, compact, c); ++c) {
            chunk_sums<BF16>(B, part, P_RHO0 + cur, P_SIG, P_RR0 + cur, P_BB, S, P_TRUE0 + cur);
            for (int m = 0; m < kChunk; ++m) {
                const int t = S.ctile[m];
                if (t < 0) break;
                int inst, rb, re;
                chunk_tile(S, m, inst, rb, re);
                const Inst I = chunk_inst(B, S, m);
                const double rho = S.sc[m][0], sigma = S.sc[m][1], rr = S.sc[m][2], bb = S.sc[m][3];
                const int rstate = BF16 ? int(S.sc[m][5] + 0.5) : 1;
                if (instance_slot(I, k, rr, bb, (rstate & 1) != 0, rstate >> 1, S.sc[m][4] != 0.0) != ITERATE)
                    continue;
                const float alpha = guarded_div(rho, sigma);
                const gp<const float> pk = cur ? I.p1 : I.p0;
                double acc[2] = {0.0, 0.0};
                for (unsigned row = unsigned(rb) + threadIdx.x; row < unsigned(re); row += blockDim.x) {
                    const f3 p = ld3g(pk, row), qq = ld3g(I.q, row);
                    f3 x = ld3g(I.x, row), r = ld3g(I.r, row);
                    x = mk(fmaf(alpha, p.x, x.x), fmaf(alpha, p.y, x.y), fmaf(alpha, p.z, x.z));
                    r = mk(fmaf(-alpha, qq.x, r.x), fmaf(-alpha, qq.y, r.y), fmaf(-alpha, qq.z, r.z));
                    const f3 zz = SL ? mv6p(gp<const float>(I.minv), row, r) : mv9p(gp<const float>(I.minv), row, r);
                    st3g(I.x, row, x);
                    st3g(I.r, row, r);
                    st3g(I.z, row, zz);
                    acc[0] += dot64(r, zz);
                    acc[1] += dot64(r, r);
                }
                block_sum(acc, S.red);
                if (threadIdx.x == 0) {
                    const gp<double> p = rec_at(part, unsigned(t), unsigned(CHX_TILE_PARTIALS));
                    p[P_RHO0 + nxt] = acc[0];
                    p[P_RR0 + nxt] = acc[1];
                    p[P_TRUE0 + nxt] = double(rstate & ~1);  // recursive residual
                }
            }
        }
        if (!grid_sync(bar, G)) return;
        trace_point(B, 2 * k + 2, active);
    }

    // ---- finalize: x_{n+1} = x_n + dx, v = dx/dt (cloth_simulator.cu:222-244) --
    if (finalize) finalize_tiles(B, G);
}

// ---------------------------------------------------------------------------
// Single-barrier form of the persistent PCG for the FP32 sliced cloth
// instances.  Same recurrence as pcg_kernel -- and as the reference
// (pcg_solver.cu:134-215) -- with each iteration's two global reductions
// merged into one phase:
//
//   phase k: with alpha_{k-1} = rho_{k-1} / sigma_{k-1} and
//            beta_{k-1} = rho_k / rho_{k-1},
//              r_k = r_{k-1} - alpha q_{k-1},  z_k = M^-1 r_k,
//              x_k = x_{k-1} + alpha p_{k-1},  p_k = z_k + beta p_{k-1},
//              q_k = A p_k,
//            tile partials rho_k = <r_k, z_k>, ||r_k||^2, sigma_k = <p_k, q_k>,
//            delta_k = <q_k, z_k>, gamma_k = <q_k, M^-1 q_k>      -> barrier
//
// rho_k is needed for beta_{k-1} before the reduction over the r_k that the
// same phase produces, so it comes from the exact identity (M^-1 symmetric)
//   <r_k, z_k> = rho_{k-1} - 2 alpha delta_{k-1} + alpha^2 gamma_{k-1}
// in FP64; every other scalar -- alpha_k's rho_k, the convergence test's
// ||r_k||^2 -- is the exact sum over the stored FP32 vectors of the phase
// before, so the identity is used for one step only and never accumulates.
// A tile forms z_k and p_k on the fly for the source window of its SpMV
// from r_{k-1}, q_{k-1}, p_{k-1} and M^-1 (double-buffered r and q), and
// updates r, x of its own rows in the SpMV's row finish.  The convergence
// test of iteration k runs at the start of phase k+1, on ||r_k||^2: the
// iteration count, the stopping iterate x_k and the final statistics are
// the two-barrier form's; the instance's last SpMV goes unused.  At the
// iteration cap the last phase updates r and x only.
// ---------------------------------------------------------------------------
// Tile partials of the single-barrier form, per parity: rho, ||r||^2,
// sigma, delta, gamma; then ||b||^2.
enum { Q_RHO0 = 0, Q_RR0 = 2, Q_SIG0 = 4, Q_DEL0 = 6, Q_GAM0 = 8, Q_BB = 10 };
static_assert(Q_BB < CHX_TILE_PARTIALS, "tile partials");

// Sums over the instance of each chunk tile of partial fields f[0..NF-1]
// into S.sc[m][0..NF-1] and the instance's raw done flag into S.sc[m][NF]
// (same lane-strided order and butterfly as chunk_sums: the same bits in
// every workgroup).
template <int NF>
__device__ __forceinline__ void chunk_sums_n(const chx_cloth_batch& B, gp<const double> part, const int (&f)[NF],
                                             PcgShared& S) {
    static_assert(NF + 1 <= 8, "S.sc holds 8 values per tile");
    __syncthreads();
    const int m = int(threadIdx.x) >> 4, l = int(threadIdx.x) & 15;
    const int t = S.ctile[m];
    double sum[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) sum[q] = 0.0;
    double done = 0.0;
    if (t >= 0) {
        const int inst = S.cinst[m];
        int tb, te;
        if (B.inst != nullptr && B.n_instances <= kCompactMax) {
            tb = S.tb[inst];
            te = tb + S.tc[inst];
        } else {
            const Inst I = inst_at(B, inst);
            tb = I.tile_begin;
            te = I.tile_end;
        }
        // kSumDepth of the lane's tiles per pass: their loads are issued
        // together (the partials were written on other XCDs, so each pass
        // waits for the Infinity Cache), then added in ascending tile order
        // -- the same sums as one tile at a time.
        for (int u0 = tb + l; u0 < te; u0 += 16 * kSumDepth) {
            double v[kSumDepth][NF];
#pragma unroll
            for (int h = 0; h < kSumDepth; ++h) {
                const int u = u0 + 16 * h;
                if (u < te) {
                    const gp<const double> p = rec_at(part, unsigned(u), unsigned(CHX_TILE_PARTIALS));
#pragma unroll
                    for (int q = 0; q < NF; ++q) v[h][q] = p[f[q]];
                }
            }
#pragma unroll
            for (int h = 0; h < kSumDepth; ++h)
                if (u0 + 16 * h < te) {
#pragma unroll
                    for (int q = 0; q < NF; ++q) sum[q] += v[h][q];
                }
        }
        if (l == 0) done = double(__hip_atomic_load(&B.inst_done[inst], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
#pragma unroll
    for (int q = 0; q < NF; ++q) sum[q] = seg_sum<16>(sum[q]);
    if (l == 0) {
#pragma unroll
        for (int q = 0; q < NF; ++q) S.sc[m][q] = sum[q];
        S.sc[m][NF] = done;
    }
    __syncthreads();
}

// Loads of a row's own update in the single-barrier iteration: r, q, p and x
// of the previous parity and the row's M^-1 block.
struct OwnRow {
    f3 r, q, p, x;
    float mi[9];
};

template <int EL, bool ST>
__global__ __launch_bounds__(kBlock) void pcg1_kernel(chx_cloth_batch B, int finalize) {
    __shared__ PcgLds<ST> S;
    const int G = int(gridDim.x);
    unsigned* sync = B.sync;
    const Barrier bar = make_barrier(B);
    const gp<double> part = glb(B.partials);

    // ---- setup: r_0 = b - A x, M^-1, z_0 = M^-1 r_0, p_0 = z_0; rho_0, ||r_0||^2, ||b||^2 ----
    if (blockIdx.x == 0 && threadIdx.x == 0) st_relaxed(&sync[SYNC_ACTIVE], unsigned(B.n_instances));
    if (blockIdx.x == 0 && B.trace != nullptr)
        for (int e = threadIdx.x; e < 2 * B.trace_entries; e += blockDim.x) glb(B.trace)[e] = 0ull;
    for (int t = blockIdx.x; t < B.n_tiles; t += G) {
        int inst, rb, re;
        get_tile(B, t, inst, rb, re);
        const Inst I = inst_at(B, inst);
        if (threadIdx.x == 0 && t == I.tile_begin)
            __hip_atomic_store(&B.inst_done[inst], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        double acc[3] = {0.0, 0.0, 0.0};
        residual_tile<4, ST, EL>(B, I, PcgMats<1>::r(I), rb, re, true, I.p0, S, acc);
        block_sum(acc, S.red);
        if (threadIdx.x == 0) {
            const gp<double> p = rec_at(part, unsigned(t), unsigned(CHX_TILE_PARTIALS));
            p[Q_RHO0 + 1] = acc[0];  // read by phase 0 as the previous parity
            p[Q_RR0 + 1] = acc[1];
            p[Q_BB] = acc[2];
        }
    }
    const bool compactable = B.inst != nullptr && B.n_instances <= kCompactMax;
    if (compactable) load_tile_ranges(B, S);
    if (!grid_sync(bar, G)) return;
    trace_point(B, 0, unsigned(B.n_instances));

    for (int k = 0;; ++k) {
        const int cur = k & 1, prv = cur ^ 1;
        if (compactable) build_active(B, S, k);
        const int fld[6] = {Q_RHO0 + prv, Q_SIG0 + prv, Q_DEL0 + prv, Q_GAM0 + prv, Q_RR0 + prv, Q_BB};
        for (int c = 0; next_chunk(B, S, compactable, c); ++c) {
            chunk_sums_n<6>(B, part, fld, S);
            for (int m = 0; m < kChunk; ++m) {
                const int t = S.ctile[m];
                if (t < 0) break;
                int inst, rb, re;
                chunk_tile(S, m, inst, rb, re);
                const Inst I = chunk_inst(B, S, m);
                const double rho = S.sc[m][0], sig = S.sc[m][1], del = S.sc[m][2], gam = S.sc[m][3];
                const double rr = S.sc[m][4], bb = S.sc[m][5];
                const int d = int(S.sc[m][6]);
                if (d != 0 && d <= k) continue;  // stopped in an earlier phase
                const int its = k > 0 ? k - 1 : 0;  // updates of x so far
                if (its >= I.max_iter || (I.tol > 0.0 && rr <= I.tol * I.tol * bb)) {
                    if (threadIdx.x == 0 && t == I.tile_begin && d == 0) {
                        __hip_atomic_store(&B.inst_done[inst], k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        B.inst_iters[inst] = its;
                        B.inst_rr[inst] = rr;
                        B.inst_bb[inst] = bb;
                        if (B.inst_rho != nullptr) B.inst_rho[inst] = rho;
                        __hip_atomic_fetch_sub(&sync[SYNC_ACTIVE], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    continue;
                }
                // buffers of parity q: r -> (r, z), q -> (q, q1), p -> (p0, p1)
                const gp<float> r_prev = prv ? I.z : I.r, r_cur = cur ? I.z : I.r;
                const gp<float> q_prev = prv ? I.q1 : I.q, q_cur = cur ? I.q1 : I.q;
                const gp<float> p_prev = prv ? I.p1 : I.p0, p_cur = cur ? I.p1 : I.p0;
                const gp<const float> minv = gp<const float>(I.minv);
                double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // rho, rr, sigma, delta, gamma
                if (k == 0) {
                    // q_0 = A p_0 (z_0 = p_0); rho_0, ||r_0||^2 carried over
                    auto pj = [&](unsigned j) { return ld3g(p_cur, j); };
                    tile_spmv<4, ST, EL>(I, PcgMats<1>::a(I), rb, re, S, pj, [&](unsigned row, f3 Ap, f3 pr) {
                        st3g(q_cur, row, Ap);
                        acc[2] += dot64(pr, Ap);
                        acc[3] += dot64(Ap, pr);
                        acc[4] += dot64(Ap, mv6p(minv, row, Ap));
                    });
                } else {
                    const float alpha = guarded_div(rho, sig);
                    const double ad = double(alpha);
                    // rho_k by the identity: its terms carry the FP32
                    // rounding of the stored vectors, a few units of 2^-24
                    // of their size.  Above 1e-5 of them (~100 such units)
                    // rho_k, and so beta, is resolved to about 1%; below,
                    // the direction restarts from z_k (beta = 0) instead of
                    // taking a rounding-noise beta.  A tighter cut (1e-3)
                    // restarts on ordinary fast-converging iterations and
                    // costs the Krylov acceleration: 13 iterations instead
                    // of 9 on a pin-driven substep of the 32x32 cloth.
                    const double t1 = 2.0 * ad * del, t2 = ad * ad * gam;
                    const double rho_k = rho - t1 + t2;
                    const bool resolved = rho_k > 1.0e-5 * (fabs(rho) + fabs(t1) + fabs(t2));
                    const float beta = resolved ? guarded_div(rho_k, rho) : 0.0f;
                    // r_k and z_k of row j from the previous parity
                    auto rz = [&](unsigned j, f3& rn, f3& zn) {
                        const f3 r = ld3g(r_prev, j), q = ld3g(q_prev, j);
                        rn = mk(fmaf(-alpha, q.x, r.x), fmaf(-alpha, q.y, r.y), fmaf(-alpha, q.z, r.z));
                        zn = mv6p(minv, j, rn);
                    };
                    // the own row's loads (issued with its block loads) and
                    // its update: r_k, x_k; returns z_k (the row finish
                    // reuses the row's M^-1 block for M^-1 q_k)
                    auto own_loads = [&](unsigned row) {
                        OwnRow o;
                        o.r = ld3g(r_prev, row);
                        o.q = ld3g(q_prev, row);
                        ld6p(minv, row, o.mi);
                        o.p = ld3g(p_prev, row);
                        o.x = ld3g(I.x, row);
                        return o;
                    };
                    auto own = [&](unsigned row, const OwnRow& o, f3& zn) {
                        const f3 rn =
                            mk(fmaf(-alpha, o.q.x, o.r.x), fmaf(-alpha, o.q.y, o.r.y), fmaf(-alpha, o.q.z, o.r.z));
                        zn = mv9(o.mi, rn);
                        const f3 x = mk(fmaf(alpha, o.p.x, o.x.x), fmaf(alpha, o.p.y, o.x.y), fmaf(alpha, o.p.z, o.x.z));
                        st3g(I.x, row, x);
                        st3g(r_cur, row, rn);
                        acc[0] += dot64(rn, zn);
                        acc[1] += dot64(rn, rn);
                    };
                    if (k < I.max_iter) {
                        auto pj = [&](unsigned j) {
                            f3 rn, zn;
                            rz(j, rn, zn);
                            const f3 pp = ld3g(p_prev, j);
                            return mk(fmaf(beta, pp.x, zn.x), fmaf(beta, pp.y, zn.y), fmaf(beta, pp.z, zn.z));
                        };
                        tile_spmv<4, ST, EL>(
                            I, PcgMats<1>::a(I), rb, re, S, pj,
                            [&](unsigned row, f3 Ap, f3 pr, const OwnRow& o) {
                                f3 zn;
                                own(row, o, zn);
                                st3g(p_cur, row, pr);
                                st3g(q_cur, row, Ap);
                                acc[2] += dot64(pr, Ap);
                                acc[3] += dot64(Ap, zn);
                                acc[4] += dot64(Ap, mv9(o.mi, Ap));
                            },
                            own_loads);
                    } else {
                        // at the cap: the last update only
                        for (unsigned row = unsigned(rb) + threadIdx.x; row < unsigned(re); row += blockDim.x) {
                            f3 zn;
                            own(row, own_loads(row), zn);
                        }
                    }
                }
                block_sum(acc, S.red);
                if (threadIdx.x == 0) {
                    const gp<double> p = rec_at(part, unsigned(t), unsigned(CHX_TILE_PARTIALS));
                    if (k == 0) {
                        p[Q_RHO0 + cur] = p[Q_RHO0 + prv];
                        p[Q_RR0 + cur] = p[Q_RR0 + prv];
                    } else {
                        p[Q_RHO0 + cur] = acc[0];
                        p[Q_RR0 + cur] = acc[1];
                    }
                    p[Q_SIG0 + cur] = acc[2];
                    p[Q_DEL0 + cur] = acc[3];
                    p[Q_GAM0 + cur] = acc[4];
                }
            }
        }
        unsigned active = 0;
        if (!grid_sync(bar, G, cur, &active)) return;
        trace_point(B, k + 1, active);
        if (active == 0u) break;
    }

    // ---- finalize: x_{n+1} = x_n + dx, v = dx/dt (cloth_simulator.cu:222-244) --
    if (finalize) finalize_tiles(B, G);
}

// ---------------------------------------------------------------------------
// Standalone SpMV: y = alpha A x + beta y (block_csr.cu:54-83).
// ---------------------------------------------------------------------------
template <int L>
__global__ __launch_bounds__(kBlock) void spmv_kernel(int n, const float* __restrict__ diag, const int* __restrict__ ro,
                                                      const int* __restrict__ ci, const float* __restrict__ val,
                                                      const float* __restrict__ x, float* __restrict__ y, float alpha,
                                                      float beta) {
    constexpr int RPB = (kBlock / kWave) * (kWave / L);  // rows per block
    const int rb = blockIdx.x * RPB;
    const int re = min(n, rb + RPB);
    const gp<const float> xg = glb(x);
    const gp<float> yg = glb(y);
    const gp<const int> rog = glb(ro), cig = glb(ci);
    spmv_core<L>(
        Mat32G{glb(diag), glb(val)}, rb, re, [&](unsigned r) { return unsigned(ldx(rog, r)); },
        [&](unsigned k) { return unsigned(ldx(cig, k)); }, [&](unsigned j) { return ld3g(xg, j); },
        [&](unsigned row, f3 Ax, f3) {
            if (beta == 0.0f) st3g(yg, row, mul(Ax, alpha));
            else st3g(yg, row, add(mul(ld3g(yg, row), beta), mul(Ax, alpha)));
        });
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
using PcgFn = void (*)(chx_cloth_batch, int);

// Bytes of a standalone solve's workspace for n block rows (the layout
// carve_workspace cuts): r, z, p0, p1, q, M^-1 (whole groups), tile
// partials, per-instance outputs and the barrier area, 256-byte aligned.
constexpr size_t al256(size_t v) { return (v + 255) & ~size_t(255); }
constexpr size_t pcg_workspace_bytes(size_t n) {
    return 5 * al256(12 * n) + al256(36 * ((n + CHX_BLOCK_GROUP - 1) / CHX_BLOCK_GROUP * CHX_BLOCK_GROUP)) +
           al256(8 * CHX_TILE_PARTIALS * ((n + CHX_TILE_ROWS - 1) / CHX_TILE_ROWS)) + al256(sizeof(int)) +
           3 * al256(sizeof(double)) + al256(sizeof(int)) + al256(SYNC_WORDS * sizeof(unsigned));
}

template <bool ST>
PcgFn pcg_fn_s(int L, int v) {
    // single-barrier FP32 sliced forms: 16-bit / 32-bit entries
    if (v == 5) return pcg1_kernel<2, ST>;
    if (v == 6) return pcg1_kernel<1, ST>;
    if (v == 0) {
        switch (L) {
            case 4: return pcg_kernel<4, 0, ST>;
            case 8: return pcg_kernel<8, 0, ST>;
            default: return pcg_kernel<16, 0, ST>;
        }
    }
    // Sliced layout (cloth instances): one lane per row, lanes_per_row
    // does not apply.
    switch (v) {
        case 1: return pcg_kernel<4, 1, ST>;
        case 2: return pcg_kernel<4, 2, ST>;
        case 3: return pcg_kernel<4, 3, ST>;
        default: return pcg_kernel<4, 4, ST>;
    }
}

// v: matrix variant (0 row-major CSR, 1 sliced FP32, 2 sliced BF16 + FP32;
// 3 / 4: 1 / 2 with 16-bit slice entries; 5 / 6: the single-barrier FP32
// form with 16-bit / 32-bit entries); staged: LDS-staged SpMV or the global
// form.
PcgFn pcg_fn(int L, int v = 1, bool staged = true) {
    if (L != 4 && L != 8 && L != 16)
        throw Error(CHX_ERR_INVALID_ARGUMENT, "chysx: lanes_per_row must be 4, 8 or 16");
    return staged ? pcg_fn_s<true>(L, v) : pcg_fn_s<false>(L, v);
}

struct DeviceLimits {
    int device = -1;
    int cus = 0;
    int max_blocks[14][17] = {{0}};  // [2 * v + staged][L]
    int api_blocks[14][17] = {{0}};
};

std::mutex g_lim_mu;
DeviceLimits g_lim[16];

// Workgroups of the persistent PCG kernel that may be in flight at once.
// The grid is capped at kBlocksPerCu per compute unit, well under the
// kernel's register-limited occupancy (7 four-wave workgroups per CU on
// gfx950), so every workgroup of the cooperative launch is resident.
constexpr int kBlocksPerCu = 4;

int max_coresident(int L, int v = 1, bool staged = true) {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_lim_mu);
    DeviceLimits& d = g_lim[dev & 15];
    if (d.device != dev) {
        d = DeviceLimits{};
        d.device = dev;
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        d.cus = prop.multiProcessorCount;
    }
    const int m = 2 * v + (staged ? 1 : 0);
    if (d.max_blocks[m][L] == 0) {
        int per_cu = 0;
        hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                      &per_cu, reinterpret_cast<const void*>(pcg_fn(L, v, staged)), kBlock, 0),
                  "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        d.api_blocks[m][L] = per_cu;
        if (per_cu < 1) per_cu = 1;
        d.max_blocks[m][L] = std::min(per_cu, kBlocksPerCu) * d.cus;
    }
    return d.max_blocks[m][L];
}

// row_major: the caller's blocks are row-major 3x3 (standalone solve);
// cloth instances store planar blocks.
// The PCG kernel variant of a launch (pcg_fn's v).
int pcg_variant(const chx_cloth_batch& B, bool row_major) {
    if (row_major) return 0;
    if (B.bf16_blocks == 0 && B.pcg_form == 0) return B.entries16 != 0 ? 5 : 6;
    return (B.bf16_blocks != 0 ? 2 : 1) + (B.entries16 != 0 ? 2 : 0);
}

void launch_pcg(const chx_cloth_batch& B, int finalize, hipStream_t stream, bool row_major = false) {
    if (B.n_tiles <= 0) return;
    const int L = B.lanes_per_row;
    const int v = pcg_variant(B, row_major);
    const bool staged = B.no_lds_staging == 0;
    PcgFn fn = pcg_fn(L, v, staged);
    int grid = max_coresident(L, v, staged);
    if (B.grid_blocks > 0) grid = std::min(grid, B.grid_blocks);
    grid = std::min(grid, B.n_tiles);
    chx_cloth_batch arg = B;
    int fin = finalize;
    void* args[] = {&arg, &fin};
    hip_check(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(fn), dim3(grid), dim3(kBlock), args, 0, stream),
              "hipLaunchCooperativeKernel(pcg)");
}

void validate_batch(const chx_cloth_batch* b) {
    require(b != nullptr, "chysx: null batch descriptor");
    require(b->n_instances >= 0 && b->n_tiles >= 0 && b->n_elem_items >= 0 && b->n_unit_items >= 0,
            "chysx: negative size in batch descriptor");
    require(b->lanes_per_row == 4 || b->lanes_per_row == 8 || b->lanes_per_row == 16,
            "chysx: lanes_per_row must be 4, 8 or 16");
    require(b->sync != nullptr, "chysx: batch descriptor without a sync area");
    if (b->inst == nullptr) {
        require(b->n_instances == 1, "chysx: a batch without an instance table holds exactly one instance");
        require(b->elem_items == nullptr && b->unit_items == nullptr,
                "chysx: single-instance element and unit lists are implied");
        const chx_instance& I = b->single;
        if (b->tiles == nullptr) {
            const int tr = b->tile_rows > 0 ? b->tile_rows : CHX_TILE_ROWS;
            require(b->tile_rows >= 0 && tr % CHX_STAGE_ROWS == 0 &&
                        b->n_tiles == (I.n_rows + tr - 1) / tr && I.tile_begin == 0 && I.tile_end == b->n_tiles,
                    "chysx: implied tile list does not match n_tiles / tile_rows");
        }
        require(I.n_rows >= 0 && I.nnz >= 0 && I.n_pins >= 0 && I.n_tris >= 0 && I.n_dihs >= 0 && I.n_stvks >= 0,
                "chysx: negative size in instance");
        if (I.n_rows > 0)
            require(I.row_offsets && I.diag && I.rhs && I.x && I.r && I.z && I.p0 && I.p1 && I.q && I.minv &&
                        (I.q1 || b->bf16_blocks || b->pcg_form != 0) &&
                        I.slice_offsets && I.slice_entries && I.upper_offsets && (I.block_slots || I.nnz == 0) &&
                        I.units && I.unit_elems && I.contrib_offsets && I.n_units > 0 &&
                        (I.slice_entries16 || b->entries16 == 0),
                    "chysx: null array in instance");
        require(I.n_padded >= I.nnz && (I.n_padded % CHX_SLICE_ROWS) == 0 && (I.n_upper % CHX_SLICE_ROWS) == 0 &&
                    2 * size_t(I.n_upper) >= size_t(I.nnz) && size_t(I.n_upper) * 36u < (size_t(1) << 32) &&
                    I.n_rows < (1 << 26),
                "chysx: sliced block layout smaller than the pattern, over 4 GiB or over 2^26 rows");
        if (b->bf16_blocks && I.n_rows > 0)
            require(I.diag_c && (I.values16 || I.nnz == 0), "chysx: BF16 mode needs the BF16 block copies");
        require(I.n_planes >= 0 && I.n_boxes >= 0 && (I.n_planes == 0 || I.planes) && (I.n_boxes == 0 || I.boxes),
                "chysx: static contact primitives missing");
    } else {
        require(b->tiles != nullptr || b->n_tiles == 0, "chysx: batch without a tile list");
        require(b->elem_items != nullptr || b->n_elem_items == 0, "chysx: batch without an element list");
        require(b->unit_items != nullptr || b->n_unit_items == 0, "chysx: batch without a unit list");
    }
    if (b->n_tiles > 0) require(b->partials && b->inst_iters && b->inst_rr && b->inst_bb && b->inst_done,
                                "chysx: batch without solver outputs");
}

int item_count(int n) { return (n + CHX_ITEM_SIZE - 1) / CHX_ITEM_SIZE; }

// Launch counts: the single instance's lists are implied by its sizes.
int elem_items(const chx_cloth_batch& B) {
    if (B.inst != nullptr) return B.n_elem_items;
    return item_count(B.single.n_pins) + item_count(B.single.n_tris) + item_count(B.single.n_dihs) +
           item_count(B.single.n_stvks);
}
int unit_items(const chx_cloth_batch& B) {
    if (B.inst != nullptr) return B.n_unit_items;
    return B.single.n_rows > 0 ? B.single.n_units : 0;
}

// The assembly kernel (it also resets the solver's control words for the
// PCG launch that follows), then in BF16 mode the compensated diagonal.
void assemble(const chx_cloth_batch& B, hipStream_t stream) {
    const int nu = unit_items(B);
    if (nu > 0) {
        assemble_kernel<<<nu, kBlock, 0, stream>>>(B);
        hip_check(hipGetLastError(), "assemble_kernel launch");
        if (B.bf16_blocks && B.n_tiles > 0) {
            compensate_kernel<<<B.n_tiles, kBlock, 0, stream>>>(B);
            hip_check(hipGetLastError(), "compensate_kernel launch");
        }
    } else {
        reset_sync_kernel<<<1, kBlock, 0, stream>>>(B.sync, 0);
        hip_check(hipGetLastError(), "reset_sync_kernel launch");
    }
}

}  // namespace
}  // namespace chx

using namespace chx;

extern "C" {

int chx_pcg_grid_limit(int lanes_per_row, int bf16_blocks, int no_lds_staging, int* max_blocks) {
    return guarded([&] {
        require(max_blocks != nullptr, "chx_pcg_grid_limit: null output");
        pcg_fn(lanes_per_row, 1);  // validates L
        *max_blocks = max_coresident(lanes_per_row, bf16_blocks ? 2 : 1, no_lds_staging == 0);
    });
}

int chx_device_info(int* cu_count, int* max_pcg_blocks, int* occupancy_blocks_per_cu) {
    return guarded([&] {
        const int mb = max_coresident(8);
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        if (cu_count) *cu_count = g_lim[dev & 15].cus;
        if (max_pcg_blocks) *max_pcg_blocks = mb;
        if (occupancy_blocks_per_cu) *occupancy_blocks_per_cu = g_lim[dev & 15].api_blocks[3][8];
    });
}

int chx_spmv_bsr3(int n_rows, const float* diag, const int* row_offsets, const int* col_indices, const float* values,
                  const float* x, float* y, float alpha, float beta, int lanes_per_row, uintptr_t stream) {
    return guarded([&] {
        require(n_rows >= 0, "spmv: negative row count");
        if (n_rows == 0) return;
        require(size_t(n_rows) * 36u < (size_t(1) << 32), "spmv: block arrays must stay below 4 GiB");
        require(diag && row_offsets && x && y, "spmv: null pointer");
        const hipStream_t s = as_stream(stream);
        auto go = [&](auto kern, int L) {
            const int rpb = (kBlock / kWave) * (kWave / L);
            kern<<<blocks_for(n_rows, rpb), kBlock, 0, s>>>(n_rows, diag, row_offsets, col_indices, values, x, y,
                                                             alpha, beta);
        };
        switch (lanes_per_row) {
            case 4: go(spmv_kernel<4>, 4); break;
            case 8: go(spmv_kernel<8>, 8); break;
            case 16: go(spmv_kernel<16>, 16); break;
            default: throw Error(CHX_ERR_INVALID_ARGUMENT, "spmv: lanes_per_row must be 4, 8 or 16");
        }
        hip_check(hipGetLastError(), "spmv_kernel launch");
    });
}

int chx_pcg_workspace_size(int n_rows, size_t* bytes) {
    return guarded([&] {
        require(n_rows >= 0 && bytes, "chx_pcg_workspace_size: bad argument");
        *bytes = pcg_workspace_bytes(size_t(n_rows));
    });
}

}  // extern "C"

namespace chx {
namespace {

// The standalone solve's descriptor over its workspace (the layout
// chx_pcg_workspace_size sizes): r, z, p0, p1, q, M^-1, tile partials, the
// per-instance outputs and the barrier area, each 256-byte aligned.
chx_cloth_batch carve_workspace(int n_rows, void* workspace) {
    const size_t n = size_t(n_rows);
    const int tiles = static_cast<int>((n + CHX_TILE_ROWS - 1) / CHX_TILE_ROWS);
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    char* w = static_cast<char*>(workspace);
    auto take = [&](size_t bytes) {
        char* p = w;
        w += al(bytes);
        return p;
    };
    chx_cloth_batch B{};
    B.n_instances = 1;
    B.n_tiles = tiles;
    chx_instance& I = B.single;
    I.n_rows = n_rows;
    I.r = reinterpret_cast<float*>(take(12 * n));
    I.z = reinterpret_cast<float*>(take(12 * n));
    I.p0 = reinterpret_cast<float*>(take(12 * n));
    I.p1 = reinterpret_cast<float*>(take(12 * n));
    I.q = reinterpret_cast<float*>(take(12 * n));
    I.minv = reinterpret_cast<float*>(take(36 * ((n + CHX_BLOCK_GROUP - 1) / CHX_BLOCK_GROUP * CHX_BLOCK_GROUP)));
    I.tile_begin = 0;
    I.tile_end = tiles;
    B.partials = reinterpret_cast<double*>(take(8 * CHX_TILE_PARTIALS * size_t(tiles)));
    B.inst_iters = reinterpret_cast<int*>(take(sizeof(int)));
    B.inst_rr = reinterpret_cast<double*>(take(sizeof(double)));
    B.inst_bb = reinterpret_cast<double*>(take(sizeof(double)));
    B.inst_done = reinterpret_cast<int*>(take(sizeof(int)));
    B.sync = reinterpret_cast<unsigned*>(take(SYNC_WORDS * sizeof(unsigned)));
    B.inst_rho = reinterpret_cast<double*>(take(sizeof(double)));
    return B;
}

// Synchronous read of a standalone solve's outputs; a timed-out solve is
// reported (and its status cleared) here.
void read_workspace_stats(const chx_cloth_batch& B, int* iterations, double* rr, double* bb, double* rho,
                          hipStream_t s) {
    int it = 0;
    double r2 = 0.0, b2 = 0.0, rz = 0.0;
    unsigned status = 0;
    hip_check(hipMemcpyAsync(&it, B.inst_iters, sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(iters)");
    hip_check(hipMemcpyAsync(&r2, B.inst_rr, sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(rr)");
    hip_check(hipMemcpyAsync(&b2, B.inst_bb, sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(bb)");
    hip_check(hipMemcpyAsync(&rz, B.inst_rho, sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(rho)");
    hip_check(hipMemcpyAsync(&status, B.sync + SYNC_STATUS, sizeof(unsigned), hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(status)");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize(pcg)");
    if (status != 0u) {
        reset_sync_kernel<<<1, kBlock, 0, s>>>(B.sync, 1);
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize(reset)");
        throw Error(CHX_ERR_DEVICE, "chysx::solver::PCGSolver: grid barrier timed out");
    }
    if (iterations) *iterations = it;
    if (rr) *rr = r2;
    if (bb) *bb = b2;
    if (rho) *rho = rz;
}

// Scratch of chx_pcg_solve_bsr3_scratch: the workspace of one standalone
// solve of at most CHX_PCG_SCRATCH_MAX_ROWS block rows as a file-scope
// device array, so that entry point allocates nothing.  Each device holds
// its own copy (6.0 MiB) until the process ends; all calls on one device
// share it, so calls that overlap on one device race.  Host code must not
// pass the array's name to a kernel: it names a host-side stand-in.  The
// device address comes from hipGetSymbolAddress, looked up once per device
// with that device current, and only that address is used.
constexpr size_t kScratchBytes = pcg_workspace_bytes(CHX_PCG_SCRATCH_MAX_ROWS);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};

void* scratch_address() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    require(dev >= 0 && dev < kScratchDevices, "chx_pcg_solve_bsr3_scratch: device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        require(p != nullptr, "chx_pcg_solve_bsr3_scratch: no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

// The standalone row-major PCG (pcg_solver.cu:134-215), with the contact
// sidecar's cross-slot terms in every A x when `op` is active
// (ContactSpMVOp::active, contact_spmv.h:81-83).
void solve_bsr3(int n_rows, const float* diag, const int* row_offsets, const int* col_indices, const float* values,
                const float* b, float* x, int max_iterations, double tol, int lanes_per_row, void* workspace,
                size_t workspace_bytes, const chx_contact_op* op, void* contact_workspace,
                size_t contact_workspace_bytes, int* iterations, double* rr, double* bb, uintptr_t stream) {
    {
        require(n_rows >= 0, "pcg: negative row count");
        require(max_iterations >= 0, "pcg: negative iteration cap");
        if (n_rows == 0) {
            if (iterations) *iterations = 0;
            if (rr) *rr = 0.0;
            if (bb) *bb = 0.0;
            return;
        }
        require(diag && row_offsets && b && x && workspace, "pcg: null pointer");
        require(size_t(n_rows) * 36u < (size_t(1) << 32), "pcg: block arrays must stay below 4 GiB");
        size_t need = 0;
        if (chx_pcg_workspace_size(n_rows, &need) != CHX_OK) throw Error(CHX_ERR_RUNTIME, chx_last_error_message());
        require(workspace_bytes >= need, "pcg: workspace too small (see chx_pcg_workspace_size)");
        chx_cloth_batch B = carve_workspace(n_rows, workspace);
        B.lanes_per_row = lanes_per_row;
        chx_instance& I = B.single;
        I.row_offsets = row_offsets;
        I.col_indices = col_indices;
        I.diag = const_cast<float*>(diag);
        I.values = const_cast<float*>(values);
        I.rhs = const_cast<float*>(b);
        I.x = x;
        I.dt = 1.0f;
        I.max_iter = max_iterations;
        I.tol = tol;
        const hipStream_t s = as_stream(stream);
        pcg_fn(lanes_per_row, 0);  // validates L
        if (op != nullptr && op->max_contacts > 0 && op->stiffness > 0.0f && op->pairs != nullptr &&
            op->count != nullptr) {
            const ContactIndex ix = build_contact_index(n_rows, *op, contact_workspace, contact_workspace_bytes, s);
            B.contact = *op;
            B.contact_rows = ix.rows;
            B.contact_entries = ix.entries;
        }
        // The workspace needs no initialisation: every control word is
        // cleared, status included (an earlier solve's abort is reported by
        // its own synchronous read or chx_pcg_last_stats before this call).
        reset_sync_kernel<<<1, kBlock, 0, s>>>(B.sync, 1);
        hip_check(hipGetLastError(), "reset_sync_kernel launch");
        launch_pcg(B, 0, s, true);
        if (iterations || rr || bb) read_workspace_stats(B, iterations, rr, bb, nullptr, s);
    }
}

}  // namespace
}  // namespace chx

extern "C" {

int chx_pcg_solve_bsr3_scratch(int n_rows, const float* diag, const int* row_offsets, const int* col_indices,
                               const float* values, const float* b, float* x, int max_iterations, double tol,
                               int lanes_per_row, int* iterations, double* rr, double* bb, uintptr_t stream) {
    return guarded([&] {
        require(n_rows >= 0, "chx_pcg_solve_bsr3_scratch: negative row count");
        require(n_rows <= CHX_PCG_SCRATCH_MAX_ROWS,
                "chx_pcg_solve_bsr3_scratch: n_rows exceeds CHX_PCG_SCRATCH_MAX_ROWS (65536); use "
                "chx_pcg_solve_bsr3 with a caller workspace");
        void* ws = nullptr;
        if (n_rows > 0) ws = scratch_address();
        solve_bsr3(n_rows, diag, row_offsets, col_indices, values, b, x, max_iterations, tol, lanes_per_row, ws,
                   kScratchBytes, nullptr, nullptr, 0, iterations, rr, bb, stream);
    });
}

int chx_pcg_solve_bsr3(int n_rows, const float* diag, const int* row_offsets, const int* col_indices,
                       const float* values, const float* b, float* x, int max_iterations, double tol, int lanes_per_row,
                       void* workspace, size_t workspace_bytes, int* iterations, double* rr, double* bb,
                       uintptr_t stream) {
    return guarded([&] {
        solve_bsr3(n_rows, diag, row_offsets, col_indices, values, b, x, max_iterations, tol, lanes_per_row,
                   workspace, workspace_bytes, nullptr, nullptr, 0, iterations, rr, bb, stream);
    });
}

int chx_pcg_solve_bsr3_contact(int n_rows, const float* diag, const int* row_offsets, const int* col_indices,
                               const float* values, const float* b, float* x, int max_iterations, double tol,
                               int lanes_per_row, void* workspace, size_t workspace_bytes,
                               const chx_contact_op* op, void* contact_workspace, size_t contact_workspace_bytes,
                               int* iterations, double* rr, double* bb, uintptr_t stream) {
    return guarded([&] {
        require(op != nullptr, "chx_pcg_solve_bsr3_contact: null contact operator");
        solve_bsr3(n_rows, diag, row_offsets, col_indices, values, b, x, max_iterations, tol, lanes_per_row,
                   workspace, workspace_bytes, op, contact_workspace, contact_workspace_bytes, iterations, rr, bb,
                   stream);
    });
}

int chx_pcg_last_stats(int n_rows, const void* workspace, size_t workspace_bytes, int* iterations, double* rr,
                       double* bb, double* rho, uintptr_t stream) {
    return guarded([&] {
        require(n_rows >= 0, "chx_pcg_last_stats: negative row count");
        if (n_rows == 0) {
            if (iterations) *iterations = 0;
            if (rr) *rr = 0.0;
            if (bb) *bb = 0.0;
            if (rho) *rho = 0.0;
            return;
        }
        require(workspace != nullptr, "chx_pcg_last_stats: null workspace");
        size_t need = 0;
        if (chx_pcg_workspace_size(n_rows, &need) != CHX_OK) throw Error(CHX_ERR_RUNTIME, chx_last_error_message());
        require(workspace_bytes >= need, "chx_pcg_last_stats: workspace too small");
        const chx_cloth_batch B = carve_workspace(n_rows, const_cast<void*>(workspace));
        read_workspace_stats(B, iterations, rr, bb, rho, as_stream(stream));
    });
}

int chx_cloth_assemble(const chx_cloth_batch* batch, uintptr_t stream) {
    return guarded([&] {
        validate_batch(batch);
        assemble(*batch, as_stream(stream));
    });
}

int chx_cloth_solve(const chx_cloth_batch* batch, int finalize, uintptr_t stream) {
    return guarded([&] {
        validate_batch(batch);
        const hipStream_t s = as_stream(stream);
        reset_sync_kernel<<<1, kBlock, 0, s>>>(batch->sync, 0);
        hip_check(hipGetLastError(), "reset_sync_kernel launch");
        launch_pcg(*batch, finalize, s);
    });
}

int chx_cloth_step(const chx_cloth_batch* batch, uintptr_t stream) {
    return guarded([&] {
        validate_batch(batch);
        const hipStream_t s = as_stream(stream);
        // host-side ranges named after the reference's step phases
        // (cloth_simulator.cu:887-1253): the assembly kernel is its
        // predictor, gradient and Hessian phases, the PCG launch its PCG and
        // finalize phases
        roctxRangePushA("step::hessian");
        assemble(*batch, s);
        roctxRangePop();
        roctxRangePushA("step::pcg");
        launch_pcg(*batch, 1, s);
        roctxRangePop();
    });
}

int chx_cloth_step_graph(const chx_cloth_batch* batch, uintptr_t stream, void** graph_exec) {
    return guarded([&] {
        require(graph_exec != nullptr, "chx_cloth_step_graph: null output");
        *graph_exec = nullptr;
        validate_batch(batch);
        const hipStream_t s = as_stream(stream);
        require(s != nullptr, "chx_cloth_step_graph: capture needs a stream other than the null stream");
        // Device limits are queried (and cached) before the capture starts.
        (void)max_coresident(batch->lanes_per_row, pcg_variant(*batch, false), batch->no_lds_staging == 0);
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipGraph_t g = nullptr;
        try {
            assemble(*batch, s);
            launch_pcg(*batch, 1, s);
        } catch (...) {
            (void)hipStreamEndCapture(s, &g);
            if (g != nullptr) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "hipGraphInstantiate");
        *graph_exec = ex;
    });
}

int chx_graph_launch(void* graph_exec, uintptr_t stream) {
    return guarded([&] {
        require(graph_exec != nullptr, "chx_graph_launch: null graph");
        hip_check(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), as_stream(stream)), "hipGraphLaunch");
    });
}

int chx_graph_destroy(void* graph_exec) {
    return guarded([&] {
        if (graph_exec != nullptr)
            hip_check(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)), "hipGraphExecDestroy");
    });
}

int chx_batch_check(const chx_cloth_batch* batch, uintptr_t stream) {
    return guarded([&] {
        require(batch != nullptr && batch->sync != nullptr, "chx_batch_check: null batch");
        unsigned words[SYNC_WORDS];
        const hipStream_t s = as_stream(stream);
        hip_check(hipMemcpyAsync(words, batch->sync, sizeof(words), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(sync)");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize(check)");
        if (words[SYNC_STATUS] != 0u) {
            reset_sync_kernel<<<1, kBlock, 0, s>>>(batch->sync, 1);
            hip_check(hipStreamSynchronize(s), "hipStreamSynchronize(reset)");
            if (batch->status_host != nullptr)
                __atomic_store_n(batch->status_host, 0u, __ATOMIC_SEQ_CST);
            const unsigned grid = words[SYNC_GRID];
            throw Error(CHX_ERR_DEVICE, "chysx::solver::PCGSolver: grid barrier timed out (grid of " +
                                            std::to_string(grid) + " workgroups: " +
                                            std::to_string(words[SYNC_SEEN]) + " of " +
                                            std::to_string(grid < kSyncGroups ? grid : kSyncGroups) +
                                            " groups complete, " + std::to_string(words[SYNC_WAITING]) +
                                            " workgroups waiting in the others)");
        }
    });
}

}  // extern "C"
