// Device-side helpers shared by the gfx950 kernels: float3 / 3x3 math in
// the reference's row-major layout, wave64 reductions, and the grid barrier
// of the persistent PCG kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "chysx_hip.h"

namespace chx {

constexpr int kWave = 64;

// Global-memory pointers.  Array pointers that the kernels read out of an
// instance descriptor are generic pointers to the compiler, which then emits
// FLAT loads and stores (64-bit VGPR addresses, counted against both the
// vector-memory and the LDS/scalar wait counters).  Retyping them into
// address space 1 gives GLOBAL instructions with an SGPR base and a 32-bit
// VGPR offset.  All arrays of this library live in device (global) memory.
template <class T>
using gp = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gp<T> glb(T* p) {
    return (gp<T>)p;
}

// ---- float3 --------------------------------------------------------------
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float a, float b, float c) { return {a, b, c}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 mul(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len(f3 a) { return sqrtf(dot(a, a)); }

// Element i / record i of a global array, addressed with a 32-bit byte
// offset: the access then compiles to a GLOBAL instruction with the array
// base in SGPRs and one 32-bit VGPR offset (64-bit index arithmetic would
// put a full 64-bit address in VGPRs per access).  The host keeps every
// array that the hot kernels index this way below 4 GiB per instance.
template <class T>
__device__ __forceinline__ gp<const T> rec_at(gp<const T> p, unsigned i, unsigned fields) {
    return (gp<const T>)((gp<const char>)p + i * (fields * unsigned(sizeof(T))));
}
template <class T>
__device__ __forceinline__ gp<T> rec_at(gp<T> p, unsigned i, unsigned fields) {
    return (gp<T>)((gp<char>)p + i * (fields * unsigned(sizeof(T))));
}
template <class T>
__device__ __forceinline__ T ldx(gp<const T> p, unsigned i) {
    return *rec_at(p, i, 1u);
}
template <class T>
__device__ __forceinline__ void stx(gp<T> p, unsigned i, T v) {
    *rec_at(p, i, 1u) = v;
}
__device__ __forceinline__ f3 ld3g(gp<const float> p, unsigned i) {
    const gp<const float> q = rec_at(p, i, 3u);
    return {q[0], q[1], q[2]};
}
__device__ __forceinline__ void st3g(gp<float> p, unsigned i, f3 v) {
    const gp<float> q = rec_at(p, i, 3u);
    q[0] = v.x;
    q[1] = v.y;
    q[2] = v.z;
}

// Loads record i of 4-byte fields through a global pointer (a C++ copy of an
// address-space-1 lvalue would bind it to a generic reference).
template <class T>
__device__ __forceinline__ T ldg(gp<const T> base, unsigned i) {
    static_assert(sizeof(T) % 4 == 0, "record of 4-byte fields");
    T v;
    const gp<const unsigned> src = (gp<const unsigned>)rec_at(base, i, 1u);
    unsigned* dst = reinterpret_cast<unsigned*>(&v);
#pragma unroll
    for (unsigned k = 0; k < sizeof(T) / 4; ++k) dst[k] = src[k];
    return v;
}

template <class P>
__device__ __forceinline__ f3 ld3(P p, long i) {
    return {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}
template <class P>
__device__ __forceinline__ void st3(P p, long i, f3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

// 3x3 row-major block times vector: m[r*3+c].  The multiply-adds are
// spelled out (no contraction left to the compiler), so every kernel that
// applies a block -- staged or global SpMV, batch or single instance --
// rounds it the same way.
template <class P>
__device__ __forceinline__ f3 mv9(P m, f3 v) {
    return {fmaf(m[2], v.z, fmaf(m[1], v.y, m[0] * v.x)), fmaf(m[5], v.z, fmaf(m[4], v.y, m[3] * v.x)),
            fmaf(m[8], v.z, fmaf(m[7], v.y, m[6] * v.x))};
}

// Adjugate / determinant inverse without a singular guard
// (ChysX/src/math/matrix.cuh:226-241).
__device__ __forceinline__ void inverse9(const float* m, float* r) {
    r[0] = m[4] * m[8] - m[5] * m[7];
    r[1] = -m[1] * m[8] + m[2] * m[7];
    r[2] = m[1] * m[5] - m[2] * m[4];
    r[3] = -m[3] * m[8] + m[5] * m[6];
    r[4] = m[0] * m[8] - m[2] * m[6];
    r[5] = -m[0] * m[5] + m[2] * m[3];
    r[6] = m[3] * m[7] - m[4] * m[6];
    r[7] = -m[0] * m[7] + m[1] * m[6];
    r[8] = m[0] * m[4] - m[1] * m[3];
    const float det = m[0] * r[0] + m[1] * r[3] + m[2] * r[6];
    const float inv = 1.0f / det;
#pragma unroll
    for (int i = 0; i < 9; ++i) r[i] *= inv;
}

// BF16 <-> FP32 (round to nearest even; NaN stays NaN).
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<uint16_t>((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(static_cast<unsigned>(h) << 16); }

__device__ __forceinline__ float to_f32(float v) { return v; }
__device__ __forceinline__ float to_f32(uint16_t v) { return bf16_to_f32(v); }

// Matrix views for the SpMV and the preconditioner set-up.  diag(row, x) and
// off(k, x) apply a diagonal / off-diagonal 3x3 block in FP32; diag_block
// loads a diagonal block as nine row-major floats.
//
// Mat32: row-major 3x3 blocks, 36 bytes each (the reference's BlockCSR3
// layout, used by the standalone SpMV / PCG entry points).
struct Mat32G {
    gp<const float> d;
    gp<const float> v;
    __device__ __forceinline__ f3 diag(unsigned row, f3 x) const { return mv9(rec_at(d, row, 9u), x); }
    __device__ __forceinline__ f3 off(unsigned k, f3 x) const { return mv9(rec_at(v, k, 9u), x); }
    __device__ __forceinline__ void diag_block(unsigned row, float (&m)[9]) const {
        const gp<const float> q = rec_at(d, row, 9u);
#pragma unroll
        for (int c = 0; c < 9; ++c) m[c] = q[c];
    }
};

// Grouped paired-plane blocks: 3x3 blocks in groups of 64, each group one
// contiguous run of 2304 bytes holding the components in planes of pairs —
// (c0,c1), (c2,c3), (c4,c5), (c6,c7) as 64 two-word pairs each, then c8 as
// 64 words (component c row-major r*3+col; CHX_BLOCK_GROUP).  One load
// instruction of a wavefront reads one pair of 64 consecutive blocks (512
// contiguous bytes), so a block takes five loads instead of nine, and the
// nine components of a group stay in one DRAM page.  Arrays hold whole
// groups (the block count rounded up to 64).
static_assert(CHX_BLOCK_GROUP == 64, "bp() indexes groups of 64 blocks");
__device__ __forceinline__ unsigned bp(unsigned c, unsigned i) {
    const unsigned g = (i >> 6) * 576u, l = i & 63u;
    return c < 8u ? g + (c >> 1) * 128u + 2u * l + (c & 1u) : g + 512u + l;
}
// Symmetric blocks (the cloth instances' diagonal blocks and block-Jacobi
// preconditioner, symmetric by construction) keep only their upper
// triangle: groups of 64 blocks as three planes of pairs (xx,xy), (xz,yy),
// (yz,zz); component q of block i at p[(i / 64) * 384 + (q / 2) * 128 +
// 2 * (i % 64) + q % 2].
__device__ __forceinline__ unsigned bp6(unsigned q, unsigned i) {
    return (i >> 6) * 384u + (q >> 1) * 128u + 2u * (i & 63u) + (q & 1u);
}
// Row-major position of upper-triangle component q.
__device__ __forceinline__ unsigned sym6_pos(unsigned q) { return q < 3u ? q : (q < 5u ? q + 1u : 8u); }
// Two consecutive components (c even) of block i as one load: two floats,
// or two BF16 words in one 32-bit load.
typedef float f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void ld_pair(gp<const float> m, unsigned off, float& a, float& b) {
    const f2v v = *(gp<const f2v>)rec_at(m, off, 1u);
    a = v.x;
    b = v.y;
}
__device__ __forceinline__ void ld_pair(gp<const uint16_t> m, unsigned off, float& a, float& b) {
    const unsigned v = *(gp<const unsigned>)rec_at(m, off, 1u);
    a = bf16_to_f32(uint16_t(v & 0xffffu));
    b = bf16_to_f32(uint16_t(v >> 16));
}
// Stores of the same pairs (assembly).
__device__ __forceinline__ void st_pair(gp<float> m, unsigned off, float a, float b) {
    f2v v;
    v.x = a;
    v.y = b;
    *(gp<f2v>)rec_at(m, off, 1u) = v;
}
__device__ __forceinline__ void st_pair(gp<uint16_t> m, unsigned off, uint16_t a, uint16_t b) {
    *(gp<unsigned>)rec_at(m, off, 1u) = unsigned(a) | (unsigned(b) << 16);
}
template <class T>
__device__ __forceinline__ void ld6p(gp<const T> m, unsigned i, float (&a)[9]) {
    float u[6];
#pragma unroll
    for (unsigned q = 0; q < 6; q += 2) ld_pair(m, bp6(q, i), u[q], u[q + 1]);
    a[0] = u[0];
    a[1] = u[1];
    a[2] = u[2];
    a[3] = u[1];
    a[4] = u[3];
    a[5] = u[4];
    a[6] = u[2];
    a[7] = u[4];
    a[8] = u[5];
}
template <class T>
__device__ __forceinline__ f3 mv6p(gp<const T> m, unsigned i, f3 v) {
    float a[9];
    ld6p(m, i, a);
    return mv9(a, v);
}
template <class T>
__device__ __forceinline__ void ld9p(gp<const T> m, unsigned i, float (&a)[9]) {
#pragma unroll
    for (unsigned c = 0; c < 8; c += 2) ld_pair(m, bp(c, i), a[c], a[c + 1]);
    a[8] = to_f32(ldx(m, bp(8u, i)));
}
template <class T>
__device__ __forceinline__ f3 mv9p(gp<const T> m, unsigned i, f3 v) {
    float a[9];
    ld9p(m, i, a);
    return mv9(a, v);
}
// Block i, or its transpose when tr (lower blocks of the symmetric layout
// are their upper partners' transposes): the same loads, swapped operands.
template <class T>
__device__ __forceinline__ f3 mv9p_t(gp<const T> m, unsigned i, f3 v, bool tr) {
    float a[9];
    ld9p(m, i, a);
    const float b[9] = {a[0], tr ? a[3] : a[1], tr ? a[6] : a[2], tr ? a[1] : a[3], a[4],
                        tr ? a[7] : a[5], tr ? a[2] : a[6], tr ? a[5] : a[7], a[8]};
    return mv9(b, v);
}
template <class T>
struct MatP {
    gp<const T> d;  // symmetric diagonal blocks (block rows, bp6)
    gp<const T> v;  // off-diagonal blocks (value slots, bp)
    __device__ __forceinline__ f3 diag(unsigned row, f3 x) const { return mv6p(d, row, x); }
    __device__ __forceinline__ f3 off(unsigned k, f3 x) const { return mv9p(v, k, x); }
    __device__ __forceinline__ f3 off_t(unsigned k, f3 x, bool tr) const { return mv9p_t(v, k, x, tr); }
    __device__ __forceinline__ void diag_block(unsigned row, float (&m)[9]) const { ld6p(d, row, m); }
};
using MatP32 = MatP<float>;

// BF16 precision mode's inner matrix: FP32 compensated diagonal blocks and
// BF16 off-diagonal blocks, FP32 arithmetic.
struct MatP16 {
    gp<const float> d;
    gp<const uint16_t> v;
    __device__ __forceinline__ f3 diag(unsigned row, f3 x) const { return mv9p(d, row, x); }
    __device__ __forceinline__ f3 off(unsigned k, f3 x) const { return mv9p(v, k, x); }
    __device__ __forceinline__ f3 off_t(unsigned k, f3 x, bool tr) const { return mv9p_t(v, k, x, tr); }
};

__device__ __forceinline__ double dot64(f3 a, f3 b) {
    return double(a.x) * double(b.x) + double(a.y) * double(b.y) + double(a.z) * double(b.z);
}

// Device view of one chx_instance: the same fields with global-memory
// pointer types.  Built from the device instance table, or from the inline
// `single` descriptor when the batch has none (both with scalar loads).
struct Inst {
    gp<float> pos, vel;
    gp<const float> inv_mass;
    gp<const int> row_offsets, col_indices;
    gp<float> diag, values;
    gp<const chx_pin> pins;
    gp<const float> pin_targets;
    gp<const chx_tri> tris;
    gp<const chx_dih> dihs;
    gp<const chx_unit> units;
    gp<const int> unit_elems;
    gp<const uint16_t> contrib_offsets, contribs;
    gp<float> rhs, x, r, z, p0, p1, q, q1, minv;
    gp<float> diag_c;
    gp<uint16_t> values16;
    gp<const int> windows;
    gp<const int> slice_offsets, slice_entries, upper_offsets, block_slots;
    gp<const uint16_t> slice_entries16;
    gp<const chx_plane> planes;
    gp<const chx_box> boxes;
    int n_planes, n_boxes;
    float contact_thickness, contact_stiffness, contact_friction, contact_friction_eps;
    gp<const chx_stvk> stvks;
    int n_stvks, n_padded, n_upper, n_units;
    int n_rows, nnz, n_pins, n_tris, n_dihs, tile_begin, tile_end;
    float dt, damping, gx, gy, gz;
    int max_iter;
    double tol;
};

template <class P>
__device__ __forceinline__ Inst make_inst(P d) {
    Inst I;
    I.pos = glb(d->pos);
    I.vel = glb(d->vel);
    I.inv_mass = glb(d->inv_mass);
    I.row_offsets = glb(d->row_offsets);
    I.col_indices = glb(d->col_indices);
    I.diag = glb(d->diag);
    I.values = glb(d->values);
    I.pins = glb(d->pins);
    I.pin_targets = glb(d->pin_targets);
    I.tris = glb(d->tris);
    I.dihs = glb(d->dihs);
    I.units = glb(d->units);
    I.unit_elems = glb(d->unit_elems);
    I.contrib_offsets = glb(d->contrib_offsets);
    I.contribs = glb(d->contribs);
    I.rhs = glb(d->rhs);
    I.x = glb(d->x);
    I.r = glb(d->r);
    I.z = glb(d->z);
    I.p0 = glb(d->p0);
    I.p1 = glb(d->p1);
    I.q = glb(d->q);
    I.q1 = glb(d->q1);
    I.minv = glb(d->minv);
    I.diag_c = glb(d->diag_c);
    I.values16 = glb(d->values16);
    I.windows = glb(d->windows);
    I.slice_offsets = glb(d->slice_offsets);
    I.slice_entries = glb(d->slice_entries);
    I.slice_entries16 = glb(d->slice_entries16);
    I.upper_offsets = glb(d->upper_offsets);
    I.block_slots = glb(d->block_slots);
    I.n_padded = d->n_padded;
    I.n_upper = d->n_upper;
    I.n_units = d->n_units;
    I.planes = glb(d->planes);
    I.boxes = glb(d->boxes);
    I.n_planes = d->n_planes;
    I.n_boxes = d->n_boxes;
    I.contact_thickness = d->contact_thickness;
    I.contact_stiffness = d->contact_stiffness;
    I.contact_friction = d->contact_friction;
    I.contact_friction_eps = d->contact_friction_eps;
    I.stvks = glb(d->stvks);
    I.n_stvks = d->n_stvks;
    I.n_rows = d->n_rows;
    I.nnz = d->nnz;
    I.n_pins = d->n_pins;
    I.n_tris = d->n_tris;
    I.n_dihs = d->n_dihs;
    I.tile_begin = d->tile_begin;
    I.tile_end = d->tile_end;
    I.dt = d->dt;
    I.damping = d->damping;
    I.gx = d->gx;
    I.gy = d->gy;
    I.gz = d->gz;
    I.max_iter = d->max_iter;
    I.tol = d->tol;
    return I;
}

// Instance i of a batch (the inline `single` one when the batch has no
// device instance table).
__device__ __forceinline__ Inst inst_at(const chx_cloth_batch& B, int i) {
    return (B.inst != nullptr) ? make_inst(rec_at(glb((const chx_instance*)B.inst), unsigned(i), 1u))
                               : make_inst(&B.single);
}

// Work-list entries (int4 arrays of the batch, or implied by the counts of
// the single instance).
__device__ __forceinline__ void get_tile(const chx_cloth_batch& B, int t, int& inst, int& rb, int& re) {
    if (B.tiles != nullptr) {
        const gp<const int> p = rec_at(glb((const int*)B.tiles), unsigned(t), 4u);
        inst = p[0];
        rb = p[1];
        re = p[2];
    } else {  // a single instance's implied list: tile_rows rows per tile
        const int tr = B.tile_rows > 0 ? B.tile_rows : CHX_TILE_ROWS;
        inst = 0;
        rb = t * tr;
        re = min(B.single.n_rows, rb + tr);
    }
}

__device__ __forceinline__ int blocks_of(int n) { return (n + CHX_ITEM_SIZE - 1) / CHX_ITEM_SIZE; }

__device__ __forceinline__ void get_elem_item(const chx_cloth_batch& B, int w, int& inst, int& kind, int& b, int& e) {
    if (B.elem_items != nullptr) {
        const gp<const int> p = rec_at(glb((const int*)B.elem_items), unsigned(w), 4u);
        inst = p[0];
        kind = p[1];
        b = p[2];
        e = p[3];
        return;
    }
    const chx_instance& I = B.single;
    inst = 0;
    const int np = blocks_of(I.n_pins), nt = blocks_of(I.n_tris), nd = blocks_of(I.n_dihs);
    int n;
    if (w < np) {
        kind = 0;
        n = I.n_pins;
    } else if (w < np + nt) {
        kind = 1;
        w -= np;
        n = I.n_tris;
    } else if (w < np + nt + nd) {
        kind = 2;
        w -= np + nt;
        n = I.n_dihs;
    } else {
        kind = 3;
        w -= np + nt + nd;
        n = I.n_stvks;
    }
    b = w * CHX_ITEM_SIZE;
    e = min(n, b + CHX_ITEM_SIZE);
}

// Assembly unit item w: {instance, unit}.
__device__ __forceinline__ void get_unit_item(const chx_cloth_batch& B, int w, int& inst, int& u) {
    if (B.unit_items != nullptr) {
        const gp<const int> p = rec_at(glb((const int*)B.unit_items), unsigned(w), 4u);
        inst = p[0];
        u = p[1];
        return;
    }
    inst = 0;
    u = w;
}

// ---- wave64 / workgroup reductions (fixed order => deterministic) ---------
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum over a segment of W consecutive lanes (W a power of two <= 64).
template <int W, class T>
__device__ __forceinline__ T seg_sum(T v) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Workgroup sum of up to 4 doubles per thread; result valid in thread 0.
// `lds` must hold 4 * (blockDim.x / 64) doubles.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* lds) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) lds[k * 4 + wid] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += lds[k * 4 + w];
            v[k] = s;
        }
    }
    __syncthreads();
}

// ---- grid barrier ----------------------------------------------------------
// Control words, each on its own 256-byte line so that polling the
// generation does not contend with arrivals:
//   sync[SYNC_COUNT] arrivals, sync[SYNC_GEN] generation, sync[SYNC_ABORT]
//   abort flag, sync[SYNC_ACTIVE] active-instance counter, sync[SYNC_STATUS]
//   status word, then, from a workgroup that gave up, the number of complete
//   workgroup groups, the grid size and the workgroups waiting in the
//   incomplete groups.
enum : int {
    SYNC_COUNT = 0,
    SYNC_GEN = 64,
    SYNC_ABORT = 128,
    SYNC_ACTIVE = 192,
    SYNC_STATUS = 256,
    SYNC_SEEN = 257,
    SYNC_GRID = 258,
    SYNC_WAITING = 259,
    SYNC_SNAP = 320,    // two snapshots of SYNC_ACTIVE, 64 words apart
    SYNC_XCOUNT = 448,  // arrival counters of the 8 workgroup groups, 64 words apart
    SYNC_FLAGS = 960,   // release flags (the generation), one per 128-byte line
    SYNC_WORDS = CHX_SYNC_WORDS
};
constexpr unsigned kSyncGroups = 8;  // workgroups g = x (mod 8): one XCD each
constexpr unsigned kSyncFlags = 64, kFlagStride = 32;  // workgroup g polls flag g mod 64
static_assert(SYNC_FLAGS + kFlagStride * kSyncFlags <= CHX_SYNC_WORDS && kSyncFlags == 64, "release flags");
static_assert(SYNC_SNAP + 64 < SYNC_XCOUNT && SYNC_XCOUNT + 64 * kSyncGroups <= SYNC_FLAGS,
              "sync area too small");
// The barrier is self-resetting: the last arriving workgroup zeroes the
// count and bumps the generation, so consecutive launches reuse it.  Every
// wait is bounded in time: a workgroup that has waited `ticks` of the
// 100 MHz real-time counter (kWaitTicks unless the batch sets a bound)
// raises the abort flag and the status word, sets the host-mapped status
// word when there is one, and all workgroups leave the kernel, so a launch
// whose workgroups are not all resident ends with an error instead of
// hanging.  Waiters back off exponentially so ~1000 workgroups polling one
// word do not starve the arrivals.
constexpr unsigned long long kWaitTicks = 100000000ull;  // 1 s at 100 MHz

struct Barrier {
    unsigned* sync;                // CHX_SYNC_WORDS control words
    unsigned* status_host;         // host-mapped abort report, or null
    unsigned long long ticks;      // bound of every wait
};
__host__ __device__ inline Barrier make_barrier(const chx_cloth_batch& B) {
    return Barrier{B.sync, B.status_host, B.wait_ticks ? (unsigned long long)B.wait_ticks : kWaitTicks};
}

__device__ __forceinline__ unsigned ld_relaxed(unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(unsigned* p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void backoff(unsigned& level) {
    switch (level) {
        case 0: __builtin_amdgcn_s_sleep(1); break;
        case 1: __builtin_amdgcn_s_sleep(2); break;
        case 2: __builtin_amdgcn_s_sleep(4); break;
        case 3: __builtin_amdgcn_s_sleep(8); break;
        default: __builtin_amdgcn_s_sleep(16); break;
    }
    if (level < 4) ++level;
}

// Fence placement matters on gfx950: each XCD has its own L2, so an
// agent-scope release writes the L2 back and an agent-scope acquire
// invalidates it.  The barrier therefore issues exactly one release and one
// acquire per workgroup, both from thread 0 (the workgroup barrier orders the
// other waves' accesses around them), arrives with a relaxed atomic, and
// polls with relaxed loads -- an acquiring poll would invalidate the L2 on
// every probe.  Each wave first drains its own outstanding stores.
//
// Arrival is two-level: agent-scope atomics on one word are performed one
// after another at the memory side (~23 ns each), so one counter for G
// workgroups costs ~G x 23 ns.  Workgroup g arrives at the counter of its
// group g mod 8 (the dispatcher's XCD); the group's last arriver resets it,
// acquires the group's arrivals, releases them on, and arrives at the top
// counter; the last of the groups acquires all of them before it releases
// the generation, so every workgroup's writes reach every other one.
//
// Release fans out the same way: the releasing workgroup's first wave writes
// the new generation into 64 flags, one per 128-byte line, after its
// release fence, and workgroup g polls flag g mod 64 -- four pollers a line
// instead of the whole grid on one word, whose loads the memory side also
// serves one after another (with one word the last workgroup saw the
// release ~2 us after the first on a grid of 256).  The wait bound is
// checked on every poll, the abort word (one line for the grid) on every
// 16th.
//
// With snap = 0 or 1 the last arriver also copies SYNC_ACTIVE into snapshot
// slot `snap` before it releases the generation, and every workgroup reads
// that slot back into *active (uniform across the workgroup).  The live
// counter may already be changing in the next phase when a slow workgroup
// leaves the barrier; the snapshot of this barrier is only rewritten two
// barriers later, which needs that workgroup's arrival first.
__device__ __forceinline__ bool grid_sync(const Barrier& bar, unsigned nblocks, int snap = -1,
                                          unsigned* active = nullptr) {
    unsigned* sync = bar.sync;
    __shared__ int s_ok;
    __shared__ unsigned s_active;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x < kWave) {
        const unsigned lane = threadIdx.x;
        unsigned* const flag = &sync[SYNC_FLAGS + kFlagStride * (blockIdx.x % kSyncFlags)];
        int ok = 1;
        bool last = false;
        unsigned gen = 0;
        if (lane == 0) {
            // The generation is read before the release fence, whose wait for
            // all outstanding memory operations completes the read before this
            // workgroup arrives.  (A read issued after the fence could still
            // be in flight when the last arrival bumps the generation; it
            // would return the next generation and this workgroup would wait
            // one barrier too long, stalling all others at the next barrier.)
            gen = ld_relaxed(flag);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            const unsigned x = blockIdx.x % kSyncGroups;
            const unsigned groups = nblocks < kSyncGroups ? nblocks : kSyncGroups;
            const unsigned gsize = (nblocks - x + kSyncGroups - 1u) / kSyncGroups;  // workgroups of group x
            unsigned* xc = &sync[SYNC_XCOUNT + 64 * x];
            if (__hip_atomic_fetch_add(xc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gsize - 1u) {
                st_relaxed(xc, 0u);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                last = __hip_atomic_fetch_add(&sync[SYNC_COUNT], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                       groups - 1u;
            }
            if (last) {
                st_relaxed(&sync[SYNC_COUNT], 0u);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                if (snap >= 0) st_relaxed(&sync[SYNC_SNAP + 64 * snap], ld_relaxed(&sync[SYNC_ACTIVE]));
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                st_relaxed(&sync[SYNC_GEN], gen + 1u);
            }
        }
        last = __builtin_amdgcn_readfirstlane(last ? 1 : 0) != 0;
        gen = __builtin_amdgcn_readfirstlane(gen);
        if (last) {
            // after lane 0's release fence in program order
            st_relaxed(&sync[SYNC_FLAGS + kFlagStride * lane], gen + 1u);
        } else if (lane == 0) {
            const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
            unsigned level = 0, polls = 0;
            while (ld_relaxed(flag) == gen) {
                if ((++polls & 15u) == 0u && ld_relaxed(&sync[SYNC_ABORT]) != 0u) {
                    ok = 0;
                    break;
                }
                if (__builtin_amdgcn_s_memrealtime() - t0 > bar.ticks) {
                    unsigned waiting = 0;
                    for (unsigned q = 0; q < kSyncGroups; ++q) waiting += ld_relaxed(&sync[SYNC_XCOUNT + 64 * q]);
                    st_relaxed(&sync[SYNC_SEEN], ld_relaxed(&sync[SYNC_COUNT]));
                    st_relaxed(&sync[SYNC_WAITING], waiting);
                    st_relaxed(&sync[SYNC_GRID], nblocks);
                    st_relaxed(&sync[SYNC_ABORT], 1u);
                    st_relaxed(&sync[SYNC_STATUS], 1u);
                    if (bar.status_host != nullptr)
                        __hip_atomic_store(bar.status_host, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    ok = 0;
                    break;
                }
                backoff(level);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        if (lane == 0) {
            s_ok = ok;
            if (snap >= 0) s_active = ld_relaxed(&sync[SYNC_SNAP + 64 * snap]);
        }
    }
    __syncthreads();
    if (active != nullptr) *active = s_active;
    return s_ok != 0;
}

}  // namespace chx
