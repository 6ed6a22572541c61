// Device-side helpers shared by the gfx950 kernels: float3 / 3x3 math in
// the reference's row-major layout, wave64 reductions, and the grid barrier
// of the persistent PCG kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "chysx_hip.h"

namespace chx {

constexpr int kWave = 64;

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

__device__ __forceinline__ f3 ld3(const float* p, long i) {
    const float* q = p + 3 * i;
    return {q[0], q[1], q[2]};
}
__device__ __forceinline__ void st3(float* p, long i, f3 v) {
    float* q = p + 3 * i;
    q[0] = v.x;
    q[1] = v.y;
    q[2] = v.z;
}

// 3x3 row-major block times vector: m[r*3+c].
__device__ __forceinline__ f3 mv9(const float* m, f3 v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
            m[6] * v.x + m[7] * v.y + m[8] * v.z};
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

// 3x3 block stored as BF16, applied in FP32.
__device__ __forceinline__ f3 mv9_bf16(const uint16_t* m, f3 v) {
    float a[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) a[i] = bf16_to_f32(m[i]);
    return mv9(a, v);
}

// Matrix views for the SpMV: FP32 blocks, or BF16 copies with FP32 math.
struct Mat32 {
    const float* d;
    const float* v;
    __device__ __forceinline__ f3 diag(long row, f3 x) const { return mv9(d + 9 * row, x); }
    __device__ __forceinline__ f3 off(long k, f3 x) const { return mv9(v + 9 * k, x); }
};
struct Mat16 {
    const uint16_t* d;
    const uint16_t* v;
    __device__ __forceinline__ f3 diag(long row, f3 x) const { return mv9_bf16(d + 9 * row, x); }
    __device__ __forceinline__ f3 off(long k, f3 x) const { return mv9_bf16(v + 9 * k, x); }
};

__device__ __forceinline__ double dot64(f3 a, f3 b) {
    return double(a.x) * double(b.x) + double(a.y) * double(b.y) + double(a.z) * double(b.z);
}

// Instance descriptor i of a batch (the inline `single` one when the batch
// has no device instance table).  Both cases go through one pointer so the
// code that reads the descriptor is the same instruction sequence.
__device__ __forceinline__ const chx_instance& inst_at(const chx_cloth_batch& B, int i) {
    const chx_instance* p = (B.inst != nullptr) ? B.inst + i : &B.single;
    return *p;
}

// Work-list entries (int4 arrays of the batch, or implied by the counts of
// the single instance).
__device__ __forceinline__ void get_tile(const chx_cloth_batch& B, int t, int& inst, int& rb, int& re) {
    if (B.tiles != nullptr) {
        const int* p = B.tiles + 4 * t;
        inst = p[0];
        rb = p[1];
        re = p[2];
    } else {
        inst = 0;
        rb = t * CHX_TILE_ROWS;
        re = min(B.single.n_rows, rb + CHX_TILE_ROWS);
    }
}

__device__ __forceinline__ int blocks_of(int n) { return (n + CHX_ITEM_SIZE - 1) / CHX_ITEM_SIZE; }

__device__ __forceinline__ void get_elem_item(const chx_cloth_batch& B, int w, int& inst, int& kind, int& b, int& e) {
    if (B.elem_items != nullptr) {
        const int* p = B.elem_items + 4 * w;
        inst = p[0];
        kind = p[1];
        b = p[2];
        e = p[3];
        return;
    }
    const chx_instance& I = B.single;
    inst = 0;
    const int np = blocks_of(I.n_pins), nt = blocks_of(I.n_tris);
    int n;
    if (w < np) {
        kind = 0;
        n = I.n_pins;
    } else if (w < np + nt) {
        kind = 1;
        w -= np;
        n = I.n_tris;
    } else {
        kind = 2;
        w -= np + nt;
        n = I.n_dihs;
    }
    b = w * CHX_ITEM_SIZE;
    e = min(n, b + CHX_ITEM_SIZE);
}

__device__ __forceinline__ void get_slot_item(const chx_cloth_batch& B, int w, int& inst, int& b, int& e) {
    if (B.slot_items != nullptr) {
        const int* p = B.slot_items + 4 * w;
        inst = p[0];
        b = p[1];
        e = p[2];
        return;
    }
    inst = 0;
    b = w * CHX_ITEM_SIZE;
    e = min(B.single.n_rows + B.single.nnz, b + CHX_ITEM_SIZE);
}

// ---- wave64 / workgroup reductions (fixed order => deterministic) ---------
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum over a segment of `width` consecutive lanes (width a power of two <= 64).
template <int W>
__device__ __forceinline__ float seg_sum(float v) {
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
//   status word, then the arrival count seen by a workgroup that gave up and
//   the grid size.
enum : int {
    SYNC_COUNT = 0,
    SYNC_GEN = 64,
    SYNC_ABORT = 128,
    SYNC_ACTIVE = 192,
    SYNC_STATUS = 256,
    SYNC_SEEN = 257,
    SYNC_GRID = 258,
    SYNC_SNAP = 320,  // two snapshots of SYNC_ACTIVE, 64 words apart
    SYNC_WORDS = CHX_SYNC_WORDS
};
static_assert(SYNC_SNAP + 64 < CHX_SYNC_WORDS, "sync area too small");
// The barrier is self-resetting: the last arriving workgroup zeroes the
// count and bumps the generation, so consecutive launches reuse it.  Every
// wait is bounded in time: a workgroup that has waited kWaitTicks of the
// 100 MHz real-time counter raises the abort flag and all workgroups leave
// the kernel, so a launch whose workgroups are not all resident ends with an
// error instead of hanging.  Waiters back off exponentially so ~1000
// workgroups polling one word do not starve the arrivals.
constexpr unsigned long long kWaitTicks = 100000000ull;  // 1 s at 100 MHz

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
// every probe.  Each wave first drains its own outstanding stores.  The last
// arriver acquires the other arrivals before it releases the generation, so
// every workgroup's writes reach every other one.
//
// With snap = 0 or 1 the last arriver also copies SYNC_ACTIVE into snapshot
// slot `snap` before it releases the generation, and every workgroup reads
// that slot back into *active (uniform across the workgroup).  The live
// counter may already be changing in the next phase when a slow workgroup
// leaves the barrier; the snapshot of this barrier is only rewritten two
// barriers later, which needs that workgroup's arrival first.
__device__ __forceinline__ bool grid_sync(unsigned* sync, unsigned nblocks, int snap = -1,
                                          unsigned* active = nullptr) {
    __shared__ int s_ok;
    __shared__ unsigned s_active;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned gen = ld_relaxed(&sync[SYNC_GEN]);
        const unsigned arrived =
            __hip_atomic_fetch_add(&sync[SYNC_COUNT], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nblocks - 1) {
            st_relaxed(&sync[SYNC_COUNT], 0u);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (snap >= 0) st_relaxed(&sync[SYNC_SNAP + 64 * snap], ld_relaxed(&sync[SYNC_ACTIVE]));
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_fetch_add(&sync[SYNC_GEN], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
            unsigned level = 0;
            while (ld_relaxed(&sync[SYNC_GEN]) == gen) {
                if (ld_relaxed(&sync[SYNC_ABORT]) != 0u) {
                    ok = 0;
                    break;
                }
                if (__builtin_amdgcn_s_memrealtime() - t0 > kWaitTicks) {
                    st_relaxed(&sync[SYNC_SEEN], ld_relaxed(&sync[SYNC_COUNT]));
                    st_relaxed(&sync[SYNC_GRID], nblocks);
                    st_relaxed(&sync[SYNC_ABORT], 1u);
                    st_relaxed(&sync[SYNC_STATUS], 1u);
                    ok = 0;
                    break;
                }
                backoff(level);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        s_ok = ok;
        if (snap >= 0) s_active = ld_relaxed(&sync[SYNC_SNAP + 64 * snap]);
    }
    __syncthreads();
    if (active != nullptr) *active = s_active;
    return s_ok != 0;
}

}  // namespace chx
