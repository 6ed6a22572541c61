// Contact SpMV sidecar: the per-row inverted contact index and the bake /
// SpMV calls of the C ABI (reference: ChysX/src/collision/contact_spmv.h:
// 59-135, contact_spmv.cu:46-268).
//
// The reference scatters every contact's 4x4 slot blocks into the diagonal
// and into y with float atomics.  Here each call first builds, on the
// device, the list of (contact, slot) entries of every particle row:
// integer-atomic counts, an exclusive scan, a fill and a per-row sort into
// ascending (contact, slot) order.  The bake and the SpMV then run one
// thread per row, summing that row's terms in that order — no float
// atomics, the same bits run to run and for any contact order on the
// device.  The persistent PCG gathers its rows' contact terms through the
// same index (chx_step.hip, `contact_row`).
#include "chx_contact.h"
#include "chx_device.h"
#include "chx_internal.h"

namespace chx {
namespace {

constexpr int kScanBlock = 1024;
constexpr int kThreads = 256;

__device__ __forceinline__ int active_contacts(const chx_contact_op& op) {
    if (op.count == nullptr || op.pairs == nullptr) return 0;
    const int c = *glb(op.count);
    return c < op.max_contacts ? (c > 0 ? c : 0) : op.max_contacts;
}

__global__ void fill_int_kernel(int* __restrict__ p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Slots of zero weight, or whose particle index lies outside [0, n), are
// not indexed (the reference skips the former and would fault on the latter).
__device__ __forceinline__ bool indexed(gp<const int> ids, gp<const float> w, int s, int n) {
    return w[s] != 0.0f && unsigned(ids[s]) < unsigned(n);
}

__global__ void count_kernel(chx_contact_op op, int* __restrict__ cnt, int n) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= active_contacts(op)) return;
    const gp<const int> ids = rec_at(glb(op.pairs), unsigned(c), 4u);
    const gp<const float> w = rec_at(glb(op.weights), unsigned(c), 8u);
    for (int s = 0; s < 4; ++s)
        if (indexed(ids, w, s, n)) atomicAdd(&cnt[ids[s]], 1);
}

// Exclusive scan of cnt[0, n) into start[0, n] in blocks of kScanBlock
// (4 per thread), block totals into sums[].
__global__ void scan_blocks_kernel(const int* __restrict__ cnt, int* __restrict__ start, int* __restrict__ sums,
                                   int n) {
    __shared__ int wsum[kThreads / 64];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * 4;
    int v[4], s = 0;
    for (int q = 0; q < 4; ++q) {
        v[q] = (base + q < n) ? cnt[base + q] : 0;
        s += v[q];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = s;
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; ++w) before += wsum[w];
    int run = before + incl - s;
    for (int q = 0; q < 4; ++q) {
        if (base + q < n) start[base + q] = run;
        run += v[q];
    }
    if (threadIdx.x == blockDim.x - 1) sums[blockIdx.x] = before + incl;
}

// Exclusive scan of the block totals (one workgroup, serial per chunk),
// total into start[n].
__global__ void scan_sums_kernel(int* __restrict__ sums, int nb, int* __restrict__ start, int n) {
    if (threadIdx.x != 0) return;
    int run = 0;
    for (int b = 0; b < nb; ++b) {
        const int t = sums[b];
        sums[b] = run;
        run += t;
    }
    start[n] = run;
}

__global__ void add_offsets_kernel(int* __restrict__ start, const int* __restrict__ sums, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) start[i] += sums[i / kScanBlock];
}

__global__ void fill_kernel(chx_contact_op op, const int* __restrict__ start, int* __restrict__ cursor,
                            int* __restrict__ entries, int n) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= active_contacts(op)) return;
    const gp<const int> ids = rec_at(glb(op.pairs), unsigned(c), 4u);
    const gp<const float> w = rec_at(glb(op.weights), unsigned(c), 8u);
    for (int s = 0; s < 4; ++s) {
        if (!indexed(ids, w, s, n)) continue;
        const int i = ids[s];
        entries[start[i] + atomicAdd(&cursor[i], 1)] = 4 * c + s;
    }
}

// Ascending (contact, slot) order within each row (insertion sort: rows
// hold a few entries).
__global__ void sort_rows_kernel(const int* __restrict__ start, int* __restrict__ entries, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = start[i], e = start[i + 1];
    for (int k = b + 1; k < e; ++k) {
        const int v = entries[k];
        int j = k - 1;
        while (j >= b && entries[j] > v) {
            entries[j + 1] = entries[j];
            --j;
        }
        entries[j + 1] = v;
    }
}

__global__ void bake_kernel(chx_contact_op op, const int* __restrict__ start, const int* __restrict__ entries,
                            float* __restrict__ diag, float alpha, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = start[i]; k < start[i + 1]; ++k) {
        const int c = entries[k] >> 2, s = entries[k] & 3;
        float h[9];
        contact_block(op, c, alpha, h);
        const float ww = contact_weight(op, c, s) * contact_weight(op, c, s);
        for (int q = 0; q < 9; ++q) acc[q] += ww * h[q];
    }
    if (start[i + 1] == start[i]) return;
    const gp<float> d = rec_at(glb(diag), unsigned(i), 9u);
    for (int q = 0; q < 9; ++q) d[q] = d[q] + acc[q];
}

__global__ void spmv_kernel(chx_contact_op op, const int* __restrict__ start, const int* __restrict__ entries,
                            const float* __restrict__ x, float* __restrict__ y, float alpha, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || start[i + 1] == start[i]) return;
    const gp<const float> xg = glb(x);
    const f3 c = contact_row(op, start, entries, unsigned(i), unsigned(n), alpha, [&](unsigned j) { return ld3g(xg, j); });
    const gp<float> yg = glb(y);
    st3g(yg, unsigned(i), add(ld3g(gp<const float>(yg), unsigned(i)), c));
}

}  // namespace

size_t contact_workspace_bytes(int n_rows, int max_contacts) {
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t n = size_t(n_rows), m = size_t(max_contacts > 0 ? max_contacts : 0);
    const size_t nb = (n + kScanBlock - 1) / kScanBlock + 1;
    return al(4 * (n + 1)) + al(4 * n) + al(4 * 4 * m + 4) + al(4 * nb);
}

ContactIndex build_contact_index(int n_rows, const chx_contact_op& op, void* workspace, size_t bytes,
                                 hipStream_t s) {
    require(n_rows >= 0 && op.max_contacts >= 0, "chysx contact: negative size");
    require(bytes >= contact_workspace_bytes(n_rows, op.max_contacts), "chysx contact: workspace too small");
    require(op.max_contacts == 0 || (op.pairs && op.weights), "chysx contact: null pairs / weights");
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t n = size_t(n_rows), m = size_t(op.max_contacts);
    char* w = static_cast<char*>(workspace);
    ContactIndex ix;
    ix.rows = reinterpret_cast<int*>(w);
    w += al(4 * (n + 1));
    int* cursor = reinterpret_cast<int*>(w);
    w += al(4 * n);
    ix.entries = reinterpret_cast<int*>(w);
    w += al(4 * 4 * m + 4);
    int* sums = reinterpret_cast<int*>(w);
    const int nb = int((n + kScanBlock - 1) / kScanBlock);
    const int mb = blocks_for(long(std::max<size_t>(m, 1)), kThreads);
    if (n_rows > 0) {
        fill_int_kernel<<<blocks_for(n_rows, kThreads), kThreads, 0, s>>>(cursor, n_rows, 0);
        count_kernel<<<mb, kThreads, 0, s>>>(op, cursor, n_rows);
        scan_blocks_kernel<<<std::max(nb, 1), kThreads, 0, s>>>(cursor, ix.rows, sums, n_rows);
        scan_sums_kernel<<<1, 64, 0, s>>>(sums, nb, ix.rows, n_rows);
        add_offsets_kernel<<<blocks_for(n_rows, kThreads), kThreads, 0, s>>>(ix.rows, sums, n_rows);
        fill_int_kernel<<<blocks_for(n_rows, kThreads), kThreads, 0, s>>>(cursor, n_rows, 0);
        fill_kernel<<<mb, kThreads, 0, s>>>(op, ix.rows, cursor, ix.entries, n_rows);
        sort_rows_kernel<<<blocks_for(n_rows, kThreads), kThreads, 0, s>>>(ix.rows, ix.entries, n_rows);
        hip_check(hipGetLastError(), "contact index launch");
    }
    return ix;
}

}  // namespace chx

using namespace chx;

extern "C" {

int chx_contact_workspace_size(int n_rows, int max_contacts, size_t* bytes) {
    return guarded([&] {
        require(n_rows >= 0 && max_contacts >= 0 && bytes, "chx_contact_workspace_size: bad argument");
        *bytes = contact_workspace_bytes(n_rows, max_contacts);
    });
}

int chx_contact_bake_diag(int n_rows, float* diag, const chx_contact_op* op, float alpha, void* workspace,
                          size_t workspace_bytes, uintptr_t stream) {
    return guarded([&] {
        require(op != nullptr && (diag != nullptr || n_rows == 0), "chx_contact_bake_diag: null argument");
        require(size_t(n_rows) * 36u < (size_t(1) << 32), "chx_contact_bake_diag: blocks must stay below 4 GiB");
        if (n_rows == 0 || op->max_contacts == 0 || op->count == nullptr) return;
        const hipStream_t s = as_stream(stream);
        const ContactIndex ix = build_contact_index(n_rows, *op, workspace, workspace_bytes, s);
        bake_kernel<<<blocks_for(n_rows, kThreads), kThreads, 0, s>>>(*op, ix.rows, ix.entries, diag, alpha, n_rows);
        hip_check(hipGetLastError(), "contact bake launch");
    });
}

int chx_contact_spmv(int n_rows, const chx_contact_op* op, const float* x, float* y, float alpha, void* workspace,
                     size_t workspace_bytes, uintptr_t stream) {
    return guarded([&] {
        require(op != nullptr && ((x && y) || n_rows == 0), "chx_contact_spmv: null argument");
        if (n_rows == 0 || op->max_contacts == 0 || op->count == nullptr) return;
        const hipStream_t s = as_stream(stream);
        const ContactIndex ix = build_contact_index(n_rows, *op, workspace, workspace_bytes, s);
        spmv_kernel<<<blocks_for(n_rows, kThreads), kThreads, 0, s>>>(*op, ix.rows, ix.entries, x, y, alpha, n_rows);
        hip_check(hipGetLastError(), "contact spmv launch");
    });
}

}  // extern "C"
