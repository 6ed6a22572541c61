// C ABI: one-time setup kernels (rest shapes, rest angles, lumped mass) and
// the diagnostic energy evaluation.
#include <algorithm>
#include <vector>

#include "chx_device.h"
#include "chx_elements.h"
#include "chx_internal.h"

namespace chx {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void rest_shape_kernel(const int* __restrict__ tris, const float* __restrict__ pos,
                                                            int n, float cos_t, float sin_t, float* __restrict__ dm_inv,
                                                            float* __restrict__ area) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const f3 a = ld3(pos, tris[3 * t]), b = ld3(pos, tris[3 * t + 1]), c = ld3(pos, tris[3 * t + 2]);
    float dm[4], ar;
    tri_rest_shape(a, b, c, cos_t, sin_t, dm, ar);
    dm_inv[4 * t + 0] = dm[0];
    dm_inv[4 * t + 1] = dm[1];
    dm_inv[4 * t + 2] = dm[2];
    dm_inv[4 * t + 3] = dm[3];
    area[t] = ar;
}

__global__ __launch_bounds__(kBlock) void rest_angle_kernel(const int* __restrict__ dih, const float* __restrict__ pos,
                                                            int n, float* __restrict__ rest) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int* d = dih + 4 * e;
    const DihFrame F = dihedral_frame(ld3(pos, d[0]), ld3(pos, d[1]), ld3(pos, d[2]), ld3(pos, d[3]));
    rest[e] = F.valid ? F.theta : 0.0f;
}

// Vertex-owner lumped mass: m_v = sum over incident triangles (ascending) of
// density/3 * area(t); inv_mass = m > 1e-12 ? 1/m : 0
// (cloth_simulator.cu:151-190: same per-triangle term, gathered instead of
// scattered with atomics).
__global__ __launch_bounds__(kBlock) void mass_kernel(const int* __restrict__ tris, const float* __restrict__ pos,
                                                      const int* __restrict__ vt_off, const int* __restrict__ vt,
                                                      int n, float d3, float* __restrict__ inv_mass) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    float m = 0.0f;
    for (int k = vt_off[v]; k < vt_off[v + 1]; ++k) {
        const int t = vt[k];
        const f3 x0 = ld3(pos, tris[3 * t]), x1 = ld3(pos, tris[3 * t + 1]), x2 = ld3(pos, tris[3 * t + 2]);
        const float area = 0.5f * len(cross(sub(x1, x0), sub(x2, x0)));
        m += d3 * area;
    }
    inv_mass[v] = (m > 1.0e-12f) ? 1.0f / m : 0.0f;
}

// FP64 energy of one element-pass work item per workgroup (fixed-order
// workgroup sum) into out[item].
__global__ __launch_bounds__(kBlock) void energy_kernel(chx_cloth_batch B, int mask, double* __restrict__ out) {
    __shared__ double lds[16];
    int inst, kind, ib, ie;
    get_elem_item(B, blockIdx.x, inst, kind, ib, ie);
    const Inst I = inst_at(B, inst);
    const int i = ib + threadIdx.x;
    double E[1] = {0.0};
    if (i < ie) {
        if (kind == 0 && (mask & 1)) {
            const chx_pin p = ldg(I.pins, unsigned(i));
            const f3 d = sub(ld3(I.pos, p.v), ld3g(I.pin_targets, unsigned(i)));
            E[0] = double(0.5f * p.k * dot(d, d));  // pin_constraint.cu:40-53
        } else if (kind == 1 && (mask & 2)) {
            const chx_tri t = ldg(I.tris, unsigned(i));
            E[0] = tri_energy(ld3(I.pos, t.v0), ld3(I.pos, t.v1), ld3(I.pos, t.v2), t.dm00, t.dm01, t.dm10, t.dm11,
                              t.kA);
        } else if (kind == 3 && (mask & 8)) {
            const chx_stvk t = ldg(I.stvks, unsigned(i));
            E[0] = stvk_energy(ld3(I.pos, t.v0), ld3(I.pos, t.v1), ld3(I.pos, t.v2), t.dm00, t.dm01, t.dm10, t.dm11,
                               t.muA, t.lamA);
        } else if (kind == 2 && (mask & 4)) {
            const chx_dih d = ldg(I.dihs, unsigned(i));
            const DihFrame F =
                dihedral_frame(ld3(I.pos, d.v0), ld3(I.pos, d.v1), ld3(I.pos, d.v2), ld3(I.pos, d.v3));
            if (F.valid) {  // bending_constraint.cu:154-172
                const float dth = F.theta - d.rest;
                E[0] = double(0.5f * d.k * dth * dth);
            }
        }
    }
    block_sum(E, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = E[0];
}

}  // namespace
}  // namespace chx

using namespace chx;

extern "C" {

int chx_triangle_rest_shape(const int* tris, const float* pos, int n, float cos_theta, float sin_theta, float* dm_inv,
                            float* area, uintptr_t stream) {
    return guarded([&] {
        require(n >= 0, "chx_triangle_rest_shape: negative count");
        if (n == 0) return;
        require(tris && pos && dm_inv && area, "chx_triangle_rest_shape: null pointer");
        rest_shape_kernel<<<blocks_for(n, kBlock), kBlock, 0, as_stream(stream)>>>(tris, pos, n, cos_theta, sin_theta,
                                                                                   dm_inv, area);
        hip_check(hipGetLastError(), "rest_shape_kernel launch");
    });
}

int chx_dihedral_rest_angle(const int* dihedrals, const float* pos, int n, float* rest, uintptr_t stream) {
    return guarded([&] {
        require(n >= 0, "chx_dihedral_rest_angle: negative count");
        if (n == 0) return;
        require(dihedrals && pos && rest, "chx_dihedral_rest_angle: null pointer");
        rest_angle_kernel<<<blocks_for(n, kBlock), kBlock, 0, as_stream(stream)>>>(dihedrals, pos, n, rest);
        hip_check(hipGetLastError(), "rest_angle_kernel launch");
    });
}

int chx_mass_area_weighted(const int* tris, const float* pos, const int* vertex_tri_offsets, const int* vertex_tris,
                           int n_vertices, float density_over_3, float* inv_mass, uintptr_t stream) {
    return guarded([&] {
        require(n_vertices >= 0, "chx_mass_area_weighted: negative count");
        if (n_vertices == 0) return;
        require(tris && pos && vertex_tri_offsets && vertex_tris && inv_mass, "chx_mass_area_weighted: null pointer");
        mass_kernel<<<blocks_for(n_vertices, kBlock), kBlock, 0, as_stream(stream)>>>(
            tris, pos, vertex_tri_offsets, vertex_tris, n_vertices, density_over_3, inv_mass);
        hip_check(hipGetLastError(), "mass_kernel launch");
    });
}

int chx_batch_energy(const chx_cloth_batch* batch, int kind_mask, double* energy, uintptr_t stream) {
    return guarded([&] {
        require(batch && energy, "chx_batch_energy: null pointer");
        const chx_cloth_batch& B = *batch;
        auto cnt = [](int n) { return (n + CHX_ITEM_SIZE - 1) / CHX_ITEM_SIZE; };
        const long total = B.inst != nullptr
                               ? long(B.n_elem_items)
                               : long(cnt(B.single.n_pins)) + cnt(B.single.n_tris) + cnt(B.single.n_dihs) +
                                     cnt(B.single.n_stvks);
        *energy = 0.0;
        if (total == 0) return;
        double* dev = nullptr;
        hip_check(hipMalloc(&dev, sizeof(double) * total), "hipMalloc(energy)");
        std::vector<double> host(total);
        try {
            energy_kernel<<<static_cast<int>(total), kBlock, 0, as_stream(stream)>>>(B, kind_mask, dev);
            hip_check(hipGetLastError(), "energy_kernel launch");
            hip_check(hipMemcpyAsync(host.data(), dev, sizeof(double) * total, hipMemcpyDeviceToHost,
                                     as_stream(stream)),
                      "hipMemcpyAsync(energy)");
            hip_check(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize(energy)");
        } catch (...) {
            (void)hipFree(dev);
            throw;
        }
        hip_check(hipFree(dev), "hipFree(energy)");
        double s = 0.0;
        for (double v : host) s += v;
        *energy = s;
    });
}

}  // extern "C"
