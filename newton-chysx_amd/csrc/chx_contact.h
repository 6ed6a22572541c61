// Contact SpMV sidecar: device helpers shared by the contact calls
// (chx_contact.hip) and the persistent PCG (chx_step.hip), and the host
// entry to the per-row inverted contact index.
#pragma once

#include "chx_device.h"
#include "chx_elements.h"

namespace chx {

// Per-row (contact, slot) lists: entries[rows[i] .. rows[i+1]) hold 4 c + s
// for every slot s of contact c on row i (nonzero weight), ascending.
struct ContactIndex {
    int* rows = nullptr;
    int* entries = nullptr;
};

size_t contact_workspace_bytes(int n_rows, int max_contacts);
// Builds the index on stream `s` (device work only, no host wait).
ContactIndex build_contact_index(int n_rows, const chx_contact_op& op, void* workspace, size_t bytes, hipStream_t s);

__device__ __forceinline__ float contact_weight(const chx_contact_op& op, int c, int s) {
    return rec_at(glb(op.weights), unsigned(c), 8u)[s];
}

// alpha H_c, row-major: (k' - a) n n^T + a I with k' = alpha k and the
// friction stiffness a = mu k' depth f1(|u_t|) (contact_spmv.cu:65-95).
__device__ __forceinline__ void contact_block(const chx_contact_op& op, int c, float alpha, float (&h)[9]) {
    const gp<const float> w = rec_at(glb(op.weights), unsigned(c), 8u);
    const float ka = alpha * op.stiffness;
    float af = 0.0f;
    if (op.friction_mu > 0.0f && op.slips != nullptr) {
        const float fn = ka * w[7];
        af = op.friction_mu * fn * ipc_f1_over_x(rec_at(glb(op.slips), unsigned(c), 4u)[3], op.friction_eps);
    }
    const float diff = ka - af, nx = w[4], ny = w[5], nz = w[6];
    const float h01 = diff * nx * ny, h02 = diff * nx * nz, h12 = diff * ny * nz;
    h[0] = diff * nx * nx + af;
    h[1] = h01;
    h[2] = h02;
    h[3] = h01;
    h[4] = diff * ny * ny + af;
    h[5] = h12;
    h[6] = h02;
    h[7] = h12;
    h[8] = diff * nz * nz + af;
}

// alpha * sum over row i's entries (c, s): w_s H_c (sum_{u != s} w_u x[id_u]),
// in entry order; src(j) reads x[j] (contact_spmv.cu:140-180).  Slots of
// zero weight or with a particle index outside [0, n) add nothing.
template <class Src>
__device__ __forceinline__ f3 contact_row(const chx_contact_op& op, const int* rows, const int* entries, unsigned i,
                                          unsigned n, float alpha, Src src) {
    const gp<const int> rg = glb(rows), eg = glb(entries);
    f3 out = mk(0.f, 0.f, 0.f);
    const int e1 = rg[i + 1u];
    for (int k = rg[i]; k < e1; ++k) {
        const int v = eg[unsigned(k)];
        const int c = v >> 2, s = v & 3;
        const gp<const int> ids = rec_at(glb(op.pairs), unsigned(c), 4u);
        const gp<const float> w = rec_at(glb(op.weights), unsigned(c), 8u);
        const float ka = alpha * op.stiffness;
        float af = 0.0f;
        if (op.friction_mu > 0.0f && op.slips != nullptr) {
            const float fn = ka * w[7];
            af = op.friction_mu * fn * ipc_f1_over_x(rec_at(glb(op.slips), unsigned(c), 4u)[3], op.friction_eps);
        }
        float tx = 0.f, ty = 0.f, tz = 0.f;
        for (int u = 0; u < 4; ++u) {
            const float wu = w[u];
            if (u == s || wu == 0.0f || unsigned(ids[u]) >= n) continue;
            const f3 xu = src(unsigned(ids[u]));
            tx += wu * xu.x;
            ty += wu * xu.y;
            tz += wu * xu.z;
        }
        const float nx = w[4], ny = w[5], nz = w[6], ws = w[s];
        const float dn = nx * tx + ny * ty + nz * tz;
        const float ncoef = ws * (ka - af) * dn;
        const float wa = ws * af;
        out = mk(out.x + (ncoef * nx + wa * tx), out.y + (ncoef * ny + wa * ty), out.z + (ncoef * nz + wa * tz));
    }
    return out;
}

}  // namespace chx
