// chysx::sparse::BlockCSR3 / spmv, chysx::collision::ContactSpMVOp and
// chysx::solver::PCGSolver — drop-ins for
// ChysX/src/sparse/block_csr.h:59-148, ChysX/src/collision/contact_spmv.h:
// 14-135 and ChysX/src/solver/pcg_solver.h:79-167 over the C ABI.
#pragma once

#include <cstdint>
#include <vector>

#include "chysx_common.h"
#include "topology.h"

namespace chysx {
namespace sparse {

class BlockCSR3 {
public:
    BlockCSR3() = default;
    BlockCSR3(const BlockCSR3&) = delete;
    BlockCSR3& operator=(const BlockCSR3&) = delete;
    BlockCSR3(BlockCSR3&&) = default;  // move-only (block_csr.h:67-68)
    BlockCSR3& operator=(BlockCSR3&&) = default;

    int num_block_rows() const noexcept { return pattern_.n; }
    int num_off_diag_blocks() const noexcept { return pattern_.nnz(); }

    DeviceArray<math::Mat3f> diag;
    DeviceArray<int> row_offsets;
    DeviceArray<int> col_indices;
    DeviceArray<math::Mat3f> values;

    // Sorted, de-duplicated off-diagonal pattern from (row, col) pairs;
    // diag and values are allocated and zeroed (block_csr.cu:91-147).
    void build_topology(int num_block_rows, const int* host_rows, const int* host_cols, int num_pairs);
    void set_zero(std::uintptr_t cuda_stream = 0);
    void resolve_slots(const int* host_rows, const int* host_cols, int* out_slots, int num_pairs) const {
        topology::resolve_slots(pattern_, host_rows, host_cols, out_slots, num_pairs);
    }
    const topology::Pattern& pattern() const noexcept { return pattern_; }

private:
    topology::Pattern pattern_;
};

// y = alpha * A * x + beta * y (beta == 0: y not read).
void spmv(const BlockCSR3& A, DeviceSpan<math::Vec3f> x, DeviceSpan<math::Vec3f> y, float alpha = 1.0f,
          float beta = 0.0f, std::uintptr_t cuda_stream = 0, int lanes_per_row = 8);

}  // namespace sparse

namespace collision {

// One contact's weights, normal and depth (ChysX/src/collision/
// self_collision.h, ContactWeights): 8 floats.
struct ContactWeights {
    float w0 = 0.f, w1 = 0.f, w2 = 0.f, w3 = 0.f;
    float nx = 0.f, ny = 0.f, nz = 0.f;
    float depth = 0.f;
};
static_assert(sizeof(ContactWeights) == 32, "8 floats");

// The contact SpMV sidecar (contact_spmv.h:14-50).  Contact c adds
// sum_{s,t} w_s w_t H_c to the blocks (pairs[c][s], pairs[c][t]) with
// H_c = k n n^T + a_c (I - n n^T), a_c = mu k depth f1(|u_t|).
struct ContactSpMVOp {
    const math::Vec4i* pairs = nullptr;
    const ContactWeights* weights = nullptr;
    const int* count_dev = nullptr;  // device int: active contacts
    int max_contacts = 0;
    float stiffness = 0.0f;
    const math::Vec4f* slips = nullptr;  // (u_t, |u_t|) per contact, or null
    float friction_mu = 0.0f;
    float friction_epsilon = 1.0e-4f;
    // Addition: device workspace of at least workspace_bytes() bytes for the
    // per-row contact index.  Null: bake / apply use a workspace the library
    // keeps per (device, stream), grown on demand and held until the process
    // ends, so steady-state calls neither allocate nor wait.
    void* workspace = nullptr;
    std::size_t workspace_size = 0;

    bool active() const noexcept { return max_contacts > 0 && stiffness > 0.0f && pairs != nullptr; }
    bool friction_active() const noexcept { return active() && friction_mu > 0.0f && slips != nullptr; }
    // Device workspace the calls need for `n_particles` rows.
    std::size_t workspace_bytes(int n_particles) const;
};

// diag_blocks[i] += alpha * sum_c w_i^2 H_c (contact_spmv.h:71-96).
// Deterministic: each row sums its contacts in (contact, slot) order.
void bake_contact_diag(math::Mat3f* diag_blocks, int n_particles, const ContactSpMVOp& op, float alpha,
                       std::uintptr_t cuda_stream);

// y += alpha * C_off x: the cross-slot terms (contact_spmv.h:98-132).
void apply_contact_spmv(const ContactSpMVOp& op, const math::Vec3f* x, math::Vec3f* y, int n_particles, float alpha,
                        std::uintptr_t cuda_stream);

}  // namespace collision

namespace solver {

struct PCGParams {
    int max_iterations = 50;
    double tolerance = 0.0;  // addition: <= 0 keeps the reference's fixed iteration count
    int lanes_per_row = 8;
};

class PCGSolver {
public:
    PCGSolver() = default;
    PCGSolver(const PCGSolver&) = delete;
    PCGSolver& operator=(const PCGSolver&) = delete;
    PCGSolver(PCGSolver&&) = default;  // move-only (pcg_solver.h:88-89)
    PCGSolver& operator=(PCGSolver&&) = default;

    void initialize(int num_block_rows);
    // Solves A x = b in place (x is the warm start), asynchronously on
    // `cuda_stream` (pcg_solver.cu:282-318): with tolerance <= 0 it runs
    // exactly max_iterations and returns that number without waiting.  With
    // a tolerance (addition) it waits for the solve and returns the
    // iterations performed.  An active `contact` turns every product into
    // (A + C_off) x; its same-slot terms are the caller's to bake into
    // A.diag (pcg_solver.h:111-131).
    int solve(const sparse::BlockCSR3& A, DeviceSpan<math::Vec3f> b, DeviceSpan<math::Vec3f> x,
              const PCGParams& params = PCGParams{}, std::uintptr_t cuda_stream = 0,
              collision::ContactSpMVOp contact = {});
    // Preconditioner-weighted residual <r, z> of the last solve's final
    // iteration, copied to the host on demand (pcg_solver.cu:319-323).
    float last_residual();
    // Additions, also read on demand: iterations performed and the final
    // relative residual ||r||_2 / ||b||_2 (FP64 norms) of the last solve.
    int last_iterations();
    double last_relative_residual();

private:
    void read_stats(int* it, double* rr, double* bb, double* rho);

    int n_ = 0;
    DeviceArray<unsigned char> workspace_;
    DeviceArray<unsigned char> contact_workspace_;
    std::uintptr_t last_stream_ = 0;
};

}  // namespace solver
}  // namespace chysx
