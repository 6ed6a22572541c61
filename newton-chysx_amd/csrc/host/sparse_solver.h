// chysx::sparse::BlockCSR3 / spmv and chysx::solver::PCGSolver — drop-ins
// for ChysX/src/sparse/block_csr.h:59-148 and
// ChysX/src/solver/pcg_solver.h:79-167 over the C ABI.
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

    void initialize(int num_block_rows);
    // Solves A x = b in place (x is the warm start); returns the number of
    // iterations performed (== max_iterations when tolerance <= 0).
    int solve(const sparse::BlockCSR3& A, DeviceSpan<math::Vec3f> b, DeviceSpan<math::Vec3f> x,
              const PCGParams& params = PCGParams{}, std::uintptr_t cuda_stream = 0);
    // Preconditioner-weighted residual of the last solve: <r, z> is not kept;
    // this returns ||r||_2 / ||b||_2 of the final iterate (FP64).
    double last_residual() const noexcept { return last_rel_residual_; }

private:
    int n_ = 0;
    DeviceArray<unsigned char> workspace_;
    double last_rel_residual_ = 0.0;
};

}  // namespace solver
}  // namespace chysx
