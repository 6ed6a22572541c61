#include "sparse_solver.h"

#include <cmath>
#include <stdexcept>

namespace chysx {
namespace sparse {

void BlockCSR3::build_topology(int num_block_rows, const int* host_rows, const int* host_cols, int num_pairs) {
    pattern_ = topology::build_pattern(num_block_rows, host_rows, host_cols, num_pairs);
    row_offsets.host() = pattern_.row_offsets;
    col_indices.host() = pattern_.col_indices;
    row_offsets.copy_to_device();
    col_indices.copy_to_device();
    diag.ensure_device(num_block_rows);
    values.ensure_device(pattern_.col_indices.size());
    set_zero();
    check(chx_device_synchronize(), "sparse::BlockCSR3", "build_topology");
}

void BlockCSR3::set_zero(std::uintptr_t cuda_stream) {
    diag.zero_device(cuda_stream);
    values.zero_device(cuda_stream);
}

void spmv(const BlockCSR3& A, DeviceSpan<math::Vec3f> x, DeviceSpan<math::Vec3f> y, float alpha, float beta,
          std::uintptr_t cuda_stream, int lanes_per_row) {
    const int n = A.num_block_rows();
    if (n == 0) return;
    if (static_cast<int>(x.size()) < n) throw std::invalid_argument("spmv: x is shorter than num_block_rows");
    if (static_cast<int>(y.size()) < n) throw std::invalid_argument("spmv: y is shorter than num_block_rows");
    check(chx_spmv_bsr3(n, reinterpret_cast<const float*>(A.diag.gpu_data()), A.row_offsets.gpu_data(),
                        A.col_indices.gpu_data(), reinterpret_cast<const float*>(A.values.gpu_data()),
                        reinterpret_cast<const float*>(x.data()), reinterpret_cast<float*>(y.data()), alpha, beta,
                        lanes_per_row, cuda_stream),
          "sparse", "spmv");
}

}  // namespace sparse

namespace solver {

void PCGSolver::initialize(int num_block_rows) {
    if (num_block_rows < 0) throw std::invalid_argument("PCGSolver::initialize: negative size");
    if (num_block_rows == n_ && (workspace_.gpu_size() > 0 || num_block_rows == 0)) return;
    std::size_t bytes = 0;
    check(chx_pcg_workspace_size(num_block_rows, &bytes), "solver::PCGSolver", "workspace size");
    workspace_.ensure_device(bytes);
    n_ = num_block_rows;
}

int PCGSolver::solve(const sparse::BlockCSR3& A, DeviceSpan<math::Vec3f> b, DeviceSpan<math::Vec3f> x,
                     const PCGParams& params, std::uintptr_t cuda_stream) {
    const int n = A.num_block_rows();
    if (n == 0) return 0;
    if (static_cast<int>(b.size()) < n || static_cast<int>(x.size()) < n)
        throw std::invalid_argument("PCGSolver::solve: b/x shorter than A rows");
    if (static_cast<int>(A.diag.gpu_size()) < n)
        throw std::invalid_argument(
            "PCGSolver::solve: A.diag has fewer than A.num_block_rows() entries; call A.build_topology(...) before "
            "solving");
    initialize(n);
    int it = 0;
    double rr = 0.0, bb = 0.0;
    check(chx_pcg_solve_bsr3(n, reinterpret_cast<const float*>(A.diag.gpu_data()), A.row_offsets.gpu_data(),
                             A.col_indices.gpu_data(), reinterpret_cast<const float*>(A.values.gpu_data()),
                             reinterpret_cast<const float*>(b.data()), reinterpret_cast<float*>(x.data()),
                             params.max_iterations, params.tolerance, params.lanes_per_row, workspace_.gpu_data(),
                             workspace_.gpu_size(), &it, &rr, &bb, cuda_stream),
          "solver::PCGSolver", "solve");
    last_rel_residual_ = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    return it;
}

}  // namespace solver
}  // namespace chysx
