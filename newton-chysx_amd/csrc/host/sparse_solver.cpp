#include "sparse_solver.h"

#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace chysx {
namespace sparse {

// Move-only, as the reference (block_csr.h:67-68, pcg_solver.h:88-89).
static_assert(std::is_move_constructible<BlockCSR3>::value && std::is_move_assignable<BlockCSR3>::value &&
                  !std::is_copy_constructible<BlockCSR3>::value,
              "BlockCSR3 is move-only");
static_assert(std::is_move_constructible<solver::PCGSolver>::value &&
                  std::is_move_assignable<solver::PCGSolver>::value &&
                  !std::is_copy_constructible<solver::PCGSolver>::value,
              "PCGSolver is move-only");

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

namespace collision {
namespace {

chx_contact_op to_abi(const ContactSpMVOp& op) {
    chx_contact_op c{};
    c.pairs = reinterpret_cast<const int*>(op.pairs);
    c.weights = reinterpret_cast<const float*>(op.weights);
    c.count = op.count_dev;
    c.slips = op.friction_active() ? reinterpret_cast<const float*>(op.slips) : nullptr;
    c.max_contacts = op.active() ? op.max_contacts : 0;
    c.stiffness = op.stiffness;
    c.friction_mu = op.friction_active() ? op.friction_mu : 0.0f;
    c.friction_eps = op.friction_epsilon;
    return c;
}

// Library-kept contact workspaces, one per (device, stream): calls on one
// stream run in order, so they may share its workspace; calls on different
// streams never share one.  A workspace only grows (waiting for its stream
// before the old one is freed) and is held until the process ends -- it is
// deliberately never freed at exit, when the runtime may be gone.
struct StreamWorkspace {
    int device;
    std::uintptr_t stream;
    void* ptr;
    std::size_t bytes;
};
std::mutex g_ws_mu;
std::vector<StreamWorkspace>* g_ws = new std::vector<StreamWorkspace>();

void* stream_workspace(std::size_t need, std::uintptr_t stream, std::size_t* bytes) {
    int dev = 0;
    check(chx_get_device(&dev), "collision", "current device");
    std::lock_guard<std::mutex> lk(g_ws_mu);
    for (StreamWorkspace& w : *g_ws) {
        if (w.device != dev || w.stream != stream) continue;
        if (w.bytes < need) {
            check(stream ? chx_stream_synchronize(stream) : chx_device_synchronize(), "collision", "workspace growth");
            check(chx_free(w.ptr), "collision", "workspace growth");
            w.ptr = nullptr;
            w.bytes = 0;
            check(chx_malloc(&w.ptr, need), "collision", "workspace growth");
            w.bytes = need;
        }
        *bytes = w.bytes;
        return w.ptr;
    }
    void* p = nullptr;
    check(chx_malloc(&p, need), "collision", "workspace");
    g_ws->push_back(StreamWorkspace{dev, stream, p, need});
    *bytes = need;
    return p;
}

// Runs `call(workspace, bytes)` with the op's workspace, or the stream's
// library-kept one.
template <class Call>
void with_workspace(const ContactSpMVOp& op, int n, std::uintptr_t cuda_stream, const char* what, Call call) {
    const std::size_t need = op.workspace_bytes(n);
    if (op.workspace != nullptr) {
        if (op.workspace_size < need) throw std::invalid_argument(std::string("chysx::collision::") + what +
                                                                  ": ContactSpMVOp workspace too small");
        call(op.workspace, op.workspace_size);
        return;
    }
    std::size_t bytes = 0;
    void* ws = stream_workspace(need, cuda_stream, &bytes);
    call(ws, bytes);
}

}  // namespace

std::size_t ContactSpMVOp::workspace_bytes(int n_particles) const {
    std::size_t bytes = 0;
    check(chx_contact_workspace_size(n_particles, max_contacts > 0 ? max_contacts : 0, &bytes), "collision",
          "contact workspace size");
    return bytes;
}

void bake_contact_diag(math::Mat3f* diag_blocks, int n_particles, const ContactSpMVOp& op, float alpha,
                       std::uintptr_t cuda_stream) {
    if (!op.active() || n_particles <= 0 || op.count_dev == nullptr) return;
    const chx_contact_op c = to_abi(op);
    with_workspace(op, n_particles, cuda_stream, "bake_contact_diag", [&](void* ws, std::size_t bytes) {
        check(chx_contact_bake_diag(n_particles, reinterpret_cast<float*>(diag_blocks), &c, alpha, ws, bytes,
                                    cuda_stream),
              "collision", "bake_contact_diag");
    });
}

void apply_contact_spmv(const ContactSpMVOp& op, const math::Vec3f* x, math::Vec3f* y, int n_particles, float alpha,
                        std::uintptr_t cuda_stream) {
    if (!op.active() || n_particles <= 0 || op.count_dev == nullptr) return;
    const chx_contact_op c = to_abi(op);
    with_workspace(op, n_particles, cuda_stream, "apply_contact_spmv", [&](void* ws, std::size_t bytes) {
        check(chx_contact_spmv(n_particles, &c, reinterpret_cast<const float*>(x), reinterpret_cast<float*>(y), alpha,
                               ws, bytes, cuda_stream),
              "collision", "apply_contact_spmv");
    });
}

}  // namespace collision

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
                     const PCGParams& params, std::uintptr_t cuda_stream, collision::ContactSpMVOp contact) {
    const int n = A.num_block_rows();
    if (n == 0) return 0;
    if (static_cast<int>(b.size()) < n || static_cast<int>(x.size()) < n)
        throw std::invalid_argument("PCGSolver::solve: b/x shorter than A rows");
    if (static_cast<int>(A.diag.gpu_size()) < n)
        throw std::invalid_argument(
            "PCGSolver::solve: A.diag has fewer than A.num_block_rows() entries; call A.build_topology(...) before "
            "solving");
    initialize(n);
    last_stream_ = cuda_stream;
    // With a tolerance the iteration count is only known on the device:
    // wait for it (an addition); otherwise the call stays asynchronous.
    const bool wait = params.tolerance > 0.0;
    int it = 0;
    int* itp = wait ? &it : nullptr;
    if (contact.active() && contact.count_dev != nullptr) {
        const chx_contact_op c = collision::to_abi(contact);
        void* cws = contact.workspace;
        std::size_t cbytes = contact.workspace_size;
        const std::size_t need = contact.workspace_bytes(n);
        if (cws == nullptr || cbytes < need) {
            contact_workspace_.ensure_device(need);
            cws = contact_workspace_.gpu_data();
            cbytes = contact_workspace_.gpu_size();
        }
        check(chx_pcg_solve_bsr3_contact(
                  n, reinterpret_cast<const float*>(A.diag.gpu_data()), A.row_offsets.gpu_data(),
                  A.col_indices.gpu_data(), reinterpret_cast<const float*>(A.values.gpu_data()),
                  reinterpret_cast<const float*>(b.data()), reinterpret_cast<float*>(x.data()), params.max_iterations,
                  params.tolerance, params.lanes_per_row, workspace_.gpu_data(), workspace_.gpu_size(), &c, cws,
                  cbytes, itp, nullptr, nullptr, cuda_stream),
              "solver::PCGSolver", "solve");
    } else {
        check(chx_pcg_solve_bsr3(n, reinterpret_cast<const float*>(A.diag.gpu_data()), A.row_offsets.gpu_data(),
                                 A.col_indices.gpu_data(), reinterpret_cast<const float*>(A.values.gpu_data()),
                                 reinterpret_cast<const float*>(b.data()), reinterpret_cast<float*>(x.data()),
                                 params.max_iterations, params.tolerance, params.lanes_per_row, workspace_.gpu_data(),
                                 workspace_.gpu_size(), itp, nullptr, nullptr, cuda_stream),
              "solver::PCGSolver", "solve");
    }
    return wait ? it : params.max_iterations;
}

void PCGSolver::read_stats(int* it, double* rr, double* bb, double* rho) {
    check(chx_pcg_last_stats(n_, workspace_.gpu_data(), workspace_.gpu_size(), it, rr, bb, rho, last_stream_),
          "solver::PCGSolver", "statistics read");
}

float PCGSolver::last_residual() {
    if (n_ == 0 || workspace_.gpu_size() == 0) return 0.0f;
    double rho = 0.0;
    read_stats(nullptr, nullptr, nullptr, &rho);
    return static_cast<float>(rho);
}

int PCGSolver::last_iterations() {
    if (n_ == 0 || workspace_.gpu_size() == 0) return 0;
    int it = 0;
    read_stats(&it, nullptr, nullptr, nullptr);
    return it;
}

double PCGSolver::last_relative_residual() {
    if (n_ == 0 || workspace_.gpu_size() == 0) return 0.0;
    double rr = 0.0, bb = 0.0;
    read_stats(nullptr, &rr, &bb, nullptr);
    return bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

}  // namespace solver
}  // namespace chysx
