#include "cloth_batch.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace chysx {
namespace cloth {

int ClothBatch::add_instance(ClothSimulator& sim) {
    if (std::find(sims_.begin(), sims_.end(), &sim) != sims_.end())
        throw std::invalid_argument("ClothBatch::add_instance: simulator already in the batch");
    sims_.push_back(&sim);
    bound_ = false;
    return static_cast<int>(sims_.size()) - 1;
}

bool ClothBatch::remove_instance(ClothSimulator& sim) {
    const auto it = std::find(sims_.begin(), sims_.end(), &sim);
    if (it == sims_.end()) return false;
    sims_.erase(it);
    bound_ = false;
    return true;
}

void ClothBatch::clear() noexcept {
    sims_.clear();
    bound_ = false;
}

std::vector<chx_instance> ClothBatch::descriptors(float dt, std::uintptr_t cuda_stream) {
    std::vector<chx_instance> ds;
    ds.reserve(sims_.size());
    for (ClothSimulator* s : sims_) {
        ds.push_back(s->prepare_instance(dt, cuda_stream));
        if (bf16_ && ds.back().n_rows > 0 && ds.back().diag_c == nullptr)
            throw std::invalid_argument(
                "ClothBatch: BF16 mode needs every instance switched with set_bf16_blocks(true)");
    }
    return ds;
}

void ClothBatch::finalize(float dt, std::uintptr_t cuda_stream) {
    if (sims_.empty()) throw std::runtime_error("ClothBatch::finalize: no instances");
    runner_.bind_many(descriptors(dt, cuda_stream), cuda_stream);
    bound_ = true;
}

void ClothBatch::step(float dt, std::uintptr_t cuda_stream) {
    if (!(dt > 0.0f)) throw std::invalid_argument("ClothBatch::step: dt must be positive");
    ProfileRange range("chysx::cloth::batch_step");
    if (!bound_) finalize(dt, cuda_stream);
    else runner_.update_many(descriptors(dt, cuda_stream), cuda_stream);
    if (graph_) runner_.step_graph(cuda_stream);
    else runner_.step(cuda_stream);
}

void ClothBatch::assemble(float dt, std::uintptr_t cuda_stream) {
    if (!(dt > 0.0f)) throw std::invalid_argument("ClothBatch::assemble: dt must be positive");
    if (!bound_) finalize(dt, cuda_stream);
    else runner_.update_many(descriptors(dt, cuda_stream), cuda_stream);
    runner_.assemble(cuda_stream);
}

void ClothBatch::solve(std::uintptr_t cuda_stream) {
    if (!bound_) throw std::runtime_error("ClothBatch::solve: call assemble() first");
    runner_.solve(true, cuda_stream);
}

int ClothBatch::num_rows() const noexcept {
    int n = 0;
    for (const ClothSimulator* s : sims_) n += s->buffers().particle_count();
    return n;
}

std::size_t ClothBatch::device_bytes() const noexcept {
    std::size_t b = 0;
    for (const ClothSimulator* s : sims_) b += s->device_bytes();
    return b;
}

std::vector<std::pair<int, int>> ClothBatch::instance_sizes() const {
    std::vector<std::pair<int, int>> out;
    for (const ClothSimulator* s : sims_) out.emplace_back(s->buffers().particle_count(), s->pattern().nnz());
    return out;
}

std::vector<int> ClothBatch::last_pcg_iterations(std::uintptr_t cuda_stream) {
    std::vector<int> it;
    std::vector<double> rr, bb;
    if (bound_) runner_.read_stats(it, rr, bb, cuda_stream);
    return it;
}

std::vector<double> ClothBatch::last_relative_residuals(std::uintptr_t cuda_stream) {
    std::vector<int> it;
    std::vector<double> rr, bb, out;
    if (!bound_) return out;
    runner_.read_stats(it, rr, bb, cuda_stream);
    out.resize(rr.size());
    for (std::size_t i = 0; i < rr.size(); ++i) out[i] = bb[i] > 0.0 ? std::sqrt(rr[i] / bb[i]) : std::sqrt(rr[i]);
    return out;
}

}  // namespace cloth
}  // namespace chysx
