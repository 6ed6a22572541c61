#include "cloth_batch.h"

#include <cmath>
#include <stdexcept>

namespace chysx {
namespace cloth {

int ClothBatch::add_instance(ClothSimulator& sim) {
    sims_.push_back(&sim);
    packed_ = false;
    return static_cast<int>(sims_.size()) - 1;
}

void ClothBatch::set_lanes_per_row(int lanes) {
    if (lanes != 4 && lanes != 8 && lanes != 16)
        throw std::invalid_argument("ClothBatch::set_lanes_per_row: lanes must be 4, 8 or 16");
    lanes_ = lanes;
    batch_.set_lanes_per_row(lanes);
}

void ClothBatch::set_grid_blocks(int blocks) {
    grid_blocks_ = blocks < 0 ? 0 : blocks;
    batch_.set_grid_blocks(grid_blocks_);
}

void ClothBatch::finalize() {
    if (sims_.empty()) throw std::runtime_error("ClothBatch::finalize: no instances");
    std::vector<engine::InstanceSource> src;
    src.reserve(sims_.size());
    for (ClothSimulator* s : sims_) {
        s->prepare_topology();
        src.push_back(s->instance_source());
    }
    batch_.build(src, lanes_);
    batch_.set_grid_blocks(grid_blocks_);
    packed_ = true;
}

void ClothBatch::prepare_step(float dt) {
    if (!(dt > 0.0f)) throw std::invalid_argument("ClothBatch::step: dt must be positive");
    if (!packed_) finalize();
    for (std::size_t i = 0; i < sims_.size(); ++i) {
        engine::InstanceSource s = sims_[i]->instance_source();
        if (s.pos == nullptr || s.vel == nullptr)
            throw std::runtime_error("ClothBatch::step: every instance needs pos / vel buffers");
        s.dt = dt;
        batch_.set_instance(static_cast<int>(i), s);
    }
}

void ClothBatch::step(float dt, std::uintptr_t cuda_stream) {
    prepare_step(dt);
    batch_.step(cuda_stream);
}

void ClothBatch::assemble(float dt, std::uintptr_t cuda_stream) {
    prepare_step(dt);
    batch_.assemble(cuda_stream);
}

void ClothBatch::solve(std::uintptr_t cuda_stream) {
    if (!packed_) throw std::runtime_error("ClothBatch::solve: call assemble() first");
    batch_.solve(true, cuda_stream);
}

std::vector<std::pair<int, int>> ClothBatch::instance_sizes() const {
    std::vector<std::pair<int, int>> out;
    for (const ClothSimulator* s : sims_) out.emplace_back(s->buffers().particle_count(), s->pattern().nnz());
    return out;
}

std::vector<int> ClothBatch::last_pcg_iterations(std::uintptr_t cuda_stream) {
    std::vector<int> it;
    std::vector<double> rr, bb;
    if (packed_) batch_.read_stats(it, rr, bb, cuda_stream);
    return it;
}

std::vector<double> ClothBatch::last_relative_residuals(std::uintptr_t cuda_stream) {
    std::vector<int> it;
    std::vector<double> rr, bb, out;
    if (!packed_) return out;
    batch_.read_stats(it, rr, bb, cuda_stream);
    out.resize(rr.size());
    for (std::size_t i = 0; i < rr.size(); ++i) out[i] = bb[i] > 0.0 ? std::sqrt(rr[i] / bb[i]) : std::sqrt(rr[i]);
    return out;
}

}  // namespace cloth
}  // namespace chysx
