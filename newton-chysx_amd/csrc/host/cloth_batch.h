// chysx::cloth::ClothBatch — many cloth instances advanced by one fused
// substep (one element pass, one gather pass and one persistent PCG launch
// for all of them).  No counterpart in the reference, which steps one
// ClothSimulator per cloth (ChysX/scenes/cloth_freefall.cu:165-170).
//
// Each instance is a fully set-up ClothSimulator (buffers, mesh, element
// sets, pins, material, PCG knobs); the batch packs them block-diagonally.
// Per-instance results are identical to stepping that simulator alone:
// every reduction is per instance and in a fixed order.
#pragma once

#include <cstdint>
#include <vector>

#include "cloth_simulator.h"
#include "packed_batch.h"

namespace chysx {
namespace cloth {

class ClothBatch {
public:
    ClothBatch() = default;
    ClothBatch(const ClothBatch&) = delete;
    ClothBatch& operator=(const ClothBatch&) = delete;

    // Registers a simulator (kept by reference; it must outlive the batch).
    // Returns its instance index.  Invalidates the packing.
    int add_instance(ClothSimulator& sim);
    // Builds every instance's topology and packs the batch (done lazily by
    // step() when needed).
    void finalize();
    // One substep of every instance; asynchronous on the stream.
    void step(float dt, std::uintptr_t cuda_stream = 0);
    // The two halves of step(): element + gather assembly, then the
    // persistent PCG solve with the fused finalize.
    void assemble(float dt, std::uintptr_t cuda_stream = 0);
    void solve(std::uintptr_t cuda_stream = 0);
    // (vertices, off-diagonal blocks) of every instance.
    std::vector<std::pair<int, int>> instance_sizes() const;

    int num_instances() const noexcept { return static_cast<int>(sims_.size()); }
    int num_rows() const noexcept { return packed_ ? batch_.n_rows() : 0; }
    std::size_t device_bytes() const noexcept { return packed_ ? batch_.device_bytes() : 0; }
    void set_lanes_per_row(int lanes);
    void set_grid_blocks(int blocks);
    std::vector<int> last_pcg_iterations(std::uintptr_t cuda_stream = 0);
    std::vector<double> last_relative_residuals(std::uintptr_t cuda_stream = 0);
    engine::PackedBatch& packed() noexcept { return batch_; }

private:
    void prepare_step(float dt);

    std::vector<ClothSimulator*> sims_;
    engine::PackedBatch batch_;
    bool packed_ = false;
    int lanes_ = 8;
    int grid_blocks_ = 0;
};

}  // namespace cloth
}  // namespace chysx
