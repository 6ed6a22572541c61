// chysx::cloth::ClothBatch — many cloth instances advanced by one fused
// substep (one element pass, one gather pass and one persistent PCG launch
// for all of them).  No counterpart in the reference, which steps one
// ClothSimulator per cloth (ChysX/scenes/cloth_freefall.cu:165-170).
//
// Each instance is a fully set-up ClothSimulator (buffers, mesh, element
// sets, pins, material, PCG knobs) whose device arrays stay where they are;
// the batch only holds a table of their descriptors and the work lists.
// Per-instance results are identical to stepping each simulator alone: the
// same kernels run on the same arrays and every reduction is per instance,
// over the same tiles, in the same order.  Instances can be added and
// removed between substeps (continuous batching) at the cost of rebuilding
// the table.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "cloth_simulator.h"
#include "engine.h"

namespace chysx {
namespace cloth {

class ClothBatch {
public:
    ClothBatch() = default;
    ClothBatch(const ClothBatch&) = delete;
    ClothBatch& operator=(const ClothBatch&) = delete;

    // Registers a simulator (kept by reference; it must outlive its
    // membership).  Returns its position.
    int add_instance(ClothSimulator& sim);
    // Removes a simulator (by identity); returns false when absent.
    bool remove_instance(ClothSimulator& sim);
    void clear() noexcept;
    // Builds every instance's device arrays and binds the table (done
    // lazily by step() after membership changes).
    void finalize(float dt = 1.0f / 60.0f, std::uintptr_t cuda_stream = 0);
    // One substep of every instance; asynchronous on the stream.
    void step(float dt, std::uintptr_t cuda_stream = 0);
    // The two halves of step(): element + gather assembly, then the
    // persistent PCG solve with the fused finalize.
    void assemble(float dt, std::uintptr_t cuda_stream = 0);
    void solve(std::uintptr_t cuda_stream = 0);

    int num_instances() const noexcept { return static_cast<int>(sims_.size()); }
    int num_rows() const noexcept;
    std::size_t device_bytes() const noexcept;
    // (vertices, off-diagonal blocks) of every instance.
    std::vector<std::pair<int, int>> instance_sizes() const;
    void set_lanes_per_row(int lanes) { runner_.set_lanes_per_row(lanes); }
    int lanes_per_row() const noexcept { return runner_.lanes_per_row(); }
    void set_grid_blocks(int blocks) { runner_.set_grid_blocks(blocks); }
    // Rows per PCG tile: a multiple of CHX_STAGE_ROWS (256).
    void set_tile_rows(int rows) { runner_.set_tile_rows(rows); }
    // SpMV form of the PCG launch: LDS-staged (default) or global gathers.
    // Substeps replayed from a hipGraph of the element pass, gather pass and
    // persistent PCG (captured on the first step and again after any change
    // of the batch or its launch shape); needs a non-null stream.
    void set_graph(bool on) { graph_ = on; }
    bool graph() const noexcept { return graph_; }
    int graph_captures() const noexcept { return runner_.graph_captures(); }
    void set_lds_staging(bool on) { runner_.set_lds_staging(on); }
    bool lds_staging() const noexcept { return runner_.lds_staging(); }
    bool entries16() const noexcept { return runner_.entries16(); }
    // Solver trace of the persistent PCG launch: capacity in barrier entries
    // (0: off); last_trace() returns {stamp in 100 MHz ticks, active
    // instances} per barrier of the last solve.
    void set_trace(int entries) { runner_.set_trace(entries); }
    std::vector<unsigned long long> last_trace(std::uintptr_t cuda_stream = 0) { return runner_.read_trace(cuda_stream); }
    int tile_rows() const noexcept { return runner_.tile_rows(); }
    // BF16 precision mode for the whole batch; every instance must have been
    // switched with ClothSimulator::set_bf16_blocks(true) as well.
    void set_bf16_blocks(bool on) {
        bf16_ = on;
        runner_.set_bf16(on);
    }
    std::vector<int> last_pcg_iterations(std::uintptr_t cuda_stream = 0);
    // Bound of each grid-barrier wait of the solver in 100 MHz ticks (0: one
    // second); an aborted solve is reported by the next step or status read.
    void set_barrier_timeout_ticks(unsigned ticks) { runner_.set_barrier_timeout_ticks(ticks); }
    // Persistent PCG form of the FP32 solve (ClothSimulator::set_pcg_form).
    void set_pcg_form(int form) { runner_.set_pcg_form(form); }
    int pcg_form() const noexcept { return runner_.pcg_form(); }
    std::vector<double> last_relative_residuals(std::uintptr_t cuda_stream = 0);

private:
    std::vector<chx_instance> descriptors(float dt, std::uintptr_t cuda_stream = 0);

    std::vector<ClothSimulator*> sims_;
    engine::BatchRunner runner_;
    bool bound_ = false;
    bool bf16_ = false;
    bool graph_ = false;
};

}  // namespace cloth
}  // namespace chysx
