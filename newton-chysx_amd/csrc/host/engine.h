// Device-side state of the cloth step, one cloth instance at a time.
//
// InstanceData owns every device array of one cloth: its BSR pattern (local
// indices, exactly the reference's), element arrays, per-step element
// records, the gather map that inverts the reference's per-element slot
// tables (constraint_n.h:115-153) into per-slot contribution lists, and the
// PCG work vectors.  BatchRunner owns what a launch needs besides: the
// single-instance descriptor or a device table of instance descriptors, the
// work lists, per-tile partial sums, per-instance solver outputs and the
// grid-barrier area.  A batch of cloths is therefore a table over instances
// that already live in HBM; admitting or retiring one rebuilds only the table.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <vector>

#include "chysx_common.h"
#include "constraints.h"
#include "topology.h"

namespace chysx {
namespace engine {

struct InstanceSource {
    int n = 0;
    float* pos = nullptr;
    float* vel = nullptr;
    const float* inv_mass = nullptr;
    float* x = nullptr;  // warm start / solution (dx), owned by the caller
    const topology::Pattern* pattern = nullptr;
    const constraint::PinConstraint* pins = nullptr;
    const constraint::TriangleStretchConstraint* stretch = nullptr;
    const constraint::TriangleStretchConstraint* shear = nullptr;
    const constraint::StvkConstraint* stvk = nullptr;
    const constraint::BendingConstraint* bending = nullptr;
    float dt = 1.0f / 60.0f;
    float damping = 0.0f;
    float gx = 0.0f, gy = 0.0f, gz = -9.81f;
    int max_iter = 50;
    double tol = 0.0;
    bool bf16_blocks = false;  // keep BF16 copies of the Hessian blocks
    // static contact primitives and parameters (active when stiffness > 0)
    const std::vector<chx_plane>* planes = nullptr;
    const std::vector<chx_box>* boxes = nullptr;
    float contact_thickness = 0.0f, contact_stiffness = 0.0f;
    float contact_friction = 0.0f, contact_friction_eps = 1.0e-4f;
};

class InstanceData {
public:
    InstanceData() = default;
    InstanceData(const InstanceData&) = delete;
    InstanceData& operator=(const InstanceData&) = delete;
    InstanceData(InstanceData&&) = default;
    InstanceData& operator=(InstanceData&&) = default;

    // Builds every device array (slot tables must be bound to s.pattern).
    // Uploads run on `stream`, which is synchronized (not the device), so
    // an instance can be set up while other streams run batches.
    void build(const InstanceSource& s, std::uintptr_t stream = 0);
    // Re-uploads the element arrays (a stiffness changed).
    void update_elements(const InstanceSource& s, std::uintptr_t stream = 0);
    // Uploads new pin targets (n AoS float3) on `stream` without waiting
    // (cloth_simulator.cu:294-313): the element pass reads them from the
    // instance's target array, so queued and captured substeps that run
    // after the copy see them.
    void upload_pin_targets(const float* host_xyz, int n, std::uintptr_t stream);
    // Re-uploads the static contact primitives (the shape list changed).
    void update_shapes(const InstanceSource& s, std::uintptr_t stream = 0);
    // Descriptor for a launch: device arrays + the step scalars of `s`.
    chx_instance descriptor(const InstanceSource& s) const;
    bool built() const noexcept { return built_; }
    int n_rows() const noexcept { return n_; }
    int nnz() const noexcept { return nnz_; }
    int n_upper() const noexcept { return nup_; }
    // Value slot of CSR block k, -1 - (partner's slot) for a lower block
    // (host copy).
    const std::vector<int>& block_slots() const noexcept { return block_slots_.host(); }
    std::size_t device_bytes() const noexcept { return bytes_; }
    std::size_t contributions() const noexcept { return contribs_.cpu_size(); }
    // Assembly units (host copies): row ranges and element counts.
    const std::vector<chx_unit>& units() const noexcept { return units_.host(); }
    const float* diag() const noexcept { return diag_.gpu_data(); }
    const float* values() const noexcept { return values_.gpu_data(); }
    const float* diag_c() const noexcept { return diag_c_.gpu_data(); }
    const uint16_t* values16() const noexcept { return values16_.gpu_data(); }
    const float* rhs() const noexcept { return rhs_.gpu_data(); }
    const std::vector<int>& stage_windows() const noexcept { return windows_.host(); }
    // True when every sliced entry fits the 16-bit form (column within 2047
    // rows of its row, block index below 16).
    bool entries16() const noexcept { return entries16_; }

private:
    void pack_elements(const InstanceSource& s);
    void build_units(const std::vector<unsigned>& goff, const std::vector<unsigned>& gent);

    bool built_ = false;
    int n_ = 0, nnz_ = 0, npad_ = 0, nup_ = 0;
    std::size_t bytes_ = 0;
    DeviceArray<int> row_offsets_, col_indices_;
    DeviceArray<chx_pin> pins_;
    DeviceArray<float> pin_targets_;  // AoS float3 per pin
    HostStager stager_;
    DeviceArray<chx_tri> tris_;
    DeviceArray<chx_dih> dihs_;
    DeviceArray<chx_stvk> stvks_;
    DeviceArray<float> diag_, values_;
    DeviceArray<chx_unit> units_;
    DeviceArray<int> unit_elems_;
    DeviceArray<uint16_t> contrib_offsets_, contribs_;
    DeviceArray<float> rhs_, r_, z_, p0_, p1_, q_, q1_, minv_;
    DeviceArray<float> diag_c_;
    DeviceArray<uint16_t> values16_;
    DeviceArray<int> windows_;  // per CHX_STAGE_ROWS-row stage: {min column, max column}
    DeviceArray<int> slice_offsets_, slice_entries_, upper_offsets_, block_slots_;  // sliced symmetric layout
    DeviceArray<uint16_t> slice_entries16_;  // the entries in 16 bits, when every one fits (else empty)
    bool entries16_ = false;
    DeviceArray<chx_plane> planes_;
    DeviceArray<chx_box> boxes_;
};

class BatchRunner {
public:
    BatchRunner() = default;
    ~BatchRunner() { release(); }
    BatchRunner(const BatchRunner&) = delete;
    BatchRunner& operator=(const BatchRunner&) = delete;
    BatchRunner(BatchRunner&& o) noexcept { *this = std::move(o); }
    BatchRunner& operator=(BatchRunner&& o) noexcept;

    // One instance: its descriptor travels in the kernel arguments.
    void bind_single(const chx_instance& d, std::uintptr_t stream = 0);
    // Many instances: builds the work lists and uploads the table.
    void bind_many(const std::vector<chx_instance>& ds, std::uintptr_t stream = 0);
    // Per-step scalars / external pointers of the bound instances.
    void update_single(const chx_instance& d) {
        const int tb = desc_.single.tile_begin, te = desc_.single.tile_end;
        desc_.entries16 = (d.n_rows > 0 && d.slice_entries16 == nullptr) ? 0 : 1;
        desc_.single = d;
        desc_.single.tile_begin = tb;
        desc_.single.tile_end = te;
    }
    void update_many(const std::vector<chx_instance>& ds, std::uintptr_t stream);

    // Launches one substep.  Raises (after synchronising) when an earlier
    // solve of this runner aborted at a grid barrier.
    void step(std::uintptr_t stream);
    // The same substep replayed from a hipGraph, captured on first use and
    // again whenever the descriptor changed since the capture (`stream` must
    // not be the null stream).
    void step_graph(std::uintptr_t stream);
    void drop_graph() noexcept;
    int graph_captures() const noexcept { return graph_captures_; }
    void assemble(std::uintptr_t stream);
    void solve(bool finalize, std::uintptr_t stream);
    double energy(int kind_mask, std::uintptr_t stream);

    void set_lanes_per_row(int L);
    int lanes_per_row() const noexcept { return lanes_; }
    // Rows per PCG tile (a multiple of CHX_STAGE_ROWS); rebuilds the tile
    // list of a bound batch.
    void set_tile_rows(int rows);
    int tile_rows() const noexcept { return tile_rows_; }
    void set_grid_blocks(int g) {
        grid_ = g < 0 ? 0 : g;
        desc_.grid_blocks = grid_;
    }
    // SpMV form of the PCG launch: LDS-staged stages (default) or gathers
    // straight from global memory (smaller LDS, more workgroups per CU).
    void set_lds_staging(bool on) { desc_.no_lds_staging = on ? 0 : 1; }
    bool lds_staging() const noexcept { return desc_.no_lds_staging == 0; }
    // Whether the bound instances' SpMV reads 16-bit sliced entries (every
    // instance has them) or 32-bit ones.
    bool entries16() const noexcept { return desc_.entries16 != 0; }
    // Precision mode of the solve: FP32 blocks, or inner CG on the BF16
    // matrix with outer refinement on the FP32 true residual.
    void set_bf16(bool on) {
        bf16_ = on;
        desc_.bf16_blocks = on ? 1 : 0;
    }
    int n_instances() const noexcept { return desc_.n_instances; }
    int n_tiles() const noexcept { return desc_.n_tiles; }
    bool bound() const noexcept { return desc_.n_instances > 0; }
    const chx_cloth_batch& desc() const noexcept { return desc_; }

    // Solver trace: capacity in barrier entries (0 switches it off).
    void set_trace(int entries);
    // {stamp (100 MHz ticks), active instances} per barrier of the last
    // solve, up to the last one written.
    std::vector<unsigned long long> read_trace(std::uintptr_t stream);

    // Synchronous read-backs.
    void read_stats(std::vector<int>& iters, std::vector<double>& rr, std::vector<double>& bb, std::uintptr_t stream);
    void check_status(std::uintptr_t stream);
    // Bound of every grid-barrier wait of the persistent solver in 100 MHz
    // ticks (0: one second).  A solve whose wait exceeds it stops and is
    // reported by the next step or status read.
    void set_barrier_timeout_ticks(unsigned ticks) { desc_.wait_ticks = ticks; }
    // Solver form of FP32 instances: 0 one grid barrier per iteration
    // (default), 1 two barriers per iteration.
    void set_pcg_form(int form) {
        if (form != 0 && form != 1) throw std::invalid_argument("chysx::engine: pcg form must be 0 or 1");
        desc_.pcg_form = form;
    }
    int pcg_form() const noexcept { return desc_.pcg_form; }

private:
    void size_outputs(int n_instances, int n_tiles, std::uintptr_t stream = 0);
    // Raises when the host-mapped status word reports an aborted solve.
    void check_host_status(std::uintptr_t stream) {
        if (status_host_ != nullptr && *static_cast<volatile unsigned*>(status_host_) != 0u) check_status(stream);
    }
    void release() noexcept;
    // Tile list of the bound instances; sets their tile_begin / tile_end.
    void build_tiles(std::vector<chx_instance>& ds, std::uintptr_t stream = 0);

    chx_cloth_batch desc_{};
    void* graph_ = nullptr;           // hipGraphExec_t of one substep
    chx_cloth_batch graph_desc_{};    // the descriptor it was captured with
    int graph_captures_ = 0;
    int lanes_ = 8;
    int grid_ = 0;
    int tile_rows_ = CHX_TILE_ROWS;
    bool bf16_ = false;
    std::vector<chx_instance> table_host_;
    DeviceArray<chx_instance> table_;
    DeviceArray<int> tiles_, elem_items_, unit_items_;
    DeviceArray<double> partials_, inst_rr_, inst_bb_, inst_rho_;
    unsigned* status_host_ = nullptr;  // page-locked, device-mapped
    DeviceArray<int> inst_iters_, inst_done_;
    DeviceArray<unsigned> sync_;
    DeviceArray<unsigned long long> trace_;
};

}  // namespace engine
}  // namespace chysx
