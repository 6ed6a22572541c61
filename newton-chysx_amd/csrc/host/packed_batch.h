// Packing of one or more cloth instances into the device layout the
// substep kernels consume (chx_cloth_batch in chysx_hip.h): a
// block-diagonal BSR system whose per-instance blocks carry exactly the
// reference pattern, packed element arrays, row tiles, and the gather map
// that turns the reference's per-element slot tables (scatter maps,
// constraint_n.h:115-153) into per-slot contribution lists for the
// atomic-free assembly.
#pragma once

#include <cstdint>
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
    const topology::Pattern* pattern = nullptr;
    const constraint::PinConstraint* pins = nullptr;
    const constraint::TriangleStretchConstraint* stretch = nullptr;
    const constraint::TriangleStretchConstraint* shear = nullptr;
    const constraint::BendingConstraint* bending = nullptr;
    float dt = 1.0f / 60.0f;
    float damping = 0.0f;
    float gx = 0.0f, gy = 0.0f, gz = -9.81f;
    int max_iter = 50;
    double tol = 0.0;
};

constexpr int kTileRows = 256;  // rows per tile; the PCG kernel's workgroup covers one tile per pass

class PackedBatch {
public:
    PackedBatch() = default;
    PackedBatch(const PackedBatch&) = delete;
    PackedBatch& operator=(const PackedBatch&) = delete;

    // Packs `src` (slot tables must already be bound against each
    // instance's pattern).  `external_x`: warm-start buffer owned by the
    // caller (single instance only); otherwise the batch owns and zeroes it.
    void build(const std::vector<InstanceSource>& src, int lanes_per_row, float* external_x = nullptr);
    // Re-packs only the element arrays (stiffness or pin targets changed).
    void update_elements(const std::vector<InstanceSource>& src);
    // Instance parameters (pos/vel pointers, dt, gravity, damping, PCG
    // knobs) for the next step; uploads the table for multi-instance batches.
    void set_instance(int i, const InstanceSource& s);
    void flush_instances(std::uintptr_t stream);

    void step(std::uintptr_t stream);
    void assemble(std::uintptr_t stream);
    void solve(bool finalize, std::uintptr_t stream);

    int n_instances() const noexcept { return desc_.n_instances; }
    int n_rows() const noexcept { return desc_.n_rows; }
    int nnz() const noexcept { return desc_.nnz; }
    int row_begin(int i) const { return inst_host_[i].row_begin; }
    int nnz_begin(int i) const { return nnz_begin_[i]; }
    const chx_cloth_batch& desc() const noexcept { return desc_; }
    chx_cloth_batch& desc() noexcept { return desc_; }
    void set_lanes_per_row(int L) { desc_.lanes_per_row = L; }
    void set_grid_blocks(int g) { desc_.grid_blocks = g; }
    std::size_t device_bytes() const noexcept { return device_bytes_; }
    std::size_t contributions() const noexcept { return contribs_.cpu_size(); }

    // Synchronous read-backs.
    void read_stats(std::vector<int>& iters, std::vector<double>& rr, std::vector<double>& bb, std::uintptr_t stream);
    void check_status(std::uintptr_t stream);
    void copy_diag(float* out, int row_begin, int n) const;
    void copy_values(float* out, int nnz_begin, int nnz) const;
    void copy_rhs(float* out, int row_begin, int n) const;
    void copy_x(float* out, int row_begin, int n) const;
    void upload_x(const float* in, int row_begin, int n);

private:
    void pack_elements(const std::vector<InstanceSource>& src);

    chx_cloth_batch desc_{};
    std::vector<chx_instance> inst_host_;
    std::vector<int> nnz_begin_;
    bool inst_dirty_ = false;
    std::size_t device_bytes_ = 0;

    DeviceArray<chx_instance> inst_;
    DeviceArray<int> row_inst_, tiles_, row_offsets_, col_indices_;
    DeviceArray<chx_pin> pins_;
    DeviceArray<chx_tri> tris_;
    DeviceArray<chx_dih> dihs_;
    DeviceArray<float> pin_rec_, tri_rec_, dih_rec_, diag_, values_;
    DeviceArray<unsigned> contrib_offsets_, contribs_;
    DeviceArray<float> rhs_, x_, r_, z_, p0_, p1_, q_, minv_;
    DeviceArray<double> partials_, inst_rr_, inst_bb_;
    DeviceArray<int> inst_iters_, inst_done_;
    DeviceArray<unsigned> sync_;
};

}  // namespace engine
}  // namespace chysx
