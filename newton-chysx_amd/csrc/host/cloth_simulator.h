// chysx::cloth::ClothSimulator — drop-in for the reference class
// (ChysX/src/cloth/cloth_simulator.h:59-719), PCG branch.
//
// Same lifecycle and contracts: the material is copied, particle buffers are
// borrowed device pointers (AoS float3 positions / velocities, optional
// inverse mass), set-up builds constraint sets from the current positions,
// and step(dt, stream) runs one implicit-Euler substep asynchronously on the
// caller's stream.  The device work goes through the C ABI in chysx_hip.h.
//
// Static plane / box contact is included; out of scope here (the
// reference's self-collision, SDF / mesh contact, tetrahedral FEM and VBD
// branches are off by default there): the corresponding methods are absent,
// and set_solver_type(VBD) makes step() throw.
#pragma once

#include <cstdint>
#include <vector>

#include "chysx_common.h"
#include "constraints.h"
#include "engine.h"
#include "topology.h"

namespace chysx {
namespace collision {

// Static contact primitives (ChysX/src/collision/static_contact.h:82-110):
// plane dot(n, x) + d = 0 with unit n pointing to the free side; oriented
// box with centre, half extents and unit frame columns ex, ey, ez.
struct PlaneShape {
    math::Vec3f n;
    float d = 0.0f;
};
struct BoxShape {
    math::Vec3f center, half_ext, ex, ey, ez;
};

}  // namespace collision

namespace cloth {

// ChysX/src/cloth/cloth_material.h:18-43.
struct ClothMaterial {
    float lame_mu = 0.0f;
    float lame_lambda = 0.0f;
    float bending = 0.0f;
    float density = 0.05f;
    float damping = 0.0f;
    float gx = 0.0f;
    float gy = 0.0f;
    float gz = -9.81f;
};

// ChysX/src/cloth/cloth_buffers.h:31-70.
struct ClothBuffers {
    DeviceSpan<math::Vec3f> pos;
    DeviceSpan<math::Vec3f> vel;
    DeviceSpan<float> inv_mass;
    DeviceSpan<math::Vec3f> ext_force;
    std::vector<math::Vec3i> tris;
    std::vector<math::Vec2i> edges;
    int particle_count() const noexcept { return static_cast<int>(pos.size()); }
};

enum class SolverType { PCG = 0, VBD = 1 };

class ClothSimulator {
public:
    ClothSimulator() = default;
    ClothSimulator(const ClothSimulator&) = delete;
    ClothSimulator& operator=(const ClothSimulator&) = delete;
    // Move-only, as the reference (cloth_simulator.h:72-75).  A simulator
    // registered with a ClothBatch must not be moved while it is a member.
    ClothSimulator(ClothSimulator&&) = default;
    ClothSimulator& operator=(ClothSimulator&&) = default;

    // ---- material ------------------------------------------------------
    void set_material(const ClothMaterial& m) noexcept { material_ = m; }
    ClothMaterial& material() noexcept { return material_; }
    const ClothMaterial& material() const noexcept { return material_; }

    // ---- buffers -------------------------------------------------------
    void set_external_buffers(std::uintptr_t pos_ptr, std::uintptr_t vel_ptr, int particle_count,
                              std::uintptr_t inv_mass_ptr = 0, std::uintptr_t force_ptr = 0) noexcept;
    ClothBuffers& buffers() noexcept { return buffers_; }
    const ClothBuffers& buffers() const noexcept { return buffers_; }

    // ---- pins -----------------------------------------------------------
    void set_pins(const int* host_indices, const math::Vec3f* host_targets, int n, float stiffness = 1.0e6f);
    void clear_pins() noexcept;
    void update_pin_targets(const math::Vec3f* host_targets, int n_pins, std::uintptr_t cuda_stream = 0);
    constraint::PinConstraint& pins() noexcept { return pins_; }
    const constraint::PinConstraint& pins() const noexcept { return pins_; }

    // ---- static plane / box contact (cloth_simulator.h:383-455) ------------
    // Penalty k/2 max(0, h - d)^2 against the deepest primitive per particle
    // at x_n, Gauss-Newton diagonal block, optional lagged Coulomb friction;
    // diagonal and right-hand side only (no BSR change).  Active when the
    // stiffness is positive and a primitive is registered.
    void add_static_plane(const collision::PlaneShape& p);
    void add_static_box(const collision::BoxShape& b);
    void clear_static_shapes();
    void set_static_contact_thickness(float t) noexcept { contact_thickness_ = t; }
    float static_contact_thickness() const noexcept { return contact_thickness_; }
    void set_static_contact_stiffness(float k) noexcept { contact_stiffness_ = k; }
    float static_contact_stiffness() const noexcept { return contact_stiffness_; }
    void set_static_contact_friction(float mu) noexcept { contact_friction_ = mu; }
    float static_contact_friction() const noexcept { return contact_friction_; }
    void set_static_contact_friction_epsilon(float eps_u) noexcept { contact_friction_eps_ = eps_u; }
    float static_contact_friction_epsilon() const noexcept { return contact_friction_eps_; }
    int static_plane_count() const noexcept { return static_cast<int>(planes_.size()); }
    int static_box_count() const noexcept { return static_cast<int>(boxes_.size()); }

    // ---- mesh and elements --------------------------------------------------
    void set_mesh(const math::Vec3i* host_triangles, int n_triangles);
    void build_springs_from_current_positions(float stiffness, std::uintptr_t cuda_stream = 0);
    void build_fem_stretch_from_current_positions(float stiffness, std::uintptr_t cuda_stream = 0);
    void build_fem_shear_from_current_positions(float stiffness, std::uintptr_t cuda_stream = 0);
    void build_bending_from_current_positions(float stiffness, std::uintptr_t cuda_stream = 0);
    constraint::SpringConstraint& springs() noexcept { return springs_; }
    constraint::TriangleStretchConstraint& fem_stretch() noexcept { return fem_stretch_; }
    const constraint::TriangleStretchConstraint& fem_stretch() const noexcept { return fem_stretch_; }
    constraint::TriangleStretchConstraint& fem_shear() noexcept { return fem_shear_; }
    const constraint::TriangleStretchConstraint& fem_shear() const noexcept { return fem_shear_; }
    // StVK membrane (no reference counterpart; the material hooks are
    // ClothMaterial::lame_mu / lame_lambda): Lame parameters mu, lambda
    // [Pa] and thickness [m], rest shape from the current positions.
    // Replaces nothing: combine it with or use it instead of the BW98
    // stretch / shear sets.
    void build_fem_stvk_from_current_positions(float mu, float lambda, float thickness,
                                               std::uintptr_t cuda_stream = 0);
    constraint::StvkConstraint& fem_stvk() noexcept { return fem_stvk_; }
    const constraint::StvkConstraint& fem_stvk() const noexcept { return fem_stvk_; }
    constraint::BendingConstraint& bending() noexcept { return bending_; }
    const constraint::BendingConstraint& bending() const noexcept { return bending_; }

    void redistribute_mass_area_weighted(float surface_density, std::uintptr_t inv_mass_ptr, int particle_count,
                                         std::uintptr_t cuda_stream = 0);

    // ---- stepping --------------------------------------------------------------
    void step(float dt, std::uintptr_t cuda_stream = 0);

    // ---- solver knobs -------------------------------------------------------------
    void set_pcg_iterations(int max_iter) noexcept { pcg_max_iterations_ = (max_iter > 0) ? max_iter : 1; }
    int pcg_iterations() const noexcept { return pcg_max_iterations_; }
    void set_solver_type(SolverType t) noexcept { solver_type_ = t; }
    SolverType solver_type() const noexcept { return solver_type_; }
    // Additions: convergence tolerance (<= 0: the reference's fixed
    // iteration count), SpMV lanes per block row, and read-backs of the last
    // solve (synchronous).
    void set_pcg_tolerance(double tol) noexcept { pcg_tolerance_ = tol; }
    double pcg_tolerance() const noexcept { return pcg_tolerance_; }
    void set_lanes_per_row(int lanes);
    // Rows per PCG tile (a multiple of 256) and persistent PCG workgroups
    // (0: every co-resident workgroup).
    void set_tile_rows(int rows) { runner_.set_tile_rows(rows); }
    // SpMV form of the PCG launch: LDS-staged (default) or global gathers.
    void set_lds_staging(bool on) { runner_.set_lds_staging(on); }
    // Solver trace of the persistent PCG launch: capacity in barrier entries
    // (0: off); last_trace() returns {stamp in 100 MHz ticks, active
    // instances} per barrier of the last solve.
    void set_trace(int entries) { runner_.set_trace(entries); }
    std::vector<unsigned long long> last_trace(std::uintptr_t cuda_stream = 0) { return runner_.read_trace(cuda_stream); }
    int tile_rows() const noexcept { return runner_.tile_rows(); }
    void set_grid_blocks(int blocks) { runner_.set_grid_blocks(blocks); }
    // {smallest, largest} column of each 256-row stage of the assembled
    // pattern (after the first step): the source rows its SpMV reads.
    const std::vector<int>& debug_stage_windows() const noexcept { return inst_.stage_windows(); }
    // Whether the SpMV reads 16-bit sliced entries (row-relative column and
    // block index) instead of 32-bit ones (after the first step).
    bool debug_entries16() const noexcept { return inst_.entries16(); }
    int lanes_per_row() const noexcept { return lanes_per_row_; }
    // Precision of the solve: false = FP32 Hessian blocks (default); true =
    // the inner CG runs on BF16 copies of the off-diagonal blocks and an FP32
    // diagonal that absorbs their rounding (FP32 accumulation); when its
    // residual meets the tolerance, the true residual (FP32 blocks, FP64
    // norms) decides: stop, or restart the inner CG from it.
    void set_bf16_blocks(bool on);
    bool bf16_blocks() const noexcept { return bf16_; }
    int last_pcg_iterations(std::uintptr_t cuda_stream = 0);
    // Bound of every grid-barrier wait of the persistent solver in 100 MHz
    // ticks (0: one second); a solve that exceeds it stops, leaves pos / vel
    // unchanged and is reported as a RuntimeError by the next step or status
    // read.
    void set_barrier_timeout_ticks(unsigned ticks) { runner_.set_barrier_timeout_ticks(ticks); }
    // Persistent PCG form of the FP32 solve: 0 (default) one grid barrier
    // per iteration, the iteration's two reductions merged; 1 two barriers
    // per iteration.  Same recurrence and stopping rule; BF16 mode always
    // uses the two-barrier form.
    void set_pcg_form(int form) { runner_.set_pcg_form(form); }
    int pcg_form() const noexcept { return runner_.pcg_form(); }
    double last_relative_residual(std::uintptr_t cuda_stream = 0);

    // ---- diagnostics -----------------------------------------------------------------
    int num_particles() const noexcept { return H_num_block_rows_; }
    int num_off_diag_blocks() const noexcept { return pattern_.nnz(); }
    void debug_copy_hessian_diag(float* out) const;
    void debug_copy_hessian_csr(int* out_row_offsets, int* out_col_indices, float* out_values) const;
    // BF16 mode: the inner matrix of the last step, row-major blocks
    // (compensated FP32 diagonal [n][3][3], BF16 bit patterns [nnz][3][3]).
    void debug_copy_bf16_matrix(float* out_diag_c, uint16_t* out_values16) const;
    void debug_copy_last_rhs(float* out) const;
    void debug_copy_last_dx(float* out) const;
    // Overwrites the PCG warm start (the previous solve's dx); n*3 floats.
    void debug_set_warm_start(const float* in);
    // Elastic energy at the current positions (FP64, synchronous);
    // kind_mask 1 pins, 2 triangles, 4 dihedrals.
    double compute_energy(int kind_mask = 15);
    const topology::Pattern& pattern() const noexcept { return pattern_; }

    // Host part of the topology build (pattern + slot tables).
    void prepare_topology();
    // Builds (or refreshes) this cloth's device arrays without stepping and
    // returns the launch descriptor for substep `dt`; ClothBatch uses it to
    // advance this cloth as one instance of a batch.
    chx_instance prepare_instance(float dt, std::uintptr_t cuda_stream = 0);
    // Builds the Hessian pattern, scatter maps and device arrays now (the
    // first step would otherwise do it), so the set-up of a cloth can run
    // ahead of its first step, e.g. on another host thread.
    // With a stream, every set-up upload and wait is on that stream only.
    void prepare(float dt, std::uintptr_t cuda_stream = 0) { (void)prepare_instance(dt, cuda_stream); }
    std::size_t device_bytes() const noexcept { return inst_.device_bytes() + dx_.gpu_size() * sizeof(math::Vec3f); }

private:
    void resize_work_buffers(int n, std::uintptr_t stream = 0);
    void ensure_hessian_topology(std::uintptr_t stream = 0);
    engine::InstanceSource source() const;
    void check_elements_current(std::uintptr_t stream = 0);

    ClothMaterial material_;
    ClothBuffers buffers_;
    constraint::PinConstraint pins_;
    constraint::SpringConstraint springs_;
    constraint::TriangleStretchConstraint fem_stretch_;
    constraint::TriangleStretchConstraint fem_shear_;
    constraint::StvkConstraint fem_stvk_;
    float packed_stvk_[3] = {0.f, 0.f, 0.f};
    constraint::BendingConstraint bending_;

    DeviceArray<math::Vec3f> dx_;
    topology::Pattern pattern_;
    engine::InstanceData inst_;
    engine::BatchRunner runner_;
    int H_num_block_rows_ = 0;
    bool topology_dirty_ = true;
    std::vector<chx_plane> planes_;
    std::vector<chx_box> boxes_;
    bool shapes_dirty_ = false;
    float contact_thickness_ = 0.0f, contact_stiffness_ = 0.0f;
    float contact_friction_ = 0.0f, contact_friction_eps_ = 1.0e-4f;
    float packed_k_[4] = {0.f, 0.f, 0.f, 0.f};
    int pcg_max_iterations_ = 50;
    double pcg_tolerance_ = 0.0;
    int lanes_per_row_ = 8;
    bool bf16_ = false;
    SolverType solver_type_ = SolverType::PCG;
};

}  // namespace cloth
}  // namespace chysx
