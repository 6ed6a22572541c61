// Constraint sets of the cloth step (host side).  Each keeps its element
// indices and rest data on the host (needed to build the packed batch) and
// computes the rest data on the device from the current positions, as the
// reference does.  Interfaces follow
//   PinConstraint               ChysX/src/constraint/cloth/pin_constraint.h:40-91
//   TriangleStretchConstraint   ChysX/src/constraint/cloth/triangle_stretch_constraint.h:60-136
//   BendingConstraint           ChysX/src/constraint/cloth/bending_constraint.h
//   SpringConstraint            ChysX/src/constraint/cloth/spring_constraint.h (kept for the API;
//                               the PCG step never evaluates springs, cloth_simulator.cu:921)
// Slot tables (the reference's Hessian scatter maps,
// constraint_n.h:115-153) are bound against the BSR pattern by
// bind_hessian_layout().
#pragma once

#include <cstdint>
#include <vector>

#include "chysx_common.h"
#include "topology.h"

namespace chysx {
namespace constraint {

class PinConstraint {
public:
    void set_pins(const int* host_indices, const math::Vec3f* host_targets, int n);
    void set_stiffness(float k) noexcept { stiffness_ = k; }
    float stiffness() const noexcept { return stiffness_; }
    int size() const noexcept { return static_cast<int>(indices_.size()); }
    const std::vector<int>& indices() const noexcept { return indices_; }
    std::vector<math::Vec3f>& targets() noexcept { return targets_; }
    const std::vector<math::Vec3f>& targets() const noexcept { return targets_; }
    void bind_hessian_layout(const topology::Pattern& p) { slots_ = topology::slot_table_pins(p, indices_); }
    const std::vector<int>& hessian_slots() const noexcept { return slots_; }

private:
    std::vector<int> indices_;
    std::vector<math::Vec3f> targets_;
    std::vector<int> slots_;
    float stiffness_ = 1.0e6f;
};

class TriangleStretchConstraint {
public:
    // Rest shape from the current device positions; material axes rotated by
    // `material_rotation_rad` (0 = BW98 stretch, pi/4 = BW98 shear).
    void set_triangles_from_positions(const math::Vec3i* host_triangles, int n, DeviceSpan<math::Vec3f> positions,
                                      std::uintptr_t stream = 0, float material_rotation_rad = 0.0f);
    void set_stiffness(float k) noexcept { stiffness_ = k; }
    float stiffness() const noexcept { return stiffness_; }
    int size() const noexcept { return static_cast<int>(indices_.size()); }
    const std::vector<math::Vec3i>& indices() const noexcept { return indices_; }
    const std::vector<math::Mat2f>& Dm_inv() const noexcept { return dm_inv_; }
    const std::vector<float>& areas() const noexcept { return areas_; }
    float material_rotation() const noexcept { return rotation_; }
    void bind_hessian_layout(const topology::Pattern& p) { slots_ = topology::slot_table3(p, indices_); }
    const std::vector<int>& hessian_slots() const noexcept { return slots_; }

private:
    std::vector<math::Vec3i> indices_;
    std::vector<math::Mat2f> dm_inv_;
    std::vector<float> areas_;
    std::vector<int> slots_;
    float stiffness_ = 0.0f;
    float rotation_ = 0.0f;
};

// StVK membrane: no reference counterpart (the reference keeps
// ClothMaterial::lame_mu / lame_lambda as unused placeholders,
// cloth_material.h:93-116).  Rest shape as the BW98 stretch triangles
// (theta = 0); per element muA = mu * area * h, lamA = lambda * area * h.
class StvkConstraint {
public:
    void set_triangles_from_positions(const math::Vec3i* host_triangles, int n, DeviceSpan<math::Vec3f> positions,
                                      std::uintptr_t stream = 0) {
        shape_.set_triangles_from_positions(host_triangles, n, positions, stream, 0.0f);
        slots_.clear();
    }
    void set_material(float mu, float lambda, float thickness);
    float mu() const noexcept { return mu_; }
    float lambda() const noexcept { return lambda_; }
    float thickness() const noexcept { return thickness_; }
    int size() const noexcept { return shape_.size(); }
    const std::vector<math::Vec3i>& indices() const noexcept { return shape_.indices(); }
    const std::vector<math::Mat2f>& Dm_inv() const noexcept { return shape_.Dm_inv(); }
    const std::vector<float>& areas() const noexcept { return shape_.areas(); }
    void bind_hessian_layout(const topology::Pattern& p) { slots_ = topology::slot_table3(p, indices()); }
    const std::vector<int>& hessian_slots() const noexcept { return slots_; }

private:
    TriangleStretchConstraint shape_;
    std::vector<int> slots_;
    float mu_ = 0.0f, lambda_ = 0.0f, thickness_ = 0.0f;
};

class BendingConstraint {
public:
    void set_dihedrals_from_positions(const math::Vec4i* host_dihedrals, int n, DeviceSpan<math::Vec3f> positions,
                                      std::uintptr_t stream = 0);
    void set_stiffness(float k) noexcept { stiffness_ = k; }
    float stiffness() const noexcept { return stiffness_; }
    int size() const noexcept { return static_cast<int>(indices_.size()); }
    const std::vector<math::Vec4i>& indices() const noexcept { return indices_; }
    const std::vector<float>& rest_angles() const noexcept { return rest_; }
    void bind_hessian_layout(const topology::Pattern& p) { slots_ = topology::slot_table4(p, indices_); }
    const std::vector<int>& hessian_slots() const noexcept { return slots_; }

private:
    std::vector<math::Vec4i> indices_;
    std::vector<float> rest_;
    std::vector<int> slots_;
    float stiffness_ = 0.0f;
};

class SpringConstraint {
public:
    void set_springs(const math::Vec2i* edges, int n);
    void set_stiffness(float k) noexcept { stiffness_ = k; }
    float stiffness() const noexcept { return stiffness_; }
    int size() const noexcept { return static_cast<int>(edges_.size()); }

private:
    std::vector<math::Vec2i> edges_;
    float stiffness_ = 0.0f;
};

}  // namespace constraint
}  // namespace chysx
