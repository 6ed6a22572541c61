#include "constraints.h"

#include <cmath>
#include <cstring>
#include <stdexcept>

namespace chysx {
namespace constraint {

void PinConstraint::set_pins(const int* host_indices, const math::Vec3f* host_targets, int n) {
    if (n < 0) throw std::invalid_argument("PinConstraint::set_pins: negative count");
    indices_.assign(host_indices ? host_indices : nullptr, host_indices ? host_indices + n : nullptr);
    targets_.assign(host_targets ? host_targets : nullptr, host_targets ? host_targets + n : nullptr);
    if (static_cast<int>(indices_.size()) != n || static_cast<int>(targets_.size()) != n) {
        indices_.clear();
        targets_.clear();
        if (n > 0) throw std::invalid_argument("PinConstraint::set_pins: null indices or targets");
    }
    slots_.clear();
}

void TriangleStretchConstraint::set_triangles_from_positions(const math::Vec3i* host_triangles, int n,
                                                             DeviceSpan<math::Vec3f> positions, std::uintptr_t stream,
                                                             float material_rotation_rad) {
    if (n < 0) throw std::invalid_argument("TriangleStretchConstraint::set_triangles_from_positions: negative count");
    rotation_ = material_rotation_rad;
    indices_.assign(host_triangles, host_triangles + n);
    dm_inv_.assign(n, math::Mat2f{});
    areas_.assign(n, 0.0f);
    slots_.clear();
    if (n == 0) return;
    // cos/sin in float on the host (triangle_stretch_constraint.cu:341-342).
    const float c = std::cos(material_rotation_rad);
    const float s = std::sin(material_rotation_rad);
    DeviceArray<math::Vec3i> d_tris;
    d_tris.host() = indices_;
    d_tris.copy_to_device(stream);
    DeviceArray<math::Mat2f> d_dm;
    DeviceArray<float> d_area;
    d_dm.ensure_device(n);
    d_area.ensure_device(n);
    check(chx_triangle_rest_shape(reinterpret_cast<const int*>(d_tris.gpu_data()),
                                  reinterpret_cast<const float*>(positions.data()), n, c, s,
                                  reinterpret_cast<float*>(d_dm.gpu_data()), d_area.gpu_data(), stream),
          "constraint::TriangleStretchConstraint", "rest shape kernel");
    d_dm.copy_to_host(stream);
    d_area.copy_to_host(stream);
    dm_inv_ = d_dm.host();
    areas_ = d_area.host();
}

void BendingConstraint::set_dihedrals_from_positions(const math::Vec4i* host_dihedrals, int n,
                                                     DeviceSpan<math::Vec3f> positions, std::uintptr_t stream) {
    if (n < 0) throw std::invalid_argument("BendingConstraint::set_dihedrals_from_positions: negative count");
    indices_.assign(host_dihedrals, host_dihedrals + n);
    rest_.assign(n, 0.0f);
    slots_.clear();
    if (n == 0) return;
    DeviceArray<math::Vec4i> d_dih;
    d_dih.host() = indices_;
    d_dih.copy_to_device(stream);
    DeviceArray<float> d_rest;
    d_rest.ensure_device(n);
    check(chx_dihedral_rest_angle(reinterpret_cast<const int*>(d_dih.gpu_data()),
                                  reinterpret_cast<const float*>(positions.data()), n, d_rest.gpu_data(), stream),
          "constraint::BendingConstraint", "rest angle kernel");
    d_rest.copy_to_host(stream);
    rest_ = d_rest.host();
}

void StvkConstraint::set_material(float mu, float lambda, float thickness) {
    if (!(mu >= 0.0f) || !(lambda >= 0.0f) || !(thickness >= 0.0f))
        throw std::invalid_argument("StvkConstraint::set_material: mu, lambda and thickness must be non-negative");
    mu_ = mu;
    lambda_ = lambda;
    thickness_ = thickness;
}

void SpringConstraint::set_springs(const math::Vec2i* edges, int n) {
    if (n < 0) throw std::invalid_argument("SpringConstraint::set_springs: negative count");
    edges_.assign(edges ? edges : nullptr, edges ? edges + n : nullptr);
}

}  // namespace constraint
}  // namespace chysx
