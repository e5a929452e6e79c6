// Conversions between the host layer's records (std::variant links, std::optional fluxes) and the
// flat C-ABI records, shared by the owning operators and the bounded workspaces.
#pragma once

#include <cstddef>
#include <optional>
#include <type_traits>
#include <variant>

#include "fluidity/host/liquid_boundary_geometry_3d.h"
#include "fluidity/host/mac_pressure_3d.h"
#include "fluidity_hip.h"

namespace fluidity::internal {

inline fl_int3 ToC3(CellIndex3D c) { return fl_int3{c.x, c.y, c.z}; }

inline fl_pressure_cell ToC(const MacPressureCell3D& c) {
  fl_pressure_cell r{};
  r.cell_index = ToC3(c.cell_index);
  r.fluid_volume = c.fluid_volume;
  r.target_divergence = c.target_divergence;
  r.bulk_modulus = c.bulk_modulus;
  r.previous_pressure = c.previous_pressure;
  r.integrated_inward_wall_flux = c.integrated_inward_wall_flux;
  return r;
}

inline MacPressureCell3D FromC(const fl_pressure_cell& c) {
  return {{c.cell_index.x, c.cell_index.y, c.cell_index.z}, c.fluid_volume, c.target_divergence,
          c.bulk_modulus, c.previous_pressure, c.integrated_inward_wall_flux};
}

// The link variant flattens to {kind, lift}: 1 internal (lift = jump), 2 Dirichlet (lift = exterior
// pressure), 0 sealed.
inline fl_pressure_face ToC(const MacPressureFace3D& f) {
  fl_pressure_face r{};
  r.axis = static_cast<std::int32_t>(f.face_index.axis);
  r.lattice = {f.face_index.lattice_index.x, f.face_index.lattice_index.y, f.face_index.lattice_index.z};
  r.open_area = f.open_area;
  r.inverse_density = f.inverse_density;
  r.pressure_gradient_distance = f.pressure_gradient_distance;
  r.predicted_normal_velocity = f.predicted_normal_velocity;
  r.wall_normal_velocity = f.wall_normal_velocity;
  std::visit(
      [&r](const auto& link) {
        using T = std::decay_t<decltype(link)>;
        if constexpr (std::is_same_v<T, InternalMacPressureLink3D>) {
          r.link_kind = 1;
          r.lift = link.pressure_jump;
        } else if constexpr (std::is_same_v<T, DirichletMacPressureLink3D>) {
          r.link_kind = 2;
          r.lift = link.exterior_pressure;
        } else {
          r.link_kind = 0;
          r.lift = 0.0;
        }
      },
      f.pressure_link);
  r.has_prescribed_wall_flux = f.prescribed_wall_flux.has_value() ? 1 : 0;
  r.prescribed_wall_flux = f.prescribed_wall_flux.value_or(0.0);
  return r;
}

inline MacPressureFace3D FromC(const fl_pressure_face& r) {
  MacPressureFace3D f{};
  f.face_index = {static_cast<FaceNormalAxis>(r.axis), {r.lattice.x, r.lattice.y, r.lattice.z}};
  f.open_area = r.open_area;
  f.inverse_density = r.inverse_density;
  f.pressure_gradient_distance = r.pressure_gradient_distance;
  f.predicted_normal_velocity = r.predicted_normal_velocity;
  f.wall_normal_velocity = r.wall_normal_velocity;
  if (r.link_kind == 1) f.pressure_link = InternalMacPressureLink3D{r.lift};
  else if (r.link_kind == 2) f.pressure_link = DirichletMacPressureLink3D{r.lift};
  else f.pressure_link = SealedMacPressureLink3D{};
  if (r.has_prescribed_wall_flux) f.prescribed_wall_flux = r.prescribed_wall_flux;
  return f;
}

inline fl_center_distance ToC(const LiquidCenterDistance3D& c) {
  return {ToC3(c.cell_index), 0, c.signed_distance};
}
inline fl_center_distance ToC(const LiquidCornerDistance3D& c) {
  return {ToC3(c.corner_index), 0, c.signed_distance};
}
inline fl_frozen_cell ToC(const FrozenLiquidCell3D& c) {
  return {ToC3(c.cell_index), 0, c.fluid_volume, c.center_distance, c.curvature};
}
inline FrozenLiquidCell3D FromC(const fl_frozen_cell& c) {
  return {{c.cell_index.x, c.cell_index.y, c.cell_index.z}, c.fluid_volume, c.center_distance, c.curvature};
}
inline fl_frozen_face ToC(const FrozenLiquidFace3D& x) {
  fl_frozen_face r{};
  r.axis = static_cast<std::uint8_t>(x.face_index.axis);
  r.lattice = {x.face_index.lattice_index.x, x.face_index.lattice_index.y, x.face_index.lattice_index.z};
  r.face_slot = x.face_slot;
  r.open_area = x.open_area;
  r.blocked_area = x.blocked_area;
  r.wall_normal_velocity = x.wall_normal_velocity;
  r.has_prescribed_wall_flux = x.prescribed_wall_flux.has_value();
  r.prescribed_wall_flux = x.prescribed_wall_flux.value_or(0.0);
  r.is_low_x_piston_face = x.is_low_x_piston_face;
  return r;
}
inline FrozenLiquidFace3D FromC(const fl_frozen_face& f) {
  FrozenLiquidFace3D r{{static_cast<FaceNormalAxis>(f.axis), {f.lattice.x, f.lattice.y, f.lattice.z}},
                       static_cast<std::size_t>(f.face_slot),
                       f.open_area,
                       f.blocked_area,
                       f.wall_normal_velocity};
  if (f.has_prescribed_wall_flux) r.prescribed_wall_flux = f.prescribed_wall_flux;
  r.is_low_x_piston_face = f.is_low_x_piston_face != 0;
  return r;
}
inline fl_cell_fields ToC(const LiquidPressureCellFields3D& c) {
  fl_cell_fields r{};
  r.target_divergence = c.target_divergence;
  r.bulk_modulus = c.bulk_modulus;
  r.previous_pressure = c.previous_pressure;
  r.has_density = c.density.has_value();
  r.density = c.density.value_or(0.0);
  return r;
}
inline fl_optional_flux ToC(const std::optional<double>& w) {
  fl_optional_flux r{};
  r.has_value = w.has_value();
  r.value = w.value_or(0.0);
  return r;
}
static_assert(sizeof(SolidCornerDistance3D) == sizeof(fl_solid_corner) &&
              offsetof(SolidCornerDistance3D, signed_distance) == offsetof(fl_solid_corner, signed_distance) &&
              offsetof(SolidCornerDistance3D, wall_velocity) == offsetof(fl_solid_corner, wall_velocity));
// The sampled-solid members of the freeze inputs (none: the plane-obstacle freeze).
inline void SetSampledSolid(const std::optional<SampledSolidBoundary3D>& solid, fl_freeze_inputs& r) {
  if (!solid) return;
  r.has_sampled_solid = 1;
  r.solid_corners = reinterpret_cast<const fl_solid_corner*>(solid->corner_distances.data());
  r.solid_corner_count = solid->corner_distances.size();
  r.face_normal_velocities = solid->face_normal_velocities.data();
  r.face_velocity_count = solid->face_normal_velocities.size();
}
inline fl_air_pressure ToC(const LiquidAirCellPressure3D& a) { return {ToC3(a.cell_index), 0, a.gas_gauge_pressure}; }
inline fl_internal_interface ToC(const LiquidInternalInterface3D& i) {
  return {i.face_ordinal, i.tail_fraction, i.surface_tension, i.curvature_tail_to_head,
          i.normal_viscous_jump_tail_to_head};
}

// The frozen-geometry obstacle record; a tank keeps its kind and its walls.
inline fl_obstacle ToC(const LiquidObstacle3D& o) {
  const auto c3 = [](Vec3 v) { return fl_vec3{v.x, v.y, v.z}; };
  fl_obstacle r{};
  if (const auto* h = std::get_if<TranslatingHalfspace3D>(&o)) {
    r = {FL_OBSTACLE_HALFSPACE, 0, c3(h->unit_normal), h->start_plane_coordinate, 0.0, c3(h->wall_velocity), {}};
  } else if (const auto* b = std::get_if<TranslatingSlab3D>(&o)) {
    r = {FL_OBSTACLE_SLAB, 0, c3(b->unit_normal), b->start_center_coordinate, b->half_width, c3(b->wall_velocity),
         {}};
  } else if (const auto* k = std::get_if<AxisAlignedLiquidTank3D>(&o)) {
    r = {FL_OBSTACLE_TANK, 0, {}, 0.0, 0.0, {}, {k->start_low_x, k->high_x, k->floor_y, k->low_z, k->high_z,
                                                  k->piston_velocity_x}};
  }
  return r;
}

}  // namespace fluidity::internal
