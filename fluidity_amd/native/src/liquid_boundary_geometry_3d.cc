// Frozen geometry and typed pressure links over the C-ABI (fl_freeze_geometry,
// fl_build_pressure_links). The records carry std::optional members, so they are converted field by
// field to the C-ABI records rather than reinterpreted.
#include "fluidity/host/liquid_boundary_geometry_3d.h"

#include <cstring>
#include <mutex>

#include "internal.h"

namespace fluidity {
namespace {

struct Staged {
  std::vector<fl_center_distance> centers;
  std::vector<fl_frozen_cell> cells;
  std::vector<fl_frozen_face> faces;
  std::vector<fl_cell_fields> fields;
  std::vector<fl_optional_flux> wetted;
  std::vector<fl_air_pressure> air;
  std::vector<fl_internal_interface> interfaces;
};

void Stage(const FrozenLiquidBoundaryGeometryView3D& g, const LiquidPressureFields3D& f, Staged& s) {
  for (const auto& c : g.center_distances)
    s.centers.push_back({{c.cell_index.x, c.cell_index.y, c.cell_index.z}, 0, c.signed_distance});
  for (const auto& c : g.cells)
    s.cells.push_back(
        {{c.cell_index.x, c.cell_index.y, c.cell_index.z}, 0, c.fluid_volume, c.center_distance, c.curvature});
  for (const auto& x : g.faces) {
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
    s.faces.push_back(r);
  }
  for (const auto& c : f.cells) {
    fl_cell_fields r{};
    r.target_divergence = c.target_divergence;
    r.bulk_modulus = c.bulk_modulus;
    r.previous_pressure = c.previous_pressure;
    r.has_density = c.density.has_value();
    r.density = c.density.value_or(0.0);
    s.fields.push_back(r);
  }
  for (const auto& w : f.wetted_piston_face_fluxes) {
    fl_optional_flux r{};
    r.has_value = w.has_value();
    r.value = w.value_or(0.0);
    s.wetted.push_back(r);
  }
  for (const auto& a : f.air_cell_pressures)
    s.air.push_back({{a.cell_index.x, a.cell_index.y, a.cell_index.z}, 0, a.gas_gauge_pressure});
  for (const auto& i : f.internal_interfaces)
    s.interfaces.push_back({i.face_ordinal, i.tail_fraction, i.surface_tension, i.curvature_tail_to_head,
                            i.normal_viscous_jump_tail_to_head});
}

MacPressureFace3D FromC(const fl_pressure_face& r) {
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

int Build(HipOperatorRuntime3D& rt, const FrozenLiquidBoundaryGeometryView3D& g, const LiquidPressureFields3D& f,
          LiquidPressureProblem3D& out, fl_measurement* m) {
  Staged s;
  Stage(g, f, s);
  const auto blocks = g.mac_grid.active_block_indices();
  fl_frozen_view v{};
  v.blocks = internal::ToC(blocks);
  v.block_count = blocks.size();
  v.geometry = internal::ToC(g.mac_grid.geometry());
  v.center_distances = s.centers.data();
  v.center_count = s.centers.size();
  v.cells = s.cells.data();
  v.cell_count = s.cells.size();
  v.faces = s.faces.data();
  v.face_count = s.faces.size();
  v.step_duration = g.step_duration;
  v.liquid_density = g.liquid_density;
  v.surface_tension = g.surface_tension;
  v.capillary_step_bound = g.capillary_step_bound;
  fl_pressure_fields pf{};
  pf.cells = s.fields.data();
  pf.cell_count = s.fields.size();
  pf.predicted_face_velocities = f.predicted_face_velocities.data();
  pf.predicted_count = f.predicted_face_velocities.size();
  pf.ambient_pressure = f.ambient_pressure;
  pf.wetted_piston_face_fluxes = s.wetted.data();
  pf.wetted_count = s.wetted.size();
  pf.air_cell_pressures = s.air.data();
  pf.air_count = s.air.size();
  pf.internal_interfaces = s.interfaces.data();
  pf.interface_count = s.interfaces.size();
  std::vector<fl_pressure_cell> cells(s.cells.size());
  std::vector<fl_pressure_face> faces(s.faces.size());
  std::uint64_t lower = 0;
  const int code = fl_build_pressure_links(rt.handle(), &v, &pf, cells.data(), faces.data(), &lower, m);
  if (code != FL_OK) return code;
  out.cells.resize(cells.size());
  for (std::size_t i = 0; i < cells.size(); ++i) {
    const auto& c = cells[i];
    out.cells[i] = {{c.cell_index.x, c.cell_index.y, c.cell_index.z}, c.fluid_volume, c.target_divergence,
                    c.bulk_modulus, c.previous_pressure, c.integrated_inward_wall_flux};
  }
  out.faces.resize(faces.size());
  for (std::size_t i = 0; i < faces.size(); ++i) out.faces[i] = FromC(faces[i]);
  out.lower_theta_regularization_count = lower;
  return FL_OK;
}

}  // namespace

StatusOr<LiquidPressureProblem3D> BuildLiquidPressureLinks3D(const FrozenLiquidBoundaryGeometryView3D& geometry,
                                                            const LiquidPressureFields3D& fields) {
  HipOperatorCapacities3D need;
  need.blocks = std::max<std::size_t>(geometry.mac_grid.active_block_indices().size(), 1);
  std::unique_lock<std::mutex> lock;
  auto rt = SharedHipRuntime3D::Acquire(need, lock);
  if (!rt.ok()) return rt.status();
  LiquidPressureProblem3D out;
  const int code = Build(**rt, geometry, fields, out, nullptr);
  if (code != FL_OK) return FromCode(code);
  return out;
}

StatusOr<LiquidPressureProblem3D> HipBuildLiquidPressureLinks3D(HipOperatorRuntime3D& runtime,
                                                               const FrozenLiquidBoundaryGeometryView3D& geometry,
                                                               const LiquidPressureFields3D& fields,
                                                               HipOperatorMeasurement& measurement) {
  fl_measurement m{};
  LiquidPressureProblem3D out;
  const int code = Build(runtime, geometry, fields, out, &m);
  if (code != FL_OK) return FromCode(code);
  std::memcpy(measurement.architecture, m.architecture, sizeof(measurement.architecture));
  measurement.kernel_ms = m.kernel_ms;
  measurement.copy_ms = m.copy_ms;
  measurement.launches = m.launches;
  return out;
}

// Stages the inputs as C-ABI records, freezes on the device and converts the frozen records back.
struct FreezeAccess {
  static int Freeze(HipOperatorRuntime3D& rt, const LiquidBoundaryGeometryInputs3D& in, std::uint64_t max_cells,
                    std::uint64_t max_faces, FrozenLiquidBoundaryGeometry3D& out, fl_measurement* m) {
    std::vector<fl_int3> candidates;
    candidates.reserve(in.cell_indices.size());
    for (const auto& c : in.cell_indices) candidates.push_back({c.x, c.y, c.z});
    std::vector<fl_center_distance> centers, corners;
    centers.reserve(in.center_distances.size());
    corners.reserve(in.corner_distances.size());
    for (const auto& c : in.center_distances)
      centers.push_back({{c.cell_index.x, c.cell_index.y, c.cell_index.z}, 0, c.signed_distance});
    for (const auto& c : in.corner_distances)
      corners.push_back({{c.corner_index.x, c.corner_index.y, c.corner_index.z}, 0, c.signed_distance});
    fl_freeze_inputs r{};
    const auto blocks = in.mac_grid.active_block_indices();
    r.blocks = internal::ToC(blocks);
    r.block_count = blocks.size();
    r.geometry = internal::ToC(in.mac_grid.geometry());
    r.cell_indices = candidates.data();
    r.cell_count = candidates.size();
    r.center_distances = centers.data();
    r.center_count = centers.size();
    r.corner_distances = corners.data();
    r.corner_count = corners.size();
    const auto c3 = [](Vec3 v) { return fl_vec3{v.x, v.y, v.z}; };
    if (const auto* h = std::get_if<TranslatingHalfspace3D>(&in.obstacle)) {
      r.obstacle = {FL_OBSTACLE_HALFSPACE, 0, c3(h->unit_normal), h->start_plane_coordinate, 0.0,
                    c3(h->wall_velocity), {}};
    } else if (const auto* b = std::get_if<TranslatingSlab3D>(&in.obstacle)) {
      r.obstacle = {FL_OBSTACLE_SLAB, 0, c3(b->unit_normal), b->start_center_coordinate, b->half_width,
                    c3(b->wall_velocity), {}};
    } else if (std::holds_alternative<AxisAlignedLiquidTank3D>(in.obstacle)) {
      r.obstacle.kind = FL_OBSTACLE_TANK;  // refused by the device freeze
    }
    r.step_duration = in.step_duration;
    r.liquid_density = in.liquid_density;
    r.surface_tension = in.surface_tension;
    r.has_exact_curvature = in.exact_outward_interface_curvature.has_value();
    r.exact_curvature = in.exact_outward_interface_curvature.value_or(0.0);
    r.max_cells = max_cells;
    r.max_faces = max_faces;
    std::vector<fl_frozen_cell> cells(candidates.size());
    std::vector<fl_frozen_face> faces(6 * candidates.size());
    std::uint64_t nc = 0, nf = 0;
    double bound = 0;
    const int code = fl_freeze_geometry(rt.handle(), &r, cells.data(), &nc, faces.data(), &nf, &bound, m);
    if (code != FL_OK) return code;
    out.centers_.assign(in.center_distances.begin(), in.center_distances.end());
    out.cells_.resize(nc);
    for (std::size_t i = 0; i < nc; ++i) {
      const auto& c = cells[i];
      out.cells_[i] = {{c.cell_index.x, c.cell_index.y, c.cell_index.z}, c.fluid_volume, c.center_distance,
                       c.curvature};
    }
    out.faces_.resize(nf);
    for (std::size_t i = 0; i < nf; ++i) {
      const auto& f = faces[i];
      out.faces_[i] = {{static_cast<FaceNormalAxis>(f.axis), {f.lattice.x, f.lattice.y, f.lattice.z}},
                       static_cast<std::size_t>(f.face_slot), f.open_area, f.blocked_area, f.wall_normal_velocity};
    }
    out.step_duration_ = in.step_duration;
    out.liquid_density_ = in.liquid_density;
    out.surface_tension_ = in.surface_tension;
    out.capillary_step_bound_ = bound;
    return FL_OK;
  }
  static FrozenLiquidBoundaryGeometry3D Empty(const SparseMacGrid3D& grid) { return FrozenLiquidBoundaryGeometry3D(grid); }
};

StatusOr<FrozenLiquidBoundaryGeometry3D> FreezeLiquidBoundaryGeometry3D(const LiquidBoundaryGeometryInputs3D& inputs) {
  // The owning freeze refuses more than 4096 candidates before any other check (:1982-1984).
  if (inputs.cell_indices.size() > 4096) return Status(StatusCode::kResourceExhausted);
  HipOperatorCapacities3D need;
  need.blocks = std::max<std::size_t>(inputs.mac_grid.active_block_indices().size(), 1);
  std::unique_lock<std::mutex> lock;
  auto rt = SharedHipRuntime3D::Acquire(need, lock);
  if (!rt.ok()) return rt.status();
  auto out = FreezeAccess::Empty(inputs.mac_grid);
  // FreezeCore's candidate and face caps (liquid_boundary_geometry_3d.cc:780-782, 918-920).
  const int code = FreezeAccess::Freeze(**rt, inputs, 4096, 24576, out, nullptr);
  if (code != FL_OK) return FromCode(code);
  return out;
}

StatusOr<FrozenLiquidBoundaryGeometry3D> HipFreezeLiquidBoundaryGeometry3D(HipOperatorRuntime3D& runtime,
                                                                          const LiquidBoundaryGeometryInputs3D& inputs,
                                                                          HipOperatorMeasurement& measurement) {
  fl_measurement m{};
  auto out = FreezeAccess::Empty(inputs.mac_grid);
  const int code = FreezeAccess::Freeze(runtime, inputs, 0, 0, out, &m);
  if (code != FL_OK) return FromCode(code);
  std::memcpy(measurement.architecture, m.architecture, sizeof(measurement.architecture));
  measurement.kernel_ms = m.kernel_ms;
  measurement.copy_ms = m.copy_ms;
  measurement.launches = m.launches;
  return out;
}

}  // namespace fluidity
