// Frozen geometry and typed pressure links over the C-ABI (fl_freeze_geometry,
// fl_build_pressure_links). The records carry std::optional members, so they are converted field by
// field to the C-ABI records rather than reinterpreted.
#include "fluidity/host/liquid_boundary_geometry_3d.h"

#include <cstring>
#include <mutex>

#include "internal.h"
#include "records.h"

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
  for (const auto& c : g.center_distances) s.centers.push_back(internal::ToC(c));
  for (const auto& c : g.cells) s.cells.push_back(internal::ToC(c));
  for (const auto& x : g.faces) s.faces.push_back(internal::ToC(x));
  for (const auto& c : f.cells) s.fields.push_back(internal::ToC(c));
  for (const auto& w : f.wetted_piston_face_fluxes) s.wetted.push_back(internal::ToC(w));
  for (const auto& a : f.air_cell_pressures) s.air.push_back(internal::ToC(a));
  for (const auto& i : f.internal_interfaces) s.interfaces.push_back(internal::ToC(i));
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
  for (std::size_t i = 0; i < cells.size(); ++i) out.cells[i] = internal::FromC(cells[i]);
  out.faces.resize(faces.size());
  for (std::size_t i = 0; i < faces.size(); ++i) out.faces[i] = internal::FromC(faces[i]);
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
    for (const auto& c : in.cell_indices) candidates.push_back(internal::ToC3(c));
    std::vector<fl_center_distance> centers, corners;
    centers.reserve(in.center_distances.size());
    corners.reserve(in.corner_distances.size());
    for (const auto& c : in.center_distances) centers.push_back(internal::ToC(c));
    for (const auto& c : in.corner_distances) corners.push_back(internal::ToC(c));
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
    r.obstacle = internal::ToC(in.obstacle);
    r.step_duration = in.step_duration;
    r.liquid_density = in.liquid_density;
    r.surface_tension = in.surface_tension;
    r.has_exact_curvature = in.exact_outward_interface_curvature.has_value();
    r.exact_curvature = in.exact_outward_interface_curvature.value_or(0.0);
    r.max_cells = max_cells;
    r.max_faces = max_faces;
    internal::SetSampledSolid(in.sampled_solid, r);
    std::vector<fl_frozen_cell> cells(candidates.size());
    std::vector<fl_frozen_face> faces(6 * candidates.size());
    std::uint64_t nc = 0, nf = 0;
    double bound = 0;
    const int code = fl_freeze_geometry(rt.handle(), &r, cells.data(), &nc, faces.data(), &nf, &bound, m);
    if (code != FL_OK) return code;
    out.centers_.assign(in.center_distances.begin(), in.center_distances.end());
    out.cells_.resize(nc);
    for (std::size_t i = 0; i < nc; ++i) out.cells_[i] = internal::FromC(cells[i]);
    out.faces_.resize(nf);
    for (std::size_t i = 0; i < nf; ++i) out.faces_[i] = internal::FromC(faces[i]);
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
