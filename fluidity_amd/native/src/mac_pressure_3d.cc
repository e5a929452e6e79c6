// ProjectMacPressure3D over the C-ABI: flattens the face link variant and wall-flux optional.
#include "fluidity/host/mac_pressure_3d.h"

#include <algorithm>
#include <cstring>
#include <type_traits>

#include "internal.h"
#include "records.h"

namespace fluidity {
namespace {

int Project(HipOperatorRuntime3D& rt, const MacPressureInputs3D& in, MacPressureSolverOptions3D options,
            const HipMacPressureSolverSettings3D& settings, HipMacPressureProjection3D& out) {
  const std::size_t nc = in.pressure_cells.size(), nf = in.pressure_faces.size();
  std::vector<fl_pressure_cell> cells(nc);
  std::vector<fl_pressure_face> faces(nf);
  std::transform(in.pressure_cells.begin(), in.pressure_cells.end(), cells.begin(),
                 [](const MacPressureCell3D& c) { return internal::ToC(c); });
  std::transform(in.pressure_faces.begin(), in.pressure_faces.end(), faces.begin(),
                 [](const MacPressureFace3D& f) { return internal::ToC(f); });
  fl_pressure_options o{};
  o.max_iterations = options.max_iterations;
  o.cpu_residual_tolerance_scale = options.cpu_residual_tolerance_scale;
  o.relative_divergence_target = settings.relative_divergence_target;
  o.levels = settings.levels;
  o.smoothing_sweeps = settings.smoothing_sweeps;
  o.coarse_sweeps = settings.coarse_sweeps;
  o.level_storage = settings.bf16_levels ? 2 : 1;
  o.preconditioner = settings.learned ? 1 : 0;
  o.smoother = settings.smoother;
  o.coarse_scale = settings.coarse_scale;
  MacPressureProjection3D& p = out.projection;
  p.cell_pressures.assign(nc, 0.0);
  p.corrected_face_velocities.assign(nf, 0.0);
  p.integrated_rhs.assign(nc, 0.0);
  p.true_residuals.assign(nc, 0.0);
  p.cell_constraint_defects.assign(nc, 0.0);
  std::vector<std::uint64_t> gauges(std::max<std::size_t>(nc, 1));
  fl_pressure_diagnostics d{};
  fl_measurement m{};
  const fl_geometry g = internal::ToC(in.mac_grid.geometry());
  const auto blocks = in.mac_grid.active_block_indices();
  const int code = fl_project(rt.handle(), &g, internal::ToC(blocks), blocks.size(), in.step_duration, cells.data(),
                              nc, faces.data(), nf, &o, p.cell_pressures.data(), p.corrected_face_velocities.data(),
                              p.integrated_rhs.data(), p.true_residuals.data(), p.cell_constraint_defects.data(),
                              gauges.data(), &d, &m);
  if (code != FL_OK) return code;
  p.gauge_cell_ordinals.assign(gauges.begin(), gauges.begin() + static_cast<std::ptrdiff_t>(d.gauge_count));
  p.diagnostics = {d.iteration_count, d.multigrid_level_count, d.true_residual_l2, d.max_volume_scaled_residual,
                   d.max_cell_constraint_defect};
  out.device_diagnostics = {d.rhs_l2, d.initial_defect_l2, d.defect_l2, d.defect_linf, d.restarts,
                            d.learned_iterations};
  std::memcpy(out.measurement.architecture, m.architecture, sizeof(out.measurement.architecture));
  out.measurement.kernel_ms = m.kernel_ms;
  out.measurement.copy_ms = m.copy_ms;
  out.measurement.launches = m.launches;
  return FL_OK;
}

HipOperatorCapacities3D ProjectionNeed(const MacPressureInputs3D& in) {
  HipOperatorCapacities3D need;
  need.blocks = std::max<std::size_t>(in.mac_grid.active_block_indices().size(), 1);
  need.lookup_blocks = internal::LookupNeed(in.mac_grid.active_block_indices());
  need.pressure_cells = std::max<std::size_t>(in.pressure_cells.size(), 1);
  need.pressure_faces = std::max<std::size_t>(in.pressure_faces.size(), 1);
  return need;
}

}  // namespace

StatusOr<MacPressureProjection3D> ProjectMacPressure3D(const MacPressureInputs3D& inputs,
                                                       MacPressureSolverOptions3D options) {
  std::unique_lock<std::mutex> lock;
  auto rt = SharedHipRuntime3D::Acquire(ProjectionNeed(inputs), lock);
  if (!rt.ok()) return rt.status();
  HipMacPressureProjection3D out;
  const int code = Project(**rt, inputs, options, HipMacPressureSolverSettings3D{}, out);
  if (code != FL_OK) return FromCode(code);
  return std::move(out.projection);
}

StatusOr<HipMacPressureProjection3D> HipProjectMacPressure3D(HipOperatorRuntime3D& runtime,
                                                             const MacPressureInputs3D& inputs,
                                                             MacPressureSolverOptions3D options,
                                                             HipMacPressureSolverSettings3D settings) {
  HipMacPressureProjection3D out;
  const int code = Project(runtime, inputs, options, settings, out);
  if (code != FL_OK) return FromCode(code);
  return out;
}

}  // namespace fluidity
