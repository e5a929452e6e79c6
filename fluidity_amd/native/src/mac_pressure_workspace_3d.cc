// Bounded pressure workspace: fl_project into banks reserved at Create, published on success.
#include "fluidity/host/mac_pressure_workspace_3d.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <cstring>
#include <mutex>
#include <new>

#include "internal.h"
#include "records.h"

namespace fluidity {
namespace {

bool CellLess(CellIndex3D a, CellIndex3D b) {
  if (a.z != b.z) return a.z < b.z;
  if (a.y != b.y) return a.y < b.y;
  return a.x < b.x;
}

// Whether `c` is one of the (z,y,x)-sorted problem cells.
bool HasCell(std::span<const MacPressureCell3D> cells, CellIndex3D c) {
  const auto it = std::lower_bound(cells.begin(), cells.end(), c,
                                   [](const MacPressureCell3D& a, CellIndex3D k) { return CellLess(a.cell_index, k); });
  return it != cells.end() && !CellLess(c, it->cell_index);
}

}  // namespace

StatusOr<MacPressureWorkspace3D> MacPressureWorkspace3D::Create(std::size_t cell_capacity, std::size_t face_capacity) {
  MacPressureWorkspace3D w;
  w.cell_capacity_ = cell_capacity;
  w.face_capacity_ = face_capacity;
  try {
    for (auto& b : w.banks_) {
      for (auto* v : {&b.pressures, &b.rhs, &b.residuals, &b.defects}) v->reserve(cell_capacity);
      b.velocities.reserve(face_capacity);
      b.gauges.reserve(std::max<std::size_t>(cell_capacity, 1));
      b.gauge_ordinals.reserve(cell_capacity);
    }
    w.cells_.reserve(cell_capacity);
    w.faces_.reserve(face_capacity);
    w.material_.reserve(cell_capacity);
    w.rows_.reserve(6 * cell_capacity);
  } catch (const std::bad_alloc&) {
    return Status(StatusCode::kResourceExhausted);
  }
  return w;
}

StatusOr<MacPressureProjectionView3D> HipProjectMacPressure3D(HipOperatorRuntime3D& runtime, SparseMacGridView3D grid,
                                                              double step_duration,
                                                              const LiquidPressureProblemView3D& problem,
                                                              MacPressureSolverOptions3D options,
                                                              MacPressureWorkspace3D& w) {
  const std::size_t nc = problem.cells.size(), nf = problem.faces.size();
  if (nc > w.cell_capacity_ || nf > w.face_capacity_) return Status(StatusCode::kResourceExhausted);
  w.cells_.clear();
  w.faces_.clear();
  for (const auto& c : problem.cells) w.cells_.push_back(internal::ToC(c));
  for (const auto& f : problem.faces) w.faces_.push_back(internal::ToC(f));
  const HipMacPressureSolverSettings3D& st = w.settings_;
  fl_pressure_options o{};
  o.max_iterations = options.max_iterations;
  o.cpu_residual_tolerance_scale = options.cpu_residual_tolerance_scale;
  o.relative_divergence_target = st.relative_divergence_target;
  o.levels = st.levels;
  o.smoothing_sweeps = st.smoothing_sweeps;
  o.coarse_sweeps = st.coarse_sweeps;
  o.level_storage = st.bf16_levels ? 2 : 1;
  o.coarse_scale = st.coarse_scale;
  o.preconditioner = st.learned ? 1 : 0;
  o.smoother = st.smoother;
  const int k = w.published_ == 0 ? 1 : 0;
  auto& b = w.banks_[k];
  for (auto* v : {&b.pressures, &b.rhs, &b.residuals, &b.defects}) v->resize(nc);
  b.velocities.resize(nf);
  b.gauges.resize(std::max<std::size_t>(nc, 1));
  fl_pressure_diagnostics d{};
  fl_measurement m{};
  const fl_geometry g = internal::ToC(grid.geometry());
  const auto blocks = grid.active_block_indices();
  const int code = fl_project(runtime.handle(), &g, internal::ToC(blocks), blocks.size(), step_duration, w.cells_.data(),
                              nc, w.faces_.data(), nf, &o, b.pressures.data(), b.velocities.data(), b.rhs.data(),
                              b.residuals.data(), b.defects.data(), b.gauges.data(), &d, &m);
  if (code != FL_OK) return FromCode(code);
  b.gauge_ordinals.resize(d.gauge_count);
  for (std::size_t i = 0; i < d.gauge_count; ++i) b.gauge_ordinals[i] = static_cast<std::size_t>(b.gauges[i]);
  b.diagnostics = {d.iteration_count, d.multigrid_level_count, d.true_residual_l2, d.max_volume_scaled_residual,
                   d.max_cell_constraint_defect};
  // Frozen observations: RHS norm, supplied volume, and the explicitly prescribed wall flux oriented into
  // the pressure cells (a face's flux runs along its positive axis, from its tail cell into its head).
  double volume = 0.0, inward = 0.0;
  for (const auto& c : problem.cells) volume += c.fluid_volume;
  for (const auto& f : problem.faces) {
    if (!f.prescribed_wall_flux) continue;
    const auto& l = f.face_index.lattice_index;
    CellIndex3D tail{l.x, l.y, l.z};
    const int axis = static_cast<int>(f.face_index.axis);
    (axis == 0 ? tail.x : axis == 1 ? tail.y : tail.z) -= 1;
    if (HasCell(problem.cells, {l.x, l.y, l.z})) inward += *f.prescribed_wall_flux;
    if (HasCell(problem.cells, tail)) inward -= *f.prescribed_wall_flux;
  }
  b.frozen = {d.rhs_l2, volume, inward};
  w.device_diagnostics_ = {d.rhs_l2, d.initial_defect_l2, d.defect_l2, d.defect_linf, d.restarts,
                           d.learned_iterations};
  std::memcpy(w.measurement_.architecture, m.architecture, sizeof(w.measurement_.architecture));
  w.measurement_.kernel_ms = m.kernel_ms;
  w.measurement_.copy_ms = m.copy_ms;
  w.measurement_.launches = m.launches;
  w.published_ = k;
  return MacPressureProjectionView3D{b.pressures, b.velocities,     b.rhs,          b.residuals,
                                     b.defects,   b.gauge_ordinals, b.diagnostics, b.frozen};
}

StatusOr<MacPressureProjectionView3D> ProjectMacPressure3D(SparseMacGridView3D grid, double step_duration,
                                                           const LiquidPressureProblemView3D& problem,
                                                           MacPressureSolverOptions3D options,
                                                           MacPressureWorkspace3D& workspace) {
  if (problem.cells.size() > workspace.cell_capacity() || problem.faces.size() > workspace.face_capacity())
    return Status(StatusCode::kResourceExhausted);
  HipOperatorCapacities3D need;
  const auto blocks = grid.active_block_indices();
  need.blocks = std::max<std::size_t>(blocks.size(), 1);
  need.lookup_blocks = internal::LookupNeed(blocks);
  need.pressure_cells = std::max<std::size_t>(problem.cells.size(), 1);
  need.pressure_faces = std::max<std::size_t>(problem.faces.size(), 1);
  std::unique_lock<std::mutex> lock;
  auto rt = SharedHipRuntime3D::Acquire(need, lock);
  if (!rt.ok()) return rt.status();
  return HipProjectMacPressure3D(**rt, grid, step_duration, problem, options, workspace);
}

StatusOr<MacViscousPressureProjectionView3D> HipProjectUniformViscosity3D(
    HipOperatorRuntime3D& runtime, const LiquidBoundaryGeometryViewInputs3D& geometry,
    const LiquidPressureProblemView3D& problem, double dynamic_viscosity,
    std::span<const double> cell_dynamic_viscosity, MacPressureSolverOptions3D options, MacPressureWorkspace3D& w) {
  const std::size_t nc = problem.cells.size(), nf = problem.faces.size();
  if (nc > w.cell_capacity_ || nf > w.face_capacity_) return Status(StatusCode::kResourceExhausted);
  if (!cell_dynamic_viscosity.empty() && cell_dynamic_viscosity.size() != nc)
    return Status(StatusCode::kInvalidArgument);
  const auto grid = geometry.mac_grid;
  const double width = grid.geometry().cell_width;
  const double full_area = width * width, full_volume = full_area * width;
  constexpr double kFullCellRoundoff = 64 * std::numeric_limits<double>::epsilon();
  bool cut = false;
  for (const auto& c : problem.cells) cut = cut || std::abs(c.fluid_volume / full_volume - 1.0) > kFullCellRoundoff;
  for (const auto& f : problem.faces) cut = cut || (f.open_area != 0.0 && f.open_area != full_area);
  if (cut) return Status(StatusCode::kUnimplemented);
  w.cells_.clear();
  w.faces_.clear();
  w.material_.clear();
  for (std::size_t i = 0; i < nc; ++i) {
    const auto& c = problem.cells[i];
    w.cells_.push_back(internal::ToC(c));
    fl_viscosity_cell m{};
    m.cell_index = {c.cell_index.x, c.cell_index.y, c.cell_index.z};
    m.dynamic_viscosity = cell_dynamic_viscosity.empty() ? dynamic_viscosity : cell_dynamic_viscosity[i];
    w.material_.push_back(m);
  }
  for (const auto& f : problem.faces) w.faces_.push_back(internal::ToC(f));
  const fl_geometry g = internal::ToC(grid.geometry());
  const auto blocks = grid.active_block_indices();
  w.rows_.resize(6 * nc);
  std::size_t rows = w.rows_.size();
  fl_measurement m{};
  int code = fl_assemble_full_cell_strain(runtime.handle(), &g, internal::ToC(blocks), blocks.size(),
                                          w.material_.data(), nc, w.rows_.data(), &rows, &m);
  if (code != FL_OK) return FromCode(code);
  const HipMacPressureSolverSettings3D& st = w.settings_;
  fl_pressure_options o{};
  o.max_iterations = options.max_iterations;
  o.cpu_residual_tolerance_scale = options.cpu_residual_tolerance_scale;
  o.levels = st.levels;
  o.smoothing_sweeps = st.smoothing_sweeps;
  o.coarse_sweeps = st.coarse_sweeps;
  o.level_storage = 1;
  o.coarse_scale = st.coarse_scale;
  o.smoother = st.smoother;
  const int k = w.published_ == 0 ? 1 : 0;
  auto& b = w.banks_[k];
  b.pressures.resize(nc);
  b.defects.resize(nc);
  b.velocities.resize(nf);
  fl_viscous_diagnostics d{};
  code = fl_project_viscous(runtime.handle(), &g, internal::ToC(blocks), blocks.size(), geometry.step_duration,
                            w.cells_.data(), nc, w.faces_.data(), nf, w.rows_.data(), rows, &o, b.pressures.data(),
                            b.velocities.data(), b.defects.data(), &d, &m);
  if (code != FL_OK) return FromCode(code);
  w.measurement_.kernel_ms = m.kernel_ms;
  w.measurement_.copy_ms = m.copy_ms;
  w.measurement_.launches = m.launches;
  std::memcpy(w.measurement_.architecture, m.architecture, sizeof(w.measurement_.architecture));
  w.published_ = k;
  MacViscousPressureProjectionView3D v;
  v.cell_pressures = b.pressures;
  v.face_velocities = b.velocities;
  v.cell_constraint_defects = b.defects;
  v.pressure_iterations = d.pressure_iterations;
  v.constraint_residual_l2 = d.constraint_residual_l2;
  v.initial_constraint_residual_l2 = d.initial_constraint_residual_l2;
  v.max_cell_constraint_defect = d.max_cell_constraint_defect;
  v.max_momentum_residual = d.max_momentum_residual;
  v.inward_prescribed_wall_flux_m3_per_s = d.inward_prescribed_wall_flux;
  return v;
}

}  // namespace fluidity
