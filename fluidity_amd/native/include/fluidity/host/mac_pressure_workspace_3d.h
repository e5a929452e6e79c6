// Bounded pressure workspace (native/include/fluidity/host/mac_pressure_workspace_3d.h:27-115): the
// projection of a typed-link problem into reserved banks, returned as a borrowed view.
//
// Create reserves the host banks and staging from the cell and face capacities; a projection writes
// the candidate bank and publishes it only on success (a failed projection leaves the previous view
// intact). The solve runs on the GPU: ProjectMacPressure3D on the process-wide runtime, the Hip twin on
// a caller's runtime, whose projection arrays and level pool fl_create reserved, so that the twin
// allocates and frees no device memory for inputs within the runtime's capacities. The host-side
// validation and hierarchy set-up of the device solve still use transient host containers.
#pragma once

#include <array>
#include <cstddef>
#include <span>
#include <vector>

#include "fluidity/host/hip_operator_runtime_3d.h"
#include "fluidity/host/liquid_boundary_workspace_3d.h"
#include "fluidity/host/mac_pressure_3d.h"
#include "fluidity/host/sparse_mac_grid_3d.h"
#include "fluidity/host/viscosity.h"
#include "fluidity/status.h"
#include "fluidity_hip.h"

namespace fluidity {

// The coupled viscous pressure projection's borrowed result (host/viscosity.h:171-182).
struct MacViscousPressureProjectionView3D {
  std::span<const double> cell_pressures;           // Pa, in pressure cell order
  std::span<const double> face_velocities;          // m/s, in incident pressure face order
  std::span<const double> cell_constraint_defects;  // outward flux minus source plus compliance, / volume
  std::size_t pressure_iterations = 0;
  double constraint_residual_l2 = 0;
  double initial_constraint_residual_l2 = 0;
  double max_cell_constraint_defect = 0;
  double max_momentum_residual = 0;
  double inward_prescribed_wall_flux_m3_per_s = 0;
};

// Scalar observations of the frozen system that take no part in the solve.
struct MacPressureFrozenMeasurement3D {
  double integrated_rhs_l2_m3_per_s = 0;
  double pressure_cell_volume_m3 = 0;
  double inward_prescribed_wall_flux_m3_per_s = 0;
};
// Read-only projection in input cell / face order, with the units of MacPressureProjection3D.
struct MacPressureProjectionView3D {
  std::span<const double> cell_pressures;
  std::span<const double> corrected_face_velocities;
  std::span<const double> integrated_rhs;
  std::span<const double> true_residuals;
  std::span<const double> cell_constraint_defects;
  std::span<const std::size_t> gauge_cell_ordinals;
  MacPressureDiagnostics3D diagnostics;
  MacPressureFrozenMeasurement3D frozen_measurement;
};

class MacPressureWorkspace3D {
 public:
  MacPressureWorkspace3D(MacPressureWorkspace3D&&) noexcept = default;
  MacPressureWorkspace3D& operator=(MacPressureWorkspace3D&&) noexcept = default;
  MacPressureWorkspace3D(const MacPressureWorkspace3D&) = delete;
  MacPressureWorkspace3D& operator=(const MacPressureWorkspace3D&) = delete;

  // Reserves two projection banks and the staging of cell_capacity cells and face_capacity faces.
  static StatusOr<MacPressureWorkspace3D> Create(std::size_t cell_capacity, std::size_t face_capacity);

  // Device solver settings beyond the reference options (levels, smoother, level storage, relative
  // divergence target, learned preconditioner).
  void set_device_settings(const HipMacPressureSolverSettings3D& settings) { settings_ = settings; }
  const HipMacPressureSolverSettings3D& device_settings() const { return settings_; }
  // Device diagnostics and measurement of the last successful projection.
  const HipMacPressureDiagnostics3D& device_diagnostics() const { return device_diagnostics_; }
  const HipOperatorMeasurement& measurement() const { return measurement_; }

  std::size_t cell_capacity() const { return cell_capacity_; }
  std::size_t face_capacity() const { return face_capacity_; }

 private:
  friend StatusOr<MacPressureProjectionView3D> HipProjectMacPressure3D(HipOperatorRuntime3D&, SparseMacGridView3D,
                                                                       double, const LiquidPressureProblemView3D&,
                                                                       MacPressureSolverOptions3D,
                                                                       MacPressureWorkspace3D&);
  friend StatusOr<MacViscousPressureProjectionView3D> HipProjectUniformViscosity3D(
      HipOperatorRuntime3D&, const LiquidBoundaryGeometryViewInputs3D&, const LiquidPressureProblemView3D&, double,
      std::span<const double>, MacPressureSolverOptions3D, MacPressureWorkspace3D&);
  MacPressureWorkspace3D() = default;
  struct Bank {
    std::vector<double> pressures, velocities, rhs, residuals, defects;
    std::vector<std::uint64_t> gauges;
    std::vector<std::size_t> gauge_ordinals;
    MacPressureDiagnostics3D diagnostics;
    MacPressureFrozenMeasurement3D frozen;
  };
  std::size_t cell_capacity_ = 0, face_capacity_ = 0;
  std::array<Bank, 2> banks_;
  std::vector<fl_pressure_cell> cells_;
  std::vector<fl_pressure_face> faces_;
  std::vector<fl_viscosity_cell> material_;  // cell_capacity
  std::vector<fl_strain_row> rows_;          // 6 * cell_capacity
  HipMacPressureSolverSettings3D settings_{};
  HipMacPressureDiagnostics3D device_diagnostics_{};
  HipOperatorMeasurement measurement_{};
  int published_ = -1;
};

// ProjectMacPressure3D on a borrowed grid and a typed-link problem (mac_pressure_workspace_3d.h:126-140):
// same validation, gates and statuses as the owning projection; the result borrows the workspace.
StatusOr<MacPressureProjectionView3D> ProjectMacPressure3D(SparseMacGridView3D grid, double step_duration,
                                                           const LiquidPressureProblemView3D& problem,
                                                           MacPressureSolverOptions3D options,
                                                           MacPressureWorkspace3D& workspace);
StatusOr<MacPressureProjectionView3D> HipProjectMacPressure3D(HipOperatorRuntime3D& runtime, SparseMacGridView3D grid,
                                                              double step_duration,
                                                              const LiquidPressureProblemView3D& problem,
                                                              MacPressureSolverOptions3D options,
                                                              MacPressureWorkspace3D& workspace);
// The substep's viscous projection member (ProjectUniformViscosity, liquid_substep_3d_cpu.cc:619-695):
// uniform (or per-cell, when cell_dynamic_viscosity is not empty) viscosity on the problem's cells,
// full-cell strain rows (fl_assemble_full_cell_strain) and the coupled viscous pressure solve
// (fl_project_viscous, the V-cycle-preconditioned GCR with the workspace's device settings) into the
// candidate bank. Cut geometry (a cell's volume off the full cell by more than 64 eps, or a face
// partly open) needs the reference's octant-weighted strain quadrature, which is not built:
// Unimplemented. The result borrows the workspace.
StatusOr<MacViscousPressureProjectionView3D> HipProjectUniformViscosity3D(
    HipOperatorRuntime3D& runtime, const LiquidBoundaryGeometryViewInputs3D& geometry,
    const LiquidPressureProblemView3D& problem, double dynamic_viscosity,
    std::span<const double> cell_dynamic_viscosity, MacPressureSolverOptions3D options,
    MacPressureWorkspace3D& workspace);

}  // namespace fluidity
