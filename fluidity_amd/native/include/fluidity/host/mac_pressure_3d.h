// MAC pressure projection by mixed-precision MGPCG on the GPU
// (native/include/fluidity/host/mac_pressure_3d.h:14-102; cuda_mac_pressure_3d.h:8-20).
//
// Same records, argument meaning, status codes and acceptance gates as ProjectMacPressure3D:
// ||b - Ap||2 <= 2e-6 + 2e-6 ||b||2 (capped at 1e-3 ||b||2 with internal jumps), max |r| / V and
// max |defect| <= 2e-5 s^-1, evaluated in FP64 on the device. The preconditioner is a geometric
// red-black Gauss-Seidel V-cycle on the 4^3 blocks (FP32, FP64 dots and residuals) instead of the
// reference's aggregation Jacobi V-cycle, so iteration counts differ while the accepted projection
// meets the same gates. The reference's 4096-cell cap is replaced by the runtime capacities.
#pragma once

#include <cstddef>
#include <optional>
#include <span>
#include <variant>
#include <vector>

#include "fluidity/base/mac_types_3d.h"
#include "fluidity/host/hip_operator_runtime_3d.h"
#include "fluidity/host/sparse_mac_grid_3d.h"
#include "fluidity/status.h"

namespace fluidity {

struct MacPressureCell3D {
  CellIndex3D cell_index;
  double fluid_volume;       // m^3, positive, at most the cell volume
  double target_divergence;  // s^-1
  double bulk_modulus;       // Pa, positive; +inf selects zero compliance
  double previous_pressure;  // Pa
  double integrated_inward_wall_flux = 0.0;  // m^3/s into the liquid
};

struct InternalMacPressureLink3D {
  double pressure_jump;  // Pa, enters as p_tail - p_head + jump
};
struct DirichletMacPressureLink3D {
  double exterior_pressure;  // Pa at the single missing endpoint
};
struct SealedMacPressureLink3D {};

struct MacPressureFace3D {
  FaceIndex3D face_index;  // lattice index names the head; the normal points from tail to head
  double open_area;        // m^2
  double inverse_density;  // m^3/kg
  double pressure_gradient_distance;  // m
  double predicted_normal_velocity;   // m/s
  double wall_normal_velocity;        // m/s
  std::variant<InternalMacPressureLink3D, DirichletMacPressureLink3D, SealedMacPressureLink3D> pressure_link;
  std::optional<double> prescribed_wall_flux = std::nullopt;  // m^3/s
};

// Cells distinct and in (z,y,x) order; faces the exact union of incident faces in owned-slot order.
struct MacPressureInputs3D {
  const SparseMacGrid3D& mac_grid;
  double step_duration;
  std::span<const MacPressureCell3D> pressure_cells;
  std::span<const MacPressureFace3D> pressure_faces;
};

struct MacPressureDiagnostics3D {
  std::size_t iteration_count = 0;
  std::size_t multigrid_level_count = 0;
  double true_residual_l2 = 0.0;
  double max_volume_scaled_residual = 0.0;
  double max_cell_constraint_defect = 0.0;
};

struct MacPressureProjection3D {
  std::vector<double> cell_pressures;
  std::vector<double> corrected_face_velocities;
  std::vector<double> integrated_rhs;
  std::vector<double> true_residuals;
  std::vector<double> cell_constraint_defects;
  std::vector<std::size_t> gauge_cell_ordinals;
  MacPressureDiagnostics3D diagnostics;
};

struct MacPressureSolverOptions3D {
  std::size_t max_iterations = 256;           // [1, 256]
  double cpu_residual_tolerance_scale = 1.0;  // [1e-12, 1]
};

// Device solver settings beyond the reference options.
struct HipMacPressureSolverSettings3D {
  double relative_divergence_target = 0.0;  // 0 disables; else also ||defect||2 <= target ||defect(p=0)||2
  int levels = 6;                           // geometric multigrid levels
  int smoothing_sweeps = 2;                 // RBGS sweeps before and after the coarse correction
  int coarse_sweeps = 2;                    // symmetric sweeps on the coarsest level
  double coarse_scale = 0.5;                // coarse operator = scale * Galerkin crossing sums
  bool bf16_levels = false;                 // V-cycle coefficients stored in BF16, FP32 accumulation
  bool learned = false;  // flexible CG preconditioned by the runtime's trained network (fl_learn_*), V-cycle fallback
  int smoother = 0;  // 0 level-wide red-black Gauss-Seidel; 1 red-black sweeps inside each 4^3 block with block
                     // Jacobi between blocks; 2 level-wide on the finest level, per block below
};

struct HipMacPressureDiagnostics3D {
  double rhs_l2 = 0.0;
  double initial_defect_l2 = 0.0;  // ||defect(p = 0)||2
  double defect_l2 = 0.0;
  double defect_linf = 0.0;
  std::size_t restarts = 0;
  std::size_t learned_iterations = 0;
};

struct HipMacPressureProjection3D {
  MacPressureProjection3D projection;
  HipMacPressureDiagnostics3D device_diagnostics;
  HipOperatorMeasurement measurement;
};

StatusOr<MacPressureProjection3D> ProjectMacPressure3D(const MacPressureInputs3D& inputs,
                                                       MacPressureSolverOptions3D options = {});

StatusOr<HipMacPressureProjection3D> HipProjectMacPressure3D(HipOperatorRuntime3D& runtime,
                                                             const MacPressureInputs3D& inputs,
                                                             MacPressureSolverOptions3D options = {},
                                                             HipMacPressureSolverSettings3D settings = {});

}  // namespace fluidity
