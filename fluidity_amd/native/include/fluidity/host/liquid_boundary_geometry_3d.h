// Typed pressure links of a frozen liquid boundary geometry, built on the GPU
// (BuildLiquidPressureLinks3D on a FrozenLiquidBoundaryGeometryView3D,
// native/include/fluidity/host/liquid_boundary_workspace_3d.h:34-48, :139;
// native/src/liquid_boundary_geometry_3d.cc:1077-1383).
//
// Same records, checks, status codes and publication rule as the reference's view-based builder,
// without its bounded workspace caps (4096 cells, 24576 faces). The freeze that produces the view
// (fluid volumes, apertures, curvature from distance samples) is not part of this layer.
#pragma once

#include <cstddef>
#include <cstdint>
#include <optional>
#include <span>
#include <vector>

#include "fluidity/host/hip_operator_runtime_3d.h"
#include "fluidity/host/mac_pressure_3d.h"
#include "fluidity/host/sparse_mac_grid_3d.h"
#include "fluidity/status.h"

namespace fluidity {

struct LiquidCenterDistance3D {
  CellIndex3D cell_index;
  double signed_distance;  // Physical liquid distance in meters, negative inside.
};
struct FrozenLiquidCell3D {
  CellIndex3D cell_index;
  double fluid_volume;     // m³.
  double center_distance;  // Physical signed distance, m.
  double curvature;        // Outward curvature, 1/m.
};
struct FrozenLiquidFace3D {
  FaceIndex3D face_index;
  std::size_t face_slot;
  double open_area;             // m².
  double blocked_area;          // Exactly h² - open_area, m².
  double wall_normal_velocity;  // Positive Cartesian face axis, m/s.
  std::optional<double> prescribed_wall_flux = std::nullopt;  // m³/s.
  bool is_low_x_piston_face = false;
};
struct LiquidPressureCellFields3D {
  double target_divergence;  // s⁻¹.
  double bulk_modulus;       // Positive Pa, or +infinity.
  double previous_pressure;  // Pa.
  std::optional<double> density = std::nullopt;  // kg/m³.
};
struct LiquidAirCellPressure3D {
  CellIndex3D cell_index;
  double gas_gauge_pressure;  // Pa.
};
struct LiquidInternalInterface3D {
  std::size_t face_ordinal;
  double tail_fraction;
  double surface_tension;
  double curvature_tail_to_head;
  double normal_viscous_jump_tail_to_head;
};
struct FrozenLiquidBoundaryGeometryView3D {
  SparseMacGridView3D mac_grid;
  std::span<const LiquidCenterDistance3D> center_distances;
  std::span<const FrozenLiquidCell3D> cells;
  std::span<const FrozenLiquidFace3D> faces;
  double step_duration = 0;
  double liquid_density = 0;
  double surface_tension = 0;
  double capillary_step_bound = 0;
};
struct LiquidPressureFields3D {
  std::span<const LiquidPressureCellFields3D> cells;  // Frozen cell order.
  std::span<const double> predicted_face_velocities;  // Frozen face order, m/s.
  double ambient_pressure;                            // Pa.
  std::span<const std::optional<double>> wetted_piston_face_fluxes{};
  std::span<const LiquidAirCellPressure3D> air_cell_pressures{};
  std::span<const LiquidInternalInterface3D> internal_interfaces{};
};
struct LiquidPressureProblem3D {
  std::vector<MacPressureCell3D> cells;
  std::vector<MacPressureFace3D> faces;
  std::size_t lower_theta_regularization_count = 0;
};

// Pressure cells and typed links in frozen order: internal links (zero jump, or an interface
// crossing's jump after its capillary bound), free-surface Dirichlet links at clamp(theta, 0.05, 1) h
// with exterior pressure gas + sigma * curvature, sealed faces. InvalidArgument for malformed records
// or fields, FailedPrecondition for missing samples, owners or incident faces, OutOfRange for
// unrepresentable arithmetic, Aborted when a step exceeds an interface's capillary bound. Errors
// publish nothing.
StatusOr<LiquidPressureProblem3D> BuildLiquidPressureLinks3D(const FrozenLiquidBoundaryGeometryView3D& geometry,
                                                            const LiquidPressureFields3D& fields);
StatusOr<LiquidPressureProblem3D> HipBuildLiquidPressureLinks3D(HipOperatorRuntime3D& runtime,
                                                               const FrozenLiquidBoundaryGeometryView3D& geometry,
                                                               const LiquidPressureFields3D& fields,
                                                               HipOperatorMeasurement& measurement);

}  // namespace fluidity
