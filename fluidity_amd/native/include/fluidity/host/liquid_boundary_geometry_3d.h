// Typed pressure links of a frozen liquid boundary geometry, built on the GPU
// (BuildLiquidPressureLinks3D on a FrozenLiquidBoundaryGeometryView3D,
// native/include/fluidity/host/liquid_boundary_workspace_3d.h:34-48, :139;
// native/src/liquid_boundary_geometry_3d.cc:1077-1383).
//
// Same records, checks, status codes and publication rule as the reference's view-based builder,
// without its bounded workspace caps (4096 cells, 24576 faces).
//
// The freeze that produces the view (FreezeLiquidBoundaryGeometry3D,
// native/include/fluidity/host/liquid_boundary_geometry_3d.h:277;
// native/src/liquid_boundary_geometry_3d.cc:765-1068) runs on the GPU too, for the absent, half-space
// and slab obstacles; the open tank, floating boxes and sampled solids stay with the reference.
#pragma once

#include <cstddef>
#include <cstdint>
#include <optional>
#include <span>
#include <variant>
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
// A corner sample: corner_index names the cell whose low corner it is.
struct LiquidCornerDistance3D {
  CellIndex3D corner_index;
  double signed_distance;  // Same physical field as the centre samples, m.
};
// A solid corner sample (liquid_boundary_geometry_3d.h:77-81 of the reference): the cell whose low corner
// it is, the solid's signed distance there (positive outside the solid, m) and the wall velocity
// integrated over blocked face parts when no dense face velocities are given (m/s).
struct SolidCornerDistance3D {
  CellIndex3D corner_index;
  double signed_distance;
  Vec3 wall_velocity{};
};
// A sampled solid for the freeze (:83-91): corner samples, distinct and (z,y,x) sorted, covering every
// candidate and frozen face; face normal velocities along each face's positive axis, one per owned slot,
// or none (then the corner velocities are integrated). A translating half-space may clip it.
struct SampledSolidBoundary3D {
  std::span<const SolidCornerDistance3D> corner_distances;
  std::span<const double> face_normal_velocities;
};
struct TranslatingHalfspace3D {
  Vec3 unit_normal;               // Outward from solid; solid is n . x <= start_plane_coordinate.
  double start_plane_coordinate;  // World meters at substep start.
  Vec3 wall_velocity;             // m/s.
};
// Solid is |unit_normal . x - start_center_coordinate| <= half_width.
struct TranslatingSlab3D {
  Vec3 unit_normal;
  double start_center_coordinate;
  double half_width;  // Positive finite, m.
  Vec3 wall_velocity;
};
// An open tank (native/include/fluidity/base/liquid_boundary_math.h:22-29): the freeze clips the liquid by
// its walls and gives the low-x face the piston's flux; the particle update resolves its contact.
struct AxisAlignedLiquidTank3D {
  double start_low_x;        // World x of the piston at substep start, m.
  double high_x;             // m.
  double floor_y;            // m.
  double low_z;              // m.
  double high_z;             // m.
  double piston_velocity_x;  // m/s.
};
using LiquidObstacle3D =
    std::variant<std::monostate, TranslatingHalfspace3D, TranslatingSlab3D, AxisAlignedLiquidTank3D>;
struct LiquidBoundaryGeometryInputs3D {
  const SparseMacGrid3D& mac_grid;
  std::span<const CellIndex3D> cell_indices;  // Candidates, strictly (z,y,x) sorted.
  std::span<const LiquidCenterDistance3D> center_distances;  // Strictly sorted.
  std::span<const LiquidCornerDistance3D> corner_distances;  // Strictly sorted.
  LiquidObstacle3D obstacle;
  double step_duration;    // s.
  double liquid_density;   // kg/m³.
  double surface_tension;  // N/m.
  std::optional<double> exact_outward_interface_curvature = std::nullopt;  // 1/m.
  std::optional<SampledSolidBoundary3D> sampled_solid = std::nullopt;
};
// Owns a frozen geometry and its centre samples; view() borrows them and the grid.
class FrozenLiquidBoundaryGeometry3D {
 public:
  const SparseMacGrid3D& mac_grid() const { return mac_grid_; }
  std::span<const FrozenLiquidCell3D> cells() const { return cells_; }
  std::span<const FrozenLiquidFace3D> faces() const { return faces_; }
  std::span<const LiquidCenterDistance3D> center_distances() const { return centers_; }
  double step_duration() const { return step_duration_; }
  double liquid_density() const { return liquid_density_; }
  double surface_tension() const { return surface_tension_; }
  // +infinity when no liquid cell has an open face to air under positive surface tension.
  double capillary_step_bound() const { return capillary_step_bound_; }
  FrozenLiquidBoundaryGeometryView3D view() const {
    return {mac_grid_.view(), centers_, cells_, faces_, step_duration_, liquid_density_, surface_tension_,
            capillary_step_bound_};
  }

 private:
  friend struct FreezeAccess;
  explicit FrozenLiquidBoundaryGeometry3D(const SparseMacGrid3D& grid) : mac_grid_(grid) {}
  SparseMacGrid3D mac_grid_;
  std::vector<LiquidCenterDistance3D> centers_;
  std::vector<FrozenLiquidCell3D> cells_;
  std::vector<FrozenLiquidFace3D> faces_;
  double step_duration_ = 0, liquid_density_ = 0, surface_tension_ = 0, capillary_step_bound_ = 0;
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

// Liquid cells of the candidates (centre distance < 0, clipped volume > 0) with their incident faces
// (open / blocked areas against the obstacle, its wall velocity) and the curvature of cells on an open
// face to air. The reference's caps apply: ResourceExhausted above 4096 candidates or 24576 faces.
// InvalidArgument for malformed inputs, FailedPrecondition for missing samples or owners, a cell split
// by a slab, a degenerate curvature gradient or a step above the capillary bound, OutOfRange for
// unrepresentable measures. Errors publish nothing.
StatusOr<FrozenLiquidBoundaryGeometry3D> FreezeLiquidBoundaryGeometry3D(const LiquidBoundaryGeometryInputs3D& inputs);
// The same without the caps (bounded by the runtime's capacities).
StatusOr<FrozenLiquidBoundaryGeometry3D> HipFreezeLiquidBoundaryGeometry3D(HipOperatorRuntime3D& runtime,
                                                                          const LiquidBoundaryGeometryInputs3D& inputs,
                                                                          HipOperatorMeasurement& measurement);

}  // namespace fluidity
