// Bounded boundary workspace of a liquid step (native/include/fluidity/host/liquid_boundary_workspace_3d.h):
// the view records the step stages pass to one another, and a workspace whose storage is reserved at
// Create and published in two banks per region.
//
// Each operation writes a candidate bank and publishes it only on success, so a failed call leaves the
// last published result (and every borrow of it) intact; a returned view stays valid until the next
// successful call of the same region, the owner's move or destruction. Operations run on the GPU:
// the reference-named methods on the process-wide runtime, the Hip* twins on a caller's runtime. With
// a runtime whose buffers were reserved (fl_reserve), the twins allocate and free nothing, host or
// device, for inputs within the workspace capacities.
#pragma once

#include <array>
#include <cstddef>
#include <optional>
#include <span>
#include <vector>

#include "fluidity/base/liquid_boundary_transfer_types_3d.h"
#include "fluidity/host/hip_operator_runtime_3d.h"
#include "fluidity/host/liquid_boundary_geometry_3d.h"
#include "fluidity/host/liquid_boundary_transfer_3d.h"
#include "fluidity/host/liquid_state_3d.h"
#include "fluidity/host/mac_pressure_3d.h"
#include "fluidity/host/sparse_mac_grid_3d.h"
#include "fluidity/status.h"
#include "fluidity_hip.h"

namespace fluidity {

// Same units, sample order and geometry contract as LiquidBoundaryGeometryInputs3D, on a borrowed grid.
struct LiquidBoundaryGeometryViewInputs3D {
  SparseMacGridView3D mac_grid;
  std::span<const CellIndex3D> cell_indices;
  std::span<const LiquidCenterDistance3D> center_distances;
  std::span<const LiquidCornerDistance3D> corner_distances;
  LiquidObstacle3D obstacle;
  double step_duration;
  double liquid_density;
  double surface_tension;
  std::optional<double> exact_outward_interface_curvature = std::nullopt;
  std::optional<SampledSolidBoundary3D> sampled_solid = std::nullopt;
};
// Pressure cells and typed links in frozen cell and face order.
struct LiquidPressureProblemView3D {
  std::span<const MacPressureCell3D> cells;
  std::span<const MacPressureFace3D> faces;
  std::size_t lower_theta_regularization_count = 0;
  bool seed_uniform_dirichlet_rest_pressure = false;
};
struct LiquidBoundaryDepositViewInputs3D {
  SparseMacGridView3D mac_grid;
  std::span<const LiquidBoundaryParticle3D> particles;
  std::span<const LiquidBoundaryFaceTag3D> face_tags;
  MacDepositQuantization3D quantization;
};
using LiquidCompatibleStencils3D = std::array<CompatibleMacFaceStencil3D, 3>;
// Integer face sums in owned-slot order and, when the workspace reserved them, the compatible stencils
// in particle order (empty otherwise).
struct LiquidBoundaryFaceDepositsView3D {
  std::span<const MacFaceDepositQuanta3D> face_quanta;
  std::span<const LiquidCompatibleStencils3D> compatible_stencils;
};
// Owned-slot fields: kilograms, m/s and the validity category.
struct NormalizedLiquidMacFacesView3D {
  std::span<const double> face_masses;
  std::span<const double> face_velocities;
  std::span<const LiquidFaceValidity3D> validity;
};
struct LiquidBoundaryGatherViewInputs3D {
  SparseMacGridView3D mac_grid;
  std::span<const LiquidBoundaryParticle3D> particles;
  double old_time;
  double new_time;
  LiquidBoundaryVelocityHistory3D old_history;
  LiquidBoundaryVelocityHistory3D new_history;
};
struct LiquidBoundaryGatherView3D {
  std::span<const LiquidBoundaryParticleGather3D> particles;
  std::size_t old_pic_fallback_particle_count = 0;
  std::size_t new_pic_fallback_particle_count = 0;
  std::size_t low_mass_fallback_particle_count = 0;
};

// Compatible stencils (5832 bytes per particle) are reserved only up to this many particles.
inline constexpr std::size_t kLiquidStencilParticleCapacity3D = 1 << 16;

class LiquidBoundaryWorkspace3D {
 public:
  LiquidBoundaryWorkspace3D(LiquidBoundaryWorkspace3D&&) noexcept = default;
  LiquidBoundaryWorkspace3D& operator=(LiquidBoundaryWorkspace3D&&) noexcept = default;
  LiquidBoundaryWorkspace3D(const LiquidBoundaryWorkspace3D&) = delete;
  LiquidBoundaryWorkspace3D& operator=(const LiquidBoundaryWorkspace3D&) = delete;

  // Reserves every bank and staging array from the capacities (cells: candidates and frozen cells,
  // 27 centre and 8 corner samples each; faces: frozen faces and links; blocks: owned slots).
  // ResourceExhausted if the reservation fails.
  static StatusOr<LiquidBoundaryWorkspace3D> Create(LiquidStateCapacities3D capacities);

  StatusOr<FrozenLiquidBoundaryGeometryView3D> Freeze(const LiquidBoundaryGeometryViewInputs3D& inputs);
  StatusOr<LiquidPressureProblemView3D> BuildLinks(const FrozenLiquidBoundaryGeometryView3D& geometry,
                                                   const LiquidPressureFields3D& fields);
  StatusOr<LiquidBoundaryFaceDepositsView3D> Deposit(const LiquidBoundaryDepositViewInputs3D& inputs);
  StatusOr<NormalizedLiquidMacFacesView3D> Normalize(std::span<const MacFaceDepositQuanta3D> quanta,
                                                     MacDepositQuantization3D quantization,
                                                     double minimum_face_mass);
  StatusOr<LiquidBoundaryGatherView3D> Gather(const LiquidBoundaryGatherViewInputs3D& inputs);

  // The same operations on an explicit runtime.
  StatusOr<FrozenLiquidBoundaryGeometryView3D> HipFreeze(HipOperatorRuntime3D& runtime,
                                                         const LiquidBoundaryGeometryViewInputs3D& inputs);
  StatusOr<LiquidPressureProblemView3D> HipBuildLinks(HipOperatorRuntime3D& runtime,
                                                      const FrozenLiquidBoundaryGeometryView3D& geometry,
                                                      const LiquidPressureFields3D& fields);
  StatusOr<LiquidBoundaryFaceDepositsView3D> HipDeposit(HipOperatorRuntime3D& runtime,
                                                        const LiquidBoundaryDepositViewInputs3D& inputs);
  StatusOr<NormalizedLiquidMacFacesView3D> HipNormalize(HipOperatorRuntime3D& runtime,
                                                        std::span<const MacFaceDepositQuanta3D> quanta,
                                                        MacDepositQuantization3D quantization,
                                                        double minimum_face_mass);
  StatusOr<LiquidBoundaryGatherView3D> HipGather(HipOperatorRuntime3D& runtime,
                                                 const LiquidBoundaryGatherViewInputs3D& inputs);

  LiquidStateCapacities3D capacities() const { return capacities_; }

 private:
  LiquidBoundaryWorkspace3D() = default;
  struct GeometryBank {
    std::vector<LiquidCenterDistance3D> centers;
    std::vector<FrozenLiquidCell3D> cells;
    std::vector<FrozenLiquidFace3D> faces;
  };
  struct LinkBank {
    std::vector<MacPressureCell3D> cells;
    std::vector<MacPressureFace3D> faces;
  };
  struct DepositBank {
    std::vector<MacFaceDepositQuanta3D> quanta;
    std::vector<LiquidCompatibleStencils3D> stencils;
  };
  struct NormalizationBank {
    std::vector<double> masses;
    std::vector<double> velocities;
    std::vector<LiquidFaceValidity3D> validity;
    std::vector<double> predicted;  // staging of the predicted velocities the call does not return
  };
  // C-ABI records staged for the device calls.
  struct Staging {
    std::vector<fl_int3> candidates;
    std::vector<fl_center_distance> centers, corners;
    std::vector<fl_frozen_cell> frozen_cells;
    std::vector<fl_frozen_face> frozen_faces;
    std::vector<fl_cell_fields> fields;
    std::vector<fl_optional_flux> wetted;
    std::vector<fl_air_pressure> air;
    std::vector<fl_internal_interface> interfaces;
    std::vector<fl_pressure_cell> link_cells;
    std::vector<fl_pressure_face> link_faces;
  };
  LiquidStateCapacities3D capacities_{};
  std::array<GeometryBank, 2> geometry_;
  std::array<LinkBank, 2> links_;
  std::array<DepositBank, 2> deposits_;
  std::array<NormalizationBank, 2> normalized_;
  std::array<std::vector<LiquidBoundaryParticleGather3D>, 2> gathered_;
  std::array<double, 2> capillary_bound_{};
  std::array<std::size_t, 2> lower_theta_{};
  std::array<std::array<std::size_t, 3>, 2> fallbacks_{};
  Staging staging_;
  int geometry_bank_ = -1, link_bank_ = -1, deposit_bank_ = -1, normalization_bank_ = -1, gather_bank_ = -1;
};

// Bounded equivalents of the owning operators: same numerics and statuses, results borrowed from the
// workspace (liquid_boundary_workspace_3d.h:203-224).
StatusOr<FrozenLiquidBoundaryGeometryView3D> FreezeLiquidBoundaryGeometry3D(
    const LiquidBoundaryGeometryViewInputs3D& inputs, LiquidBoundaryWorkspace3D& workspace);
StatusOr<LiquidPressureProblemView3D> BuildLiquidPressureLinks3D(const FrozenLiquidBoundaryGeometryView3D& geometry,
                                                                const LiquidPressureFields3D& fields,
                                                                LiquidBoundaryWorkspace3D& workspace);
StatusOr<LiquidBoundaryFaceDepositsView3D> DepositCompatibleMacQuanta3D(const LiquidBoundaryDepositViewInputs3D& inputs,
                                                                        LiquidBoundaryWorkspace3D& workspace);
StatusOr<NormalizedLiquidMacFacesView3D> NormalizeLiquidMacFaces3D(std::span<const MacFaceDepositQuanta3D> quanta,
                                                                   MacDepositQuantization3D quantization,
                                                                   double minimum_face_mass,
                                                                   LiquidBoundaryWorkspace3D& workspace);
StatusOr<LiquidBoundaryGatherView3D> GatherLiquidBoundaryParticles3D(const LiquidBoundaryGatherViewInputs3D& inputs,
                                                                     LiquidBoundaryWorkspace3D& workspace);

}  // namespace fluidity
