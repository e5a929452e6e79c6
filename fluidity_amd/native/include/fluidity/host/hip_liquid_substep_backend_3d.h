// Liquid-step plug-in table for the HIP hot path with the member types of LiquidSubstepBackend3D
// (native/include/fluidity/host/liquid_substep_backend_3d.h:47-260): a POD of C-compatible function
// pointers, each taking the backend context first, whose stage operations borrow the caller's
// workspaces and return their views.
//
// Members filled: the lifecycle hooks (initialize, prepare/commit/discard publication, prepare/commit/
// discard growth), topology, deposit, normalize, forces, freeze, links, project, gather and
// update_particles, and project_viscous for geometry without cut cells or faces. Scene sampling
// (sample_field, sample_faces and their batches), missing-face prediction, the wetted piston, the
// adaptive projection and the diagnostic phase hook
// belong to subsystems outside this library and stay null; a scheduler that needs them must refuse the
// table, as the reference's does (:39-41).
//
// The context owns one runtime whose device buffers are reserved at creation (fl_create and
// fl_reserve) from the state capacities, and the workspaces reserve their host banks at their
// Create: attempt operations within the capacities allocate and free nothing. Growth hooks run outside
// attempts and may allocate: prepare_growth builds a runtime for the new capacities, commit_growth
// adopts it and discard_growth drops it.
#pragma once

#include <memory>
#include <span>

#include "fluidity/host/hip_operator_runtime_3d.h"
#include "fluidity/host/liquid_boundary_workspace_3d.h"
#include "fluidity/host/liquid_particle_update_3d.h"
#include "fluidity/host/liquid_state_3d.h"
#include "fluidity/host/liquid_step_workspace_3d.h"
#include "fluidity/host/mac_pressure_workspace_3d.h"

namespace fluidity {

class HipLiquidSubstepContext3D {
 public:
  // One runtime for states within `capacities` (particles, blocks, cells, and faces as owned slots; a
  // projection takes at most min(6 * cells, faces) pressure faces). lookup_blocks bounds the
  // (z,y,x) box of the active blocks; 0 selects 8 * blocks. ResourceExhausted if the reservation does
  // not fit, FailedPrecondition without a device.
  static StatusOr<std::unique_ptr<HipLiquidSubstepContext3D>> Create(LiquidStateCapacities3D capacities,
                                                                     int device = 0, std::size_t lookup_blocks = 0);
  HipOperatorRuntime3D& runtime() { return *runtime_; }
  LiquidStateCapacities3D capacities() const { return capacities_; }
  // Device settings of the projection stage (levels, smoother, BF16 levels, relative divergence target).
  HipMacPressureSolverSettings3D projection_settings{};
  // Device time of the last stage call.
  HipOperatorMeasurement last_measurement{};

 private:
  friend struct HipBackendAccess;
  HipLiquidSubstepContext3D() = default;
  static StatusOr<std::unique_ptr<HipOperatorRuntime3D>> Reserve(LiquidStateCapacities3D capacities, int device,
                                                                 std::size_t lookup_blocks);
  std::unique_ptr<HipOperatorRuntime3D> runtime_, pending_runtime_;
  LiquidStateCapacities3D capacities_{}, pending_capacities_{};
  std::size_t lookup_blocks_ = 0;
  int device_ = 0;
  bool initialized_ = false;
  std::uint64_t committed_epoch_ = 0, candidate_epoch_ = 0;
  bool candidate_ = false;
};

struct HipLiquidSubstepBackend3D {
  using Initialize = Status(void* context, const LiquidState3D& initial);
  using PreparePublication = Status(void* context, const LiquidState3D& trial, LiquidClock3D accepted_clock);
  using CommitPublication = void(void* context) noexcept;
  using DiscardPublication = void(void* context) noexcept;
  using PrepareGrowth = Status(void* context, LiquidStateCapacities3D capacities, const LiquidState3D& frame_start);
  using CommitGrowth = void(void* context) noexcept;
  using DiscardGrowth = void(void* context) noexcept;
  using PrepareTopology = Status(void* context, std::span<const LiquidParticleState3D> particles,
                                 std::span<const CellIndex3D> pressure_candidates, const LiquidState3D& committed,
                                 LiquidStepWorkspace3D& workspace);
  using Deposit = StatusOr<LiquidBoundaryFaceDepositsView3D>(void* context,
                                                              const LiquidBoundaryDepositViewInputs3D& inputs,
                                                              LiquidBoundaryWorkspace3D& workspace);
  using Normalize = StatusOr<NormalizedLiquidMacFacesView3D>(void* context,
                                                              std::span<const MacFaceDepositQuanta3D> face_quanta,
                                                              MacDepositQuantization3D quantization,
                                                              double minimum_face_mass,
                                                              LiquidBoundaryWorkspace3D& workspace);
  using ApplyForces = Status(void* context, SparseMacGridView3D grid, std::span<const double> old_face_velocities,
                             Vec3 world_acceleration, double step_duration,
                             std::span<double> predicted_face_velocities);
  using Freeze = StatusOr<FrozenLiquidBoundaryGeometryView3D>(void* context,
                                                               const LiquidBoundaryGeometryViewInputs3D& inputs,
                                                               LiquidBoundaryWorkspace3D& workspace);
  using BuildLinks = StatusOr<LiquidPressureProblemView3D>(void* context,
                                                            const FrozenLiquidBoundaryGeometryView3D& geometry,
                                                            const LiquidPressureFields3D& fields,
                                                            LiquidBoundaryWorkspace3D& workspace);
  using Project = StatusOr<MacPressureProjectionView3D>(void* context, SparseMacGridView3D grid,
                                                         double step_duration,
                                                         const LiquidPressureProblemView3D& problem,
                                                         MacPressureSolverOptions3D options,
                                                         MacPressureWorkspace3D& workspace);
  using ProjectViscous = StatusOr<MacViscousPressureProjectionView3D>(
      void* context, const LiquidBoundaryGeometryViewInputs3D& geometry, const LiquidPressureProblemView3D& problem,
      double dynamic_viscosity, std::span<const double> cell_dynamic_viscosity, MacPressureSolverOptions3D options,
      MacPressureWorkspace3D& workspace);
  using Gather = StatusOr<LiquidBoundaryGatherView3D>(void* context, const LiquidBoundaryGatherViewInputs3D& inputs,
                                                       LiquidBoundaryWorkspace3D& workspace);
  using UpdateParticles = Status(void* context, const LiquidParticleUpdateInputs3D& inputs,
                                 std::span<LiquidParticleState3D> trial_particles);

  void* context = nullptr;
  void* sample_field = nullptr;  // not provided
  void* sample_faces = nullptr;  // not provided
  Initialize* initialize = nullptr;
  PreparePublication* prepare_publication = nullptr;
  CommitPublication* commit_publication = nullptr;
  DiscardPublication* discard_publication = nullptr;
  PrepareGrowth* prepare_growth = nullptr;
  CommitGrowth* commit_growth = nullptr;
  DiscardGrowth* discard_growth = nullptr;
  PrepareTopology* topology = nullptr;
  Deposit* deposit = nullptr;
  Normalize* normalize = nullptr;
  void* predict_missing_face = nullptr;  // not provided
  ApplyForces* forces = nullptr;
  Freeze* freeze = nullptr;
  void* wetted_piston = nullptr;  // not provided
  BuildLinks* links = nullptr;
  Project* project = nullptr;
  ProjectViscous* project_viscous = nullptr;  // uncut geometry (cut cells: Unimplemented)
  Gather* gather = nullptr;
  UpdateParticles* update_particles = nullptr;
};

// Table over a context (borrowed; it must outlive the table's use). Every filled member runs on the
// device; none falls back to host arithmetic.
HipLiquidSubstepBackend3D MakeHipLiquidSubstepBackend3D(HipLiquidSubstepContext3D& context);

}  // namespace fluidity
