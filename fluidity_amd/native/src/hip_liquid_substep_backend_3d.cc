// The hot-path stages of the liquid substep as a LiquidSubstepBackend3D-shaped table over one reserved
// runtime.
#include "fluidity/host/hip_liquid_substep_backend_3d.h"

#include <cstring>

#include "internal.h"

namespace fluidity {

struct HipBackendAccess {
  static HipLiquidSubstepContext3D& Ctx(void* context) { return *static_cast<HipLiquidSubstepContext3D*>(context); }

  static bool Fits(const HipLiquidSubstepContext3D& c, const LiquidState3D& s) {
    const LiquidStateCapacities3D& k = c.capacities_;
    return s.particles.size() <= k.particles && s.blocks.size() <= k.blocks && s.cells.size() <= k.cells &&
           s.faces.size() <= k.faces;
  }

  // initialize: the context holds no device copy of the state; it checks that the state fits.
  static Status Initialize(void* context, const LiquidState3D& initial) {
    auto& c = Ctx(context);
    if (!Fits(c, initial)) return Status(StatusCode::kResourceExhausted);
    c.initialized_ = true;
    c.committed_epoch_ = initial.clock.epoch;
    c.candidate_ = false;
    return Status();
  }
  static Status PreparePublication(void* context, const LiquidState3D& trial, LiquidClock3D accepted_clock) {
    auto& c = Ctx(context);
    if (!c.initialized_) return Status(StatusCode::kFailedPrecondition);
    if (!Fits(c, trial)) return Status(StatusCode::kResourceExhausted);
    c.candidate_epoch_ = accepted_clock.epoch;
    c.candidate_ = true;
    return Status();
  }
  static void CommitPublication(void* context) noexcept {
    auto& c = Ctx(context);
    if (c.candidate_) c.committed_epoch_ = c.candidate_epoch_;
    c.candidate_ = false;
  }
  static void DiscardPublication(void* context) noexcept { Ctx(context).candidate_ = false; }
  static Status PrepareGrowth(void* context, LiquidStateCapacities3D capacities, const LiquidState3D& frame_start) {
    auto& c = Ctx(context);
    auto rt = HipLiquidSubstepContext3D::Reserve(capacities, c.device_, c.lookup_blocks_);
    if (!rt.ok()) return rt.status();
    if (frame_start.particles.size() > capacities.particles || frame_start.blocks.size() > capacities.blocks ||
        frame_start.cells.size() > capacities.cells || frame_start.faces.size() > capacities.faces)
      return Status(StatusCode::kResourceExhausted);
    c.pending_runtime_ = std::move(*rt);
    c.pending_capacities_ = capacities;
    return Status();
  }
  static void CommitGrowth(void* context) noexcept {
    auto& c = Ctx(context);
    if (!c.pending_runtime_) return;
    c.runtime_ = std::move(c.pending_runtime_);
    c.capacities_ = c.pending_capacities_;
  }
  static void DiscardGrowth(void* context) noexcept { Ctx(context).pending_runtime_.reset(); }

  static Status Topology(void* context, std::span<const LiquidParticleState3D> particles,
                         std::span<const CellIndex3D> candidates, const LiquidState3D& committed,
                         LiquidStepWorkspace3D& workspace) {
    auto& c = Ctx(context);
    // A committed state without topology (the first step) prepares without history.
    const bool history = !committed.blocks.empty() || !committed.cells.empty() || !committed.faces.empty();
    return workspace.HipPrepareTrialTopology(c.runtime(), particles, candidates, history ? &committed : nullptr,
                                             c.last_measurement);
  }
  static StatusOr<LiquidBoundaryFaceDepositsView3D> Deposit(void* context,
                                                           const LiquidBoundaryDepositViewInputs3D& inputs,
                                                           LiquidBoundaryWorkspace3D& workspace) {
    return workspace.HipDeposit(Ctx(context).runtime(), inputs);
  }
  static StatusOr<NormalizedLiquidMacFacesView3D> Normalize(void* context, std::span<const MacFaceDepositQuanta3D> q,
                                                           MacDepositQuantization3D quantization,
                                                           double minimum_face_mass,
                                                           LiquidBoundaryWorkspace3D& workspace) {
    return workspace.HipNormalize(Ctx(context).runtime(), q, quantization, minimum_face_mass);
  }
  static Status Forces(void* context, SparseMacGridView3D grid, std::span<const double> u_old, Vec3 g, double dt,
                       std::span<double> predicted) {
    if (u_old.size() != grid.face_count() || predicted.size() != grid.face_count())
      return Status(StatusCode::kInvalidArgument);
    auto& c = Ctx(context);
    fl_measurement m{};
    const int code = fl_apply_forces(c.runtime().handle(), u_old.data(), u_old.size(), fl_vec3{g.x, g.y, g.z}, dt,
                                     predicted.data(), &m);
    c.last_measurement.kernel_ms = m.kernel_ms;
    c.last_measurement.copy_ms = m.copy_ms;
    c.last_measurement.launches = m.launches;
    return FromCode(code);
  }
  static StatusOr<FrozenLiquidBoundaryGeometryView3D> Freeze(void* context,
                                                            const LiquidBoundaryGeometryViewInputs3D& inputs,
                                                            LiquidBoundaryWorkspace3D& workspace) {
    return workspace.HipFreeze(Ctx(context).runtime(), inputs);
  }
  static StatusOr<LiquidPressureProblemView3D> Links(void* context, const FrozenLiquidBoundaryGeometryView3D& geometry,
                                                    const LiquidPressureFields3D& fields,
                                                    LiquidBoundaryWorkspace3D& workspace) {
    return workspace.HipBuildLinks(Ctx(context).runtime(), geometry, fields);
  }
  static StatusOr<MacPressureProjectionView3D> Project(void* context, SparseMacGridView3D grid, double dt,
                                                      const LiquidPressureProblemView3D& problem,
                                                      MacPressureSolverOptions3D options,
                                                      MacPressureWorkspace3D& workspace) {
    auto& c = Ctx(context);
    workspace.set_device_settings(c.projection_settings);
    auto r = HipProjectMacPressure3D(c.runtime(), grid, dt, problem, options, workspace);
    if (r.ok()) c.last_measurement = workspace.measurement();
    return r;
  }
  static StatusOr<MacViscousPressureProjectionView3D> ProjectViscous(
      void* context, const LiquidBoundaryGeometryViewInputs3D& geometry, const LiquidPressureProblemView3D& problem,
      double dynamic_viscosity, std::span<const double> cell_dynamic_viscosity, MacPressureSolverOptions3D options,
      MacPressureWorkspace3D& workspace) {
    auto& c = Ctx(context);
    workspace.set_device_settings(c.projection_settings);
    auto r = HipProjectUniformViscosity3D(c.runtime(), geometry, problem, dynamic_viscosity, cell_dynamic_viscosity,
                                          options, workspace);
    if (r.ok()) c.last_measurement = workspace.measurement();
    return r;
  }
  static StatusOr<LiquidBoundaryGatherView3D> Gather(void* context, const LiquidBoundaryGatherViewInputs3D& inputs,
                                                    LiquidBoundaryWorkspace3D& workspace) {
    return workspace.HipGather(Ctx(context).runtime(), inputs);
  }
  static Status Update(void* context, const LiquidParticleUpdateInputs3D& inputs,
                       std::span<LiquidParticleState3D> particles) {
    auto& c = Ctx(context);
    return HipUpdateLiquidParticles3D(c.runtime(), inputs, particles, c.last_measurement);
  }
};

StatusOr<std::unique_ptr<HipOperatorRuntime3D>> HipLiquidSubstepContext3D::Reserve(LiquidStateCapacities3D k,
                                                                                   int device,
                                                                                   std::size_t lookup_blocks) {
  HipOperatorCapacities3D caps;
  caps.particles = std::max<std::size_t>(k.particles, 1);
  caps.blocks = std::max<std::size_t>(std::min(k.blocks, k.faces / 192), 1);
  caps.lookup_blocks = lookup_blocks ? lookup_blocks : std::max<std::size_t>(8 * caps.blocks, 64);
  caps.pressure_cells = std::max<std::size_t>(k.cells, 1);
  caps.pressure_faces = std::max<std::size_t>(std::min(6 * k.cells, k.faces), 1);
  auto rt = HipOperatorRuntime3D::Create(caps, device);
  if (!rt.ok()) return rt.status();
  const int code = fl_reserve((*rt)->handle());
  if (code != FL_OK) return FromCode(code);
  return std::move(*rt);
}

StatusOr<std::unique_ptr<HipLiquidSubstepContext3D>> HipLiquidSubstepContext3D::Create(LiquidStateCapacities3D capacities,
                                                                                       int device,
                                                                                       std::size_t lookup_blocks) {
  auto rt = Reserve(capacities, device, lookup_blocks);
  if (!rt.ok()) return rt.status();
  std::unique_ptr<HipLiquidSubstepContext3D> c(new HipLiquidSubstepContext3D());
  c->runtime_ = std::move(*rt);
  c->capacities_ = capacities;
  c->lookup_blocks_ = lookup_blocks;
  c->device_ = device;
  return c;
}

HipLiquidSubstepBackend3D MakeHipLiquidSubstepBackend3D(HipLiquidSubstepContext3D& context) {
  HipLiquidSubstepBackend3D b;
  b.context = &context;
  b.initialize = &HipBackendAccess::Initialize;
  b.prepare_publication = &HipBackendAccess::PreparePublication;
  b.commit_publication = &HipBackendAccess::CommitPublication;
  b.discard_publication = &HipBackendAccess::DiscardPublication;
  b.prepare_growth = &HipBackendAccess::PrepareGrowth;
  b.commit_growth = &HipBackendAccess::CommitGrowth;
  b.discard_growth = &HipBackendAccess::DiscardGrowth;
  b.topology = &HipBackendAccess::Topology;
  b.deposit = &HipBackendAccess::Deposit;
  b.normalize = &HipBackendAccess::Normalize;
  b.forces = &HipBackendAccess::Forces;
  b.freeze = &HipBackendAccess::Freeze;
  b.links = &HipBackendAccess::Links;
  b.project = &HipBackendAccess::Project;
  b.project_viscous = &HipBackendAccess::ProjectViscous;
  b.gather = &HipBackendAccess::Gather;
  b.update_particles = &HipBackendAccess::Update;
  return b;
}

}  // namespace fluidity
