// The six hot-path stages of the liquid substep as a function-pointer table over one runtime.
#include "fluidity/host/hip_liquid_substep_backend_3d.h"

#include "internal.h"

namespace fluidity {
namespace {

HipOperatorRuntime3D& Rt(void* context) { return *static_cast<HipOperatorRuntime3D*>(context); }

Status Topology(void* context, UniformMacGeometry3D geometry, std::span<const Vec3> positions,
                std::span<LiquidParticleSortEntry3D> order) {
  if (order.size() != positions.size()) return Status(StatusCode::kInvalidArgument);
  const fl_geometry g = internal::ToC(geometry);
  return FromCode(fl_bin_sort(Rt(context).handle(), &g, reinterpret_cast<const fl_vec3*>(positions.data()),
                              positions.size(), reinterpret_cast<fl_sort_entry*>(order.data()), nullptr));
}

Status DepositStage(void* context, const SparseMacGrid3D& grid, std::span<const MacTransferParticle3D> particles,
                    MacDepositQuantization3D q, std::span<MacFaceDepositQuanta3D> face_quanta) {
  HipOperatorMeasurement m;
  return HipDepositMacQuanta3D(Rt(context), grid, particles, q, face_quanta, m);
}

Status NormalizeStage(void* context, std::span<const MacFaceDepositQuanta3D> quanta, MacDepositQuantization3D q,
                      double minimum_face_mass, std::span<double> masses, std::span<double> velocities,
                      std::span<LiquidFaceValidity3D> validity) {
  HipOperatorMeasurement m;
  return HipNormalizeApplyForces3D(Rt(context), quanta, q, minimum_face_mass, Vec3{0.0, 0.0, 0.0}, 0.0,
                                   NormalizedLiquidFaces3D{masses, velocities, validity, {}}, m);
}

Status ForcesStage(void* context, const SparseMacGrid3D& grid, std::span<const double> u_old, Vec3 g, double dt,
                   std::span<double> predicted) {
  if (u_old.size() != grid.face_count() || predicted.size() != grid.face_count())
    return Status(StatusCode::kInvalidArgument);
  return FromCode(fl_apply_forces(Rt(context).handle(), u_old.data(), u_old.size(), fl_vec3{g.x, g.y, g.z}, dt,
                                  predicted.data(), nullptr));
}

StatusOr<MacPressureProjection3D> ProjectStage(void* context, const MacPressureInputs3D& inputs,
                                               MacPressureSolverOptions3D options) {
  auto r = HipProjectMacPressure3D(Rt(context), inputs, options);
  if (!r.ok()) return r.status();
  return std::move(r.value().projection);
}

Status GatherStage(void* context, const SparseMacGrid3D& grid, std::span<const MacTransferParticle3D> particles,
                   std::span<const double> u_old, std::span<const double> u_new,
                   std::span<MacParticleGatherResult3D> results) {
  HipOperatorMeasurement m;
  return HipGatherMac3D(Rt(context), grid, particles, u_old, u_new, results, m);
}

Status UpdateStage(void* context, const LiquidParticleUpdateInputs3D& inputs,
                   std::span<LiquidParticleState3D> particles) {
  HipOperatorMeasurement m;
  return HipUpdateLiquidParticles3D(Rt(context), inputs, particles, m);
}

}  // namespace

HipLiquidSubstepBackend3D MakeHipLiquidSubstepBackend3D(HipOperatorRuntime3D& runtime) {
  HipLiquidSubstepBackend3D b;
  b.context = &runtime;
  b.topology = &Topology;
  b.deposit = &DepositStage;
  b.normalize = &NormalizeStage;
  b.forces = &ForcesStage;
  b.project = &ProjectStage;
  b.gather = &GatherStage;
  b.update_particles = &UpdateStage;
  return b;
}

}  // namespace fluidity
