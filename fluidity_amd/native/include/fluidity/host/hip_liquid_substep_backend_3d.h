// Liquid-step plug-in table for the HIP hot path, shaped like LiquidSubstepBackend3D
// (native/include/fluidity/host/liquid_substep_backend_3d.h:47-260): a POD of C-compatible function
// pointers, each taking the backend context first. The members are the stages of the path, with
// the reference's member names (:204-225): topology (bin + stable sort), deposit (P2G quanta),
// normalize, forces, project, gather (G2P) and update_particles (obstacle-free particle update). The reference's boundary-workspace records are replaced
// by the standalone operator records of this layer; scene, publication and growth hooks are not
// part of the path.
#pragma once

#include <span>

#include "fluidity/host/liquid_particle_update_3d.h"
#include "fluidity/host/mac_pressure_3d.h"
#include "fluidity/host/transfers_3d.h"

namespace fluidity {

struct HipLiquidSubstepBackend3D {
  using PrepareTopology = Status(void* context, UniformMacGeometry3D geometry,
                                 std::span<const Vec3> sorted_world_positions,
                                 std::span<LiquidParticleSortEntry3D> order);
  using Deposit = Status(void* context, const SparseMacGrid3D& grid, std::span<const MacTransferParticle3D> particles,
                         MacDepositQuantization3D quantization, std::span<MacFaceDepositQuanta3D> face_quanta);
  using Normalize = Status(void* context, std::span<const MacFaceDepositQuanta3D> face_quanta,
                           MacDepositQuantization3D quantization, double minimum_face_mass,
                           std::span<double> face_masses, std::span<double> face_velocities,
                           std::span<LiquidFaceValidity3D> validity);
  using ApplyForces = Status(void* context, const SparseMacGrid3D& grid, std::span<const double> old_face_velocities,
                             Vec3 world_acceleration, double step_duration,
                             std::span<double> predicted_face_velocities);
  using Project = StatusOr<MacPressureProjection3D>(void* context, const MacPressureInputs3D& inputs,
                                                    MacPressureSolverOptions3D options);
  using Gather = Status(void* context, const SparseMacGrid3D& grid, std::span<const MacTransferParticle3D> particles,
                        std::span<const double> old_face_velocities, std::span<const double> new_face_velocities,
                        std::span<MacParticleGatherResult3D> results);
  using UpdateParticles = Status(void* context, const LiquidParticleUpdateInputs3D& inputs,
                                 std::span<LiquidParticleState3D> trial_particles);

  void* context = nullptr;
  PrepareTopology* topology = nullptr;
  Deposit* deposit = nullptr;
  Normalize* normalize = nullptr;
  ApplyForces* forces = nullptr;
  Project* project = nullptr;
  Gather* gather = nullptr;
  UpdateParticles* update_particles = nullptr;
};

// Table whose context is the given runtime (borrowed; it must outlive the table's use). Every member
// runs on the device; none falls back to host arithmetic.
HipLiquidSubstepBackend3D MakeHipLiquidSubstepBackend3D(HipOperatorRuntime3D& runtime);

}  // namespace fluidity
