// The liquid particle update on the GPU (the update_particles stage of LiquidSubstepBackend3D,
// native/include/fluidity/host/liquid_substep_backend_3d.h:195-202; native/src/liquid_substep_3d_cpu.cc:239-405)
// for a substep without regular-cell transport, with no obstacle or a half-space or slab obstacle.
#pragma once

#include <array>
#include <cstdint>
#include <span>

#include "fluidity/base/liquid_boundary_transfer_types_3d.h"
#include "fluidity/base/mac_types_3d.h"
#include "fluidity/host/hip_operator_runtime_3d.h"
#include "fluidity/host/liquid_boundary_geometry_3d.h"
#include "fluidity/status.h"

namespace fluidity {

inline constexpr std::size_t kMaximumLiquidParticleConstituents3D = 8;
// LiquidConstituentFraction3D (native/include/fluidity/host/liquid_mixture_3d.h:33-37).
struct LiquidConstituentFraction3D {
  std::uint64_t substance_id = 0;
  double volume_fraction = 0.0;
};
// LiquidParticleState3D (native/include/fluidity/host/liquid_state_3d.h:100-120).
struct LiquidParticleState3D {
  std::uint64_t id = 0;
  std::uint32_t tag = 0;
  Vec3 world_position{};
  Vec3 world_velocity{};
  Mat3 velocity_slopes{};
  bool affine_enabled = true;
  double mass = 1.0;
  double reference_volume = 1.0;
  double volume_ratio = 1.0;
  std::int64_t side_tag = 0;
  bool pic_initialized = false;
  std::uint64_t phase_id = 0;
  std::uint8_t constituent_count = 0;
  bool quadratic_adaptive_transfer = false;
  std::array<LiquidConstituentFraction3D, kMaximumLiquidParticleConstituents3D> constituents{};
};

// LiquidParticleUpdateInputs3D without the regular-cell transport (empty); the obstacle is absent,
// a half-space, a slab or the open tank (frictionless walls and a moving piston; the piston's
// reaction force is not computed).
struct LiquidParticleUpdateInputs3D {
  std::span<const LiquidBoundaryParticleGather3D> gathered;  // Frozen old/new gather, particle order.
  std::span<const double> liquid_distances;                   // Frozen x^n field, meters, particle order.
  LiquidObstacle3D obstacle;     // Pose at substep start, constant translation during step_duration.
  double cell_width;             // Meters.
  double step_duration;          // Positive seconds.
  double particle_energy_blend;  // e in [0,1], alpha = 0.99 e and surface beta = e.
  double friction_coefficient;   // Nonnegative.
  double penetration_tolerance;  // Meters.
};

// ASFLIP stored velocity, surface-expansion advection gated by the solid distance, the certified
// swept-plane contact, half-cell guard and J update, per particle in order; the first failing
// particle's status (InvalidArgument, FailedPrecondition, OutOfRange or Aborted). Unlike the
// reference, which leaves earlier particles updated on error, the particles are written only on
// success. J may differ from the host's exp in the last place.
Status UpdateLiquidParticles3D(const LiquidParticleUpdateInputs3D& inputs,
                               std::span<LiquidParticleState3D> trial_particles);
Status HipUpdateLiquidParticles3D(HipOperatorRuntime3D& runtime, const LiquidParticleUpdateInputs3D& inputs,
                                  std::span<LiquidParticleState3D> trial_particles,
                                  HipOperatorMeasurement& measurement);

}  // namespace fluidity
