// The particle update over the C-ABI (fl_update_particles).
#include "fluidity/host/liquid_particle_update_3d.h"

#include <cstring>
#include <mutex>

#include "internal.h"
#include "records.h"

namespace fluidity {
namespace {

static_assert(sizeof(LiquidParticleState3D) == sizeof(fl_liquid_particle) && sizeof(LiquidParticleState3D) == 328);
static_assert(offsetof(LiquidParticleState3D, world_position) == offsetof(fl_liquid_particle, world_position) &&
              offsetof(LiquidParticleState3D, affine_enabled) == offsetof(fl_liquid_particle, affine_enabled) &&
              offsetof(LiquidParticleState3D, volume_ratio) == offsetof(fl_liquid_particle, volume_ratio) &&
              offsetof(LiquidParticleState3D, pic_initialized) == offsetof(fl_liquid_particle, pic_initialized) &&
              offsetof(LiquidParticleState3D, constituent_count) ==
                  offsetof(fl_liquid_particle, constituent_count) &&
              offsetof(LiquidParticleState3D, constituents) == offsetof(fl_liquid_particle, constituents));

int Update(HipOperatorRuntime3D& rt, const LiquidParticleUpdateInputs3D& in, std::span<LiquidParticleState3D> p,
           fl_measurement* m) {
  if (in.gathered.size() != p.size() || in.liquid_distances.size() != p.size()) return FL_INVALID_ARGUMENT;
  fl_particle_update_options o{in.cell_width, in.step_duration, in.particle_energy_blend, in.friction_coefficient,
                               in.penetration_tolerance, {}};
  o.obstacle = internal::ToC(in.obstacle);
  return fl_update_particles(rt.handle(), &o, reinterpret_cast<const fl_boundary_gather*>(in.gathered.data()),
                             in.liquid_distances.data(), p.size(), reinterpret_cast<fl_liquid_particle*>(p.data()), m);
}

}  // namespace

Status UpdateLiquidParticles3D(const LiquidParticleUpdateInputs3D& inputs,
                               std::span<LiquidParticleState3D> trial_particles) {
  HipOperatorCapacities3D need;
  need.particles = std::max<std::size_t>(trial_particles.size(), 1);
  std::unique_lock<std::mutex> lock;
  auto rt = SharedHipRuntime3D::Acquire(need, lock);
  if (!rt.ok()) return rt.status();
  return FromCode(Update(**rt, inputs, trial_particles, nullptr));
}

Status HipUpdateLiquidParticles3D(HipOperatorRuntime3D& runtime, const LiquidParticleUpdateInputs3D& inputs,
                                  std::span<LiquidParticleState3D> trial_particles,
                                  HipOperatorMeasurement& measurement) {
  fl_measurement m{};
  const int code = Update(runtime, inputs, trial_particles, &m);
  if (code == FL_OK) {
    std::memcpy(measurement.architecture, m.architecture, sizeof(measurement.architecture));
    measurement.kernel_ms = m.kernel_ms;
    measurement.copy_ms = m.copy_ms;
    measurement.launches = m.launches;
  }
  return FromCode(code);
}

}  // namespace fluidity
