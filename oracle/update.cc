// Oracle: the liquid particle update without an obstacle (test infrastructure only).
// Restates UpdateParticle (native/src/liquid_substep_3d_cpu.cc:239-405) for the absent obstacle and
// no regular-cell transport, with ResolveContact's absent-obstacle branch
// (native/include/fluidity/base/liquid_wall_contact_math_3d.h:365-396) and the vector helpers of
// native/src/liquid_substep_3d_staging.cc:160-171.
#include <algorithm>
#include <cmath>
#include <limits>

#include "grid_internal.h"

using namespace flo;

namespace {

flo_vec3 Add(flo_vec3 a, flo_vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
flo_vec3 Scale(flo_vec3 a, double f) { return {f * a.x, f * a.y, f * a.z}; }
double MaxMagnitude(flo_vec3 v) { return std::max({std::abs(v.x), std::abs(v.y), std::abs(v.z)}); }

int UpdateOne(const flo_update_options& o, const flo_boundary_gather& g, double liquid_distance,
              flo_liquid_particle& p) {
  const bool fallback = g.old_pic_fallback || g.new_pic_fallback || g.low_mass_fallback;
  const double alpha = p.pic_initialized && !fallback ? 0.99 * o.particle_energy_blend : 0;
  const flo_vec3 stored = Add(g.new_pic_velocity, Scale(g.particle_minus_old_pic_velocity, alpha));
  const double divergence = g.velocity_gradient.x_row.x + g.velocity_gradient.y_row.y + g.velocity_gradient.z_row.z;
  const double solid_distance = std::numeric_limits<double>::infinity();  // no obstacle
  const bool surface_expansion = divergence > 0 && liquid_distance <= 0 && liquid_distance >= -o.cell_width &&
                                 solid_distance > o.cell_width;
  const double beta = surface_expansion ? o.particle_energy_blend : 0;
  const flo_vec3 advecting = Add(g.new_pic_velocity, Scale(g.particle_minus_old_pic_velocity, beta * alpha));
  // ResolveContact: input checks, then the absent-obstacle advance.
  if (!Finite3(p.world_position) || !Finite3(stored) || !Finite3(advecting)) return FLO_INVALID_ARGUMENT;
  if (!std::isfinite(o.step_duration) || o.step_duration <= 0) return FLO_INVALID_ARGUMENT;
  if (!std::isfinite(o.friction_coefficient) || o.friction_coefficient < 0 ||
      !std::isfinite(o.penetration_tolerance) || o.penetration_tolerance < 0)
    return FLO_INVALID_ARGUMENT;
  const flo_vec3 position{std::fma(advecting.x, o.step_duration, p.world_position.x),
                          std::fma(advecting.y, o.step_duration, p.world_position.y),
                          std::fma(advecting.z, o.step_duration, p.world_position.z)};
  if (!Finite3(position)) return FLO_OUT_OF_RANGE;
  const flo_vec3 displacement{position.x - p.world_position.x, position.y - p.world_position.y,
                              position.z - p.world_position.z};
  if (MaxMagnitude(displacement) > 0.5 * o.cell_width) return FLO_ABORTED;
  const double volume_ratio = p.volume_ratio * std::exp(o.step_duration * divergence);
  if (!std::isfinite(volume_ratio) || volume_ratio <= 0) return FLO_ABORTED;
  p.world_position = position;
  p.world_velocity = stored;
  p.volume_ratio = volume_ratio;
  p.velocity_slopes = g.fitted_velocity_slopes;
  p.affine_enabled = g.affine_enabled;
  p.pic_initialized = 1;
  return FLO_OK;
}

}  // namespace

extern "C" int flo_update_particles(const flo_update_options* options, const flo_boundary_gather* gathered,
                                    const double* liquid_distances, size_t n, flo_liquid_particle* particles) {
  for (size_t i = 0; i < n; ++i)
    if (int s = UpdateOne(*options, gathered[i], liquid_distances[i], particles[i])) return s;
  return FLO_OK;
}
