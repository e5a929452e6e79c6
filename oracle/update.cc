// Oracle: the liquid particle update (test infrastructure only).
// Restates UpdateParticle (native/src/liquid_substep_3d_cpu.cc:239-405) without regular-cell transport
// for the absent, half-space, slab and open-tank obstacles: SolidDistance (native/src/liquid_substep_3d_staging.cc:
// 197-205), ResolveContact with its certified swept-plane contact
// (native/include/fluidity/base/liquid_wall_contact_math_3d.h:1-152,365-556) over the bounded projected
// sums of native/include/fluidity/base/liquid_boundary_math.h:93-263, and the vector helpers of
// native/src/liquid_substep_3d_staging.cc:160-171.
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <utility>

#include "bounded_sums.h"
#include "grid_internal.h"

using namespace flo;
using namespace flo::bounded;

namespace {

flo_vec3 Add(flo_vec3 a, flo_vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
flo_vec3 Scale(flo_vec3 a, double f) { return {f * a.x, f * a.y, f * a.z}; }
double MaxMagnitude(flo_vec3 v) { return std::max({std::abs(v.x), std::abs(v.y), std::abs(v.z)}); }
double Dot(flo_vec3 a, flo_vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
constexpr double kInf = std::numeric_limits<double>::infinity();

// ---- the swept-plane contact ----
struct SweepPlane {
  flo_vec3 normal;
  double start, half_width;
  flo_vec3 wall;
};
bool Finite3v(flo_vec3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
flo_vec3 Advance(flo_vec3 x, flo_vec3 v, double t) {
  return {std::fma(v.x, t, x.x), std::fma(v.y, t, x.y), std::fma(v.z, t, x.z)};
}
Bounds DistanceAlongPath(const SweepPlane& p, flo_vec3 x, flo_vec3 v, double duration) {
  PSum d;
  d.AddDot(p.normal, x);
  d.Add(-p.start);
  d.Add(-p.half_width);
  d.AddRelativeMotion(p.normal, v, p.wall, duration);
  return d.bounds();
}
Bounds RelativeNormalVelocity(const SweepPlane& p, flo_vec3 v) {
  PSum s;
  s.AddRelativeMotion(p.normal, v, p.wall, 1);
  return s.bounds();
}
bool RepresentablePlanes(const Bounds& c, double half_width) {
  if (!c.finite()) return false;
  if (half_width == 0) return true;
  const double low = c.estimate - half_width, high = c.estimate + half_width;
  return std::isfinite(c.lower - half_width) && std::isfinite(c.upper + half_width) && low < high;
}
// The next representable outward move with the largest normal progress.
flo_vec3 NudgeOutward(flo_vec3 v, flo_vec3 n) {
  double values[3] = {v.x, v.y, v.z};
  const double normals[3] = {n.x, n.y, n.z};
  int best_axis = 0;
  double best_value = values[0], best_progress = 0;
  for (int a = 0; a < 3; ++a) {
    if (normals[a] == 0) continue;
    const double next = std::nextafter(values[a], std::copysign(kInf, normals[a]));
    const double progress = normals[a] * (next - values[a]);
    if (std::isfinite(next) && progress > best_progress) {
      best_axis = a;
      best_value = next;
      best_progress = progress;
    }
  }
  values[best_axis] = best_value;
  return {values[0], values[1], values[2]};
}
template <typename Evaluate>
bool CertifySide(flo_vec3& candidate, flo_vec3 n, Evaluate evaluate) {
  if (!Finite3v(candidate)) return false;
  for (int k = 0; k < 2; ++k) {
    const Bounds b = evaluate(candidate);
    if (!b.finite()) return false;
    if (b.lower >= 0) return true;
    candidate = NudgeOutward(candidate, n);
  }
  const Bounds b = evaluate(candidate);
  return b.finite() && b.lower >= 0;
}
// ApplyOneWayResponse: OutOfRange when uncertified.
int OneWayResponse(const SweepPlane& p, flo_vec3 v, double relative, double nls, double friction, flo_vec3& out) {
  if (relative >= 0) {
    out = v;
    return FLO_OK;
  }
  flo_vec3 c = Advance(v, p.normal, -relative / nls);
  const Bounds rounded = RelativeNormalVelocity(p, c);
  if (!Finite3v(c) || !rounded.finite()) return FLO_OUT_OF_RANGE;
  c = Advance(c, p.normal, -rounded.estimate / nls);
  if (friction > 0) {
    const flo_vec3 t{c.x - p.wall.x, c.y - p.wall.y, c.z - p.wall.z};
    const double tangent_speed = std::hypot(t.x, t.y, t.z);
    const double inward_speed = -relative / std::sqrt(nls);
    if (!std::isfinite(tangent_speed) || inward_speed == 0) return FLO_OUT_OF_RANGE;
    if (tangent_speed == 0 || friction >= tangent_speed / inward_speed) {
      c = p.wall;
    } else {
      const double fraction = friction * inward_speed / tangent_speed;
      c = {std::lerp(c.x, p.wall.x, fraction), std::lerp(c.y, p.wall.y, fraction), std::lerp(c.z, p.wall.z, fraction)};
    }
  }
  if (!CertifySide(c, p.normal, [&](flo_vec3 u) { return RelativeNormalVelocity(p, u); })) return FLO_OUT_OF_RANGE;
  out = c;
  return FLO_OK;
}
// ResolveContact for a half-space or slab after the shared input checks: the end position and the
// stored velocity.
int PlaneContact(const flo_update_options& o, flo_vec3 x, flo_vec3 stored, flo_vec3 advecting, flo_vec3& position,
                 flo_vec3& stored_out) {
  const flo_obstacle& ob = o.obstacle;
  const bool slab = ob.kind == FLO_OBSTACLE_SLAB;
  if (slab && (!std::isfinite(ob.half_width) || ob.half_width <= 0)) return FLO_INVALID_ARGUMENT;
  SweepPlane plane{ob.unit_normal, ob.coordinate, slab ? ob.half_width : 0, ob.wall_velocity};
  const flo_vec3 n = plane.normal;
  const double nls = n.x * n.x + n.y * n.y + n.z * n.z;
  if (!Finite3v(n) || !Finite3v(plane.wall) || !std::isfinite(plane.start) || std::abs(nls - 1) > 1e-12)
    return FLO_INVALID_ARGUMENT;
  PSum translated;
  translated.Add(plane.start);
  if (!RepresentablePlanes(translated.bounds(), plane.half_width)) return FLO_OUT_OF_RANGE;
  translated.AddRelativeMotion(n, plane.wall, {0, 0, 0}, o.step_duration);
  if (!RepresentablePlanes(translated.bounds(), plane.half_width)) return FLO_OUT_OF_RANGE;
  if (plane.half_width > 0) {
    PSum centre;
    centre.AddDot(n, x);
    centre.Add(-plane.start);
    const Bounds cb = centre.bounds();
    if (!cb.finite()) return FLO_OUT_OF_RANGE;
    int side = 0;
    if (!cb.sign(&side) || side == 0) return FLO_FAILED_PRECONDITION;
    if (side < 0) {
      plane.normal = {-n.x, -n.y, -n.z};
      plane.start = -plane.start;
    }
  }
  const Bounds start_distance = DistanceAlongPath(plane, x, {0, 0, 0}, 0);
  const Bounds adv_projection = RelativeNormalVelocity(plane, advecting);
  const Bounds stored_projection = RelativeNormalVelocity(plane, stored);
  int start_side = 0, adv_direction = 0, stored_direction = 0;
  if (!start_distance.sign(&start_side) || !adv_projection.sign(&adv_direction) ||
      !stored_projection.sign(&stored_direction))
    return FLO_OUT_OF_RANGE;
  const bool penetrating = start_side < 0;
  if (penetrating && start_distance.lower < -o.penetration_tolerance) return FLO_FAILED_PRECONDITION;
  double elapsed = 0;
  bool hit = penetrating || (start_side == 0 && (adv_direction < 0 || stored_direction < 0));
  if (start_side > 0 && adv_direction < 0) {
    int end_side = 0;
    if (!DistanceAlongPath(plane, x, advecting, o.step_duration).sign(&end_side)) return FLO_OUT_OF_RANGE;
    if (end_side <= 0) {
      if (end_side < 0 && !start_distance.resolved()) return FLO_OUT_OF_RANGE;
      hit = true;
      elapsed = end_side == 0 ? o.step_duration
                              : std::min(o.step_duration, start_distance.estimate / -adv_projection.estimate);
      if (elapsed <= 0 || elapsed / o.step_duration == 0) return FLO_OUT_OF_RANGE;
    }
  }
  const auto final_distance = [&](flo_vec3 u) { return DistanceAlongPath(plane, u, {0, 0, 0}, o.step_duration); };
  if (!hit) {
    position = Advance(x, advecting, o.step_duration);
    const Bounds accepted = final_distance(position);
    if (!Finite3v(position) || !accepted.finite() || accepted.lower < 0) return FLO_OUT_OF_RANGE;
    stored_out = stored;
    return FLO_OK;
  }
  if (!stored_projection.resolved() || !adv_projection.resolved()) return FLO_OUT_OF_RANGE;
  flo_vec3 new_stored{}, new_adv{};
  if (int s = OneWayResponse(plane, stored, stored_projection.estimate, nls, o.friction_coefficient, new_stored))
    return s;
  if (int s = OneWayResponse(plane, advecting, adv_projection.estimate, nls, o.friction_coefficient, new_adv)) return s;
  flo_vec3 hit_position = Advance(x, advecting, elapsed);
  const auto hit_distance = [&](flo_vec3 u) { return DistanceAlongPath(plane, u, {0, 0, 0}, elapsed); };
  const Bounds rounded_hit = hit_distance(hit_position);
  if (!rounded_hit.finite()) return FLO_OUT_OF_RANGE;
  hit_position = Advance(hit_position, plane.normal, -rounded_hit.estimate / nls);
  if (!CertifySide(hit_position, plane.normal, hit_distance)) return FLO_OUT_OF_RANGE;
  position = Advance(hit_position, new_adv, o.step_duration - elapsed);
  if (!CertifySide(position, plane.normal, final_distance)) return FLO_OUT_OF_RANGE;
  const double nl = std::sqrt(nls);
  if (!std::isfinite(stored_projection.estimate / nl) || !std::isfinite(adv_projection.estimate / nl))
    return FLO_OUT_OF_RANGE;
  stored_out = new_stored;
  return FLO_OK;
}

// SolidDistance: the signed distance to the obstacle at its start (+infinity without one); for the tank
// the least inward clearance (SampleTankWall, liquid_boundary_math.h:319-350), NaN when one is not
// finite.
double SolidDistance(const flo_obstacle& ob, flo_vec3 x) {
  if (ob.kind == FLO_OBSTACLE_HALFSPACE) return Dot(ob.unit_normal, x) - ob.coordinate;
  if (ob.kind == FLO_OBSTACLE_SLAB) return std::abs(Dot(ob.unit_normal, x) - ob.coordinate) - ob.half_width;
  if (ob.kind == FLO_OBSTACLE_TANK) {
    const flo_tank& t = ob.tank;
    const double d[5] = {x.x - t.start_low_x, t.high_x - x.x, x.y - t.floor_y, x.z - t.low_z, t.high_z - x.z};
    double least = d[0];
    for (double v : d) {
      if (!std::isfinite(v)) return std::numeric_limits<double>::quiet_NaN();
      if (v < least) least = v;
    }
    return least;
  }
  return kInf;
}

// CheckTankStep (liquid_boundary_math.h:252-276).
int CheckTankStep(const flo_tank& t, double dt) {
  if (!std::isfinite(t.start_low_x) || !std::isfinite(t.high_x) || !std::isfinite(t.floor_y) ||
      !std::isfinite(t.low_z) || !std::isfinite(t.high_z) || !std::isfinite(t.piston_velocity_x))
    return FLO_INVALID_ARGUMENT;
  if (t.start_low_x >= t.high_x || t.low_z >= t.high_z) return FLO_INVALID_ARGUMENT;
  const double displacement = t.piston_velocity_x * dt;
  const double end_x = std::fma(t.piston_velocity_x, dt, t.start_low_x);
  const double start_width = t.high_x - t.start_low_x, end_width = t.high_x - end_x, width_z = t.high_z - t.low_z;
  if (!std::isfinite(displacement) || !std::isfinite(end_x) || !std::isfinite(start_width) ||
      !std::isfinite(end_width) || !std::isfinite(width_z))
    return FLO_OUT_OF_RANGE;
  if (end_x >= t.high_x) return FLO_FAILED_PRECONDITION;
  return FLO_OK;
}

// SweepTankAxis (liquid_wall_contact_math_3d.h:154-289): one Cartesian component against its walls,
// each expressed along local x.
int SweepTankAxis(const flo_update_options& o, const SweepPlane* walls, int count, double x0, double s0, double a0,
                  double& out_x, double& out_stored) {
  flo_vec3 start{x0, 0, 0}, stored{s0, 0, 0}, adv{a0, 0, 0};
  const SweepPlane* carrying = nullptr;
  const auto impact = [&](const SweepPlane& wall) {
    const Bounds ss = RelativeNormalVelocity(wall, stored), as = RelativeNormalVelocity(wall, adv);
    if (!ss.resolved() || !as.resolved()) return FLO_OUT_OF_RANGE;
    if (ss.estimate < 0) stored.x = wall.wall.x;
    if (as.estimate <= 0) {
      adv.x = wall.wall.x;
      carrying = &wall;
    }
    return FLO_OK;
  };
  for (int k = 0; k < count; ++k) {
    const SweepPlane& wall = walls[k];
    const Bounds sd = DistanceAlongPath(wall, start, {0, 0, 0}, 0);
    int side = 0;
    if (!sd.sign(&side)) return FLO_OUT_OF_RANGE;
    if (side > 0) continue;
    if (sd.lower < -o.penetration_tolerance) return FLO_FAILED_PRECONDITION;
    const double wall_coordinate = wall.start * wall.normal.x;
    const bool low_wall = wall.normal.x > 0;
    const bool stored_in = low_wall ? stored.x < wall.wall.x : stored.x > wall.wall.x;
    const bool adv_in = low_wall ? adv.x < wall.wall.x : adv.x > wall.wall.x;
    if (side < 0 || stored_in || adv_in) {
      if (int s = impact(wall)) return s;
      start.x = wall_coordinate;
    }
  }
  if (carrying == nullptr) {
    for (int k = 0; k < count; ++k) {
      const SweepPlane& wall = walls[k];
      const Bounds as = RelativeNormalVelocity(wall, adv);
      int direction = 0;
      if (!as.sign(&direction)) return FLO_OUT_OF_RANGE;
      if (direction >= 0) continue;
      int end_side = 0;
      if (!DistanceAlongPath(wall, start, adv, o.step_duration).sign(&end_side)) return FLO_OUT_OF_RANGE;
      if (end_side > 0) continue;
      if (end_side < 0) {
        const Bounds sd = DistanceAlongPath(wall, start, {0, 0, 0}, 0);
        if (!sd.resolved() || !as.resolved()) return FLO_OUT_OF_RANGE;
        const double t = std::min(o.step_duration, sd.estimate / -as.estimate);
        if (t <= 0 || t / o.step_duration == 0) return FLO_OUT_OF_RANGE;
      }
      if (int s = impact(wall)) return s;
      break;
    }
  }
  flo_vec3 end{0, 0, 0};
  if (carrying == nullptr) {
    end = Advance(start, adv, o.step_duration);
  } else {
    const SweepPlane& wall = *carrying;
    end = Advance({wall.start * wall.normal.x, 0, 0}, wall.wall, o.step_duration);
    if (!CertifySide(end, wall.normal,
                     [&](flo_vec3 u) { return DistanceAlongPath(wall, u, {0, 0, 0}, o.step_duration); }))
      return FLO_OUT_OF_RANGE;
  }
  for (int k = 0; k < count; ++k) {
    const Bounds ed = DistanceAlongPath(walls[k], end, {0, 0, 0}, o.step_duration);
    if (!ed.finite() || ed.lower < 0) return FLO_OUT_OF_RANGE;
  }
  out_x = end.x;
  out_stored = stored.x;
  return FLO_OK;
}

// ResolveTankContact (liquid_wall_contact_math_3d.h:290-363): frictionless, per axis x, y, z.
int TankContact(const flo_update_options& o, flo_vec3 x, flo_vec3 stored, flo_vec3 adv, flo_vec3& position,
                flo_vec3& stored_out) {
  if (o.friction_coefficient != 0) return FLO_INVALID_ARGUMENT;
  const flo_tank& t = o.obstacle.tank;
  if (int s = CheckTankStep(t, o.step_duration)) return s;
  const SweepPlane walls[3][2] = {
      {{{1, 0, 0}, t.start_low_x, 0, {t.piston_velocity_x, 0, 0}}, {{-1, 0, 0}, -t.high_x, 0, {0, 0, 0}}},
      {{{1, 0, 0}, t.floor_y, 0, {0, 0, 0}}, {{0, 0, 0}, 0, 0, {0, 0, 0}}},
      {{{1, 0, 0}, t.low_z, 0, {0, 0, 0}}, {{-1, 0, 0}, -t.high_z, 0, {0, 0, 0}}}};
  const int counts[3] = {2, 1, 2};
  const double xs[3] = {x.x, x.y, x.z}, ss[3] = {stored.x, stored.y, stored.z}, as[3] = {adv.x, adv.y, adv.z};
  double px[3], sx[3];
  for (int a = 0; a < 3; ++a)
    if (int s = SweepTankAxis(o, walls[a], counts[a], xs[a], ss[a], as[a], px[a], sx[a])) return s;
  position = {px[0], px[1], px[2]};
  stored_out = {sx[0], sx[1], sx[2]};
  return FLO_OK;
}

int UpdateOne(const flo_update_options& o, const flo_boundary_gather& g, double liquid_distance,
              flo_liquid_particle& p) {
  const bool fallback = g.old_pic_fallback || g.new_pic_fallback || g.low_mass_fallback;
  const double alpha = p.pic_initialized && !fallback ? 0.99 * o.particle_energy_blend : 0;
  const flo_vec3 stored = Add(g.new_pic_velocity, Scale(g.particle_minus_old_pic_velocity, alpha));
  const double divergence = g.velocity_gradient.x_row.x + g.velocity_gradient.y_row.y + g.velocity_gradient.z_row.z;
  const double solid_distance = SolidDistance(o.obstacle, p.world_position);
  if (std::isnan(solid_distance)) return FLO_OUT_OF_RANGE;
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
  flo_vec3 position{}, new_stored = stored;
  if (o.obstacle.kind == FLO_OBSTACLE_ABSENT) {
    position = Advance(p.world_position, advecting, o.step_duration);
    if (!Finite3(position)) return FLO_OUT_OF_RANGE;
  } else if (o.obstacle.kind == FLO_OBSTACLE_HALFSPACE || o.obstacle.kind == FLO_OBSTACLE_SLAB) {
    if (int s = PlaneContact(o, p.world_position, stored, advecting, position, new_stored)) return s;
  } else if (o.obstacle.kind == FLO_OBSTACLE_TANK) {
    if (int s = TankContact(o, p.world_position, stored, advecting, position, new_stored)) return s;
  } else {
    return FLO_INVALID_ARGUMENT;
  }
  const flo_vec3 displacement{position.x - p.world_position.x, position.y - p.world_position.y,
                              position.z - p.world_position.z};
  if (MaxMagnitude(displacement) > 0.5 * o.cell_width) return FLO_ABORTED;
  const double volume_ratio = p.volume_ratio * std::exp(o.step_duration * divergence);
  if (!std::isfinite(volume_ratio) || volume_ratio <= 0) return FLO_ABORTED;
  p.world_position = position;
  p.world_velocity = new_stored;
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
