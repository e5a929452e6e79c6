// Liquid particle update (native/src/liquid_substep_3d_cpu.cc:239-405) without regular-cell transport,
// for the absent, half-space, slab and open-tank obstacles: one lane per particle, FP64 throughout, the
// reference's operation order (no contraction except its explicit fma), first failing particle by
// atomic min of (ordinal, code). The obstacle gates the surface-expansion blend by its solid distance
// (native/src/liquid_substep_3d_staging.cc:197-205) and resolves the certified swept-plane contact
// (native/include/fluidity/base/liquid_wall_contact_math_3d.h:1-152,365-556) over bounded projected sums
// (native/include/fluidity/base/liquid_boundary_math.h:93-263). On error the device copy of the records
// is partly updated; the caller copies it back only on success.
#include "bounded.h"
#include "context.h"
#include "stdmath.h"

namespace fl {
namespace {

__device__ inline bool finite3(const fl_vec3& v) { return finite(v.x) && finite(v.y) && finite(v.z); }
__device__ inline fl_vec3 advance(fl_vec3 x, fl_vec3 v, double t) {
  return fl_vec3{fma(v.x, t, x.x), fma(v.y, t, x.y), fma(v.z, t, x.z)};
}
constexpr double kInf = __builtin_inf();
using namespace bounded;

struct Sweep {
  fl_vec3 normal;
  double start, half_width;
  fl_vec3 wall;
};
__device__ inline Bounds distance_along(const Sweep& p, fl_vec3 x, fl_vec3 v, double duration) {
  PSum d;
  d.add_dot(p.normal, x);
  d.add(-p.start);
  d.add(-p.half_width);
  d.add_motion(p.normal, v, p.wall, duration);
  return d.bounds();
}
__device__ inline Bounds normal_velocity(const Sweep& p, fl_vec3 v) {
  PSum s;
  s.add_motion(p.normal, v, p.wall, 1);
  return s.bounds();
}
__device__ inline bool representable_planes(const Bounds& c, double half_width) {
  if (!c.is_finite()) return false;
  if (half_width == 0) return true;
  const double low = c.estimate - half_width, high = c.estimate + half_width;
  return finite(c.lower - half_width) && finite(c.upper + half_width) && low < high;
}
// The next representable outward move with the largest normal progress.
__device__ inline fl_vec3 nudge_outward(fl_vec3 v, fl_vec3 n) {
  double values[3] = {v.x, v.y, v.z};
  const double normals[3] = {n.x, n.y, n.z};
  int best_axis = 0;
  double best_value = values[0], best_progress = 0;
  for (int a = 0; a < 3; ++a) {
    if (normals[a] == 0) continue;
    const double next = nextafter(values[a], copysign(kInf, normals[a]));
    const double progress = normals[a] * (next - values[a]);
    if (finite(next) && progress > best_progress) {
      best_axis = a;
      best_value = next;
      best_progress = progress;
    }
  }
  values[best_axis] = best_value;
  return fl_vec3{values[0], values[1], values[2]};
}
// At most two outward nudges to certify the kept side of the evaluated bound.
template <typename Evaluate>
__device__ inline bool certify_side(fl_vec3& c, fl_vec3 n, Evaluate evaluate) {
  if (!finite3(c)) return false;
  for (int k = 0; k < 2; ++k) {
    const Bounds b = evaluate(c);
    if (!b.is_finite()) return false;
    if (b.lower >= 0) return true;
    c = nudge_outward(c, n);
  }
  const Bounds b = evaluate(c);
  return b.is_finite() && b.lower >= 0;
}
// ApplyOneWayResponse: OutOfRange when uncertified.
__device__ inline int one_way(const Sweep& p, fl_vec3 v, double relative, double nls, double friction, fl_vec3& out) {
  if (relative >= 0) {
    out = v;
    return kOk;
  }
  fl_vec3 c = advance(v, p.normal, -relative / nls);
  const Bounds rounded = normal_velocity(p, c);
  if (!finite3(c) || !rounded.is_finite()) return kOutOfRange;
  c = advance(c, p.normal, -rounded.estimate / nls);
  if (friction > 0) {
    const fl_vec3 t{c.x - p.wall.x, c.y - p.wall.y, c.z - p.wall.z};
    const double tangent_speed = hypot3(t.x, t.y, t.z);
    const double inward_speed = -relative / sqrt(nls);
    if (!finite(tangent_speed) || inward_speed == 0) return kOutOfRange;
    if (tangent_speed == 0 || friction >= tangent_speed / inward_speed) {
      c = p.wall;
    } else {
      const double fraction = friction * inward_speed / tangent_speed;
      c = fl_vec3{lerp_std(c.x, p.wall.x, fraction), lerp_std(c.y, p.wall.y, fraction),
                  lerp_std(c.z, p.wall.z, fraction)};
    }
  }
  if (!certify_side(c, p.normal, [&](fl_vec3 u) { return normal_velocity(p, u); })) return kOutOfRange;
  out = c;
  return kOk;
}

// ResolveContact for a half-space or slab after the shared input checks: end position and stored
// velocity.
__device__ int plane_contact(const fl_particle_update_options& o, fl_vec3 x, fl_vec3 stored, fl_vec3 adv,
                             fl_vec3& position, fl_vec3& stored_out) {
  const fl_obstacle& ob = o.obstacle;
  const bool slab = ob.kind == FL_OBSTACLE_SLAB;
  if (slab && (!finite(ob.half_width) || ob.half_width <= 0)) return kInvalidArgument;
  Sweep plane{ob.unit_normal, ob.coordinate, slab ? ob.half_width : 0.0, ob.wall_velocity};
  const fl_vec3 n = plane.normal;
  const double nls = n.x * n.x + n.y * n.y + n.z * n.z;
  if (!finite3(n) || !finite3(plane.wall) || !finite(plane.start) || fabs(nls - 1) > 1e-12) return kInvalidArgument;
  PSum translated;
  translated.add(plane.start);
  if (!representable_planes(translated.bounds(), plane.half_width)) return kOutOfRange;
  translated.add_motion(n, plane.wall, fl_vec3{0, 0, 0}, o.step_duration);
  if (!representable_planes(translated.bounds(), plane.half_width)) return kOutOfRange;
  if (plane.half_width > 0) {
    PSum centre;
    centre.add_dot(n, x);
    centre.add(-plane.start);
    const Bounds cb = centre.bounds();
    if (!cb.is_finite()) return kOutOfRange;
    int side = 0;
    if (!cb.sign(&side) || side == 0) return kFailedPrecondition;
    if (side < 0) {
      plane.normal = fl_vec3{-n.x, -n.y, -n.z};
      plane.start = -plane.start;
    }
  }
  const fl_vec3 zero{0, 0, 0};
  const Bounds start_distance = distance_along(plane, x, zero, 0);
  const Bounds adv_projection = normal_velocity(plane, adv);
  const Bounds stored_projection = normal_velocity(plane, stored);
  int start_side = 0, adv_direction = 0, stored_direction = 0;
  if (!start_distance.sign(&start_side) || !adv_projection.sign(&adv_direction) ||
      !stored_projection.sign(&stored_direction))
    return kOutOfRange;
  const bool penetrating = start_side < 0;
  if (penetrating && start_distance.lower < -o.penetration_tolerance) return kFailedPrecondition;
  double elapsed = 0;
  bool hit = penetrating || (start_side == 0 && (adv_direction < 0 || stored_direction < 0));
  if (start_side > 0 && adv_direction < 0) {
    int end_side = 0;
    if (!distance_along(plane, x, adv, o.step_duration).sign(&end_side)) return kOutOfRange;
    if (end_side <= 0) {
      if (end_side < 0 && !start_distance.resolved()) return kOutOfRange;
      hit = true;
      if (end_side == 0) {
        elapsed = o.step_duration;
      } else {
        const double t = start_distance.estimate / -adv_projection.estimate;
        elapsed = t < o.step_duration ? t : o.step_duration;
      }
      if (elapsed <= 0 || elapsed / o.step_duration == 0) return kOutOfRange;
    }
  }
  const auto final_distance = [&](fl_vec3 u) { return distance_along(plane, u, zero, o.step_duration); };
  if (!hit) {
    position = advance(x, adv, o.step_duration);
    const Bounds accepted = final_distance(position);
    if (!finite3(position) || !accepted.is_finite() || accepted.lower < 0) return kOutOfRange;
    stored_out = stored;
    return kOk;
  }
  if (!stored_projection.resolved() || !adv_projection.resolved()) return kOutOfRange;
  fl_vec3 new_stored{}, new_adv{};
  if (int c = one_way(plane, stored, stored_projection.estimate, nls, o.friction_coefficient, new_stored)) return c;
  if (int c = one_way(plane, adv, adv_projection.estimate, nls, o.friction_coefficient, new_adv)) return c;
  fl_vec3 hit_position = advance(x, adv, elapsed);
  const auto hit_distance = [&](fl_vec3 u) { return distance_along(plane, u, zero, elapsed); };
  const Bounds rounded_hit = hit_distance(hit_position);
  if (!rounded_hit.is_finite()) return kOutOfRange;
  hit_position = advance(hit_position, plane.normal, -rounded_hit.estimate / nls);
  if (!certify_side(hit_position, plane.normal, hit_distance)) return kOutOfRange;
  position = advance(hit_position, new_adv, o.step_duration - elapsed);
  if (!certify_side(position, plane.normal, final_distance)) return kOutOfRange;
  const double nl = sqrt(nls);
  if (!finite(stored_projection.estimate / nl) || !finite(adv_projection.estimate / nl)) return kOutOfRange;
  stored_out = new_stored;
  return kOk;
}

// SolidDistance: the signed distance to the obstacle at its start (+infinity without one); for the tank
// the least inward clearance (SampleTankWall, liquid_boundary_math.h:319-350), NaN when one is not finite.
__device__ inline double solid_distance(const fl_obstacle& ob, fl_vec3 x) {
  const fl_vec3 n = ob.unit_normal;
  if (ob.kind == FL_OBSTACLE_HALFSPACE) return (n.x * x.x + n.y * x.y + n.z * x.z) - ob.coordinate;
  if (ob.kind == FL_OBSTACLE_SLAB) return fabs((n.x * x.x + n.y * x.y + n.z * x.z) - ob.coordinate) - ob.half_width;
  if (ob.kind == FL_OBSTACLE_TANK) {
    const fl_tank& t = ob.tank;
    const double d[5] = {x.x - t.start_low_x, t.high_x - x.x, x.y - t.floor_y, x.z - t.low_z, t.high_z - x.z};
    double least = d[0];
    for (int k = 0; k < 5; ++k) {
      if (!finite(d[k])) return __builtin_nan("");
      if (d[k] < least) least = d[k];
    }
    return least;
  }
  return kInf;
}

// CheckTankStep (liquid_boundary_math.h:252-276).
__device__ inline int check_tank_step(const fl_tank& t, double dt) {
  if (!finite(t.start_low_x) || !finite(t.high_x) || !finite(t.floor_y) || !finite(t.low_z) || !finite(t.high_z) ||
      !finite(t.piston_velocity_x))
    return kInvalidArgument;
  if (t.start_low_x >= t.high_x || t.low_z >= t.high_z) return kInvalidArgument;
  const double displacement = t.piston_velocity_x * dt;
  const double end_x = fma(t.piston_velocity_x, dt, t.start_low_x);
  const double start_width = t.high_x - t.start_low_x, end_width = t.high_x - end_x, width_z = t.high_z - t.low_z;
  if (!finite(displacement) || !finite(end_x) || !finite(start_width) || !finite(end_width) || !finite(width_z))
    return kOutOfRange;
  if (end_x >= t.high_x) return kFailedPrecondition;
  return kOk;
}

// SweepTankAxis (liquid_wall_contact_math_3d.h:154-289): one Cartesian component against its walls,
// each expressed along local x.
__device__ int sweep_tank_axis(const fl_particle_update_options& o, const Sweep* walls, int count, double x0,
                               double s0, double a0, double& out_x, double& out_stored) {
  const fl_vec3 zero{0, 0, 0};
  fl_vec3 start{x0, 0, 0}, stored{s0, 0, 0}, adv{a0, 0, 0};
  int carrying = -1;
  const auto impact = [&](int k) {
    const Sweep& wall = walls[k];
    const Bounds ss = normal_velocity(wall, stored), as = normal_velocity(wall, adv);
    if (!ss.resolved() || !as.resolved()) return kOutOfRange;
    if (ss.estimate < 0) stored.x = wall.wall.x;
    if (as.estimate <= 0) {
      adv.x = wall.wall.x;
      carrying = k;
    }
    return kOk;
  };
  for (int k = 0; k < count; ++k) {
    const Sweep& wall = walls[k];
    const Bounds sd = distance_along(wall, start, zero, 0);
    int side = 0;
    if (!sd.sign(&side)) return kOutOfRange;
    if (side > 0) continue;
    if (sd.lower < -o.penetration_tolerance) return kFailedPrecondition;
    const double wall_coordinate = wall.start * wall.normal.x;
    const bool low_wall = wall.normal.x > 0;
    const bool stored_in = low_wall ? stored.x < wall.wall.x : stored.x > wall.wall.x;
    const bool adv_in = low_wall ? adv.x < wall.wall.x : adv.x > wall.wall.x;
    if (side < 0 || stored_in || adv_in) {
      if (int c = impact(k)) return c;
      start.x = wall_coordinate;
    }
  }
  if (carrying < 0) {
    for (int k = 0; k < count; ++k) {
      const Sweep& wall = walls[k];
      const Bounds as = normal_velocity(wall, adv);
      int direction = 0;
      if (!as.sign(&direction)) return kOutOfRange;
      if (direction >= 0) continue;
      int end_side = 0;
      if (!distance_along(wall, start, adv, o.step_duration).sign(&end_side)) return kOutOfRange;
      if (end_side > 0) continue;
      if (end_side < 0) {
        const Bounds sd = distance_along(wall, start, zero, 0);
        if (!sd.resolved() || !as.resolved()) return kOutOfRange;
        const double q = sd.estimate / -as.estimate;
        const double t = q < o.step_duration ? q : o.step_duration;
        if (t <= 0 || t / o.step_duration == 0) return kOutOfRange;
      }
      if (int c = impact(k)) return c;
      break;
    }
  }
  fl_vec3 end{0, 0, 0};
  if (carrying < 0) {
    end = advance(start, adv, o.step_duration);
  } else {
    const Sweep& wall = walls[carrying];
    end = advance(fl_vec3{wall.start * wall.normal.x, 0, 0}, wall.wall, o.step_duration);
    if (!certify_side(end, wall.normal, [&](fl_vec3 u) { return distance_along(wall, u, zero, o.step_duration); }))
      return kOutOfRange;
  }
  for (int k = 0; k < count; ++k) {
    const Bounds ed = distance_along(walls[k], end, zero, o.step_duration);
    if (!ed.is_finite() || ed.lower < 0) return kOutOfRange;
  }
  out_x = end.x;
  out_stored = stored.x;
  return kOk;
}

// True when the tank step is valid and the particle stays farther from every wall than its whole
// move (and the piston's), by a margin far above the rounding of the plain differences: every wall
// distance that ResolveTankContact evaluates with bounded sums is then certainly positive at the start
// and at the end of the step, so it makes no contact and returns the plain advance.
__device__ inline bool tank_clear(const fl_particle_update_options& o, fl_vec3 x, fl_vec3 adv) {
  const fl_tank& t = o.obstacle.tank;
  if (o.friction_coefficient != 0 || check_tank_step(t, o.step_duration) != kOk) return false;
  const double dt = o.step_duration;
  const double reach = (fmax(fabs(adv.x), fmax(fabs(adv.y), fabs(adv.z))) + fabs(t.piston_velocity_x)) * dt;
  const double scale = fabs(x.x) + fabs(x.y) + fabs(x.z) + fabs(t.start_low_x) + fabs(t.high_x) + fabs(t.floor_y) +
                       fabs(t.low_z) + fabs(t.high_z) + 1.0;
  const double margin = reach * (1 + 1e-9) + 1e-12 * scale;
  if (!finite(margin)) return false;
  return x.x - t.start_low_x > margin && t.high_x - x.x > margin && x.y - t.floor_y > margin &&
         x.z - t.low_z > margin && t.high_z - x.z > margin;
}

// ResolveTankContact (liquid_wall_contact_math_3d.h:290-363): frictionless, per axis x, y, z.
__device__ int tank_contact(const fl_particle_update_options& o, fl_vec3 x, fl_vec3 stored, fl_vec3 adv,
                            fl_vec3& position, fl_vec3& stored_out) {
  if (o.friction_coefficient != 0) return kInvalidArgument;
  const fl_tank& t = o.obstacle.tank;
  if (int c = check_tank_step(t, o.step_duration)) return c;
  const fl_vec3 zero{0, 0, 0};
  const Sweep walls[3][2] = {
      {{fl_vec3{1, 0, 0}, t.start_low_x, 0, fl_vec3{t.piston_velocity_x, 0, 0}}, {fl_vec3{-1, 0, 0}, -t.high_x, 0, zero}},
      {{fl_vec3{1, 0, 0}, t.floor_y, 0, zero}, {zero, 0, 0, zero}},
      {{fl_vec3{1, 0, 0}, t.low_z, 0, zero}, {fl_vec3{-1, 0, 0}, -t.high_z, 0, zero}}};
  const int counts[3] = {2, 1, 2};
  const double xs[3] = {x.x, x.y, x.z}, ss[3] = {stored.x, stored.y, stored.z}, as[3] = {adv.x, adv.y, adv.z};
  double px[3], sx[3];
  for (int a = 0; a < 3; ++a)
    if (int c = sweep_tank_axis(o, walls[a], counts[a], xs[a], ss[a], as[a], px[a], sx[a])) return c;
  position = fl_vec3{px[0], px[1], px[2]};
  stored_out = fl_vec3{sx[0], sx[1], sx[2]};
  return kOk;
}

// One workgroup of kGroup lanes updates kGroup consecutive particles. Their 328-byte particle and
// 224-byte gather records are contiguous, 16-byte aligned blocks (kGroup * 328 and kGroup * 224 are
// multiples of 16), so they are copied to LDS and back with coalesced 16-byte accesses; each lane
// then works on its own records in LDS.
constexpr int kGroup = 64;
static_assert(kGroup * sizeof(fl_liquid_particle) % 16 == 0 && kGroup * sizeof(fl_boundary_gather) % 16 == 0,
              "16-byte blocks");

// kObstacle: 0 no obstacle (no contact code), 1 the half-space / slab instance, 2 the open-tank
// instance, 3 the open tank's clear-particle pass (no contact code: a particle near a wall returns
// kDeferred); each instance carries only its own contact code, so the plane sweep's registers do not
// weigh on the tank's kernels and the contact code's on the clear particles.
// One particle's update: ASFLIP stored velocity v = new PIC + alpha (v - old PIC), the surface-expansion
// advecting velocity (blend beta where the velocity diverges at the liquid surface away from the solid),
// the obstacle's contact or the plain advance, the half-cell guard and J exp(dt div u).
constexpr int kDeferred = -1;  // the clear-particle pass leaves this particle to the contact pass

struct UpdateIn {
  fl_vec3 x, new_pic, minus_old_pic;
  double divergence, phi, volume_ratio;
  bool pic_initialized, fallback;
};
template <int kObstacle>
__device__ inline int update_core(const fl_particle_update_options& o, const UpdateIn& in, fl_vec3& xn, fl_vec3& vn,
                                  double& j) {
  const double alpha = in.pic_initialized && !in.fallback ? 0.99 * o.particle_energy_blend : 0.0;
  const fl_vec3 np = in.new_pic, d = in.minus_old_pic;
  const fl_vec3 stored{np.x + alpha * d.x, np.y + alpha * d.y, np.z + alpha * d.z};
  const double divergence = in.divergence;
  const double solid = kObstacle ? solid_distance(o.obstacle, in.x) : kInf;
  const double phi = in.phi;
  const bool surface = divergence > 0 && phi <= 0 && phi >= -o.cell_width && solid > o.cell_width;
  const double beta = surface ? o.particle_energy_blend : 0.0;
  const double ba = beta * alpha;
  const fl_vec3 adv{np.x + ba * d.x, np.y + ba * d.y, np.z + ba * d.z};
  const fl_vec3 x = in.x;
  int code = kOk;
  if (isnan(solid)) code = kOutOfRange;
  else if (!finite3(x) || !finite3(stored) || !finite3(adv)) code = kInvalidArgument;
  else if (!finite(o.step_duration) || o.step_duration <= 0) code = kInvalidArgument;
  else if (!finite(o.friction_coefficient) || o.friction_coefficient < 0 || !finite(o.penetration_tolerance) ||
           o.penetration_tolerance < 0)
    code = kInvalidArgument;
  xn = fl_vec3{0.0, 0.0, 0.0};
  vn = stored;
  j = 0.0;
  if (code == kOk) {
    if (!kObstacle) {
      xn = advance(x, adv, o.step_duration);
      if (!finite3(xn)) code = kOutOfRange;
    } else if (kObstacle == 3) {
      // The clear-particle pass of the tank: particles near a wall go to the contact pass.
      if (!tank_clear(o, x, adv)) return kDeferred;
      xn = advance(x, adv, o.step_duration);
      if (!finite3(xn)) code = kOutOfRange;
    } else if (kObstacle == 2) {
      if (tank_clear(o, x, adv)) {
        // Far from every wall for the whole step: the per-axis sweep of ResolveTankContact finds the
        // particle inside every wall at both ends and advances it unchanged (fma of each component).
        xn = advance(x, adv, o.step_duration);
        if (!finite3(xn)) code = kOutOfRange;
      } else {
        code = tank_contact(o, x, stored, adv, xn, vn);
      }
    } else {
      code = plane_contact(o, x, stored, adv, xn, vn);
    }
  }
  if (code == kOk) {
    const double m = fmax(fmax(fabs(xn.x - x.x), fabs(xn.y - x.y)), fabs(xn.z - x.z));
    if (m > 0.5 * o.cell_width) code = kAborted;
  }
  if (code == kOk) {
    j = in.volume_ratio * exp(o.step_duration * divergence);
    if (!finite(j) || j <= 0) code = kAborted;
  }
  return code;
}

// The same update on particle columns in sort order (the device substep): positions, velocities and
// slopes in place, the gather results of GatherMac3D (24 component-major columns: new PIC, increment,
// fitted slopes, gradient; old PIC = new PIC - increment), the volume ratio and a per-particle flag
// byte (bit 0: PIC initialised). phi is the liquid field at every particle.
// rows: when set, lane t updates particle rows[t] of *count (the contact pass); deferred: when set,
// particles the clear pass leaves to the contact pass are appended there (count in deferred_count).
template <int kObstacle>
__global__ void k_update_columns(fl_particle_update_options o, uint64_t n, ParticleColumns P, const double* gather,
                                 uint64_t gcap, double phi, uint64_t* err, unsigned long long* max_speed,
                                 const uint32_t* rows, uint32_t* deferred, uint32_t* deferred_count) {
  const uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  const uint64_t limit = rows ? uint64_t(*deferred_count) : n;
  const uint64_t i = rows && t < limit ? rows[t] : t;
  double speed = 0.0;
  if (t < limit) {
    const fl_vec3 np{gather[i], gather[gcap + i], gather[2 * gcap + i]};
    const fl_vec3 inc{gather[3 * gcap + i], gather[4 * gcap + i], gather[5 * gcap + i]};
    const fl_vec3 v{P.v[0][i], P.v[1][i], P.v[2][i]};
    const fl_vec3 old{np.x - inc.x, np.y - inc.y, np.z - inc.z};
    const uint8_t flag = P.flags[i];
    const UpdateIn in{fl_vec3{P.x[0][i], P.x[1][i], P.x[2][i]},
                      np,
                      fl_vec3{v.x - old.x, v.y - old.y, v.z - old.z},
                      gather[15 * gcap + i] + gather[19 * gcap + i] + gather[23 * gcap + i],
                      phi,
                      P.volume_ratio[i],
                      (flag & 1) != 0,
                      false};
    fl_vec3 xn, vn;
    double j = 0.0;
    const int code = update_core<kObstacle>(o, in, xn, vn, j);
    if (code == kDeferred) {
      deferred[atomicAdd(deferred_count, 1u)] = static_cast<uint32_t>(i);
    } else if (code != kOk) {
      atomicMin(reinterpret_cast<unsigned long long*>(err),
                static_cast<unsigned long long>(error_word(P.ordinal ? P.ordinal[i] : i, code)));
    } else {
      P.x[0][i] = xn.x;
      P.x[1][i] = xn.y;
      P.x[2][i] = xn.z;
      P.v[0][i] = vn.x;
      P.v[1][i] = vn.y;
      P.v[2][i] = vn.z;
      for (int k = 0; k < 9; ++k) P.c[k][i] = gather[(6 + k) * gcap + i];
      P.volume_ratio[i] = j;
      P.flags[i] = static_cast<uint8_t>(flag | 1);
      speed = sqrt(vn.x * vn.x + vn.y * vn.y + vn.z * vn.z);
    }
  }
  // Largest speed of the updated particles (non-negative doubles order like their bit patterns).
  for (int off = 32; off > 0; off >>= 1) speed = fmax(speed, __shfl_xor(speed, off));
  if ((threadIdx.x & 63) == 0 && max_speed)
    atomicMax(max_speed, static_cast<unsigned long long>(__double_as_longlong(speed)));
}

template <int kObstacle>
__global__ __launch_bounds__(kGroup) void k_update_particles(fl_particle_update_options o,
                                                             const fl_boundary_gather* gathered,
                                                             const double* distances, uint64_t n,
                                                             fl_liquid_particle* particles, uint64_t* err) {
  __shared__ __align__(16) fl_liquid_particle sp[kGroup];
  __shared__ __align__(16) fl_boundary_gather sg[kGroup];
  const uint64_t i0 = blockIdx.x * uint64_t(kGroup);
  const int count = static_cast<int>(n - i0 < uint64_t(kGroup) ? n - i0 : uint64_t(kGroup));
  const int lane = threadIdx.x;
  {
    const int pw = count * int(sizeof(fl_liquid_particle)) / 8, gw = count * int(sizeof(fl_boundary_gather)) / 8;
    const uint64_t* src_p = reinterpret_cast<const uint64_t*>(particles + i0);
    const uint64_t* src_g = reinterpret_cast<const uint64_t*>(gathered + i0);
    uint64_t* dp = reinterpret_cast<uint64_t*>(sp);
    uint64_t* dg = reinterpret_cast<uint64_t*>(sg);
    if (count == kGroup) {  // whole 16-byte blocks
      for (int w = lane; w < pw / 2; w += kGroup)
        reinterpret_cast<ulonglong2*>(dp)[w] = reinterpret_cast<const ulonglong2*>(src_p)[w];
      for (int w = lane; w < gw / 2; w += kGroup)
        reinterpret_cast<ulonglong2*>(dg)[w] = reinterpret_cast<const ulonglong2*>(src_g)[w];
    } else {
      for (int w = lane; w < pw; w += kGroup) dp[w] = src_p[w];
      for (int w = lane; w < gw; w += kGroup) dg[w] = src_g[w];
    }
  }
  __syncthreads();
  if (lane < count) {
    const uint64_t i = i0 + lane;
    const fl_boundary_gather& g = sg[lane];
    fl_liquid_particle& p = sp[lane];
    const UpdateIn in{p.world_position,
                      g.new_pic_velocity,
                      g.particle_minus_old_pic_velocity,
                      g.velocity_gradient.x_row.x + g.velocity_gradient.y_row.y + g.velocity_gradient.z_row.z,
                      distances[i],
                      p.volume_ratio,
                      p.pic_initialized != 0,
                      g.old_pic_fallback || g.new_pic_fallback || g.low_mass_fallback};
    fl_vec3 xn, vn;
    double j = 0.0;
    const int code = update_core<kObstacle>(o, in, xn, vn, j);
    if (code != kOk) {
      atomicMin(reinterpret_cast<unsigned long long*>(err), static_cast<unsigned long long>(error_word(i, code)));
    } else {
      p.world_position = xn;
      p.world_velocity = vn;
      p.volume_ratio = j;
      p.velocity_slopes = g.fitted_velocity_slopes;
      p.affine_enabled = g.affine_enabled;
      p.pic_initialized = 1;
    }
  }
  __syncthreads();
  {
    const int pw = count * int(sizeof(fl_liquid_particle)) / 8;
    uint64_t* dst = reinterpret_cast<uint64_t*>(particles + i0);
    const uint64_t* sp64 = reinterpret_cast<const uint64_t*>(sp);
    if (count == kGroup) {
      for (int w = lane; w < pw / 2; w += kGroup)
        reinterpret_cast<ulonglong2*>(dst)[w] = reinterpret_cast<const ulonglong2*>(sp64)[w];
    } else {
      for (int w = lane; w < pw; w += kGroup) dst[w] = sp64[w];
    }
  }
}

}  // namespace

int update_particles(fl_ctx* ctx, const fl_particle_update_options& o, const fl_boundary_gather* gathered,
                     const double* distances, uint64_t n, fl_liquid_particle* particles) {
  if (!n) return FL_OK;
  if (o.obstacle.kind == FL_OBSTACLE_ABSENT)
    k_update_particles<0><<<grid_for(n, kGroup), kGroup, 0, ctx->stream>>>(o, gathered, distances, n, particles,
                                                                         ctx->err);
  else if (o.obstacle.kind == FL_OBSTACLE_TANK)
    k_update_particles<2><<<grid_for(n, kGroup), kGroup, 0, ctx->stream>>>(o, gathered, distances, n, particles,
                                                                         ctx->err);
  else
    k_update_particles<1><<<grid_for(n, kGroup), kGroup, 0, ctx->stream>>>(o, gathered, distances, n, particles,
                                                                         ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int update_columns(fl_ctx* ctx, const fl_particle_update_options& o, uint64_t n, const ParticleColumns& P,
                   const double* gather, uint64_t gcap, double phi, unsigned long long* max_speed, uint32_t* list,
                   uint32_t* list_count) {
  if (!n) return FL_OK;
  hipStream_t s = ctx->stream;
  if (o.obstacle.kind == FL_OBSTACLE_ABSENT) {
    k_update_columns<0><<<grid_for(n), kThreads, 0, s>>>(o, n, P, gather, gcap, phi, ctx->err, max_speed, nullptr,
                                                         nullptr, nullptr);
  } else if (o.obstacle.kind == FL_OBSTACLE_TANK) {
    // Clear particles first, then the contact pass over the ones near a wall.
    FL_HIP(hipMemsetAsync(list_count, 0, sizeof(uint32_t), s));
    k_update_columns<3><<<grid_for(n), kThreads, 0, s>>>(o, n, P, gather, gcap, phi, ctx->err, max_speed, nullptr,
                                                         list, list_count);
    uint32_t count = 0;
    FL_HIP(hipMemcpyAsync(&count, list_count, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    if (count)
      k_update_columns<2><<<grid_for(count), kThreads, 0, s>>>(o, n, P, gather, gcap, phi, ctx->err, max_speed,
                                                               list, nullptr, list_count);
  } else {
    k_update_columns<1><<<grid_for(n), kThreads, 0, s>>>(o, n, P, gather, gcap, phi, ctx->err, max_speed, nullptr,
                                                         nullptr, nullptr);
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

}  // namespace fl
