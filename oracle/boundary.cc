// Oracle: the compatible (masked) quantized deposit (test infrastructure only).
// Restates native/src/liquid_boundary_transfer_3d.cc:329-460 (DepositCore) with the helpers of
// :28-41 (IsFinite, Dot, AddScaled, Difference), :44-75 (quantization checks), :80-91
// (RoundNearestEven) and :93-95 (IsCompatible). The reference's 1024-particle / 24576-face caps
// (liquid_boundary_transfer_3d.h:16-17) are replaced by the caller's max_capacity.
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <utility>
#include <vector>

#include "grid_internal.h"

using namespace flo;

namespace {

bool Compatible(const flo_face_tag& tag, int64_t side) { return tag.structurally_eligible != 0 && tag.side_tag == side; }

// liquid_boundary_transfer_3d.cc:80-91
int64_t RoundNearestEven(double v) {
  const double mag = std::fabs(v);
  const double whole = std::floor(mag);
  const double frac = mag - whole;
  auto q = static_cast<int64_t>(whole);
  if (frac > 0.5 || (frac == 0.5 && (q % 2) != 0)) ++q;
  return v < 0.0 ? -q : q;
}

bool PowerOfTwoScale(double s) {
  int e = 0;
  return std::isnormal(s) && std::frexp(s, &e) == 0.5;
}

}  // namespace

extern "C" {

int flo_deposit_compatible(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count,
                           const flo_boundary_particle* particles, size_t n, const flo_face_tag* tags,
                           size_t tag_count, flo_quantization q, uint64_t max_capacity, flo_face_quanta* out,
                           size_t face_count, flo_compatible_sample* stencils) {
  // :330-341 capacity, quantization, tag coverage.
  if (n > max_capacity) return FLO_RESOURCE_EXHAUSTED;
  if (!PowerOfTwoScale(q.mass_per_quantum) || !PowerOfTwoScale(q.pure_momentum_per_quantum))
    return FLO_INVALID_ARGUMENT;
  if (q.declared_particle_capacity > max_capacity || n > q.declared_particle_capacity) return FLO_RESOURCE_EXHAUSTED;
  if (face_count != block_count * kFacesPerBlock) return FLO_INVALID_ARGUMENT;
  if (tag_count != face_count) return FLO_INVALID_ARGUMENT;
  std::vector<flo_face_quanta> quanta(face_count, flo_face_quanta{0, 0, 0});
  std::vector<flo_compatible_sample> staged(stencils ? n * 81 : 0, flo_compatible_sample{});
  const uint64_t cap = std::max<uint64_t>(1, q.declared_particle_capacity);
  int bits = 0;
  for (uint64_t v = cap - 1; v; v >>= 1) ++bits;  // std::bit_width(cap - 1)
  const double limit = std::ldexp(1.0, 62 - bits);
  for (size_t i = 0; i < n; ++i) {
    const flo_boundary_particle& bp = particles[i];
    const flo_particle& p = bp.transfer;
    // :346-352
    if (!std::isfinite(p.particle_mass) || p.particle_mass < 0.0 || !Finite3(p.particle_velocity))
      return FLO_INVALID_ARGUMENT;
    for (int axis = 0; axis < 3; ++axis) {
      std::array<flo_stencil_sample, 27> st;
      if (int s = OwnedStencil(*geometry, sorted_blocks, block_count, axis, p.sorted_world_position,
                               p.trial_world_position, st))
        return s;
      const flo_vec3 row = bp.affine_enabled ? Row(p.particle_velocity_slopes, axis) : flo_vec3{0.0, 0.0, 0.0};
      if (!Finite3(row)) return FLO_INVALID_ARGUMENT;
      // :376-383 compatible sums in stencil order.
      double wsum = 0.0;
      flo_vec3 gsum{0.0, 0.0, 0.0};
      bool ok[27];
      for (int k = 0; k < 27; ++k) {
        ok[k] = Compatible(tags[st[k].face_slot], bp.side_tag);
        if (!ok[k]) continue;
        wsum += st[k].weight;
        gsum.x += st[k].weight_gradient.x * 1.0;
        gsum.y += st[k].weight_gradient.y * 1.0;
        gsum.z += st[k].weight_gradient.z * 1.0;
      }
      if (!std::isfinite(wsum) || !Finite3(gsum)) return FLO_OUT_OF_RANGE;
      if (wsum <= 0.0) return FLO_FAILED_PRECONDITION;
      // :397-419 normalized basis and its quotient-rule gradient; weighted mean offset.
      flo_compatible_sample cs[27] = {};
      flo_vec3 mean{0.0, 0.0, 0.0};
      for (int k = 0; k < 27; ++k) {
        cs[k].face_slot = st[k].face_slot;
        cs[k].compatible = ok[k] ? 1 : 0;
        if (!ok[k]) continue;
        const double nw = st[k].weight / wsum;
        cs[k].normalized_weight = nw;
        cs[k].normalized_weight_gradient = {(st[k].weight_gradient.x - nw * gsum.x) / wsum,
                                            (st[k].weight_gradient.y - nw * gsum.y) / wsum,
                                            (st[k].weight_gradient.z - nw * gsum.z) / wsum};
        if (!std::isfinite(nw) || !Finite3(cs[k].normalized_weight_gradient)) return FLO_OUT_OF_RANGE;
        mean.x += st[k].face_offset_from_particle.x * nw;
        mean.y += st[k].face_offset_from_particle.y * nw;
        mean.z += st[k].face_offset_from_particle.z * nw;
      }
      if (!Finite3(mean)) return FLO_OUT_OF_RANGE;
      // :423-456 recentred offsets and the quantized contributions.
      const double v = Get(p.particle_velocity, axis);
      for (int k = 0; k < 27; ++k) {
        if (!ok[k]) continue;
        const flo_vec3 r = st[k].face_offset_from_particle;
        cs[k].centered_face_offset = {r.x - mean.x, r.y - mean.y, r.z - mean.z};
        const flo_vec3 c = cs[k].centered_face_offset;
        if (!Finite3(c)) return FLO_OUT_OF_RANGE;
        const double mass = p.particle_mass * cs[k].normalized_weight;
        const double vel = v + ((row.x * c.x + row.y * c.y) + row.z * c.z);
        const double mom = mass * vel;
        const double sm = mass / q.mass_per_quantum;
        const double sp = mom / q.pure_momentum_per_quantum;
        if (!std::isfinite(sm) || std::fabs(sm) > limit || !std::isfinite(sp) || std::fabs(sp) > limit)
          return FLO_OUT_OF_RANGE;
        quanta[st[k].face_slot].mass_quanta_sum += RoundNearestEven(sm);
        quanta[st[k].face_slot].pure_momentum_quanta_sum += RoundNearestEven(sp);
      }
      if (stencils)
        for (int k = 0; k < 27; ++k) staged[(i * 3 + axis) * 27 + k] = cs[k];
    }
  }
  std::copy(quanta.begin(), quanta.end(), out);
  if (stencils) std::copy(staged.begin(), staged.end(), stencils);
  return FLO_OK;
}

}  // extern "C"

// ---- Boundary gather (GatherLiquidBoundaryParticles3D): liquid_boundary_transfer_3d.cc:519-651 with
// ValidateHistory (:97-145), BuildOwnedAnchors (:152-177), ContinueComponent (:181-240), FillVelocity
// (:242-306) and the rank-revealing QR fit of base/liquid_boundary_transfer_math_3d.h:34-144.
namespace {

constexpr size_t kMaxDonors = 268;  // liquid_boundary_transfer_3d.cc:25

bool Usable(uint8_t v) { return v == 1 || v == 2; }  // valid or predicted

double DotL(flo_vec3 a, flo_vec3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

int ValidateHistory(const flo_history& h, size_t face_count, double stage_time) {
  if (h.velocity_sample_time != stage_time || h.wall_sample_time != stage_time) return FLO_INVALID_ARGUMENT;
  if (h.count != face_count) return FLO_INVALID_ARGUMENT;
  for (size_t s = 0; s < face_count; ++s) {
    const uint8_t v = h.validity[s];
    if (v != 0 && v != 3 && !Usable(v)) return FLO_INVALID_ARGUMENT;
    if (Usable(v) && !std::isfinite(h.face_velocities[s])) return FLO_INVALID_ARGUMENT;
    const flo_wall_sample& w = h.wall_samples[s];
    if (w.available) {
      if (!Finite3(w.wall_velocity) || !Finite3(w.unit_normal)) return FLO_INVALID_ARGUMENT;
      if (std::fabs(DotL(w.unit_normal, w.unit_normal) - 1.0) > 1e-12) return FLO_INVALID_ARGUMENT;
    }
  }
  return FLO_OK;
}

struct Anchor {
  int axis;
  flo_vec3 at;  // in cell widths
};

std::vector<Anchor> OwnedAnchors(const flo_int3* blocks, size_t count) {
  std::vector<Anchor> a;
  a.reserve(count * kFacesPerBlock);
  for (size_t b = 0; b < count; ++b)
    for (int axis = 0; axis < 3; ++axis)
      for (int z = 0; z < 4; ++z)
        for (int y = 0; y < 4; ++y)
          for (int x = 0; x < 4; ++x)
            a.push_back({axis,
                         {4.0 * blocks[b].x + x + (axis == 0 ? 0.0 : 0.5), 4.0 * blocks[b].y + y + (axis == 1 ? 0.0 : 0.5),
                          4.0 * blocks[b].z + z + (axis == 2 ? 0.0 : 0.5)}});
  return a;
}

struct Fit {
  bool full_rank = false;
  double at_target = 0.0;
  flo_vec3 slopes{0.0, 0.0, 0.0};  // per cell width
};

// FitBoundedLiquidGhostAffineCore3D<125>: column-pivoted Householder QR of [1, dx, dy, dz] against the
// velocities scaled by their largest magnitude; recomputed column norms by hypot accumulation, strict
// pivot comparison, rank loss below 1e-10 of the largest pivot. Returns InvalidArgument for more
// than 125 rows, OutOfRange for non-finite arithmetic.
int FitAffine(const std::vector<std::pair<flo_vec3, double>>& rows, Fit& out) {
  out = Fit{};
  const size_t n = rows.size();
  if (n > 125) return FLO_INVALID_ARGUMENT;
  std::vector<std::array<double, 4>> m(n);
  std::vector<double> rhs(n, 0.0);
  double scale = 0.0;
  for (size_t r = 0; r < n; ++r) {
    if (!Finite3(rows[r].first) || !std::isfinite(rows[r].second)) return FLO_INVALID_ARGUMENT;
    m[r] = {1.0, rows[r].first.x, rows[r].first.y, rows[r].first.z};
    scale = std::max(scale, std::fabs(rows[r].second));
  }
  if (n < 4) return FLO_OK;
  if (scale == 0.0) scale = 1.0;
  for (size_t r = 0; r < n; ++r) rhs[r] = rows[r].second / scale;
  std::array<size_t, 4> order{0, 1, 2, 3};
  double largest = 0.0;
  std::vector<double> refl(n);
  for (size_t pv = 0; pv < 4; ++pv) {
    size_t sel = pv;
    double sel_norm = 0.0;
    for (size_t c = pv; c < 4; ++c) {
      double norm = 0.0;
      for (size_t r = pv; r < n; ++r) norm = std::hypot(norm, m[r][c]);
      if (norm > sel_norm) {
        sel_norm = norm;
        sel = c;
      }
    }
    if (!std::isfinite(sel_norm)) return FLO_OUT_OF_RANGE;
    largest = std::max(largest, sel_norm);
    if (sel_norm == 0.0 || sel_norm < 1e-10 * largest) return FLO_OK;
    std::swap(order[pv], order[sel]);
    for (size_t r = 0; r < n; ++r) std::swap(m[r][pv], m[r][sel]);
    const double diag = -std::copysign(sel_norm, m[pv][pv]);
    double rnorm = 0.0;
    for (size_t r = pv; r < n; ++r) {
      refl[r] = m[r][pv] / sel_norm;
      if (r == pv) refl[r] -= diag / sel_norm;
      rnorm = std::hypot(rnorm, refl[r]);
    }
    for (size_t r = pv; r < n; ++r) refl[r] /= rnorm;
    for (size_t c = pv + 1; c < 4; ++c) {
      double proj = 0.0;
      for (size_t r = pv; r < n; ++r) proj += refl[r] * m[r][c];
      for (size_t r = pv; r < n; ++r) {
        m[r][c] -= 2.0 * refl[r] * proj;
        if (!std::isfinite(m[r][c])) return FLO_OUT_OF_RANGE;
      }
    }
    double rproj = 0.0;
    for (size_t r = pv; r < n; ++r) rproj += refl[r] * rhs[r];
    for (size_t r = pv; r < n; ++r) rhs[r] -= 2.0 * refl[r] * rproj;
    m[pv][pv] = diag;
  }
  std::array<double, 4> pc{}, coef{};
  for (int pv = 3; pv >= 0; --pv) {
    double res = rhs[pv];
    for (int c = pv + 1; c < 4; ++c) res -= m[pv][c] * pc[c];
    pc[pv] = res / m[pv][pv];
    coef[order[pv]] = pc[pv] * scale;
    if (!std::isfinite(coef[order[pv]])) return FLO_OUT_OF_RANGE;
  }
  out.full_rank = true;
  out.at_target = coef[0];
  out.slopes = {coef[1], coef[2], coef[3]};
  return FLO_OK;
}

struct Continued {
  double velocity = 0.0;
  bool fallback = false;
};

int ContinueComponent(const flo_history& h, const std::vector<Anchor>& anchors, flo_vec3 target, int component,
                      int64_t side, const flo_wall_sample& wall, double squared_radius, Continued& out) {
  std::vector<std::pair<flo_vec3, double>> donors;
  size_t nearest = anchors.size();
  double nearest_d2 = std::numeric_limits<double>::infinity();
  for (size_t s = 0; s < anchors.size(); ++s) {
    if (anchors[s].axis != component || !Compatible(h.face_tags[s], side) || !Usable(h.validity[s])) continue;
    const flo_vec3 off{anchors[s].at.x - target.x, anchors[s].at.y - target.y, anchors[s].at.z - target.z};
    const double d2 = DotL(off, off);
    if (d2 > squared_radius) continue;
    if (donors.size() == kMaxDonors) return FLO_RESOURCE_EXHAUSTED;
    donors.push_back({off, h.face_velocities[s]});
    if (d2 < nearest_d2) {
      nearest = s;
      nearest_d2 = d2;
    }
  }
  Fit fit;
  if (int e = FitAffine(donors, fit)) return e;
  if (fit.full_rank) {
    out = {fit.at_target, false};
    return FLO_OK;
  }
  if (nearest != anchors.size()) {
    out = {h.face_velocities[nearest], true};
    return FLO_OK;
  }
  if (wall.available && wall.side_tag == side) {
    out = {Get(wall.wall_velocity, component), true};
    return FLO_OK;
  }
  return FLO_FAILED_PRECONDITION;
}

int FillVelocity(const flo_history& h, const std::vector<Anchor>& anchors, size_t slot, int component, int64_t side,
                 Continued& out) {
  if (Compatible(h.face_tags[slot], side) && Usable(h.validity[slot])) {
    out = {h.face_velocities[slot], false};
    return FLO_OK;
  }
  const flo_wall_sample& wall = h.wall_samples[slot];
  const flo_vec3 target = anchors[slot].at;
  if (int e = ContinueComponent(h, anchors, target, component, side, wall, 4.0, out)) return e;
  const bool compatible_wall = wall.available && wall.side_tag == side;
  const double nc = Get(wall.unit_normal, component);
  if (compatible_wall && nc != 0.0) {
    flo_vec3 fitted = wall.wall_velocity;
    Set(fitted, component, out.velocity);
    for (int other = 0; other < 3; ++other) {
      if (other == component || Get(wall.unit_normal, other) == 0.0) continue;
      Continued o;
      if (int e = ContinueComponent(h, anchors, target, other, side, wall, 4.0, o)) return e;
      Set(fitted, other, o.velocity);
      out.fallback = out.fallback || o.fallback;
    }
    const flo_vec3 rel{fitted.x - wall.wall_velocity.x, fitted.y - wall.wall_velocity.y, fitted.z - wall.wall_velocity.z};
    out.velocity -= DotL(rel, wall.unit_normal) * nc;
    if (!std::isfinite(out.velocity)) return FLO_OUT_OF_RANGE;
  }
  return FLO_OK;
}

}  // namespace

extern "C" {

// internal::FitLiquidGhostAffine3D (liquid_boundary_transfer_3d.cc:311-323) on caller rows.
int flo_fit_ghost_affine(const flo_vec3* offsets, const double* velocities, size_t n, int* full_rank,
                         double* at_target, flo_vec3* slopes) {
  std::vector<std::pair<flo_vec3, double>> rows(n);
  for (size_t i = 0; i < n; ++i) rows[i] = {offsets[i], velocities[i]};
  Fit f;
  if (int e = FitAffine(rows, f)) return e;
  *full_rank = f.full_rank ? 1 : 0;
  *at_target = f.at_target;
  *slopes = f.slopes;
  return FLO_OK;
}

int flo_gather_boundary(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count,
                        const flo_boundary_particle* particles, size_t n, double old_time, double new_time,
                        const flo_history* old_history, const flo_history* new_history,
                        flo_boundary_gather* out, uint64_t* out_counts) {
  const size_t nf = block_count * kFacesPerBlock;
  if (!std::isfinite(old_time) || !std::isfinite(new_time) || new_time <= old_time) return FLO_INVALID_ARGUMENT;
  if (int e = ValidateHistory(*old_history, nf, old_time)) return e;
  if (int e = ValidateHistory(*new_history, nf, new_time)) return e;
  const bool same = old_history->face_velocities == new_history->face_velocities &&
                    old_history->validity == new_history->validity && old_history->face_tags == new_history->face_tags &&
                    old_history->wall_samples == new_history->wall_samples;
  const std::vector<Anchor> anchors = OwnedAnchors(sorted_blocks, block_count);
  std::vector<flo_boundary_gather> staged(n, flo_boundary_gather{});
  std::vector<double> fo(nf, 0.0), fn(nf, 0.0);
  uint64_t counts[3] = {0, 0, 0};
  const double h = geometry->cell_width;
  for (size_t i = 0; i < n; ++i) {
    const flo_boundary_particle& bp = particles[i];
    const flo_particle& p = bp.transfer;
    if (!Finite3(p.particle_velocity)) return FLO_INVALID_ARGUMENT;
    flo_boundary_gather& g = staged[i];
    flo_vec3 old_pic{0.0, 0.0, 0.0};
    for (int axis = 0; axis < 3; ++axis) {
      std::array<flo_stencil_sample, 27> st;
      if (int s = OwnedStencil(*geometry, sorted_blocks, block_count, axis, p.sorted_world_position,
                               p.trial_world_position, st))
        return s;
      for (const auto& smp : st) {
        Continued co, cn;
        if (int e = FillVelocity(*old_history, anchors, smp.face_slot, axis, bp.side_tag, co)) return e;
        if (same) {
          cn = co;
        } else if (int e = FillVelocity(*new_history, anchors, smp.face_slot, axis, bp.side_tag, cn)) {
          return e;
        }
        fo[smp.face_slot] = co.velocity;
        fn[smp.face_slot] = cn.velocity;
        Set(old_pic, axis, Get(old_pic, axis) + smp.weight * co.velocity);
        g.old_pic_fallback = g.old_pic_fallback || co.fallback;
        g.new_pic_fallback = g.new_pic_fallback || cn.fallback;
        g.low_mass_fallback = g.low_mass_fallback || old_history->validity[smp.face_slot] == 3 ||
                              new_history->validity[smp.face_slot] == 3;
      }
    }
    for (int axis = 0; axis < 3; ++axis) {
      std::array<flo_stencil_sample, 27> st;
      if (int s = OwnedStencil(*geometry, sorted_blocks, block_count, axis, p.sorted_world_position,
                               p.trial_world_position, st))
        return s;
      double pic, inc;
      flo_vec3 slope, grad;
      if (int e = GatherComponent(st, h, fo.data(), fn.data(), pic, inc, slope, grad)) return e;
      Set(g.new_pic_velocity, axis, pic);
      Set(g.grid_velocity_increment, axis, inc);
      SetRow(g.fitted_velocity_slopes, axis, slope);
      SetRow(g.velocity_gradient, axis, grad);
    }
    const bool fallback = g.old_pic_fallback || g.new_pic_fallback || g.low_mass_fallback;
    g.affine_enabled = bp.affine_enabled && !fallback;
    if (!fallback)
      g.particle_minus_old_pic_velocity = {p.particle_velocity.x - old_pic.x, p.particle_velocity.y - old_pic.y,
                                           p.particle_velocity.z - old_pic.z};
    if (!Finite3(g.particle_minus_old_pic_velocity)) return FLO_OUT_OF_RANGE;
    counts[0] += g.old_pic_fallback;
    counts[1] += g.new_pic_fallback;
    counts[2] += g.low_mass_fallback;
  }
  std::copy(staged.begin(), staged.end(), out);
  if (out_counts) std::copy(counts, counts + 3, out_counts);
  return FLO_OK;
}

}  // extern "C"
