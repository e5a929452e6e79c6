// Oracle: the compatible (masked) quantized deposit (test infrastructure only).
// Restates native/src/liquid_boundary_transfer_3d.cc:329-460 (DepositCore) with the helpers of
// :28-41 (IsFinite, Dot, AddScaled, Difference), :44-75 (quantization checks), :80-91
// (RoundNearestEven) and :93-95 (IsCompatible). The reference's 1024-particle / 24576-face caps
// (liquid_boundary_transfer_3d.h:16-17) are replaced by the caller's max_capacity.
#include <array>
#include <cmath>
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
