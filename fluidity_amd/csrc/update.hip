// Liquid particle update without an obstacle (native/src/liquid_substep_3d_cpu.cc:239-405): one lane
// per particle, FP64 throughout, the reference's operation order (no contraction except its explicit
// fma advance), first failing particle by atomic min of (ordinal, code). On error the device copy of
// the records is partly updated; the caller copies it back only on success.
#include "context.h"

namespace fl {
namespace {

__device__ inline bool finite3(const fl_vec3& v) { return finite(v.x) && finite(v.y) && finite(v.z); }

// One workgroup of kGroup lanes updates kGroup consecutive particles. Their 328-byte particle and
// 224-byte gather records are contiguous, 16-byte aligned blocks (kGroup * 328 and kGroup * 224 are
// multiples of 16), so they are copied to LDS and back with coalesced 16-byte accesses; each lane
// then works on its own records in LDS.
constexpr int kGroup = 64;
static_assert(kGroup * sizeof(fl_liquid_particle) % 16 == 0 && kGroup * sizeof(fl_boundary_gather) % 16 == 0,
              "16-byte blocks");

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
    const bool fallback = g.old_pic_fallback || g.new_pic_fallback || g.low_mass_fallback;
    const double alpha = p.pic_initialized && !fallback ? 0.99 * o.particle_energy_blend : 0.0;
    const fl_vec3 np = g.new_pic_velocity, d = g.particle_minus_old_pic_velocity;
    const fl_vec3 stored{np.x + alpha * d.x, np.y + alpha * d.y, np.z + alpha * d.z};
    const double divergence =
        g.velocity_gradient.x_row.x + g.velocity_gradient.y_row.y + g.velocity_gradient.z_row.z;
    // Without an obstacle the solid distance is +infinity, so it always clears the cell width.
    const double phi = distances[i];
    const bool surface = divergence > 0 && phi <= 0 && phi >= -o.cell_width;
    const double beta = surface ? o.particle_energy_blend : 0.0;
    const double ba = beta * alpha;
    const fl_vec3 adv{np.x + ba * d.x, np.y + ba * d.y, np.z + ba * d.z};
    const fl_vec3 x = p.world_position;
    int code = kOk;
    if (!finite3(x) || !finite3(stored) || !finite3(adv)) code = kInvalidArgument;
    else if (!finite(o.step_duration) || o.step_duration <= 0) code = kInvalidArgument;
    else if (!finite(o.friction_coefficient) || o.friction_coefficient < 0 || !finite(o.penetration_tolerance) ||
             o.penetration_tolerance < 0)
      code = kInvalidArgument;
    fl_vec3 xn{0.0, 0.0, 0.0};
    double j = 0.0;
    if (code == kOk) {
      xn = fl_vec3{fma(adv.x, o.step_duration, x.x), fma(adv.y, o.step_duration, x.y),
                   fma(adv.z, o.step_duration, x.z)};
      if (!finite3(xn)) code = kOutOfRange;
    }
    if (code == kOk) {
      const double m = fmax(fmax(fabs(xn.x - x.x), fabs(xn.y - x.y)), fabs(xn.z - x.z));
      if (m > 0.5 * o.cell_width) code = kAborted;
    }
    if (code == kOk) {
      j = p.volume_ratio * exp(o.step_duration * divergence);
      if (!finite(j) || j <= 0) code = kAborted;
    }
    if (code != kOk) {
      atomicMin(reinterpret_cast<unsigned long long*>(err), static_cast<unsigned long long>(error_word(i, code)));
    } else {
      p.world_position = xn;
      p.world_velocity = stored;
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
  k_update_particles<<<grid_for(n, kGroup), kGroup, 0, ctx->stream>>>(o, gathered, distances, n, particles, ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

}  // namespace fl
