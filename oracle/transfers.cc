// Oracle: APIC transfers (test infrastructure only).
// Restates native/src/transfers_3d.cc:44-261, mac_deposit_3d_internal.h:25-75,
// mac_gather_3d_internal.h:15-90, liquid_boundary_transfer_3d.cc:463-516 (normalize) and
// liquid_substep_3d_cpu.cc:156-178 (forces).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "grid_internal.h"

using namespace flo;

namespace {

bool FiniteMat(const flo_mat3& m) { return Finite3(m.x_row) && Finite3(m.y_row) && Finite3(m.z_row); }

// mac_deposit_3d_internal.h:25-37
int ValidateParticle(const flo_particle& p) {
  if (!std::isfinite(p.particle_mass) || p.particle_mass < 0.0) return FLO_INVALID_ARGUMENT;
  if (!Finite3(p.particle_velocity) || !FiniteMat(p.particle_velocity_slopes)) return FLO_INVALID_ARGUMENT;
  return FLO_OK;
}

// transfers_3d.cc:32-35
bool PositivePowerOfTwo(double s) {
  int e;
  return std::isnormal(s) && std::frexp(s, &e) == 0.5;
}

// transfers_3d.cc:44-54: round half to even, independent of the ambient rounding mode.
int64_t RoundHalfEven(double v) {
  const double mag = std::fabs(v);
  const double whole = std::floor(mag);
  const double frac = mag - whole;
  auto q = static_cast<int64_t>(whole);
  if (frac > 0.5 || (frac == 0.5 && (q % 2) != 0)) ++q;
  return v < 0.0 ? -q : q;
}

bool Overlaps(const void* a, size_t an, const void* b, size_t bn) {
  if (an == 0 || bn == 0) return false;
  const auto* pa = static_cast<const unsigned char*>(a);
  const auto* pb = static_cast<const unsigned char*>(b);
  return pa < pb + bn && pb < pa + an;
}

// One particle's three components in tensor order (mac_deposit_3d_internal.h:41-65).
// Each contribution is handed to `sink(slot, mass, momentum)`, which may fail.
template <typename Sink>
int DepositParticle(const flo_geometry& g, const flo_int3* blocks, size_t nb, const flo_particle& p, Sink&& sink) {
  if (int s = ValidateParticle(p)) return s;
  for (int axis = 0; axis < 3; ++axis) {
    std::array<flo_stencil_sample, 27> st;
    if (int s = OwnedStencil(g, blocks, nb, axis, p.sorted_world_position, p.trial_world_position, st)) return s;
    const double v = Get(p.particle_velocity, axis);
    const flo_vec3 c = Row(p.particle_velocity_slopes, axis);
    for (const auto& smp : st) {
      const flo_vec3 r = smp.face_offset_from_particle;
      const double mass = smp.weight * p.particle_mass;
      const double inc = (c.x * r.x + c.y * r.y) + c.z * r.z;
      const double vel = v + inc;
      const double mom = mass * vel;
      if (int s = sink(smp.face_slot, mass, mom)) return s;
    }
  }
  return FLO_OK;
}

struct FirstError {
  size_t index = SIZE_MAX;
  int code = FLO_OK;
};

template <typename Work>
FirstError RunChunks(size_t n, int threads, Work&& work) {
  threads = std::max(1, threads);
  if (threads == 1 || n < 1024) {
    FirstError e;
    work(0, n, 0, e);
    return e;
  }
  std::vector<FirstError> errors(threads);
  std::vector<std::thread> pool;
  const size_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const size_t lo = std::min(n, t * chunk), hi = std::min(n, lo + chunk);
    pool.emplace_back([&, lo, hi, t] { work(lo, hi, t, errors[t]); });
  }
  for (auto& th : pool) th.join();
  FirstError best;
  for (const auto& e : errors)
    if (e.code != FLO_OK && e.index < best.index) best = e;
  return best;
}

}  // namespace

extern "C" {

// transfers_3d.cc:139-145 via DepositWithFaceAccumulator (:56-104)
int flo_deposit(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count,
                const flo_particle* particles, size_t n, flo_face_deposit* out, size_t face_count) {
  if (face_count != block_count * kFacesPerBlock) return FLO_INVALID_ARGUMENT;
  if (Overlaps(out, face_count * sizeof(flo_face_deposit), particles, n * sizeof(flo_particle)))
    return FLO_INVALID_ARGUMENT;
  std::vector<flo_face_deposit> staged(face_count, flo_face_deposit{0.0, 0.0});
  for (size_t i = 0; i < n; ++i) {
    int s = DepositParticle(*geometry, sorted_blocks, block_count, particles[i],
                            [&](uint64_t slot, double m, double q) {
                              auto& f = staged[slot];
                              f.deposited_mass += m;
                              f.deposited_pure_momentum += q;
                              if (!std::isfinite(f.deposited_mass) || !std::isfinite(f.deposited_pure_momentum))
                                return static_cast<int>(FLO_OUT_OF_RANGE);
                              return static_cast<int>(FLO_OK);
                            });
    if (s) return s;
  }
  std::copy(staged.begin(), staged.end(), out);
  return FLO_OK;
}

// transfers_3d.cc:147-197. The contribution limit follows :164-168.
int flo_deposit_quanta(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count,
                       const flo_particle* particles, size_t n, flo_quantization q, uint64_t max_capacity,
                       flo_face_quanta* out, size_t face_count, int threads) {
  if (!PositivePowerOfTwo(q.mass_per_quantum) || !PositivePowerOfTwo(q.pure_momentum_per_quantum))
    return FLO_INVALID_ARGUMENT;
  if (q.declared_particle_capacity > max_capacity) return FLO_RESOURCE_EXHAUSTED;
  if (n > q.declared_particle_capacity) return FLO_RESOURCE_EXHAUSTED;
  const uint64_t cap = std::max<uint64_t>(1, q.declared_particle_capacity);
  int bits = 0;
  for (uint64_t v = cap - 1; v; v >>= 1) ++bits;  // std::bit_width(cap - 1)
  const double limit = std::ldexp(1.0, 62 - bits);
  if (face_count != block_count * kFacesPerBlock) return FLO_INVALID_ARGUMENT;
  if (Overlaps(out, face_count * sizeof(flo_face_quanta), particles, n * sizeof(flo_particle)))
    return FLO_INVALID_ARGUMENT;
  threads = std::max(1, threads);
  std::vector<std::vector<flo_face_quanta>> partial(threads == 1 || n < 1024 ? 1 : threads);
  for (auto& v : partial) v.assign(face_count, flo_face_quanta{0, 0, 0});
  const FirstError err = RunChunks(n, threads, [&](size_t lo, size_t hi, int t, FirstError& e) {
    auto& acc = partial[t];
    for (size_t i = lo; i < hi; ++i) {
      int s = DepositParticle(*geometry, sorted_blocks, block_count, particles[i],
                              [&](uint64_t slot, double m, double p) {
                                const double sm = m / q.mass_per_quantum;
                                const double sp = p / q.pure_momentum_per_quantum;
                                const bool ok_m = std::isfinite(sm) && std::fabs(sm) <= limit;
                                const bool ok_p = std::isfinite(sp) && std::fabs(sp) <= limit;
                                if (!ok_m || !ok_p) return static_cast<int>(FLO_OUT_OF_RANGE);
                                acc[slot].mass_quanta_sum += RoundHalfEven(sm);
                                acc[slot].pure_momentum_quanta_sum += RoundHalfEven(sp);
                                return static_cast<int>(FLO_OK);
                              });
      if (s) {
        e = {i, s};
        return;
      }
    }
  });
  if (err.code != FLO_OK) return err.code;
  for (size_t t = 1; t < partial.size(); ++t)
    for (size_t f = 0; f < face_count; ++f) {
      partial[0][f].mass_quanta_sum += partial[t][f].mass_quanta_sum;
      partial[0][f].pure_momentum_quanta_sum += partial[t][f].pure_momentum_quanta_sum;
    }
  std::copy(partial[0].begin(), partial[0].end(), out);
  return FLO_OK;
}

// transfers_3d.cc:199-261 with GatherMacVelocityComponentFromSlots (mac_gather_3d_internal.h:15-69)
int flo_gather(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count,
               const flo_particle* particles, size_t n, const double* old_faces, const double* new_faces,
               size_t face_count, flo_gather_result* out, int threads) {
  if (face_count != block_count * kFacesPerBlock) return FLO_INVALID_ARGUMENT;
  if (Overlaps(out, n * sizeof(flo_gather_result), particles, n * sizeof(flo_particle)) ||
      Overlaps(out, n * sizeof(flo_gather_result), old_faces, face_count * sizeof(double)) ||
      Overlaps(out, n * sizeof(flo_gather_result), new_faces, face_count * sizeof(double)))
    return FLO_INVALID_ARGUMENT;
  for (size_t f = 0; f < face_count; ++f)
    if (!std::isfinite(old_faces[f]) || !std::isfinite(new_faces[f])) return FLO_INVALID_ARGUMENT;
  const double h = geometry->cell_width;
  std::vector<flo_gather_result> staged(n);
  const FirstError err = RunChunks(n, threads, [&](size_t lo, size_t hi, int, FirstError& e) {
    for (size_t i = lo; i < hi; ++i) {
      const flo_particle& p = particles[i];
      flo_gather_result& res = staged[i];
      for (int axis = 0; axis < 3; ++axis) {
        std::array<flo_stencil_sample, 27> st;
        int s = OwnedStencil(*geometry, sorted_blocks, block_count, axis, p.sorted_world_position,
                             p.trial_world_position, st);
        if (s) {
          e = {i, s};
          return;
        }
        double new_pic, inc;
        flo_vec3 slope, grad;
        if (int g = GatherComponent(st, h, old_faces, new_faces, new_pic, inc, slope, grad)) {
          e = {i, g};
          return;
        }
        Set(res.new_pic_velocity, axis, new_pic);
        Set(res.grid_velocity_increment, axis, inc);
        SetRow(res.fitted_velocity_slopes, axis, slope);
        SetRow(res.velocity_gradient, axis, grad);
      }
    }
  });
  if (err.code != FLO_OK) return err.code;
  std::copy(staged.begin(), staged.end(), out);
  return FLO_OK;
}

// liquid_boundary_transfer_3d.cc:463-516 (validity: 0 missing, 1 valid, 3 low mass)
int flo_normalize(const flo_face_quanta* quanta, size_t face_count, flo_quantization q, double minimum_face_mass,
                  double* out_mass, double* out_velocity, uint8_t* out_validity) {
  if (!PositivePowerOfTwo(q.mass_per_quantum) || !PositivePowerOfTwo(q.pure_momentum_per_quantum))
    return FLO_INVALID_ARGUMENT;
  if (!std::isfinite(minimum_face_mass) || minimum_face_mass < 0.0) return FLO_INVALID_ARGUMENT;
  constexpr int64_t kMax = int64_t{1} << 62;
  std::vector<double> m(face_count, 0.0), u(face_count, 0.0);
  std::vector<uint8_t> v(face_count, 0);
  for (size_t f = 0; f < face_count; ++f) {
    const auto& Q = quanta[f];
    if (Q.mass_quanta_sum < 0 || Q.mass_quanta_sum > kMax || Q.pure_momentum_quanta_sum < -kMax ||
        Q.pure_momentum_quanta_sum > kMax || Q.stress_force_quanta_sum != 0)
      return FLO_INVALID_ARGUMENT;
    const double dm = static_cast<double>(Q.mass_quanta_sum) * q.mass_per_quantum;
    const double dq = static_cast<double>(Q.pure_momentum_quanta_sum) * q.pure_momentum_per_quantum;
    if (!std::isfinite(dm) || !std::isfinite(dq)) return FLO_OUT_OF_RANGE;
    m[f] = dm;
    if (dm == 0.0) continue;
    if (dm <= minimum_face_mass) {
      v[f] = 3;
      continue;
    }
    const double vel = dq / dm;
    if (!std::isfinite(vel)) return FLO_OUT_OF_RANGE;
    u[f] = vel;
    v[f] = 1;
  }
  if (out_mass) std::copy(m.begin(), m.end(), out_mass);
  if (out_velocity) std::copy(u.begin(), u.end(), out_velocity);
  if (out_validity) std::copy(v.begin(), v.end(), out_validity);
  return FLO_OK;
}

// liquid_substep_3d_cpu.cc:156-178: predicted = old + dt * g[axis(slot)]
int flo_apply_forces(const flo_int3* sorted_blocks, size_t block_count, const double* old_velocities,
                     size_t face_count, flo_vec3 acceleration, double dt, double* out_predicted) {
  (void)sorted_blocks;
  if (face_count != block_count * kFacesPerBlock) return FLO_INVALID_ARGUMENT;
  std::vector<double> out(face_count);
  for (size_t f = 0; f < face_count; ++f) {
    const int axis = static_cast<int>((f % kFacesPerBlock) / 64);
    out[f] = old_velocities[f] + dt * Get(acceleration, axis);
    if (!std::isfinite(out[f])) return FLO_OUT_OF_RANGE;
  }
  std::copy(out.begin(), out.end(), out_predicted);
  return FLO_OK;
}

}  // extern "C"
