// Span-based P2G / G2P / normalize operators over the C-ABI.
#include "fluidity/host/transfers_3d.h"

#include <algorithm>
#include <cstring>

#include "internal.h"

namespace fluidity {
namespace {

void Publish(const fl_measurement& m, HipOperatorMeasurement& out) {
  std::memcpy(out.architecture, m.architecture, sizeof(out.architecture));
  out.kernel_ms = m.kernel_ms;
  out.copy_ms = m.copy_ms;
  out.launches = m.launches;
}

fl_quantization ToC(MacDepositQuantization3D q) {
  return fl_quantization{q.mass_per_quantum, q.pure_momentum_per_quantum, q.declared_particle_capacity};
}

// Runs `call(runtime)` on the shared runtime, growing the dense tables while they are the limit.
template <class Call>
Status OnShared(HipOperatorCapacities3D need, Call call) {
  std::unique_lock<std::mutex> lock;
  for (int attempt = 0;; ++attempt) {
    auto rt = SharedHipRuntime3D::Acquire(need, lock);
    if (!rt.ok()) return rt.status();
    const int code = call(**rt);
    if (code == FL_RESOURCE_EXHAUSTED && attempt < 12 && (*rt)->capacities().lookup_blocks < (std::size_t{1} << 30)) {
      need.lookup_blocks = (*rt)->capacities().lookup_blocks * 4;
      lock.unlock();
      continue;
    }
    return FromCode(code);
  }
}

int Deposit(HipOperatorRuntime3D& rt, const SparseMacGrid3D& grid, std::span<const MacTransferParticle3D> particles,
            MacDepositQuantization3D q, std::span<MacFaceDepositQuanta3D> out, fl_measurement* m) {
  const fl_geometry g = internal::ToC(grid.geometry());
  return fl_p2g_quanta(rt.handle(), &g, internal::ToC(grid.active_block_indices()), grid.active_block_indices().size(),
                       reinterpret_cast<const fl_particle*>(particles.data()), particles.size(), ToC(q),
                       reinterpret_cast<fl_face_quanta*>(out.data()), out.size(), m);
}

int DepositF64(HipOperatorRuntime3D& rt, const SparseMacGrid3D& grid, std::span<const MacTransferParticle3D> particles,
               std::span<MacFaceDeposit3D> out, fl_measurement* m) {
  const fl_geometry g = internal::ToC(grid.geometry());
  return fl_p2g(rt.handle(), &g, internal::ToC(grid.active_block_indices()), grid.active_block_indices().size(),
                reinterpret_cast<const fl_particle*>(particles.data()), particles.size(),
                reinterpret_cast<fl_face_deposit*>(out.data()), out.size(), m);
}

int Gather(HipOperatorRuntime3D& rt, const SparseMacGrid3D& grid, std::span<const MacTransferParticle3D> particles,
           std::span<const double> u_old, std::span<const double> u_new, std::span<MacParticleGatherResult3D> out,
           fl_measurement* m) {
  if (u_old.size() != u_new.size() || out.size() != particles.size()) return FL_INVALID_ARGUMENT;
  const fl_geometry g = internal::ToC(grid.geometry());
  return fl_g2p(rt.handle(), &g, internal::ToC(grid.active_block_indices()), grid.active_block_indices().size(),
                reinterpret_cast<const fl_particle*>(particles.data()), particles.size(), u_old.data(), u_new.data(),
                u_old.size(), reinterpret_cast<fl_gather_result*>(out.data()), m);
}

HipOperatorCapacities3D TransferNeed(const SparseMacGrid3D& grid, std::size_t particles) {
  HipOperatorCapacities3D need;
  need.particles = std::max<std::size_t>(particles, 1);
  need.blocks = std::max<std::size_t>(grid.active_block_indices().size(), 1);
  need.lookup_blocks = internal::LookupNeed(grid.active_block_indices());
  return need;
}

}  // namespace

Status DepositMacQuanta3D(const SparseMacGrid3D& grid, std::span<const MacTransferParticle3D> particles,
                          MacDepositQuantization3D q, std::span<MacFaceDepositQuanta3D> out) {
  // The declared capacity must fit the runtime; the reference caps it at 2^20, this layer at 2^32 - 1.
  if (q.declared_particle_capacity > 0xffffffffULL) return Status(StatusCode::kResourceExhausted);
  return OnShared(TransferNeed(grid, std::max(particles.size(), q.declared_particle_capacity)),
                  [&](HipOperatorRuntime3D& rt) { return Deposit(rt, grid, particles, q, out, nullptr); });
}

Status DepositMac3D(const SparseMacGrid3D& grid, std::span<const MacTransferParticle3D> particles,
                    std::span<MacFaceDeposit3D> out) {
  return OnShared(TransferNeed(grid, particles.size()),
                  [&](HipOperatorRuntime3D& rt) { return DepositF64(rt, grid, particles, out, nullptr); });
}

Status HipDepositMac3D(HipOperatorRuntime3D& rt, const SparseMacGrid3D& grid,
                       std::span<const MacTransferParticle3D> particles, std::span<MacFaceDeposit3D> out,
                       HipOperatorMeasurement& measurement) {
  fl_measurement m{};
  const int code = DepositF64(rt, grid, particles, out, &m);
  if (code == FL_OK) Publish(m, measurement);
  return FromCode(code);
}

Status GatherMac3D(const SparseMacGrid3D& grid, std::span<const MacTransferParticle3D> particles,
                   std::span<const double> u_old, std::span<const double> u_new,
                   std::span<MacParticleGatherResult3D> out) {
  return OnShared(TransferNeed(grid, particles.size()),
                  [&](HipOperatorRuntime3D& rt) { return Gather(rt, grid, particles, u_old, u_new, out, nullptr); });
}

Status HipDepositMacQuanta3D(HipOperatorRuntime3D& rt, const SparseMacGrid3D& grid,
                             std::span<const MacTransferParticle3D> particles, MacDepositQuantization3D q,
                             std::span<MacFaceDepositQuanta3D> out, HipOperatorMeasurement& measurement) {
  fl_measurement m{};
  const int code = Deposit(rt, grid, particles, q, out, &m);
  if (code == FL_OK) Publish(m, measurement);
  return FromCode(code);
}

Status HipGatherMac3D(HipOperatorRuntime3D& rt, const SparseMacGrid3D& grid,
                      std::span<const MacTransferParticle3D> particles, std::span<const double> u_old,
                      std::span<const double> u_new, std::span<MacParticleGatherResult3D> out,
                      HipOperatorMeasurement& measurement) {
  fl_measurement m{};
  const int code = Gather(rt, grid, particles, u_old, u_new, out, &m);
  if (code == FL_OK) Publish(m, measurement);
  return FromCode(code);
}

Status HipNormalizeApplyForces3D(HipOperatorRuntime3D& rt, std::span<const MacFaceDepositQuanta3D> quanta,
                                 MacDepositQuantization3D q, double minimum_face_mass, Vec3 g, double dt,
                                 const NormalizedLiquidFaces3D& out, HipOperatorMeasurement& measurement) {
  const std::size_t n = quanta.size();
  auto sized = [n](std::size_t s) { return s == 0 || s == n; };
  if (!sized(out.face_masses.size()) || !sized(out.face_velocities.size()) || !sized(out.validity.size()) ||
      !sized(out.predicted_velocities.size()))
    return Status(StatusCode::kInvalidArgument);
  auto ptr = [](auto span) { return span.empty() ? nullptr : span.data(); };
  fl_measurement m{};
  const int code = fl_normalize_forces(rt.handle(), reinterpret_cast<const fl_face_quanta*>(quanta.data()), n, ToC(q),
                                       minimum_face_mass, fl_vec3{g.x, g.y, g.z}, dt, ptr(out.face_masses),
                                       ptr(out.face_velocities), reinterpret_cast<std::uint8_t*>(ptr(out.validity)),
                                       ptr(out.predicted_velocities), &m);
  if (code == FL_OK) Publish(m, measurement);
  return FromCode(code);
}

}  // namespace fluidity
