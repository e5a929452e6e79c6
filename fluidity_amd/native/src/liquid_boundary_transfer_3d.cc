// The compatible (masked) quantized deposit over the C-ABI (fl_p2g_compatible).
#include "fluidity/host/liquid_boundary_transfer_3d.h"

#include <cstring>
#include <mutex>

#include "internal.h"

namespace fluidity {
namespace {

static_assert(sizeof(LiquidBoundaryParticle3D) == sizeof(fl_boundary_particle) &&
              offsetof(LiquidBoundaryParticle3D, side_tag) == offsetof(fl_boundary_particle, side_tag) &&
              offsetof(LiquidBoundaryParticle3D, affine_enabled) == offsetof(fl_boundary_particle, affine_enabled));
static_assert(sizeof(LiquidBoundaryFaceTag3D) == sizeof(fl_face_tag) &&
              offsetof(LiquidBoundaryFaceTag3D, structurally_eligible) ==
                  offsetof(fl_face_tag, structurally_eligible));
static_assert(sizeof(CompatibleMacFaceSample3D) == sizeof(fl_compatible_sample) &&
              offsetof(CompatibleMacFaceSample3D, normalized_weight) ==
                  offsetof(fl_compatible_sample, normalized_weight) &&
              offsetof(CompatibleMacFaceSample3D, centered_face_offset) ==
                  offsetof(fl_compatible_sample, centered_face_offset));
static_assert(sizeof(std::array<CompatibleMacFaceStencil3D, 3>) == 81 * sizeof(fl_compatible_sample));
static_assert(sizeof(bool) == 1);

int Deposit(HipOperatorRuntime3D& rt, const LiquidBoundaryDepositInputs3D& in, std::span<MacFaceDepositQuanta3D> out,
            std::span<std::array<CompatibleMacFaceStencil3D, 3>> stencils, fl_measurement* m) {
  if (!stencils.empty() && stencils.size() != in.particles.size()) return FL_INVALID_ARGUMENT;
  const fl_geometry g = internal::ToC(in.mac_grid.geometry());
  const auto blocks = in.mac_grid.active_block_indices();
  const fl_quantization q{in.quantization.mass_per_quantum, in.quantization.pure_momentum_per_quantum,
                          in.quantization.declared_particle_capacity};
  return fl_p2g_compatible(rt.handle(), &g, internal::ToC(blocks), blocks.size(),
                           reinterpret_cast<const fl_boundary_particle*>(in.particles.data()), in.particles.size(),
                           reinterpret_cast<const fl_face_tag*>(in.face_tags.data()), in.face_tags.size(), q,
                           reinterpret_cast<fl_face_quanta*>(out.data()), out.size(),
                           stencils.empty() ? nullptr : reinterpret_cast<fl_compatible_sample*>(stencils.data()), m);
}

static_assert(sizeof(LiquidPrescribedWallSample3D) == sizeof(fl_wall_sample) &&
              offsetof(LiquidPrescribedWallSample3D, unit_normal) == offsetof(fl_wall_sample, unit_normal));
static_assert(sizeof(LiquidBoundaryParticleGather3D) == sizeof(fl_boundary_gather) &&
              offsetof(LiquidBoundaryParticleGather3D, affine_enabled) == offsetof(fl_boundary_gather, affine_enabled));
static_assert(sizeof(LiquidFaceValidity3D) == 1);

fl_velocity_history ToC(const LiquidBoundaryVelocityHistory3D& h, std::size_t& count_mismatch) {
  const std::size_t n = h.face_velocities.size();
  if (h.validity.size() != n || h.face_tags.size() != n || h.wall_samples.size() != n) count_mismatch = 1;
  return fl_velocity_history{h.velocity_sample_time,
                             h.wall_sample_time,
                             h.face_velocities.data(),
                             reinterpret_cast<const std::uint8_t*>(h.validity.data()),
                             reinterpret_cast<const fl_face_tag*>(h.face_tags.data()),
                             reinterpret_cast<const fl_wall_sample*>(h.wall_samples.data()),
                             n};
}

int Gather(HipOperatorRuntime3D& rt, const LiquidBoundaryGatherInputs3D& in, LiquidBoundaryGatherResults3D& out,
           fl_measurement* m) {
  std::size_t mismatch = 0;
  const fl_velocity_history oh = ToC(in.old_history, mismatch), nh = ToC(in.new_history, mismatch);
  if (mismatch) return FL_INVALID_ARGUMENT;
  const fl_geometry g = internal::ToC(in.mac_grid.geometry());
  const auto blocks = in.mac_grid.active_block_indices();
  std::vector<LiquidBoundaryParticleGather3D> staged(in.particles.size());
  std::uint64_t counts[3] = {0, 0, 0};
  const int code = fl_g2p_boundary(rt.handle(), &g, internal::ToC(blocks), blocks.size(),
                                   reinterpret_cast<const fl_boundary_particle*>(in.particles.data()),
                                   in.particles.size(), in.old_time, in.new_time, &oh, &nh,
                                   reinterpret_cast<fl_boundary_gather*>(staged.data()), counts, m);
  if (code != FL_OK) return code;
  out.particles = std::move(staged);
  out.old_pic_fallback_particle_count = counts[0];
  out.new_pic_fallback_particle_count = counts[1];
  out.low_mass_fallback_particle_count = counts[2];
  return FL_OK;
}

}  // namespace

StatusOr<LiquidBoundaryGatherResults3D> GatherLiquidBoundaryParticles3D(const LiquidBoundaryGatherInputs3D& inputs) {
  HipOperatorCapacities3D need;
  const auto blocks = inputs.mac_grid.active_block_indices();
  need.particles = std::max<std::size_t>(inputs.particles.size(), 1);
  need.blocks = std::max<std::size_t>(blocks.size(), 1);
  need.lookup_blocks = internal::LookupNeed(blocks);
  std::unique_lock<std::mutex> lock;
  for (int attempt = 0;; ++attempt) {
    auto rt = SharedHipRuntime3D::Acquire(need, lock);
    if (!rt.ok()) return rt.status();
    LiquidBoundaryGatherResults3D result;
    const int code = Gather(**rt, inputs, result, nullptr);
    if (code == FL_RESOURCE_EXHAUSTED && attempt < 12 && (*rt)->capacities().lookup_blocks < (std::size_t{1} << 30)) {
      need.lookup_blocks = (*rt)->capacities().lookup_blocks * 4;
      lock.unlock();
      continue;
    }
    if (code != FL_OK) return FromCode(code);
    return result;
  }
}

Status HipGatherLiquidBoundaryParticles3D(HipOperatorRuntime3D& runtime, const LiquidBoundaryGatherInputs3D& inputs,
                                          LiquidBoundaryGatherResults3D& results,
                                          HipOperatorMeasurement& measurement) {
  fl_measurement m{};
  const int code = Gather(runtime, inputs, results, &m);
  if (code == FL_OK) {
    std::memcpy(measurement.architecture, m.architecture, sizeof(measurement.architecture));
    measurement.kernel_ms = m.kernel_ms;
    measurement.copy_ms = m.copy_ms;
    measurement.launches = m.launches;
  }
  return FromCode(code);
}

StatusOr<LiquidBoundaryFaceDeposits3D> DepositCompatibleMacQuanta3D(const LiquidBoundaryDepositInputs3D& inputs) {
  if (inputs.quantization.declared_particle_capacity > 0xffffffffULL) return Status(StatusCode::kResourceExhausted);
  LiquidBoundaryFaceDeposits3D result;
  result.face_quanta.resize(inputs.mac_grid.face_count());
  result.compatible_stencils.resize(inputs.particles.size());
  HipOperatorCapacities3D need;
  const auto blocks = inputs.mac_grid.active_block_indices();
  need.particles = std::max<std::size_t>({inputs.particles.size(), inputs.quantization.declared_particle_capacity, 1});
  need.blocks = std::max<std::size_t>(blocks.size(), 1);
  need.lookup_blocks = internal::LookupNeed(blocks);
  std::unique_lock<std::mutex> lock;
  for (int attempt = 0;; ++attempt) {
    auto rt = SharedHipRuntime3D::Acquire(need, lock);
    if (!rt.ok()) return rt.status();
    const int code = Deposit(**rt, inputs, result.face_quanta, result.compatible_stencils, nullptr);
    if (code == FL_RESOURCE_EXHAUSTED && attempt < 12 && (*rt)->capacities().lookup_blocks < (std::size_t{1} << 30)) {
      need.lookup_blocks = (*rt)->capacities().lookup_blocks * 4;
      lock.unlock();
      continue;
    }
    if (code != FL_OK) return FromCode(code);
    return result;
  }
}

Status HipDepositCompatibleMacQuanta3D(HipOperatorRuntime3D& runtime, const LiquidBoundaryDepositInputs3D& inputs,
                                       std::span<MacFaceDepositQuanta3D> face_quanta,
                                       std::span<std::array<CompatibleMacFaceStencil3D, 3>> compatible_stencils,
                                       HipOperatorMeasurement& measurement) {
  fl_measurement m{};
  const int code = Deposit(runtime, inputs, face_quanta, compatible_stencils, &m);
  if (code == FL_OK) {
    std::memcpy(measurement.architecture, m.architecture, sizeof(measurement.architecture));
    measurement.kernel_ms = m.kernel_ms;
    measurement.copy_ms = m.copy_ms;
    measurement.launches = m.launches;
  }
  return FromCode(code);
}

}  // namespace fluidity
