// The compatible (masked) quantized deposit, executed on the GPU
// (native/include/fluidity/host/liquid_boundary_transfer_3d.h:19-48).
//
// Same signature, argument meaning, status codes and publication rule as the reference. The
// reference's bounded capacities (1024 particles, 24576 faces) are replaced by the runtime's
// capacities; the quanta and the compatible stencils are bit-identical to the reference's.
#pragma once

#include <array>
#include <span>
#include <vector>

#include "fluidity/base/liquid_boundary_transfer_types_3d.h"
#include "fluidity/host/hip_operator_runtime_3d.h"
#include "fluidity/host/sparse_mac_grid_3d.h"
#include "fluidity/status.h"

namespace fluidity {

using CompatibleMacFaceStencil3D = std::array<CompatibleMacFaceSample3D, 27>;

struct LiquidBoundaryDepositInputs3D {
  const SparseMacGrid3D& mac_grid;
  std::span<const LiquidBoundaryParticle3D> particles;
  std::span<const LiquidBoundaryFaceTag3D> face_tags;  // Indexed by owned slot of the grid.
  MacDepositQuantization3D quantization;
};

struct LiquidBoundaryFaceDeposits3D {
  std::vector<MacFaceDepositQuanta3D> face_quanta;
  // Particle order, then x/y/z components; samples in stencil (z,y,x) order, x fastest.
  std::vector<std::array<CompatibleMacFaceStencil3D, 3>> compatible_stencils;
};

// DepositCompatibleMacQuanta3D (liquid_boundary_transfer_3d.cc:329-460, 674-689): per face axis only
// the samples whose slot is structurally eligible and carries the particle's side tag deposit, with
// weights normalised over them, recentred affine offsets and nearest-even per-contribution quanta.
// InvalidArgument for invalid data, extents or scales; FailedPrecondition for missing ownership or
// motion or no positive compatible support; OutOfRange for arithmetic or range failure;
// ResourceExhausted above the capacities. Errors publish nothing.
StatusOr<LiquidBoundaryFaceDeposits3D> DepositCompatibleMacQuanta3D(const LiquidBoundaryDepositInputs3D& inputs);

// Twin with an explicit runtime, caller-owned outputs and the device measurement. compatible_stencils
// may be empty (not formed) or hold one entry per particle.
Status HipDepositCompatibleMacQuanta3D(HipOperatorRuntime3D& runtime, const LiquidBoundaryDepositInputs3D& inputs,
                                       std::span<MacFaceDepositQuanta3D> face_quanta,
                                       std::span<std::array<CompatibleMacFaceStencil3D, 3>> compatible_stencils,
                                       HipOperatorMeasurement& measurement);

// One stage snapshot; every span covers all owned face slots.
struct LiquidBoundaryVelocityHistory3D {
  double velocity_sample_time;  // Seconds, must equal the requested stage time.
  double wall_sample_time;      // Seconds, must equal velocity_sample_time.
  std::span<const double> face_velocities;
  std::span<const LiquidFaceValidity3D> validity;
  std::span<const LiquidBoundaryFaceTag3D> face_tags;
  std::span<const LiquidPrescribedWallSample3D> wall_samples;
};

struct LiquidBoundaryGatherInputs3D {
  const SparseMacGrid3D& mac_grid;
  std::span<const LiquidBoundaryParticle3D> particles;
  double old_time;  // Time of the old (before forces) face snapshot, seconds; must be finite.
  double new_time;  // Finite and strictly greater than old_time.
  LiquidBoundaryVelocityHistory3D old_history;
  LiquidBoundaryVelocityHistory3D new_history;
};

struct LiquidBoundaryGatherResults3D {
  std::vector<LiquidBoundaryParticleGather3D> particles;
  std::size_t old_pic_fallback_particle_count = 0;
  std::size_t new_pic_fallback_particle_count = 0;
  std::size_t low_mass_fallback_particle_count = 0;
};

// GatherLiquidBoundaryParticles3D (liquid_boundary_transfer_3d.cc:519-651): incompatible or unusable
// samples are continued from same-side donors within 2h (rank-revealing affine fit, nearest donor,
// compatible wall, oblique-wall normal replacement); retention, affine switch and fallbacks as the
// reference. Results agree with the reference to 1e-12 relative; flags and counts exactly.
StatusOr<LiquidBoundaryGatherResults3D> GatherLiquidBoundaryParticles3D(const LiquidBoundaryGatherInputs3D& inputs);
Status HipGatherLiquidBoundaryParticles3D(HipOperatorRuntime3D& runtime, const LiquidBoundaryGatherInputs3D& inputs,
                                          LiquidBoundaryGatherResults3D& results,
                                          HipOperatorMeasurement& measurement);

}  // namespace fluidity
