// APIC transfers on the sparse MAC grid, executed on the GPU
// (native/include/fluidity/host/transfers_3d.h:36-61 and cuda_transfers_3d.h:18-39).
//
// Same signatures, argument meaning, status codes and publication rule as the reference: outputs are
// written only on success, every failure preserves them. Differences: the reference's 2^20 particle
// cap on DepositMacQuanta3D is replaced by the runtime's particle capacity; gather results agree
// with the reference to 1e-12 relative (the device contracts the 27-sample sums axis by axis).
#pragma once

#include <span>

#include "fluidity/base/liquid_boundary_transfer_types_3d.h"
#include "fluidity/base/mac_types_3d.h"
#include "fluidity/host/hip_operator_runtime_3d.h"
#include "fluidity/host/sparse_mac_grid_3d.h"
#include "fluidity/status.h"

namespace fluidity {

// DepositMac3D (transfers_3d.cc:139-145): FP64 sums, formed in parallel, so they agree with the
// reference's particle-order sums to rounding (1e-12 relative) rather than bit for bit.
Status DepositMac3D(const SparseMacGrid3D& mac_grid, std::span<const MacTransferParticle3D> particles,
                    std::span<MacFaceDeposit3D> face_deposits);
Status HipDepositMac3D(HipOperatorRuntime3D& runtime, const SparseMacGrid3D& mac_grid,
                       std::span<const MacTransferParticle3D> particles, std::span<MacFaceDeposit3D> face_deposits,
                       HipOperatorMeasurement& measurement);

// DepositMacQuanta3D (transfers_3d.cc:147-197): per contribution rounded half to even to int64
// quanta, then summed; identical bits to the reference for any input.
Status DepositMacQuanta3D(const SparseMacGrid3D& mac_grid, std::span<const MacTransferParticle3D> particles,
                          MacDepositQuantization3D quantization, std::span<MacFaceDepositQuanta3D> face_deposits);

// GatherMac3D (transfers_3d.cc:199-261).
Status GatherMac3D(const SparseMacGrid3D& mac_grid, std::span<const MacTransferParticle3D> particles,
                   std::span<const double> old_face_velocities, std::span<const double> new_face_velocities,
                   std::span<MacParticleGatherResult3D> particle_gather_results);

// Twins with an explicit runtime and the device measurement (CudaDepositMacQuanta3D / CudaGatherMac3D).
Status HipDepositMacQuanta3D(HipOperatorRuntime3D& runtime, const SparseMacGrid3D& mac_grid,
                             std::span<const MacTransferParticle3D> particles, MacDepositQuantization3D quantization,
                             std::span<MacFaceDepositQuanta3D> face_deposits, HipOperatorMeasurement& measurement);
Status HipGatherMac3D(HipOperatorRuntime3D& runtime, const SparseMacGrid3D& mac_grid,
                      std::span<const MacTransferParticle3D> particles, std::span<const double> old_face_velocities,
                      std::span<const double> new_face_velocities,
                      std::span<MacParticleGatherResult3D> particle_gather_results,
                      HipOperatorMeasurement& measurement);

// Normalized face velocities of the liquid deposit (NormalizeLiquidMacFaces3D,
// liquid_boundary_transfer_3d.cc:463-516) followed by gravity (ApplyForces, liquid_substep_3d_cpu.cc:156-178).
struct NormalizedLiquidFaces3D {
  std::span<double> face_masses;             // kilograms, may be empty
  std::span<double> face_velocities;         // u0 = momentum / mass, m/s, may be empty
  std::span<LiquidFaceValidity3D> validity;  // may be empty
  std::span<double> predicted_velocities;    // u* = u0 + dt g_axis, may be empty
};
Status HipNormalizeApplyForces3D(HipOperatorRuntime3D& runtime, std::span<const MacFaceDepositQuanta3D> face_quanta,
                                 MacDepositQuantization3D quantization, double minimum_face_mass,
                                 Vec3 world_acceleration, double step_duration, const NormalizedLiquidFaces3D& out,
                                 HipOperatorMeasurement& measurement);

}  // namespace fluidity
