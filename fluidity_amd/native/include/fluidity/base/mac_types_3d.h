// Lattice and geometry records of the uniform staggered MAC grid.
//
// Same names, fields and layouts as native/include/fluidity/base/mac_types_3d.h:9-72 and
// native/include/fluidity/base/mac_transfer_types_3d.h:7-37, so callers of the reference operators
// compile unchanged; each record is layout-identical to its C-ABI twin in fluidity_hip.h.
#pragma once

#include <cstddef>
#include <cstdint>

namespace fluidity {

struct Vec3 {
  double x;
  double y;
  double z;
};

// Row-major, rows ordered by velocity component.
struct Mat3 {
  Vec3 x_row;
  Vec3 y_row;
  Vec3 z_row;
};

struct CellIndex3D {
  std::int32_t x;
  std::int32_t y;
  std::int32_t z;
};

// Owned 4x4x4 cell block.
struct BlockIndex3D {
  std::int32_t x;
  std::int32_t y;
  std::int32_t z;
};

// Sort bin: floor((X - 1.5) / 4) per axis, X in cells from the origin.
struct ParticleBinIndex3D {
  std::int32_t x;
  std::int32_t y;
  std::int32_t z;
};

struct FaceLatticeIndex3D {
  std::int32_t x;
  std::int32_t y;
  std::int32_t z;
};

enum class FaceNormalAxis : std::uint8_t {
  kX,
  kY,
  kZ,
};

// A cell owns the low face at its own lattice index.
struct FaceIndex3D {
  FaceNormalAxis axis;
  FaceLatticeIndex3D lattice_index;
};

struct UniformMacGeometry3D {
  Vec3 world_origin;  // lower corner of cell (0,0,0), meters
  double cell_width;  // positive normal spacing, meters
};

struct MacTransferParticle3D {
  Vec3 sorted_world_position;  // position at the last sort, meters
  Vec3 trial_world_position;   // position used by this transfer, meters
  double particle_mass;        // kilograms
  Vec3 particle_velocity;      // meters per second
  Mat3 particle_velocity_slopes;  // APIC slope rows, inverse seconds
};

struct MacDepositQuantization3D {
  double mass_per_quantum;           // kilograms per quantum, a positive normal power of two
  double pure_momentum_per_quantum;  // kilogram meters per second per quantum, a power of two
  std::size_t declared_particle_capacity;
};

struct MacFaceDeposit3D {
  double deposited_mass;           // kilograms
  double deposited_pure_momentum;  // normal momentum, kilogram meters per second
};

struct MacFaceDepositQuanta3D {
  std::int64_t mass_quanta_sum;
  std::int64_t pure_momentum_quanta_sum;
  std::int64_t stress_force_quanta_sum;  // zero for liquid
};

struct MacParticleGatherResult3D {
  Vec3 new_pic_velocity;
  Vec3 grid_velocity_increment;
  Mat3 fitted_velocity_slopes;
  Mat3 velocity_gradient;
};

// Bin and original ordinal of one particle in sorted order.
struct LiquidParticleSortEntry3D {
  ParticleBinIndex3D bin_index;
  std::uint64_t original_ordinal;
};

}  // namespace fluidity
