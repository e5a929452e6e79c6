// Liquid state records of the hot path (native/include/fluidity/host/liquid_state_3d.h:129-236): the
// block, cell and face records of the sparse topology, the clock, and the subset of LiquidState3D the
// step stages read (geometry, clock, particles and topology records; emission sources, adaptive and
// cavity records belong to subsystems outside this library).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "fluidity/base/liquid_boundary_transfer_types_3d.h"
#include "fluidity/base/mac_types_3d.h"
#include "fluidity/host/liquid_particle_update_3d.h"

namespace fluidity {

// Records of native/include/fluidity/host/liquid_state_3d.h:129-159.
struct LiquidBlockState3D {
  BlockIndex3D block_index{};
  std::uint64_t generation = 0;
  std::uint32_t activation_age = 0;  // Accepted steps since activation.
  std::uint32_t inactive_age = 0;    // Hysteresis age in accepted steps.
  bool operator==(const LiquidBlockState3D&) const;
};
enum class LiquidCellOccupancy3D : std::uint8_t { kAir, kLiquid, kSolid };
struct LiquidCellState3D {
  CellIndex3D cell_index{};
  LiquidCellOccupancy3D occupancy = LiquidCellOccupancy3D::kLiquid;
  double previous_pressure = 0.0;
  bool operator==(const LiquidCellState3D&) const;
};
struct LiquidFaceState3D {
  FaceIndex3D face_index{};
  std::size_t face_slot = 0;
  double carrier_velocity = 0.0;
  LiquidFaceValidity3D validity = LiquidFaceValidity3D::kMissing;
  std::int64_t side_tag = 0;
  bool structurally_eligible = true;
  bool operator==(const LiquidFaceState3D&) const;
};
struct LiquidStateCapacities3D {
  std::size_t particles = 0;
  std::size_t blocks = 0;
  std::size_t cells = 0;
  std::size_t faces = 0;
  std::size_t sources = 0;
};

struct LiquidClock3D {
  double time = 0.0;                        // seconds, finite and nonnegative
  std::uint64_t frame_ordinal = 0;          // completed frames
  std::uint64_t substep_ordinal = 0;        // accepted steps in the current frame
  std::uint64_t accepted_step_ordinal = 0;  // key of the accepted step
  std::uint64_t epoch = 0;                  // advances once per accepted step
};

// The state a step reads and the topology stage writes back (ApplyTrialTopology). Vectors are sized by
// the owner; their size is the valid extent.
struct LiquidState3D {
  UniformMacGeometry3D geometry{{0.0, 0.0, 0.0}, 1.0};
  LiquidClock3D clock;
  std::uint64_t random_seed = 0;
  std::uint64_t next_particle_id = 1;
  std::vector<LiquidParticleState3D> particles;
  std::vector<LiquidBlockState3D> blocks;
  std::vector<LiquidCellState3D> cells;
  std::vector<LiquidFaceState3D> faces;
};

}  // namespace fluidity
