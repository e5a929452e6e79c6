// Topology preparation of a liquid step on the GPU (LiquidStepWorkspace3D,
// native/include/fluidity/host/liquid_step_workspace_3d.h; native/src/liquid_step_workspace_3d.cc:171-341).
//
// Same publication rule as the reference: a successful PrepareTrialTopology publishes a new bank, a
// failed one leaves the last published bank untouched. Particles enter by their world positions
// (the only particle field the topology reads); the previous state enters by its topology records,
// geometry and clock epoch, whose block, cell and face checks are those of ValidateLiquidState3D
// (native/src/liquid_state_3d.cc:559-604) — the rest of that state is the caller's to validate.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <vector>

#include "fluidity/base/liquid_boundary_transfer_types_3d.h"
#include "fluidity/base/mac_types_3d.h"
#include "fluidity/host/hip_operator_runtime_3d.h"
#include "fluidity/host/sparse_mac_grid_3d.h"
#include "fluidity/status.h"

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

// The topology of a previous LiquidState3D.
struct LiquidTopologyHistory3D {
  UniformMacGeometry3D geometry{};
  std::uint64_t epoch = 0;
  std::span<const LiquidBlockState3D> blocks;
  std::span<const LiquidCellState3D> cells;
  std::span<const LiquidFaceState3D> faces;
};

struct LiquidPreparedTopologyView3D {
  SparseMacGridView3D grid;
  std::span<const LiquidParticleSortEntry3D> particle_order;
  std::span<const LiquidBlockState3D> blocks;
  std::span<const LiquidCellState3D> cells;
  std::span<const LiquidFaceState3D> faces;
};

class LiquidStepWorkspace3D {
 public:
  // kInvalidArgument for invalid geometry.
  static StatusOr<LiquidStepWorkspace3D> Create(LiquidStateCapacities3D capacities, UniformMacGeometry3D geometry);

  // PrepareTopology without an imported order: sorted particle order, active blocks (81 stencil
  // owners per particle, 7 per pressure cell), canonical unique cells and one face per owned slot,
  // each remapped from previous by identity (liquid_step_workspace_3d.cc:171-341). kResourceExhausted
  // above the capacities; kInvalidArgument / kOutOfRange for invalid positions, cells, history or
  // geometry, in the reference's order.
  Status PrepareTrialTopology(std::span<const Vec3> world_positions, std::span<const CellIndex3D> pressure_cells,
                              const LiquidTopologyHistory3D* previous = nullptr);
  // Twin on an explicit runtime with the device measurement.
  Status HipPrepareTrialTopology(HipOperatorRuntime3D& runtime, std::span<const Vec3> world_positions,
                                 std::span<const CellIndex3D> pressure_cells, const LiquidTopologyHistory3D* previous,
                                 HipOperatorMeasurement& measurement);
  // kFailedPrecondition before the first successful preparation.
  StatusOr<LiquidPreparedTopologyView3D> prepared() const;
  LiquidStateCapacities3D capacities() const { return capacities_; }

 private:
  struct Bank {
    std::vector<LiquidParticleSortEntry3D> particle_order;
    std::vector<BlockIndex3D> block_keys;
    std::vector<LiquidBlockState3D> blocks;
    std::vector<LiquidCellState3D> cells;
    std::vector<LiquidFaceState3D> faces;
  };
  LiquidStepWorkspace3D(LiquidStateCapacities3D capacities, UniformMacGeometry3D geometry)
      : capacities_(capacities), geometry_(geometry) {}
  int Prepare(HipOperatorRuntime3D& runtime, std::span<const Vec3> positions, std::span<const CellIndex3D> cells,
              const LiquidTopologyHistory3D* previous, Bank& bank, void* measurement);
  LiquidStateCapacities3D capacities_;
  UniformMacGeometry3D geometry_;
  Bank banks_[2];
  int published_ = -1;
};

}  // namespace fluidity
