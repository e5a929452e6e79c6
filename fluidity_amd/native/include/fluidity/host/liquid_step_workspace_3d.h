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
#include "fluidity/host/liquid_state_3d.h"
#include "fluidity/host/sparse_mac_grid_3d.h"
#include "fluidity/status.h"

namespace fluidity {

// The topology of a previous LiquidState3D.
struct LiquidTopologyHistory3D {
  UniformMacGeometry3D geometry{};
  std::uint64_t epoch = 0;
  std::span<const LiquidBlockState3D> blocks;
  std::span<const LiquidCellState3D> cells;
  std::span<const LiquidFaceState3D> faces;
};

// The topology of a committed state (its geometry, epoch and topology records).
LiquidTopologyHistory3D TopologyHistoryOf(const LiquidState3D& state);

struct LiquidPreparedTopologyView3D {
  SparseMacGridView3D grid;
  std::span<const LiquidParticleSortEntry3D> particle_order;
  // Particle records in sorted order; empty when the preparation took world positions only.
  std::span<const LiquidParticleState3D> particles;
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
  // The reference's overload (liquid_step_workspace_3d.h:96-100): particle records (their world
  // positions are sorted; every record is copied in sorted order into the prepared view) and an
  // optional previous state whose topology records, geometry and epoch are remapped.
  Status PrepareTrialTopology(std::span<const LiquidParticleState3D> particles,
                              std::span<const CellIndex3D> pressure_cells, const LiquidState3D* previous_state);
  Status HipPrepareTrialTopology(HipOperatorRuntime3D& runtime, std::span<const LiquidParticleState3D> particles,
                                 std::span<const CellIndex3D> pressure_cells, const LiquidState3D* previous_state,
                                 HipOperatorMeasurement& measurement);
  // kFailedPrecondition before the first successful preparation.
  StatusOr<LiquidPreparedTopologyView3D> prepared() const;
  // Copies the sorted particles and the prepared topology into the trial state's existing reservations
  // (liquid_step_workspace_3d.h:122-128): InvalidArgument for another geometry, ResourceExhausted when a
  // trial vector's capacity is too small, FailedPrecondition without a prepared view; the trial is
  // untouched on every failure.
  Status ApplyTrialTopology(LiquidState3D& trial) const;
  LiquidStateCapacities3D capacities() const { return capacities_; }

 private:
  struct Bank {
    std::vector<LiquidParticleSortEntry3D> particle_order;
    std::vector<LiquidParticleState3D> particles;
    std::vector<BlockIndex3D> block_keys;
    std::vector<LiquidBlockState3D> blocks;
    std::vector<LiquidCellState3D> cells;
    std::vector<LiquidFaceState3D> faces;
  };
  LiquidStepWorkspace3D(LiquidStateCapacities3D capacities, UniformMacGeometry3D geometry)
      : capacities_(capacities), geometry_(geometry) {}
  int Prepare(HipOperatorRuntime3D& runtime, std::span<const Vec3> positions, std::span<const CellIndex3D> cells,
              const LiquidTopologyHistory3D* previous, Bank& bank, void* measurement);
  int PrepareRecords(HipOperatorRuntime3D& runtime, std::span<const LiquidParticleState3D> particles,
                     std::span<const CellIndex3D> cells, const LiquidState3D* previous, Bank& bank,
                     void* measurement);
  LiquidStateCapacities3D capacities_;
  UniformMacGeometry3D geometry_;
  std::vector<Vec3> positions_;  // world positions of a particle-record preparation
  Bank banks_[2];
  int published_ = -1;
};

}  // namespace fluidity
