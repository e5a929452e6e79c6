// Canonical sparse staggered MAC layout (SparseMacGrid3D, native/include/fluidity/host/sparse_mac_grid_3d.h:19-171).
//
// Active 4x4x4 blocks sorted (z,y,x); block b owns 192 face slots: slot = b * 192 + axis * 64 +
// (z * 16 + y * 4 + x) for the low face of local cell (x,y,z) (sparse_mac_grid_3d.cc:388-409).
// These are host layout queries; every transfer and projection runs on the device.
#pragma once

#include <cstddef>
#include <span>
#include <vector>

#include "fluidity/base/mac_types_3d.h"
#include "fluidity/status.h"

namespace fluidity {

// Borrowed canonical layout: the keys must already be strictly increasing in (z,y,x) order
// (SparseMacGridView3D::Create, sparse_mac_grid_3d.cc:560-586).
class SparseMacGridView3D {
 public:
  SparseMacGridView3D() = default;
  static StatusOr<SparseMacGridView3D> Create(UniformMacGeometry3D geometry,
                                              std::span<const BlockIndex3D> active_block_indices);
  UniformMacGeometry3D geometry() const { return geometry_; }
  std::size_t face_count() const { return 192 * active_block_indices_.size(); }
  std::span<const BlockIndex3D> active_block_indices() const { return active_block_indices_; }
  // kInvalidArgument for an invalid axis, kFailedPrecondition if the owner block is inactive.
  StatusOr<std::size_t> FaceSlot(FaceIndex3D face_index) const;
  // kOutOfRange when face_slot >= face_count().
  StatusOr<FaceIndex3D> FaceIndex(std::size_t face_slot) const;
  // Face anchor origin + h * (lattice + 1/2 on the transverse axes), one fused multiply-add per
  // component; kOutOfRange for a non-finite result.
  StatusOr<Vec3> FaceWorldAnchor(FaceIndex3D face_index) const;
  StatusOr<Vec3> CellWorldCenter(CellIndex3D cell_index) const;

 private:
  friend class SparseMacGrid3D;
  SparseMacGridView3D(UniformMacGeometry3D geometry, std::span<const BlockIndex3D> keys)
      : geometry_(geometry), active_block_indices_(keys) {}
  UniformMacGeometry3D geometry_{{0.0, 0.0, 0.0}, 1.0};
  std::span<const BlockIndex3D> active_block_indices_;
};

// Owner block of a face lattice index: floor division by 4 on each axis.
BlockIndex3D MacOwnerBlock3D(FaceLatticeIndex3D lattice_index);

// Owning layout (SparseMacGrid3D::Create, sparse_mac_grid_3d.cc:344-374): copies and sorts the keys.
// kInvalidArgument for invalid geometry or duplicate keys, kOutOfRange for keys whose faces leave
// the int32 lattice, kResourceExhausted if the keys exceed block_capacity.
class SparseMacGrid3D {
 public:
  static StatusOr<SparseMacGrid3D> Create(UniformMacGeometry3D geometry,
                                          std::span<const BlockIndex3D> active_block_indices,
                                          std::size_t block_capacity);
  SparseMacGridView3D view() const { return SparseMacGridView3D(geometry_, active_block_indices_); }
  UniformMacGeometry3D geometry() const { return geometry_; }
  std::size_t face_count() const { return 192 * active_block_indices_.size(); }
  std::span<const BlockIndex3D> active_block_indices() const { return active_block_indices_; }
  StatusOr<std::size_t> FaceSlot(FaceIndex3D face_index) const { return view().FaceSlot(face_index); }
  StatusOr<FaceIndex3D> FaceIndex(std::size_t face_slot) const { return view().FaceIndex(face_slot); }
  StatusOr<Vec3> FaceWorldAnchor(FaceIndex3D face_index) const { return view().FaceWorldAnchor(face_index); }
  StatusOr<Vec3> CellWorldCenter(CellIndex3D cell_index) const { return view().CellWorldCenter(cell_index); }

 private:
  SparseMacGrid3D(UniformMacGeometry3D geometry, std::vector<BlockIndex3D> keys)
      : geometry_(geometry), active_block_indices_(std::move(keys)) {}
  UniformMacGeometry3D geometry_;
  std::vector<BlockIndex3D> active_block_indices_;
};

// Sort bin floor((X - 1.5) / 4) of one sorted position, computed on the device by the same kernel as
// the batch sort (ComputeParticleSortBin3D, sparse_mac_grid_3d.cc:438-449).
StatusOr<ParticleBinIndex3D> ComputeParticleSortBin3D(UniformMacGeometry3D geometry, Vec3 sorted_world_position);

// Stable (z,y,x)-bin / ordinal order of a batch of sorted positions, on the device
// (LiquidStepWorkspace3D::PrepareTopology sort, liquid_step_workspace_3d.cc:245-271). out has
// positions.size() entries in sorted order.
Status SortParticlesByBin3D(UniformMacGeometry3D geometry, std::span<const Vec3> sorted_world_positions,
                            std::span<LiquidParticleSortEntry3D> out);

}  // namespace fluidity
