// Shared helpers of the host operator layer: record conversions to the C-ABI and capacity sizing.
#pragma once

#include <algorithm>
#include <cstddef>
#include <span>
#include <type_traits>

#include "fluidity/base/mac_types_3d.h"
#include "fluidity/host/hip_operator_runtime_3d.h"
#include "fluidity/host/sparse_mac_grid_3d.h"
#include "fluidity_hip.h"

namespace fluidity::internal {

// The layer's records are layout-identical to the C-ABI records.
static_assert(sizeof(Vec3) == sizeof(fl_vec3) && sizeof(Mat3) == sizeof(fl_mat3));
static_assert(sizeof(BlockIndex3D) == sizeof(fl_int3) && sizeof(CellIndex3D) == sizeof(fl_int3));
static_assert(sizeof(MacTransferParticle3D) == sizeof(fl_particle) && sizeof(MacTransferParticle3D) == 152);
static_assert(sizeof(MacFaceDepositQuanta3D) == sizeof(fl_face_quanta));
static_assert(sizeof(MacFaceDeposit3D) == sizeof(fl_face_deposit));
static_assert(sizeof(MacParticleGatherResult3D) == sizeof(fl_gather_result));
static_assert(sizeof(LiquidParticleSortEntry3D) == sizeof(fl_sort_entry) &&
              offsetof(LiquidParticleSortEntry3D, original_ordinal) == offsetof(fl_sort_entry, original_ordinal));
static_assert(std::is_standard_layout_v<MacTransferParticle3D> && std::is_standard_layout_v<MacParticleGatherResult3D>);

inline fl_geometry ToC(UniformMacGeometry3D g) {
  return fl_geometry{{g.world_origin.x, g.world_origin.y, g.world_origin.z}, g.cell_width};
}
inline const fl_int3* ToC(std::span<const BlockIndex3D> blocks) {
  return reinterpret_cast<const fl_int3*>(blocks.data());
}

inline std::size_t NextPow2(std::size_t v) {
  std::size_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Dense-table extent for a grid: the (z,y,x) bounding box of the active blocks, one block of margin.
inline std::size_t LookupNeed(std::span<const BlockIndex3D> blocks) {
  if (blocks.empty()) return 64;
  long long lo[3] = {blocks[0].x, blocks[0].y, blocks[0].z}, hi[3] = {lo[0], lo[1], lo[2]};
  for (const auto& b : blocks) {
    const long long v[3] = {b.x, b.y, b.z};
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], v[a]);
      hi[a] = std::max(hi[a], v[a]);
    }
  }
  long double vol = 1;
  for (int a = 0; a < 3; ++a) vol *= static_cast<long double>(hi[a] - lo[a] + 3);
  return vol > 0x7fffffff ? std::size_t{0x7fffffff} : static_cast<std::size_t>(vol);
}

}  // namespace fluidity::internal
