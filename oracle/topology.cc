// Oracle: particle sort and block activation (test infrastructure only).
// Restates native/src/liquid_step_workspace_3d.cc:199-283.
#include <algorithm>
#include <vector>

#include "grid_internal.h"

using namespace flo;

extern "C" {

// Stable order by signed (z,y,x) bin, then original ordinal (liquid_step_workspace_3d.cc:208-214, :245-268).
int flo_sort_particles(const flo_geometry* geometry, const flo_vec3* positions, size_t n, flo_sort_entry* out) {
  std::vector<flo_sort_entry> e(n);
  for (size_t i = 0; i < n; ++i) {
    flo_int3 bin;
    if (int s = ComputeBin(*geometry, positions[i], bin)) return s;
    e[i] = {bin, 0u, static_cast<uint64_t>(i)};
  }
  std::sort(e.begin(), e.end(), [](const flo_sort_entry& a, const flo_sort_entry& b) {
    if (!BlockEq(a.bin_index, b.bin_index)) return BlockLess(a.bin_index, b.bin_index);
    return a.original_ordinal < b.original_ordinal;
  });
  std::copy(e.begin(), e.end(), out);
  return FLO_OK;
}

// Owners of all 81 stencil faces of every particle plus the owners of the six faces of
// every pressure cell, sorted (z,y,x) and unique (liquid_step_workspace_3d.cc:208-244, :272-276).
int flo_activate_blocks(const flo_geometry* geometry, const flo_vec3* positions, size_t n,
                        const flo_int3* cells, size_t cell_count, flo_int3* out_blocks, size_t* io_count) {
  std::vector<flo_int3> keys;
  keys.reserve(27 * n + 7 * cell_count);
  for (size_t i = 0; i < n; ++i) {
    flo_int3 bin;
    if (int s = ComputeBin(*geometry, positions[i], bin)) return s;
    for (int axis = 0; axis < 3; ++axis) {
      std::array<GeometrySample, 27> st;
      if (int s = GeometryStencil(*geometry, axis, positions[i], positions[i], st)) return s;
      for (const auto& smp : st) keys.push_back(OwnerBlock(smp.lattice));
    }
  }
  for (size_t i = 0; i < cell_count; ++i) {
    const flo_int3 c = cells[i];
    if (c.x == kI32Max || c.y == kI32Max || c.z == kI32Max) return FLO_OUT_OF_RANGE;
    keys.push_back(OwnerBlock(c));
    keys.push_back(OwnerBlock({c.x + 1, c.y, c.z}));
    keys.push_back(OwnerBlock({c.x, c.y + 1, c.z}));
    keys.push_back(OwnerBlock({c.x, c.y, c.z + 1}));
  }
  std::sort(keys.begin(), keys.end(), BlockLess);
  keys.erase(std::unique(keys.begin(), keys.end(), BlockEq), keys.end());
  if (keys.size() > *io_count) return FLO_RESOURCE_EXHAUSTED;
  std::copy(keys.begin(), keys.end(), out_blocks);
  *io_count = keys.size();
  return FLO_OK;
}

}  // extern "C"
