// Oracle: the particle-union liquid field of one substep (test infrastructure only).
//
// PrepareField's candidate cells and sample lattice for LiquidParticleSphereUnion3D
// (native/src/liquid_substep_3d_staging.cc:759-916: add_sphere at :785-824, the particle loop at
// :833-838, the 27 centres and 8 corners per candidate at :884-916) and the distance of every sample
// (FieldDistance, :210-229, at the world position of :937-957; SampleWorldBatch,
// native/src/liquid_substep_3d_cpu.cc:77-120). Each distance scans every particle, as the reference
// does; std::hypot is libstdc++'s three-argument hypot.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <set>
#include <tuple>
#include <vector>

#include "oracle.h"

namespace {
using Key = std::tuple<int32_t, int32_t, int32_t>;  // (z, y, x)
}

extern "C" int flo_particle_field(const flo_geometry* geometry, const flo_vec3* positions, size_t n, double radius,
                                  uint64_t candidate_capacity, flo_int3* out_cells, uint64_t* cell_count,
                                  flo_center_distance* out_centres, uint64_t centre_capacity, uint64_t* centre_count,
                                  flo_center_distance* out_corners, uint64_t corner_capacity,
                                  uint64_t* corner_count) {
  if (!geometry || !cell_count || !centre_count || !corner_count || (n && !positions)) return FLO_INVALID_ARGUMENT;
  // SampleWorldBatch's field and particle checks (liquid_substep_3d_cpu.cc:96-110).
  if (!std::isfinite(radius) || radius <= 0) return FLO_INVALID_ARGUMENT;
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(positions[i].x) || !std::isfinite(positions[i].y) || !std::isfinite(positions[i].z))
      return FLO_INVALID_ARGUMENT;
  const double h = geometry->cell_width;
  const double origins[3] = {geometry->world_origin.x, geometry->world_origin.y, geometry->world_origin.z};
  std::set<Key> candidates;
  for (size_t p = 0; p < n; ++p) {
    const double centres[3] = {positions[p].x, positions[p].y, positions[p].z};
    int32_t lower[3], upper[3];
    uint64_t box_count = 1;
    for (int axis = 0; axis < 3; ++axis) {
      const double low = std::floor((centres[axis] - radius - origins[axis]) / h);
      const double high = std::floor((centres[axis] + radius - origins[axis]) / h);
      if (!std::isfinite(low) || !std::isfinite(high) || low < static_cast<double>(INT32_MIN) + 1 ||
          high > static_cast<double>(INT32_MAX) - 1)
        return FLO_OUT_OF_RANGE;
      lower[axis] = static_cast<int32_t>(low);
      upper[axis] = static_cast<int32_t>(high);
      const uint64_t extent = static_cast<uint64_t>(static_cast<int64_t>(upper[axis]) - lower[axis] + 1);
      if (extent > candidate_capacity / box_count) return FLO_RESOURCE_EXHAUSTED;
      box_count *= extent;
    }
    for (int64_t z = lower[2]; z <= upper[2]; ++z)
      for (int64_t y = lower[1]; y <= upper[1]; ++y)
        for (int64_t x = lower[0]; x <= upper[0]; ++x) {
          candidates.insert(Key{int32_t(z), int32_t(y), int32_t(x)});
          if (candidates.size() > candidate_capacity) return FLO_RESOURCE_EXHAUSTED;
        }
  }
  std::set<Key> centres, corners;
  for (const Key& c : candidates) {
    const auto [cz, cy, cx] = c;
    for (int z = -1; z <= 1; ++z)
      for (int y = -1; y <= 1; ++y)
        for (int x = -1; x <= 1; ++x) centres.insert(Key{cz + z, cy + y, cx + x});
    for (int z = 0; z <= 1; ++z)
      for (int y = 0; y <= 1; ++y)
        for (int x = 0; x <= 1; ++x) corners.insert(Key{cz + z, cy + y, cx + x});
  }
  *cell_count = candidates.size();
  *centre_count = centres.size();
  *corner_count = corners.size();
  if (!out_cells || !out_centres || !out_corners) return FLO_OK;  // sizes only
  if (centres.size() > centre_capacity || corners.size() > corner_capacity) return FLO_RESOURCE_EXHAUSTED;
  const auto distance = [&](double px, double py, double pz) {
    double result = std::numeric_limits<double>::infinity();
    for (size_t p = 0; p < n; ++p)
      result = std::min(result, std::hypot(px - positions[p].x, py - positions[p].y, pz - positions[p].z) - radius);
    return result;
  };
  const auto sample = [&](const Key& k, double offset, flo_center_distance& out) -> int {
    const auto [z, y, x] = k;
    const double lx = x + offset, ly = y + offset, lz = z + offset;
    const double wx = origins[0] + lx * h, wy = origins[1] + ly * h, wz = origins[2] + lz * h;
    const double d = distance(wx, wy, wz);
    if (!std::isfinite(d)) return FLO_OUT_OF_RANGE;
    out.cell_index = flo_int3{x, y, z};
    out.pad = 0;
    out.signed_distance = d;
    return FLO_OK;
  };
  size_t i = 0;
  for (const Key& c : candidates) {
    const auto [z, y, x] = c;
    out_cells[i++] = flo_int3{x, y, z};
  }
  i = 0;
  for (const Key& k : centres)
    if (int s = sample(k, 0.5, out_centres[i++])) return s;
  i = 0;
  for (const Key& k : corners)
    if (int s = sample(k, 0.0, out_corners[i++])) return s;
  return FLO_OK;
}
