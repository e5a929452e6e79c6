// Oracle: particle sort, block activation and topology preparation with history remap (test
// infrastructure only). Restates native/src/liquid_step_workspace_3d.cc:171-341 and the topology
// checks of ValidateLiquidState3D (native/src/liquid_state_3d.cc:559-604).
#include <algorithm>
#include <cmath>
#include <cstring>
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

// Topology records of the previous state, in the order and with the codes of ValidateLiquidState3D
// (liquid_state_3d.cc:559-604): block coordinates and (z,y,x) order, cells in (z,y,x) order with an
// active owner, an occupancy and a finite pressure, faces with a valid axis, validity and velocity,
// an active owner, the owned slot of their index and increasing slots.
static int ValidateHistory(const flo_topology_history& h) {
  const auto owner_ordinal = [&](flo_int3 lattice) -> uint64_t {
    const flo_int3 key = OwnerBlock(lattice);
    const flo_block_state* b = h.blocks;
    const flo_block_state* e = h.blocks + h.block_count;
    const flo_block_state* f =
        std::lower_bound(b, e, key, [](const flo_block_state& x, const flo_int3& k) { return BlockLess(x.block_index, k); });
    return (f != e && BlockEq(f->block_index, key)) ? uint64_t(f - b) : h.block_count;
  };
  if (h.block_count > UINT64_MAX / 192) return FLO_OUT_OF_RANGE;
  for (uint64_t i = 0; i < h.block_count; ++i) {
    const flo_int3 k = h.blocks[i].block_index;
    for (int32_t c : {k.x, k.y, k.z})
      if (c < kI32Min / 4 || c > (kI32Max - 3) / 4) return FLO_OUT_OF_RANGE;
    if (i > 0 && !BlockLess(h.blocks[i - 1].block_index, k)) return FLO_INVALID_ARGUMENT;
  }
  for (uint64_t i = 0; i < h.cell_count; ++i) {
    const flo_cell_state& c = h.cells[i];
    if (i > 0 && !BlockLess(h.cells[i - 1].cell_index, c.cell_index)) return FLO_INVALID_ARGUMENT;
    if (owner_ordinal(c.cell_index) == h.block_count) return FLO_INVALID_ARGUMENT;
    if (c.occupancy > 2 || !std::isfinite(c.previous_pressure)) return FLO_INVALID_ARGUMENT;
  }
  for (uint64_t i = 0; i < h.face_count; ++i) {
    const flo_face_state& f = h.faces[i];
    if (f.face_index.axis > 2 || f.validity > 3 || !std::isfinite(f.carrier_velocity)) return FLO_INVALID_ARGUMENT;
    const flo_int3 l = f.face_index.lattice;
    const uint64_t o = owner_ordinal(l);
    if (o == h.block_count) return FLO_INVALID_ARGUMENT;
    const flo_int3 b = h.blocks[o].block_index;
    const uint64_t lx = uint64_t(int64_t(l.x) - int64_t(b.x) * 4), ly = uint64_t(int64_t(l.y) - int64_t(b.y) * 4),
                   lz = uint64_t(int64_t(l.z) - int64_t(b.z) * 4);
    const uint64_t expected = o * 192 + uint64_t(f.face_index.axis) * 64 + lz * 16 + ly * 4 + lx;
    if (f.face_slot != expected || (i > 0 && h.faces[i - 1].face_slot >= expected)) return FLO_INVALID_ARGUMENT;
  }
  return FLO_OK;
}

int flo_prepare_topology(const flo_geometry* geometry, const flo_topology_capacities* caps, const flo_vec3* positions,
                         size_t n, const flo_int3* pressure_cells, size_t cell_count,
                         const flo_topology_history* previous, flo_sort_entry* out_order, flo_block_state* out_blocks,
                         uint64_t* out_block_count, flo_cell_state* out_cells, uint64_t* out_cell_count,
                         flo_face_state* out_faces, uint64_t* out_face_count) {
  if (int s = ValidateGeometry(*geometry)) return s;
  if (n > caps->particles || cell_count > caps->cells) return FLO_RESOURCE_EXHAUSTED;
  uint64_t next_generation = 1;
  if (previous) {
    const flo_geometry& pg = previous->geometry;
    if (!(pg.world_origin.x == geometry->world_origin.x && pg.world_origin.y == geometry->world_origin.y &&
          pg.world_origin.z == geometry->world_origin.z && pg.cell_width == geometry->cell_width))
      return FLO_INVALID_ARGUMENT;
    if (int s = ValidateHistory(*previous)) return s;
    if (previous->epoch == UINT64_MAX) return FLO_OUT_OF_RANGE;
    next_generation = previous->epoch + 1;
  }
  // Particle errors in particle order (bin, then the three axis stencils), then cell range errors.
  for (size_t i = 0; i < n; ++i) {
    flo_int3 bin;
    if (int s = ComputeBin(*geometry, positions[i], bin)) return s;
    for (int axis = 0; axis < 3; ++axis) {
      std::array<GeometrySample, 27> st;
      if (int s = GeometryStencil(*geometry, axis, positions[i], positions[i], st)) return s;
    }
  }
  for (size_t i = 0; i < cell_count; ++i) {
    const flo_int3 c = pressure_cells[i];
    if (c.x == kI32Max || c.y == kI32Max || c.z == kI32Max) return FLO_OUT_OF_RANGE;
  }
  std::vector<flo_sort_entry> order(n);
  if (n) {
    if (int s = flo_sort_particles(geometry, positions, n, order.data())) return s;
  }
  std::vector<flo_int3> blocks(27 * 3 * n + 4 * cell_count + 1);
  size_t nb = blocks.size();
  if (int s = flo_activate_blocks(geometry, positions, n, pressure_cells, cell_count, blocks.data(), &nb)) return s;
  if (nb > caps->blocks || nb > caps->faces / 192) return FLO_RESOURCE_EXHAUSTED;
  // Cells: canonical (z,y,x) order without duplicates.
  std::vector<flo_int3> cells(pressure_cells, pressure_cells + cell_count);
  std::sort(cells.begin(), cells.end(), BlockLess);
  cells.erase(std::unique(cells.begin(), cells.end(), BlockEq), cells.end());
  // Blocks: the previous record with its inactive age cleared, or a new one of this generation.
  for (size_t i = 0; i < nb; ++i) {
    flo_block_state r;
    std::memset(&r, 0, sizeof r);
    r.block_index = blocks[i];
    r.generation = next_generation;
    if (previous) {
      const flo_block_state* b = previous->blocks;
      const flo_block_state* e = b + previous->block_count;
      const flo_block_state* f = std::lower_bound(
          b, e, blocks[i], [](const flo_block_state& x, const flo_int3& k) { return BlockLess(x.block_index, k); });
      if (f != e && BlockEq(f->block_index, blocks[i])) {
        r = *f;
        r.inactive_age = 0;
      }
    }
    out_blocks[i] = r;
  }
  for (size_t i = 0; i < cells.size(); ++i) {
    flo_cell_state r;
    std::memset(&r, 0, sizeof r);
    r.cell_index = cells[i];
    r.occupancy = 1;
    if (previous) {
      const flo_cell_state* b = previous->cells;
      const flo_cell_state* e = b + previous->cell_count;
      const flo_cell_state* f = std::lower_bound(
          b, e, cells[i], [](const flo_cell_state& x, const flo_int3& k) { return BlockLess(x.cell_index, k); });
      if (f != e && BlockEq(f->cell_index, cells[i])) r = *f;
    }
    out_cells[i] = r;
  }
  // Faces: one record per owned slot; the previous record of the same face (owner key, axis, lattice
  // key: owned-slot order in both states) with the new slot, or a fresh record.
  const auto face_key_less = [](const flo_face_index& a, const flo_face_index& b) {
    const flo_int3 oa = OwnerBlock(a.lattice), ob = OwnerBlock(b.lattice);
    if (!BlockEq(oa, ob)) return BlockLess(oa, ob);
    if (a.axis != b.axis) return a.axis < b.axis;
    return BlockLess(a.lattice, b.lattice);
  };
  for (uint64_t slot = 0; slot < 192 * nb; ++slot) {
    flo_face_index fi;
    if (int s = flo_slot_face(blocks.data(), nb, slot, &fi)) return s;
    flo_face_state r;
    std::memset(&r, 0, sizeof r);
    r.face_index.axis = fi.axis;
    r.face_index.lattice = fi.lattice;
    r.face_slot = slot;
    r.structurally_eligible = 1;
    if (previous) {
      const flo_face_state* b = previous->faces;
      const flo_face_state* e = b + previous->face_count;
      const flo_face_state* f = std::lower_bound(
          b, e, fi, [&](const flo_face_state& x, const flo_face_index& k) { return face_key_less(x.face_index, k); });
      if (f != e && !face_key_less(fi, f->face_index) && !face_key_less(f->face_index, fi)) {
        r = *f;
        r.face_slot = slot;
      }
    }
    out_faces[slot] = r;
  }
  std::copy(order.begin(), order.end(), out_order);
  *out_block_count = nb;
  *out_cell_count = cells.size();
  *out_face_count = 192 * nb;
  return FLO_OK;
}

}  // extern "C"
