// Oracle: grid queries (test infrastructure only). Restates native/src/sparse_mac_grid_3d.cc.
#include <algorithm>
#include <vector>

#include "grid_internal.h"

using namespace flo;

extern "C" {

// sparse_mac_grid_3d.cc:438-449
int flo_compute_bin(const flo_geometry* geometry, flo_vec3 position, flo_int3* out_bin) {
  flo_int3 bin;
  if (int s = ComputeBin(*geometry, position, bin)) return s;
  *out_bin = bin;
  return FLO_OK;
}

// SparseMacGrid3D::Create (sparse_mac_grid_3d.cc:344-374) without the capacity argument.
int flo_grid_sort_blocks(const flo_geometry* geometry, const flo_int3* blocks, size_t count, flo_int3* out_sorted) {
  if (int s = ValidateGeometry(*geometry)) return s;
  for (size_t i = 0; i < count; ++i)
    for (int a = 0; a < 3; ++a) {
      const int64_t c = Get(blocks[i], a);
      if (c < kMinBlock || c > kMaxBlock) return FLO_OUT_OF_RANGE;
    }
  std::vector<flo_int3> sorted(blocks, blocks + count);
  std::sort(sorted.begin(), sorted.end(), BlockLess);
  for (size_t i = 1; i < sorted.size(); ++i)
    if (BlockEq(sorted[i - 1], sorted[i])) return FLO_INVALID_ARGUMENT;
  std::copy(sorted.begin(), sorted.end(), out_sorted);
  return FLO_OK;
}

int flo_face_slot(const flo_int3* sorted_blocks, size_t count, flo_face_index face, uint64_t* out_slot) {
  uint64_t slot;
  if (int s = FaceSlot(sorted_blocks, count, face.axis, face.lattice, slot)) return s;
  *out_slot = slot;
  return FLO_OK;
}

// sparse_mac_grid_3d.cc:587-603
int flo_slot_face(const flo_int3* sorted_blocks, size_t count, uint64_t slot, flo_face_index* out_face) {
  if (slot >= count * kFacesPerBlock) return FLO_OUT_OF_RANGE;
  const flo_int3 b = sorted_blocks[slot / kFacesPerBlock];
  const uint64_t within = slot % kFacesPerBlock;
  const uint64_t local = within % 64;
  flo_face_index f{};
  f.axis = static_cast<uint8_t>(within / 64);
  f.lattice = {b.x * 4 + static_cast<int32_t>(local % 4), b.y * 4 + static_cast<int32_t>((local / 4) % 4),
               b.z * 4 + static_cast<int32_t>(local / 16)};
  *out_face = f;
  return FLO_OK;
}

// sparse_mac_grid_3d.cc:411-425 with ReconstructWorldPosition (:70-81)
int flo_face_world_anchor(const flo_geometry* geometry, flo_face_index face, flo_vec3* out) {
  if (face.axis > 2) return FLO_INVALID_ARGUMENT;
  flo_vec3 r;
  for (int a = 0; a < 3; ++a) {
    const double n = Get(face.lattice, a) + (a == face.axis ? 0.0 : 0.5);
    Set(r, a, std::fma(geometry->cell_width, n, Get(geometry->world_origin, a)));
  }
  if (!Finite3(r)) return FLO_OUT_OF_RANGE;
  *out = r;
  return FLO_OK;
}

int flo_owned_stencil(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t count, int axis,
                      flo_vec3 sorted_position, flo_vec3 trial_position, flo_stencil_sample* out27) {
  std::array<flo_stencil_sample, 27> st;
  if (int s = OwnedStencil(*geometry, sorted_blocks, count, axis, sorted_position, trial_position, st)) return s;
  std::copy(st.begin(), st.end(), out27);
  return FLO_OK;
}

}  // extern "C"
