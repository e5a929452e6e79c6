// Device helpers over records sorted by signed lattice key (z, then y, then x): ordering, binary
// lookup, shifted cell indices and face slots of a sorted block list, and the ordered error word of
// kernels that restate a sequential loop.
#pragma once
#include "context.h"

namespace fl {

__device__ inline bool key_less(fl_int3 a, fl_int3 b) {
  if (a.z != b.z) return a.z < b.z;
  if (a.y != b.y) return a.y < b.y;
  return a.x < b.x;
}
__device__ inline bool key_eq(fl_int3 a, fl_int3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// Index of the record with key k among n sorted records, or n.
template <typename R, typename Key>
__device__ inline uint64_t find_sorted(const R* r, uint64_t n, fl_int3 k, Key key) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (key_less(key(r[mid]), k)) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && key_eq(key(r[lo]), k)) ? lo : n;
}

// OffsetCellIndex (native/src/liquid_boundary_geometry_3d.cc:47-60): false outside the int32 lattice
// (OutOfRange).
__device__ inline bool offset(fl_int3 c, int axis, int amount, fl_int3& out) {
  int64_t w[3] = {c.x, c.y, c.z};
  w[axis] += amount;
  if (w[axis] < kI32Min || w[axis] > kI32Max) return false;
  out = fl_int3{int32_t(w[0]), int32_t(w[1]), int32_t(w[2])};
  return true;
}

// SparseMacGridView3D::FaceSlot over sorted blocks: InvalidArgument for an axis above 2,
// FailedPrecondition when the owner block is inactive.
__device__ inline int sorted_face_slot(const fl_int3* blocks, uint64_t nb, int axis, fl_int3 l, uint64_t& slot) {
  if (axis > 2) return kInvalidArgument;
  const fl_int3 o{int32_t(floor_div4(l.x)), int32_t(floor_div4(l.y)), int32_t(floor_div4(l.z))};
  const uint64_t b = find_sorted(blocks, nb, o, [](const fl_int3& x) { return x; });
  if (b == nb) return kFailedPrecondition;
  const uint64_t lx = uint64_t(int64_t(l.x) - int64_t(o.x) * 4), ly = uint64_t(int64_t(l.y) - int64_t(o.y) * 4),
                 lz = uint64_t(int64_t(l.z) - int64_t(o.z) * 4);
  slot = b * 192 + uint64_t(axis) * 64 + lz * 16 + ly * 4 + lx;
  return kOk;
}

// The first failure of a restated sequential loop: loops are phases, records are ordinals, and the
// smallest (phase, ordinal) wins.
constexpr int kPhaseShift = 40;
__device__ inline void fail(uint64_t* err, uint64_t phase, uint64_t ordinal, int code) {
  atomicMin(reinterpret_cast<unsigned long long*>(err),
            static_cast<unsigned long long>(error_word((phase << kPhaseShift) | ordinal, code)));
}

}  // namespace fl
