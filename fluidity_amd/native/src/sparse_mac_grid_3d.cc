// Host layout queries of the sparse MAC grid and the device-backed bin queries.
#include "fluidity/host/sparse_mac_grid_3d.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <tuple>

#include "internal.h"

namespace fluidity {
namespace {

constexpr std::int64_t kMinBlock = std::numeric_limits<std::int32_t>::min() / 4;
constexpr std::int64_t kMaxBlock = std::numeric_limits<std::int32_t>::max() / 4;

bool Finite(Vec3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

Status CheckGeometry(UniformMacGeometry3D g) {
  if (!Finite(g.world_origin)) return Status(StatusCode::kInvalidArgument);
  if (!(g.cell_width > 0.0) || !std::isnormal(g.cell_width)) return Status(StatusCode::kInvalidArgument);
  return OkStatus();
}

auto Key(BlockIndex3D b) { return std::tuple{b.z, b.y, b.x}; }

std::int64_t FloorDiv4(std::int64_t i) { return i >= 0 ? i / 4 : -1 - (-(i + 1)) / 4; }

bool InLattice(BlockIndex3D b) {
  for (const std::int64_t c : {std::int64_t{b.x}, std::int64_t{b.y}, std::int64_t{b.z}})
    if (c < kMinBlock || c > kMaxBlock) return false;
  return true;
}

StatusOr<Vec3> Reconstruct(UniformMacGeometry3D g, double x, double y, double z) {
  const Vec3 w{std::fma(g.cell_width, x, g.world_origin.x), std::fma(g.cell_width, y, g.world_origin.y),
               std::fma(g.cell_width, z, g.world_origin.z)};
  if (!Finite(w)) return Status(StatusCode::kOutOfRange);
  return w;
}

}  // namespace

StatusOr<SparseMacGridView3D> SparseMacGridView3D::Create(UniformMacGeometry3D geometry,
                                                          std::span<const BlockIndex3D> keys) {
  if (Status s = CheckGeometry(geometry); !s.ok()) return s;
  for (std::size_t i = 0; i < keys.size(); ++i) {
    if (!InLattice(keys[i])) return Status(StatusCode::kOutOfRange);
    if (i > 0 && !(Key(keys[i - 1]) < Key(keys[i]))) return Status(StatusCode::kInvalidArgument);
  }
  return SparseMacGridView3D(geometry, keys);
}

StatusOr<std::size_t> SparseMacGridView3D::FaceSlot(FaceIndex3D f) const {
  if (static_cast<unsigned>(f.axis) > 2) return Status(StatusCode::kInvalidArgument);
  const BlockIndex3D owner = MacOwnerBlock3D(f.lattice_index);
  const auto it = std::lower_bound(active_block_indices_.begin(), active_block_indices_.end(), owner,
                                   [](BlockIndex3D a, BlockIndex3D b) { return Key(a) < Key(b); });
  if (it == active_block_indices_.end() || Key(*it) != Key(owner)) return Status(StatusCode::kFailedPrecondition);
  const std::int64_t lx = f.lattice_index.x - 4LL * owner.x, ly = f.lattice_index.y - 4LL * owner.y,
                     lz = f.lattice_index.z - 4LL * owner.z;
  const auto block = static_cast<std::size_t>(it - active_block_indices_.begin());
  return block * 192 + static_cast<std::size_t>(f.axis) * 64 + static_cast<std::size_t>((lz * 4 + ly) * 4 + lx);
}

StatusOr<FaceIndex3D> SparseMacGridView3D::FaceIndex(std::size_t slot) const {
  if (slot >= face_count()) return Status(StatusCode::kOutOfRange);
  const BlockIndex3D b = active_block_indices_[slot / 192];
  const std::size_t within = slot % 192, local = within % 64;
  return FaceIndex3D{static_cast<FaceNormalAxis>(within / 64),
                     {b.x * 4 + static_cast<std::int32_t>(local % 4), b.y * 4 + static_cast<std::int32_t>((local / 4) % 4),
                      b.z * 4 + static_cast<std::int32_t>(local / 16)}};
}

StatusOr<Vec3> SparseMacGridView3D::FaceWorldAnchor(FaceIndex3D f) const {
  if (static_cast<unsigned>(f.axis) > 2) return Status(StatusCode::kInvalidArgument);
  const int a = static_cast<int>(f.axis);
  return Reconstruct(geometry_, f.lattice_index.x + (a == 0 ? 0.0 : 0.5), f.lattice_index.y + (a == 1 ? 0.0 : 0.5),
                     f.lattice_index.z + (a == 2 ? 0.0 : 0.5));
}

StatusOr<Vec3> SparseMacGridView3D::CellWorldCenter(CellIndex3D c) const {
  return Reconstruct(geometry_, c.x + 0.5, c.y + 0.5, c.z + 0.5);
}

BlockIndex3D MacOwnerBlock3D(FaceLatticeIndex3D l) {
  return {static_cast<std::int32_t>(FloorDiv4(l.x)), static_cast<std::int32_t>(FloorDiv4(l.y)),
          static_cast<std::int32_t>(FloorDiv4(l.z))};
}

StatusOr<SparseMacGrid3D> SparseMacGrid3D::Create(UniformMacGeometry3D geometry,
                                                  std::span<const BlockIndex3D> keys, std::size_t block_capacity) {
  if (Status s = CheckGeometry(geometry); !s.ok()) return s;
  if (keys.size() > block_capacity) return Status(StatusCode::kResourceExhausted);
  for (const BlockIndex3D& b : keys)
    if (!InLattice(b)) return Status(StatusCode::kOutOfRange);
  std::vector<BlockIndex3D> sorted(keys.begin(), keys.end());
  std::sort(sorted.begin(), sorted.end(), [](BlockIndex3D a, BlockIndex3D b) { return Key(a) < Key(b); });
  for (std::size_t i = 1; i < sorted.size(); ++i)
    if (Key(sorted[i - 1]) == Key(sorted[i])) return Status(StatusCode::kInvalidArgument);
  return SparseMacGrid3D(geometry, std::move(sorted));
}

Status SortParticlesByBin3D(UniformMacGeometry3D geometry, std::span<const Vec3> positions,
                            std::span<LiquidParticleSortEntry3D> out) {
  if (out.size() != positions.size()) return Status(StatusCode::kInvalidArgument);
  if (Status s = CheckGeometry(geometry); !s.ok()) return s;
  HipOperatorCapacities3D need;
  need.particles = internal::NextPow2(std::max<std::size_t>(positions.size(), 1));
  std::unique_lock<std::mutex> lock;
  for (int attempt = 0;; ++attempt) {
    auto rt = SharedHipRuntime3D::Acquire(need, lock);
    if (!rt.ok()) return rt.status();
    const fl_geometry g = internal::ToC(geometry);
    const int code = fl_bin_sort((*rt)->handle(), &g, reinterpret_cast<const fl_vec3*>(positions.data()),
                                 positions.size(), reinterpret_cast<fl_sort_entry*>(out.data()), nullptr);
    // The dense bin table bounds the bins' bounding box: grow it and retry while that is the limit.
    if (code == FL_RESOURCE_EXHAUSTED && need.lookup_blocks < (std::size_t{1} << 30) && attempt < 12) {
      need.lookup_blocks = (*rt)->capacities().lookup_blocks * 4;
      lock.unlock();
      continue;
    }
    return FromCode(code);
  }
}

StatusOr<ParticleBinIndex3D> ComputeParticleSortBin3D(UniformMacGeometry3D geometry, Vec3 position) {
  LiquidParticleSortEntry3D e{};
  const Status s = SortParticlesByBin3D(geometry, std::span<const Vec3>(&position, 1),
                                        std::span<LiquidParticleSortEntry3D>(&e, 1));
  if (!s.ok()) return s;
  return e.bin_index;
}

}  // namespace fluidity
