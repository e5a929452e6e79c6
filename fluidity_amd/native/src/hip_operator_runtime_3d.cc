// Device runtime ownership and the process-wide shared runtime.
#include "fluidity/host/hip_operator_runtime_3d.h"

#include <algorithm>

#include "internal.h"

namespace fluidity {

StatusOr<std::unique_ptr<HipOperatorRuntime3D>> HipOperatorRuntime3D::Create(HipOperatorCapacities3D c, int device) {
  const fl_capacities caps{c.particles, c.blocks, c.lookup_blocks, c.pressure_cells, c.pressure_faces};
  fl_ctx* ctx = nullptr;
  const int code = fl_create(&caps, device, &ctx);
  if (code != FL_OK) return FromCode(code);
  return std::unique_ptr<HipOperatorRuntime3D>(new HipOperatorRuntime3D(ctx, c, device));
}

HipOperatorRuntime3D::~HipOperatorRuntime3D() { fl_destroy(ctx_); }

namespace {
std::mutex g_shared_mutex;
std::unique_ptr<HipOperatorRuntime3D> g_shared;

bool Covers(const HipOperatorCapacities3D& have, const HipOperatorCapacities3D& need) {
  return have.particles >= need.particles && have.blocks >= need.blocks && have.lookup_blocks >= need.lookup_blocks &&
         have.pressure_cells >= need.pressure_cells && have.pressure_faces >= need.pressure_faces;
}
}  // namespace

StatusOr<HipOperatorRuntime3D*> SharedHipRuntime3D::Acquire(const HipOperatorCapacities3D& need,
                                                            std::unique_lock<std::mutex>& lock) {
  lock = std::unique_lock<std::mutex>(g_shared_mutex);
  if (g_shared && Covers(g_shared->capacities(), need)) return g_shared.get();
  HipOperatorCapacities3D want = g_shared ? g_shared->capacities() : HipOperatorCapacities3D{};
  want.particles = internal::NextPow2(std::max(want.particles, need.particles));
  want.blocks = internal::NextPow2(std::max(want.blocks, need.blocks));
  want.lookup_blocks = std::min<std::size_t>(internal::NextPow2(std::max(want.lookup_blocks, need.lookup_blocks)),
                                             0x7fffffff);
  want.pressure_cells = internal::NextPow2(std::max(want.pressure_cells, need.pressure_cells));
  want.pressure_faces = internal::NextPow2(std::max(want.pressure_faces, need.pressure_faces));
  g_shared.reset();  // release the old device arena before allocating the larger one
  auto created = HipOperatorRuntime3D::Create(want, 0);
  if (!created.ok()) return created.status();
  g_shared = std::move(created.value());
  return g_shared.get();
}

}  // namespace fluidity
