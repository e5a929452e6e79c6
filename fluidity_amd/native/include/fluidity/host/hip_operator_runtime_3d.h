// Device runtime behind the HIP operators (the role of CudaOperatorRuntime,
// native/cuda/operator_runtime.h:45-213): one context per GPU that owns every device buffer,
// allocated at Create; operator calls allocate nothing on the device.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>

#include "fluidity/status.h"

struct fl_ctx;

namespace fluidity {

// Device time of one operator call (CudaOperatorMeasurement, native/include/fluidity/host/cuda_measurement.h:5-10).
struct HipOperatorMeasurement {
  char architecture[28] = {};  // gcnArchName of the device, e.g. "gfx950:sramecc+:xnack-"
  double kernel_ms = 0.0;      // HIP-event interval of the call's kernels on the operator stream
  double copy_ms = 0.0;        // host <-> device staging of the span-based call
  std::uint32_t launches = 0;
};

struct HipOperatorCapacities3D {
  std::size_t particles = 1 << 16;
  std::size_t blocks = 1 << 14;
  std::size_t lookup_blocks = 1 << 20;  // bounding-box volume, in blocks, of the dense block and bin tables
  std::size_t pressure_cells = 1 << 16;
  std::size_t pressure_faces = 1 << 18;
};

// Single-threaded and non-reentrant, like the reference workspaces.
class HipOperatorRuntime3D {
 public:
  // kFailedPrecondition without a usable device, kResourceExhausted if the capacities do not fit.
  static StatusOr<std::unique_ptr<HipOperatorRuntime3D>> Create(HipOperatorCapacities3D capacities, int device = 0);
  ~HipOperatorRuntime3D();
  HipOperatorRuntime3D(const HipOperatorRuntime3D&) = delete;
  HipOperatorRuntime3D& operator=(const HipOperatorRuntime3D&) = delete;

  fl_ctx* handle() const { return ctx_; }
  const HipOperatorCapacities3D& capacities() const { return capacities_; }
  int device() const { return device_; }

 private:
  HipOperatorRuntime3D(fl_ctx* ctx, HipOperatorCapacities3D capacities, int device)
      : ctx_(ctx), capacities_(capacities), device_(device) {}
  fl_ctx* ctx_;
  HipOperatorCapacities3D capacities_;
  int device_;
};

// The process-wide runtime used by the operator functions that take no runtime argument. It is
// re-created with larger capacities (powers of two) when a call needs more; callers of those
// functions are serialised by the returned lock.
class SharedHipRuntime3D {
 public:
  // Returns a runtime with at least `need` of every capacity, or the creation status.
  static StatusOr<HipOperatorRuntime3D*> Acquire(const HipOperatorCapacities3D& need, std::unique_lock<std::mutex>& lock);
};

}  // namespace fluidity
