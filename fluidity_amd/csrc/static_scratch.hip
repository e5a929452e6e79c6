// fl_normalize_forces_static: fl_normalize_forces for callers that cannot allocate device memory.
//
// The call has no context. Its device memory is one file-scope __device__ array, `scratch`, of
// kStaticScratchBytes on each GPU the library runs on (the array is part of the code object, so it
// exists from module load until the process ends). Host code never passes the array's name to a
// kernel or a copy: the first call on a device, with that device current, resolves the array's device
// address with hipGetSymbolAddress and keeps it in a per-device table; every later use goes through
// that address. The call allocates and frees nothing, and runs the same kernel as fl_normalize_forces.
//
// Scratch layout, for up to kStaticMaxFaces owned face slots (F below), each region 256-byte aligned:
//   error words (2 x u64) | quanta staging (24 F) | mass, momentum, stress quanta columns (3 x 8 F) |
//   face mass (8 F) | velocity (8 F) | predicted velocity (8 F) | validity (F)
// Every call on a GPU shares that GPU's array, which keeps the last call's contents: the error words are
// preset by a kernel at the start of each call, and every other region is fully written before it is
// read. Calls from different host threads on one GPU are serialized by a per-device mutex; calls that
// overlap on one GPU by other means (another library instance in the same process) would race.
#include <cmath>
#include <mutex>

#include "context.h"

namespace {

using namespace fl;

constexpr uint64_t kStaticMaxFaces = FL_STATIC_MAX_FACES;
constexpr uint64_t kAlign = 256;

constexpr uint64_t align_up(uint64_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

constexpr uint64_t kErrOffset = 0;
constexpr uint64_t kQuantaOffset = align_up(kErrOffset + 2 * sizeof(uint64_t));
constexpr uint64_t kColumnsOffset = align_up(kQuantaOffset + kStaticMaxFaces * sizeof(fl_face_quanta));
constexpr uint64_t kMassOffset = align_up(kColumnsOffset + 3 * kStaticMaxFaces * sizeof(int64_t));
constexpr uint64_t kVelocityOffset = align_up(kMassOffset + kStaticMaxFaces * sizeof(double));
constexpr uint64_t kPredictedOffset = align_up(kVelocityOffset + kStaticMaxFaces * sizeof(double));
constexpr uint64_t kValidityOffset = align_up(kPredictedOffset + kStaticMaxFaces * sizeof(double));
constexpr uint64_t kStaticScratchBytes = align_up(kValidityOffset + kStaticMaxFaces);
static_assert(kStaticScratchBytes == FL_STATIC_SCRATCH_BYTES, "scratch size differs from the header's");

__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_setup_mutex;
unsigned char* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device, once resolved
std::mutex g_call_mutex[kMaxDevices];

// Resolves the device address of `scratch` on `device` (made current by the caller) once per device.
int scratch_address(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lock(g_setup_mutex);
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return FL_INTERNAL;
    g_scratch[device] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch[device];
  return FL_OK;
}

__global__ void k_preset_words(uint64_t* w) {
  if (threadIdx.x < 2) w[threadIdx.x] = ~uint64_t(0);
}

__global__ void k_quanta_columns(const fl_face_quanta* in, uint64_t n, int64_t* qm, int64_t* qp, int64_t* qs) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  qm[i] = in[i].mass_quanta_sum;
  qp[i] = in[i].pure_momentum_quanta_sum;
  qs[i] = in[i].stress_force_quanta_sum;
}

bool power_of_two(double q) {
  int e;
  return std::isnormal(q) && std::frexp(q, &e) == 0.5;
}

}  // namespace

extern "C" int fl_normalize_forces_static(int device, const fl_face_quanta* quanta, size_t face_count,
                                          fl_quantization quantization, double minimum_face_mass,
                                          fl_vec3 acceleration, double dt, double* out_mass, double* out_velocity,
                                          uint8_t* out_validity, double* out_predicted) {
  // fl_normalize_forces's checks, in its order, with the static capacity in place of the context's.
  if (face_count > kStaticMaxFaces) return FL_RESOURCE_EXHAUSTED;
  if (!power_of_two(quantization.mass_per_quantum) || !power_of_two(quantization.pure_momentum_per_quantum))
    return FL_INVALID_ARGUMENT;
  if (!std::isfinite(minimum_face_mass) || minimum_face_mass < 0.0) return FL_INVALID_ARGUMENT;
  if (face_count % 192 != 0) return FL_INVALID_ARGUMENT;
  if (face_count && !quanta) return FL_INVALID_ARGUMENT;
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || device < 0 || device >= devices || device >= kMaxDevices)
    return FL_FAILED_PRECONDITION;
  if (hipSetDevice(device) != hipSuccess) return FL_FAILED_PRECONDITION;
  unsigned char* base = nullptr;
  FL_TRY(scratch_address(device, &base));
  std::lock_guard<std::mutex> lock(g_call_mutex[device]);
  auto* err = reinterpret_cast<uint64_t*>(base + kErrOffset);
  auto* staged = reinterpret_cast<fl_face_quanta*>(base + kQuantaOffset);
  auto* qm = reinterpret_cast<int64_t*>(base + kColumnsOffset);
  int64_t* qp = qm + kStaticMaxFaces;
  int64_t* qs = qp + kStaticMaxFaces;
  auto* mass = reinterpret_cast<double*>(base + kMassOffset);
  auto* velocity = reinterpret_cast<double*>(base + kVelocityOffset);
  auto* predicted = reinterpret_cast<double*>(base + kPredictedOffset);
  auto* validity = reinterpret_cast<uint8_t*>(base + kValidityOffset);
  const hipStream_t s = nullptr;
  k_preset_words<<<1, 64, 0, s>>>(err);
  FL_HIP(hipGetLastError());
  if (face_count) {
    FL_HIP(hipMemcpyAsync(staged, quanta, face_count * sizeof(fl_face_quanta), hipMemcpyHostToDevice, s));
    k_quanta_columns<<<grid_for(face_count), kThreads, 0, s>>>(staged, face_count, qm, qp, qs);
    FL_HIP(hipGetLastError());
  }
  const double g[3] = {acceleration.x, acceleration.y, acceleration.z};
  FL_TRY(launch_normalize_forces(s, qm, qp, qs, face_count, quantization.mass_per_quantum,
                                 quantization.pure_momentum_per_quantum, minimum_face_mass, g, dt, mass, velocity,
                                 validity, predicted, err));
  uint64_t words[2] = {0, 0};
  FL_HIP(hipMemcpyAsync(words, err, sizeof(words), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  if (words[0] != ~uint64_t(0)) return static_cast<int>(words[0] & 0xff);
  if (words[1] != ~uint64_t(0)) return static_cast<int>(words[1] & 0xff);
  const size_t bytes = face_count * sizeof(double);
  if (face_count) {
    if (out_mass) FL_HIP(hipMemcpyAsync(out_mass, mass, bytes, hipMemcpyDeviceToHost, s));
    if (out_velocity) FL_HIP(hipMemcpyAsync(out_velocity, velocity, bytes, hipMemcpyDeviceToHost, s));
    if (out_validity) FL_HIP(hipMemcpyAsync(out_validity, validity, face_count, hipMemcpyDeviceToHost, s));
    if (out_predicted) FL_HIP(hipMemcpyAsync(out_predicted, predicted, bytes, hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
  }
  return FL_OK;
}
