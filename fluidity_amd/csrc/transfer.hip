// APIC transfers on gfx950: P2G quanta, normalize + forces, G2P.
//
// P2G (DepositMacQuanta3D, native/src/transfers_3d.cc:147-197; mac_deposit_3d_internal.h:41-65):
// one workgroup per non-empty sort bin B. Every stencil node of a particle in B lies in the 8^3-cell
// window [4B, 4B+8)^3 = blocks B + {0,1}^3 (sparse_mac_grid_3d.cc:279-283), so the workgroup reduces
// its particles' contributions into a 24 KiB LDS window (2 x 1536 int64) with LDS atomics and then
// adds the non-zero window entries to the owned face slots with one global int64 atomic each.
// Every contribution is rounded half-to-even to integer quanta before any addition, exactly as
// the reference does, so every summation order gives the same bits (F8) and the result is
// identical to the CPU restatement. Workgroup ids are remapped so that consecutive bins run on the
// same XCD and their overlapping windows meet in the same L2.
//
// G2P (GatherMac3D, native/src/transfers_3d.cc:199-261; mac_gather_3d_internal.h:15-69): one
// workgroup per bin stages the window's old and new face velocities in LDS (24 KiB) with coalesced
// block loads; each particle contracts its 3x3x3 tensor axis by axis (x, then y, then z) in FP64,
// which computes the same sums as the reference's 27-sample loop with far fewer operations.
#include "context.h"

namespace fl {
namespace {

constexpr int kWindow = 8 * 192;

// Bijective workgroup -> bin map giving each XCD a contiguous run of bins (workgroups are dispatched
// round-robin over the 8 XCDs).
__device__ inline uint32_t xcd_chunked(uint32_t wg, uint32_t n) {
  const uint32_t q = n >> 3, r = n & 7, xcd = wg & 7, slot = wg >> 3;
  return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot;
}

__device__ inline void bin_coords(const BinList& bins, uint32_t key, int32_t b[3]) {
  b[0] = bins.box.lo[0] + int32_t(key % uint32_t(bins.box.dim[0]));
  b[1] = bins.box.lo[1] + int32_t((key / uint32_t(bins.box.dim[0])) % uint32_t(bins.box.dim[1]));
  b[2] = bins.box.lo[2] + int32_t(key / (uint32_t(bins.box.dim[0]) * uint32_t(bins.box.dim[1])));
}

__device__ inline void report(uint64_t* err, uint64_t ordinal, int code) {
  atomicMin(reinterpret_cast<unsigned long long*>(err), static_cast<unsigned long long>(error_word(ordinal, code)));
}

// Window block of a coordinate's sample range, as a 2-bit mask of {low, high}.
__device__ inline uint32_t axis_blocks(int32_t first, int32_t tile) {
  const int32_t r0 = first - tile, r2 = r0 + 2;
  return (1u << (r0 >> 2)) | (1u << (r2 >> 2));
}

__device__ inline bool window_ok(const int32_t* wblock, uint32_t mx, uint32_t my, uint32_t mz) {
  bool ok = true;
#pragma unroll
  for (int z = 0; z < 2; ++z)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int x = 0; x < 2; ++x)
        ok = ok && !(((mx >> x) & 1) && ((my >> y) & 1) && ((mz >> z) & 1) && wblock[(z * 2 + y) * 2 + x] < 0);
  return ok;
}

// Samples of one coordinate: lattice index of sample 0, the three weight factors, the face offsets
// (meters) and the weight derivatives in cell units. Same arithmetic as axis_stencil; every array
// is indexed by unrolled constants so the stencil stays in registers.
struct Axis3 {
  int32_t first;
  double w[3];
  double off[3];
  double dw[3];
};

// kInterior: every lane of the wave has an interior coordinate (CoordT::interior), so the support
// range checks are known to pass and every offset takes the compensated branch.
template <bool kInterior, class Coord>
__device__ __forceinline__ int axis3(const Coord& c, double lattice_shift, int32_t bin, double h, Axis3& s) {
  const double start = lattice_shift + 0.5;
  if (!kInterior) {
    const double first_support = static_cast<double>(kI32Min) + start;
    const double past_last_support = static_cast<double>(kI32Max - 1) + start;
    if (c.below(first_support) || !c.below(past_last_support)) return kOutOfRange;
  }
  // first is an integer in [-2^31, 2^31 - 2]; the tile test [4B, 4B+8) runs on exact doubles.
  const double first = floor_shifted_d(c, start);
  const double r = first - 4.0 * static_cast<double>(bin);
  if (!(r >= 0.0 && r <= 5.0)) return kFailedPrecondition;
  const double a0 = first + lattice_shift;
  double dl, dr, cc;
  if (kInterior) {
    dl = c.offset_interior(a0 + 0.5);
    dr = fabs(c.offset_interior(a0 + 1.5));
    cc = c.offset_interior(a0 + 1.0);
  } else {
    dl = c.offset(a0 + 0.5);
    dr = fabs(c.offset(a0 + 1.5));
    cc = c.offset(a0 + 1.0);
  }
  s.first = static_cast<int32_t>(first);
  s.w[0] = 0.5 * dr * dr;
  s.w[1] = 0.75 - cc * cc;
  s.w[2] = 0.5 * dl * dl;
  s.dw[0] = -dr;
  s.dw[1] = -2.0 * cc;
  s.dw[2] = dl;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    s.off[i] = kInterior ? c.displacement_interior(a0 + static_cast<double>(i), h)
                         : c.displacement(a0 + static_cast<double>(i), h);
  return kOk;
}

// First failure among the stencils of one face axis (coordinates x, y, z in order) and its offsets.
__device__ __forceinline__ int axis_status(int cx, int cy, int cz, const Axis3& X, const Axis3& Y, const Axis3& Z) {
  int code = cx ? cx : (cy ? cy : cz);
  if (code) return code;
  bool fin = true;
#pragma unroll
  for (int i = 0; i < 3; ++i) fin = fin && finite(X.off[i]) && finite(Y.off[i]) && finite(Z.off[i]);
  return fin ? kOk : kOutOfRange;
}

// Integer quanta of one contribution x / quantum = x * scale, rounded half to even
// (transfers_3d.cc:178-187). The scale is a power of two, so x * scale is exact.
// kSmall: the contribution limit is at most 2^50, so every in-range scaled value s lies in
// [-2^50, 2^50] and s + 1.5 * 2^52 lies in [2^52, 2^53), where the FP64 spacing is 1: the single
// rounding of fma(x, scale, 1.5 * 2^52) rounds s half to even to an integer r (1.5 * 2^52 is even)
// and the result's bit pattern is that of 1.5 * 2^52 plus r. Values beyond the limit fail the call,
// so whatever bits they produce never reach an output.
template <bool kSmall>
__device__ __forceinline__ unsigned long long quanta(double x, double scale) {
  if constexpr (kSmall) {
    const double t = fma(x, scale, 0x1.8p52);
    return static_cast<unsigned long long>(__double_as_longlong(t)) - 0x4338000000000000ull;
  }
  return static_cast<unsigned long long>(static_cast<long long>(rint(x * scale)));
}

// Range test of the 27 scaled contributions of one face axis: |mass / quantum| and
// |momentum / quantum| must not exceed the limit (transfers_3d.cc:178-187); NaN and infinities fail.
// A per-particle bound settles the common case: every weight is below 1, so |mass| <= |m|, and by
// monotonic rounding |velocity| <= |v| + ((OX + OY) + OZ) with OX the rounded sum of the |cx ox_i|,
// so |mass / qm| <= |m| / qm and |momentum / qp| <= (|m| V) / qp. NaN anywhere fails the bound and
// falls through to the exact test of every contribution.
__device__ __forceinline__ bool range_ok(const Axis3& X, const Axis3& Y, const Axis3& Z, double m, double v,
                                         double cx, double cy, double cz, double inv_qm, double inv_qp,
                                         double limit) {
  double ox[3], oy[3], oz[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ox[i] = cx * X.off[i];
    oy[i] = cy * Y.off[i];
    oz[i] = cz * Z.off[i];
  }
  const double OX = (fabs(ox[0]) + fabs(ox[1])) + fabs(ox[2]);
  const double OY = (fabs(oy[0]) + fabs(oy[1])) + fabs(oy[2]);
  const double OZ = (fabs(oz[0]) + fabs(oz[1])) + fabs(oz[2]);
  const double V = fabs(v) + ((OX + OY) + OZ);
  if (fabs(m) * inv_qm <= limit && (fabs(m) * V) * inv_qp <= limit) return true;
  bool ok = true;
#pragma unroll
  for (int zi = 0; zi < 3; ++zi)
#pragma unroll
    for (int yi = 0; yi < 3; ++yi)
#pragma unroll
      for (int xi = 0; xi < 3; ++xi) {
        const double mass = ((X.w[xi] * Y.w[yi]) * Z.w[zi]) * m;
        const double mom = mass * (v + ((ox[xi] + oy[yi]) + oz[zi]));
        ok = ok && fabs(mass * inv_qm) <= limit && fabs(mom * inv_qp) <= limit;
      }
  return ok;
}

// Deposits the 27 samples of one face axis into the LDS window. Contributions are formed exactly as in
// AccumulateMacParticleComponent (mac_deposit_3d_internal.h:50-57): weight (wx wy) wz, mass w m,
// velocity v + ((cx ox + cy oy) + cz oz), momentum mass * velocity.
// kF64: the FP64 deposit of DepositMac3D (transfers_3d.cc:139-145): mass and momentum are added to
// the window as doubles (scales are 1); otherwise both are converted to integer quanta first.
template <int A, bool kSmall, bool kF64>
__device__ __forceinline__ void deposit_axis(const Axis3& X, const Axis3& Y, const Axis3& Z, const int32_t* tile,
                                             double m, double v, double cx, double cy, double cz, double inv_qm,
                                             double inv_qp, unsigned long long* wm, unsigned long long* wq) {
  // Window slot of sample (xi, yi, zi) = fz[zi] + fy[yi] + fx[xi]: window block (rz>>2, ry>>2, rx>>2)
  // times 192, face axis A times 64, cell (rz&3)*16 + (ry&3)*4 + (rx&3).
  int fx[3], fy[3], fz[3];
  double ox[3], oy[3], oz[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int rx = X.first + i - tile[0], ry = Y.first + i - tile[1], rz = Z.first + i - tile[2];
    fx[i] = (rx >> 2) * 192 + (rx & 3) + A * 64;
    fy[i] = (ry >> 2) * 384 + (ry & 3) * 4;
    fz[i] = (rz >> 2) * 768 + (rz & 3) * 16;
    ox[i] = cx * X.off[i];
    oy[i] = cy * Y.off[i];
    oz[i] = cz * Z.off[i];
  }
#pragma unroll
  for (int zi = 0; zi < 3; ++zi) {
#pragma unroll
    for (int yi = 0; yi < 3; ++yi) {
      const double wxy[3] = {X.w[0] * Y.w[yi], X.w[1] * Y.w[yi], X.w[2] * Y.w[yi]};
      const double oxy[3] = {ox[0] + oy[yi], ox[1] + oy[yi], ox[2] + oy[yi]};
#pragma unroll
      for (int xi = 0; xi < 3; ++xi) {
        const double mass = (wxy[xi] * Z.w[zi]) * m;
        const double mom = mass * (v + (oxy[xi] + oz[zi]));
        const int slot = fz[zi] + fy[yi] + fx[xi];
        if constexpr (kF64) {
          atomicAdd(reinterpret_cast<double*>(wm) + slot, mass);
          atomicAdd(reinterpret_cast<double*>(wq) + slot, mom);
        } else {
          atomicAdd(&wm[slot], quanta<kSmall>(mass, inv_qm));
          atomicAdd(&wq[slot], quanta<kSmall>(mom, inv_qp));
        }
      }
    }
  }
}

// One particle's three face axes in the reference's order (transfers_3d.cc:75-100): per axis the
// stencil status, the window ownership, then the contribution range. Returns the first failure.
template <bool kInterior, bool kSmall, bool kF64, class Coord>
__device__ __forceinline__ int deposit_particle(const Coord& cx, const Coord& cy, const Coord& cz, const int32_t* bin,
                                                const int32_t* tile, const int32_t* wblock, double h,
                                                const ParticleView& p, uint32_t k, double inv_qm, double inv_qp,
                                                double limit, unsigned long long* wm, unsigned long long* wq,
                                                int code) {
  const double m = p.m[k];
  // Normal (shift 0) and transverse (shift 1/2) samples per coordinate, each computed just before
  // its first use so that fewer stencils are live at once.
  Axis3 ty, tz;
  const int cty = axis3<kInterior>(cy, 0.5, bin[1], h, ty), ctz = axis3<kInterior>(cz, 0.5, bin[2], h, tz);
  {  // face axis x
    Axis3 nx;
    const int cnx = axis3<kInterior>(cx, 0.0, bin[0], h, nx);
    if (code == kOk) code = axis_status(cnx, cty, ctz, nx, ty, tz);
    if (code == kOk && !window_ok(wblock, axis_blocks(nx.first, tile[0]), axis_blocks(ty.first, tile[1]),
                                  axis_blocks(tz.first, tile[2])))
      code = kFailedPrecondition;
    if (code == kOk) {
      const double v = p.v[0][k], c0 = p.c[0][k], c1 = p.c[1][k], c2 = p.c[2][k];
      if (!range_ok(nx, ty, tz, m, v, c0, c1, c2, inv_qm, inv_qp, limit)) code = kOutOfRange;
      deposit_axis<0, kSmall, kF64>(nx, ty, tz, tile, m, v, c0, c1, c2, inv_qm, inv_qp, wm, wq);
    }
  }
  Axis3 tx;
  const int ctx_ = axis3<kInterior>(cx, 0.5, bin[0], h, tx);
  {  // face axis y
    Axis3 ny;
    const int cny = axis3<kInterior>(cy, 0.0, bin[1], h, ny);
    if (code == kOk) code = axis_status(ctx_, cny, ctz, tx, ny, tz);
    if (code == kOk && !window_ok(wblock, axis_blocks(tx.first, tile[0]), axis_blocks(ny.first, tile[1]),
                                  axis_blocks(tz.first, tile[2])))
      code = kFailedPrecondition;
    if (code == kOk) {
      const double v = p.v[1][k], c0 = p.c[3][k], c1 = p.c[4][k], c2 = p.c[5][k];
      if (!range_ok(tx, ny, tz, m, v, c0, c1, c2, inv_qm, inv_qp, limit)) code = kOutOfRange;
      deposit_axis<1, kSmall, kF64>(tx, ny, tz, tile, m, v, c0, c1, c2, inv_qm, inv_qp, wm, wq);
    }
  }
  {  // face axis z
    Axis3 nz;
    const int cnz = axis3<kInterior>(cz, 0.0, bin[2], h, nz);
    if (code == kOk) code = axis_status(ctx_, cty, cnz, tx, ty, nz);
    if (code == kOk && !window_ok(wblock, axis_blocks(tx.first, tile[0]), axis_blocks(ty.first, tile[1]),
                                  axis_blocks(nz.first, tile[2])))
      code = kFailedPrecondition;
    if (code == kOk) {
      const double v = p.v[2][k], c0 = p.c[6][k], c1 = p.c[7][k], c2 = p.c[8][k];
      if (!range_ok(tx, ty, nz, m, v, c0, c1, c2, inv_qm, inv_qp, limit)) code = kOutOfRange;
      deposit_axis<2, kSmall, kF64>(tx, ty, nz, tile, m, v, c0, c1, c2, inv_qm, inv_qp, wm, wq);
    }
  }
  return code;
}

__device__ inline uint32_t gcd_u32(uint32_t a, uint32_t b) {
  while (b) {
    const uint32_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// Three waves per SIMD: the unrolled FP64 stencils need 168 VGPRs, which the compiler fits without
// spilling at this occupancy; unconstrained it takes 179 and only two waves fit.
template <bool kTrial, bool kPow2, bool kSmall, bool kF64>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_p2g(ParticleView p, BinList bins, GridView grid, double inv_qm,
                                                  double inv_qp, double limit, const uint32_t* perm,
                                                  unsigned long long* qm, unsigned long long* qp, uint64_t* err) {
  __shared__ unsigned long long wm[kWindow], wq[kWindow];
  __shared__ int32_t wblock[8];
  __shared__ uint32_t stride;
  const uint32_t b = xcd_chunked(blockIdx.x, gridDim.x);
  int32_t bin[3];
  bin_coords(bins, bins.key[b], bin);
  const uint32_t off = bins.offset[b], cnt = bins.count[b];
  for (int i = threadIdx.x; i < kWindow; i += kThreads) {
    wm[i] = 0;
    wq[i] = 0;
  }
  if (threadIdx.x < 8)
    wblock[threadIdx.x] = grid.ordinal(bin[0] + (threadIdx.x & 1), bin[1] + ((threadIdx.x >> 1) & 1),
                                       bin[2] + (threadIdx.x >> 2));
  if (threadIdx.x == 0) {
    // Visit the bin with a stride coprime to its size: particles stored next to each other (often in
    // the same cell, so with the same stencil) go to different lanes, which spreads LDS atomics.
    uint32_t st = 9;
    while (cnt > 1 && gcd_u32(st, cnt) != 1) st += 2;
    stride = st;
  }
  __syncthreads();
  const Geometry g = grid.g;
  const int32_t tile[3] = {4 * bin[0], 4 * bin[1], 4 * bin[2]};
  // Lane visit order: v -> (v * stride) mod cnt, advanced by (kThreads * stride) mod cnt per pass.
  const uint32_t st = stride, step = cnt ? (kThreads * st) % cnt : 0;
  uint32_t idx = cnt ? (threadIdx.x * st) % cnt : 0;
  for (uint32_t v = threadIdx.x; v < cnt; v += kThreads) {
    const uint32_t k = off + idx;
    idx += step;
    idx = idx >= cnt ? idx - cnt : idx;
    double t[3] = {p.s[0][k], p.s[1][k], p.s[2][k]};
    int code = kOk;
    if (kTrial) {
      const double s0 = t[0], s1 = t[1], s2 = t[2];
      t[0] = p.t[0][k];
      t[1] = p.t[1][k];
      t[2] = p.t[2][k];
      if (!(within_half_cell(s0, t[0], g.h) && within_half_cell(s1, t[1], g.h) && within_half_cell(s2, t[2], g.h)))
        code = kFailedPrecondition;
    }
    const auto cx = normalize<kPow2>(t[0], g.ox, g.h, g.inv_h);
    const auto cy = normalize<kPow2>(t[1], g.oy, g.h, g.inv_h);
    const auto cz = normalize<kPow2>(t[2], g.oz, g.h, g.inv_h);
    if (__all(cx.interior() && cy.interior() && cz.interior()))
      code = deposit_particle<true, kSmall, kF64>(cx, cy, cz, bin, tile, wblock, g.h, p, k, inv_qm, inv_qp, limit,
                                                  wm, wq, code);
    else
      code = deposit_particle<false, kSmall, kF64>(cx, cy, cz, bin, tile, wblock, g.h, p, k, inv_qm, inv_qp, limit,
                                                   wm, wq, code);
    if (code) report(err, perm[k], code);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kWindow; i += kThreads) {
    const int32_t ord = wblock[i / 192];
    if (ord < 0) continue;
    const uint64_t slot = uint64_t(ord) * 192 + (i % 192);
    if constexpr (kF64) {
      if (wm[i]) unsafeAtomicAdd(reinterpret_cast<double*>(qm) + slot, __longlong_as_double(static_cast<long long>(wm[i])));
      if (wq[i]) unsafeAtomicAdd(reinterpret_cast<double*>(qp) + slot, __longlong_as_double(static_cast<long long>(wq[i])));
    } else {
      if (wm[i]) atomicAdd(&qm[slot], wm[i]);
      if (wq[i]) atomicAdd(&qp[slot], wq[i]);
    }
  }
}

// ---- FP32 transfer arithmetic (the FP32 deposit of the GPU configurations) ----
// The same bins, LDS windows, window ownership and integer quanta as k_p2g, with the stencil weights,
// offsets and contributions in FP32: each particle's lattice coordinate is taken relative to its bin's
// window in FP64 and rounded once to FP32 (a value in [1.5, 5.5), so within 5e-7 of a cell); weights
// are the quadratic B-spline of that coordinate, offsets (node - x) h, mass w m and momentum
// mass (v + c . offset) in FP32. Each contribution is rounded half to even to integer quanta (through
// an exact FP64 add of 1.5 * 2^52) before any addition, so the sums are deterministic in any order; they
// agree with the FP64 deposit to the FP32 rounding of the contributions (1e-6 relative).
struct Axis3f {
  int32_t r;  // first node relative to the window's low corner, in [0, 5]
  float w[3];
  float off[3];
};

__device__ __forceinline__ Axis3f axis3f(float xr, float shift, float h) {
  // Nodes at (first + shift + i), first = floor(xr - shift - 0.5); fx = xr - (first + shift) in [0.5, 1.5].
  const float base = floorf(xr - shift - 0.5f);
  Axis3f s;
  const float f = xr - (base + shift);
  s.r = min(max(static_cast<int32_t>(base), 0), 5);
  const float a = 1.5f - f, c = f - 1.0f, d = f - 0.5f;
  s.w[0] = 0.5f * a * a;
  s.w[1] = 0.75f - c * c;
  s.w[2] = 0.5f * d * d;
  s.off[0] = -f * h;
  s.off[1] = (1.0f - f) * h;
  s.off[2] = (2.0f - f) * h;
  return s;
}

__device__ __forceinline__ unsigned long long quanta_f32(float x) {
  const double t = static_cast<double>(x) + 0x1.8p52;
  return static_cast<unsigned long long>(__double_as_longlong(t)) - 0x4338000000000000ull;
}

template <int A>
__device__ __forceinline__ bool deposit_axis_f32(const Axis3f& X, const Axis3f& Y, const Axis3f& Z, float m, float v,
                                                 float cx, float cy, float cz, float inv_qm, float inv_qp, float limit,
                                                 unsigned long long* wm, unsigned long long* wq) {
  int fx[3], fy[3], fz[3];
  float ox[3], oy[3], oz[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int rx = X.r + i, ry = Y.r + i, rz = Z.r + i;
    fx[i] = (rx >> 2) * 192 + (rx & 3) + A * 64;
    fy[i] = (ry >> 2) * 384 + (ry & 3) * 4;
    fz[i] = (rz >> 2) * 768 + (rz & 3) * 16;
    ox[i] = cx * X.off[i];
    oy[i] = cy * Y.off[i];
    oz[i] = cz * Z.off[i];
  }
  bool ok = true;
#pragma unroll
  for (int zi = 0; zi < 3; ++zi) {
    const float mz = Z.w[zi] * m;
#pragma unroll
    for (int yi = 0; yi < 3; ++yi) {
      const float myz = Y.w[yi] * mz;
      const float vyz = v + oy[yi] + oz[zi];
#pragma unroll
      for (int xi = 0; xi < 3; ++xi) {
        const float mass = X.w[xi] * myz;
        const float qm = mass * inv_qm, qp = mass * (vyz + ox[xi]) * inv_qp;
        ok = ok && fabsf(qm) <= limit && fabsf(qp) <= limit;
        const int slot = fz[zi] + fy[yi] + fx[xi];
        atomicAdd(&wm[slot], quanta_f32(qm));
        atomicAdd(&wq[slot], quanta_f32(qp));
      }
    }
  }
  return ok;
}

// Lattice coordinate (x - o) / h relative to the window corner 4 B, rounded once to FP32.
template <bool kPow2>
__device__ __forceinline__ float window_coord(double x, double o, double h, double inv_h, int32_t corner) {
  const double X = kPow2 ? (x - o) * inv_h : (x - o) / h;
  return static_cast<float>(X - static_cast<double>(corner));
}

template <bool kPow2>
__global__ __launch_bounds__(kThreads) void k_p2g_f32(ParticleView p, BinList bins, GridView grid, double inv_qm,
                                                      double inv_qp, double limit, const uint32_t* perm,
                                                      unsigned long long* qm, unsigned long long* qp, uint64_t* err) {
  __shared__ unsigned long long wm[kWindow], wq[kWindow];
  __shared__ int32_t wblock[8];
  __shared__ uint32_t stride;
  const uint32_t b = xcd_chunked(blockIdx.x, gridDim.x);
  int32_t bin[3];
  bin_coords(bins, bins.key[b], bin);
  const uint32_t off = bins.offset[b], cnt = bins.count[b];
  for (int i = threadIdx.x; i < kWindow; i += kThreads) {
    wm[i] = 0;
    wq[i] = 0;
  }
  if (threadIdx.x < 8)
    wblock[threadIdx.x] = grid.ordinal(bin[0] + (threadIdx.x & 1), bin[1] + ((threadIdx.x >> 1) & 1),
                                       bin[2] + (threadIdx.x >> 2));
  if (threadIdx.x == 0) {
    uint32_t st = 9;
    while (cnt > 1 && gcd_u32(st, cnt) != 1) st += 2;
    stride = st;
  }
  const Geometry g = grid.g;
  __syncthreads();
  const float h = static_cast<float>(g.h);
  const float fqm = static_cast<float>(inv_qm), fqp = static_cast<float>(inv_qp), flimit = static_cast<float>(limit);
  const uint32_t st = stride, step = cnt ? (kThreads * st) % cnt : 0;
  uint32_t idx = cnt ? (threadIdx.x * st) % cnt : 0;
  for (uint32_t v = threadIdx.x; v < cnt; v += kThreads) {
    const uint32_t k = off + idx;
    idx += step;
    idx = idx >= cnt ? idx - cnt : idx;
    const float xr = window_coord<kPow2>(p.s[0][k], g.ox, g.h, g.inv_h, 4 * bin[0]);
    const float yr = window_coord<kPow2>(p.s[1][k], g.oy, g.h, g.inv_h, 4 * bin[1]);
    const float zr = window_coord<kPow2>(p.s[2][k], g.oz, g.h, g.inv_h, 4 * bin[2]);
    const float m = static_cast<float>(p.m[k]);
    const Axis3f nx = axis3f(xr, 0.0f, h), tx = axis3f(xr, 0.5f, h);
    const Axis3f ny = axis3f(yr, 0.0f, h), ty = axis3f(yr, 0.5f, h);
    const Axis3f nz = axis3f(zr, 0.0f, h), tz = axis3f(zr, 0.5f, h);
    int code = kOk;
    if (!window_ok(wblock, axis_blocks(nx.r, 0), axis_blocks(ty.r, 0), axis_blocks(tz.r, 0)) ||
        !window_ok(wblock, axis_blocks(tx.r, 0), axis_blocks(ny.r, 0), axis_blocks(tz.r, 0)) ||
        !window_ok(wblock, axis_blocks(tx.r, 0), axis_blocks(ty.r, 0), axis_blocks(nz.r, 0))) {
      code = kFailedPrecondition;
    } else {
      const float v0 = static_cast<float>(p.v[0][k]), v1 = static_cast<float>(p.v[1][k]),
                  v2 = static_cast<float>(p.v[2][k]);
      float c[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) c[q] = static_cast<float>(p.c[q][k]);
      bool ok = deposit_axis_f32<0>(nx, ty, tz, m, v0, c[0], c[1], c[2], fqm, fqp, flimit, wm, wq);
      ok = deposit_axis_f32<1>(tx, ny, tz, m, v1, c[3], c[4], c[5], fqm, fqp, flimit, wm, wq) && ok;
      ok = deposit_axis_f32<2>(tx, ty, nz, m, v2, c[6], c[7], c[8], fqm, fqp, flimit, wm, wq) && ok;
      if (!ok) code = kOutOfRange;
    }
    if (code) report(err, perm[k], code);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kWindow; i += kThreads) {
    const int32_t ord = wblock[i / 192];
    if (ord < 0) continue;
    const uint64_t slot = uint64_t(ord) * 192 + (i % 192);
    if (wm[i]) atomicAdd(&qm[slot], wm[i]);
    if (wq[i]) atomicAdd(&qp[slot], wq[i]);
  }
}

// A face sum that overflowed from finite contributions: OutOfRange, ordered after every particle's
// own failure (the running-sum position of the overflow is not tracked).
__global__ void k_check_sums(const double* m, const double* q, uint64_t nf, uint64_t n, uint64_t* err) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f < nf && !(finite(m[f]) && finite(q[f]))) report(err, n, kOutOfRange);
}

constexpr int64_t kMaxQuanta = int64_t(1) << 62;

// NormalizeCore (liquid_boundary_transfer_3d.cc:463-516) then ApplyForces (liquid_substep_3d_cpu.cc:156-178).
// err[0] holds the first normalize failure, err[1] the first forces failure.
__global__ void k_normalize_forces(const int64_t* qm, const int64_t* qp, const int64_t* qs, uint64_t nf, double sqm,
                                   double sqp, double min_mass, double gx, double gy, double gz, double dt,
                                   double* mass, double* u0, uint8_t* validity, double* ustar, uint64_t* err) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f >= nf) return;
  const int64_t Qm = qm[f], Qp = qp[f], Qs = qs ? qs[f] : 0;
  double dm = 0.0, u = 0.0;
  uint8_t v = 0;
  if (Qm < 0 || Qm > kMaxQuanta || Qp < -kMaxQuanta || Qp > kMaxQuanta || Qs != 0) {
    report(err, f, kInvalidArgument);
  } else {
    dm = static_cast<double>(Qm) * sqm;
    const double dq = static_cast<double>(Qp) * sqp;
    if (!finite(dm) || !finite(dq)) {
      report(err, f, kOutOfRange);
    } else if (dm != 0.0) {
      if (dm <= min_mass) {
        v = 3;
      } else {
        u = dq / dm;
        if (!finite(u)) report(err, f, kOutOfRange);
        v = 1;
      }
    }
  }
  const int axis = static_cast<int>((f % 192) / 64);
  const double ga = axis == 0 ? gx : (axis == 1 ? gy : gz);
  const double pred = u + dt * ga;
  if (!finite(pred)) report(err + 1, f, kOutOfRange);
  if (mass) mass[f] = dm;
  if (u0) u0[f] = u;
  if (validity) validity[f] = v;
  if (ustar) ustar[f] = pred;
}

// ApplyForces (liquid_substep_3d_cpu.cc:156-178) on its own: u* = u_old + dt * g[axis(f)], where the
// owned-slot axis is (f mod 192) / 64; err[0] holds the first face whose prediction is not finite.
__global__ void k_apply_forces(const double* u_old, uint64_t nf, double gx, double gy, double gz, double dt,
                               double* pred, uint64_t* err) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f >= nf) return;
  const int axis = static_cast<int>((f % 192) / 64);
  const double ga = axis == 0 ? gx : (axis == 1 ? gy : gz);
  const double u = u_old[f] + dt * ga;
  if (!finite(u)) report(err, f, kOutOfRange);
  pred[f] = u;
}

// Gathers one face axis from the LDS window: the separable contraction of the 3x3x3 velocities with
// the weights, the weighted offsets and the weight derivatives (x, then y, then z), in FP64 with
// fused multiply-adds. Results go straight to the component-major output columns.
template <int A>
__device__ __forceinline__ bool gather_axis(const Axis3& X, const Axis3& Y, const Axis3& Z, const int32_t* tile,
                                           double inv_h, const double* wn, const double* wo, double* out,
                                           uint64_t cap, uint64_t k) {
  double wr[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    wr[0][i] = X.w[i] * (X.off[i] * inv_h);
    wr[1][i] = Y.w[i] * (Y.off[i] * inv_h);
    wr[2][i] = Z.w[i] * (Z.off[i] * inv_h);
  }
  double pic = 0.0, pold = 0.0, mx = 0.0, my = 0.0, mz = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
  for (int zi = 0; zi < 3; ++zi) {
    double aa = 0.0, ay = 0.0, ady = 0.0, bx = 0.0, dx = 0.0, ao = 0.0;
    const int rz = Z.first + zi - tile[2];
#pragma unroll
    for (int yi = 0; yi < 3; ++yi) {
      double a = 0.0, b = 0.0, d = 0.0, o = 0.0;
      const int ry = Y.first + yi - tile[1];
#pragma unroll
      for (int xi = 0; xi < 3; ++xi) {
        const int rx = X.first + xi - tile[0];
        const int slot = (((rz >> 2) * 2 + (ry >> 2)) * 2 + (rx >> 2)) * 192 + A * 64 + ((rz & 3) * 16 + (ry & 3) * 4 + (rx & 3));
        const double un = wn[slot], uo = wo[slot];
        a = fma(X.w[xi], un, a);
        b = fma(wr[0][xi], un, b);
        d = fma(X.dw[xi], un, d);
        o = fma(X.w[xi], uo, o);
      }
      aa = fma(Y.w[yi], a, aa);
      ay = fma(wr[1][yi], a, ay);
      ady = fma(Y.dw[yi], a, ady);
      bx = fma(Y.w[yi], b, bx);
      dx = fma(Y.w[yi], d, dx);
      ao = fma(Y.w[yi], o, ao);
    }
    pic = fma(Z.w[zi], aa, pic);
    mz = fma(wr[2][zi], aa, mz);
    gz = fma(Z.dw[zi], aa, gz);
    my = fma(Z.w[zi], ay, my);
    gy = fma(Z.w[zi], ady, gy);
    mx = fma(Z.w[zi], bx, mx);
    gx = fma(Z.w[zi], dx, gx);
    pold = fma(Z.w[zi], ao, pold);
  }
  const double inc = pic - pold;
  const double sx = (4.0 * mx) * inv_h, sy = (4.0 * my) * inv_h, sz = (4.0 * mz) * inv_h;
  gx *= inv_h;
  gy *= inv_h;
  gz *= inv_h;
  out[A * cap + k] = pic;
  out[(3 + A) * cap + k] = inc;
  out[(6 + 3 * A) * cap + k] = sx;
  out[(7 + 3 * A) * cap + k] = sy;
  out[(8 + 3 * A) * cap + k] = sz;
  out[(15 + 3 * A) * cap + k] = gx;
  out[(16 + 3 * A) * cap + k] = gy;
  out[(17 + 3 * A) * cap + k] = gz;
  return finite(pic) && finite(inc) && finite(sx) && finite(sy) && finite(sz) && finite(gx) && finite(gy) &&
         finite(gz);
}

// One particle's three face axes in the reference's order (transfers_3d.cc:232-257).
template <bool kInterior, class Coord>
__device__ __forceinline__ int gather_particle(const Coord& cx, const Coord& cy, const Coord& cz, const int32_t* bin,
                                               const int32_t* tile, const int32_t* wblock, double h, double inv_h,
                                               const double* wn, const double* wo, double* out, uint64_t cap,
                                               uint64_t k, int code) {
  Axis3 ty, tz;
  const int cty = axis3<kInterior>(cy, 0.5, bin[1], h, ty), ctz = axis3<kInterior>(cz, 0.5, bin[2], h, tz);
  {
    Axis3 nx;
    const int cnx = axis3<kInterior>(cx, 0.0, bin[0], h, nx);
    if (code == kOk) code = axis_status(cnx, cty, ctz, nx, ty, tz);
    if (code == kOk && !window_ok(wblock, axis_blocks(nx.first, tile[0]), axis_blocks(ty.first, tile[1]),
                                  axis_blocks(tz.first, tile[2])))
      code = kFailedPrecondition;
    if (code == kOk && !gather_axis<0>(nx, ty, tz, tile, inv_h, wn, wo, out, cap, k)) code = kOutOfRange;
  }
  Axis3 tx;
  const int ctx_ = axis3<kInterior>(cx, 0.5, bin[0], h, tx);
  {
    Axis3 ny;
    const int cny = axis3<kInterior>(cy, 0.0, bin[1], h, ny);
    if (code == kOk) code = axis_status(ctx_, cny, ctz, tx, ny, tz);
    if (code == kOk && !window_ok(wblock, axis_blocks(tx.first, tile[0]), axis_blocks(ny.first, tile[1]),
                                  axis_blocks(tz.first, tile[2])))
      code = kFailedPrecondition;
    if (code == kOk && !gather_axis<1>(tx, ny, tz, tile, inv_h, wn, wo, out, cap, k)) code = kOutOfRange;
  }
  {
    Axis3 nz;
    const int cnz = axis3<kInterior>(cz, 0.0, bin[2], h, nz);
    if (code == kOk) code = axis_status(ctx_, cty, cnz, tx, ty, nz);
    if (code == kOk && !window_ok(wblock, axis_blocks(tx.first, tile[0]), axis_blocks(ty.first, tile[1]),
                                  axis_blocks(nz.first, tile[2])))
      code = kFailedPrecondition;
    if (code == kOk && !gather_axis<2>(tx, ty, nz, tile, inv_h, wn, wo, out, cap, k)) code = kOutOfRange;
  }
  return code;
}

template <bool kTrial, bool kPow2>
__global__ __launch_bounds__(kThreads) void k_g2p(ParticleView p, BinList bins, GridView grid, const double* uold,
                                                  const double* unew, const uint32_t* perm, double* out, uint64_t cap,
                                                  uint64_t* err) {
  __shared__ double wo[kWindow], wn[kWindow];
  __shared__ int32_t wblock[8];
  const uint32_t b = xcd_chunked(blockIdx.x, gridDim.x);
  int32_t bin[3];
  bin_coords(bins, bins.key[b], bin);
  if (threadIdx.x < 8)
    wblock[threadIdx.x] = grid.ordinal(bin[0] + (threadIdx.x & 1), bin[1] + ((threadIdx.x >> 1) & 1),
                                       bin[2] + (threadIdx.x >> 2));
  __syncthreads();
  for (int i = threadIdx.x; i < kWindow; i += kThreads) {
    const int32_t ord = wblock[i / 192];
    const uint64_t slot = uint64_t(ord) * 192 + (i % 192);
    wo[i] = ord >= 0 ? uold[slot] : 0.0;
    wn[i] = ord >= 0 ? unew[slot] : 0.0;
  }
  __syncthreads();
  const Geometry g = grid.g;
  const double inv_h = g.inv_h != 0.0 ? g.inv_h : 1.0 / g.h;
  const uint32_t off = bins.offset[b], cnt = bins.count[b];
  const int32_t tile[3] = {4 * bin[0], 4 * bin[1], 4 * bin[2]};
  for (uint32_t k = off + threadIdx.x; k < off + cnt; k += kThreads) {
    double t[3] = {p.s[0][k], p.s[1][k], p.s[2][k]};
    int code = kOk;
    if (kTrial) {
      const double s0 = t[0], s1 = t[1], s2 = t[2];
      t[0] = p.t[0][k];
      t[1] = p.t[1][k];
      t[2] = p.t[2][k];
      if (!(within_half_cell(s0, t[0], g.h) && within_half_cell(s1, t[1], g.h) && within_half_cell(s2, t[2], g.h)))
        code = kFailedPrecondition;
    }
    const auto cx = normalize<kPow2>(t[0], g.ox, g.h, g.inv_h);
    const auto cy = normalize<kPow2>(t[1], g.oy, g.h, g.inv_h);
    const auto cz = normalize<kPow2>(t[2], g.oz, g.h, g.inv_h);
    if (__all(cx.interior() && cy.interior() && cz.interior()))
      code = gather_particle<true>(cx, cy, cz, bin, tile, wblock, g.h, inv_h, wn, wo, out, cap, k, code);
    else
      code = gather_particle<false>(cx, cy, cz, bin, tile, wblock, g.h, inv_h, wn, wo, out, cap, k, code);
    if (code) report(err, perm[k], code);
  }
}

// ---- Compatible deposit (DepositCompatibleMacQuanta3D: DepositCore, liquid_boundary_transfer_3d.cc:329-460) ----
// Each particle carries a side tag and an affine switch; each owned slot a side tag and structural
// eligibility. Per face axis only the compatible samples (eligible, same side; :93-95) deposit, with
// the normalised weight N = w / S (S the compatible weight sum in stencil order), the quotient-rule
// gradient (grad w - N sum grad w) / S, and the offset recentred on the N-weighted mean offset.
// The workgroup layout is that of k_p2g; the window's slot tags are staged in LDS next to the quanta.

// A spline weight factor with the edge distance that defines an outer weight, for the exponent form
// of small gradients (sparse_mac_grid_3d.cc:236-254): outer samples carry |dw| (their distance to the
// support edge), the middle sample none.
struct WFactor {
  double r, edge;
};

__device__ __forceinline__ WFactor wfactor(const Axis3& s, int i) { return {s.w[i], i == 1 ? 0.0 : fabs(s.dw[i])}; }

__device__ inline double factor_mantissa(WFactor f, int& e) {
  if (__builtin_isnormal(f.r) || f.edge == 0.0) return frexp(f.r, &e);
  int de;
  const double dm = frexp(f.edge, &de);
  e = 2 * de;
  return 0.5 * dm * dm;
}

// PhysicalWeightGradient (sparse_mac_grid_3d.cc:312-336): (fx fy) fz / h, or, when that product is
// not normal and no factor is exactly zero, the same quotient rebuilt from mantissas and exponents.
// With a power-of-two width, g / h and g * (1 / h) are the same correctly rounded value.
template <bool kPow2>
__device__ __forceinline__ double physical_gradient(WFactor fx, WFactor fy, WFactor fz, double h, double inv_h) {
  const double g = (fx.r * fy.r) * fz.r;
  const bool zero = (fx.r == 0.0 && fx.edge == 0.0) || (fy.r == 0.0 && fy.edge == 0.0) ||
                    (fz.r == 0.0 && fz.edge == 0.0);
  if (__builtin_isnormal(g) || zero) return kPow2 ? g * inv_h : g / h;
  int ex, ey, ez, eh;
  const double mx = factor_mantissa(fx, ex), my = factor_mantissa(fy, ey), mz = factor_mantissa(fz, ez);
  const double mh = frexp(h, &eh);
  return scalbn(((mx * my) * mz) / mh, ex + ey + ez - eh);
}

// The three physical gradient components of sample (xi, yi, zi) (sparse_mac_grid_3d.cc:536-541).
template <bool kPow2>
__device__ __forceinline__ void sample_gradient(const Axis3& X, const Axis3& Y, const Axis3& Z, int xi, int yi, int zi,
                                                double h, double inv_h, double g[3]) {
  const WFactor wx = wfactor(X, xi), wy = wfactor(Y, yi), wz = wfactor(Z, zi);
  g[0] = physical_gradient<kPow2>({X.dw[xi], 0.0}, wy, wz, h, inv_h);
  g[1] = physical_gradient<kPow2>(wx, {Y.dw[yi], 0.0}, wz, h, inv_h);
  g[2] = physical_gradient<kPow2>(wx, wy, {Z.dw[zi], 0.0}, h, inv_h);
}

// The exact compatible gradient sum G and quotient-rule components (g - N G) / S of one face axis, in
// stencil order (:376-419), for the rare cases where their finiteness is not settled by a bound, and
// for the stencil records (kOut: also writes every record's zero defaults and the components).
// Returns OutOfRange for a non-finite S or G, FailedPrecondition for S <= 0, OutOfRange for a
// non-finite component.
template <bool kPow2, bool kOut>
__device__ __noinline__ int exact_compatible_gradients(const Axis3& X, const Axis3& Y, const Axis3& Z, uint32_t mask,
                                                      double S, double h, double inv_h, fl_compatible_sample* out) {
  double G[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 27; ++k) {
    if (kOut) {
      out[k].normalized_weight = 0.0;
      out[k].normalized_weight_gradient = fl_vec3{0.0, 0.0, 0.0};
      out[k].centered_face_offset = fl_vec3{0.0, 0.0, 0.0};
    }
    if (!((mask >> k) & 1)) continue;
    double g[3];
    sample_gradient<kPow2>(X, Y, Z, k % 3, (k / 3) % 3, k / 9, h, inv_h, g);
    G[0] += g[0];
    G[1] += g[1];
    G[2] += g[2];
  }
  if (!(finite(S) && finite(G[0]) && finite(G[1]) && finite(G[2]))) return kOutOfRange;
  if (!(S > 0.0)) return kFailedPrecondition;
  bool ok = true;
  for (int k = 0; k < 27; ++k) {
    if (!((mask >> k) & 1)) continue;
    const int xi = k % 3, yi = (k / 3) % 3, zi = k / 9;
    const double n = ((X.w[xi] * Y.w[yi]) * Z.w[zi]) / S;
    double g[3];
    sample_gradient<kPow2>(X, Y, Z, xi, yi, zi, h, inv_h, g);
    const double n0 = (g[0] - n * G[0]) / S, n1 = (g[1] - n * G[1]) / S, n2 = (g[2] - n * G[2]) / S;
    ok = ok && finite(n0) && finite(n1) && finite(n2);
    if (kOut) out[k].normalized_weight_gradient = fl_vec3{n0, n1, n2};
  }
  return ok ? kOk : kOutOfRange;
}

// One face axis of one particle. compat(window index) tells whether a sample is compatible. kOut: write
// the 27 CompatibleMacFaceSample3D records to out (slots from the window's block ordinals) instead of
// depositing. Returns the axis status after the stencil checks: FailedPrecondition without positive
// compatible support, OutOfRange for any non-finite or out-of-range value (:376-456).
template <int A, bool kSmall, bool kPow2, bool kOut, class Compat>
__device__ __forceinline__ int compatible_axis(const Axis3& X, const Axis3& Y, const Axis3& Z, const int32_t* tile,
                                               const int32_t* wblock, Compat compat, double m, double v, double c0,
                                               double c1, double c2, double h, double inv_h, double inv_qm,
                                               double inv_qp, double limit, unsigned long long* wm,
                                               unsigned long long* wq, fl_compatible_sample* out) {
  int fx[3], fy[3], fz[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int rx = X.first + i - tile[0], ry = Y.first + i - tile[1], rz = Z.first + i - tile[2];
    fx[i] = (rx >> 2) * 192 + (rx & 3) + A * 64;
    fy[i] = (ry >> 2) * 384 + (ry & 3) * 4;
    fz[i] = (rz >> 2) * 768 + (rz & 3) * 16;
  }
  // Compatible weight sum in stencil order (:376-383).
  uint32_t mask = 0;
  double S = 0.0;
#pragma unroll
  for (int zi = 0; zi < 3; ++zi)
#pragma unroll
    for (int yi = 0; yi < 3; ++yi)
#pragma unroll
      for (int xi = 0; xi < 3; ++xi) {
        if (!compat(fz[zi] + fy[yi] + fx[xi])) continue;
        mask |= 1u << ((zi * 3 + yi) * 3 + xi);
        S += (X.w[xi] * Y.w[yi]) * Z.w[zi];
      }
  // The gradient sum G and the quotient-rule components (g - N G) / S only matter through their
  // finiteness here. Every |g| is at most T = max over components of the product of the factor maxima
  // divided by h (monotonic rounding; an underflowed outer weight's exact value is below the middle
  // weight, which is at least 1/2), so |G| <= 28 T and, with |N| <= 1 (w is one of the summands of S),
  // every component is at most 30 T / S. When 64 T / S is finite nothing needs forming; otherwise the
  // exact sums decide.
  const double wx = fmax(X.w[0], fmax(X.w[1], X.w[2])), wy = fmax(Y.w[0], fmax(Y.w[1], Y.w[2])),
               wz = fmax(Z.w[0], fmax(Z.w[1], Z.w[2]));
  const double dx = fmax(fabs(X.dw[0]), fmax(fabs(X.dw[1]), fabs(X.dw[2]))),
               dy = fmax(fabs(Y.dw[0]), fmax(fabs(Y.dw[1]), fabs(Y.dw[2]))),
               dz = fmax(fabs(Z.dw[0]), fmax(fabs(Z.dw[1]), fabs(Z.dw[2])));
  const double P = fmax((dx * wy) * wz, fmax((wx * dy) * wz, (wx * wy) * dz));
  const double T = kPow2 ? P * inv_h : P / h;
  if (kOut || !finite((64.0 * T) / S)) {
    const int code = exact_compatible_gradients<kPow2, kOut>(X, Y, Z, mask, S, h, inv_h, out);
    if (code != kOk) return code;
  } else if (!(S > 0.0)) {
    return finite(S) ? kFailedPrecondition : kOutOfRange;
  }
  // Normalised weights and the weighted mean offset (:397-419).
  double N[27];
  double M[3] = {0.0, 0.0, 0.0};
  bool ok = true;
#pragma unroll
  for (int zi = 0; zi < 3; ++zi)
#pragma unroll
    for (int yi = 0; yi < 3; ++yi)
#pragma unroll
      for (int xi = 0; xi < 3; ++xi) {
        const int k = (zi * 3 + yi) * 3 + xi;
        N[k] = 0.0;
        if (kOut) {
          const int wi = fz[zi] + fy[yi] + fx[xi];
          out[k].face_slot = uint64_t(wblock[wi / 192]) * 192 + uint64_t(wi % 192);
          out[k].compatible = (mask >> k) & 1;
        }
        if (!((mask >> k) & 1)) continue;
        N[k] = ((X.w[xi] * Y.w[yi]) * Z.w[zi]) / S;
        if (kOut) out[k].normalized_weight = N[k];
        M[0] += X.off[xi] * N[k];
        M[1] += Y.off[yi] * N[k];
        M[2] += Z.off[zi] * N[k];
      }
  if (!(ok && finite(M[0]) && finite(M[1]) && finite(M[2]))) return kOutOfRange;
  // Recentred offsets and the quantized contributions (:420-456).
#pragma unroll
  for (int zi = 0; zi < 3; ++zi)
#pragma unroll
    for (int yi = 0; yi < 3; ++yi)
#pragma unroll
      for (int xi = 0; xi < 3; ++xi) {
        const int k = (zi * 3 + yi) * 3 + xi;
        if (!((mask >> k) & 1)) continue;
        const double rx = X.off[xi] - M[0], ry = Y.off[yi] - M[1], rz = Z.off[zi] - M[2];
        const double mass = m * N[k];
        const double mom = mass * (v + ((c0 * rx + c1 * ry) + c2 * rz));
        ok = ok && finite(rx) && finite(ry) && finite(rz) && fabs(mass * inv_qm) <= limit &&
             fabs(mom * inv_qp) <= limit;
        if constexpr (kOut) {
          out[k].centered_face_offset = fl_vec3{rx, ry, rz};
        } else {
          const int wi = fz[zi] + fy[yi] + fx[xi];
          atomicAdd(&wm[wi], quanta<kSmall>(mass, inv_qm));
          atomicAdd(&wq[wi], quanta<kSmall>(mom, inv_qp));
        }
      }
  return ok ? kOk : kOutOfRange;
}

// One particle's three face axes in the reference's order (:354-457): per axis the stencil status and
// window ownership, the enabled affine row's finiteness (InvalidArgument), then the compatible deposit.
// Mass and velocity were checked while binning (:346-352).
template <bool kInterior, bool kSmall, bool kPow2, bool kOut, class Coord, class Compat>
__device__ __forceinline__ int compatible_particle(const Coord& cx, const Coord& cy, const Coord& cz,
                                                   const int32_t* bin, const int32_t* tile, const int32_t* wblock,
                                                   double h, double inv_h, const ParticleView& p, uint64_t k,
                                                   bool affine, Compat compat, double inv_qm, double inv_qp,
                                                   double limit, unsigned long long* wm, unsigned long long* wq,
                                                   fl_compatible_sample* out, int code) {
  const double m = p.m[k];
  Axis3 ty, tz;
  const int cty = axis3<kInterior>(cy, 0.5, bin[1], h, ty), ctz = axis3<kInterior>(cz, 0.5, bin[2], h, tz);
  const auto row = [&](int a, double r[3]) {
#pragma unroll
    for (int j = 0; j < 3; ++j) r[j] = affine ? p.c[3 * a + j][k] : 0.0;
    return finite(r[0]) && finite(r[1]) && finite(r[2]);
  };
  {
    Axis3 nx;
    const int cnx = axis3<kInterior>(cx, 0.0, bin[0], h, nx);
    if (code == kOk) code = axis_status(cnx, cty, ctz, nx, ty, tz);
    if (code == kOk && !window_ok(wblock, axis_blocks(nx.first, tile[0]), axis_blocks(ty.first, tile[1]),
                                  axis_blocks(tz.first, tile[2])))
      code = kFailedPrecondition;
    double r[3];
    if (code == kOk && !row(0, r)) code = kInvalidArgument;
    if (code == kOk)
      code = compatible_axis<0, kSmall, kPow2, kOut>(nx, ty, tz, tile, wblock, compat, m, p.v[0][k], r[0], r[1], r[2],
                                                     h, inv_h, inv_qm, inv_qp, limit, wm, wq, out);
  }
  Axis3 tx;
  const int ctx_ = axis3<kInterior>(cx, 0.5, bin[0], h, tx);
  {
    Axis3 ny;
    const int cny = axis3<kInterior>(cy, 0.0, bin[1], h, ny);
    if (code == kOk) code = axis_status(ctx_, cny, ctz, tx, ny, tz);
    if (code == kOk && !window_ok(wblock, axis_blocks(tx.first, tile[0]), axis_blocks(ny.first, tile[1]),
                                  axis_blocks(tz.first, tile[2])))
      code = kFailedPrecondition;
    double r[3];
    if (code == kOk && !row(1, r)) code = kInvalidArgument;
    if (code == kOk)
      code = compatible_axis<1, kSmall, kPow2, kOut>(tx, ny, tz, tile, wblock, compat, m, p.v[1][k], r[0], r[1], r[2],
                                                     h, inv_h, inv_qm, inv_qp, limit, wm, wq, out ? out + 27 : out);
  }
  {
    Axis3 nz;
    const int cnz = axis3<kInterior>(cz, 0.0, bin[2], h, nz);
    if (code == kOk) code = axis_status(ctx_, cty, cnz, tx, ty, nz);
    if (code == kOk && !window_ok(wblock, axis_blocks(tx.first, tile[0]), axis_blocks(ty.first, tile[1]),
                                  axis_blocks(nz.first, tile[2])))
      code = kFailedPrecondition;
    double r[3];
    if (code == kOk && !row(2, r)) code = kInvalidArgument;
    if (code == kOk)
      code = compatible_axis<2, kSmall, kPow2, kOut>(tx, ty, nz, tile, wblock, compat, m, p.v[2][k], r[0], r[1], r[2],
                                                     h, inv_h, inv_qm, inv_qp, limit, wm, wq, out ? out + 54 : out);
  }
  return code;
}

// Per-slot tags and per-particle attributes of the compatible deposit (device arrays).
struct CompatTags {
  const int64_t* slot_side;   // side tag per owned slot
  const uint8_t* slot_ok;     // structural eligibility per owned slot
  const int64_t* side;        // particle side tag, caller order
  const uint8_t* affine;      // particle affine switch, caller order
};

template <bool kTrial, bool kPow2, bool kSmall>
__global__ __launch_bounds__(kThreads) void k_p2g_compatible(ParticleView p, BinList bins, GridView grid,
                                                             CompatTags tags, double inv_qm, double inv_qp,
                                                             double limit, const uint32_t* perm,
                                                             unsigned long long* qm, unsigned long long* qp,
                                                             uint64_t* err) {
  __shared__ unsigned long long wm[kWindow], wq[kWindow];
  __shared__ int64_t wside[kWindow];
  __shared__ uint8_t wok[kWindow];
  __shared__ int32_t wblock[8];
  const uint32_t b = xcd_chunked(blockIdx.x, gridDim.x);
  int32_t bin[3];
  bin_coords(bins, bins.key[b], bin);
  if (threadIdx.x < 8)
    wblock[threadIdx.x] = grid.ordinal(bin[0] + (threadIdx.x & 1), bin[1] + ((threadIdx.x >> 1) & 1),
                                       bin[2] + (threadIdx.x >> 2));
  __syncthreads();
  for (int i = threadIdx.x; i < kWindow; i += kThreads) {
    wm[i] = 0;
    wq[i] = 0;
    const int32_t ord = wblock[i / 192];
    const uint64_t slot = uint64_t(ord) * 192 + (i % 192);
    wside[i] = ord >= 0 ? tags.slot_side[slot] : 0;
    wok[i] = ord >= 0 ? tags.slot_ok[slot] : 0;
  }
  __syncthreads();
  const Geometry g = grid.g;
  const double inv_h = g.inv_h;
  const int32_t tile[3] = {4 * bin[0], 4 * bin[1], 4 * bin[2]};
  const uint32_t off = bins.offset[b], cnt = bins.count[b];
  for (uint32_t k = off + threadIdx.x; k < off + cnt; k += kThreads) {
    const uint32_t o = perm[k];
    const int64_t side = tags.side[o];
    const bool affine = tags.affine[o] != 0;
    const auto compat = [&](int wi) { return wok[wi] != 0 && wside[wi] == side; };
    double t[3] = {p.s[0][k], p.s[1][k], p.s[2][k]};
    int code = kOk;
    if (kTrial) {
      const double s0 = t[0], s1 = t[1], s2 = t[2];
      t[0] = p.t[0][k];
      t[1] = p.t[1][k];
      t[2] = p.t[2][k];
      if (!(within_half_cell(s0, t[0], g.h) && within_half_cell(s1, t[1], g.h) && within_half_cell(s2, t[2], g.h)))
        code = kFailedPrecondition;
    }
    const auto cx = normalize<kPow2>(t[0], g.ox, g.h, g.inv_h);
    const auto cy = normalize<kPow2>(t[1], g.oy, g.h, g.inv_h);
    const auto cz = normalize<kPow2>(t[2], g.oz, g.h, g.inv_h);
    if (__all(cx.interior() && cy.interior() && cz.interior()))
      code = compatible_particle<true, kSmall, kPow2, false>(cx, cy, cz, bin, tile, wblock, g.h, inv_h, p, k, affine,
                                                             compat, inv_qm, inv_qp, limit, wm, wq, nullptr, code);
    else
      code = compatible_particle<false, kSmall, kPow2, false>(cx, cy, cz, bin, tile, wblock, g.h, inv_h, p, k, affine,
                                                              compat, inv_qm, inv_qp, limit, wm, wq, nullptr, code);
    if (code) report(err, o, code);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kWindow; i += kThreads) {
    const int32_t ord = wblock[i / 192];
    if (ord < 0) continue;
    const uint64_t slot = uint64_t(ord) * 192 + (i % 192);
    if (wm[i]) atomicAdd(&qm[slot], wm[i]);
    if (wq[i]) atomicAdd(&qp[slot], wq[i]);
  }
}

// The compatible stencils of caller particles [lo, lo + count), one thread per particle, from the
// caller-order columns and bins of a call whose deposit succeeded; 81 records per particle.
template <bool kPow2>
__global__ __launch_bounds__(kThreads) void k_compatible_stencils(ParticleView p, const int32_t* bins, GridView grid,
                                                                  CompatTags tags, uint64_t lo, uint64_t count,
                                                                  fl_compatible_sample* out) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= count) return;
  const uint64_t k = lo + i;
  const int32_t bin[3] = {bins[3 * k], bins[3 * k + 1], bins[3 * k + 2]};
  int32_t wblock[8];
#pragma unroll
  for (int w = 0; w < 8; ++w) wblock[w] = grid.ordinal(bin[0] + (w & 1), bin[1] + ((w >> 1) & 1), bin[2] + (w >> 2));
  const int64_t side = tags.side[k];
  const auto compat = [&](int wi) {
    const int32_t ord = wblock[wi / 192];
    const uint64_t slot = uint64_t(ord) * 192 + (wi % 192);
    return ord >= 0 && tags.slot_ok[slot] != 0 && tags.slot_side[slot] == side;
  };
  const Geometry g = grid.g;
  const int32_t tile[3] = {4 * bin[0], 4 * bin[1], 4 * bin[2]};
  // Stencil weights at the trial position (the trial columns alias the sorted ones when equal).
  const auto cx = normalize<kPow2>(p.t[0][k], g.ox, g.h, g.inv_h);
  const auto cy = normalize<kPow2>(p.t[1][k], g.oy, g.h, g.inv_h);
  const auto cz = normalize<kPow2>(p.t[2][k], g.oz, g.h, g.inv_h);
  compatible_particle<false, false, kPow2, true>(cx, cy, cz, bin, tile, wblock, g.h, g.inv_h, p, k, tags.affine[k] != 0,
                                                 compat, 1.0, 1.0, 0x1p62, nullptr, nullptr, out + 81 * i, kOk);
}

// ---- Boundary gather (GatherLiquidBoundaryParticles3D: GatherCore, liquid_boundary_transfer_3d.cc:519-651) ----
// Fast path (k_g2p_boundary, one workgroup per bin): when every stencil sample of a particle is
// structurally eligible on its side and valid or predicted in both histories, the filled fields are the
// histories themselves, and the gather is GatherMac3D's plus the retention residual v - old PIC (old
// PIC summed sample by sample in stencil order, as the reference does) and the low-mass flag. A particle
// with any other sample goes to a list that k_g2p_ghost completes one particle per lane with the
// reference's continuation: same-side donors within two cell widths in owned-slot order, the
// rank-revealing affine fit, the nearest donor when rank-deficient, the wall, and the normal
// replacement at oblique walls.

struct HistoryView {
  const double* u;
  const uint8_t* validity;  // 0 missing, 1 valid, 2 predicted, 3 low mass
  const int64_t* side;
  const uint8_t* eligible;
  const fl_wall_sample* walls;
};

struct BoundaryGatherArgs {
  HistoryView old_h, new_h;
  int same;                // one set of samples serves both histories
  const int64_t* pside;    // particle side tags, caller order
  const uint8_t* paffine;  // particle affine switches, caller order
  double* extra;           // retention x, y, z and the flag word, sorted order, stride `extra_cap`
  uint64_t extra_cap;
  uint32_t* slow;          // sorted indices that need a continuation
  uint32_t* slow_count;
};

__device__ inline bool usable_validity(uint8_t v) { return v == 1 || v == 2; }
__device__ inline int32_t floor_div4_i32(int32_t v) { return v >> 2; }  // arithmetic shift: floor division

// Flag word: bit 0 affine enabled, 1 old-PIC fallback, 2 new-PIC fallback, 3 low-mass fallback.
__device__ inline void write_extra(const BoundaryGatherArgs& a, uint64_t k, const double delta[3], int flags) {
  a.extra[0 * a.extra_cap + k] = delta[0];
  a.extra[1 * a.extra_cap + k] = delta[1];
  a.extra[2 * a.extra_cap + k] = delta[2];
  a.extra[3 * a.extra_cap + k] = static_cast<double>(flags);
}

template <bool kTrial, bool kPow2>
__global__ __launch_bounds__(kThreads) void k_g2p_boundary(ParticleView p, BinList bins, GridView grid,
                                                           BoundaryGatherArgs a, const uint32_t* perm, double* out,
                                                           uint64_t cap, uint64_t* err) {
  __shared__ double wo[kWindow], wn[kWindow];
  __shared__ int64_t so[kWindow], sn[kWindow];
  // Per window slot: bit 0 old usable, 1 old eligible, 2 old low mass, 3-5 the same for the new history.
  __shared__ uint8_t wflag[kWindow];
  __shared__ int32_t wblock[8];
  const uint32_t b = xcd_chunked(blockIdx.x, gridDim.x);
  int32_t bin[3];
  bin_coords(bins, bins.key[b], bin);
  if (threadIdx.x < 8)
    wblock[threadIdx.x] = grid.ordinal(bin[0] + (threadIdx.x & 1), bin[1] + ((threadIdx.x >> 1) & 1),
                                       bin[2] + (threadIdx.x >> 2));
  __syncthreads();
  for (int i = threadIdx.x; i < kWindow; i += kThreads) {
    const int32_t ord = wblock[i / 192];
    const uint64_t slot = uint64_t(ord) * 192 + (i % 192);
    double uo = 0.0, un = 0.0;
    int64_t sdo = 0, sdn = 0;
    int f = 0;
    if (ord >= 0) {
      const uint8_t vo = a.old_h.validity[slot], vn = a.new_h.validity[slot];
      uo = a.old_h.u[slot];
      un = a.new_h.u[slot];
      sdo = a.old_h.side[slot];
      sdn = a.new_h.side[slot];
      f = (usable_validity(vo) ? 1 : 0) | (a.old_h.eligible[slot] ? 2 : 0) | (vo == 3 ? 4 : 0) |
          (usable_validity(vn) ? 8 : 0) | (a.new_h.eligible[slot] ? 16 : 0) | (vn == 3 ? 32 : 0);
    }
    wo[i] = uo;
    wn[i] = un;
    so[i] = sdo;
    sn[i] = sdn;
    wflag[i] = static_cast<uint8_t>(f);
  }
  __syncthreads();
  const Geometry g = grid.g;
  const double inv_h = g.inv_h != 0.0 ? g.inv_h : 1.0 / g.h;
  const uint32_t off = bins.offset[b], cnt = bins.count[b];
  const int32_t tile[3] = {4 * bin[0], 4 * bin[1], 4 * bin[2]};
  for (uint32_t k = off + threadIdx.x; k < off + cnt; k += kThreads) {
    const uint32_t o = perm[k];
    const int64_t side = a.pside[o];
    double t[3] = {p.s[0][k], p.s[1][k], p.s[2][k]};
    int code = kOk;
    if (kTrial) {
      const double s0 = t[0], s1 = t[1], s2 = t[2];
      t[0] = p.t[0][k];
      t[1] = p.t[1][k];
      t[2] = p.t[2][k];
      if (!(within_half_cell(s0, t[0], g.h) && within_half_cell(s1, t[1], g.h) && within_half_cell(s2, t[2], g.h)))
        code = kFailedPrecondition;
    }
    const auto cx = normalize<kPow2>(t[0], g.ox, g.h, g.inv_h);
    const auto cy = normalize<kPow2>(t[1], g.oy, g.h, g.inv_h);
    const auto cz = normalize<kPow2>(t[2], g.oz, g.h, g.inv_h);
    // First pass (the reference's first loop): stencil status, whether any sample needs a continuation,
    // the low-mass flag and the old PIC value sample by sample.
    bool fill = false, low = false;
    double old_pic[3] = {0.0, 0.0, 0.0};
    for (int A = 0; A < 3 && code == kOk && !fill; ++A) {
      Axis3 X, Y, Z;
      const int c0 = axis3<false>(cx, A == 0 ? 0.0 : 0.5, bin[0], g.h, X);
      const int c1 = axis3<false>(cy, A == 1 ? 0.0 : 0.5, bin[1], g.h, Y);
      const int c2 = axis3<false>(cz, A == 2 ? 0.0 : 0.5, bin[2], g.h, Z);
      code = axis_status(c0, c1, c2, X, Y, Z);
      if (code == kOk && !window_ok(wblock, axis_blocks(X.first, tile[0]), axis_blocks(Y.first, tile[1]),
                                    axis_blocks(Z.first, tile[2])))
        code = kFailedPrecondition;
      if (code != kOk) break;
      double sum = 0.0;
      for (int zi = 0; zi < 3; ++zi)
        for (int yi = 0; yi < 3; ++yi)
          for (int xi = 0; xi < 3; ++xi) {
            const int rx = X.first + xi - tile[0], ry = Y.first + yi - tile[1], rz = Z.first + zi - tile[2];
            const int wi = (((rz >> 2) * 2 + (ry >> 2)) * 2 + (rx >> 2)) * 192 + A * 64 +
                           ((rz & 3) * 16 + (ry & 3) * 4 + (rx & 3));
            const int f = wflag[wi];
            const bool ok_old = (f & 1) && (f & 2) && so[wi] == side;
            const bool ok_new = a.same ? ok_old : ((f & 8) && (f & 16) && sn[wi] == side);
            fill = fill || !ok_old || !ok_new;
            low = low || (f & 4) || (f & 32);
            sum += ((X.w[xi] * Y.w[yi]) * Z.w[zi]) * wo[wi];
          }
      old_pic[A] = sum;
    }
    if (code == kOk && fill) {
      a.slow[atomicAdd(a.slow_count, 1u)] = k;
      continue;
    }
    // Second pass: GatherMac3D on the (unfilled) histories.
    if (code == kOk) {  // every stencil passed its checks in the first pass
      Axis3 tx, ty, tz, nx, ny, nz;
      axis3<false>(cx, 0.5, bin[0], g.h, tx);
      axis3<false>(cy, 0.5, bin[1], g.h, ty);
      axis3<false>(cz, 0.5, bin[2], g.h, tz);
      axis3<false>(cx, 0.0, bin[0], g.h, nx);
      axis3<false>(cy, 0.0, bin[1], g.h, ny);
      axis3<false>(cz, 0.0, bin[2], g.h, nz);
      bool ok = gather_axis<0>(nx, ty, tz, tile, inv_h, wn, wo, out, cap, k);
      ok = gather_axis<1>(tx, ny, tz, tile, inv_h, wn, wo, out, cap, k) && ok;
      ok = gather_axis<2>(tx, ty, nz, tile, inv_h, wn, wo, out, cap, k) && ok;
      if (!ok) code = kOutOfRange;
    }
    if (code == kOk) {
      const bool affine = a.paffine[o] != 0;
      double delta[3] = {0.0, 0.0, 0.0};
      if (!low)
        for (int c = 0; c < 3; ++c) delta[c] = p.v[c][k] - old_pic[c];
      if (!(finite(delta[0]) && finite(delta[1]) && finite(delta[2]))) code = kOutOfRange;
      write_extra(a, k, delta, (affine && !low ? 1 : 0) | (low ? 8 : 0));
    }
    if (code) report(err, o, code);
  }
}

// ---- the continuation path ----
constexpr int kMaxGhostRows = 40;  // a radius-two sphere holds at most 36 staggered anchors

__device__ inline bool compatible_slot(const HistoryView& h, uint64_t s, int64_t side) {
  return h.eligible[s] != 0 && h.side[s] == side;
}

// Anchor of owned slot s in cell widths (BuildOwnedAnchors, liquid_boundary_transfer_3d.cc:152-177).
__device__ inline void slot_anchor(const GridView& grid, uint64_t s, double at[3]) {
  const uint64_t b = s / 192;
  const int r = static_cast<int>(s % 192), axis = r / 64, l = r % 64;
  const int lc[3] = {l & 3, (l >> 2) & 3, l >> 4};
  for (int d = 0; d < 3; ++d) at[d] = 4.0 * grid.blocks[3 * b + d] + lc[d] + (axis == d ? 0.0 : 0.5);
}

// FitBoundedLiquidGhostAffineCore3D (base/liquid_boundary_transfer_math_3d.h:34-144): column-pivoted
// Householder QR of rows [1, dx, dy, dz] against velocities scaled by their largest magnitude, with
// norms recomputed by hypot accumulation, strict pivot comparison and rank loss below 1e-10 of the
// largest pivot. *full is false for fewer than 4 rows or rank loss; returns OutOfRange for non-finite
// arithmetic.
__device__ int ghost_fit(int n, double (*m)[4], double* rhs, double* refl, double* at_target, bool* full) {
  *full = false;
  double scale = 0.0;
  for (int r = 0; r < n; ++r) scale = fmax(scale, fabs(rhs[r]));
  if (n < 4) return kOk;
  if (scale == 0.0) scale = 1.0;
  for (int r = 0; r < n; ++r) rhs[r] = rhs[r] / scale;
  int order[4] = {0, 1, 2, 3};
  double largest = 0.0;
  for (int pv = 0; pv < 4; ++pv) {
    int sel = pv;
    double sel_norm = 0.0;
    for (int c = pv; c < 4; ++c) {
      double norm = 0.0;
      for (int r = pv; r < n; ++r) norm = hypot(norm, m[r][c]);
      if (norm > sel_norm) {
        sel_norm = norm;
        sel = c;
      }
    }
    if (!finite(sel_norm)) return kOutOfRange;
    largest = fmax(largest, sel_norm);
    if (sel_norm == 0.0 || sel_norm < 1e-10 * largest) return kOk;
    const int tmp = order[pv];
    order[pv] = order[sel];
    order[sel] = tmp;
    for (int r = 0; r < n; ++r) {
      const double v = m[r][pv];
      m[r][pv] = m[r][sel];
      m[r][sel] = v;
    }
    const double diag = -copysign(sel_norm, m[pv][pv]);
    double rnorm = 0.0;
    for (int r = pv; r < n; ++r) {
      refl[r] = m[r][pv] / sel_norm;
      if (r == pv) refl[r] -= diag / sel_norm;
      rnorm = hypot(rnorm, refl[r]);
    }
    for (int r = pv; r < n; ++r) refl[r] /= rnorm;
    for (int c = pv + 1; c < 4; ++c) {
      double proj = 0.0;
      for (int r = pv; r < n; ++r) proj += refl[r] * m[r][c];
      for (int r = pv; r < n; ++r) {
        m[r][c] -= 2.0 * refl[r] * proj;
        if (!finite(m[r][c])) return kOutOfRange;
      }
    }
    double rproj = 0.0;
    for (int r = pv; r < n; ++r) rproj += refl[r] * rhs[r];
    for (int r = pv; r < n; ++r) rhs[r] -= 2.0 * refl[r] * rproj;
    m[pv][pv] = diag;
  }
  double pc[4] = {0.0, 0.0, 0.0, 0.0}, coef[4] = {0.0, 0.0, 0.0, 0.0};
  for (int pv = 3; pv >= 0; --pv) {
    double res = rhs[pv];
    for (int c = pv + 1; c < 4; ++c) res -= m[pv][c] * pc[c];
    pc[pv] = res / m[pv][pv];
    coef[order[pv]] = pc[pv] * scale;
    if (!finite(coef[order[pv]])) return kOutOfRange;
  }
  *full = true;
  *at_target = coef[0];
  return kOk;
}

__device__ inline double vec_get(const fl_vec3& v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); }

// ContinueComponent (liquid_boundary_transfer_3d.cc:181-240) on the lattice neighbourhood of the target:
// candidates are visited block by block in ordinal order and cell by cell in (z,y,x) order, which is
// the owned-slot order the reference scans.
__device__ __noinline__ int continue_component(const GridView& grid, const HistoryView& h, const double t[3], int c,
                                               int64_t side,
                                               const fl_wall_sample& wall, double* velocity, bool* fallback) {
  double m[kMaxGhostRows][4], rhs[kMaxGhostRows], refl[kMaxGhostRows];
  int n = 0;
  uint64_t nearest = ~uint64_t(0);
  double nearest_d2 = __builtin_huge_val();
  int64_t lo[3], hi[3];
  int32_t blo[3], bhi[3];
  for (int d = 0; d < 3; ++d) {
    const double sh = d == c ? 0.0 : 0.5;
    lo[d] = static_cast<int64_t>(ceil(t[d] - sh - 2.0));
    hi[d] = static_cast<int64_t>(floor(t[d] - sh + 2.0));
    blo[d] = floor_div4_i32(static_cast<int32_t>(lo[d]));
    bhi[d] = floor_div4_i32(static_cast<int32_t>(hi[d]));
  }
  int32_t ords[8];
  int32_t bk[8][3];
  int nb = 0;
  for (int32_t bz = blo[2]; bz <= bhi[2]; ++bz)
    for (int32_t by = blo[1]; by <= bhi[1]; ++by)
      for (int32_t bx = blo[0]; bx <= bhi[0]; ++bx) {
        const int32_t ord = grid.ordinal(bx, by, bz);
        if (ord < 0 || nb == 8) continue;
        int pos = nb++;
        while (pos > 0 && ords[pos - 1] > ord) {  // insertion by ordinal
          ords[pos] = ords[pos - 1];
          bk[pos][0] = bk[pos - 1][0];
          bk[pos][1] = bk[pos - 1][1];
          bk[pos][2] = bk[pos - 1][2];
          --pos;
        }
        ords[pos] = ord;
        bk[pos][0] = bx;
        bk[pos][1] = by;
        bk[pos][2] = bz;
      }
  for (int q = 0; q < nb; ++q)
    for (int lz = 0; lz < 4; ++lz)
      for (int ly = 0; ly < 4; ++ly)
        for (int lx = 0; lx < 4; ++lx) {
          const int64_t L[3] = {4 * int64_t(bk[q][0]) + lx, 4 * int64_t(bk[q][1]) + ly, 4 * int64_t(bk[q][2]) + lz};
          if (L[0] < lo[0] || L[0] > hi[0] || L[1] < lo[1] || L[1] > hi[1] || L[2] < lo[2] || L[2] > hi[2]) continue;
          const uint64_t s = uint64_t(ords[q]) * 192 + uint64_t(c) * 64 + uint64_t(lz * 16 + ly * 4 + lx);
          if (!compatible_slot(h, s, side) || !usable_validity(h.validity[s])) continue;
          const double ax = 4.0 * bk[q][0] + lx + (c == 0 ? 0.0 : 0.5);
          const double ay = 4.0 * bk[q][1] + ly + (c == 1 ? 0.0 : 0.5);
          const double az = 4.0 * bk[q][2] + lz + (c == 2 ? 0.0 : 0.5);
          const double ox = ax - t[0], oy = ay - t[1], oz = az - t[2];
          const double d2 = (ox * ox + oy * oy) + oz * oz;
          if (d2 > 4.0) continue;
          if (n == kMaxGhostRows) return kInternal;
          m[n][0] = 1.0;
          m[n][1] = ox;
          m[n][2] = oy;
          m[n][3] = oz;
          rhs[n] = h.u[s];
          ++n;
          if (d2 < nearest_d2) {
            nearest = s;
            nearest_d2 = d2;
          }
        }
  bool full = false;
  double at = 0.0;
  if (int e = ghost_fit(n, m, rhs, refl, &at, &full)) return e;
  if (full) {
    *velocity = at;
    *fallback = false;
    return kOk;
  }
  if (nearest != ~uint64_t(0)) {
    *velocity = h.u[nearest];
    *fallback = true;
    return kOk;
  }
  if (wall.available && wall.side_tag == side) {
    *velocity = vec_get(wall.wall_velocity, c);
    *fallback = true;
    return kOk;
  }
  return kFailedPrecondition;
}

// FillVelocity (liquid_boundary_transfer_3d.cc:242-306).
__device__ __noinline__ int fill_velocity(const GridView& grid, const HistoryView& h, uint64_t s, int c, int64_t side,
                                          double* velocity, bool* fallback) {
  if (compatible_slot(h, s, side) && usable_validity(h.validity[s])) {
    *velocity = h.u[s];
    *fallback = false;
    return kOk;
  }
  const fl_wall_sample wall = h.walls[s];
  double t[3];
  slot_anchor(grid, s, t);
  if (int e = continue_component(grid, h, t, c, side, wall, velocity, fallback)) return e;
  const double nc = vec_get(wall.unit_normal, c);
  if (wall.available && wall.side_tag == side && nc != 0.0) {
    double fitted[3] = {wall.wall_velocity.x, wall.wall_velocity.y, wall.wall_velocity.z};
    fitted[c] = *velocity;
    for (int other = 0; other < 3; ++other) {
      if (other == c || vec_get(wall.unit_normal, other) == 0.0) continue;
      double v;
      bool fb;
      if (int e = continue_component(grid, h, t, other, side, wall, &v, &fb)) return e;
      fitted[other] = v;
      *fallback = *fallback || fb;
    }
    const double rx = fitted[0] - wall.wall_velocity.x, ry = fitted[1] - wall.wall_velocity.y,
                 rz = fitted[2] - wall.wall_velocity.z;
    *velocity -= ((rx * wall.unit_normal.x + ry * wall.unit_normal.y) + rz * wall.unit_normal.z) * nc;
    if (!finite(*velocity)) return kOutOfRange;
  }
  return kOk;
}

// One particle per lane: GatherCore with continuations (both loops of the reference), the gather
// sums in the reference's sample order.
template <bool kPow2>
__global__ __launch_bounds__(64) void k_g2p_ghost(ParticleView p, const int32_t* bins, GridView grid,
                                                  BoundaryGatherArgs a, const uint32_t* perm, double* out,
                                                  uint64_t cap, uint64_t* err) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= *a.slow_count) return;
  const uint32_t k = a.slow[idx];
  const uint32_t o = perm[k];
  const int64_t side = a.pside[o];
  const Geometry g = grid.g;
  // The bin is recomputed from the sorted position (the checks already passed in the fast path).
  int32_t bin[3];
  const auto sx = normalize<kPow2>(p.s[0][k], g.ox, g.h, g.inv_h);
  const auto sy = normalize<kPow2>(p.s[1][k], g.oy, g.h, g.inv_h);
  const auto sz = normalize<kPow2>(p.s[2][k], g.oz, g.h, g.inv_h);
  checked_bin(sx, bin[0]);
  checked_bin(sy, bin[1]);
  checked_bin(sz, bin[2]);
  const auto cx = normalize<kPow2>(p.t[0][k], g.ox, g.h, g.inv_h);
  const auto cy = normalize<kPow2>(p.t[1][k], g.oy, g.h, g.inv_h);
  const auto cz = normalize<kPow2>(p.t[2][k], g.oz, g.h, g.inv_h);
  double fo[3][27], fn[3][27];
  double old_pic[3] = {0.0, 0.0, 0.0};
  bool ofb = false, nfb = false, low = false;
  int code = kOk;
  for (int A = 0; A < 3 && code == kOk; ++A) {
    Axis3 X, Y, Z;
    const int c0 = axis3<false>(cx, A == 0 ? 0.0 : 0.5, bin[0], g.h, X);
    const int c1 = axis3<false>(cy, A == 1 ? 0.0 : 0.5, bin[1], g.h, Y);
    const int c2 = axis3<false>(cz, A == 2 ? 0.0 : 0.5, bin[2], g.h, Z);
    code = axis_status(c0, c1, c2, X, Y, Z);
    if (code != kOk) break;
    // Ownership of every sample's block (the stencil's owned-slot resolution).
    int32_t ord[27];
    for (int q = 0; q < 27 && code == kOk; ++q) {
      const int32_t lx = X.first + q % 3, ly = Y.first + (q / 3) % 3, lz = Z.first + q / 9;
      ord[q] = grid.ordinal(floor_div4_i32(lx), floor_div4_i32(ly), floor_div4_i32(lz));
      if (ord[q] < 0) code = kFailedPrecondition;
    }
    for (int q = 0; q < 27 && code == kOk; ++q) {
      const int xi = q % 3, yi = (q / 3) % 3, zi = q / 9;
      const int32_t lx = X.first + xi, ly = Y.first + yi, lz = Z.first + zi;
      const uint64_t s = uint64_t(ord[q]) * 192 + uint64_t(A) * 64 +
                         uint64_t((lz & 3) * 16 + (ly & 3) * 4 + (lx & 3));
      double vo, vn;
      bool bo, bn;
      code = fill_velocity(grid, a.old_h, s, A, side, &vo, &bo);
      if (code != kOk) break;
      if (a.same) {
        vn = vo;
        bn = bo;
      } else {
        code = fill_velocity(grid, a.new_h, s, A, side, &vn, &bn);
        if (code != kOk) break;
      }
      fo[A][q] = vo;
      fn[A][q] = vn;
      old_pic[A] += ((X.w[xi] * Y.w[yi]) * Z.w[zi]) * vo;
      ofb = ofb || bo;
      nfb = nfb || bn;
      low = low || a.old_h.validity[s] == 3 || a.new_h.validity[s] == 3;
    }
  }
  // The complete gather on the filled values (mac_gather_3d_internal.h:15-69), samples in stencil order.
  for (int A = 0; A < 3 && code == kOk; ++A) {
    Axis3 X, Y, Z;
    axis3<false>(cx, A == 0 ? 0.0 : 0.5, bin[0], g.h, X);
    axis3<false>(cy, A == 1 ? 0.0 : 0.5, bin[1], g.h, Y);
    axis3<false>(cz, A == 2 ? 0.0 : 0.5, bin[2], g.h, Z);
    double opic = 0.0, npic = 0.0, mom[3] = {0.0, 0.0, 0.0}, grad[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < 27; ++q) {
      const int xi = q % 3, yi = (q / 3) % 3, zi = q / 9;
      const double w = (X.w[xi] * Y.w[yi]) * Z.w[zi];
      const double uo = fo[A][q], un = fn[A][q];
      const double wun = w * un;
      opic += w * uo;
      npic += wun;
      const double r[3] = {X.off[xi], Y.off[yi], Z.off[zi]};
      double gw[3];
      sample_gradient<kPow2>(X, Y, Z, xi, yi, zi, g.h, g.inv_h, gw);
      for (int d = 0; d < 3; ++d) {
        const double rn = kPow2 ? r[d] * g.inv_h : r[d] / g.h;
        mom[d] += rn * wun;
        grad[d] += gw[d] * un;
      }
    }
    double slope[3];
    for (int d = 0; d < 3; ++d) slope[d] = kPow2 ? (4.0 * mom[d]) * g.inv_h : (4.0 * mom[d]) / g.h;
    const double inc = npic - opic;
    bool ok = finite(npic) && finite(inc);
    for (int d = 0; d < 3; ++d) ok = ok && finite(slope[d]) && finite(grad[d]);
    if (!ok) {
      code = kOutOfRange;
      break;
    }
    out[A * cap + k] = npic;
    out[(3 + A) * cap + k] = inc;
    for (int d = 0; d < 3; ++d) {
      out[(6 + 3 * A + d) * cap + k] = slope[d];
      out[(15 + 3 * A + d) * cap + k] = grad[d];
    }
  }
  if (code == kOk) {
    const bool fallback = ofb || nfb || low;
    double delta[3] = {0.0, 0.0, 0.0};
    if (!fallback)
      for (int c = 0; c < 3; ++c) delta[c] = p.v[c][k] - old_pic[c];
    if (!(finite(delta[0]) && finite(delta[1]) && finite(delta[2]))) code = kOutOfRange;
    write_extra(a, k, delta,
                (a.paffine[o] != 0 && !fallback ? 1 : 0) | (ofb ? 2 : 0) | (nfb ? 4 : 0) | (low ? 8 : 0));
  }
  if (code) report(err, o, code);
}

__global__ void k_unpack(const fl_particle* in, uint64_t n, double* base, uint64_t cap) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const double* r = reinterpret_cast<const double*>(in + i);
  for (int c = 0; c < kParticleDoubles; ++c) base[c * cap + i] = r[c];
}

}  // namespace

int unpack_particles(fl_ctx* ctx, uint64_t n) {
  if (!n) return FL_OK;
  k_unpack<<<grid_for(n), kThreads, 0, ctx->stream>>>(ctx->staging_aos, n, ctx->raw.base, ctx->raw.capacity);
  FL_HIP(hipGetLastError());
  ctx->raw.trial_aliases_sorted = false;
  return FL_OK;
}

int p2g_quanta(fl_ctx* ctx, uint64_t n, double inv_qm, double inv_qp, double limit, hipEvent_t before,
               hipEvent_t after, bool f64) {
  if (ctx->transfer_bits == 32 && !f64 && ctx->sorted.trial_aliases_sorted) {
    const uint64_t nf = ctx->nb * 192;
    FL_TRY(fill_i64(ctx, ctx->qm, nf, 0));
    FL_TRY(fill_i64(ctx, ctx->qp, nf, 0));
    if (before) FL_HIP(hipEventRecord(before, ctx->stream));
    if (ctx->nbins && n) {
      auto* kernel = ctx->geometry.inv_h != 0.0 ? k_p2g_f32<true> : k_p2g_f32<false>;
      kernel<<<ctx->nbins, kThreads, 0, ctx->stream>>>(ctx->sorted.view(), bin_list(ctx), grid_view(ctx), inv_qm,
                                                       inv_qp, limit, ctx->perm,
                                                       reinterpret_cast<unsigned long long*>(ctx->qm),
                                                       reinterpret_cast<unsigned long long*>(ctx->qp), ctx->err);
      FL_HIP(hipGetLastError());
    }
    if (after) FL_HIP(hipEventRecord(after, ctx->stream));
    return FL_OK;
  }
  const uint64_t nf = ctx->nb * 192;
  FL_TRY(fill_i64(ctx, ctx->qm, nf, 0));
  FL_TRY(fill_i64(ctx, ctx->qp, nf, 0));
  if (before) FL_HIP(hipEventRecord(before, ctx->stream));
  if (after && (ctx->nbins == 0 || n == 0)) FL_HIP(hipEventRecord(after, ctx->stream));
  if (ctx->nbins == 0 || n == 0) return FL_OK;
  // Instances: trial positions distinct from the sorted ones, power-of-two cell width, limit <= 2^50.
  using K = decltype(&k_p2g<false, true, true, false>);
  static const K table[8] = {k_p2g<false, false, false, false>, k_p2g<false, false, true, false>,
                             k_p2g<false, true, false, false>,  k_p2g<false, true, true, false>,
                             k_p2g<true, false, false, false>,  k_p2g<true, false, true, false>,
                             k_p2g<true, true, false, false>,   k_p2g<true, true, true, false>};
  static const K table_f64[4] = {k_p2g<false, false, false, true>, k_p2g<false, true, false, true>,
                                 k_p2g<true, false, false, true>, k_p2g<true, true, false, true>};
  const bool trial = !ctx->sorted.trial_aliases_sorted, pow2 = ctx->geometry.inv_h != 0.0, small = limit <= 0x1p50;
  auto* kernel = f64 ? table_f64[(trial ? 2 : 0) + (pow2 ? 1 : 0)]
                     : table[(trial ? 4 : 0) + (pow2 ? 2 : 0) + (small ? 1 : 0)];
  kernel<<<ctx->nbins, kThreads, 0, ctx->stream>>>(ctx->sorted.view(), bin_list(ctx), grid_view(ctx), inv_qm, inv_qp,
                                                   limit, ctx->perm, reinterpret_cast<unsigned long long*>(ctx->qm),
                                                   reinterpret_cast<unsigned long long*>(ctx->qp), ctx->err);
  FL_HIP(hipGetLastError());
  if (f64) {
    k_check_sums<<<grid_for(nf), kThreads, 0, ctx->stream>>>(reinterpret_cast<const double*>(ctx->qm),
                                                             reinterpret_cast<const double*>(ctx->qp), nf, n,
                                                             ctx->err);
    FL_HIP(hipGetLastError());
  }
  if (after) FL_HIP(hipEventRecord(after, ctx->stream));
  return FL_OK;
}

int p2g_compatible(fl_ctx* ctx, uint64_t n, const int64_t* slot_side, const uint8_t* slot_ok, const int64_t* side,
                   const uint8_t* affine, double inv_qm, double inv_qp, double limit) {
  const uint64_t nf = ctx->nb * 192;
  FL_TRY(fill_i64(ctx, ctx->qm, nf, 0));
  FL_TRY(fill_i64(ctx, ctx->qp, nf, 0));
  if (ctx->nbins == 0 || n == 0) return FL_OK;
  using K = decltype(&k_p2g_compatible<false, true, true>);
  static const K table[8] = {k_p2g_compatible<false, false, false>, k_p2g_compatible<false, false, true>,
                             k_p2g_compatible<false, true, false>,  k_p2g_compatible<false, true, true>,
                             k_p2g_compatible<true, false, false>,  k_p2g_compatible<true, false, true>,
                             k_p2g_compatible<true, true, false>,   k_p2g_compatible<true, true, true>};
  const bool trial = !ctx->sorted.trial_aliases_sorted, pow2 = ctx->geometry.inv_h != 0.0, small = limit <= 0x1p50;
  const CompatTags tags{slot_side, slot_ok, side, affine};
  table[(trial ? 4 : 0) + (pow2 ? 2 : 0) + (small ? 1 : 0)]<<<ctx->nbins, kThreads, 0, ctx->stream>>>(
      ctx->sorted.view(), bin_list(ctx), grid_view(ctx), tags, inv_qm, inv_qp, limit, ctx->perm,
      reinterpret_cast<unsigned long long*>(ctx->qm), reinterpret_cast<unsigned long long*>(ctx->qp), ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int compatible_stencils(fl_ctx* ctx, uint64_t lo, uint64_t count, const int64_t* slot_side, const uint8_t* slot_ok,
                        const int64_t* side, const uint8_t* affine, fl_compatible_sample* d_out) {
  if (!count) return FL_OK;
  const CompatTags tags{slot_side, slot_ok, side, affine};
  auto* kernel = ctx->geometry.inv_h != 0.0 ? k_compatible_stencils<true> : k_compatible_stencils<false>;
  kernel<<<grid_for(count), kThreads, 0, ctx->stream>>>(ctx->raw.view(), ctx->bin, grid_view(ctx), tags, lo, count,
                                                        d_out);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int g2p_boundary(fl_ctx* ctx, uint64_t n, const BoundaryDevice& d, uint32_t* slow_count) {
  *slow_count = 0;
  if (ctx->nbins == 0 || n == 0) return FL_OK;
  BoundaryGatherArgs a;
  a.old_h = HistoryView{d.u[0], d.validity[0], d.side[0], d.eligible[0], d.walls[0]};
  a.new_h = HistoryView{d.u[1], d.validity[1], d.side[1], d.eligible[1], d.walls[1]};
  a.same = d.same;
  a.pside = d.pside;
  a.paffine = d.paffine;
  a.extra = d.extra;
  a.extra_cap = d.extra_cap;
  a.slow = d.slow;
  a.slow_count = d.slow_count;
  FL_HIP(hipMemsetAsync(d.slow_count, 0, sizeof(uint32_t), ctx->stream));
  const bool trial = !ctx->sorted.trial_aliases_sorted, pow2 = ctx->geometry.inv_h != 0.0;
  auto* fast = trial ? (pow2 ? k_g2p_boundary<true, true> : k_g2p_boundary<true, false>)
                     : (pow2 ? k_g2p_boundary<false, true> : k_g2p_boundary<false, false>);
  fast<<<ctx->nbins, kThreads, 0, ctx->stream>>>(ctx->sorted.view(), bin_list(ctx), grid_view(ctx), a, ctx->perm,
                                                 ctx->gather, ctx->sorted.capacity, ctx->err);
  FL_HIP(hipGetLastError());
  FL_HIP(hipMemcpyAsync(slow_count, d.slow_count, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  if (*slow_count) {
    auto* slow = pow2 ? k_g2p_ghost<true> : k_g2p_ghost<false>;
    slow<<<(*slow_count + 63) / 64, 64, 0, ctx->stream>>>(ctx->sorted.view(), nullptr, grid_view(ctx), a, ctx->perm,
                                                         ctx->gather, ctx->sorted.capacity, ctx->err);
    FL_HIP(hipGetLastError());
  }
  return FL_OK;
}

int normalize_forces(fl_ctx* ctx, uint64_t nf, double qm, double qp, double min_mass, const double g[3], double dt) {
  if (!nf) return FL_OK;
  k_normalize_forces<<<grid_for(nf), kThreads, 0, ctx->stream>>>(ctx->qm, ctx->qp, nullptr, nf, qm, qp, min_mass,
                                                                  g[0], g[1], g[2], dt, ctx->face_mass, ctx->u0,
                                                                  ctx->validity, ctx->ustar, ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int apply_forces(fl_ctx* ctx, const double* d_old, uint64_t nf, const double g[3], double dt, double* d_pred) {
  if (!nf) return FL_OK;
  k_apply_forces<<<grid_for(nf), kThreads, 0, ctx->stream>>>(d_old, nf, g[0], g[1], g[2], dt, d_pred, ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int normalize_forces_stress(fl_ctx* ctx, const int64_t* qs, uint64_t nf, double qm, double qp, double min_mass,
                            const double g[3], double dt) {
  return launch_normalize_forces(ctx->stream, ctx->qm, ctx->qp, qs, nf, qm, qp, min_mass, g, dt, ctx->face_mass,
                                 ctx->u0, ctx->validity, ctx->ustar, ctx->err);
}

int launch_normalize_forces(hipStream_t s, const int64_t* qm, const int64_t* qp, const int64_t* qs, uint64_t nf,
                            double sqm, double sqp, double min_mass, const double g[3], double dt, double* mass,
                            double* u0, uint8_t* validity, double* ustar, uint64_t* err) {
  if (!nf) return FL_OK;
  k_normalize_forces<<<grid_for(nf), kThreads, 0, s>>>(qm, qp, qs, nf, sqm, sqp, min_mass, g[0], g[1], g[2], dt, mass,
                                                       u0, validity, ustar, err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int g2p(fl_ctx* ctx, uint64_t n, const double* d_old, const double* d_new, hipEvent_t before, hipEvent_t after) {
  if (before) FL_HIP(hipEventRecord(before, ctx->stream));
  if (after && (ctx->nbins == 0 || n == 0)) FL_HIP(hipEventRecord(after, ctx->stream));
  if (ctx->nbins == 0 || n == 0) return FL_OK;
  const bool trial = !ctx->sorted.trial_aliases_sorted, pow2 = ctx->geometry.inv_h != 0.0;
  auto* kernel = trial ? (pow2 ? k_g2p<true, true> : k_g2p<true, false>)
                       : (pow2 ? k_g2p<false, true> : k_g2p<false, false>);
  kernel<<<ctx->nbins, kThreads, 0, ctx->stream>>>(ctx->sorted.view(), bin_list(ctx), grid_view(ctx), d_old, d_new,
                                                   ctx->perm, ctx->gather, ctx->sorted.capacity, ctx->err);
  FL_HIP(hipGetLastError());
  if (after) FL_HIP(hipEventRecord(after, ctx->stream));
  return FL_OK;
}

}  // namespace fl
