// Device lattice arithmetic of the sparse staggered MAC grid, bit-exact with the reference.
//
// Follows native/src/sparse_mac_grid_3d.cc: compensated coordinate normalisation (:83-184),
// floor and bin rules (:50-56, :186-220), half-cell motion (:222-234), checked quadratic axis
// stencils with exact support-edge weights (:236-310), and the scaled weight gradient (:312-336).
// Every operation is a correctly rounded IEEE FP64 operation and the library is compiled with
// -ffp-contract=off, so a bin, weight or offset computed here equals the host value bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fl {

enum Status : int {
  kOk = 0,
  kInvalidArgument = 3,
  kResourceExhausted = 8,
  kFailedPrecondition = 9,
  kAborted = 10,
  kOutOfRange = 11,
  kInternal = 13,
};

constexpr int64_t kI32Min = -2147483648LL;
constexpr int64_t kI32Max = 2147483647LL;
constexpr double kEps = 2.220446049250313080847e-16;

__device__ __host__ inline bool finite(double v) { return __builtin_isfinite(v); }
__device__ __host__ inline bool normal(double v) { return __builtin_isnormal(v); }

__device__ __host__ inline int64_t floor_div4(int64_t i) { return i >= 0 ? i / 4 : -1 - (-(i + 1)) / 4; }

__device__ inline double sub_roundoff(double a, double b, double d) {
  if (fabs(a) >= fabs(b)) return (a - d) - b;
  return a - (d + b);
}

// One coordinate in cell units with the world subtraction kept for exact edge decisions.
// kPow2: the cell width is a power of two and invH is its exact reciprocal, so s / H and s * invH
// are the same correctly rounded value; otherwise every quotient is a division.
template <bool kPow2>
struct CoordT {
  double X, D, E, H, invH;
  int shift;

  __device__ double div_h(double s) const {
    if constexpr (kPow2) return s * invH;
    return s / H;
  }
  __device__ double scaled_offset(double edge) const { return fma(-edge, H, D) + E; }
  __device__ bool below(double edge) const {
    const double off = X - edge;
    const double bound = fabs(X) * (2.0 * kEps);
    if (fabs(off) > bound) return off < 0.0;
    return scaled_offset(edge) < 0.0;
  }
  __device__ double offset(double edge) const {
    const double off = X - edge;
    if (fabs(off) >= fabs(X) && off != 0.0) return off;
    const double s = scaled_offset(edge);
    return finite(s) ? div_h(s) : off;
  }
  __device__ double displacement(double anchor, double h) const {
    const double n = -offset(anchor);
    const double d = n * h;
    if (normal(n) && finite(d)) return d;
    const double s = scaled_offset(anchor);
    return s == 0.0 ? 0.0 : -ldexp(s, -shift);
  }

  // Interior coordinates: 4 < |X| < 2^30 and |D| < 2^1000. Every edge and anchor of a transfer stencil
  // lies within 1.5 cells of X, so X - edge is smaller than X in magnitude and offset() takes its
  // compensated branch, whose scaled subtraction cannot overflow (|edge H| is about |D|); and X is far
  // inside the int32 lattice, so the support range checks pass. The *_interior forms below skip
  // exactly the decisions this settles and return the same values.
  __device__ bool interior() const {
    const double a = fabs(X);
    return a > 4.0 && a < 0x1p30 && fabs(D) < 0x1p1000;
  }
  __device__ double offset_interior(double edge) const { return div_h(scaled_offset(edge)); }
  __device__ double displacement_interior(double anchor, double h) const {
    const double n = -offset_interior(anchor);
    const double d = n * h;
    if (normal(n) && finite(d)) return d;
    const double s = scaled_offset(anchor);
    return s == 0.0 ? 0.0 : -ldexp(s, -shift);
  }
};

// inv_h: exact 1/h when kPow2 (h a power of two), unused otherwise.
template <bool kPow2>
__device__ inline CoordT<kPow2> normalize(double x, double o, double h, double inv_h) {
  double diff = x - o;
  int e = 0;
  if (!finite(diff)) {
    x *= 0.5;
    o *= 0.5;
    diff = x - o;
    e = -1;
  }
  double H = ldexp(h, e);
  double invH = kPow2 ? ldexp(inv_h, -e) : 0.0;
  const double X = kPow2 ? diff * invH : diff / H;
  double err = sub_roundoff(x, o, diff);
  if (finite(X) && H < 0x1p-960) {
    e += 64;
    diff = ldexp(diff, 64);
    err = ldexp(err, 64);
    H = ldexp(H, 64);
    invH = kPow2 ? ldexp(invH, -64) : 0.0;
  }
  CoordT<kPow2> c;
  c.X = X;
  c.D = diff;
  c.E = err;
  c.H = H;
  c.invH = invH;
  c.shift = e;
  return c;
}

// floor(true coordinate - shift) as an exact integer-valued double (|X| < 2^52 on every caller's path).
template <class Coord>
__device__ inline double floor_shifted_d(const Coord& c, double shift) {
  const double k = floor(c.X - shift);
  const double edge = k + shift;
  if (c.below(edge)) return k - 1.0;
  if (c.below(edge + 1.0)) return k;
  return k + 1.0;
}

template <class Coord>
__device__ inline int64_t floor_shifted(const Coord& c, double shift) {
  const int64_t k = static_cast<int64_t>(floor(c.X - shift));
  const double edge = static_cast<double>(k) + shift;
  if (c.below(edge)) return k - 1;
  if (c.below(edge + 1.0)) return k;
  return k + 1;
}

// Sort bin of one coordinate, kOutOfRange outside the int32 block lattice.
template <class Coord>
__device__ inline int checked_bin(const Coord& c, int32_t& bin) {
  const double first_edge = static_cast<double>(kI32Min * 4) + 1.5;
  const double past_last_edge = static_cast<double>((kI32Max + 1) * 4) + 1.5;
  if (c.below(first_edge) || !c.below(past_last_edge)) return kOutOfRange;
  bin = static_cast<int32_t>(floor_div4(floor_shifted(c, 1.5)));
  return kOk;
}

__device__ inline bool within_half_cell(double s, double t, double h) {
  const double hi = fmax(s, t);
  const double lo = fmin(s, t);
  const double d = hi - lo;
  const double dd = 2.0 * d;
  if (dd != h) return dd < h;
  return sub_roundoff(hi, lo, d) <= 0.0;
}

// Weight factor plus the support-edge distance that defines an outer weight.
struct Factor {
  double r, edge;
  __device__ bool zero() const { return r == 0.0 && edge == 0.0; }
  __device__ double mantissa(int& e) const {
    if (normal(r) || edge == 0.0) return frexp(r, &e);
    int de;
    const double dm = frexp(edge, &de);
    e = 2 * de;
    return 0.5 * dm * dm;
  }
};

// The three samples of one coordinate axis.
struct AxisStencil {
  int32_t first;     // lattice index of sample 0
  Factor w[3];
  double offset[3];  // face anchor minus particle, meters
  double dw[3];      // derivative in cell units
};

// kOutOfRange for unrepresentable lattice indices, kFailedPrecondition when the support leaves the
// sorted bin's transfer tile [4B, 4B+8).
template <class Coord>
__device__ inline int axis_stencil(const Coord& c, double lattice_shift, int32_t bin, double h, AxisStencil& s) {
  const double start = lattice_shift + 0.5;
  const double first_support = static_cast<double>(kI32Min) + start;
  const double past_last_support = static_cast<double>(kI32Max - 1) + start;
  if (c.below(first_support) || !c.below(past_last_support)) return kOutOfRange;
  const int64_t first = floor_shifted(c, start);
  const int64_t tile = static_cast<int64_t>(bin) * 4;
  if (first < tile || first + 2 >= tile + 8) return kFailedPrecondition;
  const double a0 = static_cast<double>(first) + lattice_shift;
  const double dl = c.offset(a0 + 0.5);
  const double dr = fabs(c.offset(a0 + 1.5));
  const double cc = c.offset(a0 + 1.0);
  s.first = static_cast<int32_t>(first);
  s.w[0] = {0.5 * dr * dr, dr};
  s.w[1] = {0.75 - cc * cc, 0.0};
  s.w[2] = {0.5 * dl * dl, dl};
  s.dw[0] = -dr;
  s.dw[1] = -2.0 * cc;
  s.dw[2] = dl;
  for (int i = 0; i < 3; ++i) s.offset[i] = c.displacement(static_cast<double>(first + i) + lattice_shift, h);
  return kOk;
}

__device__ inline double physical_gradient(Factor fx, Factor fy, Factor fz, double h) {
  const double g = fx.r * fy.r * fz.r;
  if (normal(g) || fx.zero() || fy.zero() || fz.zero()) return g / h;
  int ex, ey, ez, eh;
  const double mx = fx.mantissa(ex);
  const double my = fy.mantissa(ey);
  const double mz = fz.mantissa(ez);
  const double mh = frexp(h, &eh);
  return ldexp((mx * my * mz) / mh, ex + ey + ez - eh);
}

// Validated geometry scalars shared by all kernels.
struct Geometry {
  double ox, oy, oz, h;
  double inv_h;  // exact 1/h when h is a power of two, else 0
};

// Kernels are instantiated for power-of-two cell widths (exact reciprocal) and for general widths;
// the host picks the instance from Geometry::inv_h.

__device__ inline double comp(const Geometry& g, int a) { return a == 0 ? g.ox : (a == 1 ? g.oy : g.oz); }

// Error word ordered by particle ordinal, then by status code.
__device__ __host__ inline uint64_t error_word(uint64_t ordinal, int code) {
  return (ordinal << 8) | static_cast<uint64_t>(code & 0xff);
}

}  // namespace fl
