// Device restatements of libstdc++'s std::lerp, std::midpoint (double) and three-argument std::hypot
// (GCC 11 <cmath> / <numeric>), so device results equal the host's bit for bit. Every operation is a
// correctly rounded IEEE operation; kernels are built without contraction.
#pragma once
#include "context.h"

namespace fl {

__device__ inline double lerp_std(double a, double b, double t) {
  if ((a <= 0 && b >= 0) || (a >= 0 && b <= 0)) return t * b + (1 - t) * a;
  if (t == 1) return b;
  const double x = a + t * (b - a);
  return (t > 1) == (b > a) ? (b < x ? x : b) : (b > x ? x : b);
}

__device__ inline double midpoint_std(double a, double b) {
  const double lo = 2.2250738585072014e-308 * 2, hi = 1.7976931348623157e308 / 2;
  const double aa = a < 0 ? -a : a, ab = b < 0 ? -b : b;
  if (aa <= hi && ab <= hi) return (a + b) / 2;
  if (aa < lo) return a + b / 2;
  if (ab < lo) return a / 2 + b;
  return a / 2 + b / 2;
}

// The largest magnitude times the root of the scaled squares.
__device__ inline double hypot3(double x, double y, double z) {
  x = fabs(x);
  y = fabs(y);
  z = fabs(z);
  const double a = x < y ? (y < z ? z : y) : (x < z ? z : x);
  if (a == 0) return 0;
  return a * sqrt((x / a) * (x / a) + (y / a) * (y / a) + (z / a) * (z / a));
}

}  // namespace fl
