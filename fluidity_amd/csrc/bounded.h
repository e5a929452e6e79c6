// Bounded projected sums on the device (native/include/fluidity/base/liquid_boundary_math.h:93-263): a
// running binary64 sum whose rounding corrections are bounded below and above, with certified signs.
// Shared by the particle update's contact and the frozen geometry's tank clipping.
#pragma once

#include "context.h"
#include "stdmath.h"

namespace fl {
namespace bounded {

constexpr double kBoundedInf = __builtin_inf();

struct Rounded {
  double value, residual;
};
// Sum and exact residual, the larger magnitude first.
__device__ inline Rounded add_residual(double l, double r) {
  if (fabs(l) < fabs(r)) {
    const double t = l;
    l = r;
    r = t;
  }
  const double s = l + r;
  if (!finite(s)) return {s, 0.0};
  return {s, r - (s - l)};
}
__device__ inline double add_down(double l, double r) {
  const Rounded s = add_residual(l, r);
  return s.residual < 0 ? nextafter(s.value, -kBoundedInf) : s.value;
}
__device__ inline double add_up(double l, double r) {
  const Rounded s = add_residual(l, r);
  return s.residual > 0 ? nextafter(s.value, kBoundedInf) : s.value;
}

struct Bounds {
  double estimate, lower, upper;
  __device__ bool is_finite() const { return finite(estimate) && finite(lower) && finite(upper); }
  __device__ bool resolved() const {
    return is_finite() && upper - lower <= 16 * 2.220446049250313e-16 * fabs(estimate);
  }
  // The certified sign into *out, or false when it is not certified.
  __device__ bool sign(int* out) const {
    if (!is_finite()) return false;
    if (lower > 0) *out = 1;
    else if (upper < 0) *out = -1;
    else if (lower == 0 && upper == 0) *out = 0;
    else return false;
    return true;
  }
};

// A running sum with separately bounded rounding corrections.
struct PSum {
  double sum = 0, lo = 0, hi = 0;
  __device__ void correct(double c) {
    if (!finite(lo) || !finite(hi)) {
      sum = kBoundedInf;
      return;
    }
    lo = add_down(lo, c);
    hi = add_up(hi, c);
  }
  __device__ void add(double term) {
    if (!finite(sum) || !finite(term)) {
      sum = kBoundedInf;
      return;
    }
    const Rounded s = add_residual(sum, term);
    sum = s.value;
    correct(s.residual);
  }
  __device__ void product(double l, double r) {
    const double p = l * r;
    add(p);
    if (finite(p) && l != 0 && r != 0 && fabs(l) != 1 && fabs(r) != 1) {
      correct(fma(l, r, -p));
      if (fabs(p) <= 0x1p-969) {
        const double d = 4.9406564584124654e-324;
        lo = add_down(lo, -d);
        hi = add_up(hi, d);
      }
    }
  }
  __device__ void scaled(double l, double r, double scale) {
    if (l == 0 || r == 0 || scale == 0) return;
    if (scale == 1) {
      product(l, r);
      return;
    }
    if (fabs(l) == 1 || fabs(r) == 1) {
      product(l * r, scale);
      return;
    }
    const double p = l * r;
    if (!finite(p) || fabs(p) <= 0x1p-969) {
      sum = kBoundedInf;
      return;
    }
    product(p, scale);
    product(fma(l, r, -p), scale);
  }
  __device__ void relative(double n, double v, double wall, double duration) {
    if (n == 0 || duration == 0) return;
    const Rounded r = add_residual(v, -wall);
    if (!finite(r.value)) {
      sum = kBoundedInf;
      return;
    }
    scaled(n, r.value, duration);
    scaled(n, r.residual, duration);
  }
  __device__ void add_dot(fl_vec3 n, fl_vec3 v) {
    scaled(n.x, v.x, 1);
    scaled(n.y, v.y, 1);
    scaled(n.z, v.z, 1);
  }
  __device__ void add_motion(fl_vec3 n, fl_vec3 v, fl_vec3 wall, double duration) {
    relative(n.x, v.x, wall.x, duration);
    relative(n.y, v.y, wall.y, duration);
    relative(n.z, v.z, wall.z, duration);
  }
  __device__ Bounds bounds() const {
    if (!finite(sum) || !finite(lo) || !finite(hi)) return {kBoundedInf, kBoundedInf, kBoundedInf};
    return {sum + midpoint_std(lo, hi), add_down(sum, lo), add_up(sum, hi)};
  }
};

}  // namespace bounded
}  // namespace fl
