// Oracle: bounded projected sums (test infrastructure only), shared by the particle update's contact and
// the frozen geometry's tank clipping: the ProjectedSum of
// native/include/fluidity/base/liquid_boundary_math.h:93-263 (a running binary64 sum whose rounding
// corrections are bounded below and above, with certified signs).
#pragma once

#include <cmath>
#include <limits>
#include <numeric>
#include <utility>

#include "oracle.h"

namespace flo::bounded {

constexpr double kInf = std::numeric_limits<double>::infinity();

// Binary64, round to nearest.
struct Rounded {
  double value, residual;
};
// Sum and exact residual; the larger magnitude first.
inline Rounded AddWithResidual(double left, double right) {
  if (std::abs(left) < std::abs(right)) std::swap(left, right);
  const double sum = left + right;
  if (!std::isfinite(sum)) return {sum, 0};
  return {sum, right - (sum - left)};
}
inline double AddDown(double left, double right) {
  const Rounded s = AddWithResidual(left, right);
  return s.residual < 0 ? std::nextafter(s.value, -kInf) : s.value;
}
inline double AddUp(double left, double right) {
  const Rounded s = AddWithResidual(left, right);
  return s.residual > 0 ? std::nextafter(s.value, kInf) : s.value;
}
struct Bounds {
  double estimate, lower, upper;
  bool finite() const { return std::isfinite(estimate) && std::isfinite(lower) && std::isfinite(upper); }
  bool resolved() const {
    return finite() && upper - lower <= 16 * std::numeric_limits<double>::epsilon() * std::abs(estimate);
  }
  // The certified sign: 1, -1 or 0 in *out, or false when it is not certified.
  bool sign(int* out) const {
    if (!finite()) return false;
    if (lower > 0) *out = 1;
    else if (upper < 0) *out = -1;
    else if (lower == 0 && upper == 0) *out = 0;
    else return false;
    return true;
  }
};
// A running sum with separately bounded rounding corrections (ProjectedSum).
class PSum {
 public:
  void Add(double term) {
    if (!std::isfinite(sum_) || !std::isfinite(term)) {
      sum_ = kInf;
      return;
    }
    const Rounded s = AddWithResidual(sum_, term);
    sum_ = s.value;
    Correct(s.residual);
  }
  void AddDot(flo_vec3 n, flo_vec3 v) {
    Scaled(n.x, v.x, 1);
    Scaled(n.y, v.y, 1);
    Scaled(n.z, v.z, 1);
  }
  void AddRelativeMotion(flo_vec3 n, flo_vec3 v, flo_vec3 wall, double duration) {
    Relative(n.x, v.x, wall.x, duration);
    Relative(n.y, v.y, wall.y, duration);
    Relative(n.z, v.z, wall.z, duration);
  }
  void AddProduct(double left, double right) {
    const double product = left * right;
    Add(product);
    if (std::isfinite(product) && left != 0 && right != 0 && std::abs(left) != 1 && std::abs(right) != 1) {
      Correct(std::fma(left, right, -product));
      if (std::abs(product) <= 0x1p-969) {
        const double r = std::numeric_limits<double>::denorm_min();
        lower_ = AddDown(lower_, -r);
        upper_ = AddUp(upper_, r);
      }
    }
  }
  Bounds bounds() const {
    if (!std::isfinite(sum_) || !std::isfinite(lower_) || !std::isfinite(upper_)) return {kInf, kInf, kInf};
    return {sum_ + std::midpoint(lower_, upper_), AddDown(sum_, lower_), AddUp(sum_, upper_)};
  }

 private:
  void Correct(double c) {
    if (!std::isfinite(lower_) || !std::isfinite(upper_)) {
      sum_ = kInf;
      return;
    }
    lower_ = AddDown(lower_, c);
    upper_ = AddUp(upper_, c);
  }
  void Scaled(double left, double right, double scale) {
    if (left == 0 || right == 0 || scale == 0) return;
    if (scale == 1) {
      AddProduct(left, right);
      return;
    }
    if (std::abs(left) == 1 || std::abs(right) == 1) {
      AddProduct(left * right, scale);
      return;
    }
    const double product = left * right;
    if (!std::isfinite(product) || std::abs(product) <= 0x1p-969) {
      sum_ = kInf;
      return;
    }
    AddProduct(product, scale);
    AddProduct(std::fma(left, right, -product), scale);
  }
  void Relative(double n, double v, double wall, double duration) {
    if (n == 0 || duration == 0) return;
    const Rounded r = AddWithResidual(v, -wall);
    if (!std::isfinite(r.value)) {
      sum_ = kInf;
      return;
    }
    Scaled(n, r.value, duration);
    Scaled(n, r.residual, duration);
  }
  double sum_ = 0, lower_ = 0, upper_ = 0;
};

}  // namespace flo::bounded
