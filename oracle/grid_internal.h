// Oracle internals: the reference's exact lattice arithmetic (test infrastructure only).
//
// Restates native/src/sparse_mac_grid_3d.cc: compensated coordinate normalisation
// (:83-184), floor/bin rules (:50-56, :186-220), half-cell motion (:222-234), the checked
// quadratic axis stencil (:236-310), the scaled weight gradient (:312-336), the owned slot
// map (:388-409, :587-603) and the 27-sample geometry stencil (:451-548, :627-671).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "oracle.h"

namespace flo {

constexpr int64_t kBlock = 4;
constexpr int64_t kTile = 8;
constexpr int64_t kFacesPerBlock = 192;
constexpr int64_t kI32Min = std::numeric_limits<int32_t>::min();
constexpr int64_t kI32Max = std::numeric_limits<int32_t>::max();
constexpr int64_t kMinBlock = kI32Min / kBlock;
constexpr int64_t kMaxBlock = kI32Max / kBlock;

inline bool Finite3(flo_vec3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
inline double Get(flo_vec3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
inline int32_t Get(flo_int3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
inline void Set(flo_vec3& v, int a, double s) { (a == 0 ? v.x : (a == 1 ? v.y : v.z)) = s; }
inline flo_vec3 Row(const flo_mat3& m, int a) { return a == 0 ? m.x_row : (a == 1 ? m.y_row : m.z_row); }
inline void SetRow(flo_mat3& m, int a, flo_vec3 r) { (a == 0 ? m.x_row : (a == 1 ? m.y_row : m.z_row)) = r; }

// sparse_mac_grid_3d.cc:32-41
inline int ValidateGeometry(const flo_geometry& g) {
  if (!Finite3(g.world_origin)) return FLO_INVALID_ARGUMENT;
  if (!(g.cell_width > 0.0 && std::isnormal(g.cell_width))) return FLO_INVALID_ARGUMENT;
  return FLO_OK;
}

// Mathematical floor division by the block width (sparse_mac_grid_3d.cc:50-56).
inline int64_t FloorDiv4(int64_t i) { return i >= 0 ? i / kBlock : -1 - (-(i + 1)) / kBlock; }

// FastTwoSum residual of a rounded subtraction (sparse_mac_grid_3d.cc:85-91).
inline double SubRoundoff(double a, double b, double d) {
  if (std::fabs(a) >= std::fabs(b)) return (a - d) - b;
  return a - (d + b);
}

// One coordinate in cell units plus the retained world subtraction
// (sparse_mac_grid_3d.cc:95-165).
struct Coord {
  double X = 0;      // rounded (x - o)/h
  double D = 0;      // scaled world difference
  double E = 0;      // its subtraction roundoff
  double H = 1;      // scaled cell width
  int shift = 0;     // world scale exponent

  double ScaledOffset(double edge) const { return std::fma(-edge, H, D) + E; }
  bool Below(double edge) const {
    const double off = X - edge;
    const double bound = std::fabs(X) * (2.0 * std::numeric_limits<double>::epsilon());
    if (std::fabs(off) > bound) return off < 0.0;
    return ScaledOffset(edge) < 0.0;
  }
  double Offset(double edge) const {
    const double off = X - edge;
    if (std::fabs(off) >= std::fabs(X) && off != 0.0) return off;
    const double s = ScaledOffset(edge);
    return std::isfinite(s) ? s / H : off;
  }
  double Displacement(double anchor, double h) const {
    const double n = -Offset(anchor);
    const double d = n * h;
    if (std::isnormal(n) && std::isfinite(d)) return d;
    const double s = ScaledOffset(anchor);
    return s == 0.0 ? 0.0 : -std::scalbn(s, -shift);
  }
};

inline Coord Normalize(double x, double o, double h) {
  Coord c;
  double diff = x - o;
  int e = 0;
  if (!std::isfinite(diff)) {
    x *= 0.5;
    o *= 0.5;
    diff = x - o;
    e = -1;
  }
  double H = std::scalbn(h, e);
  const double X = diff / H;
  double err = SubRoundoff(x, o, diff);
  if (std::isfinite(X) && H < 0x1p-960) {
    e += 64;
    diff = std::scalbn(diff, 64);
    err = std::scalbn(err, 64);
    H = std::scalbn(H, 64);
  }
  c.X = X;
  c.D = diff;
  c.E = err;
  c.H = H;
  c.shift = e;
  return c;
}

// sparse_mac_grid_3d.cc:169-184
inline int NormalizePosition(const flo_geometry& g, flo_vec3 p, std::array<Coord, 3>& out) {
  if (!Finite3(p)) return FLO_INVALID_ARGUMENT;
  for (int a = 0; a < 3; ++a) {
    out[a] = Normalize(Get(p, a), Get(g.world_origin, a), g.cell_width);
    if (!std::isfinite(out[a].X)) return FLO_OUT_OF_RANGE;
  }
  return FLO_OK;
}

// sparse_mac_grid_3d.cc:187-198
inline int64_t FloorShifted(const Coord& c, double shift) {
  const auto k = static_cast<int64_t>(std::floor(c.X - shift));
  const double edge = static_cast<double>(k) + shift;
  if (c.Below(edge)) return k - 1;
  if (c.Below(edge + 1.0)) return k;
  return k + 1;
}

// sparse_mac_grid_3d.cc:200-220
inline int CheckedBin(const std::array<Coord, 3>& p, flo_int3& bin) {
  constexpr double first_edge = static_cast<double>(kI32Min * kBlock) + 1.5;
  constexpr double past_last_edge = static_cast<double>((kI32Max + 1) * kBlock) + 1.5;
  int32_t b[3];
  for (int a = 0; a < 3; ++a) {
    if (p[a].Below(first_edge) || !p[a].Below(past_last_edge)) return FLO_OUT_OF_RANGE;
    b[a] = static_cast<int32_t>(FloorDiv4(FloorShifted(p[a], 1.5)));
  }
  bin = {b[0], b[1], b[2]};
  return FLO_OK;
}

// sparse_mac_grid_3d.cc:222-234
inline bool WithinHalfCell(double s, double t, double h) {
  const double hi = std::max(s, t);
  const double lo = std::min(s, t);
  const double d = hi - lo;
  const double dd = 2.0 * d;
  if (dd != h) return dd < h;
  return SubRoundoff(hi, lo, d) <= 0.0;
}

// A weight factor and, for outer weights, the edge distance that defines it
// (sparse_mac_grid_3d.cc:236-254).
struct Factor {
  double r = 0;
  double edge = 0;
  bool Zero() const { return r == 0.0 && edge == 0.0; }
  double Mantissa(int& e) const {
    if (std::isnormal(r) || edge == 0.0) return std::frexp(r, &e);
    int de;
    const double dm = std::frexp(edge, &de);
    e = 2 * de;
    return 0.5 * dm * dm;
  }
};

struct AxisSample {
  int32_t lattice;
  Factor w;
  double offset;
  double dw;  // derivative in normalised coordinates
};

// sparse_mac_grid_3d.cc:266-310
inline int AxisStencil(const Coord& c, double lattice_shift, int32_t bin, double h, std::array<AxisSample, 3>& out) {
  const double start = lattice_shift + 0.5;
  const double first_support = static_cast<double>(kI32Min) + start;
  const double past_last_support = static_cast<double>(kI32Max - 1) + start;
  if (c.Below(first_support) || !c.Below(past_last_support)) return FLO_OUT_OF_RANGE;
  const int64_t first = FloorShifted(c, start);
  const int64_t tile_start = static_cast<int64_t>(bin) * kBlock;
  if (first < tile_start || first + 2 >= tile_start + kTile) return FLO_FAILED_PRECONDITION;
  const double a0 = static_cast<double>(first) + lattice_shift;
  const double dl = c.Offset(a0 + 0.5);
  const double dr = std::fabs(c.Offset(a0 + 1.5));
  const double cc = c.Offset(a0 + 1.0);
  const Factor w[3] = {{0.5 * dr * dr, dr}, {0.75 - cc * cc, 0.0}, {0.5 * dl * dl, dl}};
  const double dw[3] = {-dr, -2.0 * cc, dl};
  for (int i = 0; i < 3; ++i) {
    const int64_t li = first + i;
    const double anchor = static_cast<double>(li) + lattice_shift;
    out[i] = {static_cast<int32_t>(li), w[i], c.Displacement(anchor, h), dw[i]};
  }
  return FLO_OK;
}

// sparse_mac_grid_3d.cc:312-336
inline double PhysicalGradient(Factor fx, Factor fy, Factor fz, double h) {
  const double g = fx.r * fy.r * fz.r;
  if (std::isnormal(g) || fx.Zero() || fy.Zero() || fz.Zero()) return g / h;
  int ex, ey, ez, eh;
  const double mx = fx.Mantissa(ex);
  const double my = fy.Mantissa(ey);
  const double mz = fz.Mantissa(ez);
  const double mh = std::frexp(h, &eh);
  return std::scalbn((mx * my * mz) / mh, ex + ey + ez - eh);
}

struct GeometrySample {
  flo_int3 lattice;
  double w;
  flo_vec3 r;
  flo_vec3 grad;
};

// sparse_mac_grid_3d.cc:451-548
inline int GeometryStencil(const flo_geometry& g, int axis, flo_vec3 sorted, flo_vec3 trial,
                           std::array<GeometrySample, 27>& out) {
  if (axis < 0 || axis > 2) return FLO_INVALID_ARGUMENT;
  if (int s = ValidateGeometry(g)) return s;
  std::array<Coord, 3> sp, tp;
  if (int s = NormalizePosition(g, sorted, sp)) return s;
  if (int s = NormalizePosition(g, trial, tp)) return s;
  flo_int3 bin;
  if (int s = CheckedBin(sp, bin)) return s;
  const double h = g.cell_width;
  if (!(WithinHalfCell(sorted.x, trial.x, h) && WithinHalfCell(sorted.y, trial.y, h) &&
        WithinHalfCell(sorted.z, trial.z, h)))
    return FLO_FAILED_PRECONDITION;
  std::array<std::array<AxisSample, 3>, 3> ax;
  for (int a = 0; a < 3; ++a) {
    const double shift = a == axis ? 0.0 : 0.5;
    if (int s = AxisStencil(tp[a], shift, Get(bin, a), h, ax[a])) return s;
  }
  int k = 0;
  for (const auto& sz : ax[2])
    for (const auto& sy : ax[1])
      for (const auto& sx : ax[0]) {
        GeometrySample& o = out[k++];
        o.lattice = {sx.lattice, sy.lattice, sz.lattice};
        o.r = {sx.offset, sy.offset, sz.offset};
        o.grad = {PhysicalGradient({sx.dw, 0.0}, sy.w, sz.w, h), PhysicalGradient(sx.w, {sy.dw, 0.0}, sz.w, h),
                  PhysicalGradient(sx.w, sy.w, {sz.dw, 0.0}, h)};
        if (!Finite3(o.r) || !Finite3(o.grad)) return FLO_OUT_OF_RANGE;
        o.w = sx.w.r * sy.w.r * sz.w.r;
      }
  return FLO_OK;
}

inline bool BlockLess(const flo_int3& a, const flo_int3& b) {
  if (a.z != b.z) return a.z < b.z;
  if (a.y != b.y) return a.y < b.y;
  return a.x < b.x;
}
inline bool BlockEq(const flo_int3& a, const flo_int3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

inline flo_int3 OwnerBlock(flo_int3 lattice) {
  return {static_cast<int32_t>(FloorDiv4(lattice.x)), static_cast<int32_t>(FloorDiv4(lattice.y)),
          static_cast<int32_t>(FloorDiv4(lattice.z))};
}

// Owned slot of a face on canonical sorted blocks (sparse_mac_grid_3d.cc:388-409).
inline int FaceSlot(const flo_int3* blocks, size_t count, int axis, flo_int3 lattice, uint64_t& slot) {
  if (axis < 0 || axis > 2) return FLO_INVALID_ARGUMENT;
  const flo_int3 owner = OwnerBlock(lattice);
  const flo_int3* end = blocks + count;
  const flo_int3* it = std::lower_bound(blocks, end, owner, BlockLess);
  if (it == end || !BlockEq(*it, owner)) return FLO_FAILED_PRECONDITION;
  const int64_t lx = lattice.x - kBlock * owner.x;
  const int64_t ly = lattice.y - kBlock * owner.y;
  const int64_t lz = lattice.z - kBlock * owner.z;
  slot = static_cast<uint64_t>(it - blocks) * kFacesPerBlock + static_cast<uint64_t>(axis) * 64 +
         static_cast<uint64_t>((lz * kBlock + ly) * kBlock + lx);
  return FLO_OK;
}

// sparse_mac_grid_3d.cc:627-657
inline int OwnedStencil(const flo_geometry& g, const flo_int3* blocks, size_t count, int axis, flo_vec3 sorted,
                        flo_vec3 trial, std::array<flo_stencil_sample, 27>& out) {
  std::array<GeometrySample, 27> geo;
  if (int s = GeometryStencil(g, axis, sorted, trial, geo)) return s;
  for (int i = 0; i < 27; ++i) {
    uint64_t slot;
    if (int s = FaceSlot(blocks, count, axis, geo[i].lattice, slot)) return s;
    out[i] = {slot, geo[i].w, geo[i].r, geo[i].grad};
  }
  return FLO_OK;
}

// One velocity component of GatherMac3D on a resolved stencil (mac_gather_3d_internal.h:15-69): new
// PIC value, grid increment, the complete-second-moment slope row and the true gradient row.
inline int GatherComponent(const std::array<flo_stencil_sample, 27>& st, double h, const double* old_faces,
                           const double* new_faces, double& new_pic, double& inc, flo_vec3& slope, flo_vec3& grad) {
  double old_pic = 0.0;
  new_pic = 0.0;
  flo_vec3 moment{0, 0, 0};
  grad = {0, 0, 0};
  for (const auto& smp : st) {
    const double uo = old_faces[smp.face_slot];
    const double un = new_faces[smp.face_slot];
    const double wun = smp.weight * un;
    old_pic += smp.weight * uo;
    new_pic += wun;
    const flo_vec3 rn{smp.face_offset_from_particle.x / h, smp.face_offset_from_particle.y / h,
                      smp.face_offset_from_particle.z / h};
    moment.x += rn.x * wun;
    moment.y += rn.y * wun;
    moment.z += rn.z * wun;
    grad.x += smp.weight_gradient.x * un;
    grad.y += smp.weight_gradient.y * un;
    grad.z += smp.weight_gradient.z * un;
  }
  slope = {(4.0 * moment.x) / h, (4.0 * moment.y) / h, (4.0 * moment.z) / h};
  inc = new_pic - old_pic;
  if (!std::isfinite(new_pic) || !std::isfinite(inc) || !Finite3(slope) || !Finite3(grad)) return FLO_OUT_OF_RANGE;
  return FLO_OK;
}

inline int ComputeBin(const flo_geometry& g, flo_vec3 p, flo_int3& bin) {
  if (int s = ValidateGeometry(g)) return s;
  std::array<Coord, 3> np;
  if (int s = NormalizePosition(g, p, np)) return s;
  return CheckedBin(np, bin);
}

}  // namespace flo
