// Frozen boundary geometry on the device (FreezeCore, native/src/liquid_boundary_geometry_3d.cc:765-1068)
// for the absent, half-space, slab and open-tank obstacles: per candidate cell the liquid volume on each side of
// the obstacle from six body-diagonal tetrahedra clipped by the piecewise-linear liquid distance and
// the obstacle planes; per incident face the open and blocked area of its two triangles; the centred
// curvature of cells on an open liquid-air face (base/liquid_boundary_geometry_math_3d.h:160-417,
// 625-848). The open tank clips the liquid by its axis-aligned walls located in the lattice frame with
// bounded sums; its faces carry the aperture of the open walls and the piston's flux on the low-x face
// (:420-623).
//
// With a sampled solid (base/sampled_solid_liquid_geometry_3d.h) the cell query measures the liquid outside
// the piecewise-linear solid of the corner samples (each tetrahedron clipped by the liquid plane and by
// the affine extension of its solid samples, optionally by an analytic half-space), checks that the wet
// pieces are connected, and the face query measures the open area and integrates the corner wall
// velocities over the blocked part of the face.
//
// Each sequential loop of the reference is one kernel over its records; a failure writes
// min(phase, ordinal, code) into the context's error word. Phases: 0 centre samples, 1 corner samples
// (then the solid samples), 2 candidate cells, 3 faces, 4 curvature needs, 5 curvature. Candidate cells that no plane cuts are
// settled by a light staging kernel; only cut cells are clipped. Kept cells are compacted in candidate
// order by a scan; their faces are made unique and sorted by slot through a bitmap over the grid's face
// slots.
//
// Clipping keeps fixed-capacity polygons in registers and private memory. A liquid plane and one
// obstacle plane cut a tetrahedron into at most six faces of at most five vertices, with at most ten
// crossings before duplicates are removed; the capacities below hold these with room to spare, and an
// overflow gives a NaN measure (OutOfRange), as the reference's bounded lists do.
#include "bounded.h"
#include "keys.h"
#include "stdmath.h"

namespace fl {
namespace {

constexpr int kVerts = 8;     // vertices per polygon
constexpr int kCross = 16;    // crossings collected by one polyhedron cut
constexpr int kFaces = 8;     // polygons per polyhedron
constexpr double kNaN = __builtin_nan("");

__device__ inline bool normal_or_zero(double v) { return v == 0.0 || (finite(v) && fabs(v) >= 2.2250738585072014e-308); }
// std::max / std::min of two values: the first unless the second compares greater (smaller).
__device__ inline double max_std(double a, double b) { return a < b ? b : a; }
__device__ inline double min_std(double a, double b) { return b < a ? b : a; }
// std::clamp: lo if v < lo, else hi if hi < v, else v.
__device__ inline double clamp_std(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

struct V3 {
  double c[3];
};
__device__ inline V3 add(V3 a, V3 b) { return {{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}}; }
__device__ inline V3 sub(V3 a, V3 b) { return {{a.c[0] - b.c[0], a.c[1] - b.c[1], a.c[2] - b.c[2]}}; }
__device__ inline V3 scale(V3 a, double f) { return {{a.c[0] * f, a.c[1] * f, a.c[2] * f}}; }
__device__ inline double dot(V3 a, V3 b) { return a.c[0] * b.c[0] + a.c[1] * b.c[1] + a.c[2] * b.c[2]; }
__device__ inline V3 cross(V3 a, V3 b) {
  return {{a.c[1] * b.c[2] - a.c[2] * b.c[1], a.c[2] * b.c[0] - a.c[0] * b.c[2], a.c[0] * b.c[1] - a.c[1] * b.c[0]}};
}
__device__ inline bool finite3(V3 a) { return finite(a.c[0]) && finite(a.c[1]) && finite(a.c[2]); }
__device__ inline bool same(V3 a, V3 b) { return a.c[0] == b.c[0] && a.c[1] == b.c[1] && a.c[2] == b.c[2]; }

struct Vtx {
  V3 p;        // cell-width coordinates
  double phi;  // interpolated liquid distance
};
template <int N>
struct Poly {
  Vtx v[N];
  int n;
  bool overflow;
  __device__ void clear() {
    n = 0;
    overflow = false;
  }
  __device__ void push(const Vtx& x) {
    if (n == N) {
      overflow = true;
      return;
    }
    v[n++] = x;
  }
};
using Polygon = Poly<kVerts>;
struct Solid {
  Polygon f[kFaces];
  int n;
  bool overflow;
};
struct Plane {
  V3 normal;
  double coordinate;
  bool liquid;
  int snap = -1;  // an axis-aligned cut pins its crossings' coordinate on this axis to the plane's
};

__device__ inline double distance(const Vtx& x, const Plane& p) {
  return p.liquid ? x.phi : dot(p.normal, x.p) - p.coordinate;
}

// ClipPolygon: keeps distance <= 0; each crossing takes both weights computed directly.
template <int N>
__device__ void clip(const Polygon& in, const Plane& plane, Polygon& out, Poly<N>* crossings) {
  out.clear();
  if (in.overflow) {
    out.overflow = true;
    return;
  }
  if (in.n == 0) return;
  Vtx prev = in.v[in.n - 1];
  double dp = distance(prev, plane);
  for (int i = 0; i < in.n; ++i) {
    const Vtx cur = in.v[i];
    const double dc = distance(cur, plane);
    if ((dp <= 0) != (dc <= 0)) {
      const double m = max_std(fabs(dp), fabs(dc));
      const double ps = dp / m, cs = dc / m;
      const double diff = ps - cs;
      const double wp = -cs / diff, wc = ps / diff;
      Vtx x{add(scale(prev.p, wp), scale(cur.p, wc)), wp * prev.phi + wc * cur.phi};
      if (plane.snap >= 0) x.p.c[plane.snap] = plane.coordinate * plane.normal.c[plane.snap];
      out.push(x);
      if (crossings) crossings->push(x);
    }
    if (dc <= 0) out.push(cur);
    prev = cur;
    dp = dc;
  }
}

// UnitDirection: binary-exponent normalisation, then division by the length.
__device__ inline V3 unit_direction(V3 d) {
  const double largest = max_std(max_std(fabs(d.c[0]), fabs(d.c[1])), fabs(d.c[2]));
  if (largest == 0) return {{0, 0, 0}};
  int e = 0;
  (void)frexp(largest, &e);
  for (double& c : d.c) c = ldexp(c, -e);
  return scale(d, 1 / hypot3(d.c[0], d.c[1], d.c[2]));
}

// ClipPolyhedron: clips every face and closes the cut with the distinct crossings ordered by angle
// about their centroid (insertion sort on atan2, as the reference).
__device__ void clip_solid(const Solid& in, const Plane& plane, Solid& out) {
  out.n = 0;
  out.overflow = false;
  if (in.overflow) {
    out.overflow = true;
    return;
  }
  Poly<kCross> crossings;
  crossings.clear();
  for (int k = 0; k < in.n; ++k) {
    Polygon tmp;
    clip(in.f[k], plane, tmp, &crossings);
    if (tmp.overflow || crossings.overflow) {
      out.overflow = true;
      return;
    }
    if (tmp.n >= 3) {
      if (out.n == kFaces) out.overflow = true;
      else out.f[out.n++] = tmp;
    }
  }
  Polygon cap;
  cap.clear();
  for (int k = 0; k < crossings.n; ++k) {
    bool seen = false;
    for (int j = 0; j < cap.n; ++j) seen = seen || same(cap.v[j].p, crossings.v[k].p);
    if (!seen) cap.push(crossings.v[k]);
  }
  if (cap.n < 3) return;
  V3 center{{0, 0, 0}};
  for (int k = 0; k < cap.n; ++k) center = add(center, scale(cap.v[k].p, 1.0 / cap.n));
  const V3 tangent = unit_direction(sub(cap.v[0].p, center));
  V3 normal{{0, 0, 0}};
  for (int k = 1; k < cap.n; ++k) {
    normal = unit_direction(cross(tangent, unit_direction(sub(cap.v[k].p, center))));
    if (hypot3(normal.c[0], normal.c[1], normal.c[2]) > 0) break;
  }
  const V3 bitangent = cross(normal, tangent);
  const double tl = hypot3(tangent.c[0], tangent.c[1], tangent.c[2]);
  const double bl = hypot3(bitangent.c[0], bitangent.c[1], bitangent.c[2]);
  if (!(tl > 0 && bl > 0)) return;
  const V3 ut = scale(tangent, 1 / tl), ub = scale(bitangent, 1 / bl);
  double angle[kVerts];
  for (int k = 0; k < cap.n; ++k) {
    const V3 d = sub(cap.v[k].p, center);
    angle[k] = atan2(dot(d, ub), dot(d, ut));
  }
  for (int k = 1; k < cap.n; ++k) {
    const Vtx x = cap.v[k];
    const double a = angle[k];
    int j = k;
    while (j > 0 && a < angle[j - 1]) {
      cap.v[j] = cap.v[j - 1];
      angle[j] = angle[j - 1];
      --j;
    }
    cap.v[j] = x;
    angle[j] = a;
  }
  if (out.n == kFaces) {
    out.overflow = true;
    return;
  }
  out.f[out.n++] = cap;
}

// NormalizeCoordinateRow over n edges.
template <int N>
__device__ inline int normalize_row(V3 (&e)[N], int axis) {
  double largest = 0;
  for (int k = 0; k < N; ++k) largest = max_std(largest, fabs(e[k].c[axis]));
  if (largest == 0) return 0;
  int x = 0;
  (void)frexp(largest, &x);
  for (int k = 0; k < N; ++k) e[k].c[axis] = ldexp(e[k].c[axis], -x);
  return x;
}

// PhysicalMeasure: restores the local exponent and h^dimension; NaN for an unrepresentable nonzero.
__device__ inline double physical_measure(double det, int local_exponent, double h, int dimension) {
  if (det == 0) return 0;
  int we = 0;
  const double wm = frexp(h, &we);
  double m = det;
  for (int a = 0; a < dimension; ++a) m *= wm;
  const double r = ldexp(m, local_exponent + dimension * we);
  if (!finite(r) || r == 0) return kNaN;
  return r;
}

__device__ double volume(const Solid& s, double h) {
  if (s.overflow) return kNaN;
  V3 apex{{0, 0, 0}};
  int count = 0;
  for (int k = 0; k < s.n; ++k) {
    if (s.f[k].overflow) return kNaN;
    for (int j = 0; j < s.f[k].n; ++j) {
      apex = add(apex, s.f[k].v[j].p);
      ++count;
    }
  }
  if (count == 0) return 0;
  apex = scale(apex, 1.0 / count);
  double total = 0;
  for (int k = 0; k < s.n; ++k) {
    const Polygon& g = s.f[k];
    for (int j = 1; j + 1 < g.n; ++j) {
      V3 e[3] = {sub(g.v[0].p, apex), sub(g.v[j].p, apex), sub(g.v[j + 1].p, apex)};
      int ex = 0;
      for (int a = 0; a < 3; ++a) ex += normalize_row(e, a);
      const double scaled = fabs(dot(e[0], cross(e[1], e[2]))) / 6;
      total += physical_measure(scaled, ex, h, 3);
    }
  }
  return total;
}

__device__ double area(const Polygon& g, double h) {
  if (g.overflow) return kNaN;
  double total = 0;
  for (int j = 1; j + 1 < g.n; ++j) {
    const V3 e0 = sub(g.v[j].p, g.v[0].p), e1 = sub(g.v[j + 1].p, g.v[0].p);
    double w[3];
    for (int a = 0; a < 3; ++a) {
      V3 n[2] = {e0, e1};
      const int f = (a + 1) % 3, s = (a + 2) % 3;
      const int fe = normalize_row(n, f);
      const int se = normalize_row(n, s);
      const double c = (n[0].c[f] * n[1].c[s] - n[0].c[s] * n[1].c[f]) / 2;
      w[a] = physical_measure(c, fe + se, h, 2);
    }
    total += hypot3(w[0], w[1], w[2]);
  }
  return total;
}

// ComputeOpenFaceArea: zero unless a vertex lies strictly on the open side.
__device__ double open_area(const Polygon& g, const Plane& plane, double h) {
  if (g.overflow) return kNaN;
  bool any = false;
  for (int j = 0; j < g.n; ++j) any = any || distance(g.v[j], plane) < 0;
  if (!any) return 0;
  Polygon c;
  clip<kCross>(g, plane, c, nullptr);
  return area(c, h);
}

struct Planes {
  bool present, slab;
  V3 normal;
  double low, high;
};

// LocalObstacle: the planes in the unit frame of the element whose low corner is lattice index c.
__device__ inline int local_planes(const FreezeDevice& d, fl_int3 c, Planes& out) {
  const fl_obstacle& o = d.obstacle;
  out.present = o.kind == FL_OBSTACLE_HALFSPACE || o.kind == FL_OBSTACLE_SLAB;
  out.slab = o.kind == FL_OBSTACLE_SLAB;
  out.normal = {{o.unit_normal.x, o.unit_normal.y, o.unit_normal.z}};
  out.low = o.coordinate - (out.slab ? o.half_width : 0);
  out.high = o.coordinate + (out.slab ? o.half_width : 0);
  if (!out.present) return kOk;
  const V3 anchor{{d.origin.x + d.h * c.x, d.origin.y + d.h * c.y, d.origin.z + d.h * c.z}};
  const double projection = dot(out.normal, anchor);
  out.low = (out.low - projection) / d.h;
  out.high = (out.high - projection) / d.h;
  if (!finite3(anchor) || !finite(out.low) || !finite(out.high) || (out.slab && out.low >= out.high))
    return kOutOfRange;
  return kOk;
}

__constant__ int kTetrahedra[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7},
                                      {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};

// ---- the open tank (base/liquid_boundary_math.h:278-292; liquid_boundary_geometry_math_3d.h:420-623) ----

struct TankLattice {
  double origin, index, width;
};
__device__ inline TankLattice tank_axis(const FreezeDevice& d, fl_int3 c, int axis) {
  const double o[3] = {d.origin.x, d.origin.y, d.origin.z};
  const int32_t i[3] = {c.x, c.y, c.z};
  return {o[axis], double(i[axis]), d.h};
}

// TankCellHasResolvedWorldExtent: each axis's world interval is ordered, finite and resolved by the
// spacing toward its interior at both ends.
__device__ inline bool tank_cell_resolved(const FreezeDevice& d, fl_int3 c) {
  const double anchor[3] = {d.origin.x + d.h * c.x, d.origin.y + d.h * c.y, d.origin.z + d.h * c.z};
  for (int a = 0; a < 3; ++a) {
    const TankLattice L = tank_axis(d, c, a);
    const double low = anchor[a], high = L.origin + L.width * (L.index + 1);
    if (!finite(low) || !finite(high) || low >= high) return false;
    const double low_spacing = nextafter(low, high) - low;
    const double high_spacing = high - nextafter(high, low);
    if (low_spacing > L.width || high_spacing > L.width) return false;
  }
  return true;
}

// TankPlaneOffset: world - origin - h index as a bounded sum in the lattice frame.
__device__ inline bounded::PSum tank_plane_offset(double world, const TankLattice& L) {
  bounded::PSum o;
  o.add(world);
  o.add(-L.origin);
  o.product(-L.width, L.index);
  return o;
}

__device__ inline int tank_plane_sign(double world, const TankLattice& L, int& sign) {
  const double translated = world - L.origin;
  if (isinf(translated)) {
    sign = translated > 0 ? 1 : -1;
    return kOk;
  }
  return tank_plane_offset(world, L).bounds().sign(&sign) ? kOk : kOutOfRange;
}

// TankClipCoordinate: the plane's coordinate in the unit cell, 0 or 1 beyond a corner.
__device__ inline int tank_clip_coordinate(double world, const TankLattice& L, double& out) {
  const double translated = world - L.origin;
  if (isinf(translated)) {
    out = translated > 0 ? 1.0 : 0.0;
    return kOk;
  }
  bounded::PSum offset = tank_plane_offset(world, L);
  const bounded::Bounds low = offset.bounds();
  int low_side = 0;
  if (!low.sign(&low_side)) return kOutOfRange;
  if (low_side <= 0) {
    out = 0.0;
    return kOk;
  }
  offset.add(-L.width);
  int high_side = 0;
  if (!offset.bounds().sign(&high_side)) return kOutOfRange;
  if (high_side >= 0) {
    out = 1.0;
    return kOk;
  }
  const double cell = low.estimate / L.width;
  if (cell <= 0 || cell >= 1) return kOutOfRange;
  out = cell;
  return kOk;
}

struct TankBounds {
  double low[3], high[3];
};

// LocalTankBounds: the walls in the unit frame of the element at c (a face skips its normal axis).
__device__ inline int local_tank_bounds(const FreezeDevice& d, fl_int3 c, int face_axis, TankBounds& b) {
  const fl_tank& t = d.obstacle.tank;
  const double low_world[3] = {t.start_low_x, t.floor_y, t.low_z};
  const double high_world[3] = {t.high_x, __builtin_inf(), t.high_z};
  for (int a = 0; a < 3; ++a) {
    b.low[a] = 0;
    b.high[a] = 1;
  }
  for (int a = 0; a < 3; ++a) {
    if (a == face_axis) continue;
    const TankLattice L = tank_axis(d, c, a);
    double lo = 0, hi = 0;
    const int s1 = tank_clip_coordinate(low_world[a], L, lo);
    const int s2 = tank_clip_coordinate(high_world[a], L, hi);
    if (s1 || s2) return kOutOfRange;
    b.low[a] = lo;
    b.high[a] = hi;
    if (lo == hi && lo > 0 && hi < 1) return kOutOfRange;
  }
  return kOk;
}

// TankBoxMeasure: the product of the box's extents, scaled before multiplying.
__device__ inline double tank_box_measure(const TankBounds& b, double h, int excluded) {
  double mantissa = 1;
  int exponent = 0, dimension = 0;
  for (int a = 0; a < 3; ++a) {
    if (a == excluded) continue;
    const double extent = b.high[a] - b.low[a];
    if (extent <= 0) return 0;
    int e = 0;
    mantissa *= frexp(extent, &e);
    exponent += e;
    ++dimension;
  }
  return physical_measure(mantissa, exponent, h, dimension);
}

// ClipLiquidToTank: the liquid inside the walls that cut the unit cell; in and tmp swap roles, the
// result is left in `in`.
__device__ inline void clip_liquid_to_tank(Solid& in, Solid& tmp, const TankBounds& b) {
  for (int a = 0; a < 3; ++a) {
    V3 n{{0, 0, 0}};
    if (b.low[a] > 0) {
      n.c[a] = -1;
      clip_solid(in, Plane{n, -b.low[a], false, a}, tmp);
      in = tmp;
    }
    if (b.high[a] < 1) {
      n.c[a] = 1;
      clip_solid(in, Plane{n, b.high[a], false, a}, tmp);
      in = tmp;
    }
  }
}

// ComputeTankFace: the aperture of an open face; on a closed low-x face the piston's rectangle and flux.
__device__ int tank_face(const FreezeDevice& d, fl_int3 l, int axis, const TankBounds& b, fl_frozen_face& r) {
  const fl_tank& t = d.obstacle.tank;
  const double h = d.h, full_area = h * h;
  const double low_world[3] = {t.start_low_x, t.floor_y, t.low_z};
  const double high_world[3] = {t.high_x, __builtin_inf(), t.high_z};
  const TankLattice L = tank_axis(d, l, axis);
  int low_sign = 0, high_sign = 0;
  const int s1 = tank_plane_sign(low_world[axis], L, low_sign);
  const int s2 = tank_plane_sign(high_world[axis], L, high_sign);
  if (s1 || s2) return kOutOfRange;
  const bool piston_side = axis == 0 && low_sign >= 0;
  const bool open = low_sign < 0 && high_sign > 0;
  double o = 0, flux = 0, v = 0;
  uint8_t low_x = 0;
  if (open || piston_side) {
    const double transverse = tank_box_measure(b, h, axis);
    if (!normal_or_zero(transverse) || transverse > full_area) return kOutOfRange;
    if (open) {
      o = transverse;
      for (int a = 0; a < 3; ++a)
        if (a != axis && transverse == full_area && (b.low[a] > 0 || b.high[a] < 1)) return kOutOfRange;
    } else if (transverse > 0) {
      low_x = 1;
      v = t.piston_velocity_x;
      flux = transverse * v;
      if (!normal_or_zero(flux) || (v != 0 && flux == 0)) return kOutOfRange;
    }
  }
  const double blocked = full_area - o;
  if (!normal_or_zero(blocked)) return kOutOfRange;
  r.open_area = o;
  r.blocked_area = blocked;
  r.wall_normal_velocity = v;
  r.prescribed_wall_flux = flux;
  r.has_prescribed_wall_flux = 1;
  r.is_low_x_piston_face = low_x;
  return kOk;
}

// ---- sampled solids ----

constexpr double kMeasureSlack = 64 * 2.220446049250313e-16;

// SampledSolidExteriorPlane: the affine extension of the solid samples along a chain of corners whose
// consecutive bits differ in one axis; its distance is minus the interpolated solid distance.
template <int N>
__device__ inline Plane sampled_exterior(const double (&s)[8], const int (&corners)[N]) {
  Plane p{{{0, 0, 0}}, s[corners[0]], false};
  for (int e = 1; e < N; ++e) {
    const int step = corners[e] ^ corners[e - 1];
    const int axis = (step & 1) != 0 ? 0 : ((step & 2) != 0 ? 1 : 2);
    p.normal.c[axis] = s[corners[e - 1]] - s[corners[e]];
  }
  return p;
}

// ClipSampledAvailableFace: the part of a polygon on the open side of the analytic boundary.
__device__ inline void available_face(const Polygon& g, const Plane* analytic, Polygon& out) {
  if (analytic == nullptr || g.overflow) {
    out = g;
    return;
  }
  bool any = false;
  for (int j = 0; j < g.n; ++j) any = any || distance(g.v[j], *analytic) < 0;
  if (!any) {
    out.clear();
    return;
  }
  clip<kCross>(g, *analytic, out, nullptr);
}

__device__ inline void triangle(const Vtx& a, const Vtx& b, const Vtx& c, Polygon& g) {
  g.clear();
  g.push(a);
  g.push(b);
  g.push(c);
}

// CheckSampledWetConnectivity: wet tetrahedra joined through positive wet openings of the triangles they
// share around the body diagonal must form one piece.
__device__ int wet_connectivity(const Vtx (&x)[8], const double (&s)[8], const bool (&wet)[6], double width,
                                const Plane* analytic) {
  const auto wet_corner = [&](int k) {
    return x[k].phi < 0 && s[k] > 0 && (analytic == nullptr || distance(x[k], *analytic) < 0);
  };
  if (wet_corner(0) || wet_corner(7)) return kOk;
  int components = 0;
  for (int t = 0; t < 6; ++t) components += wet[t] ? 1 : 0;
  if (components <= 1) return kOk;
  const int ring[6] = {0, 1, 4, 5, 3, 2};
  const int shared[6] = {1, 5, 4, 6, 2, 3};
  int group[6] = {0, 1, 2, 3, 4, 5};
  for (int e = 0; e < 6; ++e) {
    const int first = ring[e], second = ring[(e + 1) % 6];
    if (!wet[first] || !wet[second] || group[first] == group[second]) continue;
    const int middle = shared[e];
    if (!(x[0].phi < 0 || x[middle].phi < 0 || x[7].phi < 0)) continue;
    const int corners[3] = {0, middle, 7};
    if (s[0] <= 0 && s[middle] <= 0 && s[7] <= 0) continue;
    Polygon tri, liquid, avail;
    triangle(x[0], x[middle], x[7], tri);
    clip<kCross>(tri, Plane{{{0, 0, 0}}, 0, true}, liquid, nullptr);
    available_face(liquid, analytic, avail);
    const double wet_area = open_area(avail, sampled_exterior(s, corners), width);
    if (!normal_or_zero(wet_area)) return kOutOfRange;
    if (wet_area > 0) {
      const int old_group = group[second], new_group = group[first];
      for (int k = 0; k < 6; ++k)
        if (group[k] == old_group) group[k] = new_group;
      --components;
    }
  }
  return components == 1 ? kOk : kFailedPrecondition;
}

// MeasureSampledSolidLiquidCell: wet volume and the six open face areas (low/high x, y, z).
// split_rule: 1 the reference's wet-piece connectivity refusal; 2 only where a corner lies in the solid
// (sampled distance <= 0 or on the analytic wall's side), so liquid that the particle field itself
// splits inside a solid-free cell is measured as in the plane-obstacle freeze instead of refused; 3 no
// refusal (the wet pieces are one cell).
__device__ int measure_sampled_cell(double width, const double (&liquid)[8], const double (&solid)[8],
                                    const Plane* analytic, double& wet_volume, double (&open)[6],
                                    int split_rule = 1) {
  const bool solid_split_only = split_rule == 2;
  if (!finite(width) || width <= 0) return kInvalidArgument;
  const double full_area = width * width, full_volume = full_area * width;
  if (!normal(full_area) || !normal(full_volume)) return kOutOfRange;
  Vtx x[8];
  double largest = 0;
  for (int k = 0; k < 8; ++k) {
    if (!finite(liquid[k]) || !finite(solid[k])) return kInvalidArgument;
    x[k].p = {{double(k & 1), double((k >> 1) & 1), double((k >> 2) & 1)}};
    x[k].phi = liquid[k];
    largest = max_std(largest, fabs(solid[k]));
  }
  int exponent = 0;
  (void)frexp(largest, &exponent);
  double scaled[8];
  for (int k = 0; k < 8; ++k) {
    scaled[k] = ldexp(solid[k], -exponent);
    if (solid[k] != 0 && !normal(scaled[k])) return kOutOfRange;
  }
  double wet = 0;
  bool wet_tet[6] = {false, false, false, false, false, false};
  Solid a_solid, b_solid;
  for (int t = 0; t < 6; ++t) {
    const int corners[4] = {0, kTetrahedra[t][1], kTetrahedra[t][2], 7};
    bool has_liquid = false, has_exterior = false;
    for (int k = 0; k < 4; ++k) {
      has_liquid = has_liquid || x[corners[k]].phi < 0;
      has_exterior = has_exterior || scaled[corners[k]] > 0;
    }
    if (!has_liquid || !has_exterior) continue;
    const Vtx a = x[corners[0]], b = x[corners[1]], c = x[corners[2]], dd = x[corners[3]];
    a_solid.n = 4;
    a_solid.overflow = false;
    triangle(a, b, c, a_solid.f[0]);
    triangle(a, b, dd, a_solid.f[1]);
    triangle(a, c, dd, a_solid.f[2]);
    triangle(b, c, dd, a_solid.f[3]);
    clip_solid(a_solid, Plane{{{0, 0, 0}}, 0, true}, b_solid);           // liquid part
    clip_solid(b_solid, sampled_exterior(scaled, corners), a_solid);    // outside the solid
    double v;
    if (analytic == nullptr) {
      v = volume(a_solid, width);
    } else {
      clip_solid(a_solid, *analytic, b_solid);
      v = volume(b_solid, width);
    }
    wet_tet[t] = v > 0;
    wet += v;
  }
  if (!normal_or_zero(wet) || wet / full_volume > 1 + kMeasureSlack) return kOutOfRange;
  wet = min_std(wet, full_volume);
  bool touched = !solid_split_only;
  for (int k = 0; k < 8 && !touched; ++k)
    touched = scaled[k] <= 0 || (analytic != nullptr && distance(x[k], *analytic) >= 0);
  if (touched && split_rule != 3)
    if (int code = wet_connectivity(x, scaled, wet_tet, width, analytic)) return code;
  double areas[6];
  for (int axis = 0; axis < 3; ++axis) {
    const int bit = 1 << axis;
    for (int side = 0; side < 2; ++side) {
      const int first = side * bit, last = first | (7 ^ bit);
      double o = 0;
      for (int tr = 0; tr < 3; ++tr) {
        if (tr == axis) continue;
        const int middle = first | (1 << tr);
        const int corners[3] = {first, middle, last};
        Polygon tri, avail;
        triangle(x[first], x[middle], x[last], tri);
        available_face(tri, analytic, avail);
        o += open_area(avail, sampled_exterior(scaled, corners), width);
      }
      if (!normal_or_zero(o) || o / full_area > 1 + kMeasureSlack) return kOutOfRange;
      areas[2 * axis + side] = min_std(o, full_area);
    }
  }
  wet_volume = wet;
  for (int k = 0; k < 6; ++k) open[k] = areas[k];
  return kOk;
}

// MeasureSampledSolidFaceWeights: the blocked part of the face's two triangles, split over the four
// corners by the centroids of its fan pieces.
__device__ int sampled_face_weights(double width, const double (&solid)[4], const Plane* analytic,
                                    double (&out)[4]) {
  if (!finite(width) || width <= 0) return kInvalidArgument;
  const double full_area = width * width;
  if (!normal(full_area)) return kOutOfRange;
  double largest = 0;
  for (int k = 0; k < 4; ++k) {
    if (!finite(solid[k])) return kInvalidArgument;
    largest = max_std(largest, fabs(solid[k]));
  }
  int exponent = 0;
  (void)frexp(largest, &exponent);
  double scaled[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  Vtx x[4];
  for (int k = 0; k < 4; ++k) {
    scaled[k] = ldexp(solid[k], -exponent);
    if (solid[k] != 0 && !normal(scaled[k])) return kOutOfRange;
    x[k].p = {{double(k & 1), double((k >> 1) & 1), 0.0}};
    x[k].phi = 0;
  }
  double w[4] = {0, 0, 0, 0};
  for (int axis = 0; axis < 2; ++axis) {
    const int middle = 1 << axis;
    const int corners[3] = {0, middle, 3};
    Polygon tri, avail, blocked;
    triangle(x[0], x[middle], x[3], tri);
    const Plane exterior = sampled_exterior(scaled, corners);
    const Plane interior{scale(exterior.normal, -1), -exterior.coordinate, false};
    available_face(tri, analytic, avail);
    clip<kCross>(avail, interior, blocked, nullptr);
    if (blocked.overflow) return kOutOfRange;
    for (int v = 1; v + 1 < blocked.n; ++v) {
      Polygon piece;
      triangle(blocked.v[0], blocked.v[v], blocked.v[v + 1], piece);
      const double a = area(piece, width);
      if (!normal_or_zero(a)) return kOutOfRange;
      const V3 centroid = scale(add(add(piece.v[0].p, piece.v[1].p), piece.v[2].p), 1.0 / 3);
      const double along = clamp_std(centroid.c[axis], 0.0, 1.0);
      const double across = clamp_std(centroid.c[1 - axis], 0.0, along);
      w[0] += a * (1 - along);
      w[middle] += a * (along - across);
      w[3] += a * across;
    }
  }
  const double measured = w[0] + w[1] + w[2] + w[3];
  if (!normal_or_zero(measured) || measured / full_area > 1 + kMeasureSlack) return kOutOfRange;
  for (int k = 0; k < 4; ++k) out[k] = w[k];
  return kOk;
}

__device__ inline double lookup(const fl_center_distance* s, uint64_t n, fl_int3 k, bool& found) {
  const uint64_t i = find_sorted(s, n, k, [](const fl_center_distance& x) { return x.cell_index; });
  found = i != n;
  return found ? s[i].signed_distance : 0.0;
}

__global__ void k_check_samples(const fl_center_distance* s, uint64_t n, uint64_t phase, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  if (!finite(s[i].signed_distance) || (i > 0 && !key_less(s[i - 1].cell_index, s[i].cell_index)))
    fail(err, phase, i, kInvalidArgument);
}

// The solid samples' checks, after the corner samples' in phase 1.
__global__ void k_check_solids(const fl_solid_corner* s, uint64_t n, uint64_t first, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  if (!finite(s[i].signed_distance) || (i > 0 && !key_less(s[i - 1].corner_index, s[i].corner_index)))
    fail(err, 1, first + i, kInvalidArgument);
}

__device__ inline const fl_solid_corner* lookup_solid(const fl_solid_corner* s, uint64_t n, fl_int3 k) {
  const uint64_t i = find_sorted(s, n, k, [](const fl_solid_corner& x) { return x.corner_index; });
  return i != n ? s + i : nullptr;
}

// The liquid volume on each side of the obstacle of a cell with corner distances phi (kCell).
// (with the open tank's bounds: the liquid inside its walls).
__device__ void cell_volumes(const double (&phi)[8], const Planes& P, double h, double& low, double& high,
                             const TankBounds* tank = nullptr) {
  Vtx x[8];
  for (int k = 0; k < 8; ++k) {
    x[k].phi = phi[k];
    for (int a = 0; a < 3; ++a) x[k].p.c[a] = (k >> a) & 1;
  }
  low = high = 0;
  Solid a_solid, liquid;  // the tetrahedron, then each obstacle part of its liquid
  Solid& tet = a_solid;
  Solid& part = a_solid;
  for (int t = 0; t < 6; ++t) {
    bool negative = false;
    for (int k = 0; k < 4; ++k) negative = negative || x[kTetrahedra[t][k]].phi < 0;
    if (!negative) continue;
    const Vtx a = x[kTetrahedra[t][0]], b = x[kTetrahedra[t][1]], cc = x[kTetrahedra[t][2]], dd = x[kTetrahedra[t][3]];
    const Vtx tri[4][3] = {{a, b, cc}, {a, b, dd}, {a, cc, dd}, {b, cc, dd}};
    tet.n = 4;
    tet.overflow = false;
    for (int k = 0; k < 4; ++k) {
      tet.f[k].clear();
      for (int j = 0; j < 3; ++j) tet.f[k].push(tri[k][j]);
    }
    clip_solid(tet, Plane{{{0, 0, 0}}, 0, true}, liquid);
    if (tank) {
      clip_liquid_to_tank(liquid, a_solid, *tank);
      high += volume(liquid, h);
    } else if (!P.present) {
      high += volume(liquid, h);
    } else {
      clip_solid(liquid, Plane{scale(P.normal, -1), -P.high, false}, part);
      high += volume(part, h);
      if (P.slab) {
        clip_solid(liquid, Plane{P.normal, P.low, false}, part);
        low += volume(part, h);
      }
    }
  }
}

// The consume step of a candidate: split refusal, volume checks, and for a liquid cell its record and
// its six face slots.
__device__ void consume_cell(const FreezeDevice& d, uint64_t i, fl_int3 c, double center, double low, double high,
                             uint64_t* err) {
  if (low > 0 && high > 0 && center < 0) return fail(err, 2, i, kFailedPrecondition);
  const double unclamped = low + high;
  if (!finite(unclamped) || unclamped < 0 || unclamped / d.full_volume > 1 + 64 * 2.220446049250313e-16)
    return fail(err, 2, i, kOutOfRange);
  const double vol = min_std(d.full_volume, unclamped);
  if (!normal_or_zero(vol)) return fail(err, 2, i, kOutOfRange);
  if (center >= 0 || vol == 0) return;
  for (int a = 0; a < 3; ++a)
    for (int side = 0; side < 2; ++side) {
      fl_int3 l;
      if (!offset(c, a, side, l)) return fail(err, 2, i, kOutOfRange);
      uint64_t slot = 0;
      if (int code = sorted_face_slot(d.blocks, d.nb, a, l, slot)) return fail(err, 2, i, code);
      d.slots[6 * i + 2 * a + side] = slot;
    }
  fl_frozen_cell r{};
  r.cell_index = c;
  r.fluid_volume = vol;
  r.center_distance = center;
  r.curvature = 0;
  d.staged[i] = r;
  d.keep[i] = 1;
}

// The clipped volume of an all-liquid cell that no obstacle plane cuts: it depends on h alone, so one
// lane computes it with the general clipping for every such cell.
__global__ void k_uncut_volume(FreezeDevice d) {
  const double phi[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  Planes P{};
  P.present = false;
  double low = 0, high = 0;
  cell_volumes(phi, P, d.h, low, high);
  d.uncut_volume[0] = high;
  if (d.sampled) {  // the sampled measure of a cell that no plane cuts: its clips return their inputs
    const double solid[8] = {1, 1, 1, 1, 1, 1, 1, 1};
    double wet = 0, open[6];
    d.uncut_volume[1] = measure_sampled_cell(d.h, phi, solid, nullptr, wet, open, 1) == kOk ? wet : -1.0;
  }
}

// Whether every corner of the unit cell lies on the kept side (distance <= 0) of the plane.
__device__ inline bool all_kept(const Plane& plane) {
  for (int k = 0; k < 8; ++k) {
    const Vtx x{{{double(k & 1), double((k >> 1) & 1), double((k >> 2) & 1)}}, 0.0};
    if (distance(x, plane) > 0) return false;
  }
  return true;
}
// Whether every corner lies strictly on the removed side of the plane (nothing kept).
__device__ inline bool none_kept(const Plane& plane) {
  for (int k = 0; k < 8; ++k) {
    const Vtx x{{{double(k & 1), double((k >> 1) & 1), double((k >> 2) & 1)}}, 0.0};
    if (distance(x, plane) <= 0) return false;
  }
  return true;
}

// One lane per candidate: order check and sample lookups, then the kCell query for the cells it settles
// without clipping, which give exactly what clipping gives: no negative corner (every tetrahedron
// skipped, no volume), or every corner negative with each obstacle plane keeping all corners or none
// (each clip returns its input or nothing, so the volume is the uncut cell's). Other cells go to the
// work list of k_freeze_cut with their samples in phi[9 i ...] (eight corners, then the centre).
__global__ void k_freeze_stage(FreezeDevice d, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= d.ncand) return;
  d.keep[i] = 0;
  const fl_int3 c = d.candidates[i];
  if (i > 0 && !key_less(d.candidates[i - 1], c)) return fail(err, 2, i, kInvalidArgument);
  bool found = false;
  const double center = lookup(d.centers, d.ncenters, c, found);
  if (!found) return fail(err, 2, i, kFailedPrecondition);
  double phi[8];
  bool any_negative = false, all_negative = true;
  for (int k = 0; k < 8; ++k) {
    fl_int3 q = c;
    for (int a = 0; a < 3; ++a)
      if (!offset(q, a, (k >> a) & 1, q)) return fail(err, 2, i, kOutOfRange);
    phi[k] = lookup(d.corners, d.ncorners, q, found);
    if (!found) return fail(err, 2, i, kFailedPrecondition);
    if (d.sampled) {
      const fl_solid_corner* solid = lookup_solid(d.solids, d.nsolids, q);
      if (!solid) return fail(err, 2, i, kFailedPrecondition);
      d.solid_phi[8 * i + k] = solid->signed_distance;
    }
    any_negative = any_negative || phi[k] < 0;
    all_negative = all_negative && phi[k] < 0;
  }
  if (d.sampled) {
    // A cell wholly liquid, with every solid sample clearly outside the solid and every corner strictly
    // on the fluid side of the analytic wall, is one that no clip of the sampled measure cuts (the
    // liquid, sampled-exterior and wall planes keep every vertex), so its volume is the uncut cell's
    // sampled measure, computed once. Every other cell is measured against the sampled solid.
    if (all_negative && d.uncut_volume[1] >= 0) {
      double smax = 0, smin = __builtin_inf();
      for (int k = 0; k < 8; ++k) {
        smax = max_std(smax, fabs(d.solid_phi[8 * i + k]));
        smin = min_std(smin, d.solid_phi[8 * i + k]);
      }
      bool clear = smin > 1e-9 * smax;
      if (clear) {
        Planes P;
        if (int code = local_planes(d, c, P)) return fail(err, 2, i, code);
        if (P.present) {
          const Plane analytic{scale(P.normal, -1), -P.high, false};
          for (int k = 0; k < 8 && clear; ++k) {
            const Vtx x{{{double(k & 1), double((k >> 1) & 1), double((k >> 2) & 1)}}, 0.0};
            clear = distance(x, analytic) < 0;
          }
        }
      }
      if (clear) return consume_cell(d, i, c, center, 0.0, d.uncut_volume[1], err);
    }
    for (int k = 0; k < 8; ++k) d.phi[9 * i + k] = phi[k];
    d.phi[9 * i + 8] = center;
    const unsigned long long w = atomicAdd(reinterpret_cast<unsigned long long*>(d.counts + 3), 1ull);
    d.work[w] = i;
    return;
  }
  Planes P;
  if (int code = local_planes(d, c, P)) return fail(err, 2, i, code);
  if (d.obstacle.kind == FL_OBSTACLE_TANK) {  // the tank cell's box, or the liquid inside it
    if (!tank_cell_resolved(d, c)) return fail(err, 2, i, kOutOfRange);
    TankBounds tb;
    if (int code = local_tank_bounds(d, c, -1, tb)) return fail(err, 2, i, code);
    const double box = tank_box_measure(tb, d.h, -1);
    if (!normal_or_zero(box)) return fail(err, 2, i, kOutOfRange);
    if (box == 0 || !any_negative) return consume_cell(d, i, c, center, 0.0, 0.0, err);
    if (all_negative) return consume_cell(d, i, c, center, 0.0, box, err);
    for (int k = 0; k < 8; ++k) d.phi[9 * i + k] = phi[k];
    d.phi[9 * i + 8] = center;
    const unsigned long long w = atomicAdd(reinterpret_cast<unsigned long long*>(d.counts + 3), 1ull);
    d.work[w] = i;
    return;
  }
  if (!any_negative) return consume_cell(d, i, c, center, 0.0, 0.0, err);
  if (all_negative) {
    const bool high_whole = !P.present || all_kept(Plane{scale(P.normal, -1), -P.high, false});
    const bool low_empty = !P.slab || none_kept(Plane{P.normal, P.low, false});
    if (high_whole && low_empty) return consume_cell(d, i, c, center, 0.0, d.uncut_volume[0], err);
  }
  for (int k = 0; k < 8; ++k) d.phi[9 * i + k] = phi[k];
  d.phi[9 * i + 8] = center;
  const unsigned long long w = atomicAdd(reinterpret_cast<unsigned long long*>(d.counts + 3), 1ull);
  d.work[w] = i;
}

// One lane per listed candidate: the clipped volumes and the consume step.
__global__ void __launch_bounds__(64) k_freeze_cut(FreezeDevice d, uint64_t* err) {
  const uint64_t w = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (w >= d.counts[3]) return;
  const uint64_t i = d.work[w];
  const fl_int3 c = d.candidates[i];
  double phi[8];
  for (int k = 0; k < 8; ++k) phi[k] = d.phi[9 * i + k];
  Planes P;
  double low = 0, high = 0;
  if (d.sampled) {  // ComputeSampledSolidGeometry for a cell: the wet volume as the high side's
    if (int code = local_planes(d, c, P)) return fail(err, 2, i, code);
    if (d.obstacle.kind != FL_OBSTACLE_ABSENT && d.obstacle.kind != FL_OBSTACLE_HALFSPACE)
      return fail(err, 2, i, kInvalidArgument);
    double solid[8], open[6];
    for (int k = 0; k < 8; ++k) solid[k] = d.solid_phi[8 * i + k];
    const Plane analytic{scale(P.normal, -1), -P.high, false};
    if (int code = measure_sampled_cell(d.h, phi, solid, P.present ? &analytic : nullptr, high, open, d.sampled))
      return fail(err, 2, i, code);
    return consume_cell(d, i, c, d.phi[9 * i + 8], 0.0, high, err);
  }
  (void)local_planes(d, c, P);  // succeeded in k_freeze_stage
  if (d.obstacle.kind == FL_OBSTACLE_TANK) {
    TankBounds tb;
    (void)local_tank_bounds(d, c, -1, tb);  // succeeded in k_freeze_stage
    cell_volumes(phi, P, d.h, low, high, &tb);
    return consume_cell(d, i, c, d.phi[9 * i + 8], low, high, err);
  }
  cell_volumes(phi, P, d.h, low, high);
  consume_cell(d, i, c, d.phi[9 * i + 8], low, high, err);
}

// Kept cells to their scanned positions; their face slots into the slot bitmap; the kept count.
__global__ void k_compact_cells(FreezeDevice d) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= d.ncand) return;
  const uint64_t pos = d.keep_scan[i];
  if (i == d.ncand - 1) d.counts[0] = pos + d.keep[i];
  if (!d.keep[i]) return;
  d.cells[pos] = d.staged[i];
  d.needs[pos] = 0;
  for (int k = 0; k < 6; ++k) {
    const uint64_t s = d.slots[6 * i + k];
    atomicOr(d.words + s / 64, 1ull << (s % 64));
  }
}

__global__ void k_clear(unsigned long long* w, uint64_t n) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) w[i] = 0;
}

__global__ void k_word_popcounts(const unsigned long long* w, uint64_t n, uint64_t* counts) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) counts[i] = __popcll(w[i]);
}

// Every set slot in increasing order, at its word's scanned position; the face count.
__global__ void k_compact_slots(FreezeDevice d) {
  const uint64_t w = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (w >= d.nwords) return;
  unsigned long long bits = d.words[w];
  uint64_t pos = d.word_scan[w];
  if (w == d.nwords - 1) d.counts[1] = pos + __popcll(bits);
  while (bits) {
    const int b = __ffsll(bits) - 1;
    bits &= bits - 1;
    d.face_slots[pos++] = w * 64 + uint64_t(b);
  }
}

// kFace: open and blocked areas of the face whose low corner is lattice l (two triangles).
__device__ int face_areas(double h, const Planes& P, int axis, double& open_out, double& blocked_out) {
  const double full_area = h * h;
  int tr[2], n = 0;
  for (int a = 0; a < 3; ++a)
    if (a != axis) tr[n++] = a;
  Vtx x[4] = {};
  x[1].p.c[tr[0]] = 1;
  x[2].p.c[tr[0]] = 1;
  x[2].p.c[tr[1]] = 1;
  x[3].p.c[tr[1]] = 1;
  double blocked = 0, open = 0;
  for (int t = 0; t < 2; ++t) {
    Polygon g;
    triangle(x[0], x[t + 1], x[t + 2], g);
    if (!P.present) {
      open += area(g, h);
      continue;
    }
    Polygon bl, tmp;
    clip<kCross>(g, Plane{P.normal, P.high, false}, bl, nullptr);
    if (P.slab) {
      clip<kCross>(bl, Plane{scale(P.normal, -1), -P.low, false}, tmp, nullptr);
      bl = tmp;
    }
    blocked += area(bl, h);
    open += open_area(g, Plane{scale(P.normal, -1), -P.high, false}, h);
    if (P.slab) open += open_area(g, Plane{P.normal, P.low, false}, h);
  }
  open_out = min_std(open, full_area);
  blocked_out = full_area - open_out;
  if (!finite(open) || !finite(blocked) || open / full_area > 1 + 64 * 2.220446049250313e-16 ||
      !normal_or_zero(open_out) || !normal_or_zero(blocked_out) || (blocked > 0 && blocked_out == 0))
    return kOutOfRange;
  return kOk;
}

// ComputeSampledSolidGeometry for the face at lattice l normal to axis: open and blocked areas, the wall
// flux of the corner normal velocities through the blocked-area weights plus an analytic wall's, and
// their mean over the weighted area. The solid samples of the four low-side corners are looked up here.
__device__ int sampled_face(const FreezeDevice& d, fl_int3 l, int axis, uint64_t slot, fl_frozen_face& r) {
  double solid[8] = {0, 0, 0, 0, 0, 0, 0, 0}, velocity[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const bool halfspace = d.obstacle.kind == FL_OBSTACLE_HALFSPACE;
  for (int k = 0; k < 8; ++k) {
    if ((k & (1 << axis)) != 0) continue;
    fl_int3 q = l;
    for (int a = 0; a < 3; ++a)
      if (!offset(q, a, (k >> a) & 1, q)) return kOutOfRange;
    const fl_solid_corner* sample = lookup_solid(d.solids, d.nsolids, q);
    if (!sample) return kFailedPrecondition;
    solid[k] = sample->signed_distance;
    const double wv[3] = {sample->wall_velocity.x, sample->wall_velocity.y, sample->wall_velocity.z};
    if (d.nface_velocities == 0) velocity[k] = wv[axis];
    else if (halfspace) velocity[k] = d.face_velocities[slot];
  }
  const double width = d.h;
  Planes P;
  if (int code = local_planes(d, l, P)) return code;
  if (d.obstacle.kind != FL_OBSTACLE_ABSENT && !halfspace) return kInvalidArgument;
  const Plane analytic3{scale(P.normal, -1), -P.high, false};
  const double liquid[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double wet = 0, open[6];
  if (int code = measure_sampled_cell(width, liquid, solid, P.present ? &analytic3 : nullptr, wet, open,
                                      d.sampled))
    return code;
  const double face_open = open[2 * axis];
  const double face_blocked = width * width - face_open;
  double face_solid[4], face_velocity[4];
  int fc = 0;
  for (int k = 0; k < 8; ++k) {
    if ((k & (1 << axis)) != 0) continue;
    face_solid[fc] = solid[k];
    face_velocity[fc] = velocity[k];
    if (!finite(face_velocity[fc])) return kInvalidArgument;
    ++fc;
  }
  // MeasureSampledSolidAnalyticFace: the analytic wall's blocked area and flux, and its open side in the
  // face's transverse coordinates.
  Plane fluid_side{{{0, 0, 0}}, -P.high, false};
  double analytic_blocked = 0, analytic_flux = 0;
  if (P.present) {
    double o = 0, b = 0;
    if (int code = face_areas(width, P, axis, o, b)) return code;
    const double v[3] = {d.obstacle.wall_velocity.x, d.obstacle.wall_velocity.y, d.obstacle.wall_velocity.z};
    analytic_blocked = b;
    analytic_flux = b * v[axis];
    int t = 0;
    for (int a = 0; a < 3; ++a)
      if (a != axis) fluid_side.normal.c[t++] = -P.normal.c[a];
    if (!normal_or_zero(analytic_flux)) return kOutOfRange;
  }
  double w[4];
  if (int code = sampled_face_weights(width, face_solid, P.present ? &fluid_side : nullptr, w)) return code;
  double weighted = 0, flux = 0;
  for (int k = 0; k < 4; ++k) {
    weighted += w[k];
    flux += w[k] * face_velocity[k];
  }
  flux += analytic_flux;
  weighted += analytic_blocked;
  const double slack = 64 * 2.220446049250313e-16 * width * width;
  if (!normal_or_zero(face_blocked) || !normal_or_zero(flux) || fabs(face_blocked - weighted) > slack ||
      (weighted > 0 && face_blocked == 0))
    return kOutOfRange;
  double mean = 0;
  if (weighted > 0) {
    mean = flux / weighted;
    if (!finite(mean)) return kOutOfRange;
  }
  r.open_area = face_open;
  r.blocked_area = face_blocked;
  if (halfspace || d.nface_velocities == 0) {
    r.wall_normal_velocity = mean;
    r.prescribed_wall_flux = flux;
    r.has_prescribed_wall_flux = 1;
  } else {
    r.wall_normal_velocity = d.face_velocities[slot];
  }
  return kOk;
}

// kFace per unique face: open and blocked areas of the two triangles and the wall velocity.
__global__ void __launch_bounds__(64) k_freeze_faces(FreezeDevice d, uint64_t nf, uint64_t* err) {
  const uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (j >= nf) return;
  const uint64_t slot = d.face_slots[j];
  const fl_int3 b = d.blocks[slot / 192];
  const uint64_t within = slot % 192, local = within % 64;
  const int axis = int(within / 64);
  const fl_int3 l{int32_t(int64_t(b.x) * 4 + int64_t(local % 4)), int32_t(int64_t(b.y) * 4 + int64_t((local / 4) % 4)),
                  int32_t(int64_t(b.z) * 4 + int64_t(local / 16))};
  fl_frozen_face r{};
  r.axis = uint8_t(axis);
  r.lattice = l;
  r.face_slot = slot;
  d.faces[j] = r;  // a closed placeholder until the areas are known, so later phases read a valid record
  if (d.obstacle.kind == FL_OBSTACLE_TANK) {
    TankBounds tb;
    if (int code = local_tank_bounds(d, l, axis, tb)) return fail(err, 3, j, code);
    if (int code = tank_face(d, l, axis, tb, r)) return fail(err, 3, j, code);
    d.faces[j] = r;
    return;
  }
  if (d.sampled) {
    if (int code = sampled_face(d, l, axis, slot, r)) return fail(err, 3, j, code);
    d.faces[j] = r;
    return;
  }
  Planes P;
  if (int code = local_planes(d, l, P)) return fail(err, 3, j, code);
  if (int code = face_areas(d.h, P, axis, r.open_area, r.blocked_area)) return fail(err, 3, j, code);
  const double v[3] = {d.obstacle.wall_velocity.x, d.obstacle.wall_velocity.y, d.obstacle.wall_velocity.z};
  r.wall_normal_velocity = P.present ? v[axis] : 0.0;
  d.faces[j] = r;
}

// Cells on an open face whose other side is air (a centre sample >= 0) need curvature.
__global__ void k_curvature_needs(FreezeDevice d, uint64_t nc, uint64_t nf, uint64_t* err) {
  const uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (j >= nf) return;
  const fl_frozen_face f = d.faces[j];
  if (f.open_area == 0) return;
  const fl_int3 head = f.lattice;
  fl_int3 tail;
  if (!offset(head, f.axis, -1, tail)) return fail(err, 4, j, kOutOfRange);
  const auto key = [](const fl_frozen_cell& c) { return c.cell_index; };
  const uint64_t hc = find_sorted(d.cells, nc, head, key), tc = find_sorted(d.cells, nc, tail, key);
  const bool has_head = hc != nc, has_tail = tc != nc;
  if (has_head && has_tail) return;
  bool found = false;
  const double air = lookup(d.centers, d.ncenters, has_head ? tail : head, found);
  if (!found) return fail(err, 4, j, kFailedPrecondition);
  if (air >= 0) {
    d.needs[has_head ? hc : tc] = 1;
    d.counts[2] = 1;
  }
}

__global__ void k_exact_curvature(FreezeDevice d, uint64_t nc, double curvature) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < nc && d.needs[i]) d.cells[i].curvature = curvature;
}

// kCurvature from the 19-point centred halo (x fastest, index 13 the centre).
__global__ void k_curvature(FreezeDevice d, uint64_t nc, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nc || !d.needs[i]) return;
  double s[27] = {};
  const fl_int3 c = d.cells[i].cell_index;
  for (int z = -1; z <= 1; ++z)
    for (int y = -1; y <= 1; ++y)
      for (int x = -1; x <= 1; ++x) {
        if (x != 0 && y != 0 && z != 0) continue;
        fl_int3 q = c;
        const int o[3] = {x, y, z};
        for (int a = 0; a < 3; ++a)
          if (!offset(q, a, o[a], q)) return fail(err, 5, i, kOutOfRange);
        bool found = false;
        s[(z + 1) * 9 + (y + 1) * 3 + x + 1] = lookup(d.centers, d.ncenters, q, found);
        if (!found) return fail(err, 5, i, kFailedPrecondition);
      }
  const double h = d.h;
  const int stride[3] = {1, 3, 9};
  V3 grad{{0, 0, 0}};
  V3 H[3] = {{{0, 0, 0}}, {{0, 0, 0}}, {{0, 0, 0}}};
  for (int a = 0; a < 3; ++a) {
    const double lo = s[13 - stride[a]], hi = s[13 + stride[a]];
    grad.c[a] = (hi - lo) / (2 * h);
    H[a].c[a] = ((hi - s[13]) + (lo - s[13])) / h / h;
    for (int o = 0; o < a; ++o) {
      double mixed = 0;
      for (int f = -1; f <= 1; f += 2)
        for (int g = -1; g <= 1; g += 2) mixed += double(f * g) * (s[13 + f * stride[a] + g * stride[o]] / 4);
      H[a].c[o] = mixed / h / h;
      H[o].c[a] = H[a].c[o];
    }
  }
  const double norm = hypot3(grad.c[0], grad.c[1], grad.c[2]);
  if (!finite3(grad) || !finite(norm)) return fail(err, 5, i, kOutOfRange);
  if (norm < 1e-8) return fail(err, 5, i, kFailedPrecondition);
  const V3 n = scale(grad, 1 / norm);
  double k = 0;
  for (int a = 0; a < 3; ++a) {
    if (!finite3(H[a])) return fail(err, 5, i, kOutOfRange);
    k += (H[a].c[a] - n.c[a] * dot(H[a], n)) / norm;
  }
  if (!finite(k)) return fail(err, 5, i, kOutOfRange);
  d.cells[i].curvature = k;
}

}  // namespace

int freeze_cells_device(fl_ctx* ctx, const FreezeDevice& d) {
  hipStream_t s = ctx->stream;
  if (d.ncenters) k_check_samples<<<grid_for(d.ncenters), kThreads, 0, s>>>(d.centers, d.ncenters, 0, ctx->err);
  if (d.ncorners) k_check_samples<<<grid_for(d.ncorners), kThreads, 0, s>>>(d.corners, d.ncorners, 1, ctx->err);
  if (d.sampled && d.nsolids)
    k_check_solids<<<grid_for(d.nsolids), kThreads, 0, s>>>(d.solids, d.nsolids, d.ncorners, ctx->err);
  k_clear<<<grid_for(d.nwords), kThreads, 0, s>>>(d.words, d.nwords);
  k_clear<<<1, 64, 0, s>>>(reinterpret_cast<unsigned long long*>(d.counts), 4);
  if (d.ncand) {
    k_uncut_volume<<<1, 1, 0, s>>>(d);
    k_freeze_stage<<<grid_for(d.ncand), kThreads, 0, s>>>(d, ctx->err);
    k_freeze_cut<<<grid_for(d.ncand, 64), 64, 0, s>>>(d, ctx->err);
    FL_TRY(exclusive_scan_u64(ctx, d.keep, d.keep_scan, d.ncand));
    k_compact_cells<<<grid_for(d.ncand), kThreads, 0, s>>>(d);
  }
  if (d.nwords) {
    k_word_popcounts<<<grid_for(d.nwords), kThreads, 0, s>>>(d.words, d.nwords, d.word_counts);
    FL_TRY(exclusive_scan_u64(ctx, d.word_counts, d.word_scan, d.nwords));
    k_compact_slots<<<grid_for(d.nwords), kThreads, 0, s>>>(d);
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int freeze_faces_device(fl_ctx* ctx, const FreezeDevice& d, uint64_t nc, uint64_t nf, bool needs_curvature) {
  hipStream_t s = ctx->stream;
  if (nf) k_freeze_faces<<<grid_for(nf, 64), 64, 0, s>>>(d, nf, ctx->err);
  if (nf && needs_curvature) k_curvature_needs<<<grid_for(nf), kThreads, 0, s>>>(d, nc, nf, ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int freeze_curvature_device(fl_ctx* ctx, const FreezeDevice& d, uint64_t nc, bool exact, double curvature) {
  hipStream_t s = ctx->stream;
  if (!nc) return FL_OK;
  if (exact) k_exact_curvature<<<grid_for(nc), kThreads, 0, s>>>(d, nc, curvature);
  else k_curvature<<<grid_for(nc), kThreads, 0, s>>>(d, nc, ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

}  // namespace fl
