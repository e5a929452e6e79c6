// Oracle: frozen boundary geometry (test infrastructure only).
// Restates FreezeCore (native/src/liquid_boundary_geometry_3d.cc:765-1068) with its sequential
// geometry queries (RunGeometryQueries without a batch, :217-295) for the absent, half-space and slab
// obstacles (ValidateObstacle :132-175, PackObstacle and LocalObstacle
// native/include/fluidity/base/liquid_boundary_geometry_math_3d.h:625-650), and the scalar geometry
// arithmetic of that header: polygon and polyhedron clipping (:160-260), clipped volume and area
// (:296-417), the cell, face, curvature and capillary queries (:658-848); and the sampled-solid mode
// (native/include/fluidity/base/sampled_solid_liquid_geometry_3d.h: MeasureSampledSolidLiquidCell,
// MeasureSampledSolidFaceWeights, MeasureSampledSolidAnalyticFace, ComputeSampledSolidGeometry) that
// FreezeCore selects when the inputs carry solid corner samples.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "bounded_sums.h"
#include "grid_internal.h"

using namespace flo;

namespace {

using V3 = std::array<double, 3>;

V3 Add(V3 a, V3 b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
V3 Sub(V3 a, V3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 Scale(V3 a, double f) { return {a[0] * f, a[1] * f, a[2] * f}; }
double Dot(V3 a, V3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3 Cross(V3 a, V3 b) { return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}; }
bool Finite3(V3 a) { return std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]); }
bool PositiveFinite(double v) { return std::isfinite(v) && v > 0; }
bool NormalOrZero(double v) { return v == 0.0 || std::isnormal(v); }

// Fixed-capacity list with the reference's overflow flag (base/liquid_boundary_geometry_math_3d.h:7-55).
template <typename T, size_t N>
struct List {
  T v[N]{};
  size_t n = 0;
  bool overflow = false;
  void push(const T& x) {
    if (n == N) {
      overflow = true;
      return;
    }
    v[n++] = x;
  }
};
struct Vertex {
  V3 p;        // cell-width coordinates
  double phi;  // interpolated liquid distance
};
using Polygon = List<Vertex, 32>;
using Polyhedron = List<Polygon, 10>;
struct Plane {
  V3 normal;
  double coordinate;
  bool liquid;    // distance is the vertex's liquid value
  int snap = -1;  // an axis-aligned cut pins its crossings' coordinate on this axis to the plane's
};

double PlaneDistance(const Vertex& x, const Plane& p) { return p.liquid ? x.phi : Dot(p.normal, x.p) - p.coordinate; }

// ClipPolygon: keeps the side with distance <= 0; crossings use both directly computed weights.
Polygon Clip(const Polygon& in, const Plane& plane, Polygon* crossings) {
  Polygon out;
  if (in.overflow) {
    out.overflow = true;
    return out;
  }
  if (in.n == 0) return out;
  Vertex prev = in.v[in.n - 1];
  double dp = PlaneDistance(prev, plane);
  for (size_t i = 0; i < in.n; ++i) {
    const Vertex cur = in.v[i];
    const double dc = PlaneDistance(cur, plane);
    if ((dp <= 0) != (dc <= 0)) {
      const double m = std::max(std::abs(dp), std::abs(dc));
      const double ps = dp / m, cs = dc / m;
      const double diff = ps - cs;
      const double wp = -cs / diff, wc = ps / diff;
      Vertex x{Add(Scale(prev.p, wp), Scale(cur.p, wc)), wp * prev.phi + wc * cur.phi};
      if (plane.snap >= 0) x.p[plane.snap] = plane.coordinate * plane.normal[plane.snap];
      out.push(x);
      if (crossings) crossings->push(x);
    }
    if (dc <= 0) out.push(cur);
    prev = cur;
    dp = dc;
  }
  return out;
}

// libstdc++'s three-argument std::hypot (the largest magnitude times the root of the scaled squares).
double Hypot3(double x, double y, double z) { return std::hypot(x, y, z); }

// UnitDirection: binary-exponent normalisation, then division by the length.
V3 UnitDirection(V3 d) {
  const double largest = std::max({std::abs(d[0]), std::abs(d[1]), std::abs(d[2])});
  if (largest == 0) return {0, 0, 0};
  int e = 0;
  (void)std::frexp(largest, &e);
  for (double& c : d) c = std::ldexp(c, -e);
  return Scale(d, 1 / Hypot3(d[0], d[1], d[2]));
}

// ClipPolyhedron: clips every polygon and closes the cut with the deduplicated crossings, ordered by
// angle about their centroid.
Polyhedron ClipSolid(const Polyhedron& in, const Plane& plane) {
  Polyhedron out;
  if (in.overflow) {
    out.overflow = true;
    return out;
  }
  Polygon crossings;
  for (size_t k = 0; k < in.n; ++k) {
    Polygon c = Clip(in.v[k], plane, &crossings);
    if (c.overflow || crossings.overflow) {
      out.overflow = true;
      return out;
    }
    if (c.n >= 3) out.push(c);
  }
  Polygon cap;
  for (size_t k = 0; k < crossings.n; ++k) {
    bool seen = false;
    for (size_t j = 0; j < cap.n; ++j) seen = seen || cap.v[j].p == crossings.v[k].p;
    if (!seen) cap.push(crossings.v[k]);
  }
  if (cap.n >= 3) {
    V3 center{0, 0, 0};
    for (size_t k = 0; k < cap.n; ++k) center = Add(center, Scale(cap.v[k].p, 1.0 / cap.n));
    const V3 tangent = UnitDirection(Sub(cap.v[0].p, center));
    V3 normal{0, 0, 0};
    for (size_t k = 1; k < cap.n; ++k) {
      normal = UnitDirection(Cross(tangent, UnitDirection(Sub(cap.v[k].p, center))));
      if (Hypot3(normal[0], normal[1], normal[2]) > 0) break;
    }
    const V3 bitangent = Cross(normal, tangent);
    const double tl = Hypot3(tangent[0], tangent[1], tangent[2]);
    const double bl = Hypot3(bitangent[0], bitangent[1], bitangent[2]);
    if (tl > 0 && bl > 0) {
      const V3 ut = Scale(tangent, 1 / tl), ub = Scale(bitangent, 1 / bl);
      const auto angle = [&](const Vertex& x) {
        const V3 d = Sub(x.p, center);
        return std::atan2(Dot(d, ub), Dot(d, ut));
      };
      for (size_t k = 1; k < cap.n; ++k) {  // insertion sort, as the reference
        const Vertex x = cap.v[k];
        size_t j = k;
        while (j > 0 && angle(x) < angle(cap.v[j - 1])) {
          cap.v[j] = cap.v[j - 1];
          --j;
        }
        cap.v[j] = x;
      }
      out.push(cap);
    }
  }
  return out;
}

// NormalizeCoordinateRow: scales one coordinate of every edge by the largest one's binary exponent.
template <size_t N>
int NormalizeRow(std::array<V3, N>& edges, int axis) {
  double largest = 0;
  for (const V3& e : edges) largest = std::max(largest, std::abs(e[axis]));
  if (largest == 0) return 0;
  int e = 0;
  (void)std::frexp(largest, &e);
  for (V3& x : edges) x[axis] = std::ldexp(x[axis], -e);
  return e;
}

// PhysicalMeasure: restores the local exponent and h^dimension; NaN when a nonzero measure is not
// representable.
double PhysicalMeasure(double det, int local_exponent, double h, int dimension) {
  if (det == 0) return 0;
  int we = 0;
  const double wm = std::frexp(h, &we);
  double m = det;
  for (int a = 0; a < dimension; ++a) m *= wm;
  const double r = std::ldexp(m, local_exponent + dimension * we);
  if (!std::isfinite(r) || r == 0) return std::numeric_limits<double>::quiet_NaN();
  return r;
}

double Volume(const Polyhedron& s, double h) {
  if (s.overflow) return std::numeric_limits<double>::quiet_NaN();
  V3 apex{0, 0, 0};
  size_t count = 0;
  for (size_t k = 0; k < s.n; ++k) {
    if (s.v[k].overflow) return std::numeric_limits<double>::quiet_NaN();
    for (size_t j = 0; j < s.v[k].n; ++j) {
      apex = Add(apex, s.v[k].v[j].p);
      ++count;
    }
  }
  if (count == 0) return 0;
  apex = Scale(apex, 1.0 / count);
  double total = 0;
  for (size_t k = 0; k < s.n; ++k) {
    const Polygon& g = s.v[k];
    for (size_t j = 1; j + 1 < g.n; ++j) {
      std::array<V3, 3> e{Sub(g.v[0].p, apex), Sub(g.v[j].p, apex), Sub(g.v[j + 1].p, apex)};
      int ex = 0;
      for (int a = 0; a < 3; ++a) ex += NormalizeRow(e, a);
      const double scaled = std::abs(Dot(e[0], Cross(e[1], e[2]))) / 6;
      total += PhysicalMeasure(scaled, ex, h, 3);
    }
  }
  return total;
}

double Area(const Polygon& g, double h) {
  if (g.overflow) return std::numeric_limits<double>::quiet_NaN();
  double total = 0;
  for (size_t j = 1; j + 1 < g.n; ++j) {
    const std::array<V3, 2> e{Sub(g.v[j].p, g.v[0].p), Sub(g.v[j + 1].p, g.v[0].p)};
    V3 w{0, 0, 0};
    for (int a = 0; a < 3; ++a) {
      auto n = e;
      const int f = (a + 1) % 3, s = (a + 2) % 3;
      const int fe = NormalizeRow(n, f);
      const int se = NormalizeRow(n, s);
      const double c = (n[0][f] * n[1][s] - n[0][s] * n[1][f]) / 2;
      w[a] = PhysicalMeasure(c, fe + se, h, 2);
    }
    total += Hypot3(w[0], w[1], w[2]);
  }
  return total;
}

// ComputeOpenFaceArea: zero unless some vertex lies strictly on the open side.
double OpenArea(const Polygon& g, const Plane& plane, double h) {
  if (g.overflow) return std::numeric_limits<double>::quiet_NaN();
  bool any = false;
  for (size_t j = 0; j < g.n; ++j) any = any || PlaneDistance(g.v[j], plane) < 0;
  if (!any) return 0;
  return Area(Clip(g, plane, nullptr), h);
}

struct Planes {
  bool present = false, slab = false;
  V3 normal{0, 0, 0};
  double low = 0, high = 0;
  V3 velocity{0, 0, 0};
};

V3 ToV3(flo_vec3 v) { return {v.x, v.y, v.z}; }

// ---- the open tank (base/liquid_boundary_math.h:252-292; liquid_boundary_geometry_math_3d.h:420-623) ----

// CheckTankStep: finite walls and piston, ordered x and z walls, a representable piston pose that stays
// strictly inside the high-x wall.
int CheckTankStep(const flo_tank& t, double dt) {
  if (!std::isfinite(t.start_low_x) || !std::isfinite(t.high_x) || !std::isfinite(t.floor_y) ||
      !std::isfinite(t.low_z) || !std::isfinite(t.high_z) || !std::isfinite(t.piston_velocity_x))
    return FLO_INVALID_ARGUMENT;
  if (t.start_low_x >= t.high_x || t.low_z >= t.high_z) return FLO_INVALID_ARGUMENT;
  const double displacement = t.piston_velocity_x * dt;
  const double end_x = std::fma(t.piston_velocity_x, dt, t.start_low_x);
  const double start_width = t.high_x - t.start_low_x, end_width = t.high_x - end_x, width_z = t.high_z - t.low_z;
  if (!std::isfinite(displacement) || !std::isfinite(end_x) || !std::isfinite(start_width) ||
      !std::isfinite(end_width) || !std::isfinite(width_z))
    return FLO_OUT_OF_RANGE;
  if (end_x >= t.high_x) return FLO_FAILED_PRECONDITION;
  return FLO_OK;
}

struct TankLattice {
  double origin, index, width;
};
TankLattice TankAxis(const flo_geometry& g, flo_int3 c, int axis) {
  const double o[3] = {g.world_origin.x, g.world_origin.y, g.world_origin.z};
  const int32_t i[3] = {c.x, c.y, c.z};
  return {o[axis], double(i[axis]), g.cell_width};
}

// TankCellHasResolvedWorldExtent: each axis's world interval is ordered, finite and resolved by the
// spacing toward its interior at both ends.
bool TankCellResolved(const flo_geometry& g, flo_int3 c) {
  const double anchor[3] = {g.world_origin.x + g.cell_width * c.x, g.world_origin.y + g.cell_width * c.y,
                            g.world_origin.z + g.cell_width * c.z};
  for (int a = 0; a < 3; ++a) {
    const TankLattice L = TankAxis(g, c, a);
    const double low = anchor[a], high = L.origin + L.width * (L.index + 1);
    if (!std::isfinite(low) || !std::isfinite(high) || low >= high) return false;
    const double low_spacing = std::nextafter(low, high) - low;
    const double high_spacing = high - std::nextafter(high, low);
    if (low_spacing > L.width || high_spacing > L.width) return false;
  }
  return true;
}

// TankPlaneOffset: world - origin - h index as a bounded sum (in the lattice frame).
bounded::PSum TankPlaneOffset(double world, const TankLattice& L) {
  bounded::PSum o;
  o.Add(world);
  o.Add(-L.origin);
  o.AddProduct(-L.width, L.index);
  return o;
}

int TankPlaneOffsetSign(double world, const TankLattice& L, int& sign) {
  const double translated = world - L.origin;
  if (std::isinf(translated)) {
    sign = translated > 0 ? 1 : -1;
    return FLO_OK;
  }
  if (!TankPlaneOffset(world, L).bounds().sign(&sign)) return FLO_OUT_OF_RANGE;
  return FLO_OK;
}

// TankClipCoordinate: the plane's coordinate in the unit cell, 0 or 1 when it lies beyond a corner.
int TankClipCoordinate(double world, const TankLattice& L, double& out) {
  const double translated = world - L.origin;
  if (std::isinf(translated)) {
    out = translated > 0 ? 1.0 : 0.0;
    return FLO_OK;
  }
  bounded::PSum offset = TankPlaneOffset(world, L);
  const bounded::Bounds low = offset.bounds();
  int low_side = 0;
  if (!low.sign(&low_side)) return FLO_OUT_OF_RANGE;
  if (low_side <= 0) {
    out = 0.0;
    return FLO_OK;
  }
  offset.Add(-L.width);
  int high_side = 0;
  if (!offset.bounds().sign(&high_side)) return FLO_OUT_OF_RANGE;
  if (high_side >= 0) {
    out = 1.0;
    return FLO_OK;
  }
  const double cell = low.estimate / L.width;
  if (cell <= 0 || cell >= 1) return FLO_OUT_OF_RANGE;
  out = cell;
  return FLO_OK;
}

struct TankBounds {
  V3 low{0, 0, 0}, high{1, 1, 1};
};

// LocalTankBounds: the tank's walls in the unit frame of the element at c (a face skips its normal axis).
int LocalTankBounds(const flo_tank& t, const flo_geometry& g, flo_int3 c, int face_axis, TankBounds& out) {
  const double low_world[3] = {t.start_low_x, t.floor_y, t.low_z};
  const double high_world[3] = {t.high_x, std::numeric_limits<double>::infinity(), t.high_z};
  TankBounds b;
  for (int a = 0; a < 3; ++a) {
    if (a == face_axis) continue;
    const TankLattice L = TankAxis(g, c, a);
    double lo = 0, hi = 0;
    const int s1 = TankClipCoordinate(low_world[a], L, lo);
    const int s2 = TankClipCoordinate(high_world[a], L, hi);
    if (s1 || s2) return FLO_OUT_OF_RANGE;
    b.low[a] = lo;
    b.high[a] = hi;
    if (lo == hi && lo > 0 && hi < 1) return FLO_OUT_OF_RANGE;
  }
  out = b;
  return FLO_OK;
}

double PhysicalMeasure(double det, int local_exponent, double h, int dimension);

// TankBoxMeasure: the product of the box extents, scaled before multiplying.
double TankBoxMeasure(const TankBounds& b, double h, int excluded) {
  double mantissa = 1;
  int exponent = 0, dimension = 0;
  for (int a = 0; a < 3; ++a) {
    if (a == excluded) continue;
    const double extent = b.high[a] - b.low[a];
    if (extent <= 0) return 0;
    int e = 0;
    mantissa *= std::frexp(extent, &e);
    exponent += e;
    ++dimension;
  }
  return PhysicalMeasure(mantissa, exponent, h, dimension);
}

Polyhedron ClipSolid(const Polyhedron& in, const Plane& plane);

// ClipLiquidToTank: the liquid inside the tank's walls that cut the unit cell.
Polyhedron ClipLiquidToTank(Polyhedron liquid, const TankBounds& b) {
  for (int a = 0; a < 3; ++a) {
    V3 n{0, 0, 0};
    if (b.low[a] > 0) {
      n[a] = -1;
      liquid = ClipSolid(liquid, Plane{n, -b.low[a], false, a});
    }
    if (b.high[a] < 1) {
      n[a] = 1;
      liquid = ClipSolid(liquid, Plane{n, b.high[a], false, a});
    }
  }
  return liquid;
}

// ComputeTankFace: the aperture of an open face, the piston's rectangle and flux on a closed low-x face.
int TankFace(const flo_tank& t, const flo_geometry& g, flo_int3 l, int axis, const TankBounds& b, double& open_area,
             double& blocked_area, double& flux, double& piston, uint8_t& low_x) {
  const double h = g.cell_width, full_area = h * h;
  const double low_world[3] = {t.start_low_x, t.floor_y, t.low_z};
  const double high_world[3] = {t.high_x, std::numeric_limits<double>::infinity(), t.high_z};
  const TankLattice L = TankAxis(g, l, axis);
  int low_sign = 0, high_sign = 0;
  const int s1 = TankPlaneOffsetSign(low_world[axis], L, low_sign);
  const int s2 = TankPlaneOffsetSign(high_world[axis], L, high_sign);
  if (s1 || s2) return FLO_OUT_OF_RANGE;
  const bool piston_side = axis == 0 && low_sign >= 0;
  const bool open = low_sign < 0 && high_sign > 0;
  double o = 0, f = 0, v = 0;
  uint8_t lx = 0;
  if (open || piston_side) {
    const double transverse = TankBoxMeasure(b, h, axis);
    if (!NormalOrZero(transverse) || transverse > full_area) return FLO_OUT_OF_RANGE;
    if (open) {
      o = transverse;
      for (int a = 0; a < 3; ++a)
        if (a != axis && transverse == full_area && (b.low[a] > 0 || b.high[a] < 1)) return FLO_OUT_OF_RANGE;
    } else if (transverse > 0) {
      lx = 1;
      v = t.piston_velocity_x;
      f = transverse * v;
      if (!NormalOrZero(f) || (v != 0 && f == 0)) return FLO_OUT_OF_RANGE;
    }
  }
  const double blocked = full_area - o;
  if (!NormalOrZero(blocked)) return FLO_OUT_OF_RANGE;
  open_area = o;
  blocked_area = blocked;
  flux = f;
  piston = v;
  low_x = lx;
  return FLO_OK;
}

// ValidateObstacle.
int ValidateObstacle(const flo_obstacle& o, double dt, Planes& out) {
  Planes r;
  if (o.kind == FLO_OBSTACLE_HALFSPACE) {
    r = {true, false, ToV3(o.unit_normal), o.coordinate, o.coordinate, ToV3(o.wall_velocity)};
  } else if (o.kind == FLO_OBSTACLE_SLAB) {
    if (!PositiveFinite(o.half_width)) return FLO_INVALID_ARGUMENT;
    r = {true, true, ToV3(o.unit_normal), o.coordinate - o.half_width, o.coordinate + o.half_width,
         ToV3(o.wall_velocity)};
    if (!std::isfinite(o.coordinate)) return FLO_INVALID_ARGUMENT;
    if (!std::isfinite(r.low) || !std::isfinite(r.high) || r.low >= r.high) return FLO_OUT_OF_RANGE;
  } else if (o.kind == FLO_OBSTACLE_TANK) {
    if (int s = CheckTankStep(o.tank, dt)) return s;
  } else if (o.kind != FLO_OBSTACLE_ABSENT) {
    return FLO_INVALID_ARGUMENT;
  }
  if (r.present && (!Finite3(r.normal) || !Finite3(r.velocity) || !std::isfinite(r.low) ||
                    std::abs(Dot(r.normal, r.normal) - 1.0) > 1e-12))
    return FLO_INVALID_ARGUMENT;
  out = r;
  return FLO_OK;
}

// LocalObstacle: the planes in the unit frame of the element whose low corner is lattice index c.
int LocalPlanes(const flo_obstacle& o, const flo_geometry& g, flo_int3 c, Planes& out) {
  Planes r;
  r.present = o.kind == FLO_OBSTACLE_HALFSPACE || o.kind == FLO_OBSTACLE_SLAB;
  r.slab = o.kind == FLO_OBSTACLE_SLAB;
  r.normal = ToV3(o.unit_normal);
  r.low = o.coordinate - (r.slab ? o.half_width : 0);
  r.high = o.coordinate + (r.slab ? o.half_width : 0);
  r.velocity = ToV3(o.wall_velocity);
  if (!r.present) {
    out = r;
    return FLO_OK;
  }
  const V3 anchor{g.world_origin.x + g.cell_width * c.x, g.world_origin.y + g.cell_width * c.y,
                  g.world_origin.z + g.cell_width * c.z};
  const double projection = Dot(r.normal, anchor);
  r.low = (r.low - projection) / g.cell_width;
  r.high = (r.high - projection) / g.cell_width;
  if (!Finite3(anchor) || !std::isfinite(r.low) || !std::isfinite(r.high) || (r.slab && r.low >= r.high))
    return FLO_OUT_OF_RANGE;
  out = r;
  return FLO_OK;
}

constexpr int kTetrahedra[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7}, {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};

// kCell: liquid volume on each side of the obstacle, by six body-diagonal tetrahedra.
int CellVolumes(const flo_freeze_inputs& in, flo_int3 c, const double corners[8], double& low, double& high) {
  const double h = in.geometry.cell_width;
  const bool tank = in.obstacle.kind == FLO_OBSTACLE_TANK;
  TankBounds tb;
  low = high = 0;
  if (tank) {  // the tank cell's box, or the liquid inside it when not every corner is liquid
    if (!TankCellResolved(in.geometry, c)) return FLO_OUT_OF_RANGE;
    if (int s = LocalTankBounds(in.obstacle.tank, in.geometry, c, -1, tb)) return s;
    const double box = TankBoxMeasure(tb, h, -1);
    if (!NormalOrZero(box)) return FLO_OUT_OF_RANGE;
    if (box == 0) return FLO_OK;
    bool all = true;
    for (int k = 0; k < 8; ++k) all = all && corners[k] < 0;
    if (all) {
      high = box;
      return FLO_OK;
    }
  }
  Planes P;
  if (int s = LocalPlanes(in.obstacle, in.geometry, c, P)) return s;
  Vertex x[8];
  for (int k = 0; k < 8; ++k) {
    for (int a = 0; a < 3; ++a) x[k].p[a] = (k >> a) & 1;
    x[k].phi = corners[k];
  }
  low = high = 0;
  for (int t = 0; t < 6; ++t) {
    const int* q = kTetrahedra[t];
    bool negative = false;
    for (int k = 0; k < 4; ++k) negative = negative || x[q[k]].phi < 0;
    if (!negative) continue;
    const Vertex a = x[q[0]], b = x[q[1]], cc = x[q[2]], d = x[q[3]];
    Polyhedron tet;
    for (const auto& tri : {std::array<Vertex, 3>{a, b, cc}, std::array<Vertex, 3>{a, b, d},
                            std::array<Vertex, 3>{a, cc, d}, std::array<Vertex, 3>{b, cc, d}}) {
      Polygon g;
      for (const Vertex& v : tri) g.push(v);
      tet.push(g);
    }
    const Polyhedron liquid = ClipSolid(tet, {{0, 0, 0}, 0, true});
    if (tank) {
      high += Volume(ClipLiquidToTank(liquid, tb), h);
    } else if (!P.present) {
      high += Volume(liquid, h);
    } else {
      high += Volume(ClipSolid(liquid, {Scale(P.normal, -1), -P.high, false}), h);
      if (P.slab) low += Volume(ClipSolid(liquid, {P.normal, P.low, false}), h);
    }
  }
  return FLO_OK;
}

// kFace: open and blocked area of the face whose low corner is lattice l, normal to axis.
int FaceAreas(const flo_freeze_inputs& in, flo_int3 l, int axis, double& open_area, double& blocked_area) {
  Planes P;
  if (int s = LocalPlanes(in.obstacle, in.geometry, l, P)) return s;
  const double h = in.geometry.cell_width, full_area = h * h;
  int tr[2], n = 0;
  for (int a = 0; a < 3; ++a)
    if (a != axis) tr[n++] = a;
  Vertex x[4]{};
  x[1].p[tr[0]] = 1;
  x[2].p[tr[0]] = 1;
  x[2].p[tr[1]] = 1;
  x[3].p[tr[1]] = 1;
  double blocked = 0, open = 0;
  for (int t = 0; t < 2; ++t) {
    Polygon g;
    g.push(x[0]);
    g.push(x[t + 1]);
    g.push(x[t + 2]);
    if (!P.present) {
      open += Area(g, h);
      continue;
    }
    Polygon b = Clip(g, {P.normal, P.high, false}, nullptr);
    if (P.slab) b = Clip(b, {Scale(P.normal, -1), -P.low, false}, nullptr);
    blocked += Area(b, h);
    open += OpenArea(g, {Scale(P.normal, -1), -P.high, false}, h);
    if (P.slab) open += OpenArea(g, {P.normal, P.low, false}, h);
  }
  open_area = std::min(open, full_area);
  blocked_area = full_area - open_area;
  if (!std::isfinite(open) || !std::isfinite(blocked) ||
      open / full_area > 1 + 64 * std::numeric_limits<double>::epsilon() || !NormalOrZero(open_area) ||
      !NormalOrZero(blocked_area) || (blocked > 0 && blocked_area == 0))
    return FLO_OUT_OF_RANGE;
  return FLO_OK;
}

// ---- sampled solids (base/sampled_solid_liquid_geometry_3d.h) ----

constexpr double kMeasureSlack = 64 * std::numeric_limits<double>::epsilon();

// SampledSolidExteriorPlane (:18-31): the affine extension of the solid samples along a chain of corners
// whose consecutive bits differ in one axis; its distance is minus the interpolated solid distance, so
// clipping keeps the exterior (distance >= 0 outside the solid).
template <size_t N>
Plane SampledExterior(const double (&s)[8], const int (&corners)[N]) {
  Plane p{{0, 0, 0}, s[corners[0]], false};
  for (size_t e = 1; e < N; ++e) {
    const int step = corners[e] ^ corners[e - 1];
    const int axis = (step & 1) != 0 ? 0 : ((step & 2) != 0 ? 1 : 2);
    p.normal[axis] = s[corners[e - 1]] - s[corners[e]];
  }
  return p;
}

// ClipSampledAvailableFace (:33-46): the part of a polygon on the open side of the analytic boundary.
Polygon AvailableFace(const Polygon& g, const Plane* analytic) {
  if (analytic == nullptr || g.overflow) return g;
  bool any = false;
  for (size_t j = 0; j < g.n; ++j) any = any || PlaneDistance(g.v[j], *analytic) < 0;
  if (!any) return Polygon{};
  return Clip(g, *analytic, nullptr);
}

// CheckSampledWetConnectivity (:48-115): wet tetrahedra joined through positive wet openings of the
// triangles they share around the body diagonal must form one piece.
int WetConnectivity(const Vertex (&x)[8], const double (&s)[8], const bool (&wet)[6], double width,
                    const Plane* analytic) {
  const auto wet_corner = [&](int k) {
    return x[k].phi < 0 && s[k] > 0 && (analytic == nullptr || PlaneDistance(x[k], *analytic) < 0);
  };
  if (wet_corner(0) || wet_corner(7)) return FLO_OK;
  int components = 0;
  for (const bool w : wet) components += w ? 1 : 0;
  if (components <= 1) return FLO_OK;
  constexpr int kRing[6] = {0, 1, 4, 5, 3, 2};
  constexpr int kShared[6] = {1, 5, 4, 6, 2, 3};
  int group[6] = {0, 1, 2, 3, 4, 5};
  for (int e = 0; e < 6; ++e) {
    const int first = kRing[e], second = kRing[(e + 1) % 6];
    if (!wet[first] || !wet[second] || group[first] == group[second]) continue;
    const int middle = kShared[e];
    Polygon tri;
    tri.push(x[0]);
    tri.push(x[middle]);
    tri.push(x[7]);
    bool any_liquid = false;
    for (size_t j = 0; j < tri.n; ++j) any_liquid = any_liquid || tri.v[j].phi < 0;
    if (!any_liquid) continue;
    const int corners[3] = {0, middle, 7};
    if (s[0] <= 0 && s[middle] <= 0 && s[7] <= 0) continue;
    const Polygon liquid = Clip(tri, {{0, 0, 0}, 0, true}, nullptr);
    const double wet_area = OpenArea(AvailableFace(liquid, analytic), SampledExterior(s, corners), width);
    if (!NormalOrZero(wet_area)) return FLO_OUT_OF_RANGE;
    if (wet_area > 0) {
      const int old_group = group[second], new_group = group[first];
      for (int& m : group)
        if (m == old_group) m = new_group;
      --components;
    }
  }
  return components == 1 ? FLO_OK : FLO_FAILED_PRECONDITION;
}

// MeasureSampledSolidLiquidCell (:118-218). split_rule (has_sampled_solid; 2 and 3 are relaxations with
// no reference counterpart): 1 the wet-piece refusal everywhere, 2 only where some corner is in the
// solid, 3 nowhere.
int MeasureSampledCell(double width, const double (&liquid)[8], const double (&solid)[8], const Plane* analytic,
                       double& wet_volume, double (&open_area)[6], int split_rule = 1) {
  const bool solid_split_only = split_rule == 2;
  if (!std::isfinite(width) || width <= 0) return FLO_INVALID_ARGUMENT;
  const double full_area = width * width, full_volume = full_area * width;
  if (!std::isnormal(full_area) || !std::isnormal(full_volume)) return FLO_OUT_OF_RANGE;
  Vertex x[8];
  double largest = 0;
  for (int k = 0; k < 8; ++k) {
    if (!std::isfinite(liquid[k]) || !std::isfinite(solid[k])) return FLO_INVALID_ARGUMENT;
    x[k].p = {double(k & 1), double((k >> 1) & 1), double((k >> 2) & 1)};
    x[k].phi = liquid[k];
    largest = std::max(largest, std::abs(solid[k]));
  }
  int exponent = 0;
  (void)std::frexp(largest, &exponent);
  double scaled[8];
  for (int k = 0; k < 8; ++k) {
    scaled[k] = std::ldexp(solid[k], -exponent);
    if (solid[k] != 0 && !std::isnormal(scaled[k])) return FLO_OUT_OF_RANGE;
  }
  double wet = 0;
  bool wet_tet[6] = {};
  for (int t = 0; t < 6; ++t) {
    const int corners[4] = {0, kTetrahedra[t][1], kTetrahedra[t][2], 7};
    bool has_liquid = false, has_exterior = false;
    for (const int k : corners) {
      has_liquid = has_liquid || x[k].phi < 0;
      has_exterior = has_exterior || scaled[k] > 0;
    }
    if (!has_liquid || !has_exterior) continue;
    const Vertex a = x[corners[0]], b = x[corners[1]], c = x[corners[2]], d = x[corners[3]];
    Polyhedron tet;
    for (const auto& tri : {std::array<Vertex, 3>{a, b, c}, std::array<Vertex, 3>{a, b, d},
                            std::array<Vertex, 3>{a, c, d}, std::array<Vertex, 3>{b, c, d}}) {
      Polygon g;
      for (const Vertex& v : tri) g.push(v);
      tet.push(g);
    }
    const Polyhedron liquid_part = ClipSolid(tet, {{0, 0, 0}, 0, true});
    const Polyhedron exterior = ClipSolid(liquid_part, SampledExterior(scaled, corners));
    const double v = analytic == nullptr ? Volume(exterior, width) : Volume(ClipSolid(exterior, *analytic), width);
    wet_tet[t] = v > 0;
    wet += v;
  }
  if (!NormalOrZero(wet) || wet / full_volume > 1 + kMeasureSlack) return FLO_OUT_OF_RANGE;
  wet = std::min(wet, full_volume);
  bool touched = !solid_split_only;
  for (int k = 0; k < 8 && !touched; ++k)
    touched = scaled[k] <= 0 || (analytic != nullptr && PlaneDistance(x[k], *analytic) >= 0);
  if (touched && split_rule != 3)
    if (int c = WetConnectivity(x, scaled, wet_tet, width, analytic)) return c;
  double open[6];
  for (int axis = 0; axis < 3; ++axis) {
    const int bit = 1 << axis;
    for (int side = 0; side < 2; ++side) {
      const int first = side * bit, last = first | (7 ^ bit);
      double o = 0;
      for (int tr = 0; tr < 3; ++tr) {
        if (tr == axis) continue;
        const int middle = first | (1 << tr);
        const int corners[3] = {first, middle, last};
        Polygon tri;
        tri.push(x[first]);
        tri.push(x[middle]);
        tri.push(x[last]);
        o += OpenArea(AvailableFace(tri, analytic), SampledExterior(scaled, corners), width);
      }
      if (!NormalOrZero(o) || o / full_area > 1 + kMeasureSlack) return FLO_OUT_OF_RANGE;
      open[2 * axis + side] = std::min(o, full_area);
    }
  }
  wet_volume = wet;
  for (int k = 0; k < 6; ++k) open_area[k] = open[k];
  return FLO_OK;
}

// MeasureSampledSolidFaceWeights (:220-318): the blocked part of the face's two triangles (solid side
// of each triangle's plane, inside the analytic open side), split over the four corners by the
// centroids of its fan pieces.
int SampledFaceWeights(double width, const double (&solid)[4], const Plane* analytic, double (&out)[4]) {
  if (!std::isfinite(width) || width <= 0) return FLO_INVALID_ARGUMENT;
  const double full_area = width * width;
  if (!std::isnormal(full_area)) return FLO_OUT_OF_RANGE;
  double largest = 0;
  for (const double d : solid) {
    if (!std::isfinite(d)) return FLO_INVALID_ARGUMENT;
    largest = std::max(largest, std::abs(d));
  }
  int exponent = 0;
  (void)std::frexp(largest, &exponent);
  double scaled[8] = {};
  Vertex x[4]{};
  for (int k = 0; k < 4; ++k) {
    scaled[k] = std::ldexp(solid[k], -exponent);
    if (solid[k] != 0 && !std::isnormal(scaled[k])) return FLO_OUT_OF_RANGE;
    x[k].p = {double(k & 1), double((k >> 1) & 1), 0};
  }
  double w[4] = {};
  for (int axis = 0; axis < 2; ++axis) {
    const int middle = 1 << axis;
    const int corners[3] = {0, middle, 3};
    Polygon tri;
    tri.push(x[0]);
    tri.push(x[middle]);
    tri.push(x[3]);
    const Plane exterior = SampledExterior(scaled, corners);
    const Plane interior{Scale(exterior.normal, -1), -exterior.coordinate, false};
    const Polygon blocked = Clip(AvailableFace(tri, analytic), interior, nullptr);
    if (blocked.overflow) return FLO_OUT_OF_RANGE;
    for (size_t v = 1; v + 1 < blocked.n; ++v) {
      Polygon piece;
      piece.push(blocked.v[0]);
      piece.push(blocked.v[v]);
      piece.push(blocked.v[v + 1]);
      const double a = Area(piece, width);
      if (!NormalOrZero(a)) return FLO_OUT_OF_RANGE;
      const V3 centroid = Scale(Add(Add(piece.v[0].p, piece.v[1].p), piece.v[2].p), 1.0 / 3);
      const double along = std::clamp(centroid[axis], 0.0, 1.0);
      const double across = std::clamp(centroid[1 - axis], 0.0, along);
      w[0] += a * (1 - along);
      w[middle] += a * (along - across);
      w[3] += a * across;
    }
  }
  const double measured = w[0] + w[1] + w[2] + w[3];
  if (!NormalOrZero(measured) || measured / full_area > 1 + kMeasureSlack) return FLO_OUT_OF_RANGE;
  for (int k = 0; k < 4; ++k) out[k] = w[k];
  return FLO_OK;
}

// ComputeSampledSolidGeometry (:386-471) for a cell: the wet volume (liquid outside the sampled solid and
// on the open side of an analytic half-space), reported as the high-side volume.
int SampledCellVolume(const flo_freeze_inputs& in, flo_int3 c, const double (&liquid)[8], const double (&solid)[8],
                      double& wet) {
  Planes P;
  if (int s = LocalPlanes(in.obstacle, in.geometry, c, P)) return s;
  if (in.obstacle.kind != FLO_OBSTACLE_ABSENT && in.obstacle.kind != FLO_OBSTACLE_HALFSPACE) return FLO_INVALID_ARGUMENT;
  const Plane analytic{Scale(P.normal, -1), -P.high, false};
  double open[6];
  return MeasureSampledCell(in.geometry.cell_width, liquid, solid, P.present ? &analytic : nullptr, wet, open,
                            in.has_sampled_solid);
}

// ComputeSampledSolidGeometry for a face (the low face of the element at lattice l): open and blocked
// areas, the wall flux of the corner normal velocities through the blocked-area weights plus the
// analytic wall's, and the mean normal velocity over the weighted area. solid and velocity hold the
// four low-side corners (zeros elsewhere).
int SampledFace(const flo_freeze_inputs& in, flo_int3 l, int axis, const double (&solid)[8], const double (&velocity)[8],
                double& open_area, double& blocked_area, double& flux, double& piston) {
  if (axis < 0 || axis >= 3) return FLO_INVALID_ARGUMENT;
  const double width = in.geometry.cell_width;
  const double liquid[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  Planes P;
  if (int s = LocalPlanes(in.obstacle, in.geometry, l, P)) return s;
  if (in.obstacle.kind != FLO_OBSTACLE_ABSENT && in.obstacle.kind != FLO_OBSTACLE_HALFSPACE) return FLO_INVALID_ARGUMENT;
  const Plane analytic3{Scale(P.normal, -1), -P.high, false};
  double wet = 0, open[6];
  if (int s = MeasureSampledCell(width, liquid, solid, P.present ? &analytic3 : nullptr, wet, open)) return s;
  const double face_open = open[2 * axis];
  const double face_blocked = width * width - face_open;
  double face_solid[4], face_velocity[4];
  int fc = 0;
  for (int k = 0; k < 8; ++k) {
    if ((k & (1 << axis)) != 0) continue;
    face_solid[fc] = solid[k];
    face_velocity[fc] = velocity[k];
    if (!std::isfinite(face_velocity[fc])) return FLO_INVALID_ARGUMENT;
    ++fc;
  }
  // MeasureSampledSolidAnalyticFace (:362-384): the analytic wall's blocked area and flux, and its open
  // side in the face's transverse coordinates.
  Plane fluid_side{{0, 0, 0}, -P.high, false};
  double analytic_blocked = 0, analytic_flux = 0;
  if (P.present) {
    double o = 0, b = 0;
    if (int s = FaceAreas(in, l, axis, o, b)) return s;
    analytic_blocked = b;
    analytic_flux = b * P.velocity[axis];
    int t = 0;
    for (int a = 0; a < 3; ++a)
      if (a != axis) fluid_side.normal[t++] = -P.normal[a];
    if (!NormalOrZero(analytic_flux)) return FLO_OUT_OF_RANGE;
  }
  double w[4];
  if (int s = SampledFaceWeights(width, face_solid, P.present ? &fluid_side : nullptr, w)) return s;
  double weighted = 0, f = 0;
  for (int k = 0; k < 4; ++k) {
    weighted += w[k];
    f += w[k] * face_velocity[k];
  }
  f += analytic_flux;
  weighted += analytic_blocked;
  const double slack = 64 * std::numeric_limits<double>::epsilon() * width * width;
  if (!NormalOrZero(face_blocked) || !NormalOrZero(f) || std::abs(face_blocked - weighted) > slack ||
      (weighted > 0 && face_blocked == 0))
    return FLO_OUT_OF_RANGE;
  double v = 0;
  if (weighted > 0) {
    v = f / weighted;
    if (!std::isfinite(v)) return FLO_OUT_OF_RANGE;
  }
  open_area = face_open;
  blocked_area = face_blocked;
  flux = f;
  piston = v;
  return FLO_OK;
}

// kCurvature: centred differences of the 19-point halo (x fastest, index 13 the centre).
int Curvature(const double d[27], double h, double& out) {
  V3 grad{0, 0, 0};
  std::array<V3, 3> H{};
  constexpr int stride[3] = {1, 3, 9};
  for (int a = 0; a < 3; ++a) {
    const double lo = d[13 - stride[a]], hi = d[13 + stride[a]];
    grad[a] = (hi - lo) / (2 * h);
    H[a][a] = ((hi - d[13]) + (lo - d[13])) / h / h;
    for (int o = 0; o < a; ++o) {
      double mixed = 0;
      for (int f = -1; f <= 1; f += 2)
        for (int s = -1; s <= 1; s += 2) mixed += f * s * (d[13 + f * stride[a] + s * stride[o]] / 4);
      H[a][o] = mixed / h / h;
      H[o][a] = H[a][o];
    }
  }
  const double norm = Hypot3(grad[0], grad[1], grad[2]);
  if (!Finite3(grad) || !std::isfinite(norm)) return FLO_OUT_OF_RANGE;
  if (norm < 1e-8) return FLO_FAILED_PRECONDITION;
  const V3 n = Scale(grad, 1 / norm);
  double k = 0;
  for (int a = 0; a < 3; ++a) {
    if (!Finite3(H[a])) return FLO_OUT_OF_RANGE;
    k += (H[a][a] - n[a] * Dot(H[a], n)) / norm;
  }
  if (!std::isfinite(k)) return FLO_OUT_OF_RANGE;
  out = k;
  return FLO_OK;
}

int Offset(flo_int3 c, int axis, int amount, flo_int3& out) {
  int64_t v[3] = {c.x, c.y, c.z};
  v[axis] += amount;
  if (v[axis] < kI32Min || v[axis] > kI32Max) return FLO_OUT_OF_RANGE;
  out = {int32_t(v[0]), int32_t(v[1]), int32_t(v[2])};
  return FLO_OK;
}

int ValidateSamples(const flo_center_distance* s, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i)
    if (!std::isfinite(s[i].signed_distance) || (i > 0 && !BlockLess(s[i - 1].cell_index, s[i].cell_index)))
      return FLO_INVALID_ARGUMENT;
  return FLO_OK;
}

int Lookup(const flo_center_distance* s, uint64_t n, flo_int3 key, double& out) {
  const flo_center_distance* e = s + n;
  const flo_center_distance* f = std::lower_bound(
      s, e, key, [](const flo_center_distance& x, const flo_int3& k) { return BlockLess(x.cell_index, k); });
  if (f == e || !BlockEq(f->cell_index, key)) return FLO_FAILED_PRECONDITION;
  out = f->signed_distance;
  return FLO_OK;
}

int ValidateSolidSamples(const flo_solid_corner* s, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i)
    if (!std::isfinite(s[i].signed_distance) || (i > 0 && !BlockLess(s[i - 1].corner_index, s[i].corner_index)))
      return FLO_INVALID_ARGUMENT;
  return FLO_OK;
}

const flo_solid_corner* LookupSolid(const flo_solid_corner* s, uint64_t n, flo_int3 key) {
  const flo_solid_corner* e = s + n;
  const flo_solid_corner* f = std::lower_bound(
      s, e, key, [](const flo_solid_corner& x, const flo_int3& k) { return BlockLess(x.corner_index, k); });
  return (f == e || !BlockEq(f->corner_index, key)) ? nullptr : f;
}

uint64_t FindCell(const std::vector<flo_frozen_cell>& cells, flo_int3 key) {
  const auto f = std::lower_bound(cells.begin(), cells.end(), key, [](const flo_frozen_cell& c, const flo_int3& k) {
    return BlockLess(c.cell_index, k);
  });
  return (f != cells.end() && BlockEq(f->cell_index, key)) ? uint64_t(f - cells.begin()) : cells.size();
}

}  // namespace

extern "C" int flo_freeze_geometry(const flo_freeze_inputs* inputs, flo_frozen_cell* out_cells,
                                   uint64_t* out_cell_count, flo_frozen_face* out_faces, uint64_t* out_face_count,
                                   double* out_capillary_step_bound) {
  if (!inputs || !out_cell_count || !out_face_count || !out_capillary_step_bound) return FLO_INVALID_ARGUMENT;
  const flo_freeze_inputs& in = *inputs;
  if (!PositiveFinite(in.step_duration) || !PositiveFinite(in.liquid_density) || !std::isfinite(in.surface_tension) ||
      in.surface_tension < 0)
    return FLO_INVALID_ARGUMENT;
  if (in.has_exact_curvature && !std::isfinite(in.exact_curvature)) return FLO_INVALID_ARGUMENT;
  if (in.max_cells && in.cell_count > in.max_cells) return FLO_RESOURCE_EXHAUSTED;
  Planes obstacle;
  if (int s = ValidateObstacle(in.obstacle, in.step_duration, obstacle)) return s;
  const double h = in.geometry.cell_width, full_area = h * h, full_volume = full_area * h;
  if (!std::isnormal(full_area) || !std::isnormal(full_volume) || !std::isnormal(1 / in.liquid_density))
    return FLO_OUT_OF_RANGE;
  if (int s = ValidateSamples(in.center_distances, in.center_count)) return s;
  if (int s = ValidateSamples(in.corner_distances, in.corner_count)) return s;
  const bool sampled = in.has_sampled_solid != 0;
  if (in.has_sampled_solid > 3) return FLO_INVALID_ARGUMENT;
  const bool sampled_halfspace = sampled && in.obstacle.kind == FLO_OBSTACLE_HALFSPACE;
  if (sampled) {
    const uint64_t face_count = 192 * in.block_count;
    bool bad = (!sampled_halfspace && in.obstacle.kind != FLO_OBSTACLE_ABSENT) ||
               (in.face_velocity_count != 0 && in.face_velocity_count != face_count);
    for (uint64_t f = 0; !bad && f < in.face_velocity_count; ++f) bad = !std::isfinite(in.face_normal_velocities[f]);
    if (bad) return FLO_INVALID_ARGUMENT;
    if (int s = ValidateSolidSamples(in.solid_corners, in.solid_corner_count)) return s;
  }

  std::vector<flo_frozen_cell> cells;
  std::vector<flo_frozen_face> faces;
  for (uint64_t i = 0; i < in.cell_count; ++i) {
    const flo_int3 c = in.cell_indices[i];
    if (i > 0 && !BlockLess(in.cell_indices[i - 1], c)) return FLO_INVALID_ARGUMENT;
    double center = 0;
    if (int s = Lookup(in.center_distances, in.center_count, c, center)) return s;
    double corners[8], solid[8];
    for (int k = 0; k < 8; ++k) {
      flo_int3 q = c;
      for (int a = 0; a < 3; ++a)
        if (int s = Offset(q, a, (k >> a) & 1, q)) return s;
      if (int s = Lookup(in.corner_distances, in.corner_count, q, corners[k])) return s;
      if (sampled) {
        const flo_solid_corner* f = LookupSolid(in.solid_corners, in.solid_corner_count, q);
        if (!f) return FLO_FAILED_PRECONDITION;
        solid[k] = f->signed_distance;
      }
    }
    double low = 0, high = 0;
    if (sampled) {
      if (int s = SampledCellVolume(in, c, corners, solid, high)) return s;
    } else if (int s = CellVolumes(in, c, corners, low, high)) {
      return s;
    }
    if (low > 0 && high > 0 && center < 0) return FLO_FAILED_PRECONDITION;
    const double unclamped = low + high;
    if (!std::isfinite(unclamped) || unclamped < 0 ||
        unclamped / full_volume > 1 + 64 * std::numeric_limits<double>::epsilon())
      return FLO_OUT_OF_RANGE;
    const double volume = std::min(full_volume, unclamped);
    if (!NormalOrZero(volume)) return FLO_OUT_OF_RANGE;
    if (center >= 0 || volume == 0) continue;
    flo_frozen_cell fc;
    std::memset(&fc, 0, sizeof fc);
    fc.cell_index = c;
    fc.fluid_volume = volume;
    fc.center_distance = center;
    cells.push_back(fc);
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side) {
        flo_int3 l{};
        if (int s = Offset(c, a, side, l)) return s;
        uint64_t slot = 0;
        if (int s = flo_face_slot(in.blocks, in.block_count, flo_face_index{uint8_t(a), l}, &slot)) return s;
        flo_frozen_face f;
        std::memset(&f, 0, sizeof f);
        f.face_index.axis = uint8_t(a);
        f.face_index.lattice = l;
        f.face_slot = slot;
        faces.push_back(f);
      }
  }
  std::sort(faces.begin(), faces.end(),
            [](const flo_frozen_face& a, const flo_frozen_face& b) { return a.face_slot < b.face_slot; });
  faces.erase(std::unique(faces.begin(), faces.end(),
                          [](const flo_frozen_face& a, const flo_frozen_face& b) { return a.face_slot == b.face_slot; }),
              faces.end());
  if (in.max_faces && faces.size() > in.max_faces) return FLO_RESOURCE_EXHAUSTED;
  for (flo_frozen_face& f : faces) {
    if (in.obstacle.kind == FLO_OBSTACLE_TANK) {
      TankBounds tb;
      if (int s = LocalTankBounds(in.obstacle.tank, in.geometry, f.face_index.lattice, f.face_index.axis, tb)) return s;
      double flux = 0, piston = 0;
      if (int s = TankFace(in.obstacle.tank, in.geometry, f.face_index.lattice, f.face_index.axis, tb, f.open_area,
                           f.blocked_area, flux, piston, f.is_low_x_piston_face))
        return s;
      f.wall_normal_velocity = piston;
      f.prescribed_wall_flux = flux;
      f.has_prescribed_wall_flux = 1;
      continue;
    }
    if (!sampled) {
      if (int s = FaceAreas(in, f.face_index.lattice, f.face_index.axis, f.open_area, f.blocked_area)) return s;
      f.wall_normal_velocity = obstacle.velocity[f.face_index.axis];
      continue;
    }
    const int axis = f.face_index.axis;
    double solid[8] = {}, velocity[8] = {};
    for (int k = 0; k < 8; ++k) {
      if ((k & (1 << axis)) != 0) continue;
      flo_int3 q = f.face_index.lattice;
      for (int a = 0; a < 3; ++a)
        if (int s = Offset(q, a, (k >> a) & 1, q)) return s;
      const flo_solid_corner* sample = LookupSolid(in.solid_corners, in.solid_corner_count, q);
      if (!sample) return FLO_FAILED_PRECONDITION;
      solid[k] = sample->signed_distance;
      const double wv[3] = {sample->wall_velocity.x, sample->wall_velocity.y, sample->wall_velocity.z};
      if (in.face_velocity_count == 0) velocity[k] = wv[axis];
      else if (sampled_halfspace) velocity[k] = in.face_normal_velocities[f.face_slot];
    }
    double flux = 0, piston = 0;
    if (int s = SampledFace(in, f.face_index.lattice, axis, solid, velocity, f.open_area, f.blocked_area, flux, piston))
      return s;
    if (sampled_halfspace || in.face_velocity_count == 0) {
      f.wall_normal_velocity = piston;
      f.prescribed_wall_flux = flux;
      f.has_prescribed_wall_flux = 1;
    } else {
      f.wall_normal_velocity = in.face_normal_velocities[f.face_slot];
    }
  }
  // Cells on an open face whose other side is air (a centre sample >= 0) need curvature.
  std::vector<uint8_t> needs(cells.size(), 0);
  for (const flo_frozen_face& f : faces) {
    if (f.open_area == 0 || in.surface_tension == 0) continue;
    const flo_int3 head = f.face_index.lattice;
    flo_int3 tail{};
    if (int s = Offset(head, f.face_index.axis, -1, tail)) return s;
    const uint64_t hc = FindCell(cells, head), tc = FindCell(cells, tail);
    const bool has_head = hc != cells.size(), has_tail = tc != cells.size();
    if (has_head && has_tail) continue;
    double air = 0;
    if (int s = Lookup(in.center_distances, in.center_count, has_head ? tail : head, air)) return s;
    if (air >= 0) needs[has_head ? hc : tc] = 1;
  }
  const bool capillary = std::any_of(needs.begin(), needs.end(), [](uint8_t x) { return x != 0; });
  if (in.has_exact_curvature) {
    if (capillary && in.exact_curvature != 0 && !std::isnormal(in.surface_tension * in.exact_curvature))
      return FLO_OUT_OF_RANGE;
    for (size_t i = 0; i < cells.size(); ++i)
      if (needs[i]) cells[i].curvature = in.exact_curvature;
  } else {
    for (size_t i = 0; i < cells.size(); ++i) {
      if (!needs[i]) continue;
      double d[27] = {};
      for (int z = -1; z <= 1; ++z)
        for (int y = -1; y <= 1; ++y)
          for (int x = -1; x <= 1; ++x) {
            if (x != 0 && y != 0 && z != 0) continue;
            flo_int3 q = cells[i].cell_index;
            const int o[3] = {x, y, z};
            for (int a = 0; a < 3; ++a)
              if (int s = Offset(q, a, o[a], q)) return s;
            if (int s = Lookup(in.center_distances, in.center_count, q, d[(z + 1) * 9 + (y + 1) * 3 + x + 1]))
              return s;
          }
      if (int s = Curvature(d, h, cells[i].curvature)) return s;
    }
  }
  double bound = std::numeric_limits<double>::infinity();
  if (capillary) {
    constexpr double kPi = 3.141592653589793238462643383279502884;
    const double square = in.liquid_density * full_volume / (4 * kPi * in.surface_tension);
    if (!std::isnormal(square)) return FLO_OUT_OF_RANGE;
    bound = std::sqrt(square);
    if (in.step_duration > bound) return FLO_FAILED_PRECONDITION;
  }
  if (!cells.empty()) std::memcpy(out_cells, cells.data(), cells.size() * sizeof(flo_frozen_cell));
  if (!faces.empty()) std::memcpy(out_faces, faces.data(), faces.size() * sizeof(flo_frozen_face));
  *out_cell_count = cells.size();
  *out_face_count = faces.size();
  *out_capillary_step_bound = bound;
  return FLO_OK;
}

extern "C" int flo_measure_sampled_cell(double width, const double liquid[8], const double solid[8],
                                        const double* analytic, double* out_wet_volume, double out_open_area[6]) {
  double l[8], so[8], open[6], wet = 0;
  for (int k = 0; k < 8; ++k) {
    l[k] = liquid[k];
    so[k] = solid[k];
  }
  Plane p{};
  if (analytic) p = Plane{{analytic[0], analytic[1], analytic[2]}, analytic[3], false};
  const int s = MeasureSampledCell(width, l, so, analytic ? &p : nullptr, wet, open);
  if (s) return s;
  *out_wet_volume = wet;
  for (int k = 0; k < 6; ++k) out_open_area[k] = open[k];
  return FLO_OK;
}

extern "C" int flo_sampled_face_weights(double width, const double solid[4], const double* analytic,
                                        double out_weights[4]) {
  double so[4] = {solid[0], solid[1], solid[2], solid[3]}, w[4];
  Plane p{};
  if (analytic) p = Plane{{analytic[0], analytic[1], analytic[2]}, analytic[3], false};
  const int s = SampledFaceWeights(width, so, analytic ? &p : nullptr, w);
  if (s) return s;
  for (int k = 0; k < 4; ++k) out_weights[k] = w[k];
  return FLO_OK;
}
