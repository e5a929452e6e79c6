// Oracle: MAC pressure projection (test infrastructure only).
//
// Preparation restates native/src/mac_pressure_3d_arithmetic.h:110-399 (FreezeCell,
// FreezeLinkFromIncidence, AssembleIntegratedRhsInto with Neumaier sums, SelectGaugesInto)
// and mac_pressure_workspace_3d.cc:68-166 (Prepare). Evaluate restates
// mac_pressure_3d_arithmetic.h:404-503 and mac_pressure_workspace_3d.cc:399-461. The
// stopping rule restates mac_pressure_3d_multigrid.cc:303-367. The outer mixed-precision
// PCG follows mac_pressure_workspace_3d.cc:544-658 (FP32 vectors, FP64 dots, FP64 restart).
//
// The preconditioner is the geometric V-cycle the HIP path runs (2x2x2 aggregation with
// piecewise-constant prolongation, restriction R = P^T, Galerkin crossing-sum coarse
// operators, red-black Gauss-Seidel smoothing), so the oracle times and checks the same
// algorithm. The reference's own preconditioner (aggregation + weighted Jacobi + dense
// Cholesky, mac_pressure_workspace_3d.cc:298-397) converges to the same acceptance gates;
// parity is judged on the accepted projection, not iterate by iterate.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "grid_internal.h"

using namespace flo;

namespace {

constexpr double kEps = std::numeric_limits<double>::epsilon();
constexpr double kCompatibilityRoundoff = 32.0 * kEps;
constexpr double kIdentityRoundoff = 64.0 * kEps;
constexpr size_t kNone = SIZE_MAX;

bool FinitePositive(double v) { return std::isfinite(v) && v > 0.0; }

// mac_pressure_3d_arithmetic.h:33-49
bool Representable(double a, double b, double p) {
  if (!std::isfinite(p)) return false;
  if (a == 0.0 || b == 0.0 || std::fabs(p) >= std::numeric_limits<double>::min()) return true;
  if (p == 0.0) return false;
  int ea = 0, eb = 0;
  const double ma = std::frexp(a, &ea);
  const double mb = std::frexp(b, &eb);
  const double np = std::scalbn(p, -(ea + eb));
  return std::fma(ma, mb, -np) == 0.0;
}

// Neumaier summation (mac_pressure_3d_arithmetic.h:57-74)
struct Comp {
  double sum = 0.0, corr = 0.0;
  void Add(double t) {
    const double n = sum + t;
    if (std::fabs(sum) >= std::fabs(t))
      corr += (sum - n) + t;
    else
      corr += (t - n) + sum;
    sum = n;
  }
  double Total() const { return sum + corr; }
};
struct FluxRow {
  Comp flux, abs;
  void Add(double t) {
    flux.Add(t);
    abs.Add(std::fabs(t));
  }
};

struct Cell {
  flo_int3 index;
  double volume, target_flux, wall_flux, compliance, previous;
};
struct Link {
  uint64_t slot;
  size_t tail, head;  // kNone when absent
  double open_area, blocked_area, scale, conductance, predicted, wall_flux;
  int kind;
  double lift;
};

struct System {
  double dt = 0;
  std::vector<Cell> cells;
  std::vector<Link> links;
  std::vector<double> rhs, rhs_abs;
  std::vector<size_t> component, gauges;
};

bool CellLess(const flo_int3& a, const flo_int3& b) { return BlockLess(a, b); }

size_t FindCell(const std::vector<Cell>& cells, flo_int3 idx) {
  auto it = std::lower_bound(cells.begin(), cells.end(), idx,
                             [](const Cell& c, const flo_int3& k) { return CellLess(c.index, k); });
  if (it == cells.end() || !BlockEq(it->index, idx)) return kNone;
  return static_cast<size_t>(it - cells.begin());
}

template <typename F>
void ForIncident(const Link& l, F f) {
  if (l.tail != kNone) f(l.tail, 1.0);
  if (l.head != kNone) f(l.head, -1.0);
}

int FreezeCell(const flo_pressure_cell& c, double dt, double full_volume, Cell& out) {
  if (!FinitePositive(c.fluid_volume) || c.fluid_volume > full_volume) return FLO_INVALID_ARGUMENT;
  if (!std::isfinite(c.target_divergence) || !std::isfinite(c.previous_pressure) ||
      !std::isfinite(c.integrated_inward_wall_flux))
    return FLO_INVALID_ARGUMENT;
  double compliance = 0.0;
  if (c.bulk_modulus != std::numeric_limits<double>::infinity()) {
    if (!FinitePositive(c.bulk_modulus)) return FLO_INVALID_ARGUMENT;
    const double kdt = c.bulk_modulus * dt;
    if (!FinitePositive(kdt)) return FLO_OUT_OF_RANGE;
    compliance = c.fluid_volume / kdt;
    if (!FinitePositive(compliance)) return FLO_OUT_OF_RANGE;
  }
  const double tf = c.fluid_volume * c.target_divergence;
  if (!Representable(c.fluid_volume, c.target_divergence, tf)) return FLO_OUT_OF_RANGE;
  out = {c.cell_index, c.fluid_volume, tf, c.integrated_inward_wall_flux, compliance, c.previous_pressure};
  return FLO_OK;
}

// mac_pressure_3d_arithmetic.h:174-245
int FreezeLink(const flo_pressure_face& f, uint64_t slot, double full_area, double dt, size_t tail, size_t head,
               Link& out) {
  if (!std::isfinite(f.open_area) || f.open_area < 0.0 || f.open_area > full_area) return FLO_INVALID_ARGUMENT;
  if (!std::isfinite(f.inverse_density) || !std::isfinite(f.pressure_gradient_distance) ||
      !std::isfinite(f.predicted_normal_velocity) || !std::isfinite(f.wall_normal_velocity))
    return FLO_INVALID_ARGUMENT;
  const bool both = tail != kNone && head != kNone;
  const bool one = (tail != kNone) != (head != kNone);
  if (f.link_kind == 1) {
    if (!both || !std::isfinite(f.lift)) return FLO_INVALID_ARGUMENT;
  } else if (f.link_kind == 2) {
    if (!one || f.open_area == 0.0 || !std::isfinite(f.lift)) return FLO_INVALID_ARGUMENT;
  } else if (f.link_kind == 0) {
    if (f.open_area != 0.0) return FLO_INVALID_ARGUMENT;
  } else {
    return FLO_INVALID_ARGUMENT;
  }
  double scale = 0.0, conductance = 0.0;
  if (f.open_area > 0.0) {
    if (f.inverse_density <= 0.0 || f.pressure_gradient_distance <= 0.0) return FLO_INVALID_ARGUMENT;
    const double mobility = dt * f.inverse_density;
    scale = mobility / f.pressure_gradient_distance;
    conductance = f.open_area * scale;
    if (!FinitePositive(scale) || !FinitePositive(conductance)) return FLO_OUT_OF_RANGE;
  }
  const double blocked = full_area - f.open_area;
  double wall_flux = 0.0;
  if (f.has_prescribed_wall_flux) {
    wall_flux = f.prescribed_wall_flux;
    if (!std::isfinite(wall_flux) || (blocked == 0 && wall_flux != 0)) return FLO_INVALID_ARGUMENT;
  } else {
    wall_flux = blocked * f.wall_normal_velocity;
    if (!Representable(blocked, f.wall_normal_velocity, wall_flux)) return FLO_OUT_OF_RANGE;
  }
  out = {slot, tail, head, f.open_area, blocked, scale, conductance, f.predicted_normal_velocity,
         wall_flux, f.link_kind, f.link_kind == 0 ? 0.0 : f.lift};
  return FLO_OK;
}

size_t Root(std::vector<size_t>& parent, size_t c) {
  while (parent[c] != c) c = parent[c];
  return c;
}

// mac_pressure_workspace_3d.cc:68-166
int Prepare(const flo_geometry& g, const flo_int3* blocks, size_t nb, double dt, const flo_pressure_cell* in_cells,
            size_t nc, const flo_pressure_face* in_faces, size_t nf, System& sys) {
  if (nc == 0 || nf == 0 || !FinitePositive(dt)) return FLO_INVALID_ARGUMENT;
  const double h = g.cell_width;
  const double area = h * h;
  const double volume = area * h;
  if (!FinitePositive(area) || !FinitePositive(volume)) return FLO_OUT_OF_RANGE;
  sys.dt = dt;
  sys.cells.resize(nc);
  std::vector<uint64_t> incident;
  incident.reserve(6 * nc);
  for (size_t i = 0; i < nc; ++i) {
    const auto& c = in_cells[i];
    if (i != 0 && !CellLess(in_cells[i - 1].cell_index, c.cell_index)) return FLO_INVALID_ARGUMENT;
    if (int s = FreezeCell(c, dt, volume, sys.cells[i])) return s;
    for (int a = 0; a < 3; ++a) {
      if (Get(c.cell_index, a) == kI32Max) return FLO_OUT_OF_RANGE;
      for (int side = 0; side < 2; ++side) {
        flo_int3 l = c.cell_index;
        if (a == 0) l.x += side;
        if (a == 1) l.y += side;
        if (a == 2) l.z += side;
        uint64_t slot;
        if (int s = FaceSlot(blocks, nb, a, l, slot)) return s;
        incident.push_back(slot);
      }
    }
  }
  std::sort(incident.begin(), incident.end());
  incident.erase(std::unique(incident.begin(), incident.end()), incident.end());
  if (incident.size() != nf) return FLO_INVALID_ARGUMENT;
  sys.links.resize(nf);
  for (size_t i = 0; i < nf; ++i) {
    const auto& f = in_faces[i];
    uint64_t slot;
    if (int s = FaceSlot(blocks, nb, f.axis, f.lattice, slot)) return s;
    if (slot != incident[i]) return FLO_INVALID_ARGUMENT;
    flo_int3 tail = f.lattice;
    const int64_t tc = static_cast<int64_t>(Get(f.lattice, f.axis)) - 1;
    size_t tail_ordinal = kNone;
    if (tc >= kI32Min) {
      if (f.axis == 0) tail.x = static_cast<int32_t>(tc);
      if (f.axis == 1) tail.y = static_cast<int32_t>(tc);
      if (f.axis == 2) tail.z = static_cast<int32_t>(tc);
      tail_ordinal = FindCell(sys.cells, tail);
    }
    if (int s = FreezeLink(f, slot, area, dt, tail_ordinal, FindCell(sys.cells, f.lattice), sys.links[i])) return s;
  }
  // AssembleIntegratedRhsInto (mac_pressure_3d_arithmetic.h:263-313)
  std::vector<FluxRow> rows(nc);
  for (size_t c = 0; c < nc; ++c) {
    rows[c].Add(sys.cells[c].target_flux);
    rows[c].Add(sys.cells[c].wall_flux);
    const double pc = sys.cells[c].compliance * sys.cells[c].previous;
    if (!Representable(sys.cells[c].compliance, sys.cells[c].previous, pc)) return FLO_OUT_OF_RANGE;
    rows[c].Add(pc);
  }
  for (const auto& l : sys.links) {
    const double pf = l.open_area * l.predicted;
    if (!Representable(l.open_area, l.predicted, pf)) return FLO_OUT_OF_RANGE;
    ForIncident(l, [&](size_t c, double s) {
      rows[c].Add(-s * pf);
      rows[c].Add(-s * l.wall_flux);
    });
    if (l.kind == 1) {
      const double jf = l.conductance * l.lift;
      if (!Representable(l.conductance, l.lift, jf)) return FLO_OUT_OF_RANGE;
      ForIncident(l, [&](size_t c, double s) { rows[c].Add(-s * jf); });
    } else if (l.kind == 2) {
      const double ef = l.conductance * l.lift;
      if (!Representable(l.conductance, l.lift, ef)) return FLO_OUT_OF_RANGE;
      ForIncident(l, [&](size_t c, double) { rows[c].Add(ef); });
    }
  }
  sys.rhs.resize(nc);
  sys.rhs_abs.resize(nc);
  for (size_t c = 0; c < nc; ++c) {
    sys.rhs[c] = rows[c].flux.Total();
    sys.rhs_abs[c] = rows[c].abs.Total();
    if (!std::isfinite(sys.rhs[c]) || !std::isfinite(sys.rhs_abs[c])) return FLO_OUT_OF_RANGE;
  }
  // SelectGaugesInto (mac_pressure_3d_arithmetic.h:328-399)
  auto& comp = sys.component;
  comp.resize(nc);
  std::iota(comp.begin(), comp.end(), size_t{0});
  for (const auto& l : sys.links)
    if (l.conductance > 0.0 && l.tail != kNone && l.head != kNone) {
      const size_t a = Root(comp, l.tail), b = Root(comp, l.head);
      comp[std::max(a, b)] = std::min(a, b);
    }
  std::vector<char> anchored(nc, 0);
  std::vector<FluxRow> balance(nc);
  for (size_t c = 0; c < nc; ++c) {
    const size_t r = Root(comp, c);
    comp[c] = r;
    if (sys.cells[c].compliance > 0.0) anchored[r] = 1;
  }
  for (const auto& l : sys.links)
    if (l.kind == 2) ForIncident(l, [&](size_t c, double) { anchored[comp[c]] = 1; });
  for (size_t c = 0; c < nc; ++c)
    if (!anchored[comp[c]]) {
      balance[comp[c]].Add(sys.cells[c].target_flux);
      balance[comp[c]].Add(sys.cells[c].wall_flux);
    }
  for (const auto& l : sys.links) {
    const bool stays = l.tail != kNone && l.head != kNone && comp[l.tail] == comp[l.head];
    if (stays) continue;
    ForIncident(l, [&](size_t c, double s) {
      if (!anchored[comp[c]]) balance[comp[c]].Add(-s * l.wall_flux);
    });
  }
  sys.gauges.clear();
  for (size_t c = 0; c < nc; ++c) {
    if (comp[c] != c || anchored[c]) continue;
    const double b = balance[c].flux.Total(), a = balance[c].abs.Total();
    if (!std::isfinite(b) || !std::isfinite(a)) return FLO_OUT_OF_RANGE;
    if (std::fabs(b) > kCompatibilityRoundoff * a) return FLO_FAILED_PRECONDITION;
    sys.gauges.push_back(c);
  }
  return FLO_OK;
}

struct Diag {
  double l2 = 0, max_res = 0, max_defect = 0, defect_l2 = 0;
};

// Physical defect rows only (mac_pressure_3d_arithmetic.h:441-503).
int PhysicalRows(const System& sys, const std::vector<double>& p, std::vector<double>* vel, std::vector<FluxRow>& rows) {
  const size_t nc = sys.cells.size();
  rows.assign(nc, FluxRow{});
  for (size_t c = 0; c < nc; ++c) {
    const auto& k = sys.cells[c];
    rows[c].Add(-k.target_flux);
    rows[c].Add(-k.wall_flux);
    if (k.compliance > 0.0) {
      const double dp = p[c] - k.previous;
      const double f = k.compliance * dp;
      if (!Representable(k.compliance, dp, f)) return FLO_OUT_OF_RANGE;
      rows[c].Add(f);
    }
  }
  for (size_t i = 0; i < sys.links.size(); ++i) {
    const auto& l = sys.links[i];
    double u = l.predicted;
    if (l.open_area > 0.0) {
      const double ext = l.kind == 2 ? l.lift : 0.0;
      const double jump = l.kind == 1 ? l.lift : 0.0;
      const double pt = l.tail != kNone ? p[l.tail] : ext;
      const double ph = l.head != kNone ? p[l.head] : ext;
      const double dp = pt - ph + jump;
      const double du = l.scale * dp;
      if (!Representable(l.scale, dp, du)) return FLO_OUT_OF_RANGE;
      u += du;
    }
    if (!std::isfinite(u)) return FLO_OUT_OF_RANGE;
    if (vel) (*vel)[i] = u;
    const double ff = l.open_area * u;
    if (!Representable(l.open_area, u, ff)) return FLO_OUT_OF_RANGE;
    ForIncident(l, [&](size_t c, double s) {
      rows[c].Add(s * ff);
      rows[c].Add(s * l.wall_flux);
    });
  }
  return FLO_OK;
}

// mac_pressure_workspace_3d.cc:399-461
int Evaluate(const System& sys, const std::vector<double>& p, std::vector<double>& vel, std::vector<double>& res,
             std::vector<double>& def, Diag& d) {
  const size_t nc = sys.cells.size();
  for (double v : p)
    if (!std::isfinite(v)) return FLO_INVALID_ARGUMENT;
  std::vector<FluxRow> op(nc);
  for (size_t c = 0; c < nc; ++c) {
    const double f = sys.cells[c].compliance * p[c];
    if (!Representable(sys.cells[c].compliance, p[c], f)) return FLO_OUT_OF_RANGE;
    op[c].Add(f);
  }
  for (const auto& l : sys.links) {
    if (l.conductance == 0.0) continue;
    const double pt = l.tail != kNone ? p[l.tail] : 0.0;
    const double ph = l.head != kNone ? p[l.head] : 0.0;
    const double dp = pt - ph;
    const double f = l.conductance * dp;
    if (!Representable(l.conductance, dp, f)) return FLO_OUT_OF_RANGE;
    ForIncident(l, [&](size_t c, double s) { op[c].Add(s * f); });
  }
  std::vector<FluxRow> phys;
  vel.resize(sys.links.size());
  if (int s = PhysicalRows(sys, p, &vel, phys)) return s;
  d = {};
  res.resize(nc);
  def.resize(nc);
  for (size_t c = 0; c < nc; ++c) {
    const double r = sys.rhs[c] - op[c].flux.Total();
    const double V = sys.cells[c].volume;
    const double vr = std::fabs(r) / V;
    const double df = phys[c].flux.Total() / V;
    const double ie = V * df + r;
    const double terms = sys.rhs_abs[c] + op[c].abs.Total() + phys[c].abs.Total();
    if (!std::isfinite(vr) || !std::isfinite(df) || !std::isfinite(ie) || !std::isfinite(terms))
      return FLO_OUT_OF_RANGE;
    if (std::fabs(ie) > kIdentityRoundoff * terms) return FLO_INTERNAL;
    res[c] = r;
    def[c] = df;
    d.l2 = std::hypot(d.l2, r);
    d.defect_l2 = std::hypot(d.defect_l2, df);
    d.max_res = std::max(d.max_res, vr);
    d.max_defect = std::max(d.max_defect, std::fabs(df));
  }
  if (!std::isfinite(d.l2)) return FLO_OUT_OF_RANGE;
  return FLO_OK;
}

// ---------------- geometric V-cycle ----------------

template <typename T>
struct Level {
  size_t n = 0;
  std::vector<flo_int3> coord;
  std::vector<uint8_t> color;
  std::vector<std::array<int32_t, 6>> nbr;  // -x,+x,-y,+y,-z,+z
  std::vector<std::array<T, 6>> k;
  std::vector<T> anchor, diag;
  std::vector<int32_t> parent;  // into the next level
  std::vector<T> x, b, r;
};

struct KeyHash {
  size_t operator()(const flo_int3& c) const {
    return (static_cast<size_t>(static_cast<uint32_t>(c.x)) * 73856093u) ^
           (static_cast<size_t>(static_cast<uint32_t>(c.y)) * 19349663u) ^
           (static_cast<size_t>(static_cast<uint32_t>(c.z)) * 83492791u);
  }
};
struct KeyEq {
  bool operator()(const flo_int3& a, const flo_int3& b) const { return BlockEq(a, b); }
};
using CoordMap = std::unordered_map<flo_int3, int32_t, KeyHash, KeyEq>;

int32_t FloorHalf(int32_t v) { return v >= 0 ? v / 2 : -((-(v + 1)) / 2) - 1; }

// Fine edges and anchors in FP64 as built from the frozen links (anchors: compliance,
// Dirichlet conductances and conductances to zero-correction gauge cells).
struct FineGraph {
  std::vector<size_t> unknown_to_cell;
  std::vector<size_t> cell_to_unknown;
  std::vector<flo_int3> coord;
  std::vector<double> anchor;
  std::vector<std::array<double, 6>> k;
  std::vector<std::array<int32_t, 6>> nbr;
};

FineGraph BuildFine(const System& sys) {
  FineGraph g;
  const size_t nc = sys.cells.size();
  std::vector<char> is_gauge(nc, 0);
  for (size_t c : sys.gauges) is_gauge[c] = 1;
  g.cell_to_unknown.assign(nc, kNone);
  for (size_t c = 0; c < nc; ++c)
    if (!is_gauge[c]) {
      g.cell_to_unknown[c] = g.unknown_to_cell.size();
      g.unknown_to_cell.push_back(c);
      g.coord.push_back(sys.cells[c].index);
    }
  const size_t n = g.unknown_to_cell.size();
  g.anchor.assign(n, 0.0);
  g.k.assign(n, {0, 0, 0, 0, 0, 0});
  g.nbr.assign(n, {-1, -1, -1, -1, -1, -1});
  for (size_t u = 0; u < n; ++u) g.anchor[u] = sys.cells[g.unknown_to_cell[u]].compliance;
  for (const auto& l : sys.links) {
    if (l.conductance <= 0.0) continue;
    const size_t ut = l.tail != kNone ? g.cell_to_unknown[l.tail] : kNone;
    const size_t uh = l.head != kNone ? g.cell_to_unknown[l.head] : kNone;
    // Recover the face axis from the slot: (slot % 192) / 64.
    const int axis = static_cast<int>((l.slot % 192) / 64);
    if (ut != kNone && uh != kNone) {
      g.k[ut][2 * axis + 1] = l.conductance;
      g.nbr[ut][2 * axis + 1] = static_cast<int32_t>(uh);
      g.k[uh][2 * axis] = l.conductance;
      g.nbr[uh][2 * axis] = static_cast<int32_t>(ut);
    } else {
      if (ut != kNone) g.anchor[ut] += l.conductance;
      if (uh != kNone) g.anchor[uh] += l.conductance;
    }
  }
  return g;
}

template <typename T>
void FinishLevel(Level<T>& L) {
  L.color.resize(L.n);
  L.diag.resize(L.n);
  for (size_t i = 0; i < L.n; ++i) {
    L.color[i] = static_cast<uint8_t>((L.coord[i].x + L.coord[i].y + L.coord[i].z) & 1);
    double d = static_cast<double>(L.anchor[i]);
    for (int f = 0; f < 6; ++f) d += static_cast<double>(L.k[i][f]);
    L.diag[i] = static_cast<T>(d);
  }
  L.x.assign(L.n, T(0));
  L.b.assign(L.n, T(0));
  L.r.assign(L.n, T(0));
}

template <typename T>
std::vector<Level<T>> BuildHierarchy(const FineGraph& g, int levels, double coarse_scale) {
  std::vector<Level<T>> H(1);
  Level<T>& L0 = H[0];
  L0.n = g.coord.size();
  L0.coord = g.coord;
  L0.nbr = g.nbr;
  L0.k.resize(L0.n);
  L0.anchor.resize(L0.n);
  for (size_t i = 0; i < L0.n; ++i) {
    for (int f = 0; f < 6; ++f) L0.k[i][f] = static_cast<T>(g.k[i][f]);
    L0.anchor[i] = static_cast<T>(g.anchor[i]);
  }
  FinishLevel(L0);
  for (int lv = 1; lv < levels; ++lv) {
    Level<T>& F = H[lv - 1];
    CoordMap map;
    Level<T> C;
    F.parent.resize(F.n);
    for (size_t i = 0; i < F.n; ++i) {
      const flo_int3 pc{FloorHalf(F.coord[i].x), FloorHalf(F.coord[i].y), FloorHalf(F.coord[i].z)};
      auto it = map.find(pc);
      int32_t id;
      if (it == map.end()) {
        id = static_cast<int32_t>(C.coord.size());
        map.emplace(pc, id);
        C.coord.push_back(pc);
      } else {
        id = it->second;
      }
      F.parent[i] = id;
    }
    C.n = C.coord.size();
    if (C.n == F.n || C.n == 0) break;
    std::vector<double> anchor(C.n, 0.0);
    std::vector<std::array<double, 6>> k(C.n, {0, 0, 0, 0, 0, 0});
    C.nbr.assign(C.n, {-1, -1, -1, -1, -1, -1});
    for (size_t i = 0; i < F.n; ++i) {
      const int32_t I = F.parent[i];
      anchor[I] += coarse_scale * static_cast<double>(F.anchor[i]);
      for (int axis = 0; axis < 3; ++axis) {
        const int32_t j = F.nbr[i][2 * axis + 1];
        if (j < 0) continue;
        const int32_t J = F.parent[j];
        if (J == I) continue;
        const double kk = coarse_scale * static_cast<double>(F.k[i][2 * axis + 1]);
        k[I][2 * axis + 1] += kk;
        k[J][2 * axis] += kk;
        C.nbr[I][2 * axis + 1] = J;
        C.nbr[J][2 * axis] = I;
      }
    }
    C.k.resize(C.n);
    C.anchor.resize(C.n);
    for (size_t i = 0; i < C.n; ++i) {
      for (int f = 0; f < 6; ++f) C.k[i][f] = static_cast<T>(k[i][f]);
      C.anchor[i] = static_cast<T>(anchor[i]);
    }
    FinishLevel(C);
    H.push_back(std::move(C));
  }
  return H;
}

// Edge-difference operator (mac_pressure_workspace_3d.cc:228-246).
template <typename T>
void Apply(const Level<T>& L, const std::vector<T>& x, std::vector<T>& y) {
  for (size_t i = 0; i < L.n; ++i) {
    T s = L.anchor[i] * x[i];
    for (int f = 0; f < 6; ++f) {
      const int32_t j = L.nbr[i][f];
      if (j >= 0) s += L.k[i][f] * (x[i] - x[j]);
    }
    y[i] = s;
  }
}

template <typename T>
void HalfSweep(Level<T>& L, uint8_t color) {
  for (size_t i = 0; i < L.n; ++i) {
    if (L.color[i] != color || L.diag[i] == T(0)) continue;
    T s = L.b[i];
    for (int f = 0; f < 6; ++f) {
      const int32_t j = L.nbr[i][f];
      if (j >= 0) s += L.k[i][f] * L.x[j];
    }
    L.x[i] = s / L.diag[i];
  }
}

// Symmetric V-cycle: x = M^{-1} b on level lv, starting from zero.
template <typename T>
void VCycle(std::vector<Level<T>>& H, size_t lv, int nu, int coarse_sweeps) {
  Level<T>& L = H[lv];
  std::fill(L.x.begin(), L.x.end(), T(0));
  if (lv + 1 == H.size()) {
    for (int s = 0; s < coarse_sweeps; ++s) {
      HalfSweep(L, 0);
      HalfSweep(L, 1);
    }
    for (int s = 0; s < coarse_sweeps; ++s) {
      HalfSweep(L, 1);
      HalfSweep(L, 0);
    }
    return;
  }
  for (int s = 0; s < nu; ++s) {
    HalfSweep(L, 0);
    HalfSweep(L, 1);
  }
  Apply(L, L.x, L.r);
  Level<T>& C = H[lv + 1];
  std::fill(C.b.begin(), C.b.end(), T(0));
  for (size_t i = 0; i < L.n; ++i) C.b[L.parent[i]] += L.b[i] - L.r[i];
  VCycle(H, lv + 1, nu, coarse_sweeps);
  for (size_t i = 0; i < L.n; ++i) L.x[i] += C.x[L.parent[i]];
  for (int s = 0; s < nu; ++s) {
    HalfSweep(L, 1);
    HalfSweep(L, 0);
  }
}

template <typename T>
double Dot(const std::vector<T>& a, const std::vector<T>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += static_cast<double>(a[i]) * static_cast<double>(b[i]);
  return s;
}

struct Criteria {
  double integrated = 0, volume = 2e-5, relative = 0, initial_defect = 0;
};

bool Accept(const Diag& d, const Criteria& c) {
  if (!(std::isfinite(d.l2) && std::isfinite(d.max_res) && std::isfinite(d.max_defect))) return false;
  if (!(d.l2 <= c.integrated && d.max_res <= c.volume && d.max_defect <= c.volume)) return false;
  if (c.relative > 0.0 && !(d.defect_l2 <= c.relative * c.initial_defect)) return false;
  return true;
}

template <typename T>
int Solve(const System& sys, const flo_pressure_options& opt, const Criteria& crit, std::vector<double>& p,
          std::vector<double>& vel, std::vector<double>& res, std::vector<double>& def, Diag& diag,
          flo_pressure_diagnostics& out) {
  const FineGraph g = BuildFine(sys);
  auto H = BuildHierarchy<T>(g, std::max(1, opt.levels), opt.coarse_scale);
  out.multigrid_level_count = H.size();
  const size_t n = g.unknown_to_cell.size();
  std::vector<double> base = p;
  if (int s = Evaluate(sys, p, vel, res, def, diag)) return s;
  out.iteration_count = 0;
  if (Accept(diag, crit)) return FLO_OK;
  if (n == 0) return FLO_ABORTED;
  Level<T>& L0 = H[0];
  std::vector<T> r(n), x(n, T(0)), d(n, T(0)), q(n), z(n);
  auto load_residual = [&]() {
    for (size_t u = 0; u < n; ++u) r[u] = static_cast<T>(res[g.unknown_to_cell[u]]);
  };
  load_residual();
  double rho_prev = 0.0;
  bool restart = true;
  const int refresh = 10;
  int since_refresh = 0;
  for (uint64_t it = 1; it <= opt.max_iterations; ++it) {
    L0.b = r;
    VCycle(H, 0, opt.smoothing_sweeps, opt.coarse_sweeps);
    z = L0.x;
    const double rho = Dot(r, z);
    if (!(rho > 0.0)) return FLO_ABORTED;
    if (restart) {
      d = z;
    } else {
      const T beta = static_cast<T>(rho / rho_prev);
      for (size_t u = 0; u < n; ++u) d[u] = z[u] + beta * d[u];
    }
    Apply(L0, d, q);
    const double curv = Dot(d, q);
    if (!(curv > 0.0)) return FLO_ABORTED;
    const T alpha = static_cast<T>(rho / curv);
    double rn = 0.0, rmax = 0.0, dn = 0.0;
    for (size_t u = 0; u < n; ++u) {
      x[u] += alpha * d[u];
      r[u] -= alpha * q[u];
      const double rv = static_cast<double>(r[u]);
      const double V = sys.cells[g.unknown_to_cell[u]].volume;
      rn = std::hypot(rn, rv);
      rmax = std::max(rmax, std::fabs(rv) / V);
      dn = std::hypot(dn, rv / V);
    }
    restart = false;
    rho_prev = rho;
    ++since_refresh;
    const bool last = it == opt.max_iterations;
    bool recursive = rn <= crit.integrated && rmax <= crit.volume;
    if (crit.relative > 0.0) recursive = recursive && dn <= crit.relative * crit.initial_defect;
    if (last || recursive || since_refresh >= refresh) {
      p = base;
      for (size_t u = 0; u < n; ++u) p[g.unknown_to_cell[u]] += static_cast<double>(x[u]);
      if (int s = Evaluate(sys, p, vel, res, def, diag)) return s;
      out.iteration_count = it;
      if (Accept(diag, crit)) return FLO_OK;
      if (last) break;
      base = p;
      std::fill(x.begin(), x.end(), T(0));
      load_residual();
      restart = true;
      since_refresh = 0;
      ++out.restarts;
    }
  }
  return FLO_ABORTED;
}

// ---------------- the reference's own preconditioner: aggregation multigrid ----------------
//
// BuildMacPressureHierarchy3D (mac_pressure_3d_multigrid.cc:218-301): fine edges and anchors
// rounded once to FP32 (:20-31, :238-275), levels completed by CompleteLevel (:42-87), coarsened by
// floor-half aggregation inside connected components with the linked-pair fallback on a signed
// plateau (Coarsen, :102-165) until every component of a coarse level has at most 32 unknowns (the
// fine level: one), and the last level's components factored by FP64 Cholesky
// (FactorTerminalBlocks, :167-214). The V-cycle is ApplyVcycle (mac_pressure_workspace_3d.cc:298-397):
// one 2/3-weighted Jacobi sweep before and after the piecewise-constant coarse correction, the fine
// level in the PCG's scalar, every coarse level in FP64. The PCG is SolvePrepared's
// (mac_pressure_workspace_3d.cc:544-658): the true residual is recomputed only at recursive
// convergence or on the last iteration. The dense shortcut for at most 16 cells with a tightened
// tolerance (:493-543) is not restated; this solver is the CPU baseline of large systems.

struct AggEdge {
  size_t lo, hi;
  double k;
};
struct AggNeighbour {
  size_t other, edge;
};
struct AggLevel {
  std::vector<flo_int3> coord;
  std::vector<double> anchor, diag;
  std::vector<AggEdge> edges;
  std::vector<size_t> row;  // CSR offsets, size n + 1
  std::vector<AggNeighbour> nbr;
  std::vector<size_t> comp;  // connected-component root of each unknown
  std::vector<size_t> agg;   // aggregate in the next level
  std::vector<size_t> term_row, term_unknowns;  // terminal blocks of the last level
  std::vector<size_t> term_offset;              // first factor entry of each block
  std::vector<double> term_factor;              // lower Cholesky factors, row-major per block
  size_t size() const { return coord.size(); }
};
struct AggHierarchy {
  std::vector<size_t> unknown_to_cell, cell_to_unknown;
  std::vector<AggLevel> levels;
};

// mac_pressure_3d_multigrid.cc:20-31
int RoundToFine(double v, double& out) {
  if (!FinitePositive(v) || v > std::numeric_limits<float>::max()) return FLO_OUT_OF_RANGE;
  const float f = static_cast<float>(v);
  if (!(std::isfinite(f) && f > 0.0f)) return FLO_OUT_OF_RANGE;
  out = static_cast<double>(f);
  return FLO_OK;
}

size_t HalvingRoot(std::vector<size_t>& r, size_t u) {
  while (r[u] != u) {
    r[u] = r[r[u]];
    u = r[u];
  }
  return u;
}

int32_t FloorHalfSigned(int32_t v) { return v / 2 - (v % 2 < 0 ? 1 : 0); }

// Diagonal, CSR rows (neighbours in ascending ordinal) and components from the edge list.
int CompleteAggLevel(AggLevel& L) {
  const size_t n = L.size();
  L.diag = L.anchor;
  L.row.assign(n + 1, 0);
  L.comp.resize(n);
  std::iota(L.comp.begin(), L.comp.end(), size_t{0});
  for (const AggEdge& e : L.edges) {
    L.diag[e.lo] += e.k;
    L.diag[e.hi] += e.k;
    ++L.row[e.lo + 1];
    ++L.row[e.hi + 1];
    const size_t a = HalvingRoot(L.comp, e.lo), b = HalvingRoot(L.comp, e.hi);
    L.comp[std::max(a, b)] = std::min(a, b);
  }
  for (size_t u = 0; u < n; ++u) {
    if (!FinitePositive(L.diag[u]) || !std::isfinite(L.anchor[u])) return FLO_OUT_OF_RANGE;
    L.comp[u] = HalvingRoot(L.comp, u);
    L.row[u + 1] += L.row[u];
  }
  std::vector<size_t> next(L.row.begin(), L.row.end() - 1);
  L.nbr.resize(L.row.back());
  for (size_t i = 0; i < L.edges.size(); ++i) {
    L.nbr[next[L.edges[i].lo]++] = {L.edges[i].hi, i};
    L.nbr[next[L.edges[i].hi]++] = {L.edges[i].lo, i};
  }
  for (size_t u = 0; u < n; ++u)
    std::sort(L.nbr.begin() + static_cast<std::ptrdiff_t>(L.row[u]), L.nbr.begin() + static_cast<std::ptrdiff_t>(L.row[u + 1]),
              [](const AggNeighbour& a, const AggNeighbour& b) { return a.other < b.other; });
  return FLO_OK;
}

// Connected components, ordered by root ordinal, members ascending.
std::vector<std::vector<size_t>> AggComponents(const AggLevel& L) {
  std::vector<std::vector<size_t>> by_root(L.size());
  for (size_t u = 0; u < L.size(); ++u) by_root[L.comp[u]].push_back(u);
  std::vector<std::vector<size_t>> out;
  for (auto& c : by_root)
    if (!c.empty()) out.push_back(std::move(c));
  return out;
}

int CoarsenAgg(AggLevel& F, AggLevel& C) {
  constexpr size_t kFree = SIZE_MAX;
  F.agg.assign(F.size(), kFree);
  for (const auto& comp : AggComponents(F)) {
    // Spatial groups in (z, y, x) order of the halved coordinates.
    std::map<std::array<int32_t, 3>, std::vector<size_t>> groups;
    for (size_t u : comp) {
      const flo_int3 c = F.coord[u];
      groups[{FloorHalfSigned(c.z), FloorHalfSigned(c.y), FloorHalfSigned(c.x)}].push_back(u);
    }
    if (groups.size() < comp.size() || comp.size() == 1) {
      for (const auto& [key, members] : groups) {
        const size_t a = C.coord.size();
        C.coord.push_back({key[2], key[1], key[0]});
        for (size_t u : members) F.agg[u] = a;
      }
      continue;
    }
    // Every unknown halves to its own coordinate (a signed plateau): pair each free unknown with its
    // lowest free neighbour so that the level still shrinks.
    for (size_t u : comp) {
      if (F.agg[u] != kFree) continue;
      const size_t a = C.coord.size();
      const flo_int3 c = F.coord[u];
      C.coord.push_back({FloorHalfSigned(c.x), FloorHalfSigned(c.y), FloorHalfSigned(c.z)});
      F.agg[u] = a;
      for (size_t o = F.row[u]; o < F.row[u + 1]; ++o) {
        const size_t v = F.nbr[o].other;
        if (F.agg[v] == kFree) {
          F.agg[v] = a;
          break;
        }
      }
    }
  }
  C.anchor.assign(C.size(), 0.0);
  for (size_t u = 0; u < F.size(); ++u) C.anchor[F.agg[u]] += F.anchor[u];
  // Galerkin operator of piecewise-constant prolongation: edges inside an aggregate cancel, crossing
  // conductances add once per aggregate pair.
  std::map<std::pair<size_t, size_t>, double> crossing;
  for (const AggEdge& e : F.edges) {
    const size_t a = F.agg[e.lo], b = F.agg[e.hi];
    if (a != b) crossing[{std::min(a, b), std::max(a, b)}] += e.k;
  }
  C.edges.reserve(crossing.size());
  for (const auto& [ends, k] : crossing) {
    if (!FinitePositive(k)) return FLO_OUT_OF_RANGE;
    C.edges.push_back({ends.first, ends.second, k});
  }
  return CompleteAggLevel(C);
}

// Cholesky-Banachiewicz (row by row, lower factor) of each terminal component's dense matrix.
int FactorAggTerminal(AggLevel& L) {
  std::vector<size_t> local(L.size());
  L.term_row.assign(1, 0);
  L.term_offset.clear();
  L.term_unknowns.clear();
  L.term_factor.clear();
  for (const auto& comp : AggComponents(L)) {
    const size_t m = comp.size();
    if (m > 32) return FLO_INTERNAL;
    for (size_t i = 0; i < m; ++i) local[comp[i]] = i;
    std::vector<double> a(m * m, 0.0);
    for (size_t r = 0; r < m; ++r) {
      const size_t u = comp[r];
      a[r * m + r] = L.diag[u];
      for (size_t o = L.row[u]; o < L.row[u + 1]; ++o) a[r * m + local[L.nbr[o].other]] = -L.edges[L.nbr[o].edge].k;
    }
    for (size_t r = 0; r < m; ++r) {
      for (size_t c = 0; c <= r; ++c) {
        double s = a[r * m + c];
        for (size_t q = 0; q < c; ++q) s -= a[r * m + q] * a[c * m + q];
        if (r == c) {
          if (!FinitePositive(s)) return FLO_OUT_OF_RANGE;
          a[r * m + r] = std::sqrt(s);
        } else {
          a[r * m + c] = s / a[c * m + c];
          if (!std::isfinite(a[r * m + c])) return FLO_OUT_OF_RANGE;
        }
      }
      for (size_t c = r + 1; c < m; ++c) a[r * m + c] = 0.0;
    }
    L.term_unknowns.insert(L.term_unknowns.end(), comp.begin(), comp.end());
    L.term_row.push_back(L.term_unknowns.size());
    L.term_offset.push_back(L.term_factor.size());
    L.term_factor.insert(L.term_factor.end(), a.begin(), a.end());
  }
  return FLO_OK;
}

int BuildAggHierarchy(const System& sys, AggHierarchy& H) {
  const size_t nc = sys.cells.size();
  std::vector<char> pinned(nc, 0);
  for (size_t g : sys.gauges) pinned[g] = 1;
  H.cell_to_unknown.assign(nc, kNone);
  H.unknown_to_cell.clear();
  AggLevel fine;
  for (size_t c = 0; c < nc; ++c) {
    if (pinned[c]) continue;
    H.cell_to_unknown[c] = H.unknown_to_cell.size();
    H.unknown_to_cell.push_back(c);
    fine.coord.push_back(sys.cells[c].index);
    fine.anchor.push_back(sys.cells[c].compliance);
  }
  for (const Link& l : sys.links) {
    if (l.conductance == 0.0) continue;
    const size_t t = l.tail != kNone ? H.cell_to_unknown[l.tail] : kNone;
    const size_t h = l.head != kNone ? H.cell_to_unknown[l.head] : kNone;
    if (t != kNone && h != kNone) {
      double k;
      if (int s = RoundToFine(l.conductance, k)) return s;
      fine.edges.push_back({std::min(t, h), std::max(t, h), k});
    } else if (t != kNone) {
      fine.anchor[t] += l.conductance;
    } else if (h != kNone) {
      fine.anchor[h] += l.conductance;
    }
  }
  for (double& a : fine.anchor)
    if (a != 0.0)
      if (int s = RoundToFine(a, a)) return s;
  std::sort(fine.edges.begin(), fine.edges.end(),
            [](const AggEdge& a, const AggEdge& b) { return a.lo != b.lo ? a.lo < b.lo : a.hi < b.hi; });
  if (int s = CompleteAggLevel(fine)) return s;
  for (double d : fine.diag) {
    double r;
    if (int s = RoundToFine(d, r)) return s;
  }
  H.levels.clear();
  H.levels.push_back(std::move(fine));
  for (;;) {
    const size_t limit = H.levels.size() == 1 ? 1 : 32;
    bool coarsen = false;
    for (const auto& c : AggComponents(H.levels.back())) coarsen |= c.size() > limit;
    if (!coarsen) break;
    AggLevel coarse;
    if (int s = CoarsenAgg(H.levels.back(), coarse)) return s;
    H.levels.push_back(std::move(coarse));
  }
  return FactorAggTerminal(H.levels.back());
}

int ToFine(double v, float& out) {  // FineValue (mac_pressure_workspace_3d.cc:212-223)
  if (!std::isfinite(v) || std::fabs(v) > std::numeric_limits<float>::max()) return FLO_OUT_OF_RANGE;
  const float r = static_cast<float>(v);
  if (!std::isfinite(r) || (v != 0.0 && r == 0.0f)) return FLO_OUT_OF_RANGE;
  out = r;
  return FLO_OK;
}
int ToFine(double v, double& out) {
  out = v;
  return FLO_OK;
}

// ApplyLevel (mac_pressure_workspace_3d.cc:228-246) in scalar S.
template <typename S>
int AggApply(const AggLevel& L, const std::vector<S>& x, std::vector<S>& y) {
  y.resize(L.size());
  for (size_t u = 0; u < L.size(); ++u) {
    S acc = static_cast<S>(L.anchor[u]) * x[u];
    for (size_t o = L.row[u]; o < L.row[u + 1]; ++o) {
      const S diff = x[u] - x[L.nbr[o].other];
      const S flux = static_cast<S>(L.edges[L.nbr[o].edge].k) * diff;
      acc += flux;
    }
    if (!std::isfinite(acc)) return FLO_OUT_OF_RANGE;
    y[u] = acc;
  }
  return FLO_OK;
}

template <typename S>
int AggVcycle(const AggHierarchy& H, size_t l, const std::vector<S>& r, std::vector<S>& x) {
  const AggLevel& L = H.levels[l];
  x.assign(L.size(), S(0));
  if (l + 1 == H.levels.size()) {
    double y[32];
    for (size_t b = 0; b + 1 < L.term_row.size(); ++b) {
      const size_t begin = L.term_row[b], m = L.term_row[b + 1] - begin;
      const size_t* unk = L.term_unknowns.data() + begin;
      const double* f = L.term_factor.data() + L.term_offset[b];
      for (size_t i = 0; i < m; ++i) {  // L y = r
        double s = r[unk[i]];
        for (size_t j = 0; j < i; ++j) s -= f[i * m + j] * y[j];
        y[i] = s / f[i * m + i];
      }
      for (size_t k = m; k > 0; --k) {  // L^T z = y
        const size_t i = k - 1;
        double s = y[i];
        for (size_t j = i + 1; j < m; ++j) s -= f[j * m + i] * y[j];
        y[i] = s / f[i * m + i];
        if (!std::isfinite(y[i])) return FLO_OUT_OF_RANGE;
        if (int st = ToFine(y[i], x[unk[i]])) return st;
      }
    }
    return FLO_OK;
  }
  const S w = S(2) / S(3);
  for (size_t u = 0; u < L.size(); ++u) x[u] = w * (r[u] / static_cast<S>(L.diag[u]));
  std::vector<S> flux;
  if (int s = AggApply(L, x, flux)) return s;
  const AggLevel& C = H.levels[l + 1];
  std::vector<double> cr(C.size(), 0.0), cx;
  for (size_t u = 0; u < L.size(); ++u) {
    const S rem = r[u] - flux[u];
    cr[L.agg[u]] += static_cast<double>(rem);
  }
  for (double v : cr)
    if (!std::isfinite(v)) return FLO_OUT_OF_RANGE;
  if (int s = AggVcycle<double>(H, l + 1, cr, cx)) return s;
  for (size_t u = 0; u < L.size(); ++u) {
    S c;
    if (int s = ToFine(cx[L.agg[u]], c)) return s;
    x[u] += c;
  }
  if (int s = AggApply(L, x, flux)) return s;
  for (size_t u = 0; u < L.size(); ++u) {
    const S rem = r[u] - flux[u];
    x[u] += w * (rem / static_cast<S>(L.diag[u]));
    if (!std::isfinite(x[u])) return FLO_OUT_OF_RANGE;
  }
  return FLO_OK;
}

template <typename T>
int DotChecked(const std::vector<T>& a, const std::vector<T>& b, double& out) {
  out = Dot(a, b);
  return std::isfinite(out) ? FLO_OK : FLO_OUT_OF_RANGE;
}

// SolvePrepared's PCG (mac_pressure_workspace_3d.cc:544-658) with the aggregation V-cycle.
template <typename T>
int SolveAggregation(const System& sys, const flo_pressure_options& opt, const Criteria& crit, std::vector<double>& p,
                     std::vector<double>& vel, std::vector<double>& res, std::vector<double>& def, Diag& diag,
                     flo_pressure_diagnostics& out) {
  AggHierarchy H;
  if (int s = BuildAggHierarchy(sys, H)) return s;
  out.multigrid_level_count = H.levels.size();
  std::vector<double> base = p;
  if (int s = Evaluate(sys, p, vel, res, def, diag)) return s;
  out.iteration_count = 0;
  if (Accept(diag, crit)) return FLO_OK;
  const size_t n = H.unknown_to_cell.size();
  if (n == 0) return FLO_ABORTED;
  std::vector<T> r(n), x(n, T(0)), d(n, T(0)), q, z;
  auto load_residual = [&]() -> int {
    for (size_t u = 0; u < n; ++u)
      if (int s = ToFine(res[H.unknown_to_cell[u]], r[u])) return s;
    return FLO_OK;
  };
  if (int s = load_residual()) return s;
  double rho_prev = 0.0;
  bool restart = true;
  for (uint64_t it = 1; it <= opt.max_iterations; ++it) {
    if (int s = AggVcycle<T>(H, 0, r, z)) return s;
    double rho;
    if (int s = DotChecked(r, z, rho)) return s;
    if (rho <= 0.0) return FLO_ABORTED;
    if (restart) {
      d = z;
    } else {
      T beta;
      if (int s = ToFine(rho / rho_prev, beta)) return s;
      for (size_t u = 0; u < n; ++u) {
        const T prev = beta * d[u];
        d[u] = z[u] + prev;
      }
    }
    if (int s = AggApply(H.levels[0], d, q)) return s;
    double curv;
    if (int s = DotChecked(d, q, curv)) return s;
    if (curv <= 0.0) return FLO_ABORTED;
    T alpha;
    if (int s = ToFine(rho / curv, alpha)) return s;
    double rn = 0.0, rmax = 0.0, dn = 0.0;
    for (size_t u = 0; u < n; ++u) {
      const T dp = alpha * d[u];
      const T dq = alpha * q[u];
      x[u] += dp;
      r[u] -= dq;
      if (!std::isfinite(x[u]) || !std::isfinite(r[u])) return FLO_OUT_OF_RANGE;
      const double rv = static_cast<double>(r[u]);
      const double V = sys.cells[H.unknown_to_cell[u]].volume;
      rn = std::hypot(rn, rv);
      rmax = std::max(rmax, std::fabs(rv) / V);
      dn = std::hypot(dn, rv / V);
    }
    restart = false;
    rho_prev = rho;
    const bool last = it == opt.max_iterations;
    bool recursive = rn <= crit.integrated && rmax <= crit.volume;
    if (crit.relative > 0.0) recursive = recursive && dn <= crit.relative * crit.initial_defect;
    if (last || recursive) {
      p = base;
      for (size_t u = 0; u < n; ++u) {
        p[H.unknown_to_cell[u]] += static_cast<double>(x[u]);
        if (!std::isfinite(p[H.unknown_to_cell[u]])) return FLO_OUT_OF_RANGE;
      }
      if (int s = Evaluate(sys, p, vel, res, def, diag)) return s;
      out.iteration_count = it;
      if (Accept(diag, crit)) return FLO_OK;
      if (last) break;
      base = p;
      std::fill(x.begin(), x.end(), T(0));
      if (int s = load_residual()) return s;
      restart = true;
      ++out.restarts;
    }
  }
  return FLO_ABORTED;
}

int PrepareAndInit(const flo_geometry* g, const flo_int3* blocks, size_t nb, double dt, const flo_pressure_cell* cells,
                   size_t nc, const flo_pressure_face* faces, size_t nf, System& sys, std::vector<double>& p) {
  if (int s = Prepare(*g, blocks, nb, dt, cells, nc, faces, nf, sys)) return s;
  // InitializeMacPressure3D (mac_pressure_3d_multigrid.cc:303-321)
  std::vector<double> pin(nc, 0.0);
  for (size_t c : sys.gauges) pin[sys.component[c]] = sys.cells[c].previous;
  p.resize(nc);
  for (size_t c = 0; c < nc; ++c) {
    p[c] = sys.cells[c].previous - pin[sys.component[c]];
    if (!std::isfinite(p[c])) return FLO_OUT_OF_RANGE;
  }
  return FLO_OK;
}


// ---- coupled viscous pressure (mac_viscous_pressure_3d.cc:151-459) ----

// AccumulateConstraint (mac_viscous_pressure_3d.cc:83-134): integrated D u + compliance flux per cell;
// the target and wall fluxes and the pressure history enter only for the physical candidate.
int CoupledConstraint(const System& sys, const std::vector<double>& p, const std::vector<double>& u, bool physical,
                      std::vector<double>& out) {
  const size_t nc = sys.cells.size();
  std::vector<Comp> flux(nc);
  for (size_t c = 0; c < nc; ++c) {
    if (physical) flux[c].Add(-sys.cells[c].target_flux);
    if (sys.cells[c].compliance > 0.0) {
      const double dp = p[c] - (physical ? sys.cells[c].previous : 0.0);
      const double f = sys.cells[c].compliance * dp;
      if (!Representable(sys.cells[c].compliance, dp, f)) return FLO_OUT_OF_RANGE;
      flux[c].Add(f);
    }
  }
  for (size_t f = 0; f < sys.links.size(); ++f) {
    const Link& l = sys.links[f];
    const double open = l.open_area * u[f];
    if (!Representable(l.open_area, u[f], open)) return FLO_OUT_OF_RANGE;
    ForIncident(l, [&](size_t c, double sign) {
      flux[c].Add(sign * open);
      if (physical) flux[c].Add(sign * l.wall_flux);
    });
  }
  out.resize(nc);
  for (size_t c = 0; c < nc; ++c) {
    out[c] = flux[c].Total();
    if (!std::isfinite(out[c])) return FLO_OUT_OF_RANGE;
  }
  return FLO_OK;
}

int CoupledNorm(const std::vector<double>& v, double& norm) {  // CoupledConstraintNorm (:140-149)
  norm = 0.0;
  for (double x : v) norm = std::hypot(norm, x);
  return std::isfinite(norm) ? FLO_OK : FLO_OUT_OF_RANGE;
}

double PlainDot(const std::vector<double>& a, const std::vector<double>& b) {
  return std::inner_product(a.begin(), a.end(), b.begin(), 0.0);
}

constexpr size_t kCoupledRestart = 32;  // kCoupledPressureRestartLength3D

// SolveCoupled (mac_viscous_pressure_3d.cc:151-459): GCR on the pressure with the Jacobi diagonal of
// the viscous Schur complement, each pressure iterate's velocity from the velocity block
// (M + dt E^T W E) u = base + dt A (p_tail - p_head) solved to 1e-10 of its largest right-hand side.
int SolveCoupled(const System& sys, double dt, std::vector<flo_strain_row> rows, const std::vector<double>& wall_u,
                 uint64_t max_iterations, std::vector<double>& p, std::vector<double>& velocity,
                 std::vector<double>& defects, flo_viscous_diagnostics& out) {
  const size_t nc = sys.cells.size(), nf = sys.links.size();
  // RemapStrain (:50-77): grid slots to link ordinals (the links are slot-sorted).
  for (auto& row : rows) {
    if (row.term_count == 0 || row.term_count > 8) return FLO_INVALID_ARGUMENT;
    if (!std::isfinite(row.quadrature_weight) || row.quadrature_weight < 0.0 || !std::isfinite(row.prescribed_strain))
      return FLO_INVALID_ARGUMENT;
    for (uint64_t t = 0; t < row.term_count; ++t) {
      auto& term = row.terms[t];
      auto it = std::lower_bound(sys.links.begin(), sys.links.end(), term.face_slot,
                                 [](const Link& l, uint64_t slot) { return l.slot < slot; });
      if (it == sys.links.end() || it->slot != term.face_slot || !std::isfinite(term.coefficient))
        return FLO_INVALID_ARGUMENT;
      term.face_slot = static_cast<uint64_t>(it - sys.links.begin());
    }
  }
  std::vector<double> inertia(nf, 0.0), base(nf, 0.0), vdiag(nf, 0.0), pdiag(nc, 0.0);
  for (size_t c = 0; c < nc; ++c) pdiag[c] = sys.cells[c].compliance;
  for (size_t f = 0; f < nf; ++f) {
    const Link& l = sys.links[f];
    if (l.open_area == 0.0) continue;
    const double step_area = dt * l.open_area;
    inertia[f] = step_area / l.scale;
    if (!std::isnormal(inertia[f]) || inertia[f] <= 0.0) return FLO_OUT_OF_RANGE;
    double lift = 0.0;
    if (l.kind == 1) lift = l.lift;
    else if (l.kind == 2) lift = l.tail != kNone ? -l.lift : l.lift;
    base[f] = inertia[f] * l.predicted + step_area * lift;
  }
  vdiag = inertia;
  for (const auto& row : rows)
    for (uint64_t t = 0; t < row.term_count; ++t) {
      const auto& term = row.terms[t];
      if (inertia[term.face_slot] > 0.0)
        vdiag[term.face_slot] += dt * row.quadrature_weight * term.coefficient * term.coefficient;
    }
  for (size_t f = 0; f < nf; ++f) {
    if (inertia[f] == 0.0) continue;
    const Link& l = sys.links[f];
    const double effective = l.conductance * (inertia[f] / vdiag[f]);
    if (!FinitePositive(effective)) return FLO_OUT_OF_RANGE;
    ForIncident(l, [&](size_t c, double) { pdiag[c] += effective; });
  }
  // Eliminate prescribed velocities of the strain rows (:226-246).
  for (const auto& row : rows) {
    if (row.quadrature_weight == 0.0) continue;
    double prescribed = row.prescribed_strain;
    for (uint64_t t = 0; t < row.term_count; ++t)
      if (inertia[row.terms[t].face_slot] == 0.0) prescribed += row.terms[t].coefficient * wall_u[row.terms[t].face_slot];
    if (prescribed == 0.0) continue;
    for (uint64_t t = 0; t < row.term_count; ++t)
      if (inertia[row.terms[t].face_slot] > 0.0)
        base[row.terms[t].face_slot] -= dt * row.quadrature_weight * prescribed * row.terms[t].coefficient;
  }
  for (double b : base)
    if (!std::isfinite(b)) return FLO_OUT_OF_RANGE;
  std::vector<char> pinned(nc, 0);
  for (size_t c : sys.gauges) pinned[c] = 1;
  std::vector<double> momentum(nf), u(nf);
  double momentum_residual = 0.0;
  auto solve_velocity = [&](const std::vector<double>& pressure, bool include_base) -> int {
    for (size_t f = 0; f < nf; ++f) {
      const Link& l = sys.links[f];
      const double tail = l.tail != kNone ? pressure[l.tail] : 0.0;
      const double head = l.head != kNone ? pressure[l.head] : 0.0;
      momentum[f] = include_base ? base[f] : 0.0;
      if (l.open_area > 0.0) momentum[f] += dt * l.open_area * (tail - head);
    }
    uint64_t iterations = 0;
    ++out.velocity_solves;
    return flo_solve_viscous_block(rows.data(), rows.size(), inertia.data(), nf, dt, momentum.data(), 256, 1e-10,
                                   u.data(), &iterations, &momentum_residual);
  };
  // InitializeMacPressure3D (mac_pressure_3d_multigrid.cc:303-321)
  std::vector<double> pin(nc, 0.0);
  for (size_t c : sys.gauges) pin[sys.component[c]] = sys.cells[c].previous;
  p.assign(nc, 0.0);
  bool history = false;
  for (size_t c = 0; c < nc; ++c) {
    p[c] = sys.cells[c].previous - pin[sys.component[c]];
    if (!std::isfinite(p[c])) return FLO_OUT_OF_RANGE;
    history |= p[c] != 0.0;
  }
  std::vector<double> residual, response;
  double rhs_norm = 0.0;
  if (history) {
    const std::vector<double> zero(nc, 0.0);
    if (int s = solve_velocity(zero, true)) return s;
    if (int s = CoupledConstraint(sys, zero, u, true, residual)) return s;
    if (int s = CoupledNorm(residual, rhs_norm)) return s;
    if (rhs_norm == 0.0) std::fill(p.begin(), p.end(), 0.0);
  }
  const size_t capacity = std::min<size_t>(kCoupledRestart, max_iterations);
  std::vector<std::vector<double>> directions(capacity, std::vector<double>(nc)), responses(capacity,
                                                                                           std::vector<double>(nc));
  size_t basis = 0;
  double integrated_limit = 0.0;
  const double volume_limit = 2e-5;
  for (uint64_t it = 0; it <= max_iterations; ++it) {
    if (int s = solve_velocity(p, true)) return s;
    if (int s = CoupledConstraint(sys, p, u, true, residual)) return s;
    double norm = 0.0;
    if (int s = CoupledNorm(residual, norm)) return s;
    if (it == 0) {
      if (!history) rhs_norm = norm;
      // MacPressureStoppingCriteria (mac_pressure_3d_multigrid.cc:341-357)
      integrated_limit = 2e-6 + 2e-6 * rhs_norm;
      if (!std::isfinite(rhs_norm) || rhs_norm < 0.0 || !std::isfinite(integrated_limit)) return FLO_OUT_OF_RANGE;
      bool jump = false;
      for (const auto& l : sys.links) jump |= l.conductance > 0.0 && l.kind == 1 && l.lift != 0.0;
      if (rhs_norm > 0.0 && jump)
        integrated_limit =
            std::min(integrated_limit, std::max(std::numeric_limits<double>::denorm_min(), 1e-3 * rhs_norm));
    }
    double max_defect = 0.0;
    for (size_t c = 0; c < nc; ++c) {
      const double d = residual[c] / sys.cells[c].volume;
      if (!std::isfinite(d)) return FLO_OUT_OF_RANGE;
      max_defect = std::max(max_defect, std::fabs(d));
    }
    out.pressure_iterations = it;
    out.constraint_residual_l2 = norm;
    out.initial_constraint_residual_l2 = rhs_norm;
    out.max_cell_constraint_defect = max_defect;
    out.max_momentum_residual = momentum_residual;
    if (max_defect <= volume_limit && norm <= integrated_limit) {
      defects.resize(nc);
      for (size_t c = 0; c < nc; ++c) defects[c] = residual[c] / sys.cells[c].volume;
      velocity = u;
      for (size_t f = 0; f < nf; ++f)
        if (inertia[f] == 0.0) velocity[f] = wall_u[f];
      return FLO_OK;
    }
    if (it == max_iterations) break;
    if (basis == capacity) basis = 0;
    std::vector<double>& dir = directions[basis];
    std::vector<double>& resp = responses[basis];
    for (size_t c = 0; c < nc; ++c) {
      if (pinned[c]) {
        residual[c] = 0.0;
        dir[c] = 0.0;
      } else {
        if (!FinitePositive(pdiag[c])) return FLO_OUT_OF_RANGE;
        dir[c] = residual[c] / pdiag[c];
      }
    }
    if (int s = solve_velocity(dir, false)) return s;
    if (int s = CoupledConstraint(sys, dir, u, false, resp)) return s;
    for (size_t c = 0; c < nc; ++c)
      if (pinned[c]) resp[c] = 0.0;
    for (size_t q = 0; q < basis; ++q) {
      const double overlap = PlainDot(resp, responses[q]);
      for (size_t c = 0; c < nc; ++c) {
        resp[c] -= overlap * responses[q][c];
        dir[c] -= overlap * directions[q][c];
      }
    }
    const double rn = std::sqrt(PlainDot(resp, resp));
    if (!std::isfinite(rn) || rn <= 0.0) return FLO_ABORTED;
    for (size_t c = 0; c < nc; ++c) {
      resp[c] /= rn;
      dir[c] /= rn;
    }
    const double step = PlainDot(residual, resp);
    for (size_t c = 0; c < nc; ++c) {
      p[c] -= step * dir[c];
      if (!std::isfinite(p[c])) return FLO_OUT_OF_RANGE;
    }
    ++basis;
  }
  return FLO_ABORTED;
}

}  // namespace

extern "C" {

int flo_project(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count, double dt,
                const flo_pressure_cell* cells, size_t cell_count, const flo_pressure_face* faces, size_t face_count,
                const flo_pressure_options* options, double* out_pressures, double* out_velocities, double* out_rhs,
                double* out_residuals, double* out_defects, uint64_t* out_gauges,
                flo_pressure_diagnostics* out_diagnostics) {
  const flo_pressure_options opt = *options;
  if (opt.max_iterations < 1 || opt.preconditioner < 0 || opt.preconditioner > 1) return FLO_INVALID_ARGUMENT;
  if (!(opt.cpu_residual_tolerance_scale >= 1e-12 && opt.cpu_residual_tolerance_scale <= 1.0))
    return FLO_INVALID_ARGUMENT;
  System sys;
  std::vector<double> p;
  if (int s = PrepareAndInit(geometry, sorted_blocks, block_count, dt, cells, cell_count, faces, face_count, sys, p))
    return s;
  // MacPressureStoppingCriteria (mac_pressure_3d_multigrid.cc:333-357)
  double rhs_norm = 0.0;
  for (double b : sys.rhs) rhs_norm = std::hypot(rhs_norm, b);
  double limit = 2e-6 + 2e-6 * rhs_norm;
  if (!std::isfinite(rhs_norm) || !std::isfinite(limit)) return FLO_OUT_OF_RANGE;
  bool has_jump = false;
  for (const auto& l : sys.links) has_jump |= (l.conductance > 0.0 && l.kind == 1 && l.lift != 0.0);
  if (rhs_norm > 0.0 && has_jump) limit = std::min(limit, std::max(std::numeric_limits<double>::denorm_min(), 1e-3 * rhs_norm));
  Criteria crit;
  crit.integrated = limit * opt.cpu_residual_tolerance_scale;
  crit.volume = 2e-5 * opt.cpu_residual_tolerance_scale;
  crit.relative = opt.relative_divergence_target;
  std::vector<double> zero(cell_count, 0.0);
  {
    std::vector<FluxRow> rows;
    if (int s = PhysicalRows(sys, zero, nullptr, rows)) return s;
    double n0 = 0.0;
    for (size_t c = 0; c < cell_count; ++c) n0 = std::hypot(n0, rows[c].flux.Total() / sys.cells[c].volume);
    crit.initial_defect = n0;
  }
  std::vector<double> vel, res, def;
  Diag diag;
  flo_pressure_diagnostics out{};
  int s;
  if (opt.preconditioner == 1)
    s = opt.fp32_vcycle ? SolveAggregation<float>(sys, opt, crit, p, vel, res, def, diag, out)
                        : SolveAggregation<double>(sys, opt, crit, p, vel, res, def, diag, out);
  else
    s = opt.fp32_vcycle ? Solve<float>(sys, opt, crit, p, vel, res, def, diag, out)
                        : Solve<double>(sys, opt, crit, p, vel, res, def, diag, out);
  out.true_residual_l2 = diag.l2;
  out.max_volume_scaled_residual = diag.max_res;
  out.max_cell_constraint_defect = diag.max_defect;
  out.rhs_l2 = rhs_norm;
  out.initial_defect_l2 = crit.initial_defect;
  out.defect_l2 = diag.defect_l2;
  out.defect_linf = diag.max_defect;
  out.gauge_count = sys.gauges.size();
  if (out_diagnostics) *out_diagnostics = out;
  if (s) return s;
  if (out_pressures) std::copy(p.begin(), p.end(), out_pressures);
  if (out_velocities) std::copy(vel.begin(), vel.end(), out_velocities);
  if (out_rhs) std::copy(sys.rhs.begin(), sys.rhs.end(), out_rhs);
  if (out_residuals) std::copy(res.begin(), res.end(), out_residuals);
  if (out_defects) std::copy(def.begin(), def.end(), out_defects);
  if (out_gauges) std::copy(sys.gauges.begin(), sys.gauges.end(), out_gauges);
  return FLO_OK;
}

int flo_evaluate(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count, double dt,
                 const flo_pressure_cell* cells, size_t cell_count, const flo_pressure_face* faces, size_t face_count,
                 const double* pressures, double* out_velocities, double* out_residuals, double* out_defects,
                 flo_pressure_diagnostics* out_diagnostics) {
  System sys;
  std::vector<double> p0;
  if (int s = PrepareAndInit(geometry, sorted_blocks, block_count, dt, cells, cell_count, faces, face_count, sys, p0))
    return s;
  std::vector<double> p(pressures, pressures + cell_count), vel, res, def;
  Diag d;
  if (int s = Evaluate(sys, p, vel, res, def, d)) return s;
  std::vector<double> zero(cell_count, 0.0);
  std::vector<FluxRow> rows;
  double n0 = 0.0;
  if (PhysicalRows(sys, zero, nullptr, rows) == FLO_OK)
    for (size_t c = 0; c < cell_count; ++c) n0 = std::hypot(n0, rows[c].flux.Total() / sys.cells[c].volume);
  if (out_velocities) std::copy(vel.begin(), vel.end(), out_velocities);
  if (out_residuals) std::copy(res.begin(), res.end(), out_residuals);
  if (out_defects) std::copy(def.begin(), def.end(), out_defects);
  if (out_diagnostics) {
    flo_pressure_diagnostics o{};
    o.true_residual_l2 = d.l2;
    o.max_volume_scaled_residual = d.max_res;
    o.max_cell_constraint_defect = d.max_defect;
    o.initial_defect_l2 = n0;
    o.defect_l2 = d.defect_l2;
    o.defect_linf = d.max_defect;
    o.gauge_count = sys.gauges.size();
    *out_diagnostics = o;
  }
  return FLO_OK;
}

int flo_reference_hierarchy(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count, double dt,
                            const flo_pressure_cell* cells, size_t cell_count, const flo_pressure_face* faces,
                            size_t face_count, uint64_t level, uint64_t capacity, flo_hierarchy_level* out_level,
                            int32_t* coords, double* anchors, double* diagonal, uint64_t* edge_lo, uint64_t* edge_hi,
                            double* edge_k, uint64_t* aggregates, uint64_t* unknown_to_cell) {
  System sys;
  std::vector<double> p;
  if (int s = PrepareAndInit(geometry, sorted_blocks, block_count, dt, cells, cell_count, faces, face_count, sys, p))
    return s;
  AggHierarchy H;
  if (int s = BuildAggHierarchy(sys, H)) return s;
  if (level >= H.levels.size()) return FLO_OUT_OF_RANGE;
  const AggLevel& L = H.levels[level];
  if (L.size() > capacity || L.edges.size() > 3 * capacity) return FLO_RESOURCE_EXHAUSTED;
  flo_hierarchy_level o{};
  o.level_count = H.levels.size();
  o.unknowns = L.size();
  o.edges = L.edges.size();
  if (level + 1 == H.levels.size()) {
    o.terminal_blocks = L.term_row.size() - 1;
    for (size_t b = 0; b + 1 < L.term_row.size(); ++b)
      o.max_terminal = std::max<uint64_t>(o.max_terminal, L.term_row[b + 1] - L.term_row[b]);
  }
  if (out_level) *out_level = o;
  for (size_t u = 0; u < L.size(); ++u) {
    if (coords) {
      coords[3 * u] = L.coord[u].x;
      coords[3 * u + 1] = L.coord[u].y;
      coords[3 * u + 2] = L.coord[u].z;
    }
    if (anchors) anchors[u] = L.anchor[u];
    if (diagonal) diagonal[u] = L.diag[u];
    if (aggregates) aggregates[u] = L.agg.empty() ? UINT64_MAX : L.agg[u];
    if (unknown_to_cell && level == 0) unknown_to_cell[u] = H.unknown_to_cell[u];
  }
  for (size_t e = 0; e < L.edges.size(); ++e) {
    if (edge_lo) edge_lo[e] = L.edges[e].lo;
    if (edge_hi) edge_hi[e] = L.edges[e].hi;
    if (edge_k) edge_k[e] = L.edges[e].k;
  }
  return FLO_OK;
}

int flo_reference_vcycle(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count, double dt,
                         const flo_pressure_cell* cells, size_t cell_count, const flo_pressure_face* faces,
                         size_t face_count, const double* r, double* z) {
  System sys;
  std::vector<double> p;
  if (int s = PrepareAndInit(geometry, sorted_blocks, block_count, dt, cells, cell_count, faces, face_count, sys, p))
    return s;
  AggHierarchy H;
  if (int s = BuildAggHierarchy(sys, H)) return s;
  const size_t n = H.unknown_to_cell.size();
  std::vector<double> rv(r, r + n), zv;
  if (int s = AggVcycle<double>(H, 0, rv, zv)) return s;
  std::copy(zv.begin(), zv.end(), z);
  return FLO_OK;
}

int flo_project_viscous(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count, double dt,
                        const flo_pressure_cell* cells, size_t cell_count, const flo_pressure_face* faces,
                        size_t face_count, const flo_strain_row* rows, size_t row_count, uint64_t max_iterations,
                        double* out_pressures, double* out_velocities, double* out_defects,
                        flo_viscous_diagnostics* out_diagnostics) {
  // ProjectStandaloneUniform (mac_viscous_pressure_3d.cc:558-575), ProjectUniformWithVelocityView (:461-507)
  if (max_iterations == 0 || max_iterations > 256) return FLO_INVALID_ARGUMENT;
  System sys;
  if (int s = Prepare(*geometry, sorted_blocks, block_count, dt, cells, cell_count, faces, face_count, sys)) return s;
  std::vector<double> wall_u(face_count);
  Comp inward;
  for (size_t f = 0; f < face_count; ++f) {
    wall_u[f] = faces[f].wall_normal_velocity;
    const Link& l = sys.links[f];
    // AccumulateInwardExplicitWallFlux (mac_pressure_3d_arithmetic.h:97-108)
    if (faces[f].has_prescribed_wall_flux && (l.tail != kNone) != (l.head != kNone))
      ForIncident(l, [&](size_t, double sign) { inward.Add(-sign * l.wall_flux); });
  }
  flo_viscous_diagnostics out{};
  std::vector<double> p, u, d;
  const int s = SolveCoupled(sys, dt, std::vector<flo_strain_row>(rows, rows + row_count), wall_u, max_iterations, p,
                             u, d, out);
  out.inward_prescribed_wall_flux = inward.Total();
  if (out_diagnostics) *out_diagnostics = out;
  if (s) return s;
  if (out_pressures) std::copy(p.begin(), p.end(), out_pressures);
  if (out_velocities) std::copy(u.begin(), u.end(), out_velocities);
  if (out_defects) std::copy(d.begin(), d.end(), out_defects);
  return FLO_OK;
}

const char* flo_compiler(void) {
#if defined(__clang__)
  return "clang " __clang_version__;
#else
  return "g++ " __VERSION__;
#endif
}

}  // extern "C"
