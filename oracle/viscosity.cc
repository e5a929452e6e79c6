// Oracle: the variational MAC viscosity operators (test infrastructure only). Restates
// native/src/viscosity.cc: full-cell strain rows (AssembleMacStrain, :762-946, with HarmonicEdgeViscosity
// :91-103, StrainWeight3D :150-158, SetStrainTerms :160-178, FactorProduct :25-55) and the Jacobi-PCG
// velocity block (ValidateVelocityBlockInputs :1331-1358, BuildWeightedVelocityRows :1360-1406,
// SolvePreparedVelocityBlock :1408-1512). Sums run sequentially in the reference's order.
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>

#include "grid_internal.h"

using namespace flo;

namespace {

struct Binary {
  double fraction;
  int exponent;
  double value(int shift = 0) const { return std::scalbn(fraction, exponent + shift); }
};

// FactorProduct (viscosity.cc:36-51): fractions multiplied, exponents summed separately.
Binary Factor(double a, double b, double c, double divisor = 1.0) {
  if (a == 0.0 || b == 0.0 || c == 0.0) return {0.0, 0};
  int ea, eb, ec, ed;
  const double fa = std::frexp(a, &ea), fb = std::frexp(b, &eb), fc = std::frexp(c, &ec), fd = std::frexp(divisor, &ed);
  return {(fa * (fb * fc)) / fd, ea + eb + ec - ed};
}
double Scaled(double a, double b, double c, double divisor = 1.0) { return Factor(a, b, c, divisor).value(); }

// StrainWeight3D (:150-158): mu h^3 2^k without an intermediate rounding of mu h^2.
double StrainWeight(double mu, double h, int k) {
  int eh;
  const double fh = std::frexp(h, &eh);
  const Binary area = Factor(mu, h, h);
  return std::scalbn(area.fraction * fh, area.exponent + eh + k);
}

// HarmonicEdgeViscosity (:91-103).
double HarmonicEdge(const double mu[4]) {
  const double m = std::min(std::min(mu[0], mu[1]), std::min(mu[2], mu[3]));
  if (m == 0.0) return 0.0;
  double sum = 0.0;
  for (int i = 0; i < 4; ++i) sum += m / mu[i];
  return m * (4.0 / sum);
}

bool CellLess(const flo_int3& a, const flo_int3& b) {
  if (a.z != b.z) return a.z < b.z;
  if (a.y != b.y) return a.y < b.y;
  return a.x < b.x;
}

const flo_viscosity_cell* Find(const flo_viscosity_cell* cells, size_t n, flo_int3 c) {
  const flo_viscosity_cell* end = cells + n;
  const flo_viscosity_cell* it =
      std::lower_bound(cells, end, c, [](const flo_viscosity_cell& a, const flo_int3& k) { return CellLess(a.cell_index, k); });
  if (it == end || CellLess(c, it->cell_index)) return nullptr;
  return it;
}

flo_int3 Next(flo_int3 c, int axis) {
  if (axis == 0) ++c.x;
  if (axis == 1) ++c.y;
  if (axis == 2) ++c.z;
  return c;
}

// SetStrainTerms (:160-178): slots of the faces, coefficients -c, +c, -c, +c.
int Terms(const flo_int3* blocks, size_t nb, const int* axes, const flo_int3* lattices, int count, double c,
          flo_strain_row& row) {
  for (int t = 0; t < count; ++t) {
    uint64_t slot;
    if (int s = FaceSlot(blocks, nb, static_cast<uint8_t>(axes[t]), lattices[t], slot)) return s;
    row.terms[t] = {slot, t % 2 == 0 ? -c : c};
  }
  row.term_count = static_cast<uint64_t>(count);
  return FLO_OK;
}

}  // namespace

extern "C" {

int flo_assemble_full_cell_strain(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count,
                                  const flo_viscosity_cell* cells, size_t n, flo_strain_row* out, size_t capacity,
                                  size_t* out_count) {
  for (size_t i = 0; i < n; ++i) {
    const flo_viscosity_cell& c = cells[i];
    if (i > 0 && !CellLess(cells[i - 1].cell_index, c.cell_index)) return FLO_INVALID_ARGUMENT;
    if (!std::isfinite(c.dynamic_viscosity) || c.dynamic_viscosity < 0.0) return FLO_INVALID_ARGUMENT;
    constexpr int32_t kMax = std::numeric_limits<int32_t>::max();
    if (c.cell_index.x == kMax || c.cell_index.y == kMax || c.cell_index.z == kMax) return FLO_OUT_OF_RANGE;
  }
  if (n == 0) {
    *out_count = 0;
    return FLO_OK;
  }
  const double h = geometry->cell_width;
  const double normal = 1.0 / h, shear = 0.5 * normal;
  if (!std::isnormal(normal)) return FLO_OUT_OF_RANGE;
  std::vector<flo_strain_row> rows;
  rows.reserve(6 * n);
  for (size_t i = 0; i < n; ++i) {
    const flo_viscosity_cell& c = cells[i];
    const double w = StrainWeight(c.dynamic_viscosity, h, 1);
    if (c.dynamic_viscosity > 0.0 && !std::isnormal(w)) return FLO_OUT_OF_RANGE;
    for (int axis = 0; axis < 3; ++axis) {
      flo_strain_row row{};
      row.low_cell = c.cell_index;
      row.first_axis = row.second_axis = static_cast<uint8_t>(axis);
      row.quadrature_weight = w;
      const int axes[2] = {axis, axis};
      const flo_int3 lat[2] = {c.cell_index, Next(c.cell_index, axis)};
      if (int s = Terms(sorted_blocks, block_count, axes, lat, 2, normal, row)) return s;
      rows.push_back(row);
    }
  }
  for (size_t i = 0; i < n; ++i) {
    const flo_viscosity_cell& c = cells[i];
    for (int a = 0; a < 2; ++a) {
      const flo_int3 fi = Next(c.cell_index, a);
      const flo_viscosity_cell* first = Find(cells, n, fi);
      if (!first) continue;
      for (int b = a + 1; b < 3; ++b) {
        const flo_int3 si = Next(c.cell_index, b), di = Next(fi, b);
        const flo_viscosity_cell* second = Find(cells, n, si);
        const flo_viscosity_cell* diagonal = Find(cells, n, di);
        if (!second || !diagonal) continue;
        const double mu[4] = {c.dynamic_viscosity, first->dynamic_viscosity, second->dynamic_viscosity,
                              diagonal->dynamic_viscosity};
        if (!std::isnormal(shear)) return FLO_OUT_OF_RANGE;
        const double edge = HarmonicEdge(mu);
        if (!std::isfinite(edge)) return FLO_OUT_OF_RANGE;
        const double w = StrainWeight(edge, h, 2);
        if (edge > 0.0 && !std::isnormal(w)) return FLO_OUT_OF_RANGE;
        flo_strain_row row{};
        row.low_cell = c.cell_index;
        row.first_axis = static_cast<uint8_t>(a);
        row.second_axis = static_cast<uint8_t>(b);
        row.quadrature_weight = w;
        const int axes[4] = {a, a, b, b};
        const flo_int3 lat[4] = {fi, di, si, di};
        if (int s = Terms(sorted_blocks, block_count, axes, lat, 4, shear, row)) return s;
        rows.push_back(row);
      }
    }
  }
  if (rows.size() > capacity) return FLO_RESOURCE_EXHAUSTED;
  std::copy(rows.begin(), rows.end(), out);
  *out_count = rows.size();
  return FLO_OK;
}

int flo_solve_viscous_block(const flo_strain_row* rows, size_t row_count, const double* inertia, size_t face_count,
                            double dt, const double* rhs, uint64_t max_iterations, double relative_tolerance,
                            double* out_velocity, uint64_t* out_iterations, double* out_max_residual) {
  // ValidateVelocityBlockInputs
  if (!std::isfinite(dt) || dt < 0.0) return FLO_INVALID_ARGUMENT;
  if (max_iterations == 0 || !std::isfinite(relative_tolerance) || relative_tolerance <= 0.0 ||
      relative_tolerance >= 1.0)
    return FLO_INVALID_ARGUMENT;
  double max_rhs = 0.0;
  for (size_t f = 0; f < face_count; ++f) {
    if (!std::isfinite(inertia[f]) || inertia[f] < 0.0 || !std::isfinite(rhs[f])) return FLO_INVALID_ARGUMENT;
    if (inertia[f] == 0.0 && rhs[f] != 0.0) return FLO_INVALID_ARGUMENT;
    max_rhs = std::max(max_rhs, std::abs(rhs[f]));
  }
  // BuildWeightedVelocityRows
  std::vector<flo_strain_row> w(rows, rows + row_count);
  std::vector<double> diag(inertia, inertia + face_count);
  const double sqrt_dt = std::sqrt(dt);
  for (auto& row : w) {
    if (row.term_count == 0 || row.term_count > 8 || !std::isfinite(row.quadrature_weight) ||
        row.quadrature_weight < 0.0)
      return FLO_INVALID_ARGUMENT;
    const double sqrt_w = std::sqrt(row.quadrature_weight);
    for (uint64_t t = 0; t < row.term_count; ++t) {
      flo_strain_term& term = row.terms[t];
      if (term.face_slot >= face_count || !std::isfinite(term.coefficient)) return FLO_INVALID_ARGUMENT;
      for (uint64_t e = 0; e < t; ++e)
        if (row.terms[e].face_slot == term.face_slot) return FLO_INVALID_ARGUMENT;
      if (inertia[term.face_slot] == 0.0 || sqrt_dt == 0.0 || sqrt_w == 0.0 || term.coefficient == 0.0) {
        term.coefficient = 0.0;
        continue;
      }
      const double c = Scaled(sqrt_dt, sqrt_w, term.coefficient);
      if (!std::isnormal(c)) return FLO_OUT_OF_RANGE;
      term.coefficient = c;
      diag[term.face_slot] += c * c;
      if (!std::isfinite(diag[term.face_slot])) return FLO_OUT_OF_RANGE;
    }
  }
  // SolvePreparedVelocityBlock
  const size_t nf = face_count;
  std::vector<double> v(nf, 0.0), r(rhs, rhs + nf), z(nf), d(nf), q(nf);
  auto apply = [&](const std::vector<double>& x, std::vector<double>& y) {
    for (size_t f = 0; f < nf; ++f) y[f] = inertia[f] * x[f];
    for (const auto& row : w) {
      double s = 0.0;
      for (uint64_t t = 0; t < row.term_count; ++t) s += row.terms[t].coefficient * x[row.terms[t].face_slot];
      for (uint64_t t = 0; t < row.term_count; ++t) y[row.terms[t].face_slot] += row.terms[t].coefficient * s;
    }
    return std::all_of(y.begin(), y.end(), [](double a) { return std::isfinite(a); });
  };
  auto precondition = [&]() {
    for (size_t f = 0; f < nf; ++f) z[f] = inertia[f] == 0.0 ? 0.0 : r[f] / diag[f];
    return std::inner_product(r.begin(), r.end(), z.begin(), 0.0);
  };
  uint64_t iterations = 0;
  double max_residual = 0.0;
  if (max_rhs != 0.0) {
    const double tolerance = relative_tolerance * max_rhs;
    double rz = precondition();
    d = z;
    bool done = false;
    while (iterations < max_iterations) {
      if (!std::isfinite(rz) || rz <= 0.0 || !apply(d, q)) return FLO_OUT_OF_RANGE;
      const double curvature = std::inner_product(d.begin(), d.end(), q.begin(), 0.0);
      if (!std::isfinite(curvature) || curvature <= 0.0) return FLO_ABORTED;
      const double step = rz / curvature;
      if (!std::isnormal(step)) return FLO_OUT_OF_RANGE;
      double mr = 0.0;
      for (size_t f = 0; f < nf; ++f) {
        v[f] += step * d[f];
        r[f] -= step * q[f];
        if (!std::isfinite(v[f]) || !std::isfinite(r[f])) return FLO_OUT_OF_RANGE;
        mr = std::max(mr, std::abs(r[f]));
      }
      ++iterations;
      const bool check = mr <= tolerance || iterations == max_iterations;
      if (check) {
        if (!apply(v, q)) return FLO_OUT_OF_RANGE;
        mr = 0.0;
        for (size_t f = 0; f < nf; ++f) {
          r[f] = rhs[f] - q[f];
          if (!std::isfinite(r[f])) return FLO_OUT_OF_RANGE;
          mr = std::max(mr, std::abs(r[f]));
        }
        if (mr <= tolerance) {
          max_residual = mr;
          done = true;
          break;
        }
      }
      const double next = precondition();
      const double beta = check ? 0.0 : next / rz;
      for (size_t f = 0; f < nf; ++f) d[f] = z[f] + beta * d[f];
      rz = next;
    }
    if (!done) return FLO_ABORTED;
  }
  std::copy(v.begin(), v.end(), out_velocity);
  *out_iterations = iterations;
  *out_max_residual = max_residual;
  return FLO_OK;
}

}  // extern "C"
