// Variational MAC viscosity on gfx950 (native/src/viscosity.cc).
//
// Strain assembly (AssembleFullCellMacStrain3D, :762-946): one lane per cell writes its three normal rows
// (rows 3i .. 3i+2), counts its interior shear rows (cell, +first, +second and the diagonal cell all
// present, pairs xy, xz, yz), and after an exclusive scan of the counts writes them at 3n + offset, so
// the rows come out in the reference's order. Cells are found by binary search in the sorted cell list,
// face slots through the grid's dense block table. Every failure reports an ordinal that follows the
// reference's sequential order (all cell checks, then the normal rows cell by cell, then the shear rows),
// and the smallest ordinal is the call's status.
//
// Velocity block (SolveMacViscousVelocityBlock3D, :1331-1650): Jacobi PCG on (M + dt E^T W E) u = rhs in
// FP64. Two passes over the caller's rows (staged through LDS) pack the operator: the nonzero weighted
// terms in row order (face slot, coefficient) and, per face, its incidence in row order (row,
// coefficient). An apply is then a row pass forming every weighted strain and a face pass adding to M u
// the strains of the rows that hold the face. The PCG scalars (r.z, curvature, step, beta) and the
// iteration's checks stay on the device; the host reads them once per iteration for the stopping test.
// Dots and maxima reduce per workgroup and then in one fixed-order workgroup pass, so every run gives
// the same bits.
#include "context.h"

namespace fl {
namespace {

constexpr int kReduceGroups = 1024;

__device__ inline bool normal_d(double v) { return __builtin_isnormal(v); }

__device__ inline void report_min(uint64_t* err, uint64_t ordinal, int code) {
  atomicMin(reinterpret_cast<unsigned long long*>(err), static_cast<unsigned long long>(error_word(ordinal, code)));
}

// FactorProduct (viscosity.cc:36-51) with divisor 1: fractions multiplied, exponents summed.
struct Binary {
  double f;
  int e;
};
__device__ inline Binary factor3(double a, double b, double c) {
  if (a == 0.0 || b == 0.0 || c == 0.0) return {0.0, 0};
  int ea, eb, ec, ed;
  const double fa = frexp(a, &ea), fb = frexp(b, &eb), fc = frexp(c, &ec), fd = frexp(1.0, &ed);
  return {(fa * (fb * fc)) / fd, ea + eb + ec - ed};
}
__device__ inline double scaled3(double a, double b, double c) {
  const Binary p = factor3(a, b, c);
  return scalbn(p.f, p.e);
}
// StrainWeight3D (:150-158): mu h^3 2^k.
__device__ inline double strain_weight(double mu, double h, int k) {
  int eh;
  const double fh = frexp(h, &eh);
  const Binary area = factor3(mu, h, h);
  return scalbn(area.f * fh, area.e + eh + k);
}
// HarmonicEdgeViscosity (:91-103).
__device__ inline double harmonic_edge(const double mu[4]) {
  const double m = fmin(fmin(mu[0], mu[1]), fmin(mu[2], mu[3]));
  if (m == 0.0) return 0.0;
  double sum = 0.0;
  for (int i = 0; i < 4; ++i) sum += m / mu[i];
  return m * (4.0 / sum);
}

__device__ inline bool cell_less(const fl_int3& a, const fl_int3& b) {
  if (a.z != b.z) return a.z < b.z;
  if (a.y != b.y) return a.y < b.y;
  return a.x < b.x;
}

// Ordinal of cell c in the (z,y,x)-sorted list, or -1.
__device__ inline int64_t find_cell(const fl_viscosity_cell* cells, uint64_t n, fl_int3 c) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (cell_less(cells[mid].cell_index, c)) lo = mid + 1;
    else hi = mid;
  }
  if (lo < n && !cell_less(c, cells[lo].cell_index)) return static_cast<int64_t>(lo);
  return -1;
}

__device__ inline fl_int3 next(fl_int3 c, int axis) {
  if (axis == 0) ++c.x;
  if (axis == 1) ++c.y;
  if (axis == 2) ++c.z;
  return c;
}

__device__ inline int32_t floor4(int32_t v) { return v >= 0 ? v / 4 : -1 - (-(v + 1)) / 4; }

// SparseMacGridView3D::FaceSlot through the dense block table; FailedPrecondition without an owner.
__device__ inline int face_slot(const GridView& g, int axis, fl_int3 l, uint64_t& slot) {
  const int32_t bx = floor4(l.x), by = floor4(l.y), bz = floor4(l.z);
  const int32_t ord = g.ordinal(bx, by, bz);
  if (ord < 0) return kFailedPrecondition;
  slot = uint64_t(ord) * 192 + uint64_t(axis) * 64 +
         uint64_t(((l.z - 4 * bz) * 4 + (l.y - 4 * by)) * 4 + (l.x - 4 * bx));
  return kOk;
}

// SetStrainTerms (:160-178): -c, +c, -c, +c.
__device__ inline int set_terms(const GridView& g, const int* axes, const fl_int3* lattices, int count, double c,
                                fl_strain_row& row) {
  for (int t = 0; t < count; ++t) {
    uint64_t slot;
    if (int code = face_slot(g, axes[t], lattices[t], slot)) return code;
    row.terms[t].face_slot = slot;
    row.terms[t].coefficient = t % 2 == 0 ? -c : c;
  }
  row.term_count = static_cast<uint64_t>(count);
  return kOk;
}

// Phase 0 (ordinals 0 .. n-1): order, viscosity and representable high faces, per cell.
__global__ void k_strain_check(const fl_viscosity_cell* cells, uint64_t n, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const fl_viscosity_cell c = cells[i];
  int code = kOk;
  if (i > 0 && !cell_less(cells[i - 1].cell_index, c.cell_index)) code = kInvalidArgument;
  else if (!finite(c.dynamic_viscosity) || c.dynamic_viscosity < 0.0) code = kInvalidArgument;
  else if (c.cell_index.x == INT32_MAX || c.cell_index.y == INT32_MAX || c.cell_index.z == INT32_MAX)
    code = kOutOfRange;
  if (code) report_min(err, i, code);
}

// Phase 1 (ordinals n + 4i + k): the weight check, then the three normal rows of cell i.
__global__ void k_strain_normal(GridView g, const fl_viscosity_cell* cells, uint64_t n, double h, double normal,
                                fl_strain_row* rows, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const fl_viscosity_cell c = cells[i];
  const double w = strain_weight(c.dynamic_viscosity, h, 1);
  if (c.dynamic_viscosity > 0.0 && !normal_d(w)) {
    report_min(err, n + 4 * i, kOutOfRange);
    return;
  }
  for (int axis = 0; axis < 3; ++axis) {
    fl_strain_row row{};
    row.low_cell = c.cell_index;
    row.first_axis = row.second_axis = static_cast<uint8_t>(axis);
    row.quadrature_weight = w;
    const int axes[2] = {axis, axis};
    const fl_int3 lat[2] = {c.cell_index, next(c.cell_index, axis)};
    if (int code = set_terms(g, axes, lat, 2, normal, row)) {
      report_min(err, n + 4 * i + 1 + axis, code);
      return;
    }
    rows[3 * i + axis] = row;
  }
}

// The interior shear pairs of cell i as a bit set: bit p for pair p of (x,y), (x,z), (y,z).
__device__ inline uint32_t shear_pairs(const fl_viscosity_cell* cells, uint64_t n, fl_int3 c, int64_t nb[3][3]) {
  uint32_t bits = 0;
  int p = 0;
  for (int a = 0; a < 2; ++a) {
    const fl_int3 fi = next(c, a);
    const int64_t first = find_cell(cells, n, fi);
    for (int b = a + 1; b < 3; ++b, ++p) {
      if (first < 0) continue;
      const int64_t second = find_cell(cells, n, next(c, b));
      const int64_t diagonal = find_cell(cells, n, next(fi, b));
      if (second < 0 || diagonal < 0) continue;
      bits |= 1u << p;
      nb[p][0] = first;
      nb[p][1] = second;
      nb[p][2] = diagonal;
    }
  }
  return bits;
}

__global__ void k_strain_shear_count(const fl_viscosity_cell* cells, uint64_t n, uint64_t* counts) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  int64_t nb[3][3];
  counts[i] = __popc(shear_pairs(cells, n, cells[i].cell_index, nb));
}

// Phase 2 (ordinals 5n + 3i + p): the shear rows of cell i at 3n + offset[i].
__global__ void k_strain_shear(GridView g, const fl_viscosity_cell* cells, uint64_t n, double h, double shear,
                               const uint64_t* offset, fl_strain_row* rows, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const fl_viscosity_cell c = cells[i];
  int64_t nb[3][3];
  const uint32_t bits = shear_pairs(cells, n, c.cell_index, nb);
  uint64_t out = 3 * n + offset[i];
  int p = 0;
  for (int a = 0; a < 2; ++a)
    for (int b = a + 1; b < 3; ++b, ++p) {
      if (!((bits >> p) & 1)) continue;
      const uint64_t ordinal = 5 * n + 3 * i + p;
      const double mu[4] = {c.dynamic_viscosity, cells[nb[p][0]].dynamic_viscosity, cells[nb[p][1]].dynamic_viscosity,
                            cells[nb[p][2]].dynamic_viscosity};
      if (!normal_d(shear)) {
        report_min(err, ordinal, kOutOfRange);
        return;
      }
      const double edge = harmonic_edge(mu);
      if (!finite(edge)) {
        report_min(err, ordinal, kOutOfRange);
        return;
      }
      const double w = strain_weight(edge, h, 2);
      if (edge > 0.0 && !normal_d(w)) {
        report_min(err, ordinal, kOutOfRange);
        return;
      }
      fl_strain_row row{};
      row.low_cell = c.cell_index;
      row.first_axis = static_cast<uint8_t>(a);
      row.second_axis = static_cast<uint8_t>(b);
      row.quadrature_weight = w;
      const fl_int3 fi = next(c.cell_index, a), si = next(c.cell_index, b), di = next(fi, b);
      const int axes[4] = {a, a, b, b};
      const fl_int3 lat[4] = {fi, di, si, di};
      if (int code = set_terms(g, axes, lat, 4, shear, row)) {
        report_min(err, ordinal, code);
        return;
      }
      rows[out++] = row;
    }
}

// ---------------------------------------------------------------- velocity block

// Rows of one workgroup staged through LDS with coalesced 8-byte loads (a row is 21 words), so that the
// row passes read HBM in full lines instead of one 168-byte stride per lane.
constexpr int kRowWords = sizeof(fl_strain_row) / sizeof(uint64_t);
static_assert(sizeof(fl_strain_row) == 168 && kRowWords == 21);
__device__ inline const fl_strain_row& stage_row(const fl_strain_row* rows, uint64_t nr, fl_strain_row* lds) {
  const uint64_t base = blockIdx.x * uint64_t(blockDim.x);
  const uint64_t count = base < nr ? min(uint64_t(blockDim.x), nr - base) : 0;
  const uint64_t* src = reinterpret_cast<const uint64_t*>(rows + base);
  uint64_t* dst = reinterpret_cast<uint64_t*>(lds);
  for (uint64_t i = threadIdx.x; i < count * kRowWords; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
  return lds[threadIdx.x];
}

// Weighted coefficient of term t of a checked row (BuildWeightedVelocityRows, :1360-1406): sqrt(dt)
// sqrt(w) c, zero on an eliminated face or a zero factor; an unrepresentable product reports OutOfRange
// at the term and counts as zero.
__device__ inline double weighted_coefficient(const fl_strain_row& row, int t, double sqrt_w, const double* inertia,
                                              double sqrt_dt, uint64_t r, uint64_t* err) {
  const uint64_t slot = row.terms[t].face_slot;
  const double c = row.terms[t].coefficient;
  if (inertia[slot] == 0.0 || sqrt_dt == 0.0 || sqrt_w == 0.0 || c == 0.0) return 0.0;
  const double wc = scaled3(sqrt_dt, sqrt_w, c);
  if (!normal_d(wc)) {
    if (err) report_min(err, 9 * r + 1 + t, kOutOfRange);
    return 0.0;
  }
  return wc;
}

// First row pass, per row, ordinals 9r + (0 row, 1 + t term): the row and term checks, then the nonzero
// weighted terms counted per row and per face. The rows are not modified.
__global__ __launch_bounds__(kThreads) void k_count_rows(const fl_strain_row* rows, uint64_t nr, const double* inertia,
                                                         uint64_t nf, double sqrt_dt, uint32_t* inc_count,
                                                         uint64_t* row_nnz, uint64_t* err) {
  __shared__ fl_strain_row lds[kThreads];
  const fl_strain_row& row = stage_row(rows, nr, lds);
  const uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (r >= nr) return;
  row_nnz[r] = 0;
  if (row.term_count == 0 || row.term_count > 8 || !finite(row.quadrature_weight) || row.quadrature_weight < 0.0) {
    report_min(err, 9 * r, kInvalidArgument);
    return;
  }
  const int count = static_cast<int>(row.term_count);
  for (int t = 0; t < count; ++t) {
    const uint64_t slot = row.terms[t].face_slot;
    if (slot >= nf || !finite(row.terms[t].coefficient)) {
      report_min(err, 9 * r + 1 + t, kInvalidArgument);
      return;
    }
    for (int e = 0; e < t; ++e)
      if (row.terms[e].face_slot == slot) {
        report_min(err, 9 * r + 1 + t, kInvalidArgument);
        return;
      }
  }
  const double sqrt_w = sqrt(row.quadrature_weight);
  uint64_t nnz = 0;
  for (int t = 0; t < count; ++t)
    if (weighted_coefficient(row, t, sqrt_w, inertia, sqrt_dt, r, err) != 0.0) {
      atomicAdd(&inc_count[row.terms[t].face_slot], 1u);
      ++nnz;
    }
  row_nnz[r] = nnz;
}

// ptr[n] = ptr[n - 1] + count[n - 1]: the total after an exclusive scan.
__global__ void k_scan_tail(const uint64_t* count, uint64_t n, uint64_t* ptr) {
  if (threadIdx.x == 0) ptr[n] = n ? ptr[n - 1] + count[n - 1] : 0;
}

// Second row pass (checked rows): the nonzero weighted terms packed in row and term order, and each
// packed term k of row r entered into its face's incidence as (k << 32 | r), unordered.
__global__ __launch_bounds__(kThreads) void k_pack_rows(const fl_strain_row* rows, uint64_t nr, const double* inertia,
                                                        double sqrt_dt, const uint64_t* row_ptr,
                                                        const uint64_t* inc_ptr, uint32_t* cursor,
                                                        uint32_t* term_slot, double* term_coef, uint64_t* entries) {
  __shared__ fl_strain_row lds[kThreads];
  const fl_strain_row& row = stage_row(rows, nr, lds);
  const uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (r >= nr) return;
  const double sqrt_w = sqrt(row.quadrature_weight);
  uint64_t k = row_ptr[r];
  const int count = static_cast<int>(row.term_count);
  for (int t = 0; t < count && t < 8; ++t) {
    const double c = weighted_coefficient(row, t, sqrt_w, inertia, sqrt_dt, r, nullptr);
    if (c == 0.0) continue;
    const uint64_t slot = row.terms[t].face_slot;
    term_slot[k] = static_cast<uint32_t>(slot);
    term_coef[k] = c;
    const uint32_t i = atomicAdd(&cursor[slot], 1u);
    entries[inc_ptr[slot] + i] = (k << 32) | r;
    ++k;
  }
}

// Sorts each face's incidence by packed term (row, then term order), stores it as (row, coefficient)
// and forms the Jacobi diagonal inertia + sum c^2 in that order; a diagonal that leaves the finite range
// reports OutOfRange at the term that made it so.
__global__ void k_incidence_diagonal(const fl_strain_row* rows, const double* inertia, uint64_t nf,
                                     const uint64_t* inc_ptr, uint64_t* entries, const double* term_coef,
                                     uint32_t* inc_row, double* inc_coef, double* diag, uint64_t* err) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f >= nf) return;
  const uint64_t base = inc_ptr[f];
  uint64_t* e = entries + base;
  const uint32_t m = static_cast<uint32_t>(inc_ptr[f + 1] - base);
  for (uint32_t i = 1; i < m; ++i) {
    const uint64_t key = e[i];
    uint32_t j = i;
    while (j > 0 && e[j - 1] > key) {
      e[j] = e[j - 1];
      --j;
    }
    e[j] = key;
  }
  double d = inertia[f];
  for (uint32_t i = 0; i < m; ++i) {
    const uint32_t r = static_cast<uint32_t>(e[i]);
    const double c = term_coef[e[i] >> 32];
    inc_row[base + i] = r;
    inc_coef[base + i] = c;
    d += c * c;
    if (!finite(d)) {
      int t = 0;
      while (t < 7 && rows[r].terms[t].face_slot != f) ++t;
      report_min(err, 9 * uint64_t(r) + 1 + t, kOutOfRange);
      break;
    }
  }
  diag[f] = d;
}

// Weighted strain of every row: sum of c x[slot] over its packed terms in term order (a zero term of the
// row would add an exact zero and is not packed).
__global__ void k_row_strain(const uint64_t* row_ptr, const uint32_t* term_slot, const double* term_coef, uint64_t nr,
                             const double* x, double* strain) {
  const uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (r >= nr) return;
  const uint64_t end = row_ptr[r + 1];
  double s = 0.0;
  for (uint64_t k = row_ptr[r]; k < end; ++k) s += term_coef[k] * x[term_slot[k]];
  strain[r] = s;
}

// Workgroup partial sums written to partial[blockIdx] (256 lanes, LDS tree).
__device__ inline void block_sum(double v, double* partial) {
  __shared__ double red[kThreads];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
__device__ inline void block_max(double v, double* partial) {
  __shared__ double red[kThreads];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// y = M x + E^T (strain), the strains of a face's rows added in row order; bit 0 of flag[1] if any y is
// not finite. With DOT, partial[blockIdx] holds the workgroup's sum of x . y (grid-stride over faces).
template <bool DOT>
__global__ __launch_bounds__(kThreads) void k_face_apply(const double* inertia, uint64_t nf, const uint64_t* inc_ptr,
                                                         const uint32_t* inc_row, const double* inc_coef,
                                                         const double* strain, const double* x, double* y,
                                                         uint32_t* flag, double* partial) {
  double dot = 0.0;
  for (uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; f < nf; f += uint64_t(gridDim.x) * blockDim.x) {
    const double xf = x[f];
    double m = inertia[f] * xf;
    const uint64_t end = inc_ptr[f + 1];
    for (uint64_t i = inc_ptr[f]; i < end; ++i) m += inc_coef[i] * strain[inc_row[i]];
    y[f] = m;
    if (!finite(m)) atomicOr(flag + 1, 1u);
    dot += xf * m;
  }
  if (DOT) block_sum(dot, partial);
}

// Device scalars of the PCG: scal[0] r.z, [1] curvature d.q, [2] max |r|, [3] step, [4] beta. flag[0] is
// the first failure status (0 none), flag[1] bit 0 a non-finite iterate since the last conversion.
__device__ inline void fail(uint32_t* flag, uint32_t code) { atomicCAS(flag, 0u, code); }

// Fixed-order reduction of g partials by one workgroup: lane i folds i, i + 256, ..., then an LDS tree.
__device__ inline double reduce_partials(const double* partial, int g, bool max_mode) {
  __shared__ double red[kThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < g; i += kThreads) acc = max_mode ? fmax(acc, partial[i]) : acc + partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      red[threadIdx.x] = max_mode ? fmax(red[threadIdx.x], red[threadIdx.x + s]) : red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

// Sum of g partials into scal[slot].
__global__ __launch_bounds__(kThreads) void k_finish(const double* partial, int g, int max_mode, double* scal,
                                                     int slot) {
  const double v = reduce_partials(partial, g, max_mode != 0);
  if (threadIdx.x == 0) scal[slot] = v;
}

// After the direction's apply: the curvature, the iteration's checks in the reference's order (r.z
// positive and finite with a finite apply, else OutOfRange; a positive finite curvature, else Aborted;
// a normal step, else OutOfRange) and the step rz / curvature.
__global__ __launch_bounds__(kThreads) void k_finish_curvature(const double* partial, int g, double* scal,
                                                               uint32_t* flag) {
  const double curvature = reduce_partials(partial, g, false);
  if (threadIdx.x != 0) return;
  scal[1] = curvature;
  scal[3] = 0.0;
  const double rz = scal[0];
  if (!isfinite(rz) || rz <= 0.0 || (flag[1] & 1u)) return fail(flag, kOutOfRange);
  if (!isfinite(curvature) || curvature <= 0.0) return fail(flag, kAborted);
  const double step = rz / curvature;
  if (!normal_d(step)) return fail(flag, kOutOfRange);
  scal[3] = step;
}

// v += step d, r -= step q (OutOfRange on a non-finite iterate); then z = r / diag on positive-inertia
// faces (0 elsewhere). Partials: max |r| in partial[0, g), r . z in partial[g, 2g).
__global__ __launch_bounds__(kThreads) void k_step_precondition(const double* scal, const double* d, const double* q,
                                                                const double* inertia, const double* diag,
                                                                uint64_t nf, double* v, double* r, double* z,
                                                                double* partial, uint32_t* flag) {
  const double step = scal[3];
  double mx = 0.0, rz = 0.0;
  for (uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; f < nf; f += uint64_t(gridDim.x) * blockDim.x) {
    const double vf = v[f] + step * d[f];
    const double rf = r[f] - step * q[f];
    v[f] = vf;
    r[f] = rf;
    if (!finite(vf) || !finite(rf)) atomicOr(flag + 1, 1u);
    mx = fmax(mx, fabs(rf));
    const double zf = inertia[f] == 0.0 ? 0.0 : rf / diag[f];
    z[f] = zf;
    rz += rf * zf;
  }
  block_max(mx, partial);
  block_sum(rz, partial + gridDim.x);
}

// After the step: max |r| into scal[2], a non-finite iterate into the status, and the next r.z with
// beta = next / previous (scal[0] becomes the next r.z).
__global__ __launch_bounds__(kThreads) void k_finish_step(const double* partial, int g, double* scal, uint32_t* flag) {
  const double mr = reduce_partials(partial, g, true);
  __syncthreads();
  const double rz = reduce_partials(partial + g, g, false);
  if (threadIdx.x != 0) return;
  scal[2] = mr;
  if (flag[1] & 1u) fail(flag, kOutOfRange);
  scal[4] = rz / scal[0];
  scal[0] = rz;
}

// z = r / diag on positive-inertia faces (0 elsewhere); partial r . z.
__global__ void k_precondition(const double* r, const double* inertia, const double* diag, uint64_t nf, double* z,
                               double* partial) {
  double acc = 0.0;
  for (uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; f < nf; f += uint64_t(gridDim.x) * blockDim.x) {
    const double zf = inertia[f] == 0.0 ? 0.0 : r[f] / diag[f];
    z[f] = zf;
    acc += r[f] * zf;
  }
  block_sum(acc, partial);
}

// r = rhs - q; partial max |r|; bit 0 of flag[1] on a non-finite residual.
__global__ void k_true_residual_f64(const double* rhs, const double* q, uint64_t nf, double* r, double* partial,
                                    uint32_t* flag) {
  double mx = 0.0;
  for (uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; f < nf; f += uint64_t(gridDim.x) * blockDim.x) {
    const double rf = rhs[f] - q[f];
    r[f] = rf;
    if (!finite(rf)) atomicOr(flag + 1, 1u);
    mx = fmax(mx, fabs(rf));
  }
  block_max(mx, partial);
}

// d = z + beta d, beta = scal[4] (0 after a residual recomputation).
__global__ void k_direction_f64(const double* z, const double* scal, int use_beta, uint64_t nf, double* d) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  const double beta = use_beta ? scal[4] : 0.0;
  if (f < nf) d[f] = z[f] + beta * d[f];
}

__global__ void k_copy_d(const double* a, uint64_t n, double* b) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_fill_d(double* a, uint64_t n, double v) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_fill_u32(uint32_t* a, uint64_t n) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void k_widen(const uint32_t* a, uint64_t n, uint64_t* b) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) b[i] = a[i];
}

}  // namespace

int assemble_strain(fl_ctx* ctx, const ViscosityDevice& V, uint64_t n, double h, uint64_t* total) {
  hipStream_t s = ctx->stream;
  const double normal = 1.0 / h, shear = 0.5 * normal;
  *total = 0;
  if (n == 0) return FL_OK;
  FL_TRY(reset_error(ctx));
  k_strain_check<<<grid_for(n), kThreads, 0, s>>>(V.cells, n, ctx->err);
  FL_HIP(hipGetLastError());
  uint64_t word = 0;
  FL_TRY(read_error(ctx, &word));
  if (word != ~uint64_t(0)) return static_cast<int>(word & 0xff);
  if (!std::isnormal(normal)) return FL_OUT_OF_RANGE;
  const GridView g = grid_view(ctx);
  k_strain_normal<<<grid_for(n), kThreads, 0, s>>>(g, V.cells, n, h, normal, V.rows, ctx->err);
  k_strain_shear_count<<<grid_for(n), kThreads, 0, s>>>(V.cells, n, V.counts);
  FL_HIP(hipGetLastError());
  FL_TRY(exclusive_scan_u64(ctx, V.counts, V.offsets, n));
  uint64_t last[2] = {0, 0};
  FL_HIP(hipMemcpyAsync(&last[0], V.counts + n - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FL_HIP(hipMemcpyAsync(&last[1], V.offsets + n - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  const uint64_t nshear = last[0] + last[1];
  if (3 * n + nshear > V.row_capacity) return FL_RESOURCE_EXHAUSTED;
  k_strain_shear<<<grid_for(n), kThreads, 0, s>>>(g, V.cells, n, h, shear, V.offsets, V.rows, ctx->err);
  FL_HIP(hipGetLastError());
  FL_TRY(read_error(ctx, &word));
  if (word != ~uint64_t(0)) return static_cast<int>(word & 0xff);
  *total = 3 * n + nshear;
  return FL_OK;
}

int weight_viscous_rows(fl_ctx* ctx, const ViscosityDevice& V, uint64_t nr, uint64_t nf, double dt, uint64_t* terms) {
  hipStream_t s = ctx->stream;
  *terms = 0;
  FL_TRY(reset_error(ctx));
  k_fill_u32<<<grid_for(nf), kThreads, 0, s>>>(V.inc_count, nf);
  if (nr) k_count_rows<<<grid_for(nr), kThreads, 0, s>>>(V.rows, nr, V.inertia, nf, std::sqrt(dt), V.inc_count,
                                                         V.row_nnz, ctx->err);
  FL_HIP(hipGetLastError());
  uint64_t word = 0;
  FL_TRY(read_error(ctx, &word));
  if (word != ~uint64_t(0)) return static_cast<int>(word & 0xff);
  FL_TRY(exclusive_scan_u64(ctx, V.row_nnz, V.row_ptr, nr));
  k_scan_tail<<<1, 64, 0, s>>>(V.row_nnz, nr, V.row_ptr);
  k_widen<<<grid_for(nf), kThreads, 0, s>>>(V.inc_count, nf, V.counts);
  FL_TRY(exclusive_scan_u64(ctx, V.counts, V.inc_ptr, nf));
  k_scan_tail<<<1, 64, 0, s>>>(V.counts, nf, V.inc_ptr);
  FL_HIP(hipGetLastError());
  FL_HIP(hipMemcpyAsync(terms, V.row_ptr + nr, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  return FL_OK;
}

int pack_viscous_block(fl_ctx* ctx, const ViscosityDevice& V, uint64_t nr, uint64_t nf, uint64_t nt, double dt) {
  hipStream_t s = ctx->stream;
  // Packed term ordinals and row ordinals share one 64-bit incidence key.
  if (nt > V.term_capacity || nt >= (uint64_t(1) << 32)) return FL_RESOURCE_EXHAUSTED;
  uint64_t word = 0;
  if (nf) {
    k_fill_u32<<<grid_for(nf), kThreads, 0, s>>>(V.cursor, nf);
    if (nr)
      k_pack_rows<<<grid_for(nr), kThreads, 0, s>>>(V.rows, nr, V.inertia, std::sqrt(dt), V.row_ptr, V.inc_ptr,
                                                    V.cursor, V.term_slot, V.term_coef, V.entries);
    k_incidence_diagonal<<<grid_for(nf), kThreads, 0, s>>>(V.rows, V.inertia, nf, V.inc_ptr, V.entries, V.term_coef,
                                                           V.inc_row, V.inc_coef, V.diag, ctx->err);
    FL_HIP(hipGetLastError());
    FL_TRY(read_error(ctx, &word));
    if (word != ~uint64_t(0)) return static_cast<int>(word & 0xff);
  }
  return FL_OK;
}

int solve_viscous_block(fl_ctx* ctx, const ViscosityDevice& V, uint64_t nr, uint64_t nf, uint64_t nt, double dt,
                        double max_rhs, uint64_t max_iterations, double relative_tolerance, uint64_t* iterations,
                        double* max_residual) {
  *iterations = 0;
  *max_residual = 0.0;
  FL_TRY(pack_viscous_block(ctx, V, nr, nf, nt, dt));
  return pcg_viscous_block(ctx, V, nr, nf, max_rhs, max_iterations, relative_tolerance, iterations, max_residual);
}

int pcg_viscous_block(fl_ctx* ctx, const ViscosityDevice& V, uint64_t nr, uint64_t nf, double max_rhs,
                      uint64_t max_iterations, double relative_tolerance, uint64_t* iterations,
                      double* max_residual) {
  hipStream_t s = ctx->stream;
  *iterations = 0;
  *max_residual = 0.0;
  if (nf) k_fill_d<<<grid_for(nf), kThreads, 0, s>>>(V.v, nf, 0.0);
  if (max_rhs == 0.0) return FL_OK;
  const double tolerance = relative_tolerance * max_rhs;
  const unsigned rg = static_cast<unsigned>(std::min<uint64_t>(kReduceGroups, (nf + kThreads - 1) / kThreads));
  const unsigned fg = grid_for(nf);
  if (2 * uint64_t(rg) > V.partial_capacity) return FL_RESOURCE_EXHAUSTED;
  double* scal = V.scal;
  uint32_t status[2] = {0, 0};
  double host[5] = {};
  // Reads the status words and the scalars; one synchronization.
  auto sync = [&]() -> int {
    FL_HIP(hipMemcpyAsync(status, V.flag, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    FL_HIP(hipMemcpyAsync(host, scal, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    return FL_OK;
  };
  auto apply = [&](const double* x, double* y, bool dot) {
    if (nr) k_row_strain<<<grid_for(nr), kThreads, 0, s>>>(V.row_ptr, V.term_slot, V.term_coef, nr, x, V.strain);
    if (dot)
      k_face_apply<true><<<rg, kThreads, 0, s>>>(V.inertia, nf, V.inc_ptr, V.inc_row, V.inc_coef, V.strain, x, y,
                                                 V.flag, V.partial);
    else
      k_face_apply<false><<<fg, kThreads, 0, s>>>(V.inertia, nf, V.inc_ptr, V.inc_row, V.inc_coef, V.strain, x, y,
                                                  V.flag, V.partial);
  };
  k_fill_u32<<<1, 64, 0, s>>>(V.flag, 2);
  k_copy_d<<<grid_for(nf), kThreads, 0, s>>>(V.rhs, nf, V.r);
  k_precondition<<<rg, kThreads, 0, s>>>(V.r, V.inertia, V.diag, nf, V.z, V.partial);
  k_finish<<<1, kThreads, 0, s>>>(V.partial, static_cast<int>(rg), 0, scal, 0);
  k_copy_d<<<grid_for(nf), kThreads, 0, s>>>(V.z, nf, V.d);
  FL_HIP(hipGetLastError());
  while (*iterations < max_iterations) {
    apply(V.d, V.q, true);
    k_finish_curvature<<<1, kThreads, 0, s>>>(V.partial, static_cast<int>(rg), scal, V.flag);
    k_step_precondition<<<rg, kThreads, 0, s>>>(scal, V.d, V.q, V.inertia, V.diag, nf, V.v, V.r, V.z, V.partial,
                                                V.flag);
    k_finish_step<<<1, kThreads, 0, s>>>(V.partial, static_cast<int>(rg), scal, V.flag);
    FL_HIP(hipGetLastError());
    FL_TRY(sync());
    if (status[0]) return static_cast<int>(status[0]);
    ++*iterations;
    const double mr = host[2];
    const bool check = mr <= tolerance || *iterations == max_iterations;
    if (check) {
      apply(V.v, V.q, false);
      k_true_residual_f64<<<rg, kThreads, 0, s>>>(V.rhs, V.q, nf, V.r, V.partial, V.flag);
      k_finish<<<1, kThreads, 0, s>>>(V.partial, static_cast<int>(rg), 1, scal, 2);
      FL_HIP(hipGetLastError());
      FL_TRY(sync());
      if (status[1] & 1u) return FL_OUT_OF_RANGE;
      if (host[2] <= tolerance) {
        *max_residual = host[2];
        return FL_OK;
      }
      k_precondition<<<rg, kThreads, 0, s>>>(V.r, V.inertia, V.diag, nf, V.z, V.partial);
      k_finish<<<1, kThreads, 0, s>>>(V.partial, static_cast<int>(rg), 0, scal, 0);
    }
    k_direction_f64<<<grid_for(nf), kThreads, 0, s>>>(V.z, scal, check ? 0 : 1, nf, V.d);
    FL_HIP(hipGetLastError());
  }
  return FL_ABORTED;
}

}  // namespace fl
