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
// FP64. The operator is applied in gather form: a row pass forms every weighted strain, then a face pass
// adds to M u the strains of the rows that hold the face, in row order, from a face -> (row, term)
// incidence built on the device. Dots and maxima reduce per workgroup and then in one fixed-order pass,
// so every run gives the same bits.
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

// BuildWeightedVelocityRows (:1360-1406), per row, ordinals 9r + (0 row, 1 + t term): the row and term
// checks, then each term's coefficient sqrt(dt) sqrt(w) c (zero on an eliminated face or a zero
// factor); terms with a nonzero coefficient are counted against their face.
__global__ void k_weight_rows(fl_strain_row* rows, uint64_t nr, const double* inertia, uint64_t nf, double sqrt_dt,
                              uint32_t* inc_count, uint64_t* err) {
  const uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (r >= nr) return;
  fl_strain_row& row = rows[r];
  if (row.term_count == 0 || row.term_count > 8 || !finite(row.quadrature_weight) || row.quadrature_weight < 0.0) {
    report_min(err, 9 * r, kInvalidArgument);
    return;
  }
  const double sqrt_w = sqrt(row.quadrature_weight);
  const int count = static_cast<int>(row.term_count);
  for (int t = 0; t < count; ++t) {
    const uint64_t slot = row.terms[t].face_slot;
    const double c = row.terms[t].coefficient;
    if (slot >= nf || !finite(c)) {
      report_min(err, 9 * r + 1 + t, kInvalidArgument);
      return;
    }
    for (int e = 0; e < t; ++e)
      if (row.terms[e].face_slot == slot) {
        report_min(err, 9 * r + 1 + t, kInvalidArgument);
        return;
      }
  }
  for (int t = 0; t < count; ++t) {
    const uint64_t slot = row.terms[t].face_slot;
    const double c = row.terms[t].coefficient;
    if (inertia[slot] == 0.0 || sqrt_dt == 0.0 || sqrt_w == 0.0 || c == 0.0) {
      row.terms[t].coefficient = 0.0;
      continue;
    }
    const double wc = scaled3(sqrt_dt, sqrt_w, c);
    if (!normal_d(wc)) {
      report_min(err, 9 * r + 1 + t, kOutOfRange);
      row.terms[t].coefficient = 0.0;
      continue;
    }
    row.terms[t].coefficient = wc;
    atomicAdd(&inc_count[slot], 1u);
  }
}

__global__ void k_fill_incidence(const fl_strain_row* rows, uint64_t nr, const uint64_t* inc_offset, uint32_t* cursor,
                                 uint64_t* entries) {
  const uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (r >= nr) return;
  const fl_strain_row& row = rows[r];
  const int count = static_cast<int>(row.term_count);
  for (int t = 0; t < count && t < 8; ++t) {
    if (row.terms[t].coefficient == 0.0) continue;
    const uint64_t slot = row.terms[t].face_slot;
    const uint32_t k = atomicAdd(&cursor[slot], 1u);
    entries[inc_offset[slot] + k] = (r << 3) | uint64_t(t);
  }
}

// Sorts each face's incidence by (row, term) and forms the Jacobi diagonal inertia + sum c^2 in that
// order; a diagonal that leaves the finite range reports OutOfRange at the term that made it so.
__global__ void k_incidence_diagonal(const fl_strain_row* rows, const double* inertia, uint64_t nf,
                                     const uint64_t* inc_offset, const uint32_t* inc_count, uint64_t* entries,
                                     double* diag, uint64_t* err) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f >= nf) return;
  uint64_t* e = entries + inc_offset[f];
  const uint32_t m = inc_count[f];
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
    const uint64_t r = e[i] >> 3;
    const int t = static_cast<int>(e[i] & 7);
    const double c = rows[r].terms[t].coefficient;
    d += c * c;
    if (!finite(d)) {
      report_min(err, 9 * r + 1 + t, kOutOfRange);
      break;
    }
  }
  diag[f] = d;
}

// Weighted strain of every row: sum of c x[slot] in term order.
__global__ void k_row_strain(const fl_strain_row* rows, uint64_t nr, const double* x, double* strain) {
  const uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (r >= nr) return;
  const fl_strain_row& row = rows[r];
  double s = 0.0;
  const int count = static_cast<int>(row.term_count);
  for (int t = 0; t < count; ++t) s += row.terms[t].coefficient * x[row.terms[t].face_slot];
  strain[r] = s;
}

// y = M x + E^T (strain), the strains of a face's rows added in row order; flag 1 if any y is not finite.
__global__ void k_face_apply(const fl_strain_row* rows, const double* inertia, uint64_t nf, const uint64_t* inc_offset,
                             const uint32_t* inc_count, const uint64_t* entries, const double* strain, const double* x,
                             double* y, uint32_t* flag) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f >= nf) return;
  double m = inertia[f] * x[f];
  const uint64_t* e = entries + inc_offset[f];
  for (uint32_t i = 0; i < inc_count[f]; ++i) {
    const uint64_t r = e[i] >> 3;
    m += rows[r].terms[e[i] & 7].coefficient * strain[r];
  }
  y[f] = m;
  if (!finite(m)) atomicOr(flag, 1u);
}

// Workgroup partial sums written to partial[blockIdx] (256 lanes, LDS tree).
__device__ inline void block_sum(double v, double* partial) {
  __shared__ double red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
__device__ inline void block_max(double v, double* partial) {
  __shared__ double red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
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

__global__ void k_dot_f64(const double* a, const double* b, uint64_t n, double* partial) {
  double acc = 0.0;
  for (uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; f < n; f += uint64_t(gridDim.x) * blockDim.x)
    acc += a[f] * b[f];
  block_sum(acc, partial);
}

// v += step d, r -= step q; partial max |r|; flag 1 on a non-finite iterate.
__global__ void k_step(double step, const double* d, const double* q, uint64_t nf, double* v, double* r,
                       double* partial, uint32_t* flag) {
  double mx = 0.0;
  for (uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; f < nf; f += uint64_t(gridDim.x) * blockDim.x) {
    const double vf = v[f] + step * d[f];
    const double rf = r[f] - step * q[f];
    v[f] = vf;
    r[f] = rf;
    if (!finite(vf) || !finite(rf)) atomicOr(flag, 1u);
    mx = fmax(mx, fabs(rf));
  }
  block_max(mx, partial);
}

// r = rhs - q; partial max |r|; flag 1 on a non-finite residual.
__global__ void k_true_residual_f64(const double* rhs, const double* q, uint64_t nf, double* r, double* partial,
                                    uint32_t* flag) {
  double mx = 0.0;
  for (uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; f < nf; f += uint64_t(gridDim.x) * blockDim.x) {
    const double rf = rhs[f] - q[f];
    r[f] = rf;
    if (!finite(rf)) atomicOr(flag, 1u);
    mx = fmax(mx, fabs(rf));
  }
  block_max(mx, partial);
}

__global__ void k_direction_f64(const double* z, double beta, uint64_t nf, double* d) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f < nf) d[f] = z[f] + beta * d[f];
}

// Fixed-order finish of a reduction: partial[0..g) summed (or maximised) by one lane into out.
__global__ void k_finish(const double* partial, int g, int max_mode, double* out) {
  if (threadIdx.x != 0) return;
  double acc = 0.0;
  for (int i = 0; i < g; ++i) acc = max_mode ? fmax(acc, partial[i]) : acc + partial[i];
  *out = acc;
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

int solve_viscous_block(fl_ctx* ctx, const ViscosityDevice& V, uint64_t nr, uint64_t nf, double dt, double max_rhs,
                        uint64_t max_iterations, double relative_tolerance, uint64_t* iterations,
                        double* max_residual) {
  hipStream_t s = ctx->stream;
  *iterations = 0;
  *max_residual = 0.0;
  FL_TRY(reset_error(ctx));
  k_fill_u32<<<grid_for(nf), kThreads, 0, s>>>(V.inc_count, nf);
  k_fill_u32<<<grid_for(nf), kThreads, 0, s>>>(V.cursor, nf);
  if (nr) k_weight_rows<<<grid_for(nr), kThreads, 0, s>>>(V.rows, nr, V.inertia, nf, std::sqrt(dt), V.inc_count, ctx->err);
  FL_HIP(hipGetLastError());
  uint64_t word = 0;
  FL_TRY(read_error(ctx, &word));
  if (word != ~uint64_t(0)) return static_cast<int>(word & 0xff);
  if (nf) {
    k_widen<<<grid_for(nf), kThreads, 0, s>>>(V.inc_count, nf, V.counts);
    FL_TRY(exclusive_scan_u64(ctx, V.counts, V.inc_offset, nf));
    if (nr) k_fill_incidence<<<grid_for(nr), kThreads, 0, s>>>(V.rows, nr, V.inc_offset, V.cursor, V.entries);
    k_incidence_diagonal<<<grid_for(nf), kThreads, 0, s>>>(V.rows, V.inertia, nf, V.inc_offset, V.inc_count, V.entries,
                                                           V.diag, ctx->err);
    FL_HIP(hipGetLastError());
    FL_TRY(read_error(ctx, &word));
    if (word != ~uint64_t(0)) return static_cast<int>(word & 0xff);
    k_fill_d<<<grid_for(nf), kThreads, 0, s>>>(V.v, nf, 0.0);
  }
  if (max_rhs == 0.0) return FL_OK;
  const double tolerance = relative_tolerance * max_rhs;
  const unsigned rg = static_cast<unsigned>(std::min<uint64_t>(kReduceGroups, (nf + kThreads - 1) / kThreads));
  double* scal = V.scal;  // [0] r.z, [1] curvature, [2] max |r|
  auto read = [&](int i, double* out) -> int {
    FL_HIP(hipMemcpyAsync(out, scal + i, sizeof(double), hipMemcpyDeviceToHost, s));
    return FL_OK;
  };
  uint32_t flag = 0;
  auto apply = [&](const double* x, double* y) -> int {
    if (nr) k_row_strain<<<grid_for(nr), kThreads, 0, s>>>(V.rows, nr, x, V.strain);
    k_face_apply<<<grid_for(nf), kThreads, 0, s>>>(V.rows, V.inertia, nf, V.inc_offset, V.inc_count, V.entries,
                                                   V.strain, x, y, V.flag);
    FL_HIP(hipGetLastError());
    return FL_OK;
  };
  k_copy_d<<<grid_for(nf), kThreads, 0, s>>>(V.rhs, nf, V.r);
  k_precondition<<<rg, kThreads, 0, s>>>(V.r, V.inertia, V.diag, nf, V.z, V.partial);
  k_finish<<<1, 64, 0, s>>>(V.partial, static_cast<int>(rg), 0, scal + 0);
  k_copy_d<<<grid_for(nf), kThreads, 0, s>>>(V.z, nf, V.d);
  FL_HIP(hipGetLastError());
  double rz = 0;
  FL_TRY(read(0, &rz));
  FL_HIP(hipStreamSynchronize(s));
  while (*iterations < max_iterations) {
    k_fill_u32<<<1, 64, 0, s>>>(V.flag, 1);
    FL_TRY(apply(V.d, V.q));
    k_dot_f64<<<rg, kThreads, 0, s>>>(V.d, V.q, nf, V.partial);
    k_finish<<<1, 64, 0, s>>>(V.partial, static_cast<int>(rg), 0, scal + 1);
    double curvature = 0;
    FL_TRY(read(1, &curvature));
    FL_HIP(hipMemcpyAsync(&flag, V.flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    if (!std::isfinite(rz) || rz <= 0.0 || flag) return FL_OUT_OF_RANGE;
    if (!std::isfinite(curvature) || curvature <= 0.0) return FL_ABORTED;
    const double step = rz / curvature;
    if (!std::isnormal(step)) return FL_OUT_OF_RANGE;
    k_step<<<rg, kThreads, 0, s>>>(step, V.d, V.q, nf, V.v, V.r, V.partial, V.flag);
    k_finish<<<1, 64, 0, s>>>(V.partial, static_cast<int>(rg), 1, scal + 2);
    double mr = 0;
    FL_TRY(read(2, &mr));
    FL_HIP(hipMemcpyAsync(&flag, V.flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    if (flag) return FL_OUT_OF_RANGE;
    ++*iterations;
    const bool check = mr <= tolerance || *iterations == max_iterations;
    if (check) {
      FL_TRY(apply(V.v, V.q));
      k_true_residual_f64<<<rg, kThreads, 0, s>>>(V.rhs, V.q, nf, V.r, V.partial, V.flag);
      k_finish<<<1, 64, 0, s>>>(V.partial, static_cast<int>(rg), 1, scal + 2);
      FL_TRY(read(2, &mr));
      FL_HIP(hipMemcpyAsync(&flag, V.flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      FL_HIP(hipStreamSynchronize(s));
      if (flag) return FL_OUT_OF_RANGE;
      if (mr <= tolerance) {
        *max_residual = mr;
        return FL_OK;
      }
    }
    k_precondition<<<rg, kThreads, 0, s>>>(V.r, V.inertia, V.diag, nf, V.z, V.partial);
    k_finish<<<1, 64, 0, s>>>(V.partial, static_cast<int>(rg), 0, scal + 0);
    double next_rz = 0;
    FL_TRY(read(0, &next_rz));
    FL_HIP(hipStreamSynchronize(s));
    const double beta = check ? 0.0 : next_rz / rz;
    k_direction_f64<<<grid_for(nf), kThreads, 0, s>>>(V.z, beta, nf, V.d);
    FL_HIP(hipGetLastError());
    rz = next_rz;
  }
  return FL_ABORTED;
}

}  // namespace fl
