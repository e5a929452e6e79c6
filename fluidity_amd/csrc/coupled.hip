// Coupled viscous pressure projection on the GPU: ProjectMacViscousPressure3D's SolveCoupled
// (native/src/mac_viscous_pressure_3d.cc:151-459).
//
// The unknown is the cell pressure p. Each pressure iterate's face velocities solve the implicit viscous
// momentum block (M + dt E^T W E) u = M u* + dt A (lift + p_tail - p_head) (M the face inertia
// dt A / pressure_to_velocity_scale, E the strain rows, W their quadrature weights), by the Jacobi PCG
// of viscosity.hip to 1e-10 of its largest right-hand side; the cell constraint is the integrated flux
// D u + compliance (p - p_prev) - target + wall fluxes. The pressure iteration is the reference's GCR with
// restart length 32: the preconditioned residual is made orthogonal to the previous responses by
// modified Gram-Schmidt, the measured response of every direction is kept, and the physical residual
// is recomputed from a fresh velocity solve at every iterate, so acceptance never relies on an
// inexact inner solve being linear. Acceptance is the reference's: every |defect| / volume within 2e-5
// and the residual L2 within 2e-6 (1 + |initial residual|) (tightened when a weighted jump is present).
//
// The preconditioner: the reference divides by the Jacobi diagonal of the Schur complement
// (compliance + the conductances scaled by inertia / velocity diagonal) (jacobi = true); the default
// applies one V-cycle of the pressure hierarchy of the same links with those effective conductances,
// so its fine operator is the pressure operator whose diagonal is the reference's Jacobi diagonal.
// Both give the same accepted state to the gates; the V-cycle needs far fewer iterations (and
// velocity solves) on large grids.
//
// Reductions run over fixed partial slots, so results repeat run to run; the per-cell constraint sums
// are Neumaier-compensated in the cell's face order, not in the reference's global link order, and
// the product representability checks of the reference are not restated (overflow still fails as
// OutOfRange through the finiteness checks).
#include <cstring>

#include "context.h"

namespace fl {
namespace {

constexpr int kRestart = 32;          // kCoupledPressureRestartLength3D
constexpr unsigned kGroups = 1024;    // reduction workgroups
constexpr int kScalars = 16;

struct CoupledBuffers {
  double *base = nullptr, *eff = nullptr, *pdiag = nullptr, *res = nullptr, *p = nullptr, *zero = nullptr,
         *u = nullptr, *defects = nullptr, *dirs = nullptr, *resps = nullptr, *partial = nullptr, *scal = nullptr,
         *host = nullptr;
  uint64_t* err = nullptr;
  uint64_t cap_c = 0, cap_f = 0;
  void release() {
    void* ptrs[] = {base, eff, pdiag, res, p, zero, u, defects, dirs, resps, partial, scal, err};
    for (void* q : ptrs)
      if (q) (void)hipFree(q);
    if (host) (void)hipHostFree(host);
    *this = CoupledBuffers{};
  }
};

template <typename T>
int alloc(T** p, uint64_t n) {
  *p = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(p), std::max<uint64_t>(n, 1) * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return FL_RESOURCE_EXHAUSTED;
  }
  return FL_OK;
}

int ensure(fl_ctx* ctx, uint64_t nc, uint64_t nf) {
  if (!ctx->coupled) ctx->coupled = new (std::nothrow) CoupledBuffers();
  auto* B = static_cast<CoupledBuffers*>(ctx->coupled);
  if (!B) return FL_RESOURCE_EXHAUSTED;
  if (nc <= B->cap_c && nf <= B->cap_f && B->base) return FL_OK;
  B->release();
  FL_TRY(alloc(&B->base, nf));
  FL_TRY(alloc(&B->eff, nf));
  FL_TRY(alloc(&B->u, nf));
  FL_TRY(alloc(&B->pdiag, nc));
  FL_TRY(alloc(&B->res, nc));
  FL_TRY(alloc(&B->p, nc));
  FL_TRY(alloc(&B->zero, nc));
  FL_TRY(alloc(&B->defects, nc));
  FL_TRY(alloc(&B->dirs, kRestart * nc));
  FL_TRY(alloc(&B->resps, kRestart * nc));
  FL_TRY(alloc(&B->partial, 4 * uint64_t(kGroups)));
  FL_TRY(alloc(&B->scal, kScalars));
  FL_TRY(alloc(&B->err, 2));
  if (hipHostMalloc(reinterpret_cast<void**>(&B->host), 32 * sizeof(double)) != hipSuccess) {
    B->host = nullptr;
    return FL_RESOURCE_EXHAUSTED;
  }
  B->cap_c = nc;
  B->cap_f = nf;
  return FL_OK;
}

__device__ inline void report(uint64_t* err, uint64_t ordinal, int code) {
  atomicMin(reinterpret_cast<unsigned long long*>(err), (static_cast<unsigned long long>(ordinal) << 8) | code);
}

// Workgroup reduction of (sum, sum, max, max) into partial[4 * group + k].
__device__ inline void group_reduce4(double v[4], double* partial) {
  __shared__ double red[4][kThreads / 64];
  for (int o = 32; o > 0; o >>= 1) {
    v[0] += __shfl_xor(v[0], o);
    v[1] += __shfl_xor(v[1], o);
    v[2] = fmax(v[2], __shfl_xor(v[2], o));
    v[3] = fmax(v[3], __shfl_xor(v[3], o));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 4; ++k) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    double a[4] = {red[0][0], red[1][0], red[2][0], red[3][0]};
    for (int q = 1; q < kThreads / 64; ++q) {
      a[0] += red[0][q];
      a[1] += red[1][q];
      a[2] = fmax(a[2], red[2][q]);
      a[3] = fmax(a[3], red[3][q]);
    }
    for (int k = 0; k < 4; ++k) partial[4 * blockIdx.x + k] = a[k];
  }
}

// scal[slot + k] = the reduction of the groups' partials (sums in group order, maxima).
__global__ void k_finish4(const double* partial, unsigned groups, double* scal, int slot) {
  __shared__ double red[4][kThreads];
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (unsigned g = threadIdx.x; g < groups; g += blockDim.x) {
    a[0] += partial[4 * g];
    a[1] += partial[4 * g + 1];
    a[2] = fmax(a[2], partial[4 * g + 2]);
    a[3] = fmax(a[3], partial[4 * g + 3]);
  }
  for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
      red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + s]);
      red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) scal[slot + k] = red[k][0];
}

__global__ void k_reset(uint64_t* err) {
  if (threadIdx.x < 2) err[threadIdx.x] = ~uint64_t(0);
}

// RemapStrain (:50-77): each term's grid slot to its link ordinal by binary search over the ascending
// link slots; a row with no terms, more than 8, a negative or non-finite weight, a non-finite prescribed
// strain, or a term off the links or with a non-finite coefficient is InvalidArgument.
__global__ void k_remap(fl_strain_row* rows, uint64_t nr, const uint64_t* slot, uint64_t nf, uint64_t* err) {
  const uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (r >= nr) return;
  fl_strain_row& row = rows[r];
  if (row.term_count == 0 || row.term_count > 8 || !isfinite(row.quadrature_weight) || row.quadrature_weight < 0.0 ||
      !isfinite(row.prescribed_strain)) {
    report(err, r, FL_INVALID_ARGUMENT);
    return;
  }
  for (uint64_t t = 0; t < row.term_count; ++t) {
    const uint64_t key = row.terms[t].face_slot;
    uint64_t lo = 0, hi = nf;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) / 2;
      if (slot[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    if (lo == nf || slot[lo] != key || !isfinite(row.terms[t].coefficient)) {
      report(err, r, FL_INVALID_ARGUMENT);
      return;
    }
    row.terms[t].face_slot = lo;
  }
}

// Face inertia and the base momentum M u* + dt A lift (:185-205); lift is the internal jump, or the
// exterior pressure signed by the missing endpoint.
__global__ void k_setup(uint64_t nf, double dt, const double* area, const double* scale, const double* ustar,
                        const double* lift, const int32_t* kind, const int32_t* tail, double* inertia, double* base,
                        uint64_t* err) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f >= nf) return;
  double m = 0.0, b = 0.0;
  if (area[f] != 0.0) {
    const double step_area = dt * area[f];
    m = step_area / scale[f];
    if (!(isnormal(m) && m > 0.0)) report(err, f, FL_OUT_OF_RANGE);
    double l = 0.0;
    if (kind[f] == 1) l = lift[f];
    else if (kind[f] == 2) l = tail[f] >= 0 ? -lift[f] : lift[f];
    b = m * ustar[f] + step_area * l;
  }
  inertia[f] = m;
  base[f] = b;
}

// Prescribed strain of the rows' eliminated (zero-inertia) faces moved to the right-hand side (:226-246).
__global__ void k_eliminate(const fl_strain_row* rows, uint64_t nr, const double* inertia, const double* uwall,
                            double dt, double* base) {
  const uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (r >= nr) return;
  const fl_strain_row& row = rows[r];
  if (row.quadrature_weight == 0.0) return;
  double prescribed = row.prescribed_strain;
  for (uint64_t t = 0; t < row.term_count; ++t) {
    const uint64_t f = row.terms[t].face_slot;
    if (inertia[f] == 0.0) prescribed += row.terms[t].coefficient * uwall[f];
  }
  if (prescribed == 0.0) return;
  for (uint64_t t = 0; t < row.term_count; ++t) {
    const uint64_t f = row.terms[t].face_slot;
    if (inertia[f] > 0.0) atomicAdd(&base[f], -(dt * row.quadrature_weight * prescribed * row.terms[t].coefficient));
  }
}

// Effective conductance k M / (M + dt sum w c^2) of each open face (:213-224); the base momentum must
// be finite.
__global__ void k_effective(uint64_t nf, const double* conductance, const double* inertia, const double* vdiag,
                            const double* base, double* eff, uint64_t* err) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f >= nf) return;
  double e = 0.0;
  if (inertia[f] != 0.0) {
    e = conductance[f] * (inertia[f] / vdiag[f]);
    if (!(isfinite(e) && e > 0.0)) report(err, f, FL_OUT_OF_RANGE);
  }
  if (!isfinite(base[f])) report(err, f, FL_OUT_OF_RANGE);
  eff[f] = e;
}

__global__ void k_pressure_diagonal(uint64_t nc, const int32_t* cell_faces, const double* comp, const double* eff,
                                    double* pdiag) {
  const uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (c >= nc) return;
  double d = comp[c];
  for (int j = 0; j < 6; ++j) d += eff[cell_faces[6 * c + j]];
  pdiag[c] = d;
}

// Velocity block right-hand side for pressure p: base (physical only) + dt A (p_tail - p_head), missing
// endpoints at zero (:281-297); partial maxima of |rhs| in slot 2.
__global__ void k_momentum(uint64_t nf, double dt, const double* base, const double* area, const int32_t* tail,
                           const int32_t* head, const double* p, int physical, double* rhs, double* partial) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; f < nf; f += uint64_t(gridDim.x) * blockDim.x) {
    double m = physical ? base[f] : 0.0;
    if (area[f] > 0.0) {
      const double pt = tail[f] >= 0 ? p[tail[f]] : 0.0;
      const double ph = head[f] >= 0 ? p[head[f]] : 0.0;
      m += dt * area[f] * (pt - ph);
    }
    rhs[f] = m;
    v[2] = fmax(v[2], fabs(m));
  }
  group_reduce4(v, partial);
}

struct Neumaier {
  double sum = 0.0, corr = 0.0;
  __device__ void add(double t) {
    const double n = sum + t;
    corr += fabs(sum) >= fabs(t) ? (sum - n) + t : (t - n) + sum;
    sum = n;
  }
  __device__ double total() const { return sum + corr; }
};

// AccumulateConstraint (:83-134) per cell: -target (physical), compliance (p - p_prev or p), and the
// outward open flux A u of the six incident links plus their wall flux (physical). Non-physical calls
// zero the gauge cells (the responses). Partials: slot 0 sum r^2, slot 3 max |r| / V.
__global__ void k_constraint(uint64_t nc, const int32_t* cell_faces, const int32_t* tail, const double* area,
                             const double* wall, const double* target, const double* comp, const double* prev,
                             const double* volume, const uint8_t* flag, const double* p, const double* u,
                             int physical, double* out, double* partial, uint64_t* err) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < nc; i += uint64_t(gridDim.x) * blockDim.x) {
    const int32_t c = static_cast<int32_t>(i);
    Neumaier flux;
    if (physical) flux.add(-target[c]);
    if (comp[c] > 0.0) flux.add(comp[c] * (p[c] - (physical ? prev[c] : 0.0)));
    for (int j = 0; j < 6; ++j) {
      const int32_t f = cell_faces[6 * c + j];
      const double sign = tail[f] == c ? 1.0 : -1.0;
      flux.add(sign * (area[f] * u[f]));
      if (physical) flux.add(sign * wall[f]);
    }
    double r = flux.total();
    if (!isfinite(r)) report(err, i, FL_OUT_OF_RANGE);
    if (!physical && flag[c] == 2) r = 0.0;
    out[c] = r;
    v[0] += r * r;
    v[3] = fmax(v[3], fabs(r / volume[c]));
  }
  group_reduce4(v, partial);
}

// Pinned (gauge) residuals to zero; the Jacobi direction residual / diagonal (:409-419).
__global__ void k_direction(uint64_t nc, const uint8_t* flag, const double* pdiag, double* res, double* dir,
                            int jacobi, uint64_t* err) {
  const uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (c >= nc) return;
  if (flag[c] == 2) {
    res[c] = 0.0;
    if (jacobi) dir[c] = 0.0;
    return;
  }
  if (jacobi) {
    if (!(isfinite(pdiag[c]) && pdiag[c] > 0.0)) report(err, c, FL_OUT_OF_RANGE);
    dir[c] = res[c] / pdiag[c];
  }
}

// Partial sums of a . b (slot 0) and c . b (slot 1).
__global__ void k_dots(uint64_t n, const double* a, const double* b, const double* c, double* partial) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    v[0] += a[i] * b[i];
    if (c) v[1] += c[i] * b[i];
  }
  group_reduce4(v, partial);
}

// resp -= s rq, dir -= s dq with s = scal[0] (one modified Gram-Schmidt step, :433-441).
__global__ void k_orthogonalize(uint64_t n, const double* scal, const double* rq, const double* dq, double* resp,
                                double* dir) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const double s = scal[0];
  resp[i] -= s * rq[i];
  dir[i] -= s * dq[i];
}

// Normalise the response and direction by |resp| = sqrt(scal[0]) and step the pressure by
// (res . resp) / |resp| = scal[1] / |resp| (:442-456).
__global__ void k_step(uint64_t n, const double* scal, double* resp, double* dir, double* p, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  const double norm = sqrt(scal[0]);
  if (!(isfinite(norm) && norm > 0.0)) {
    if (i == 0) report(err + 1, 0, FL_ABORTED);
    return;
  }
  if (i >= n) return;
  const double step = scal[1] / norm;
  const double r = resp[i] / norm, d = dir[i] / norm;
  resp[i] = r;
  dir[i] = d;
  const double q = p[i] - step * d;
  if (!isfinite(q)) report(err, i, FL_OUT_OF_RANGE);
  p[i] = q;
}

// Accepted velocities: the solved ones on open faces, the wall normal velocity on eliminated ones; the
// defects r / V.
__global__ void k_accept(uint64_t nf, const double* inertia, const double* v, const double* uwall, double* u) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f < nf) u[f] = inertia[f] == 0.0 ? uwall[f] : v[f];
}
__global__ void k_defects(uint64_t nc, const double* res, const double* volume, double* d) {
  const uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (c < nc) d[c] = res[c] / volume[c];
}
__global__ void k_copy(uint64_t n, const double* a, double* b) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) b[i] = a[i];
}
__global__ void k_zero(uint64_t n, double* a) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// Whether a positive-conductance internal link carries a nonzero jump (tightens the L2 gate); bit 0 of *flag.
__global__ void k_jump(uint64_t nf, const double* conductance, const int32_t* kind, const double* lift,
                       unsigned int* flag) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f < nf && conductance[f] > 0.0 && kind[f] == 1 && lift[f] != 0.0) atomicOr(flag, 1u);
}

// Recursive GCR residual r - (r . q) q with q = resp / |resp|, before k_step normalises resp:
// res -= (scal[1] / scal[0]) resp (scal[0] = resp . resp, scal[1] = res . resp); partials of sum r^2
// (slot 0) and max |r| / V (slot 3). Gauge cells stay zero (their residual and response are zero).
__global__ void k_recursive(uint64_t nc, const double* scal, const double* resp, const double* volume, double* res,
                            double* partial) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  const double step = scal[1] / scal[0];
  for (uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; c < nc; c += uint64_t(gridDim.x) * blockDim.x) {
    const double r = res[c] - step * resp[c];
    res[c] = r;
    v[0] += r * r;
    v[3] = fmax(v[3], fabs(r / volume[c]));
  }
  group_reduce4(v, partial);
}

unsigned groups_for(uint64_t n) { return static_cast<unsigned>(std::min<uint64_t>(kGroups, grid_for(n))); }

int read_err(fl_ctx* ctx, CoupledBuffers& B) {
  FL_HIP(hipMemcpyAsync(B.host + 16, B.err, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  const uint64_t* w = reinterpret_cast<const uint64_t*>(B.host + 16);
  for (int i = 0; i < 2; ++i)
    if (w[i] != ~uint64_t(0)) return static_cast<int>(w[i] & 0xff);
  return FL_OK;
}

}  // namespace

int coupled_rows(fl_ctx* ctx, const CoupledSystem& S, const ViscosityDevice& V, uint64_t nr, double dt) {
  FL_TRY(ensure(ctx, S.nc, S.nf));
  auto& B = *static_cast<CoupledBuffers*>(ctx->coupled);
  hipStream_t s = ctx->stream;
  k_reset<<<1, 64, 0, s>>>(B.err);
  if (nr) k_remap<<<grid_for(nr), kThreads, 0, s>>>(V.rows, nr, S.slot, S.nf, B.err);
  FL_HIP(hipGetLastError());
  FL_TRY(read_err(ctx, B));
  k_setup<<<grid_for(S.nf), kThreads, 0, s>>>(S.nf, dt, S.area, S.scale, S.ustar, S.lift, S.kind, S.tail,
                                              const_cast<double*>(V.inertia), B.base, B.err);
  if (nr) k_eliminate<<<grid_for(nr), kThreads, 0, s>>>(V.rows, nr, V.inertia, S.uwall, dt, B.base);
  FL_HIP(hipGetLastError());
  return read_err(ctx, B);
}

int coupled_solve(fl_ctx* ctx, const CoupledSystem& S, const ViscosityDevice& V, uint64_t nr, uint64_t nt, double dt,
                  const fl_pressure_options& o, bool jacobi, CoupledResult* out) {
  auto& B = *static_cast<CoupledBuffers*>(ctx->coupled);
  hipStream_t s = ctx->stream;
  const uint64_t nc = S.nc, nf = S.nf;
  *out = CoupledResult{};
  FL_TRY(pack_viscous_block(ctx, V, nr, nf, nt, dt));  // V.diag = M + dt sum w c^2
  k_reset<<<1, 64, 0, s>>>(B.err);
  k_effective<<<grid_for(nf), kThreads, 0, s>>>(nf, S.conductance, V.inertia, V.diag, B.base, B.eff, B.err);
  k_pressure_diagonal<<<grid_for(nc), kThreads, 0, s>>>(nc, S.cell_faces, S.comp, B.eff, B.pdiag);
  if (!jacobi) FL_TRY(coupled_coefficients(ctx, B.eff));
  k_zero<<<grid_for(nc), kThreads, 0, s>>>(nc, B.zero);
  k_copy<<<grid_for(nc), kThreads, 0, s>>>(nc, S.p_init, B.p);
  FL_HIP(hipGetLastError());
  FL_TRY(read_err(ctx, B));
  const unsigned gf = groups_for(nf), gc = groups_for(nc);

  // One velocity solve for pressure p; V.v holds the velocities.
  auto solve_velocity = [&](const double* p, bool physical) -> int {
    k_momentum<<<gf, kThreads, 0, s>>>(nf, dt, B.base, S.area, S.tail, S.head, p, physical ? 1 : 0,
                                       const_cast<double*>(V.rhs), B.partial);
    k_finish4<<<1, kThreads, 0, s>>>(B.partial, gf, B.scal, 0);
    FL_HIP(hipGetLastError());
    FL_HIP(hipMemcpyAsync(B.host, B.scal, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    const double max_rhs = B.host[2];
    if (!std::isfinite(max_rhs)) return FL_OUT_OF_RANGE;
    uint64_t it = 0;
    double mr = 0.0;
    // The physical solves to the reference's 1e-10; on the V-cycle path a direction's solve to 1e-5, as
    // GCR keeps the response it measures and acceptance rests on a fresh physical solve.
    const double tolerance = physical || jacobi ? 1e-10 : 1e-5;
    FL_TRY(pcg_viscous_block(ctx, V, nr, nf, max_rhs, 256, tolerance, &it, &mr));
    ++out->velocity_solves;
    out->velocity_iterations += it;
    if (physical) out->max_momentum_residual = mr;
    return FL_OK;
  };
  // Constraint of (p, V.v) into dst; returns the residual L2 and max |r| / V (physical) in B.host[0..1].
  auto constraint = [&](const double* p, bool physical, double* dst) -> int {
    k_constraint<<<gc, kThreads, 0, s>>>(nc, S.cell_faces, S.tail, S.area, S.wall, S.target, S.comp, S.prev,
                                         S.volume, S.flag, p, V.v, physical ? 1 : 0, dst, B.partial, B.err);
    k_finish4<<<1, kThreads, 0, s>>>(B.partial, gc, B.scal, 4);
    FL_HIP(hipGetLastError());
    return FL_OK;
  };
  auto read_norms = [&](double* norm, double* max_defect) -> int {
    FL_HIP(hipMemcpyAsync(B.host, B.scal + 4, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    FL_HIP(hipMemcpyAsync(B.host + 16, B.err, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    const uint64_t* w = reinterpret_cast<const uint64_t*>(B.host + 16);
    for (int i = 0; i < 2; ++i)
      if (w[i] != ~uint64_t(0)) return static_cast<int>(w[i] & 0xff);
    *norm = std::sqrt(B.host[0]);
    *max_defect = B.host[3];
    if (!std::isfinite(*norm)) return FL_OUT_OF_RANGE;
    if (!std::isfinite(*max_defect)) return FL_OUT_OF_RANGE;
    return FL_OK;
  };

  double rhs_norm = 0.0, dummy = 0.0;
  const bool history = !S.p_init_zero;
  if (history) {
    FL_TRY(solve_velocity(B.zero, true));
    FL_TRY(constraint(B.zero, true, B.res));
    FL_TRY(read_norms(&rhs_norm, &dummy));
    if (rhs_norm == 0.0) k_zero<<<grid_for(nc), kThreads, 0, s>>>(nc, B.p);
  }
  bool jump = false;  // a positive-conductance internal link with a nonzero jump tightens the L2 gate
  {
    unsigned int* flag = reinterpret_cast<unsigned int*>(B.scal + 15);
    FL_HIP(hipMemsetAsync(flag, 0, sizeof(unsigned int), s));
    k_jump<<<grid_for(nf), kThreads, 0, s>>>(nf, S.conductance, S.kind, S.lift, flag);
    FL_HIP(hipGetLastError());
    unsigned int word = 0;
    FL_HIP(hipMemcpyAsync(B.host + 20, flag, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    std::memcpy(&word, B.host + 20, sizeof(word));
    jump = (word & 1u) != 0;
  }
  const uint64_t capacity = std::min<uint64_t>(kRestart, o.max_iterations);
  uint64_t basis = 0;
  double limit = 0.0;
  // The reference recomputes the physical residual from a fresh velocity solve at every iterate
  // (jacobi: that iteration exactly). The V-cycle path steps on the recursive GCR residual
  // r - (r . q) q of the normalised response q, and recomputes the physical residual only where the
  // recursive one meets the gates (or at the budget's end); a failed check restarts the basis from the
  // fresh residual. Both accept only on a fresh residual.
  bool fresh = true;
  double norm = 0.0, max_defect = 0.0;
  for (uint64_t it = 0; it <= o.max_iterations; ++it) {
    if (fresh) {
      FL_TRY(solve_velocity(B.p, true));
      FL_TRY(constraint(B.p, true, B.res));
      FL_TRY(read_norms(&norm, &max_defect));
    }
    if (it == 0) {
      if (!history) rhs_norm = norm;
      limit = 2e-6 + 2e-6 * rhs_norm;
      if (!std::isfinite(rhs_norm) || !std::isfinite(limit)) return FL_OUT_OF_RANGE;
      if (rhs_norm > 0.0 && jump) limit = std::min(limit, std::max(std::numeric_limits<double>::denorm_min(), 1e-3 * rhs_norm));
    }
    out->iterations = it;
    out->residual_l2 = norm;
    out->initial_residual_l2 = rhs_norm;
    out->max_defect = max_defect;
    const bool gates = max_defect <= 2e-5 && norm <= limit;
    if (!fresh && (gates || it == o.max_iterations)) {
      fresh = true;  // verify on the physical residual, without counting an iteration
      basis = capacity;  // a failed check restarts the basis from the fresh residual
      --it;
      continue;
    }
    if (gates) {
      k_accept<<<grid_for(nf), kThreads, 0, s>>>(nf, V.inertia, V.v, S.uwall, B.u);
      k_defects<<<grid_for(nc), kThreads, 0, s>>>(nc, B.res, S.volume, B.defects);
      FL_HIP(hipGetLastError());
      FL_HIP(hipStreamSynchronize(s));
      return FL_OK;
    }
    if (it == o.max_iterations) break;
    if (basis == capacity) basis = 0;
    double* dir = B.dirs + basis * nc;
    double* resp = B.resps + basis * nc;
    k_direction<<<grid_for(nc), kThreads, 0, s>>>(nc, S.flag, B.pdiag, B.res, dir, jacobi ? 1 : 0, B.err);
    FL_HIP(hipGetLastError());
    if (!jacobi) FL_TRY(coupled_vcycle(ctx, B.res, dir, o));
    FL_TRY(solve_velocity(dir, false));
    FL_TRY(constraint(dir, false, resp));
    for (uint64_t q = 0; q < basis; ++q) {
      k_dots<<<gc, kThreads, 0, s>>>(nc, resp, B.resps + q * nc, nullptr, B.partial);
      k_finish4<<<1, kThreads, 0, s>>>(B.partial, gc, B.scal, 8);
      k_orthogonalize<<<grid_for(nc), kThreads, 0, s>>>(nc, B.scal + 8, B.resps + q * nc, B.dirs + q * nc, resp, dir);
    }
    k_dots<<<gc, kThreads, 0, s>>>(nc, resp, resp, B.res, B.partial);
    k_finish4<<<1, kThreads, 0, s>>>(B.partial, gc, B.scal, 8);
    if (!jacobi) {  // the recursive residual before k_step normalises the response in place
      k_recursive<<<gc, kThreads, 0, s>>>(nc, B.scal + 8, resp, S.volume, B.res, B.partial);
      k_finish4<<<1, kThreads, 0, s>>>(B.partial, gc, B.scal, 4);
    }
    k_step<<<grid_for(nc), kThreads, 0, s>>>(nc, B.scal + 8, resp, dir, B.p, B.err);
    FL_HIP(hipGetLastError());
    ++basis;
    fresh = jacobi;
    if (!jacobi) FL_TRY(read_norms(&norm, &max_defect));
  }
  FL_TRY(coupled_error(ctx));
  return FL_ABORTED;
}

int coupled_outputs(fl_ctx* ctx, const double** p, const double** u, const double** defects) {
  auto* B = static_cast<CoupledBuffers*>(ctx->coupled);
  if (!B) return FL_FAILED_PRECONDITION;
  *p = B->p;
  *u = B->u;
  *defects = B->defects;
  return FL_OK;
}

void release_coupled(fl_ctx* ctx) {
  auto* B = static_cast<CoupledBuffers*>(ctx->coupled);
  if (!B) return;
  B->release();
  delete B;
  ctx->coupled = nullptr;
}

}  // namespace fl
