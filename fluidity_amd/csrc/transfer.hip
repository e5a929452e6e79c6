// APIC transfers on gfx950: P2G quanta, normalize + forces, G2P.
//
// P2G (DepositMacQuanta3D, native/src/transfers_3d.cc:147-197; mac_deposit_3d_internal.h:41-65):
// one workgroup per non-empty sort bin B. Every stencil node of a particle in B lies in the 8^3-cell
// window [4B, 4B+8)^3 = blocks B + {0,1}^3 (sparse_mac_grid_3d.cc:279-283), so the workgroup reduces
// its particles' contributions into a 24 KiB LDS window (2 x 1536 int64) with LDS atomics and then
// adds the non-zero window entries to the owned face slots with one global int64 atomic each.
// Every contribution is rounded half-to-even to integer quanta before any addition, exactly as
// the reference does, so every summation order gives the same bits (F8) and the result is
// identical to the CPU restatement. Workgroup ids are remapped so that consecutive bins run on the
// same XCD and their overlapping windows meet in the same L2.
//
// G2P (GatherMac3D, native/src/transfers_3d.cc:199-261; mac_gather_3d_internal.h:15-69): one
// workgroup per bin stages the window's old and new face velocities in LDS (24 KiB) with coalesced
// block loads; each particle contracts its 3x3x3 tensor axis by axis (x, then y, then z) in FP64,
// which computes the same sums as the reference's 27-sample loop with far fewer operations.
#include "context.h"

namespace fl {
namespace {

constexpr int kWindow = 8 * 192;

// Bijective workgroup -> bin map giving each XCD a contiguous run of bins (workgroups are dispatched
// round-robin over the 8 XCDs).
__device__ inline uint32_t xcd_chunked(uint32_t wg, uint32_t n) {
  const uint32_t q = n >> 3, r = n & 7, xcd = wg & 7, slot = wg >> 3;
  return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot;
}

__device__ inline void bin_coords(const BinList& bins, uint32_t key, int32_t b[3]) {
  b[0] = bins.box.lo[0] + int32_t(key % uint32_t(bins.box.dim[0]));
  b[1] = bins.box.lo[1] + int32_t((key / uint32_t(bins.box.dim[0])) % uint32_t(bins.box.dim[1]));
  b[2] = bins.box.lo[2] + int32_t(key / (uint32_t(bins.box.dim[0]) * uint32_t(bins.box.dim[1])));
}

__device__ inline void report(uint64_t* err, uint64_t ordinal, int code) {
  atomicMin(reinterpret_cast<unsigned long long*>(err), static_cast<unsigned long long>(error_word(ordinal, code)));
}

// Window block of a coordinate's sample range, as a 2-bit mask of {low, high}.
__device__ inline uint32_t axis_blocks(int32_t first, int32_t tile) {
  const int32_t r0 = first - tile, r2 = r0 + 2;
  return (1u << (r0 >> 2)) | (1u << (r2 >> 2));
}

__device__ inline bool window_ok(const int32_t* wblock, uint32_t mx, uint32_t my, uint32_t mz) {
  for (int z = 0; z < 2; ++z)
    for (int y = 0; y < 2; ++y)
      for (int x = 0; x < 2; ++x)
        if (((mx >> x) & 1) && ((my >> y) & 1) && ((mz >> z) & 1) && wblock[(z * 2 + y) * 2 + x] < 0) return false;
  return true;
}

__global__ __launch_bounds__(kThreads) void k_p2g(ParticleView p, BinList bins, GridView grid, double inv_qm,
                                                  double inv_qp, double limit, const uint32_t* perm,
                                                  unsigned long long* qm, unsigned long long* qp, uint64_t* err) {
  __shared__ unsigned long long wm[kWindow], wq[kWindow];
  __shared__ int32_t wblock[8];
  const uint32_t b = xcd_chunked(blockIdx.x, gridDim.x);
  int32_t bin[3];
  bin_coords(bins, bins.key[b], bin);
  for (int i = threadIdx.x; i < kWindow; i += kThreads) {
    wm[i] = 0;
    wq[i] = 0;
  }
  if (threadIdx.x < 8)
    wblock[threadIdx.x] = grid.ordinal(bin[0] + (threadIdx.x & 1), bin[1] + ((threadIdx.x >> 1) & 1),
                                       bin[2] + (threadIdx.x >> 2));
  __syncthreads();
  const Geometry g = grid.g;
  const uint32_t off = bins.offset[b], cnt = bins.count[b];
  const int32_t tile[3] = {4 * bin[0], 4 * bin[1], 4 * bin[2]};
  for (uint32_t k = off + threadIdx.x; k < off + cnt; k += kThreads) {
    const uint64_t ordinal = perm[k];
    const double s[3] = {p.s[0][k], p.s[1][k], p.s[2][k]};
    const double t[3] = {p.t[0][k], p.t[1][k], p.t[2][k]};
    const double m = p.m[k];
    bool motion = true;
    for (int a = 0; a < 3; ++a) motion = motion && within_half_cell(s[a], t[a], g.h);
    if (!motion) {
      report(err, ordinal, kFailedPrecondition);
      continue;
    }
    Coord tc[3];
    for (int a = 0; a < 3; ++a) tc[a] = normalize(t[a], comp(g, a), g.h);
    int code = kOk;
    for (int axis = 0; axis < 3 && code == kOk; ++axis) {
      AxisStencil st[3];
      for (int c = 0; c < 3 && code == kOk; ++c) code = axis_stencil(tc[c], c == axis ? 0.0 : 0.5, bin[c], g.h, st[c]);
      if (code) break;
      for (int c = 0; c < 3 && code == kOk; ++c)
        for (int i = 0; i < 3; ++i)
          if (!finite(st[c].offset[i])) code = kOutOfRange;
      if (code) break;
      if (!window_ok(wblock, axis_blocks(st[0].first, tile[0]), axis_blocks(st[1].first, tile[1]),
                     axis_blocks(st[2].first, tile[2]))) {
        code = kFailedPrecondition;
        break;
      }
      const double v = p.v[axis][k];
      const double cx = p.c[3 * axis][k], cy = p.c[3 * axis + 1][k], cz = p.c[3 * axis + 2][k];
      for (int zi = 0; zi < 3 && code == kOk; ++zi) {
        const int32_t rz = st[2].first + zi - tile[2];
        const double wz = st[2].w[zi].r, ofz = st[2].offset[zi];
        for (int yi = 0; yi < 3 && code == kOk; ++yi) {
          const int32_t ry = st[1].first + yi - tile[1];
          const double wy = st[1].w[yi].r, ofy = st[1].offset[yi];
          for (int xi = 0; xi < 3; ++xi) {
            const int32_t rx = st[0].first + xi - tile[0];
            const double ofx = st[0].offset[xi];
            const double w = st[0].w[xi].r * wy * wz;
            const double mass = w * m;
            const double inc = (cx * ofx + cy * ofy) + cz * ofz;
            const double vel = v + inc;
            const double mom = mass * vel;
            const double sm = mass * inv_qm;  // equals mass / quantum: the quantum is a power of two
            const double sp = mom * inv_qp;
            if (!(finite(sm) && fabs(sm) <= limit && finite(sp) && fabs(sp) <= limit)) {
              code = kOutOfRange;
              break;
            }
            const int wb = ((rz >> 2) * 2 + (ry >> 2)) * 2 + (rx >> 2);
            const int slot = wb * 192 + axis * 64 + ((rz & 3) * 16 + (ry & 3) * 4 + (rx & 3));
            const long long im = static_cast<long long>(rint(sm));
            const long long ip = static_cast<long long>(rint(sp));
            if (im) atomicAdd(&wm[slot], static_cast<unsigned long long>(im));
            if (ip) atomicAdd(&wq[slot], static_cast<unsigned long long>(ip));
          }
        }
      }
    }
    if (code) report(err, ordinal, code);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kWindow; i += kThreads) {
    const int32_t ord = wblock[i / 192];
    if (ord < 0) continue;
    const uint64_t slot = uint64_t(ord) * 192 + (i % 192);
    if (wm[i]) atomicAdd(&qm[slot], wm[i]);
    if (wq[i]) atomicAdd(&qp[slot], wq[i]);
  }
}

constexpr int64_t kMaxQuanta = int64_t(1) << 62;

// NormalizeCore (liquid_boundary_transfer_3d.cc:463-516) then ApplyForces (liquid_substep_3d_cpu.cc:156-178).
// err[0] holds the first normalize failure, err[1] the first forces failure.
__global__ void k_normalize_forces(const int64_t* qm, const int64_t* qp, const int64_t* qs, uint64_t nf, double sqm,
                                   double sqp, double min_mass, double gx, double gy, double gz, double dt,
                                   double* mass, double* u0, uint8_t* validity, double* ustar, uint64_t* err) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f >= nf) return;
  const int64_t Qm = qm[f], Qp = qp[f], Qs = qs ? qs[f] : 0;
  double dm = 0.0, u = 0.0;
  uint8_t v = 0;
  if (Qm < 0 || Qm > kMaxQuanta || Qp < -kMaxQuanta || Qp > kMaxQuanta || Qs != 0) {
    report(err, f, kInvalidArgument);
  } else {
    dm = static_cast<double>(Qm) * sqm;
    const double dq = static_cast<double>(Qp) * sqp;
    if (!finite(dm) || !finite(dq)) {
      report(err, f, kOutOfRange);
    } else if (dm != 0.0) {
      if (dm <= min_mass) {
        v = 3;
      } else {
        u = dq / dm;
        if (!finite(u)) report(err, f, kOutOfRange);
        v = 1;
      }
    }
  }
  const int axis = static_cast<int>((f % 192) / 64);
  const double ga = axis == 0 ? gx : (axis == 1 ? gy : gz);
  const double pred = u + dt * ga;
  if (!finite(pred)) report(err + 1, f, kOutOfRange);
  if (mass) mass[f] = dm;
  if (u0) u0[f] = u;
  if (validity) validity[f] = v;
  if (ustar) ustar[f] = pred;
}

// Contract a 3x3x3 tensor T[z][y][x] with one vector per axis.
__device__ inline double contract(const double T[27], const double x[3], const double y[3], const double z[3]) {
  double sz = 0.0;
  for (int k = 0; k < 3; ++k) {
    double sy = 0.0;
    for (int j = 0; j < 3; ++j) {
      const double* r = T + 9 * k + 3 * j;
      sy += y[j] * ((x[0] * r[0] + x[1] * r[1]) + x[2] * r[2]);
    }
    sz += z[k] * sy;
  }
  return sz;
}

__global__ __launch_bounds__(kThreads) void k_g2p(ParticleView p, BinList bins, GridView grid, const double* uold,
                                                  const double* unew, const uint32_t* perm, double* out, uint64_t cap,
                                                  uint64_t* err) {
  __shared__ double wo[kWindow], wn[kWindow];
  __shared__ int32_t wblock[8];
  const uint32_t b = xcd_chunked(blockIdx.x, gridDim.x);
  int32_t bin[3];
  bin_coords(bins, bins.key[b], bin);
  if (threadIdx.x < 8)
    wblock[threadIdx.x] = grid.ordinal(bin[0] + (threadIdx.x & 1), bin[1] + ((threadIdx.x >> 1) & 1),
                                       bin[2] + (threadIdx.x >> 2));
  __syncthreads();
  for (int i = threadIdx.x; i < kWindow; i += kThreads) {
    const int32_t ord = wblock[i / 192];
    const uint64_t slot = uint64_t(ord) * 192 + (i % 192);
    wo[i] = ord >= 0 ? uold[slot] : 0.0;
    wn[i] = ord >= 0 ? unew[slot] : 0.0;
  }
  __syncthreads();
  const Geometry g = grid.g;
  const double inv_h = 1.0 / g.h;
  const uint32_t off = bins.offset[b], cnt = bins.count[b];
  const int32_t tile[3] = {4 * bin[0], 4 * bin[1], 4 * bin[2]};
  for (uint32_t k = off + threadIdx.x; k < off + cnt; k += kThreads) {
    const uint64_t ordinal = perm[k];
    const double s[3] = {p.s[0][k], p.s[1][k], p.s[2][k]};
    const double t[3] = {p.t[0][k], p.t[1][k], p.t[2][k]};
    bool motion = true;
    for (int a = 0; a < 3; ++a) motion = motion && within_half_cell(s[a], t[a], g.h);
    if (!motion) {
      report(err, ordinal, kFailedPrecondition);
      continue;
    }
    Coord tc[3];
    for (int a = 0; a < 3; ++a) tc[a] = normalize(t[a], comp(g, a), g.h);
    double res[kGatherDoubles];
    int code = kOk;
    for (int axis = 0; axis < 3 && code == kOk; ++axis) {
      AxisStencil st[3];
      for (int c = 0; c < 3 && code == kOk; ++c) code = axis_stencil(tc[c], c == axis ? 0.0 : 0.5, bin[c], g.h, st[c]);
      if (code) break;
      for (int c = 0; c < 3 && code == kOk; ++c)
        for (int i = 0; i < 3; ++i)
          if (!finite(st[c].offset[i])) code = kOutOfRange;
      if (code) break;
      if (!window_ok(wblock, axis_blocks(st[0].first, tile[0]), axis_blocks(st[1].first, tile[1]),
                     axis_blocks(st[2].first, tile[2]))) {
        code = kFailedPrecondition;
        break;
      }
      double U[27], O[27];
      for (int zi = 0; zi < 3; ++zi)
        for (int yi = 0; yi < 3; ++yi)
          for (int xi = 0; xi < 3; ++xi) {
            const int32_t rx = st[0].first + xi - tile[0], ry = st[1].first + yi - tile[1],
                          rz = st[2].first + zi - tile[2];
            const int wb = ((rz >> 2) * 2 + (ry >> 2)) * 2 + (rx >> 2);
            const int slot = wb * 192 + axis * 64 + ((rz & 3) * 16 + (ry & 3) * 4 + (rx & 3));
            U[9 * zi + 3 * yi + xi] = wn[slot];
            O[9 * zi + 3 * yi + xi] = wo[slot];
          }
      double w[3][3], wr[3][3], dw[3][3];
      for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 3; ++i) {
          w[c][i] = st[c].w[i].r;
          wr[c][i] = w[c][i] * (st[c].offset[i] * inv_h);
          dw[c][i] = st[c].dw[i];
        }
      const double pic = contract(U, w[0], w[1], w[2]);
      const double pic_old = contract(O, w[0], w[1], w[2]);
      const double mx = contract(U, wr[0], w[1], w[2]);
      const double my = contract(U, w[0], wr[1], w[2]);
      const double mz = contract(U, w[0], w[1], wr[2]);
      const double gx = contract(U, dw[0], w[1], w[2]) * inv_h;
      const double gy = contract(U, w[0], dw[1], w[2]) * inv_h;
      const double gz = contract(U, w[0], w[1], dw[2]) * inv_h;
      const double inc = pic - pic_old;
      const double sx = (4.0 * mx) * inv_h, sy = (4.0 * my) * inv_h, sz = (4.0 * mz) * inv_h;
      if (!(finite(pic) && finite(inc) && finite(sx) && finite(sy) && finite(sz) && finite(gx) && finite(gy) &&
            finite(gz))) {
        code = kOutOfRange;
        break;
      }
      res[axis] = pic;
      res[3 + axis] = inc;
      res[6 + 3 * axis + 0] = sx;
      res[6 + 3 * axis + 1] = sy;
      res[6 + 3 * axis + 2] = sz;
      res[15 + 3 * axis + 0] = gx;
      res[15 + 3 * axis + 1] = gy;
      res[15 + 3 * axis + 2] = gz;
    }
    if (code) {
      report(err, ordinal, code);
      continue;
    }
    for (int c = 0; c < kGatherDoubles; ++c) out[c * cap + k] = res[c];
  }
}

__global__ void k_unpack(const fl_particle* in, uint64_t n, double* base, uint64_t cap) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const double* r = reinterpret_cast<const double*>(in + i);
  for (int c = 0; c < kParticleDoubles; ++c) base[c * cap + i] = r[c];
}

}  // namespace

int unpack_particles(fl_ctx* ctx, uint64_t n) {
  if (!n) return FL_OK;
  k_unpack<<<grid_for(n), kThreads, 0, ctx->stream>>>(ctx->staging_aos, n, ctx->raw.base, ctx->raw.capacity);
  FL_HIP(hipGetLastError());
  ctx->raw.trial_aliases_sorted = false;
  return FL_OK;
}

int p2g_quanta(fl_ctx* ctx, uint64_t n, double inv_qm, double inv_qp, double limit) {
  const uint64_t nf = ctx->nb * 192;
  FL_TRY(fill_i64(ctx, ctx->qm, nf, 0));
  FL_TRY(fill_i64(ctx, ctx->qp, nf, 0));
  if (ctx->nbins == 0 || n == 0) return FL_OK;
  k_p2g<<<ctx->nbins, kThreads, 0, ctx->stream>>>(ctx->sorted.view(), bin_list(ctx), grid_view(ctx), inv_qm, inv_qp,
                                                  limit, ctx->perm, reinterpret_cast<unsigned long long*>(ctx->qm),
                                                  reinterpret_cast<unsigned long long*>(ctx->qp), ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int normalize_forces(fl_ctx* ctx, uint64_t nf, double qm, double qp, double min_mass, const double g[3], double dt) {
  if (!nf) return FL_OK;
  k_normalize_forces<<<grid_for(nf), kThreads, 0, ctx->stream>>>(ctx->qm, ctx->qp, nullptr, nf, qm, qp, min_mass,
                                                                  g[0], g[1], g[2], dt, ctx->face_mass, ctx->u0,
                                                                  ctx->validity, ctx->ustar, ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int normalize_forces_stress(fl_ctx* ctx, const int64_t* qs, uint64_t nf, double qm, double qp, double min_mass,
                            const double g[3], double dt) {
  if (!nf) return FL_OK;
  k_normalize_forces<<<grid_for(nf), kThreads, 0, ctx->stream>>>(ctx->qm, ctx->qp, qs, nf, qm, qp, min_mass, g[0],
                                                                  g[1], g[2], dt, ctx->face_mass, ctx->u0,
                                                                  ctx->validity, ctx->ustar, ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int g2p(fl_ctx* ctx, uint64_t n, const double* d_old, const double* d_new) {
  if (ctx->nbins == 0 || n == 0) return FL_OK;
  k_g2p<<<ctx->nbins, kThreads, 0, ctx->stream>>>(ctx->sorted.view(), bin_list(ctx), grid_view(ctx), d_old, d_new,
                                                  ctx->perm, ctx->gather, ctx->sorted.capacity, ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

}  // namespace fl
