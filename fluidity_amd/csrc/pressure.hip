// MAC pressure projection on gfx950: mixed-precision MGPCG.
//
// Follows ProjectMacPressure3D (native/src/mac_pressure_workspace_3d.cc:462-659): the system is frozen
// and assembled in FP64 with compensated sums (mac_pressure_3d_arithmetic.h:110-313), the outer PCG keeps
// FP32 vectors with FP64 dot products, alpha and beta rounded to FP32, and restarts from the FP64 true
// residual computed by Evaluate (mac_pressure_workspace_3d.cc:399-461), which also checks the residual /
// defect identity r = -V defect. Acceptance uses the reference gates (mac_pressure_3d_multigrid.cc:341-367)
// plus an optional relative-divergence target; the true residual is recomputed at least every 10
// iterations.
//
// Layout: pressure cells live in their 4^3 blocks; one wavefront owns one block, one lane one cell.
// Every multigrid level is a set of 4^3 blocks with a 6-neighbour block table, so all levels share the
// same kernels. Level l+1 aggregates 2x2x2 cells of level l (parent = floor(c/2)); restriction sums the
// 8 children (pull, no atomics, deterministic), prolongation is piecewise constant, and the coarse
// operator is the scaled Galerkin crossing sum of the fine conductances and anchors. Smoothing is
// red-black Gauss-Seidel, symmetric across the cycle (red then black before, black then red after).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "context.h"

namespace fl {
namespace {

constexpr double kEpsD = 2.220446049250313080847e-16;
constexpr int kCellsPerBlock = 64;
constexpr int kPartials = 2048;  // reduction partial slots

struct Key3 {
  int32_t x, y, z;
  bool operator==(const Key3& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct Key3Hash {
  size_t operator()(const Key3& k) const {
    return (size_t(uint32_t(k.x)) * 73856093u) ^ (size_t(uint32_t(k.y)) * 19349663u) ^ (size_t(uint32_t(k.z)) * 83492791u);
  }
};
bool key_less(const Key3& a, const Key3& b) {
  if (a.z != b.z) return a.z < b.z;
  if (a.y != b.y) return a.y < b.y;
  return a.x < b.x;
}
int32_t floor_half(int32_t v) { return v >= 0 ? v / 2 : -((-(v + 1)) / 2) - 1; }

struct Level {
  uint64_t nb = 0;
  std::vector<Key3> keys;
  int32_t* nbr = nullptr;    // [nb][6]: -x,+x,-y,+y,-z,+z
  int32_t* child = nullptr;  // [nb][8] blocks of the finer level (levels > 0)
  float* x = nullptr;
  float* b = nullptr;
  float* r = nullptr;
  float* k = nullptr;       // [3][nb*64] low-face conductances along x, y, z
  float* anchor = nullptr;  // [nb*64]
  float* diag = nullptr;    // [nb*64], 0 marks a non-unknown cell
};

struct Pressure {
  // capacities
  uint64_t cap_cells = 0, cap_faces = 0, cap_fine = 0;
  // structure of the current system
  uint64_t nc = 0, nf = 0;
  std::vector<Level> levels;
  int32_t* fine_cell = nullptr;   // [nb0*64] input cell ordinal or -1
  int32_t* cell_fine = nullptr;   // [nc] fine index
  uint8_t* cell_flag = nullptr;   // [nc] 1 unknown, 2 gauge
  int32_t* cell_faces = nullptr;  // [nc][6]
  int32_t* face_tail = nullptr;   // [nf] cell ordinal or -1
  int32_t* face_head = nullptr;
  int32_t* face_kind = nullptr;
  uint64_t* face_slot = nullptr;
  double* face_area = nullptr;
  double* face_beta = nullptr;
  double* face_dist = nullptr;
  double* face_ustar = nullptr;
  double* face_uwall = nullptr;
  double* face_flux = nullptr;
  int32_t* face_has_flux = nullptr;
  double* face_lift = nullptr;
  double* face_scale = nullptr;
  double* face_k = nullptr;
  double* face_wall = nullptr;
  double* face_u = nullptr;
  double* cell_volume = nullptr;
  double* cell_s = nullptr;
  double* cell_K = nullptr;
  double* cell_prev = nullptr;
  double* cell_inwall = nullptr;
  double* cell_comp = nullptr;
  double* cell_target = nullptr;
  double* rhs = nullptr;
  double* rhs_abs = nullptr;
  double* p = nullptr;
  double* p_base = nullptr;
  double* res = nullptr;
  double* def = nullptr;
  double* zero = nullptr;
  float* fine_vol = nullptr;  // [nb0*64] fluid volume (0 if none)
  // PCG (level-0 sized)
  float* r32 = nullptr;
  float* x32 = nullptr;
  float* d32 = nullptr;
  float* q32 = nullptr;
  double* partial = nullptr;  // [8][kPartials]
  double* scal = nullptr;     // device scalars
  float* fscal = nullptr;
  double* hscal = nullptr;    // pinned host mirror
  uint64_t* perr = nullptr;   // error words
  // host copies
  std::vector<int32_t> h_cell_comp_root;
  std::vector<uint64_t> h_gauges;
  bool has_jump = false;
  double full_area = 0, full_volume = 0;
  double scale_coarse = 0.5;
};

template <typename T>
int dal(T** p, uint64_t n) {
  *p = nullptr;
  if (!n) return FL_OK;
  FL_HIP(hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)));
  return FL_OK;
}
template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void free_levels(Pressure& P) {
  for (auto& L : P.levels) {
    dfree(L.nbr);
    dfree(L.child);
    dfree(L.x);
    dfree(L.b);
    dfree(L.r);
    dfree(L.k);
    dfree(L.anchor);
    dfree(L.diag);
  }
  P.levels.clear();
}

// ------------------------------------------------------------------ device helpers

struct Neumaier {
  double sum = 0.0, corr = 0.0;
  __device__ void add(double t) {
    const double n = sum + t;
    if (fabs(sum) >= fabs(t))
      corr += (sum - n) + t;
    else
      corr += (t - n) + sum;
    sum = n;
  }
  __device__ double total() const { return sum + corr; }
};
struct Row {
  Neumaier flux, abs;
  __device__ void add(double t) {
    flux.add(t);
    abs.add(fabs(t));
  }
};

__device__ inline void perr_report(uint64_t* err, uint64_t ordinal, int code) {
  atomicMin(reinterpret_cast<unsigned long long*>(err), static_cast<unsigned long long>(error_word(ordinal, code)));
}

// Neighbour cell index within a level (block-local cells, (z,y,x) local order).
__device__ inline int64_t neighbour(const int32_t* nbr, uint64_t blk, int lx, int ly, int lz, int dir) {
  int nx = lx, ny = ly, nz = lz;
  int64_t b = static_cast<int64_t>(blk);
  switch (dir) {
    case 0: nx -= 1; break;
    case 1: nx += 1; break;
    case 2: ny -= 1; break;
    case 3: ny += 1; break;
    case 4: nz -= 1; break;
    default: nz += 1; break;
  }
  if (nx < 0 || nx > 3 || ny < 0 || ny > 3 || nz < 0 || nz > 3) {
    b = nbr[6 * blk + dir];
    if (b < 0) return -1;
    nx &= 3;
    ny &= 3;
    nz &= 3;
  }
  return b * 64 + (nz * 16 + ny * 4 + nx);
}

// Conductance of face `dir` of cell i: low faces are stored on the cell, high faces on the neighbour.
__device__ inline float face_k(const float* k, uint64_t n, int64_t i, int64_t j, int dir) {
  const int axis = dir >> 1;
  return (dir & 1) ? k[axis * n + j] : k[axis * n + i];
}

// ------------------------------------------------------------------ system kernels (FP64)

__global__ void k_freeze_faces(uint64_t nf, double dt, double full_area, const double* area, const double* beta,
                               const double* dist, const double* uwall, const int32_t* has_flux, const double* flux,
                               double* scale, double* k, double* wall) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f >= nf) return;
  double s = 0.0, c = 0.0;
  if (area[f] > 0.0) {
    const double mobility = dt * beta[f];
    s = mobility / dist[f];
    c = area[f] * s;
  }
  scale[f] = s;
  k[f] = c;
  const double blocked = full_area - area[f];
  wall[f] = has_flux[f] ? flux[f] : blocked * uwall[f];
}

__global__ void k_freeze_cells(uint64_t nc, double dt, const double* V, const double* s, const double* K,
                               double* comp, double* target) {
  const uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (c >= nc) return;
  comp[c] = isinf(K[c]) ? 0.0 : V[c] / (K[c] * dt);
  target[c] = V[c] * s[c];
}

// Integrated RHS (mac_pressure_3d_arithmetic.h:263-313) with Neumaier sums per cell.
__global__ void k_rhs(uint64_t nc, const int32_t* cell_faces, const int32_t* tail, const int32_t* kind,
                      const double* area, const double* ustar, const double* wall, const double* k, const double* lift,
                      const double* target, const double* inwall, const double* comp, const double* prev, double* rhs,
                      double* rhs_abs) {
  const uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (c >= nc) return;
  Row row;
  row.add(target[c]);
  row.add(inwall[c]);
  row.add(comp[c] * prev[c]);
  for (int j = 0; j < 6; ++j) {
    const int32_t f = cell_faces[6 * c + j];
    const double sign = tail[f] == static_cast<int32_t>(c) ? 1.0 : -1.0;
    row.add(-sign * (area[f] * ustar[f]));
    row.add(-sign * wall[f]);
    if (kind[f] == 1) row.add(-sign * (k[f] * lift[f]));
    else if (kind[f] == 2) row.add(k[f] * lift[f]);
  }
  rhs[c] = row.flux.total();
  rhs_abs[c] = row.abs.total();
}

// Corrected face velocity u = u* + s (p_tail - p_head + jump), exterior pressure at a missing endpoint.
__device__ inline double corrected(int32_t f, const int32_t* tail, const int32_t* head, const int32_t* kind,
                                   const double* area, const double* ustar, const double* scale, const double* lift,
                                   const double* p) {
  double u = ustar[f];
  if (area[f] > 0.0) {
    const double ext = kind[f] == 2 ? lift[f] : 0.0;
    const double jump = kind[f] == 1 ? lift[f] : 0.0;
    const double pt = tail[f] >= 0 ? p[tail[f]] : ext;
    const double ph = head[f] >= 0 ? p[head[f]] : ext;
    u += scale[f] * (pt - ph + jump);
  }
  return u;
}

__global__ void k_face_velocity(uint64_t nf, const int32_t* tail, const int32_t* head, const int32_t* kind,
                                const double* area, const double* ustar, const double* scale, const double* lift,
                                const double* p, double* u) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f >= nf) return;
  u[f] = corrected(static_cast<int32_t>(f), tail, head, kind, area, ustar, scale, lift, p);
}

// Partial reductions: 0 sum r^2, 1 max |r|/V, 2 max |defect|, 3 sum defect^2 ; flag 4 identity failures.
__device__ inline void block_reduce(double v[4], double* partial, unsigned slot) {
  __shared__ double red[4][kThreads / 64];
  for (int o = 32; o > 0; o >>= 1) {
    v[0] += __shfl_xor(v[0], o);
    v[1] = fmax(v[1], __shfl_xor(v[1], o));
    v[2] = fmax(v[2], __shfl_xor(v[2], o));
    v[3] += __shfl_xor(v[3], o);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int q = 0; q < 4; ++q) red[q][wave] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = red[0][0], b = red[1][0], c = red[2][0], d = red[3][0];
    for (int w = 1; w < kThreads / 64; ++w) {
      a += red[0][w];
      b = fmax(b, red[1][w]);
      c = fmax(c, red[2][w]);
      d += red[3][w];
    }
    partial[0 * kPartials + slot] = a;
    partial[1 * kPartials + slot] = b;
    partial[2 * kPartials + slot] = c;
    partial[3 * kPartials + slot] = d;
  }
}

// Evaluate (mac_pressure_workspace_3d.cc:399-461): FP64 residual of the original operator, physical
// defect from corrected fluxes, and the identity |V defect + r| <= 64 eps (sum of absolute terms).
__global__ void k_evaluate(uint64_t nc, const int32_t* cell_faces, const int32_t* tail, const int32_t* head,
                           const int32_t* kind, const double* area, const double* ustar, const double* wall,
                           const double* scale, const double* k, const double* lift, const double* V,
                           const double* target, const double* inwall, const double* comp, const double* prev,
                           const double* rhs, const double* rhs_abs, const double* p, double* res, double* def,
                           double* partial, uint64_t* err) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; c < nc; c += uint64_t(gridDim.x) * blockDim.x) {
    Row op, phys;
    op.add(comp[c] * p[c]);
    phys.add(-target[c]);
    phys.add(-inwall[c]);
    if (comp[c] > 0.0) phys.add(comp[c] * (p[c] - prev[c]));
    for (int j = 0; j < 6; ++j) {
      const int32_t f = cell_faces[6 * c + j];
      const double sign = tail[f] == static_cast<int32_t>(c) ? 1.0 : -1.0;
      if (k[f] != 0.0) {
        const double pt = tail[f] >= 0 ? p[tail[f]] : 0.0;
        const double ph = head[f] >= 0 ? p[head[f]] : 0.0;
        op.add(sign * (k[f] * (pt - ph)));
      }
      const double u = corrected(f, tail, head, kind, area, ustar, scale, lift, p);
      phys.add(sign * (area[f] * u));
      phys.add(sign * wall[f]);
    }
    const double r = rhs[c] - op.flux.total();
    const double d = phys.flux.total() / V[c];
    const double ie = V[c] * d + r;
    const double terms = rhs_abs[c] + op.abs.total() + phys.abs.total();
    if (!finite(r) || !finite(d) || !finite(ie) || !finite(terms)) perr_report(err, c, kOutOfRange);
    else if (fabs(ie) > 64.0 * kEpsD * terms) perr_report(err, c, kInternal);
    res[c] = r;
    def[c] = d;
    v[0] += r * r;
    v[1] = fmax(v[1], fabs(r) / V[c]);
    v[2] = fmax(v[2], fabs(d));
    v[3] += d * d;
  }
  block_reduce(v, partial, blockIdx.x);
}

__global__ void k_sum_sq(uint64_t n, const double* a, double* partial) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; c < n; c += uint64_t(gridDim.x) * blockDim.x)
    v[0] += a[c] * a[c];
  block_reduce(v, partial, blockIdx.x);
}

// Reduces the 4 partial rows in a fixed order into scal[base..base+3].
__global__ void k_finalize4(const double* partial, unsigned slots, double* scal, int base) {
  __shared__ double red[4][kThreads];
  double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
  for (unsigned i = threadIdx.x; i < slots; i += blockDim.x) {
    a += partial[i];
    b = fmax(b, partial[kPartials + i]);
    c = fmax(c, partial[2 * kPartials + i]);
    d += partial[3 * kPartials + i];
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  red[2][threadIdx.x] = c;
  red[3][threadIdx.x] = d;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + s]);
      red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + s]);
      red[3][threadIdx.x] += red[3][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; ++q) scal[base + q] = red[q][0];
}

// ------------------------------------------------------------------ level coefficients

// Level 0 from the frozen FP64 links: edges between unknowns keep their conductance, links to a
// Dirichlet exterior or to a gauge cell become anchors, compliance is an anchor.
__global__ void k_level0(uint64_t n, const int32_t* fine_cell, const uint8_t* flag, const int32_t* cell_faces,
                         const int32_t* tail, const int32_t* head, const double* k, const double* comp, const double* V,
                         float* kk, float* anchor, float* diag, float* vol) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const int32_t c = fine_cell[i];
  float kl[3] = {0.f, 0.f, 0.f};
  double a = 0.0, ksum[6] = {0, 0, 0, 0, 0, 0};
  bool unknown = c >= 0 && flag[c] == 1;
  if (unknown) {
    a = comp[c];
    for (int j = 0; j < 6; ++j) {
      const int32_t f = cell_faces[6 * c + j];
      if (!(k[f] > 0.0)) continue;
      const int32_t other = tail[f] == c ? head[f] : tail[f];
      if (other >= 0 && flag[other] == 1) {
        ksum[j] = k[f];
        if ((j & 1) == 0) kl[j >> 1] = static_cast<float>(k[f]);
      } else {
        a += k[f];
      }
    }
  }
  for (int ax = 0; ax < 3; ++ax) kk[ax * n + i] = kl[ax];
  const float af = static_cast<float>(a);
  anchor[i] = unknown ? af : 0.f;
  double d = static_cast<double>(af);
  for (int j = 0; j < 6; ++j) d += static_cast<double>(static_cast<float>(ksum[j]));
  diag[i] = unknown ? static_cast<float>(d) : 0.f;
  vol[i] = c >= 0 ? static_cast<float>(V[c]) : 0.f;
}

// Coarse conductances and anchors: scaled sums over the 2x2x2 children (Galerkin crossing sums).
__global__ void k_coarsen(uint64_t nc_cells, const int32_t* child, uint64_t nf_cells, const float* fk,
                          const float* fanchor, const float* fdiag, float scale, float* ck, float* canchor,
                          float* cdiag_mark) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nc_cells) return;
  const uint64_t blk = i >> 6;
  const int l = static_cast<int>(i & 63);
  const int cx = l & 3, cy = (l >> 2) & 3, cz = l >> 4;
  double a = 0.0, kx = 0.0, ky = 0.0, kz = 0.0;
  bool any = false;
  for (int oz = 0; oz < 2; ++oz)
    for (int oy = 0; oy < 2; ++oy)
      for (int ox = 0; ox < 2; ++ox) {
        const int fx = 2 * cx + ox, fy = 2 * cy + oy, fz = 2 * cz + oz;
        const int32_t fb = child[8 * blk + ((fz >> 2) * 2 + (fy >> 2)) * 2 + (fx >> 2)];
        if (fb < 0) continue;
        const uint64_t fi = uint64_t(fb) * 64 + ((fz & 3) * 16 + (fy & 3) * 4 + (fx & 3));
        if (fdiag[fi] == 0.f) continue;
        any = true;
        a += static_cast<double>(scale) * fanchor[fi];
        if (ox == 0) kx += static_cast<double>(scale) * fk[0 * nf_cells + fi];
        if (oy == 0) ky += static_cast<double>(scale) * fk[1 * nf_cells + fi];
        if (oz == 0) kz += static_cast<double>(scale) * fk[2 * nf_cells + fi];
      }
  ck[0 * nc_cells + i] = static_cast<float>(kx);
  ck[1 * nc_cells + i] = static_cast<float>(ky);
  ck[2 * nc_cells + i] = static_cast<float>(kz);
  canchor[i] = static_cast<float>(a);
  cdiag_mark[i] = any ? 1.f : 0.f;
}

__global__ void k_diag(uint64_t n, const int32_t* nbr, const float* k, const float* anchor, float* diag) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  if (diag[i] == 0.f) return;
  const uint64_t blk = i >> 6;
  const int l = static_cast<int>(i & 63);
  const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
  double d = static_cast<double>(anchor[i]);
  for (int dir = 0; dir < 6; ++dir) {
    const int64_t j = neighbour(nbr, blk, lx, ly, lz, dir);
    if (j < 0) continue;
    d += static_cast<double>(face_k(k, n, static_cast<int64_t>(i), j, dir));
  }
  diag[i] = static_cast<float>(d);
}

// ------------------------------------------------------------------ V-cycle kernels (FP32)

// One red-black Gauss-Seidel half sweep. `first` means x == 0 on entry: the colour is set to b/d
// and the other colour to zero, which saves a fill pass.
__global__ void k_smooth(uint64_t n, const int32_t* nbr, const float* k, const float* diag, const float* b, float* x,
                         int color, int first) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const int l = static_cast<int>(i & 63);
  const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
  const float d = diag[i];
  if (((lx + ly + lz) & 1) != color) {
    if (first) x[i] = 0.f;
    return;
  }
  if (d == 0.f) {
    if (first) x[i] = 0.f;
    return;
  }
  float s = b[i];
  if (!first) {
    const uint64_t blk = i >> 6;
    for (int dir = 0; dir < 6; ++dir) {
      const int64_t j = neighbour(nbr, blk, lx, ly, lz, dir);
      if (j < 0) continue;
      s += face_k(k, n, static_cast<int64_t>(i), j, dir) * x[j];
    }
  }
  x[i] = s / d;
}

// r = b - A x in the edge-difference form (mac_pressure_workspace_3d.cc:228-246).
__global__ void k_residual(uint64_t n, const int32_t* nbr, const float* k, const float* anchor, const float* diag,
                           const float* b, const float* x, float* r) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  if (diag[i] == 0.f) {
    r[i] = 0.f;
    return;
  }
  const uint64_t blk = i >> 6;
  const int l = static_cast<int>(i & 63);
  const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
  const float xi = x[i];
  float s = anchor[i] * xi;
  for (int dir = 0; dir < 6; ++dir) {
    const int64_t j = neighbour(nbr, blk, lx, ly, lz, dir);
    if (j < 0) continue;
    s += face_k(k, n, static_cast<int64_t>(i), j, dir) * (xi - x[j]);
  }
  r[i] = b[i] - s;
}

__global__ void k_restrict(uint64_t nc_cells, const int32_t* child, const float* fr, float* cb) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nc_cells) return;
  const uint64_t blk = i >> 6;
  const int l = static_cast<int>(i & 63);
  const int cx = l & 3, cy = (l >> 2) & 3, cz = l >> 4;
  float s = 0.f;
  for (int oz = 0; oz < 2; ++oz)
    for (int oy = 0; oy < 2; ++oy)
      for (int ox = 0; ox < 2; ++ox) {
        const int fx = 2 * cx + ox, fy = 2 * cy + oy, fz = 2 * cz + oz;
        const int32_t fb = child[8 * blk + ((fz >> 2) * 2 + (fy >> 2)) * 2 + (fx >> 2)];
        if (fb < 0) continue;
        s += fr[uint64_t(fb) * 64 + ((fz & 3) * 16 + (fy & 3) * 4 + (fx & 3))];
      }
  cb[i] = s;
}

// x_fine += x_coarse(parent); parent block = child-table inverse passed as (parent block, octant).
__global__ void k_prolong(uint64_t nf_cells, const int32_t* parent, const float* diag, const float* cx, float* fx) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nf_cells || diag[i] == 0.f) return;
  const uint64_t blk = i >> 6;
  const int32_t pb = parent[2 * blk], oct = parent[2 * blk + 1];
  const int l = static_cast<int>(i & 63);
  const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
  const int px = ((oct & 1) * 4 + lx) >> 1, py = (((oct >> 1) & 1) * 4 + ly) >> 1, pz = ((oct >> 2) * 4 + lz) >> 1;
  fx[i] += cx[uint64_t(pb) * 64 + (pz * 16 + py * 4 + px)];
}

// Coarsest level: `sweeps` forward (red, black) then `sweeps` backward (black, red) sweeps from zero,
// all inside one workgroup.
__global__ void k_coarsest(uint64_t n, const int32_t* nbr, const float* k, const float* diag, const float* b, float* x,
                           int sweeps) {
  for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) x[i] = 0.f;
  __syncthreads();
  for (int pass = 0; pass < 4 * sweeps; ++pass) {
    const int fwd = pass < 2 * sweeps;
    const int color = fwd ? (pass & 1) : 1 - (pass & 1);
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
      const int l = static_cast<int>(i & 63);
      const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
      if (((lx + ly + lz) & 1) != color || diag[i] == 0.f) continue;
      const uint64_t blk = i >> 6;
      float s = b[i];
      for (int dir = 0; dir < 6; ++dir) {
        const int64_t j = neighbour(nbr, blk, lx, ly, lz, dir);
        if (j < 0) continue;
        s += face_k(k, n, static_cast<int64_t>(i), j, dir) * x[j];
      }
      x[i] = s / diag[i];
    }
    __threadfence_block();
    __syncthreads();
  }
}

// ------------------------------------------------------------------ PCG kernels

__device__ inline void block_sum3(double a, double b, double c, double* partial, unsigned slot) {
  double v[4] = {a, 0.0, 0.0, c};
  // reuse block_reduce: slot 0 sums a, slot 1 max b, slot 3 sums c
  v[1] = b;
  block_reduce(v, partial, slot);
}

__global__ void k_dot(uint64_t n, const float* a, const float* b, double* partial) {
  double s = 0.0;
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    s += static_cast<double>(a[i]) * static_cast<double>(b[i]);
  double v[4] = {s, 0.0, 0.0, 0.0};
  block_reduce(v, partial, blockIdx.x);
}

// scal: 0 rho, 1 rho_prev, 2 curvature, 3 alpha, 4 beta, 8.. norms ; fscal: 0 alpha, 1 beta ; err flags.
__device__ inline bool fine_value(double v, float& out) {
  if (!finite(v) || fabs(v) > 3.4028234663852886e38) return false;
  const float r = static_cast<float>(v);
  if (!isfinite(r) || (v != 0.0 && r == 0.f)) return false;
  out = r;
  return true;
}

__global__ void k_after_rho(double* scal, float* fscal, uint64_t* err, int restart) {
  if (threadIdx.x) return;
  const double rho = scal[20];
  scal[0] = rho;
  if (!(rho > 0.0)) {
    perr_report(err + 1, 0, kAborted);
    return;
  }
  if (restart) {
    fscal[1] = 0.f;
  } else {
    float beta;
    if (!fine_value(rho / scal[1], beta)) {
      perr_report(err + 1, 1, kOutOfRange);
      return;
    }
    fscal[1] = beta;
  }
}

__global__ void k_after_curv(double* scal, float* fscal, uint64_t* err) {
  if (threadIdx.x) return;
  const double curv = scal[20];
  scal[2] = curv;
  if (!(curv > 0.0)) {
    perr_report(err + 1, 2, kAborted);
    return;
  }
  float alpha;
  if (!fine_value(scal[0] / curv, alpha)) {
    perr_report(err + 1, 3, kOutOfRange);
    return;
  }
  fscal[0] = alpha;
  scal[1] = scal[0];
}

__global__ void k_update_d(uint64_t n, const float* z, float* d, const float* fscal, int restart) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  if (restart) {
    d[i] = z[i];
  } else {
    const float prev = fscal[1] * d[i];
    d[i] = z[i] + prev;
  }
}

__global__ void k_update_xr(uint64_t n, const float* d, const float* q, const float* diag, const float* vol, float* x,
                            float* r, const float* fscal, double* partial, uint64_t* err) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  const float alpha = fscal[0];
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    if (diag[i] == 0.f) continue;
    const float pi = alpha * d[i];
    const float fi = alpha * q[i];
    const float xn = x[i] + pi;
    const float rn = r[i] - fi;
    if (!isfinite(xn) || !isfinite(rn)) perr_report(err + 1, 4, kOutOfRange);
    x[i] = xn;
    r[i] = rn;
    const double rv = rn, V = vol[i];
    v[0] += rv * rv;
    v[1] = fmax(v[1], fabs(rv) / V);
    v[3] += (rv / V) * (rv / V);
  }
  block_reduce(v, partial, blockIdx.x);
}

// p = base + correction on the unknown cells.
__global__ void k_apply_correction(uint64_t nc, const int32_t* cell_fine, const uint8_t* flag, const double* base,
                                   const float* x, double* p, uint64_t* err) {
  const uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (c >= nc) return;
  double v = base[c];
  if (flag[c] == 1) v += static_cast<double>(x[cell_fine[c]]);
  if (!finite(v)) perr_report(err + 1, 5, kOutOfRange);
  p[c] = v;
}

// Loads the FP64 true residual into the FP32 recursive residual (FineValue rule,
// mac_pressure_workspace_3d.cc:214-223, :262-275) and clears the correction.
__global__ void k_load_residual(uint64_t n, const int32_t* fine_cell, const uint8_t* flag, const double* res, float* r,
                                float* x, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const int32_t c = fine_cell[i];
  float v = 0.f;
  if (c >= 0 && flag[c] == 1) {
    if (!fine_value(res[c], v)) perr_report(err + 1, 6, kOutOfRange);
  }
  r[i] = v;
  x[i] = 0.f;
}

__global__ void k_gather_slots(uint64_t nf, const uint64_t* slot, const double* src, double* dst) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f < nf) dst[f] = src[slot[f]];
}
__global__ void k_scatter_slots(uint64_t nf, const uint64_t* slot, const double* src, double* dst) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f < nf) dst[slot[f]] = src[f];
}
__global__ void k_copy_f64(uint64_t n, const double* src, double* dst) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ------------------------------------------------------------------ host side

Pressure* state(fl_ctx* ctx) { return static_cast<Pressure*>(ctx->pressure); }

int ensure_capacity(fl_ctx* ctx, uint64_t nc, uint64_t nf, uint64_t nfine) {
  if (!ctx->pressure) ctx->pressure = new Pressure();
  Pressure& P = *state(ctx);
  if (nc <= P.cap_cells && nf <= P.cap_faces && nfine <= P.cap_fine) return FL_OK;
  // Grow once to the context capacities (all pressure allocation happens on first use).
  const uint64_t cc = std::max<uint64_t>({nc, ctx->cap.pressure_cells, 1});
  const uint64_t cf = std::max<uint64_t>({nf, ctx->cap.pressure_faces, 1});
  const uint64_t cn = std::max<uint64_t>(nfine, 64);
  if (nc > ctx->cap.pressure_cells || nf > ctx->cap.pressure_faces) return FL_RESOURCE_EXHAUSTED;
  void* ptrs[] = {P.fine_cell, P.cell_fine, P.cell_flag, P.cell_faces, P.face_tail, P.face_head, P.face_kind,
                  P.face_slot, P.face_area, P.face_beta, P.face_dist, P.face_ustar, P.face_uwall, P.face_flux,
                  P.face_has_flux, P.face_lift, P.face_scale, P.face_k, P.face_wall, P.face_u, P.cell_volume,
                  P.cell_s, P.cell_K, P.cell_prev, P.cell_inwall, P.cell_comp, P.cell_target, P.rhs, P.rhs_abs,
                  P.p, P.p_base, P.res, P.def, P.zero, P.fine_vol, P.r32, P.x32, P.d32, P.q32, P.partial, P.scal,
                  P.fscal, P.perr};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (P.hscal) (void)hipHostFree(P.hscal);
  FL_TRY(dal(&P.cell_fine, cc));
  FL_TRY(dal(&P.cell_flag, cc));
  FL_TRY(dal(&P.cell_faces, 6 * cc));
  FL_TRY(dal(&P.face_tail, cf));
  FL_TRY(dal(&P.face_head, cf));
  FL_TRY(dal(&P.face_kind, cf));
  FL_TRY(dal(&P.face_slot, cf));
  FL_TRY(dal(&P.face_area, cf));
  FL_TRY(dal(&P.face_beta, cf));
  FL_TRY(dal(&P.face_dist, cf));
  FL_TRY(dal(&P.face_ustar, cf));
  FL_TRY(dal(&P.face_uwall, cf));
  FL_TRY(dal(&P.face_flux, cf));
  FL_TRY(dal(&P.face_has_flux, cf));
  FL_TRY(dal(&P.face_lift, cf));
  FL_TRY(dal(&P.face_scale, cf));
  FL_TRY(dal(&P.face_k, cf));
  FL_TRY(dal(&P.face_wall, cf));
  FL_TRY(dal(&P.face_u, cf));
  for (double** a : {&P.cell_volume, &P.cell_s, &P.cell_K, &P.cell_prev, &P.cell_inwall, &P.cell_comp, &P.cell_target,
                     &P.rhs, &P.rhs_abs, &P.p, &P.p_base, &P.res, &P.def, &P.zero})
    FL_TRY(dal(a, cc));
  FL_TRY(dal(&P.fine_cell, cn));
  FL_TRY(dal(&P.fine_vol, cn));
  FL_TRY(dal(&P.r32, cn));
  FL_TRY(dal(&P.x32, cn));
  FL_TRY(dal(&P.d32, cn));
  FL_TRY(dal(&P.q32, cn));
  FL_TRY(dal(&P.partial, 4 * kPartials));
  FL_TRY(dal(&P.scal, 64));
  FL_TRY(dal(&P.fscal, 16));
  FL_TRY(dal(&P.perr, 4));
  FL_HIP(hipHostMalloc(reinterpret_cast<void**>(&P.hscal), 64 * sizeof(double), hipHostMallocDefault));
  FL_HIP(hipMemsetAsync(P.zero, 0, cc * sizeof(double), ctx->stream));
  P.cap_cells = cc;
  P.cap_faces = cf;
  P.cap_fine = cn;
  return FL_OK;
}

unsigned reduce_grid(uint64_t n) {
  const uint64_t g = (n + kThreads - 1) / kThreads;
  return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(g, kPartials)));
}

// Host-side structure for one pressure system (records in the reference's order).
struct HostSystem {
  std::vector<Key3> cell_keys;  // input cells
  std::vector<int32_t> cell_faces, face_tail, face_head;
  std::vector<uint8_t> flag;
  std::vector<uint64_t> gauges;
  std::vector<int32_t> component;
  std::vector<int32_t> cell_fine;
  std::vector<Key3> blocks0;
};

// Builds the block hierarchy on the host and uploads neighbour, child and parent tables.
int build_levels(fl_ctx* ctx, Pressure& P, const std::vector<Key3>& blocks0, int levels_wanted) {
  free_levels(P);
  std::vector<std::vector<Key3>> keys{blocks0};
  std::vector<std::vector<int32_t>> parents;  // per level (except coarsest): [nb][2] parent block, octant
  while (static_cast<int>(keys.size()) < levels_wanted) {
    const auto& fine = keys.back();
    std::vector<Key3> coarse;
    coarse.reserve(fine.size() / 4 + 1);
    for (const Key3& k : fine) coarse.push_back({floor_half(k.x), floor_half(k.y), floor_half(k.z)});
    std::sort(coarse.begin(), coarse.end(), key_less);
    coarse.erase(std::unique(coarse.begin(), coarse.end()), coarse.end());
    std::unordered_map<Key3, int32_t, Key3Hash> index;
    for (size_t i = 0; i < coarse.size(); ++i) index[coarse[i]] = static_cast<int32_t>(i);
    std::vector<int32_t> par(2 * fine.size());
    for (size_t i = 0; i < fine.size(); ++i) {
      const Key3 pk{floor_half(fine[i].x), floor_half(fine[i].y), floor_half(fine[i].z)};
      par[2 * i] = index[pk];
      par[2 * i + 1] = ((fine[i].z - 2 * pk.z) * 2 + (fine[i].y - 2 * pk.y)) * 2 + (fine[i].x - 2 * pk.x);
    }
    parents.push_back(std::move(par));
    const bool shrinks = coarse.size() < fine.size();
    keys.push_back(std::move(coarse));
    if (!shrinks && keys.back().size() <= 1) break;
  }
  P.levels.resize(keys.size());
  for (size_t l = 0; l < keys.size(); ++l) {
    Level& L = P.levels[l];
    L.keys = keys[l];
    L.nb = keys[l].size();
    const uint64_t n = L.nb * 64;
    std::unordered_map<Key3, int32_t, Key3Hash> index;
    for (size_t i = 0; i < L.nb; ++i) index[L.keys[i]] = static_cast<int32_t>(i);
    std::vector<int32_t> nbr(6 * L.nb, -1);
    const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (size_t i = 0; i < L.nb; ++i)
      for (int dir = 0; dir < 6; ++dir) {
        auto it = index.find({L.keys[i].x + d[dir][0], L.keys[i].y + d[dir][1], L.keys[i].z + d[dir][2]});
        if (it != index.end()) nbr[6 * i + dir] = it->second;
      }
    FL_TRY(dal(&L.nbr, 6 * L.nb));
    FL_HIP(hipMemcpy(L.nbr, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (l > 0) {
      std::vector<int32_t> child(8 * L.nb, -1);
      const auto& par = parents[l - 1];
      for (size_t i = 0; i < keys[l - 1].size(); ++i) child[8 * par[2 * i] + par[2 * i + 1]] = static_cast<int32_t>(i);
      FL_TRY(dal(&L.child, 8 * L.nb));
      FL_HIP(hipMemcpy(L.child, child.data(), child.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    FL_TRY(dal(&L.x, n));
    FL_TRY(dal(&L.b, n));
    FL_TRY(dal(&L.r, n));
    FL_TRY(dal(&L.k, 3 * n));
    FL_TRY(dal(&L.anchor, n));
    FL_TRY(dal(&L.diag, n));
  }
  // Parent tables live in the fine level's `child` slot of a separate array: store in r? Use a dedicated buffer.
  for (size_t l = 0; l + 1 < keys.size(); ++l) {
    int32_t* par = nullptr;
    FL_TRY(dal(&par, parents[l].size()));
    FL_HIP(hipMemcpy(par, parents[l].data(), parents[l].size() * sizeof(int32_t), hipMemcpyHostToDevice));
    // The finest level has no child table; keep the parent table there for levels > 0 as well by
    // chaining through a side vector.
    P.levels[l].child = l == 0 ? par : P.levels[l].child;
    if (l > 0) {
      // Level l already owns a child table; append its parent table after it in a new allocation.
      int32_t* both = nullptr;
      const uint64_t nb = P.levels[l].nb;
      FL_TRY(dal(&both, 8 * nb + 2 * nb));
      FL_HIP(hipMemcpy(both, P.levels[l].child, 8 * nb * sizeof(int32_t), hipMemcpyDeviceToDevice));
      FL_HIP(hipMemcpy(both + 8 * nb, par, 2 * nb * sizeof(int32_t), hipMemcpyDeviceToDevice));
      dfree(P.levels[l].child);
      dfree(par);
      P.levels[l].child = both;
    }
  }
  return FL_OK;
}

// Parent table of level l (valid for l + 1 < levels).
const int32_t* parent_table(const Pressure& P, size_t l) {
  return l == 0 ? P.levels[0].child : P.levels[l].child + 8 * P.levels[l].nb;
}

int setup_coefficients(fl_ctx* ctx, Pressure& P) {
  hipStream_t s = ctx->stream;
  Level& L0 = P.levels[0];
  const uint64_t n0 = L0.nb * 64;
  k_level0<<<grid_for(n0), kThreads, 0, s>>>(n0, P.fine_cell, P.cell_flag, P.cell_faces, P.face_tail, P.face_head,
                                             P.face_k, P.cell_comp, P.cell_volume, L0.k, L0.anchor, L0.diag,
                                             P.fine_vol);
  FL_HIP(hipGetLastError());
  for (size_t l = 1; l < P.levels.size(); ++l) {
    Level& F = P.levels[l - 1];
    Level& C = P.levels[l];
    const uint64_t nfc = F.nb * 64, ncc = C.nb * 64;
    k_coarsen<<<grid_for(ncc), kThreads, 0, s>>>(ncc, C.child, nfc, F.k, F.anchor, F.diag,
                                                 static_cast<float>(P.scale_coarse), C.k, C.anchor, C.diag);
    k_diag<<<grid_for(ncc), kThreads, 0, s>>>(ncc, C.nbr, C.k, C.anchor, C.diag);
    FL_HIP(hipGetLastError());
  }
  return FL_OK;
}

int vcycle(fl_ctx* ctx, Pressure& P, size_t l, int nu, int coarse_sweeps) {
  hipStream_t s = ctx->stream;
  Level& L = P.levels[l];
  const uint64_t n = L.nb * 64;
  const unsigned g = grid_for(n);
  if (l + 1 == P.levels.size()) {
    if (L.nb <= 64) {
      k_coarsest<<<1, 1024, 0, s>>>(n, L.nbr, L.k, L.diag, L.b, L.x, coarse_sweeps);
    } else {
      for (int pass = 0; pass < 4 * coarse_sweeps; ++pass) {
        const int fwd = pass < 2 * coarse_sweeps;
        const int color = fwd ? (pass & 1) : 1 - (pass & 1);
        k_smooth<<<g, kThreads, 0, s>>>(n, L.nbr, L.k, L.diag, L.b, L.x, color, pass == 0);
      }
    }
    FL_HIP(hipGetLastError());
    return FL_OK;
  }
  for (int it = 0; it < nu; ++it) {
    k_smooth<<<g, kThreads, 0, s>>>(n, L.nbr, L.k, L.diag, L.b, L.x, 0, it == 0);
    k_smooth<<<g, kThreads, 0, s>>>(n, L.nbr, L.k, L.diag, L.b, L.x, 1, 0);
  }
  k_residual<<<g, kThreads, 0, s>>>(n, L.nbr, L.k, L.anchor, L.diag, L.b, L.x, L.r);
  Level& C = P.levels[l + 1];
  const uint64_t nc = C.nb * 64;
  k_restrict<<<grid_for(nc), kThreads, 0, s>>>(nc, C.child, L.r, C.b);
  FL_HIP(hipGetLastError());
  FL_TRY(vcycle(ctx, P, l + 1, nu, coarse_sweeps));
  k_prolong<<<g, kThreads, 0, s>>>(n, parent_table(P, l), L.diag, C.x, L.x);
  for (int it = 0; it < nu; ++it) {
    k_smooth<<<g, kThreads, 0, s>>>(n, L.nbr, L.k, L.diag, L.b, L.x, 1, 0);
    k_smooth<<<g, kThreads, 0, s>>>(n, L.nbr, L.k, L.diag, L.b, L.x, 0, 0);
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

struct Gates {
  double integrated, volume, relative, initial_defect;
};

struct EvalResult {
  double l2, max_res, max_def, def_l2;
  int code;
};

int evaluate(fl_ctx* ctx, Pressure& P, const double* p, EvalResult& out) {
  hipStream_t s = ctx->stream;
  const unsigned g = reduce_grid(P.nc);
  k_evaluate<<<g, kThreads, 0, s>>>(P.nc, P.cell_faces, P.face_tail, P.face_head, P.face_kind, P.face_area,
                                    P.face_ustar, P.face_wall, P.face_scale, P.face_k, P.face_lift, P.cell_volume,
                                    P.cell_target, P.cell_inwall, P.cell_comp, P.cell_prev, P.rhs, P.rhs_abs, p, P.res,
                                    P.def, P.partial, P.perr);
  k_finalize4<<<1, kThreads, 0, s>>>(P.partial, g, P.scal, 8);
  FL_HIP(hipGetLastError());
  FL_HIP(hipMemcpyAsync(P.hscal, P.scal, 16 * sizeof(double), hipMemcpyDeviceToHost, s));
  FL_HIP(hipMemcpyAsync(ctx->host_words, P.perr, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  out.l2 = std::sqrt(P.hscal[8]);
  out.max_res = P.hscal[9];
  out.max_def = P.hscal[10];
  out.def_l2 = std::sqrt(P.hscal[11]);
  out.code = ctx->host_words[0] != ~uint64_t(0) ? static_cast<int>(ctx->host_words[0] & 0xff) : FL_OK;
  if (out.code == FL_OK && ctx->host_words[1] != ~uint64_t(0)) out.code = static_cast<int>(ctx->host_words[1] & 0xff);
  return FL_OK;
}

bool accept(const EvalResult& e, const Gates& g) {
  if (!(std::isfinite(e.l2) && std::isfinite(e.max_res) && std::isfinite(e.max_def))) return false;
  if (!(e.l2 <= g.integrated && e.max_res <= g.volume && e.max_def <= g.volume)) return false;
  if (g.relative > 0.0 && !(e.def_l2 <= g.relative * g.initial_defect)) return false;
  return true;
}

__global__ void k_reset_err(uint64_t* e) {
  if (threadIdx.x < 4) e[threadIdx.x] = ~uint64_t(0);
}

// The mixed-precision PCG with FP64 restarts. Requires the system to be frozen (faces, cells, RHS,
// initial pressure in P.p_base) and the hierarchy coefficients set.
int solve(fl_ctx* ctx, Pressure& P, const fl_pressure_options& opt, const Gates& gates,
          fl_pressure_diagnostics& diag) {
  hipStream_t s = ctx->stream;
  Level& L0 = P.levels[0];
  const uint64_t n = L0.nb * 64;
  const unsigned g = grid_for(n);
  const unsigned rg = reduce_grid(n);
  diag.multigrid_level_count = P.levels.size();
  diag.iteration_count = 0;
  diag.restarts = 0;
  EvalResult e{};
  k_copy_f64<<<grid_for(P.nc), kThreads, 0, s>>>(P.nc, P.p_base, P.p);
  FL_TRY(evaluate(ctx, P, P.p, e));
  auto publish = [&](const EvalResult& r) {
    diag.true_residual_l2 = r.l2;
    diag.max_volume_scaled_residual = r.max_res;
    diag.max_cell_constraint_defect = r.max_def;
    diag.defect_l2 = r.def_l2;
    diag.defect_linf = r.max_def;
  };
  publish(e);
  if (e.code) return e.code;
  if (accept(e, gates)) return FL_OK;
  uint64_t unknowns = 0;
  for (uint8_t f : std::vector<uint8_t>()) unknowns += f;  // counted on the host at setup
  k_load_residual<<<g, kThreads, 0, s>>>(n, P.fine_cell, P.cell_flag, P.res, P.r32, P.x32, P.perr);
  bool restart = true;
  int since = 0;
  for (uint64_t it = 1; it <= opt.max_iterations; ++it) {
    hipMemcpyAsync(L0.b, P.r32, n * sizeof(float), hipMemcpyDeviceToDevice, s);
    FL_TRY(vcycle(ctx, P, 0, opt.smoothing_sweeps, opt.coarse_sweeps));
    k_dot<<<rg, kThreads, 0, s>>>(n, P.r32, L0.x, P.partial);
    k_finalize4<<<1, kThreads, 0, s>>>(P.partial, rg, P.scal, 20);
    k_after_rho<<<1, 64, 0, s>>>(P.scal, P.fscal, P.perr, restart ? 1 : 0);
    k_update_d<<<g, kThreads, 0, s>>>(n, L0.x, P.d32, P.fscal, restart ? 1 : 0);
    k_residual<<<g, kThreads, 0, s>>>(n, L0.nbr, L0.k, L0.anchor, L0.diag, P.zero == nullptr ? nullptr : nullptr,
                                      P.d32, P.q32);
    k_dot<<<rg, kThreads, 0, s>>>(n, P.d32, P.q32, P.partial);
    k_finalize4<<<1, kThreads, 0, s>>>(P.partial, rg, P.scal, 20);
    k_after_curv<<<1, 64, 0, s>>>(P.scal, P.fscal, P.perr);
    k_update_xr<<<rg, kThreads, 0, s>>>(n, P.d32, P.q32, L0.diag, P.fine_vol, P.x32, P.r32, P.fscal, P.partial,
                                        P.perr);
    k_finalize4<<<1, kThreads, 0, s>>>(P.partial, rg, P.scal, 8);
    FL_HIP(hipGetLastError());
    FL_HIP(hipMemcpyAsync(P.hscal, P.scal, 16 * sizeof(double), hipMemcpyDeviceToHost, s));
    FL_HIP(hipMemcpyAsync(ctx->host_words, P.perr, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    if (ctx->host_words[1] != ~uint64_t(0)) {
      const uint64_t w = ctx->host_words[1];
      return static_cast<int>(w & 0xff);
    }
    restart = false;
    ++since;
    const double rn = std::sqrt(P.hscal[8]), rmax = P.hscal[9], dn = std::sqrt(P.hscal[11]);
    bool recursive = rn <= gates.integrated && rmax <= gates.volume;
    if (gates.relative > 0.0) recursive = recursive && dn <= gates.relative * gates.initial_defect;
    const bool last = it == opt.max_iterations;
    if (last || recursive || since >= 10) {
      k_apply_correction<<<grid_for(P.nc), kThreads, 0, s>>>(P.nc, P.cell_fine, P.cell_flag, P.p_base, P.x32, P.p,
                                                             P.perr);
      FL_TRY(evaluate(ctx, P, P.p, e));
      diag.iteration_count = it;
      publish(e);
      if (e.code) return e.code;
      if (accept(e, gates)) return FL_OK;
      if (last) break;
      k_copy_f64<<<grid_for(P.nc), kThreads, 0, s>>>(P.nc, P.p, P.p_base);
      k_load_residual<<<g, kThreads, 0, s>>>(n, P.fine_cell, P.cell_flag, P.res, P.r32, P.x32, P.perr);
      restart = true;
      since = 0;
      ++diag.restarts;
    }
  }
  return FL_ABORTED;
}

}  // namespace

void destroy_pressure(fl_ctx* ctx) {
  Pressure* P = state(ctx);
  if (!P) return;
  free_levels(*P);
  void* ptrs[] = {P->fine_cell, P->cell_fine, P->cell_flag, P->cell_faces, P->face_tail, P->face_head, P->face_kind,
                  P->face_slot, P->face_area, P->face_beta, P->face_dist, P->face_ustar, P->face_uwall, P->face_flux,
                  P->face_has_flux, P->face_lift, P->face_scale, P->face_k, P->face_wall, P->face_u, P->cell_volume,
                  P->cell_s, P->cell_K, P->cell_prev, P->cell_inwall, P->cell_comp, P->cell_target, P->rhs,
                  P->rhs_abs, P->p, P->p_base, P->res, P->def, P->zero, P->fine_vol, P->r32, P->x32, P->d32, P->q32,
                  P->partial, P->scal, P->fscal, P->perr};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (P->hscal) (void)hipHostFree(P->hscal);
  delete P;
  ctx->pressure = nullptr;
}

}  // namespace fl
