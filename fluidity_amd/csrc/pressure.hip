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
#include "learn.h"

namespace fl {
namespace {

constexpr double kEpsD = 2.220446049250313080847e-16;
constexpr int kPartials = 16384;  // reduction partial slots (one per workgroup of a reducing kernel)
constexpr unsigned kStreamPartials = 2048;  // workgroups of the grid-stride streaming reductions

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

// Frozen link of one face, packed in one 64-byte line for the cell-centred gathers.
struct alignas(64) FaceRec {
  int32_t tail, head, kind, pad;
  double area, ustar, wall, scale, k, lift;
};

struct Level {
  uint64_t nb = 0;
  std::vector<Key3> keys;
  int32_t* nbr = nullptr;    // [nb][6]: -x,+x,-y,+y,-z,+z
  int32_t* child = nullptr;   // [nb][8] blocks of the finer level (levels > 0)
  int32_t* parent = nullptr;  // [nb][2] coarser block and octant (all but the coarsest level)
  float* x = nullptr;
  float* b = nullptr;
  float* r = nullptr;       // [nb*64] residual (red half used)
  uint16_t* kb = nullptr;   // [3][nb*64] BF16 copies of k for the V-cycle (BF16 level storage)
  float* diagb = nullptr;   // [nb*64] diagonal of the BF16-coefficient operator
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
  FaceRec* face_rec = nullptr;
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
  double* p_init = nullptr;  // warm start of every solve (previous pressure minus gauge pin)
  double* res = nullptr;
  double* def = nullptr;
  float* fine_vol = nullptr;  // [nb0*64] fluid volume (0 if none)
  // 8^3 tiles of level 0 for the learned preconditioner: 2x2x2 blocks each, and a class per cell
  int32_t* tile_blocks = nullptr;  // [ntiles][8] level-0 block or -1, octant (z,y,x) order
  uint8_t* cell_class = nullptr;   // [nb0*64] 1 liquid unknown, 2 air, 3 solid
  uint64_t ntiles = 0;
  // PCG (level-0 sized)
  float* r32 = nullptr;
  float* x32 = nullptr;
  float* d32 = nullptr;
  float* dn32 = nullptr;  // the next direction (swapped with d32 every iteration)
  float* q32 = nullptr;
  double* partial = nullptr;  // [8][kPartials]
  double* scal = nullptr;     // device scalars
  float* fscal = nullptr;
  double* hscal = nullptr;    // pinned host mirror
  uint64_t* perr = nullptr;   // error words
  // host copies
  std::vector<int32_t> h_cell_comp_root;
  std::vector<uint64_t> h_gauges;
  std::vector<Key3> h_unknown_cells;  // level-0 unknown cell coordinates
  // Multigrid level storage: one device pool reserved with the other arrays, carved level by level.
  unsigned char* level_pool = nullptr;
  uint64_t pool_bytes = 0, pool_used = 0;
  int levels_requested = 0;
  int last_storage = 1, last_nu = 2, last_coarse_sweeps = 2;  // V-cycle settings of the last solve
  bool p_init_zero = false;  // warm start is identically zero: its evaluation equals the p = 0 one
  bool has_jump = false;
  // Frozen link coefficients and hierarchy coefficients of the current system for step duration
  // coeff_dt (valid until the system, its levels or its coefficients change); BF16 level copies valid.
  bool coeff_valid = false, bf16_valid = false;
  double coeff_dt = 0.0;
  double full_area = 0, full_volume = 0;
  double scale_coarse = 0.5;
  // Scratch of device-built systems (prepare_system_device): level-0 block marks and their scan over
  // the grid's blocks, the face of each owned slot, the union-find forest and anchors of the cells,
  // small counters, a block lookup and unknown-cell masks per level, and the block keys of every level.
  uint64_t* dv_mark = nullptr;       // 2 * cap_dv_boxes
  int32_t* dv_slot_face = nullptr;   // 192 * cap_dv_grid
  int32_t* dv_parent = nullptr;      // cap_dv_cells
  int32_t* dv_root = nullptr;        // cap_dv_cells
  uint8_t* dv_anchor = nullptr;      // cap_dv_cells
  uint64_t* dv_words = nullptr;      // 8 counters
  int32_t* dv_lut = nullptr;         // cap_dv_boxes
  uint64_t* dv_occ[2] = {nullptr, nullptr};  // cap_dv_grid each
  int32_t* dv_keys = nullptr;        // 3 * 2 * cap_dv_grid
  uint64_t cap_dv_grid = 0, cap_dv_cells = 0, cap_dv_boxes = 0;
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

// Forgets the levels; their storage stays in the pool for the next hierarchy.
void free_levels(Pressure& P) {
  P.levels.clear();
  P.pool_used = 0;
}

constexpr uint64_t kPoolAlign = 256;

uint64_t pool_round(uint64_t b) { return (b + kPoolAlign - 1) / kPoolAlign * kPoolAlign; }

// Pool bytes of one level of nb blocks (tables and the ten per-cell arrays).
uint64_t level_bytes(uint64_t nb) {
  const uint64_t n = 64 * nb;
  return pool_round(6 * nb * 4) + pool_round(8 * nb * 4) + pool_round(2 * nb * 4) + 6 * pool_round(n * 4) +
         pool_round(3 * n * 2) + pool_round(3 * n * 4);
}

template <typename T>
T* carve(Pressure& P, uint64_t count) {
  T* p = reinterpret_cast<T*>(P.level_pool + P.pool_used);
  P.pool_used += pool_round(count * sizeof(T));
  return p;
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

// Cells of a 4^3 block are stored colour-major: the 32 cells with x+y+z even first, then the 32 odd
// ones, each half in (z,y,x) order. l = z*16 + y*4 + x is the natural local index. A red-black half
// sweep then reads and writes one contiguous half of every block.
__host__ __device__ inline int cslot(int l) {
  return ((((l & 3) + ((l >> 2) & 3) + (l >> 4)) & 1) << 5) | (l >> 1);
}
__host__ __device__ inline int cnat(int s) {
  const int l0 = (s & 31) << 1;
  const int parity = ((l0 & 3) + ((l0 >> 2) & 3) + (l0 >> 4)) & 1;
  return l0 + (parity ^ (s >> 5));
}

// Neighbour cell index within a level.
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
  return b * 64 + cslot(nz * 16 + ny * 4 + nx);
}

// Conductance of face `dir` of cell i: low faces are stored on the cell, high faces on the neighbour.
__device__ inline float face_k(const float* k, uint64_t n, int64_t i, int64_t j, int dir) {
  const int axis = dir >> 1;
  return (dir & 1) ? k[axis * n + j] : k[axis * n + i];
}

// Coefficients are FP32 or BF16 (the high half of an FP32 pattern); all arithmetic is FP32.
struct Bf16 {
  uint16_t bits;
};
__device__ inline float kload(const float* k, int64_t i) { return k[i]; }
__device__ inline float kload(const Bf16* k, int64_t i) { return __uint_as_float(uint32_t(k[i].bits) << 16); }
__device__ __host__ inline uint16_t to_bf16(float v) {
  uint32_t u;
  memcpy(&u, &v, 4);
  return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // round to nearest even (finite v)
}


// Conductance of face `dir` of cell i (low faces on the cell, high faces on the neighbour).
template <class KT>
__device__ inline float face_kt(const KT* k, uint64_t n, int64_t i, int64_t j, int dir) {
  const int axis = dir >> 1;
  return (dir & 1) ? kload(k, axis * int64_t(n) + j) : kload(k, axis * int64_t(n) + i);
}

// ------------------------------------------------------------------ system kernels (FP64)


__global__ void k_freeze_faces(uint64_t nf, double dt, double full_area, const int32_t* tail, const int32_t* head,
                               const int32_t* kind, const double* area, const double* beta, const double* dist,
                               const double* ustar, const double* uwall, const int32_t* has_flux, const double* flux,
                               const double* lift, double* scale, double* k, double* wall, FaceRec* rec) {
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
  const double w = has_flux[f] ? flux[f] : blocked * uwall[f];
  wall[f] = w;
  FaceRec r;
  r.tail = tail[f];
  r.head = head[f];
  r.kind = kind[f];
  r.pad = 0;
  r.area = area[f];
  r.ustar = ustar[f];
  r.wall = w;
  r.scale = s;
  r.k = c;
  r.lift = lift[f];
  rec[f] = r;
}

// The predicted normal velocity of every face record (the only record field a new u* changes).
__global__ void k_refresh_ustar(uint64_t nf, const double* ustar, FaceRec* rec) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f < nf) rec[f].ustar = ustar[f];
}

__global__ void k_freeze_cells(uint64_t nc, double dt, const double* V, const double* s, const double* K,
                               double* comp, double* target) {
  const uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (c >= nc) return;
  comp[c] = isinf(K[c]) ? 0.0 : V[c] / (K[c] * dt);
  target[c] = V[c] * s[c];
}

// Integrated RHS (mac_pressure_3d_arithmetic.h:263-313) with Neumaier sums per cell, fused with the
// evaluation at p = 0 (the defect of the unprojected field, which fixes the relative-divergence gate):
// the terms of k_evaluate in the same order with every cell pressure zero, so the operator part
// vanishes and r = b. One lane per cell, in input (z,y,x) cell order.
__global__ void k_rhs_zero_eval(uint64_t nc, const int32_t* cell_faces, const FaceRec* faces, const double* V,
                                const double* target, const double* inwall, const double* comp, const double* prev,
                                double* rhs, double* rhs_abs, double* res, double* def, double* partial,
                                uint64_t* err);

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

// IsRepresentablePressureProduct (mac_pressure_3d_arithmetic.h:33-49): a product of the evaluation is
// usable when it is finite and, if it fell below the normal range with both factors nonzero, it is still
// the exact product (checked on the factors' mantissas, with the product brought back to their scale).
__device__ inline bool exact_product(double a, double b, double prod) {
  if (!finite(prod)) return false;
  if (a == 0.0 || b == 0.0 || fabs(prod) >= __DBL_MIN__) return true;
  if (prod == 0.0) return false;
  int ea = 0, eb = 0;
  const double ma = frexp(a, &ea), mb = frexp(b, &eb);
  return fma(ma, mb, -scalbn(prod, -(ea + eb))) == 0.0;
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
// Cells are visited in input order, one per lane (the per-cell work is a chain of dependent loads, so
// the launch is as wide as the system).
__global__ void k_evaluate(uint64_t nc, const int32_t* cell_faces, const FaceRec* faces,
                           const double* V, const double* target, const double* inwall, const double* comp,
                           const double* prev, const double* rhs, const double* rhs_abs, const double* p, double* res,
                           double* def, double* partial, uint64_t* err) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < nc; i += uint64_t(gridDim.x) * blockDim.x) {
    const int32_t c = static_cast<int32_t>(i);
    Row op, phys;
    const double pc = p[c];
    const double cflux = comp[c] * pc;
    bool exact = exact_product(comp[c], pc, cflux);
    op.add(cflux);
    phys.add(-target[c]);
    phys.add(-inwall[c]);
    if (comp[c] > 0.0) {
      const double dp = pc - prev[c], dflux = comp[c] * dp;
      exact = exact && exact_product(comp[c], dp, dflux);
      phys.add(dflux);
    }
    for (int j = 0; j < 6; ++j) {
      const FaceRec& f = faces[cell_faces[6 * c + j]];
      const double sign = f.tail == c ? 1.0 : -1.0;
      const double pt_op = f.tail >= 0 ? p[f.tail] : 0.0;
      const double ph_op = f.head >= 0 ? p[f.head] : 0.0;
      if (f.k != 0.0) {
        const double dpf = pt_op - ph_op, kflux = f.k * dpf;
        exact = exact && exact_product(f.k, dpf, kflux);
        op.add(sign * kflux);
      }
      double u = f.ustar;
      if (f.area > 0.0) {
        const double ext = f.kind == 2 ? f.lift : 0.0;
        const double jump = f.kind == 1 ? f.lift : 0.0;
        const double pt = f.tail >= 0 ? pt_op : ext;
        const double ph = f.head >= 0 ? ph_op : ext;
        const double lifted = pt - ph + jump, du = f.scale * lifted;
        exact = exact && exact_product(f.scale, lifted, du);
        u += du;
      }
      const double aflux = f.area * u;
      exact = exact && finite(u) && exact_product(f.area, u, aflux);
      phys.add(sign * aflux);
      phys.add(sign * f.wall);
    }
    const double r = rhs[c] - op.flux.total();
    const double d = phys.flux.total() / V[c];
    const double ie = V[c] * d + r;
    const double terms = rhs_abs[c] + op.abs.total() + phys.abs.total();
    if (!exact || !finite(r) || !finite(d) || !finite(ie) || !finite(terms)) perr_report(err, c, kOutOfRange);
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

// The operator residual r = b - A p of k_evaluate alone (same terms, same order, same compensated
// sum), for the periodic restarts of the PCG that are not convergence checks: it reads the 16 B of
// conductance and endpoints per face instead of the 64 B face record, and visits the cells in level-0
// block order so that a wave's faces are one contiguous run of the face arrays.
__global__ void k_true_residual(uint64_t n0, const int32_t* fine_cell, const int32_t* cell_faces,
                                const double* face_k, const int32_t* face_tail, const int32_t* face_head,
                                const double* comp, const double* rhs, const double* p, double* res) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n0) return;
  const int32_t c = fine_cell[i];
  if (c < 0) return;
  Row op;
  const double pc = p[c];
  op.add(comp[c] * pc);
  for (int j = 0; j < 6; ++j) {
    const int32_t f = cell_faces[6 * c + j];
    const double fk = face_k[f];
    const int32_t tail = face_tail[f], head = face_head[f];
    const double sign = tail == c ? 1.0 : -1.0;
    const double pt = tail >= 0 ? p[tail] : 0.0;
    const double ph = head >= 0 ? p[head] : 0.0;
    if (fk != 0.0) op.add(sign * (fk * (pt - ph)));
  }
  res[c] = rhs[c] - op.flux.total();
}

__global__ void k_rhs_zero_eval(uint64_t nc, const int32_t* cell_faces, const FaceRec* faces, const double* V,
                                const double* target, const double* inwall, const double* comp, const double* prev,
                                double* rhs, double* rhs_abs, double* res, double* def, double* partial,
                                uint64_t* err) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < nc; i += uint64_t(gridDim.x) * blockDim.x) {
    const int32_t c = static_cast<int32_t>(i);
    Row row, phys;
    row.add(target[c]);
    row.add(inwall[c]);
    row.add(comp[c] * prev[c]);
    phys.add(-target[c]);
    phys.add(-inwall[c]);
    bool exact = true;
    if (comp[c] > 0.0) {
      const double dp = 0.0 - prev[c], dflux = comp[c] * dp;
      exact = exact_product(comp[c], dp, dflux);
      phys.add(dflux);
    }
    for (int j = 0; j < 6; ++j) {
      const FaceRec& f = faces[cell_faces[6 * c + j]];
      const double sign = f.tail == c ? 1.0 : -1.0;
      row.add(-sign * (f.area * f.ustar));
      row.add(-sign * f.wall);
      if (f.kind == 1) row.add(-sign * (f.k * f.lift));
      else if (f.kind == 2) row.add(f.k * f.lift);
      double u = f.ustar;
      if (f.area > 0.0) {
        const double ext = f.kind == 2 ? f.lift : 0.0;
        const double jump = f.kind == 1 ? f.lift : 0.0;
        const double pt = f.tail >= 0 ? 0.0 : ext;
        const double ph = f.head >= 0 ? 0.0 : ext;
        const double lifted = pt - ph + jump, du = f.scale * lifted;
        exact = exact && exact_product(f.scale, lifted, du);
        u += du;
      }
      const double aflux = f.area * u;
      exact = exact && finite(u) && exact_product(f.area, u, aflux);
      phys.add(sign * aflux);
      phys.add(sign * f.wall);
    }
    const double b = row.flux.total(), babs = row.abs.total();
    rhs[c] = b;
    rhs_abs[c] = babs;
    const double r = b - 0.0;
    const double d = phys.flux.total() / V[c];
    const double ie = V[c] * d + r;
    const double terms = babs + 0.0 + phys.abs.total();
    if (!exact || !finite(r) || !finite(d) || !finite(ie) || !finite(terms)) perr_report(err, c, kOutOfRange);
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
// Reads the face conductances and endpoints from their own arrays (16 B per face instead of the 64 B
// face record).
__global__ void k_level0(uint64_t n, const int32_t* fine_cell, const uint8_t* flag, const int32_t* cell_faces,
                         const double* face_k, const int32_t* face_tail, const int32_t* face_head, const double* comp,
                         const double* V, float* kk, float* anchor, float* diag, float* vol) {
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
      const double fk = face_k[f];
      if (!(fk > 0.0)) continue;
      const int32_t tail = face_tail[f];
      const int32_t other = tail == c ? face_head[f] : tail;
      if (other >= 0 && flag[other] == 1) {
        ksum[j] = fk;
        if ((j & 1) == 0) kl[j >> 1] = static_cast<float>(fk);
      } else {
        a += fk;
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
  const int l = cnat(static_cast<int>(i & 63));
  const int cx = l & 3, cy = (l >> 2) & 3, cz = l >> 4;
  double a = 0.0, kx = 0.0, ky = 0.0, kz = 0.0;
  bool any = false;
  for (int oz = 0; oz < 2; ++oz)
    for (int oy = 0; oy < 2; ++oy)
      for (int ox = 0; ox < 2; ++ox) {
        const int fx = 2 * cx + ox, fy = 2 * cy + oy, fz = 2 * cz + oz;
        const int32_t fb = child[8 * blk + ((fz >> 2) * 2 + (fy >> 2)) * 2 + (fx >> 2)];
        if (fb < 0) continue;
        const uint64_t fi = uint64_t(fb) * 64 + cslot((fz & 3) * 16 + (fy & 3) * 4 + (fx & 3));
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

// BF16 copies of a level's conductances and the diagonal of the operator they define.
__global__ void k_to_bf16(const float* k, uint64_t n, uint16_t* kb) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) kb[i] = to_bf16(k[i]);
}
__global__ void k_diag_bf16(uint64_t n, const int32_t* nbr, const uint16_t* kb, const float* anchor, const float* diag,
                            float* diagb) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  if (diag[i] == 0.f) {
    diagb[i] = 0.f;
    return;
  }
  const uint64_t blk = i >> 6;
  const int l = cnat(static_cast<int>(i & 63));
  const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
  float d = anchor[i];
  for (int dir = 0; dir < 6; ++dir) {
    const int64_t j = neighbour(nbr, blk, lx, ly, lz, dir);
    if (j < 0) continue;
    d += face_kt(reinterpret_cast<const Bf16*>(kb), n, static_cast<int64_t>(i), j, dir);
  }
  diagb[i] = d;
}

__global__ void k_diag(uint64_t n, const int32_t* nbr, const float* k, const float* anchor, float* diag) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  if (diag[i] == 0.f) return;
  const uint64_t blk = i >> 6;
  const int l = cnat(static_cast<int>(i & 63));
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

// Per-cell V-cycle operations, shared by the grid-wide kernels and the single-workgroup tail cycle.
template <class KT>
struct LevelViewT {
  uint64_t n;             // cells (64 per block)
  const int32_t* nbr;     // [nb][6]
  const int32_t* child;   // [nb][8] finer blocks (levels > 0)
  const int32_t* parent;  // [nb][2] coarser block and octant
  float* x;
  const float* b;
  const KT* k;
  const float* anchor;
  const float* diag;
};
using LevelView = LevelViewT<float>;


// Gauss-Seidel update of the cell at colour-half position t (block t/32, slot colour*32 + t%32).
// `first` means x == 0 on entry, so the update is b/d. Cells that are not unknowns keep x = 0.
template <class KT>
__device__ inline void smooth_cell(const LevelViewT<KT>& L, uint64_t t, int color, int first) {
  const uint64_t blk = t >> 5;
  const int sl = (color << 5) | static_cast<int>(t & 31);
  const uint64_t i = blk * 64 + static_cast<uint64_t>(sl);
  // Every load is issued before any of them is used (no early exit on non-unknown cells), so the
  // independent reads overlap; missing neighbours read slot i with a zero coefficient.
  const float d = L.diag[i];
  float s = L.b[i];
  if (!first) {
    const int l = cnat(sl);
    const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
    float kk[6], xx[6];
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      const int64_t j = neighbour(L.nbr, blk, lx, ly, lz, dir);
      const int64_t jj = j < 0 ? static_cast<int64_t>(i) : j;
      const float kv = face_kt(L.k, L.n, static_cast<int64_t>(i), jj, dir);
      const float xv = L.x[jj];
      kk[dir] = j < 0 ? 0.f : kv;
      xx[dir] = j < 0 ? 0.f : xv;
    }
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) s += kk[dir] * xx[dir];
  }
  L.x[i] = d == 0.f ? 0.f : s / d;
}

// Coarse right-hand side of coarse cell ci: the sum of the fine residuals b - A x of its 2x2x2
// children. Pre-smoothing ends with a black half sweep, which leaves every black (odd) cell's
// equation satisfied, so only the four red children carry a residual.
template <class KT>
__device__ inline float restrict_residual(const LevelViewT<KT>& F, const int32_t* child, uint64_t ci) {
  const uint64_t blk = ci >> 6;
  const int l = cnat(static_cast<int>(ci & 63));
  const int cx = l & 3, cy = (l >> 2) & 3, cz = l >> 4;
  float sum = 0.f;
  for (int o = 0; o < 8; ++o) {
    const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
    if (((ox + oy + oz) & 1) != 0) continue;  // fine parity = own parity (2c is even): odd octants are black
    const int fx = 2 * cx + ox, fy = 2 * cy + oy, fz = 2 * cz + oz;
    const int32_t fb = child[8 * blk + ((fz >> 2) * 2 + (fy >> 2)) * 2 + (fx >> 2)];
    if (fb < 0) continue;
    const int lx = fx & 3, ly = fy & 3, lz = fz & 3;
    const uint64_t fi = uint64_t(fb) * 64 + cslot(lz * 16 + ly * 4 + lx);
    if (F.diag[fi] == 0.f) continue;
    const float xi = F.x[fi];
    float a = F.anchor[fi] * xi;
    for (int dir = 0; dir < 6; ++dir) {
      const int64_t j = neighbour(F.nbr, uint64_t(fb), lx, ly, lz, dir);
      if (j < 0) continue;
      a += face_kt(F.k, F.n, static_cast<int64_t>(fi), j, dir) * (xi - F.x[j]);
    }
    sum += F.b[fi] - a;
  }
  return sum;
}

// x_fine += x_coarse(parent).
template <class KT>
__device__ inline void prolong_cell(const LevelViewT<KT>& F, const float* cx, uint64_t i) {
  if (F.diag[i] == 0.f) return;
  const uint64_t blk = i >> 6;
  const int32_t pb = F.parent[2 * blk], oct = F.parent[2 * blk + 1];
  const int l = cnat(static_cast<int>(i & 63));
  const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
  const int px = ((oct & 1) * 4 + lx) >> 1, py = (((oct >> 1) & 1) * 4 + ly) >> 1, pz = ((oct >> 2) * 4 + lz) >> 1;
  F.x[i] += cx[uint64_t(pb) * 64 + cslot(pz * 16 + py * 4 + px)];
}

// One red-black Gauss-Seidel half sweep over the `color` half of every block.
template <class KT>
__global__ void k_smooth(LevelViewT<KT> L, int color, int first) {
  const uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (t < L.n / 2) smooth_cell(L, t, color, first);
}

// Residual b - A x of the red (even) cells after pre-smoothing, one lane per red cell, into r's red
// half; the black cells' residual is zero after the closing black half sweep.
template <class KT>
__global__ void k_residual_red(LevelViewT<KT> L, float* r) {
  const uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (t >= L.n / 2) return;
  const uint64_t blk = t >> 5;
  const int sl = static_cast<int>(t & 31);
  const uint64_t i = blk * 64 + static_cast<uint64_t>(sl);
  if (L.diag[i] == 0.f) {
    r[i] = 0.f;
    return;
  }
  const int l = cnat(sl);
  const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
  const float xi = L.x[i];
  float a = L.anchor[i] * xi;
  for (int dir = 0; dir < 6; ++dir) {
    const int64_t j = neighbour(L.nbr, blk, lx, ly, lz, dir);
    if (j < 0) continue;
    a += face_kt(L.k, L.n, static_cast<int64_t>(i), j, dir) * (xi - L.x[j]);
  }
  r[i] = L.b[i] - a;
}

// Coarse right-hand side: the sum of the red children's residuals (pull, fixed order).
template <class KT>
__global__ void k_restrict_red(const float* r, LevelViewT<KT> C) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= C.n) return;
  const uint64_t blk = i >> 6;
  const int l = cnat(static_cast<int>(i & 63));
  const int cx = l & 3, cy = (l >> 2) & 3, cz = l >> 4;
  float sum = 0.f;
  for (int o = 0; o < 8; ++o) {
    const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
    if (((ox + oy + oz) & 1) != 0) continue;
    const int fx = 2 * cx + ox, fy = 2 * cy + oy, fz = 2 * cz + oz;
    const int32_t fb = C.child[8 * blk + ((fz >> 2) * 2 + (fy >> 2)) * 2 + (fx >> 2)];
    if (fb < 0) continue;
    sum += r[uint64_t(fb) * 64 + cslot((fz & 3) * 16 + (fy & 3) * 4 + (fx & 3))];
  }
  const_cast<float*>(C.b)[i] = sum;
}

template <class KT>
__global__ void k_prolong(LevelViewT<KT> F, const float* cx) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < F.n) prolong_cell(F, cx, i);
}

// ---- Block smoother: red-black Gauss-Seidel inside each 4^3 block, block Jacobi between blocks ----
// One wave per block, one lane per cell (colour-major slots: lanes 0-31 the even cells, 32-63 the
// odd ones). The block's values live in LDS for the whole smoothing step; values across the block
// faces are read once, at the start, and held fixed. A pre-smoothing step starts from x = 0 and
// sweeps red then black `sweeps` times; a post-smoothing step first adds the coarse correction
// (prolongation) to the block and to its halo, sweeps black then red (the adjoint order, so the
// V-cycle stays symmetric), and writes to a second buffer so that no block reads a neighbour's
// halo while that neighbour writes it. Per cell this reads x, b, diag and the coefficients once,
// instead of once per half sweep.
template <class KT>
__device__ inline float prolonged(const LevelViewT<KT>& F, const float* xc, uint64_t blk, int l) {
  const int32_t pb = F.parent[2 * blk], oct = F.parent[2 * blk + 1];
  const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
  const int px = ((oct & 1) * 4 + lx) >> 1, py = (((oct >> 1) & 1) * 4 + ly) >> 1, pz = ((oct >> 2) * 4 + lz) >> 1;
  return xc[uint64_t(pb) * 64 + cslot(pz * 16 + py * 4 + px)];
}

template <class KT, bool kPost>
__global__ __launch_bounds__(256) void k_block_smooth(LevelViewT<KT> L, float* xout, const float* xc, int sweeps) {
  __shared__ float sx[4][64];
  const int w = threadIdx.x >> 6, sl = threadIdx.x & 63;
  const uint64_t blk = blockIdx.x * 4ull + w;
  const bool live = blk < L.n / 64;
  const uint64_t i = blk * 64 + sl;
  const int l = cnat(sl);
  const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
  float d = 0.f, bb = 0.f, x0 = 0.f;
  float kk[6], halo[6];
  int inner[6];
#pragma unroll
  for (int dir = 0; dir < 6; ++dir) {
    kk[dir] = 0.f;
    halo[dir] = 0.f;
    inner[dir] = -1;
  }
  if (live) {
    d = L.diag[i];
    bb = L.b[i];
    if (kPost && d != 0.f) x0 = L.x[i] + prolonged(L, xc, blk, l);
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      int nx = lx, ny = ly, nz = lz;
      if (dir == 0) nx -= 1;
      if (dir == 1) nx += 1;
      if (dir == 2) ny -= 1;
      if (dir == 3) ny += 1;
      if (dir == 4) nz -= 1;
      if (dir == 5) nz += 1;
      const bool outside = nx < 0 || nx > 3 || ny < 0 || ny > 3 || nz < 0 || nz > 3;
      const int64_t nb = outside ? L.nbr[6 * blk + dir] : static_cast<int64_t>(blk);
      if (nb < 0) continue;
      const int nl = (nz & 3) * 16 + (ny & 3) * 4 + (nx & 3);
      const int64_t j = nb * 64 + cslot(nl);
      kk[dir] = face_kt(L.k, L.n, static_cast<int64_t>(i), j, dir);
      if (!outside) {
        inner[dir] = cslot(nl);
      } else if (kPost) {
        const float xj = L.x[j];
        halo[dir] = L.diag[j] == 0.f ? xj : xj + prolonged(L, xc, static_cast<uint64_t>(nb), nl);
      }
    }
  }
  sx[w][sl] = x0;
  __syncthreads();
  for (int it = 0; it < sweeps; ++it) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int colour = kPost ? 1 - h : h;
      if ((sl >> 5) == colour) {
        float acc = bb;
#pragma unroll
        for (int dir = 0; dir < 6; ++dir) acc += kk[dir] * (inner[dir] >= 0 ? sx[w][inner[dir]] : halo[dir]);
        sx[w][sl] = d == 0.f ? 0.f : acc / d;
      }
      __syncthreads();
    }
  }
  if (live) (kPost ? xout : L.x)[i] = sx[w][sl];
}

// Residual b - A x of every cell (after block smoothing the odd cells at block faces carry a residual
// too), then the coarse right-hand side as the sum over all eight children.
template <class KT>
__global__ void k_residual_all(LevelViewT<KT> L, float* r) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= L.n) return;
  if (L.diag[i] == 0.f) {
    r[i] = 0.f;
    return;
  }
  const uint64_t blk = i >> 6;
  const int l = cnat(static_cast<int>(i & 63));
  const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
  const float xi = L.x[i];
  float a = L.anchor[i] * xi;
  for (int dir = 0; dir < 6; ++dir) {
    const int64_t j = neighbour(L.nbr, blk, lx, ly, lz, dir);
    if (j < 0) continue;
    a += face_kt(L.k, L.n, static_cast<int64_t>(i), j, dir) * (xi - L.x[j]);
  }
  r[i] = L.b[i] - a;
}

template <class KT>
__global__ void k_restrict_all(const float* r, LevelViewT<KT> C) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= C.n) return;
  const uint64_t blk = i >> 6;
  const int l = cnat(static_cast<int>(i & 63));
  const int cx = l & 3, cy = (l >> 2) & 3, cz = l >> 4;
  float sum = 0.f;
  for (int o = 0; o < 8; ++o) {
    const int fx = 2 * cx + (o & 1), fy = 2 * cy + ((o >> 1) & 1), fz = 2 * cz + (o >> 2);
    const int32_t fb = C.child[8 * blk + ((fz >> 2) * 2 + (fy >> 2)) * 2 + (fx >> 2)];
    if (fb < 0) continue;
    sum += r[uint64_t(fb) * 64 + cslot((fz & 3) * 16 + (fy & 3) * 4 + (fx & 3))];
  }
  const_cast<float*>(C.b)[i] = sum;
}

constexpr int kMaxTail = 16;
constexpr uint64_t kTailBlocks = 32;  // levels with at most this many blocks run in the tail cycle
template <class KT>
struct TailCycle {
  int count, nu, coarse_sweeps;
  LevelViewT<KT> lv[kMaxTail];
};

// The V-cycle below and including the first small level, in one workgroup: pre-smoothing and
// residual restriction down to the coarsest level, symmetric sweeps there from zero, then
// prolongation and post-smoothing back up. Barriers separate the dependent steps; every wave of the
// workgroup shares the CU's vector cache, so each step sees the previous one's stores.
template <class KT>
__global__ __launch_bounds__(1024) void k_tail_vcycle(TailCycle<KT> T) {
  const uint64_t tid = threadIdx.x, nt = blockDim.x;
  auto sync = [] {
    __threadfence_block();
    __syncthreads();
  };
  for (int q = 0; q + 1 < T.count; ++q) {
    const LevelViewT<KT>& L = T.lv[q];
    for (int it = 0; it < T.nu; ++it) {
      for (uint64_t t = tid; t < L.n / 2; t += nt) smooth_cell(L, t, 0, it == 0);
      sync();
      for (uint64_t t = tid; t < L.n / 2; t += nt) smooth_cell(L, t, 1, 0);
      sync();
    }
    const LevelViewT<KT>& C = T.lv[q + 1];
    for (uint64_t i = tid; i < C.n; i += nt) const_cast<float*>(C.b)[i] = restrict_residual(L, C.child, i);
    sync();
  }
  const LevelViewT<KT>& B = T.lv[T.count - 1];
  for (int pass = 0; pass < 4 * T.coarse_sweeps; ++pass) {
    const int fwd = pass < 2 * T.coarse_sweeps;
    const int color = fwd ? (pass & 1) : 1 - (pass & 1);
    for (uint64_t t = tid; t < B.n / 2; t += nt) smooth_cell(B, t, color, pass == 0);
    sync();
  }
  for (int q = T.count - 2; q >= 0; --q) {
    const LevelViewT<KT>& L = T.lv[q];
    for (uint64_t i = tid; i < L.n; i += nt) prolong_cell(L, T.lv[q + 1].x, i);
    sync();
    for (int it = 0; it < T.nu; ++it) {
      for (uint64_t t = tid; t < L.n / 2; t += nt) smooth_cell(L, t, 1, 0);
      sync();
      for (uint64_t t = tid; t < L.n / 2; t += nt) smooth_cell(L, t, 0, 0);
      sync();
    }
  }
}

// y = A x in the edge-difference form.
__global__ void k_apply(uint64_t n, const int32_t* nbr, const float* k, const float* anchor, const float* diag,
                        const float* x, float* y) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  if (diag[i] == 0.f) {
    y[i] = 0.f;
    return;
  }
  const uint64_t blk = i >> 6;
  const int l = cnat(static_cast<int>(i & 63));
  const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
  const float xi = x[i];
  float s = anchor[i] * xi;
  for (int dir = 0; dir < 6; ++dir) {
    const int64_t j = neighbour(nbr, blk, lx, ly, lz, dir);
    if (j < 0) continue;
    s += face_k(k, n, static_cast<int64_t>(i), j, dir) * (xi - x[j]);
  }
  y[i] = s;
}

// ------------------------------------------------------------------ learned preconditioner tiles

// Level-0 storage index of voxel v of tile t (8^3, (z*8 + y)*8 + x), or -1 outside the level's blocks.
__device__ inline int64_t tile_cell(const int32_t* tile_blocks, uint64_t t, int v) {
  const int z = v >> 6, y = (v >> 3) & 7, x = v & 7;
  const int32_t b = tile_blocks[8 * t + ((z >> 2) * 2 + (y >> 2)) * 2 + (x >> 2)];
  if (b < 0) return -1;
  return int64_t(b) * 64 + cslot((z & 3) * 16 + (y & 3) * 4 + (x & 3));
}

// Network inputs of every tile from the level-0 residual: the residual divided by its RMS over the
// tile's liquid cells, and one-hot liquid / air / solid classes (cells outside the level's blocks are
// air). The output scale of the tile is RMS / (mean diagonal over its liquid cells), so the network maps
// a dimensionless residual to a dimensionless correction. One workgroup of 512 lanes per tile.
__global__ __launch_bounds__(512) void k_tile_gather(const float* r, const float* diag, const uint8_t* cls,
                                                     const int32_t* tile_blocks, float* act, float* scale) {
  __shared__ float red[3][512];
  const uint64_t t = blockIdx.x;
  const int v = threadIdx.x;
  const int64_t i = tile_cell(tile_blocks, t, v);
  const int c = i >= 0 ? cls[i] : 2;
  const float rv = (i >= 0 && c == 1) ? r[i] : 0.f;
  red[0][v] = rv * rv;
  red[1][v] = c == 1 ? 1.f : 0.f;
  red[2][v] = (i >= 0 && c == 1) ? diag[i] : 0.f;
  __syncthreads();
  for (int s = 256; s > 0; s >>= 1) {
    if (v < s) {
      red[0][v] += red[0][v + s];
      red[1][v] += red[1][v + s];
      red[2][v] += red[2][v + s];
    }
    __syncthreads();
  }
  const float sum = red[0][0], count = red[1][0], dsum = red[2][0];
  const float rms = (count > 0.f && sum > 0.f) ? sqrtf(sum / count) : 0.f;
  float* in = act + t * kActFloats + kActIn;
  in[v] = rms > 0.f ? rv / rms : 0.f;
  in[512 + v] = c == 1 ? 1.f : 0.f;
  in[1024 + v] = c == 2 ? 1.f : 0.f;
  in[1536 + v] = c == 3 ? 1.f : 0.f;
  if (v == 0) scale[t] = (rms > 0.f && dsum > 0.f) ? rms * count / dsum : 0.f;
}

// z = network output times the tile's scale on liquid unknowns, zero elsewhere. Tiles partition the
// level's blocks, so every cell is written once.
__global__ void k_tile_scatter(const float* act, const float* scale, const uint8_t* cls, const int32_t* tile_blocks,
                               uint64_t ntiles, float* z) {
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (gid >= ntiles * 512) return;
  const uint64_t t = gid >> 9;
  const int v = static_cast<int>(gid & 511);
  const int64_t i = tile_cell(tile_blocks, t, v);
  if (i < 0) return;
  const float zv = act[t * kActFloats + kActOut + v] * scale[t];
  z[i] = (cls[i] == 1 && isfinite(zv)) ? zv : 0.f;  // a non-finite output counts as no correction
}

// Training pairs: the gathered inputs and the V-cycle's correction divided by the same tile scale.
__global__ void k_tile_pairs(const float* act, const float* scale, const float* corr, const uint8_t* cls,
                             const int32_t* tile_blocks, uint64_t ntiles, float* data_in, float* data_target) {
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (gid >= ntiles * 512) return;
  const uint64_t t = gid >> 9;
  const int v = static_cast<int>(gid & 511);
  for (int c = 0; c < 4; ++c) data_in[t * 2048 + c * 512 + v] = act[t * kActFloats + kActIn + c * 512 + v];
  const int64_t i = tile_cell(tile_blocks, t, v);
  const float sc = scale[t];
  data_target[t * 512 + v] = (i >= 0 && cls[i] == 1 && sc > 0.f) ? corr[i] / sc : 0.f;
}

// ------------------------------------------------------------------ PCG kernels

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

__device__ void after_rho(double* scal, float* fscal, uint64_t* err, int restart) {
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

// Flexible CG (one-term orthogonalisation): beta = -(z . q_prev) / (d_prev . q_prev), with z . q_prev in
// scal[28] and the previous curvature in scal[2]. The step length stays (r . z) / (d . q), which the
// orthogonality of r to d_prev makes equal to (r . d) / (d . q); it may be of either sign.
__device__ void after_rho_fcg(double* scal, float* fscal, uint64_t* err, int restart) {
  const double rho = scal[20];
  scal[0] = rho;
  scal[14] = 0.0;
  if (!finite(rho)) {
    perr_report(err + 1, 0, kOutOfRange);
    return;
  }
  if (restart) {
    fscal[1] = 0.f;
  } else {
    float beta;
    if (!fine_value(-scal[28] / scal[2], beta)) {
      perr_report(err + 1, 1, kOutOfRange);
      return;
    }
    fscal[1] = beta;
  }
}

// Flexible-CG curvature: a zero direction (a network output orthogonal to everything) is not an error
// here; the step is skipped and scal[14] tells the host to hand over to the V-cycle.
__device__ void after_curv_fcg(double* scal, float* fscal, uint64_t* err) {
  const double curv = scal[20];
  float alpha = 0.f;
  if (!(curv > 0.0) || !fine_value(scal[0] / curv, alpha)) {
    scal[14] = 1.0;
    fscal[0] = 0.f;
    return;
  }
  scal[2] = curv;
  fscal[0] = alpha;
  scal[1] = scal[0];
}

__device__ void after_curv(double* scal, float* fscal, uint64_t* err) {
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

// Fixed-order reduction of the dot-product partials into scal[20..23], then the scalar step of the
// iteration on one lane: 1 rho (PCG), 2 rho (flexible CG), 3 curvature (PCG), 4 curvature (flexible CG).
__global__ void k_finalize_step(const double* partial, unsigned slots, double* scal, float* fscal, uint64_t* err,
                                int post, int restart) {
  __shared__ double red[kThreads];
  double a = 0.0;
  for (unsigned i = threadIdx.x; i < slots; i += blockDim.x) a += partial[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x) return;
  scal[20] = red[0];
  if (post == 1) after_rho(scal, fscal, err, restart);
  if (post == 2) after_rho_fcg(scal, fscal, err, restart);
  if (post == 3) after_curv(scal, fscal, err);
  if (post == 4) after_curv_fcg(scal, fscal, err);
}

// One PCG direction step in one pass: d' = z + beta d (d' = z on a restart), q = A d' in the
// edge-difference form, and the block partials of d' . q. The neighbours' d' are formed from their z
// and d the same way, and d' goes to a second buffer, so no lane reads a direction that another lane
// has already replaced.
__global__ void k_direction_apply(uint64_t n, const int32_t* nbr, const float* k, const float* anchor,
                                  const float* diag, const float* z, const float* d, float* dn, float* q,
                                  const float* fscal, int restart, double* partial) {
  const float beta = restart ? 0.f : fscal[1];
  auto direction = [&](uint64_t j) {
    if (restart) return z[j];
    const float prev = beta * d[j];
    return z[j] + prev;
  };
  double acc = 0.0;
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    const float di = direction(i);
    dn[i] = di;
    if (diag[i] == 0.f) {
      q[i] = 0.f;
      continue;
    }
    const uint64_t blk = i >> 6;
    const int l = cnat(static_cast<int>(i & 63));
    const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
    float sum = anchor[i] * di;
    for (int dir = 0; dir < 6; ++dir) {
      const int64_t j = neighbour(nbr, blk, lx, ly, lz, dir);
      if (j < 0) continue;
      sum += face_k(k, n, static_cast<int64_t>(i), j, dir) * (di - direction(static_cast<uint64_t>(j)));
    }
    q[i] = sum;
    acc += static_cast<double>(di) * static_cast<double>(sum);
  }
  double v[4] = {acc, 0.0, 0.0, 0.0};
  block_reduce(v, partial, blockIdx.x);
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
  const uint64_t cn = std::max<uint64_t>({nfine, 64 * ctx->cap.blocks, 64});
  if (nc > ctx->cap.pressure_cells || nf > ctx->cap.pressure_faces) return FL_RESOURCE_EXHAUSTED;
  free_levels(P);
  dfree(P.level_pool);
  P.pool_bytes = 0;
  dfree(P.tile_blocks);
  dfree(P.cell_class);
  P.ntiles = 0;
  void* ptrs[] = {P.fine_cell, P.cell_fine, P.cell_flag, P.cell_faces, P.face_tail, P.face_head, P.face_kind,
                  P.face_slot, P.face_area, P.face_beta, P.face_dist, P.face_ustar, P.face_uwall, P.face_flux,
                  P.face_has_flux, P.face_lift, P.face_scale, P.face_k, P.face_wall, P.face_u, P.face_rec, P.cell_volume,
                  P.cell_s, P.cell_K, P.cell_prev, P.cell_inwall, P.cell_comp, P.cell_target, P.rhs, P.rhs_abs,
                  P.p, P.p_base, P.p_init, P.res, P.def, P.fine_vol, P.r32, P.x32, P.d32, P.dn32, P.q32, P.partial, P.scal,
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
  FL_TRY(dal(&P.face_rec, cf));
  for (double** a : {&P.cell_volume, &P.cell_s, &P.cell_K, &P.cell_prev, &P.cell_inwall, &P.cell_comp, &P.cell_target,
                     &P.rhs, &P.rhs_abs, &P.p, &P.p_base, &P.p_init, &P.res, &P.def})
    FL_TRY(dal(a, cc));
  FL_TRY(dal(&P.fine_cell, cn));
  FL_TRY(dal(&P.fine_vol, cn));
  FL_TRY(dal(&P.cell_class, cn));
  FL_TRY(dal(&P.tile_blocks, 8 * (cn / 64)));
  FL_TRY(dal(&P.r32, cn));
  FL_TRY(dal(&P.x32, cn));
  FL_TRY(dal(&P.d32, cn));
  FL_TRY(dal(&P.dn32, cn));
  FL_TRY(dal(&P.q32, cn));
  FL_TRY(dal(&P.partial, 4 * kPartials));
  FL_TRY(dal(&P.scal, 64));
  FL_TRY(dal(&P.fscal, 16));
  FL_TRY(dal(&P.perr, 4));
  FL_HIP(hipHostMalloc(reinterpret_cast<void**>(&P.hscal), 64 * sizeof(double), hipHostMallocDefault));
  // Level pool: the fine level at its full block capacity plus as much again for the coarser levels
  // (each coarser level has at most as many blocks as the one below; a hierarchy that would overrun
  // the pool stops adding levels), with room for the per-level alignment.
  P.pool_bytes = 2 * level_bytes(cn / 64) + 64 * 16 * kPoolAlign;
  FL_TRY(dal(&P.level_pool, P.pool_bytes));
  P.cap_cells = cc;
  P.cap_faces = cf;
  P.cap_fine = cn;
  return FL_OK;
}

unsigned reduce_grid(uint64_t n, unsigned cap = kStreamPartials) {
  const uint64_t g = (n + kThreads - 1) / kThreads;
  return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(g, cap)));
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

uint64_t pack_key(const Key3& k) {
  return (uint64_t(uint32_t(k.z) ^ 0x80000000u) << 42) ^ (uint64_t(uint32_t(k.y) ^ 0x80000000u) << 21) ^
         uint64_t(uint32_t(k.x) ^ 0x80000000u);
}

// Builds the block hierarchy on the host and uploads neighbour, child and parent tables. Levels are
// added while aggregation still shrinks the set of unknown cells (the oracle's rule), up to
// `levels_wanted`. Level 0 keeps every block that holds a pressure cell.
int build_levels(fl_ctx* ctx, Pressure& P, const std::vector<Key3>& blocks0, int levels_wanted) {
  P.coeff_valid = false;
  P.bf16_valid = false;
  free_levels(P);
  P.levels_requested = levels_wanted;
  std::vector<std::vector<Key3>> keys{blocks0};
  std::vector<Key3> cells = P.h_unknown_cells;
  std::vector<std::vector<int32_t>> parents;  // per level except the coarsest: [nb][2] parent block, octant
  uint64_t planned = level_bytes(blocks0.size());
  if (planned > P.pool_bytes) return FL_RESOURCE_EXHAUSTED;
  while (static_cast<int>(keys.size()) < levels_wanted && !cells.empty()) {
    std::vector<Key3> coarse_cells(cells.size());
    for (size_t i = 0; i < cells.size(); ++i)
      coarse_cells[i] = {floor_half(cells[i].x), floor_half(cells[i].y), floor_half(cells[i].z)};
    std::sort(coarse_cells.begin(), coarse_cells.end(), key_less);
    coarse_cells.erase(std::unique(coarse_cells.begin(), coarse_cells.end()), coarse_cells.end());
    if (coarse_cells.size() == cells.size()) break;
    std::vector<Key3> coarse;
    coarse.reserve(coarse_cells.size() / 8 + 1);
    for (const Key3& c : coarse_cells)
      coarse.push_back({int32_t(floor_div4(c.x)), int32_t(floor_div4(c.y)), int32_t(floor_div4(c.z))});
    std::sort(coarse.begin(), coarse.end(), key_less);
    coarse.erase(std::unique(coarse.begin(), coarse.end()), coarse.end());
    if (planned + level_bytes(coarse.size()) > P.pool_bytes) break;  // the pool bounds the depth
    planned += level_bytes(coarse.size());
    std::unordered_map<Key3, int32_t, Key3Hash> index;
    for (size_t i = 0; i < coarse.size(); ++i) index[coarse[i]] = static_cast<int32_t>(i);
    const auto& fine = keys.back();
    std::vector<int32_t> par(2 * fine.size(), -1);
    for (size_t i = 0; i < fine.size(); ++i) {
      const Key3 pk{floor_half(fine[i].x), floor_half(fine[i].y), floor_half(fine[i].z)};
      auto it = index.find(pk);
      if (it == index.end()) continue;  // a block without unknowns has no coarse parent
      par[2 * i] = it->second;
      par[2 * i + 1] = ((fine[i].z - 2 * pk.z) * 2 + (fine[i].y - 2 * pk.y)) * 2 + (fine[i].x - 2 * pk.x);
    }
    parents.push_back(std::move(par));
    keys.push_back(std::move(coarse));
    cells = std::move(coarse_cells);
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
    L.nbr = carve<int32_t>(P, 6 * L.nb);
    FL_HIP(hipMemcpy(L.nbr, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    L.child = carve<int32_t>(P, 8 * L.nb);
    if (l > 0) {
      std::vector<int32_t> child(8 * L.nb, -1);
      const auto& par = parents[l - 1];
      for (size_t i = 0; i < keys[l - 1].size(); ++i)
        if (par[2 * i] >= 0) child[8 * par[2 * i] + par[2 * i + 1]] = static_cast<int32_t>(i);
      FL_HIP(hipMemcpy(L.child, child.data(), child.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    } else {
      L.child = nullptr;
    }
    L.parent = carve<int32_t>(P, 2 * L.nb);
    if (l + 1 < keys.size()) {
      FL_HIP(hipMemcpy(L.parent, parents[l].data(), parents[l].size() * sizeof(int32_t), hipMemcpyHostToDevice));
    } else {
      L.parent = nullptr;
    }
    L.x = carve<float>(P, n);
    L.r = carve<float>(P, n);
    L.diagb = carve<float>(P, n);
    L.b = carve<float>(P, n);
    L.anchor = carve<float>(P, n);
    L.diag = carve<float>(P, n);
    L.kb = carve<uint16_t>(P, 3 * n);
    L.k = carve<float>(P, 3 * n);
    if (P.pool_used > P.pool_bytes) return FL_INTERNAL;
  }
  (void)pack_key;
  return FL_OK;
}

// Level coefficients from the links' conductances (or from `conductance` when given: the coupled
// viscous solve's effective conductances).
int setup_coefficients(fl_ctx* ctx, Pressure& P, const double* conductance = nullptr) {
  hipStream_t s = ctx->stream;
  P.coeff_valid = false;
  P.bf16_valid = false;
  Level& L0 = P.levels[0];
  const uint64_t n0 = L0.nb * 64;
  k_level0<<<grid_for(n0), kThreads, 0, s>>>(n0, P.fine_cell, P.cell_flag, P.cell_faces,
                                             conductance ? conductance : P.face_k, P.face_tail,
                                             P.face_head, P.cell_comp,
                                             P.cell_volume, L0.k, L0.anchor, L0.diag, P.fine_vol);
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

template <class KT>
LevelViewT<KT> view(Pressure& P, size_t l);
template <>
LevelViewT<float> view<float>(Pressure& P, size_t l) {
  Level& L = P.levels[l];
  return {L.nb * 64, L.nbr, L.child, L.parent, L.x, l == 0 ? P.r32 : L.b, L.k, L.anchor, L.diag};
}
template <>
LevelViewT<Bf16> view<Bf16>(Pressure& P, size_t l) {
  Level& L = P.levels[l];
  return {L.nb * 64,           L.nbr,  L.child, L.parent, L.x, l == 0 ? P.r32 : L.b, reinterpret_cast<const Bf16*>(L.kb),
          L.anchor,            L.diagb};
}

// V-cycle from level l. Level 0 smooths the recursive FP32 residual of the outer PCG in place of its
// own right-hand side. Levels of at most kTailBlocks blocks (and everything below them) run in one
// single-workgroup launch.
template <class KT>
int vcycle(fl_ctx* ctx, Pressure& P, size_t l, int nu, int coarse_sweeps, int smoother) {
  hipStream_t s = ctx->stream;
  const size_t last = P.levels.size() - 1;
  if (l == last && P.levels[l].nb > kTailBlocks) {  // a large coarsest level: grid-wide sweeps from zero
    const LevelViewT<KT> B = view<KT>(P, l);
    for (int pass = 0; pass < 4 * coarse_sweeps; ++pass) {
      const int fwd = pass < 2 * coarse_sweeps;
      k_smooth<<<grid_for(B.n / 2), kThreads, 0, s>>>(B, fwd ? (pass & 1) : 1 - (pass & 1), pass == 0);
    }
    FL_HIP(hipGetLastError());
    return FL_OK;
  }
  if (P.levels[l].nb <= kTailBlocks) {
    TailCycle<KT> T{};
    T.count = static_cast<int>(last - l + 1);
    if (T.count > kMaxTail) return FL_INTERNAL;
    T.nu = nu;
    T.coarse_sweeps = coarse_sweeps;
    for (size_t q = l; q <= last; ++q) T.lv[q - l] = view<KT>(P, q);
    k_tail_vcycle<<<1, 1024, 0, s>>>(T);
    FL_HIP(hipGetLastError());
    return FL_OK;
  }
  if (smoother == 1 || (smoother == 2 && l > 0)) {  // block smoother: pre-smooth, restrict, coarse cycle, post
    const LevelViewT<KT> L = view<KT>(P, l), C = view<KT>(P, l + 1);
    const unsigned gb = static_cast<unsigned>((P.levels[l].nb + 3) / 4);
    k_block_smooth<KT, false><<<gb, 256, 0, s>>>(L, nullptr, nullptr, nu);
    k_residual_all<<<grid_for(L.n), kThreads, 0, s>>>(L, P.levels[l].r);
    k_restrict_all<<<grid_for(C.n), kThreads, 0, s>>>(P.levels[l].r, C);
    FL_HIP(hipGetLastError());
    FL_TRY(vcycle<KT>(ctx, P, l + 1, nu, coarse_sweeps, smoother));
    const LevelViewT<KT> Lp = view<KT>(P, l);
    k_block_smooth<KT, true><<<gb, 256, 0, s>>>(Lp, P.levels[l].r, P.levels[l + 1].x, nu);
    FL_HIP(hipGetLastError());
    std::swap(P.levels[l].x, P.levels[l].r);  // the post-smoothed values are the level's x from here on
    return FL_OK;
  }
  const LevelViewT<KT> L = view<KT>(P, l), C = view<KT>(P, l + 1);
  const unsigned gh = grid_for(L.n / 2);
  for (int it = 0; it < nu; ++it) {
    k_smooth<<<gh, kThreads, 0, s>>>(L, 0, it == 0);
    k_smooth<<<gh, kThreads, 0, s>>>(L, 1, 0);
  }
  k_residual_red<<<gh, kThreads, 0, s>>>(L, P.levels[l].r);
  k_restrict_red<<<grid_for(C.n), kThreads, 0, s>>>(P.levels[l].r, C);
  FL_HIP(hipGetLastError());
  FL_TRY(vcycle<KT>(ctx, P, l + 1, nu, coarse_sweeps, smoother));
  k_prolong<<<grid_for(L.n), kThreads, 0, s>>>(L, P.levels[l + 1].x);  // the coarse level may have swapped x
  for (int it = 0; it < nu; ++it) {
    k_smooth<<<gh, kThreads, 0, s>>>(L, 1, 0);
    k_smooth<<<gh, kThreads, 0, s>>>(L, 0, 0);
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

// Reduces the evaluation partials of `groups` workgroups and reads the norms and error words back.
int read_evaluation(fl_ctx* ctx, Pressure& P, unsigned g, EvalResult& out) {
  hipStream_t s = ctx->stream;
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

int evaluate(fl_ctx* ctx, Pressure& P, const double* p, EvalResult& out) {
  const unsigned g = reduce_grid(P.nc, kPartials);
  k_evaluate<<<g, kThreads, 0, ctx->stream>>>(P.nc, P.cell_faces, P.face_rec, P.cell_volume, P.cell_target,
                                              P.cell_inwall, P.cell_comp, P.cell_prev, P.rhs, P.rhs_abs, p, P.res,
                                              P.def, P.partial, P.perr);
  return read_evaluation(ctx, P, g, out);
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
int solve(fl_ctx* ctx, Pressure& P, const fl_pressure_options& opt, const Gates& gates, const EvalResult& zero_eval,
          fl_pressure_diagnostics& diag) {
  hipStream_t s = ctx->stream;
  Level& L0 = P.levels[0];
  const uint64_t n = L0.nb * 64;
  const unsigned g = grid_for(n);
  const unsigned rg = reduce_grid(n);
  diag.multigrid_level_count = P.levels.size();
  diag.iteration_count = 0;
  P.last_storage = opt.level_storage;
  P.last_nu = opt.smoothing_sweeps;
  P.last_coarse_sweeps = opt.coarse_sweeps;
  diag.restarts = 0;
  EvalResult e{};
  k_copy_f64<<<grid_for(P.nc), kThreads, 0, s>>>(P.nc, P.p_base, P.p);
  if (P.p_init_zero) {
    e = zero_eval;  // P.res and P.def still hold the p = 0 evaluation
  } else {
    FL_TRY(evaluate(ctx, P, P.p, e));
  }
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
  k_load_residual<<<g, kThreads, 0, s>>>(n, P.fine_cell, P.cell_flag, P.res, P.r32, P.x32, P.perr);
  if (opt.level_storage == 2 && !P.bf16_valid) {
    for (auto& L : P.levels) {
      const uint64_t nl = L.nb * 64;
      k_to_bf16<<<grid_for(3 * nl), kThreads, 0, s>>>(L.k, 3 * nl, L.kb);
      k_diag_bf16<<<grid_for(nl), kThreads, 0, s>>>(nl, L.nbr, L.kb, L.anchor, L.diag, L.diagb);
    }
    FL_HIP(hipGetLastError());
    P.bf16_valid = true;
  }
  bool restart = true;
  int since = 0;
  // Learned preconditioner: flexible CG with the network until 3 consecutive iterations fail to reduce
  // the recursive residual by 10 %, then the V-cycle from the current iterate (restart).
  LearnState* net = static_cast<LearnState*>(ctx->learn);
  bool use_net = opt.preconditioner == 1;
  if (use_net && (!net || P.ntiles > net->tile_cap)) return FL_FAILED_PRECONDITION;
  int stagnant = 0;
  double rn_prev = std::numeric_limits<double>::infinity();
  diag.learned_iterations = 0;
  for (uint64_t it = 1; it <= opt.max_iterations; ++it) {
    if (use_net) {
      k_tile_gather<<<static_cast<unsigned>(P.ntiles), 512, 0, s>>>(P.r32, L0.diag, P.cell_class, P.tile_blocks,
                                                                    net->act, net->scale);
      FL_TRY(net_infer(s, net->params, net->frags, net->act, P.ntiles));
      k_tile_scatter<<<grid_for(P.ntiles * 512), kThreads, 0, s>>>(net->act, net->scale, P.cell_class, P.tile_blocks,
                                                                    P.ntiles, L0.x);
      ++diag.learned_iterations;
    } else if (opt.level_storage == 2) {
      FL_TRY(vcycle<Bf16>(ctx, P, 0, opt.smoothing_sweeps, opt.coarse_sweeps, opt.smoother));
    } else {
      FL_TRY(vcycle<float>(ctx, P, 0, opt.smoothing_sweeps, opt.coarse_sweeps, opt.smoother));
    }
    if (!use_net && net && net->collect_every > 0 && it % uint64_t(net->collect_every) == 0 &&
        P.ntiles <= net->tile_cap && net->ndata + P.ntiles <= net->data_cap) {
      // Training pairs: this iteration's residual and its V-cycle correction, per tile.
      k_tile_gather<<<static_cast<unsigned>(P.ntiles), 512, 0, s>>>(P.r32, L0.diag, P.cell_class, P.tile_blocks,
                                                                    net->act, net->scale);
      k_tile_pairs<<<grid_for(P.ntiles * 512), kThreads, 0, s>>>(net->act, net->scale, L0.x, P.cell_class,
                                                                  P.tile_blocks, P.ntiles,
                                                                  net->data_in + net->ndata * 2048,
                                                                  net->data_target + net->ndata * 512);
      net->ndata += P.ntiles;
    }
    if (use_net && !restart) {  // z . q_prev for the flexible-CG orthogonalisation, before q is overwritten
      k_dot<<<rg, kThreads, 0, s>>>(n, L0.x, P.q32, P.partial);
      k_finalize4<<<1, kThreads, 0, s>>>(P.partial, rg, P.scal, 28);
    }
    k_dot<<<rg, kThreads, 0, s>>>(n, P.r32, L0.x, P.partial);
    k_finalize_step<<<1, kThreads, 0, s>>>(P.partial, rg, P.scal, P.fscal, P.perr, use_net ? 2 : 1, restart ? 1 : 0);
    k_direction_apply<<<rg, kThreads, 0, s>>>(n, L0.nbr, L0.k, L0.anchor, L0.diag, L0.x, P.d32, P.dn32, P.q32,
                                              P.fscal, restart ? 1 : 0, P.partial);
    std::swap(P.d32, P.dn32);
    k_finalize_step<<<1, kThreads, 0, s>>>(P.partial, rg, P.scal, P.fscal, P.perr, use_net ? 4 : 3, 0);
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
    if (use_net) {
      stagnant = (rn > 0.9 * rn_prev || !(P.hscal[0] > 0.0)) ? stagnant + 1 : 0;
      if (stagnant >= 3 || P.hscal[14] != 0.0) {
        use_net = false;
        restart = true;
      }
    }
    rn_prev = rn;
    bool recursive = rn <= gates.integrated && rmax <= gates.volume;
    if (gates.relative > 0.0) recursive = recursive && dn <= gates.relative * gates.initial_defect;
    const bool last = it == opt.max_iterations;
    if (!last && !recursive && since >= 10) {
      // Periodic restart from the FP64 true residual; acceptance is only tested at recursive
      // convergence or on the last iteration.
      k_apply_correction<<<grid_for(P.nc), kThreads, 0, s>>>(P.nc, P.cell_fine, P.cell_flag, P.p_base, P.x32, P.p,
                                                             P.perr);
      k_true_residual<<<g, kThreads, 0, s>>>(n, P.fine_cell, P.cell_faces, P.face_k, P.face_tail, P.face_head,
                                              P.cell_comp, P.rhs, P.p, P.res);
      k_copy_f64<<<grid_for(P.nc), kThreads, 0, s>>>(P.nc, P.p, P.p_base);
      k_load_residual<<<g, kThreads, 0, s>>>(n, P.fine_cell, P.cell_flag, P.res, P.r32, P.x32, P.perr);
      FL_HIP(hipGetLastError());
      restart = true;
      since = 0;
      ++diag.restarts;
      continue;
    }
    if (last || recursive) {
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
  dfree(P->level_pool);
  dfree(P->tile_blocks);
  dfree(P->cell_class);
  void* ptrs[] = {P->fine_cell, P->cell_fine, P->cell_flag, P->cell_faces, P->face_tail, P->face_head, P->face_kind,
                  P->face_slot, P->face_area, P->face_beta, P->face_dist, P->face_ustar, P->face_uwall, P->face_flux,
                  P->face_has_flux, P->face_lift, P->face_scale, P->face_k, P->face_wall, P->face_u, P->face_rec, P->cell_volume,
                  P->cell_s, P->cell_K, P->cell_prev, P->cell_inwall, P->cell_comp, P->cell_target, P->rhs,
                  P->rhs_abs, P->p, P->p_base, P->p_init, P->res, P->def, P->fine_vol, P->r32, P->x32, P->d32, P->dn32, P->q32,
                  P->partial, P->scal, P->fscal, P->perr};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (P->hscal) (void)hipHostFree(P->hscal);
  void* dv[] = {P->dv_mark, P->dv_slot_face, P->dv_parent, P->dv_root, P->dv_anchor, P->dv_words, P->dv_lut,
                P->dv_occ[0], P->dv_occ[1], P->dv_keys};
  for (void* p : dv)
    if (p) (void)hipFree(p);
  delete P;
  ctx->pressure = nullptr;
}

namespace {

// Mean device time of `reps` back-to-back launches of `launch` on the context stream.
template <class Launch>
int mean_ms(fl_ctx* ctx, int reps, Launch&& launch, double* ms) {
  hipStream_t s = ctx->stream;
  launch();  // warm-up
  FL_HIP(hipGetLastError());
  FL_HIP(hipEventRecord(ctx->stage[6], s));
  for (int i = 0; i < reps; ++i) launch();
  FL_HIP(hipEventRecord(ctx->stage[7], s));
  FL_HIP(hipEventSynchronize(ctx->stage[7]));
  float t = 0;
  FL_HIP(hipEventElapsedTime(&t, ctx->stage[6], ctx->stage[7]));
  *ms = static_cast<double>(t) / reps;
  return FL_OK;
}

template <class KT>
int time_levels(fl_ctx* ctx, Pressure& P, int reps, fl_kernel_time* out, size_t capacity, size_t* count) {
  hipStream_t s = ctx->stream;
  size_t k = 0;
  auto put = [&](uint32_t kernel, uint32_t level, uint64_t slots, double ms) {
    if (k < capacity) out[k] = fl_kernel_time{kernel, level, slots, ms};
    ++k;
  };
  double ms = 0;
  Level& L0 = P.levels[0];
  const uint64_t n0 = L0.nb * 64;
  const LevelViewT<KT> V0 = view<KT>(P, 0);
  const int nu = std::max(P.last_nu, 1);
  // Level-0 V-cycle kernels.
  FL_TRY(mean_ms(ctx, reps, [&] { k_smooth<KT><<<grid_for(n0 / 2), kThreads, 0, s>>>(V0, 0, 0); }, &ms));
  put(FL_KERNEL_SMOOTH, 0, n0, ms);
  FL_TRY(mean_ms(ctx, reps, [&] { k_residual_red<KT><<<grid_for(n0 / 2), kThreads, 0, s>>>(V0, L0.r); }, &ms));
  put(FL_KERNEL_RESIDUAL_RED, 0, n0, ms);
  if (P.levels.size() > 1) {
    const LevelViewT<KT> V1 = view<KT>(P, 1);
    FL_TRY(mean_ms(ctx, reps, [&] { k_restrict_red<KT><<<grid_for(V1.n), kThreads, 0, s>>>(L0.r, V1); }, &ms));
    put(FL_KERNEL_RESTRICT_RED, 0, n0, ms);
    FL_TRY(mean_ms(ctx, reps, [&] { k_prolong<KT><<<grid_for(n0), kThreads, 0, s>>>(V0, P.levels[1].x); }, &ms));
    put(FL_KERNEL_PROLONG, 0, n0, ms);
    if (P.levels[1].nb > kTailBlocks && P.levels.size() > 2) {
      const LevelViewT<KT> V2 = view<KT>(P, 2);
      const unsigned gb = static_cast<unsigned>((P.levels[1].nb + 3) / 4);
      FL_TRY(mean_ms(ctx, reps, [&] { k_block_smooth<KT, false><<<gb, 256, 0, s>>>(V1, nullptr, nullptr, nu); }, &ms));
      put(FL_KERNEL_BLOCK_SMOOTH_PRE, 1, V1.n, ms);
      FL_TRY(mean_ms(ctx, reps,
                     [&] { k_block_smooth<KT, true><<<gb, 256, 0, s>>>(V1, P.levels[1].r, P.levels[2].x, nu); },
                     &ms));
      put(FL_KERNEL_BLOCK_SMOOTH_POST, 1, V1.n, ms);
      FL_TRY(mean_ms(ctx, reps, [&] { k_residual_all<KT><<<grid_for(V1.n), kThreads, 0, s>>>(V1, P.levels[1].r); },
                     &ms));
      put(FL_KERNEL_RESIDUAL_ALL, 1, V1.n, ms);
      FL_TRY(mean_ms(ctx, reps, [&] { k_restrict_all<KT><<<grid_for(V2.n), kThreads, 0, s>>>(P.levels[1].r, V2); },
                     &ms));
      put(FL_KERNEL_RESTRICT_ALL, 1, V1.n, ms);
    }
  }
  // The single-workgroup tail from the first level of at most kTailBlocks blocks.
  size_t first = P.levels.size();
  for (size_t l = 0; l < P.levels.size(); ++l)
    if (P.levels[l].nb <= kTailBlocks) {
      first = l;
      break;
    }
  if (first < P.levels.size() && P.levels.size() - first <= static_cast<size_t>(kMaxTail)) {
    TailCycle<KT> T{};
    T.count = static_cast<int>(P.levels.size() - first);
    T.nu = nu;
    T.coarse_sweeps = P.last_coarse_sweeps;
    for (size_t q = first; q < P.levels.size(); ++q) T.lv[q - first] = view<KT>(P, q);
    FL_TRY(mean_ms(ctx, reps, [&] { k_tail_vcycle<<<1, 1024, 0, s>>>(T); }, &ms));
    put(FL_KERNEL_TAIL_VCYCLE, static_cast<uint32_t>(first), P.levels[first].nb * 64, ms);
  }
  // PCG kernels on level 0 (FP32 vectors and coefficients).
  const unsigned rg = reduce_grid(n0);
  FL_TRY(mean_ms(ctx, reps, [&] {
    k_direction_apply<<<rg, kThreads, 0, s>>>(n0, L0.nbr, L0.k, L0.anchor, L0.diag, L0.x, P.d32, P.dn32, P.q32,
                                              P.fscal, 0, P.partial);
  }, &ms));
  put(FL_KERNEL_DIRECTION_APPLY, 0, n0, ms);
  FL_TRY(mean_ms(ctx, reps, [&] {
    k_update_xr<<<rg, kThreads, 0, s>>>(n0, P.d32, P.q32, L0.diag, P.fine_vol, P.x32, P.r32, P.fscal, P.partial,
                                        P.perr);
  }, &ms));
  put(FL_KERNEL_UPDATE_XR, 0, n0, ms);
  FL_TRY(mean_ms(ctx, reps, [&] { k_dot<<<rg, kThreads, 0, s>>>(n0, P.r32, L0.x, P.partial); }, &ms));
  put(FL_KERNEL_DOT, 0, n0, ms);
  // The FP64 evaluation of the accepted pressure (recomputes the same residuals and defects).
  const unsigned g = reduce_grid(P.nc, kPartials);
  FL_TRY(mean_ms(ctx, reps, [&] {
    k_evaluate<<<g, kThreads, 0, s>>>(P.nc, P.cell_faces, P.face_rec, P.cell_volume, P.cell_target, P.cell_inwall,
                                      P.cell_comp, P.cell_prev, P.rhs, P.rhs_abs, P.p, P.res, P.def, P.partial,
                                      P.perr);
  }, &ms));
  put(FL_KERNEL_EVALUATE, 0, P.nc, ms);
  *count = k;
  return k <= capacity ? FL_OK : FL_RESOURCE_EXHAUSTED;
}

}  // namespace

int time_projection_kernels(fl_ctx* ctx, int reps, fl_kernel_time* out, size_t capacity, size_t* count) {
  Pressure* P = state(ctx);
  if (!P || P->levels.empty() || P->nc == 0) return FL_FAILED_PRECONDITION;
  (void)hipSetDevice(ctx->device);
  if (P->last_storage == 2) return time_levels<Bf16>(ctx, *P, reps, out, capacity, count);
  return time_levels<float>(ctx, *P, reps, out, capacity, count);
}

int reserve_pressure(fl_ctx* ctx) {
  if (ctx->cap.pressure_cells == 0 || ctx->cap.pressure_faces == 0) return FL_OK;
  return ensure_capacity(ctx, ctx->cap.pressure_cells, ctx->cap.pressure_faces, 64 * ctx->cap.blocks);
}

}  // namespace fl

namespace fl {
namespace {

bool representable(double a, double b, double p) {
  if (!std::isfinite(p)) return false;
  if (a == 0.0 || b == 0.0 || std::fabs(p) >= std::numeric_limits<double>::min()) return true;
  if (p == 0.0) return false;
  int ea = 0, eb = 0;
  const double ma = std::frexp(a, &ea), mb = std::frexp(b, &eb);
  return std::fma(ma, mb, -std::scalbn(p, -(ea + eb))) == 0.0;
}
bool finite_positive(double v) { return std::isfinite(v) && v > 0.0; }

struct HostComp {
  double sum = 0.0, corr = 0.0;
  void add(double t) {
    const double n = sum + t;
    if (std::fabs(sum) >= std::fabs(t)) corr += (sum - n) + t; else corr += (t - n) + sum;
    sum = n;
  }
  double total() const { return sum + corr; }
};

int64_t host_floor_div4(int64_t i) { return i >= 0 ? i / 4 : -1 - (-(i + 1)) / 4; }

// Validation and structure of one pressure system, in the order of Prepare
// (mac_pressure_workspace_3d.cc:68-166) and SelectGaugesInto (mac_pressure_3d_arithmetic.h:328-399).
// `check_values` also applies the per-value checks of FreezeLink and the RHS assembly to u*.
int prepare_system(fl_ctx* ctx, Pressure& P, const fl_geometry* g, const fl_int3* grid_blocks, size_t nb, double dt,
                   const fl_pressure_cell* cells, size_t nc, const fl_pressure_face* faces, size_t nf,
                   bool check_values) {
  P.coeff_valid = false;
  P.bf16_valid = false;
  if (nc == 0 || nf == 0 || !finite_positive(dt)) return FL_INVALID_ARGUMENT;
  const double h = g->cell_width, area = h * h, volume = area * h;
  if (!finite_positive(area) || !finite_positive(volume)) return FL_OUT_OF_RANGE;
  if (nc > ctx->cap.pressure_cells || nf > ctx->cap.pressure_faces) return FL_RESOURCE_EXHAUSTED;
  if (nc > 0x7fffffffULL || nf > 0x7fffffffULL) return FL_RESOURCE_EXHAUSTED;
  P.full_area = area;
  P.full_volume = volume;
  std::unordered_map<Key3, int32_t, Key3Hash> grid;
  grid.reserve(nb * 2);
  for (size_t i = 0; i < nb; ++i) grid[{grid_blocks[i].x, grid_blocks[i].y, grid_blocks[i].z}] = int32_t(i);
  auto slot_of = [&](int axis, int64_t x, int64_t y, int64_t z, uint64_t& slot) -> int {
    if (axis < 0 || axis > 2) return FL_INVALID_ARGUMENT;
    const Key3 o{int32_t(host_floor_div4(x)), int32_t(host_floor_div4(y)), int32_t(host_floor_div4(z))};
    auto it = grid.find(o);
    if (it == grid.end()) return FL_FAILED_PRECONDITION;
    slot = uint64_t(it->second) * 192 + uint64_t(axis) * 64 +
           uint64_t(((z - 4 * int64_t(o.z)) * 4 + (y - 4 * int64_t(o.y))) * 4 + (x - 4 * int64_t(o.x)));
    return FL_OK;
  };
  // cells
  std::vector<double> comp(nc), target(nc);
  std::vector<uint64_t> incident;
  incident.reserve(6 * nc);
  std::unordered_map<Key3, int32_t, Key3Hash> cell_index;
  cell_index.reserve(nc * 2);
  for (size_t i = 0; i < nc; ++i) {
    const fl_pressure_cell& c = cells[i];
    if (i) {
      const fl_int3 a = cells[i - 1].cell_index, b = c.cell_index;
      const bool less = a.z != b.z ? a.z < b.z : (a.y != b.y ? a.y < b.y : a.x < b.x);
      if (!less) return FL_INVALID_ARGUMENT;
    }
    if (!finite_positive(c.fluid_volume) || c.fluid_volume > volume) return FL_INVALID_ARGUMENT;
    if (!std::isfinite(c.target_divergence) || !std::isfinite(c.previous_pressure) ||
        !std::isfinite(c.integrated_inward_wall_flux))
      return FL_INVALID_ARGUMENT;
    double cmp = 0.0;
    if (c.bulk_modulus != std::numeric_limits<double>::infinity()) {
      if (!finite_positive(c.bulk_modulus)) return FL_INVALID_ARGUMENT;
      const double kdt = c.bulk_modulus * dt;
      if (!finite_positive(kdt)) return FL_OUT_OF_RANGE;
      cmp = c.fluid_volume / kdt;
      if (!finite_positive(cmp)) return FL_OUT_OF_RANGE;
    }
    const double tf = c.fluid_volume * c.target_divergence;
    if (!representable(c.fluid_volume, c.target_divergence, tf)) return FL_OUT_OF_RANGE;
    comp[i] = cmp;
    target[i] = tf;
    const int64_t xyz[3] = {c.cell_index.x, c.cell_index.y, c.cell_index.z};
    for (int a = 0; a < 3; ++a) {
      if (xyz[a] == 2147483647LL) return FL_OUT_OF_RANGE;
      for (int side = 0; side < 2; ++side) {
        int64_t l[3] = {xyz[0], xyz[1], xyz[2]};
        l[a] += side;
        uint64_t slot;
        FL_TRY(slot_of(a, l[0], l[1], l[2], slot));
        incident.push_back(slot);
      }
    }
    cell_index[{c.cell_index.x, c.cell_index.y, c.cell_index.z}] = int32_t(i);
  }
  std::sort(incident.begin(), incident.end());
  incident.erase(std::unique(incident.begin(), incident.end()), incident.end());
  if (incident.size() != nf) return FL_INVALID_ARGUMENT;
  // faces
  std::vector<int32_t> tail(nf), head(nf), kind(nf);
  std::vector<double> kcond(nf), wall(nf);
  std::vector<uint64_t> slots(nf);
  std::unordered_map<uint64_t, int32_t> face_of_slot;
  face_of_slot.reserve(nf * 2);
  bool has_jump = false;
  for (size_t i = 0; i < nf; ++i) {
    const fl_pressure_face& f = faces[i];
    uint64_t slot;
    FL_TRY(slot_of(f.axis, f.lattice.x, f.lattice.y, f.lattice.z, slot));
    if (slot != incident[i]) return FL_INVALID_ARGUMENT;
    slots[i] = slot;
    face_of_slot[slot] = int32_t(i);
    int64_t t[3] = {f.lattice.x, f.lattice.y, f.lattice.z};
    t[f.axis] -= 1;
    int32_t to = -1;
    if (t[f.axis] >= -2147483648LL) {
      auto it = cell_index.find({int32_t(t[0]), int32_t(t[1]), int32_t(t[2])});
      if (it != cell_index.end()) to = it->second;
    }
    auto ih = cell_index.find({f.lattice.x, f.lattice.y, f.lattice.z});
    const int32_t ho = ih == cell_index.end() ? -1 : ih->second;
    // FreezeLinkFromIncidence (mac_pressure_3d_arithmetic.h:174-245)
    if (!std::isfinite(f.open_area) || f.open_area < 0.0 || f.open_area > area) return FL_INVALID_ARGUMENT;
    if (!std::isfinite(f.inverse_density) || !std::isfinite(f.pressure_gradient_distance) ||
        !std::isfinite(f.predicted_normal_velocity) || !std::isfinite(f.wall_normal_velocity))
      return FL_INVALID_ARGUMENT;
    const bool both = to >= 0 && ho >= 0, one = (to >= 0) != (ho >= 0);
    if (f.link_kind == 1) {
      if (!both || !std::isfinite(f.lift)) return FL_INVALID_ARGUMENT;
    } else if (f.link_kind == 2) {
      if (!one || f.open_area == 0.0 || !std::isfinite(f.lift)) return FL_INVALID_ARGUMENT;
    } else if (f.link_kind == 0) {
      if (f.open_area != 0.0) return FL_INVALID_ARGUMENT;
    } else {
      return FL_INVALID_ARGUMENT;
    }
    double scale = 0.0, cond = 0.0;
    if (f.open_area > 0.0) {
      if (f.inverse_density <= 0.0 || f.pressure_gradient_distance <= 0.0) return FL_INVALID_ARGUMENT;
      scale = (dt * f.inverse_density) / f.pressure_gradient_distance;
      cond = f.open_area * scale;
      if (!finite_positive(scale) || !finite_positive(cond)) return FL_OUT_OF_RANGE;
    }
    const double blocked = area - f.open_area;
    double wf;
    if (f.has_prescribed_wall_flux) {
      wf = f.prescribed_wall_flux;
      if (!std::isfinite(wf) || (blocked == 0 && wf != 0)) return FL_INVALID_ARGUMENT;
    } else {
      wf = blocked * f.wall_normal_velocity;
      if (!representable(blocked, f.wall_normal_velocity, wf)) return FL_OUT_OF_RANGE;
    }
    tail[i] = to;
    head[i] = ho;
    kind[i] = f.link_kind;
    kcond[i] = cond;
    wall[i] = wf;
    has_jump |= (cond > 0.0 && f.link_kind == 1 && f.lift != 0.0);
  }
  // RHS representability (AssembleIntegratedRhsInto, mac_pressure_3d_arithmetic.h:263-313)
  if (check_values) {
    for (size_t i = 0; i < nc; ++i)
      if (!representable(comp[i], cells[i].previous_pressure, comp[i] * cells[i].previous_pressure))
        return FL_OUT_OF_RANGE;
    for (size_t i = 0; i < nf; ++i) {
      const fl_pressure_face& f = faces[i];
      if (!representable(f.open_area, f.predicted_normal_velocity, f.open_area * f.predicted_normal_velocity))
        return FL_OUT_OF_RANGE;
      if ((kind[i] == 1 || kind[i] == 2) && !representable(kcond[i], f.lift, kcond[i] * f.lift)) return FL_OUT_OF_RANGE;
    }
  }
  // gauges: components over positive internal links, anchored by compliance or a Dirichlet link
  std::vector<int32_t> parent(nc);
  std::iota(parent.begin(), parent.end(), 0);
  auto root = [&](int32_t c) {
    while (parent[c] != c) c = parent[c];
    return c;
  };
  for (size_t i = 0; i < nf; ++i)
    if (kcond[i] > 0.0 && tail[i] >= 0 && head[i] >= 0) {
      const int32_t a = root(tail[i]), b = root(head[i]);
      parent[std::max(a, b)] = std::min(a, b);
    }
  std::vector<char> anchored(nc, 0);
  std::vector<int32_t> component(nc);
  for (size_t c = 0; c < nc; ++c) {
    component[c] = root(int32_t(c));
    if (comp[c] > 0.0) anchored[component[c]] = 1;
  }
  for (size_t i = 0; i < nf; ++i)
    if (kind[i] == 2) {
      if (tail[i] >= 0) anchored[component[tail[i]]] = 1;
      if (head[i] >= 0) anchored[component[head[i]]] = 1;
    }
  std::vector<HostComp> bal(nc), babs(nc);
  for (size_t c = 0; c < nc; ++c)
    if (!anchored[component[c]]) {
      for (double t : {target[c], cells[c].integrated_inward_wall_flux}) {
        bal[component[c]].add(t);
        babs[component[c]].add(std::fabs(t));
      }
    }
  for (size_t i = 0; i < nf; ++i) {
    const bool stays = tail[i] >= 0 && head[i] >= 0 && component[tail[i]] == component[head[i]];
    if (stays) continue;
    for (int side = 0; side < 2; ++side) {
      const int32_t c = side == 0 ? tail[i] : head[i];
      if (c < 0 || anchored[component[c]]) continue;
      const double sgn = side == 0 ? 1.0 : -1.0;
      bal[component[c]].add(-sgn * wall[i]);
      babs[component[c]].add(std::fabs(-sgn * wall[i]));
    }
  }
  std::vector<uint8_t> flag(nc, 1);
  P.h_gauges.clear();
  for (size_t c = 0; c < nc; ++c) {
    if (component[c] != int32_t(c) || anchored[c]) continue;
    const double b = bal[c].total(), a = babs[c].total();
    if (!std::isfinite(b) || !std::isfinite(a)) return FL_OUT_OF_RANGE;
    if (std::fabs(b) > 32.0 * 2.220446049250313e-16 * a) return FL_FAILED_PRECONDITION;
    P.h_gauges.push_back(c);
    flag[c] = 2;
  }
  P.h_cell_comp_root = component;
  P.has_jump = has_jump;
  P.h_unknown_cells.clear();
  for (size_t c = 0; c < nc; ++c)
    if (flag[c] == 1) P.h_unknown_cells.push_back({cells[c].cell_index.x, cells[c].cell_index.y, cells[c].cell_index.z});
  // block layout of the fine level
  std::vector<Key3> blocks0;
  blocks0.reserve(nc / 8 + 1);
  for (size_t c = 0; c < nc; ++c) {
    const fl_int3 ci = cells[c].cell_index;
    blocks0.push_back({int32_t(host_floor_div4(ci.x)), int32_t(host_floor_div4(ci.y)), int32_t(host_floor_div4(ci.z))});
  }
  std::sort(blocks0.begin(), blocks0.end(), key_less);
  blocks0.erase(std::unique(blocks0.begin(), blocks0.end()), blocks0.end());
  std::unordered_map<Key3, int32_t, Key3Hash> bidx;
  bidx.reserve(blocks0.size() * 2);
  for (size_t i = 0; i < blocks0.size(); ++i) bidx[blocks0[i]] = int32_t(i);
  const uint64_t nfine = blocks0.size() * 64;
  FL_TRY(ensure_capacity(ctx, nc, nf, nfine));
  std::vector<int32_t> cell_fine(nc), fine_cell(nfine, -1), cell_faces(6 * nc);
  for (size_t c = 0; c < nc; ++c) {
    const fl_int3 ci = cells[c].cell_index;
    const Key3 b{int32_t(host_floor_div4(ci.x)), int32_t(host_floor_div4(ci.y)), int32_t(host_floor_div4(ci.z))};
    const int64_t fi = int64_t(bidx[b]) * 64 + cslot((ci.z - 4 * b.z) * 16 + (ci.y - 4 * b.y) * 4 + (ci.x - 4 * b.x));
    cell_fine[c] = int32_t(fi);
    fine_cell[fi] = int32_t(c);
    const int64_t xyz[3] = {ci.x, ci.y, ci.z};
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side) {
        int64_t l[3] = {xyz[0], xyz[1], xyz[2]};
        l[a] += side;
        uint64_t slot;
        slot_of(a, l[0], l[1], l[2], slot);
        cell_faces[6 * c + 2 * a + side] = face_of_slot[slot];
      }
  }
  P.nc = nc;
  P.nf = nf;
  hipStream_t s = ctx->stream;
  std::vector<double> V(nc), S(nc), K(nc), prev(nc), inw(nc);
  for (size_t c = 0; c < nc; ++c) {
    V[c] = cells[c].fluid_volume;
    S[c] = cells[c].target_divergence;
    K[c] = cells[c].bulk_modulus;
    prev[c] = cells[c].previous_pressure;
    inw[c] = cells[c].integrated_inward_wall_flux;
  }
  // Warm start (InitializeMacPressure3D, mac_pressure_3d_multigrid.cc:303-321).
  std::vector<double> pin(nc, 0.0), p0(nc);
  for (uint64_t gg : P.h_gauges) pin[component[gg]] = prev[gg];
  for (size_t c = 0; c < nc; ++c) {
    p0[c] = prev[c] - pin[component[c]];
    if (!std::isfinite(p0[c])) return FL_OUT_OF_RANGE;
  }
  std::vector<double> fa(nf), fb(nf), fd(nf), fu(nf), fw(nf), ff(nf), fl(nf);
  std::vector<int32_t> fh(nf);
  for (size_t i = 0; i < nf; ++i) {
    fa[i] = faces[i].open_area;
    fb[i] = faces[i].inverse_density;
    fd[i] = faces[i].pressure_gradient_distance;
    fu[i] = faces[i].predicted_normal_velocity;
    fw[i] = faces[i].wall_normal_velocity;
    ff[i] = faces[i].prescribed_wall_flux;
    fh[i] = faces[i].has_prescribed_wall_flux;
    fl[i] = kind[i] == 0 ? 0.0 : faces[i].lift;
  }
#define UP(dst, src) FL_HIP(hipMemcpyAsync(dst, (src).data(), (src).size() * sizeof((src)[0]), hipMemcpyHostToDevice, s))
  UP(P.cell_fine, cell_fine);
  UP(P.cell_flag, flag);
  UP(P.cell_faces, cell_faces);
  UP(P.fine_cell, fine_cell);
  UP(P.face_tail, tail);
  UP(P.face_head, head);
  UP(P.face_kind, kind);
  UP(P.face_slot, slots);
  UP(P.face_area, fa);
  UP(P.face_beta, fb);
  UP(P.face_dist, fd);
  UP(P.face_ustar, fu);
  UP(P.face_uwall, fw);
  UP(P.face_flux, ff);
  UP(P.face_has_flux, fh);
  UP(P.face_lift, fl);
  UP(P.cell_volume, V);
  UP(P.cell_s, S);
  UP(P.cell_K, K);
  UP(P.cell_prev, prev);
  UP(P.cell_inwall, inw);
  UP(P.p_init, p0);
  P.p_init_zero = std::all_of(p0.begin(), p0.end(), [](double v) { return v == 0.0; });
  // Learned-preconditioner tiles (2x2x2 level-0 blocks) and cell classes: unknowns are liquid; a cell
  // across a sealed face of an unknown is solid; every other cell is air.
  {
    std::vector<uint8_t> cls(nfine, 2);
    for (size_t c = 0; c < nc; ++c) {
      if (flag[c] != 1) continue;
      const fl_int3 ci = cells[c].cell_index;
      cls[cell_fine[c]] = 1;
      for (int a = 0; a < 3; ++a)
        for (int side = 0; side < 2; ++side) {
          if (kind[cell_faces[6 * c + 2 * a + side]] != 0) continue;
          int64_t q[3] = {ci.x, ci.y, ci.z};
          q[a] += side ? 1 : -1;
          const Key3 b{int32_t(host_floor_div4(q[0])), int32_t(host_floor_div4(q[1])), int32_t(host_floor_div4(q[2]))};
          auto it = bidx.find(b);
          if (it == bidx.end()) continue;
          const int64_t fi = int64_t(it->second) * 64 +
                             cslot(int((q[2] - 4 * b.z) * 16 + (q[1] - 4 * b.y) * 4 + (q[0] - 4 * b.x)));
          if (cls[fi] != 1) cls[fi] = 3;
        }
    }
    std::vector<Key3> tiles;
    tiles.reserve(blocks0.size());
    for (const Key3& b : blocks0) tiles.push_back({floor_half(b.x), floor_half(b.y), floor_half(b.z)});
    std::sort(tiles.begin(), tiles.end(), key_less);
    tiles.erase(std::unique(tiles.begin(), tiles.end()), tiles.end());
    std::vector<int32_t> tb(8 * tiles.size(), -1);
    std::unordered_map<Key3, int32_t, Key3Hash> tidx;
    for (size_t t = 0; t < tiles.size(); ++t) tidx[tiles[t]] = int32_t(t);
    for (size_t i = 0; i < blocks0.size(); ++i) {
      const Key3& b = blocks0[i];
      const Key3 tk{floor_half(b.x), floor_half(b.y), floor_half(b.z)};
      const int o = ((b.z - 2 * tk.z) * 2 + (b.y - 2 * tk.y)) * 2 + (b.x - 2 * tk.x);
      tb[8 * tidx[tk] + o] = int32_t(i);
    }
    P.ntiles = tiles.size();
    UP(P.cell_class, cls);
    UP(P.tile_blocks, tb);
  }
#undef UP
  FL_HIP(hipStreamSynchronize(s));
  return build_levels(ctx, P, blocks0, 1);
}

// Freezes the system for dt on the device, assembles the RHS and the hierarchy coefficients, and
// computes the gates. The warm start must be in P.p_base.
// reuse: a caller whose system, levels and step duration may be unchanged since the last freeze (the
// device-resident step on the scene's fixed structure): then only the predicted velocities of the face
// records are refreshed; every coefficient is what the full freeze would write again.
int freeze_and_assemble(fl_ctx* ctx, Pressure& P, double dt, const fl_pressure_options& opt, Gates& gates,
                        EvalResult& e0, fl_pressure_diagnostics& diag, bool reuse = false) {
  hipStream_t s = ctx->stream;
  k_reset_err<<<1, 64, 0, s>>>(P.perr);
  k_copy_f64<<<grid_for(P.nc), kThreads, 0, s>>>(P.nc, P.p_init, P.p_base);
  const bool cached = reuse && P.coeff_valid && P.coeff_dt == dt;
  if (cached) {
    k_refresh_ustar<<<grid_for(P.nf), kThreads, 0, s>>>(P.nf, P.face_ustar, P.face_rec);
  } else {
    k_freeze_faces<<<grid_for(P.nf), kThreads, 0, s>>>(P.nf, dt, P.full_area, P.face_tail, P.face_head, P.face_kind,
                                                       P.face_area, P.face_beta, P.face_dist, P.face_ustar,
                                                       P.face_uwall, P.face_has_flux, P.face_flux, P.face_lift,
                                                       P.face_scale, P.face_k, P.face_wall, P.face_rec);
    k_freeze_cells<<<grid_for(P.nc), kThreads, 0, s>>>(P.nc, dt, P.cell_volume, P.cell_s, P.cell_K, P.cell_comp,
                                                       P.cell_target);
  }
  // RHS and the evaluation of the unprojected field (p = 0) in one pass; r = b there, so the residual
  // norm is the RHS norm.
  const unsigned g = reduce_grid(P.nc, kPartials);
  k_rhs_zero_eval<<<g, kThreads, 0, s>>>(P.nc, P.cell_faces, P.face_rec, P.cell_volume, P.cell_target, P.cell_inwall,
                                         P.cell_comp, P.cell_prev, P.rhs, P.rhs_abs, P.res, P.def, P.partial, P.perr);
  FL_HIP(hipGetLastError());
  if (!cached) {
    FL_TRY(setup_coefficients(ctx, P));
    P.coeff_valid = true;
    P.coeff_dt = dt;
  }
  FL_TRY(read_evaluation(ctx, P, g, e0));
  const double rhs_norm = e0.l2;
  double limit = 2e-6 + 2e-6 * rhs_norm;
  if (!std::isfinite(rhs_norm) || !std::isfinite(limit)) return FL_OUT_OF_RANGE;
  if (rhs_norm > 0.0 && P.has_jump)
    limit = std::min(limit, std::max(std::numeric_limits<double>::denorm_min(), 1e-3 * rhs_norm));
  gates.integrated = limit * opt.cpu_residual_tolerance_scale;
  gates.volume = 2e-5 * opt.cpu_residual_tolerance_scale;
  gates.relative = opt.relative_divergence_target;
  gates.initial_defect = e0.def_l2;
  diag.rhs_l2 = rhs_norm;
  diag.initial_defect_l2 = e0.def_l2;
  diag.gauge_count = P.h_gauges.size();
  k_reset_err<<<1, 64, 0, s>>>(P.perr);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int check_options(const fl_pressure_options* o) {
  if (!o) return FL_INVALID_ARGUMENT;
  if (o->max_iterations < 1 || o->max_iterations > 256) return FL_INVALID_ARGUMENT;
  if (!(o->cpu_residual_tolerance_scale >= 1e-12 && o->cpu_residual_tolerance_scale <= 1.0)) return FL_INVALID_ARGUMENT;
  if (o->preconditioner < 0 || o->preconditioner > 1 || o->smoother < 0 || o->smoother > 2) return FL_INVALID_ARGUMENT;
  if (o->level_storage < 1 || o->level_storage > 2 || o->levels < 1 || o->levels > 16 || o->smoothing_sweeps < 1 || o->coarse_sweeps < 1)
    return FL_INVALID_ARGUMENT;
  if (!(o->coarse_scale > 0.0 && o->coarse_scale <= 1.0)) return FL_INVALID_ARGUMENT;
  if (!(o->relative_divergence_target >= 0.0)) return FL_INVALID_ARGUMENT;
  return FL_OK;
}

}  // namespace

int project_pressure(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count, double dt,
                     const fl_pressure_cell* cells, size_t cell_count, const fl_pressure_face* faces,
                     size_t face_count, const fl_pressure_options* options, double* out_pressures,
                     double* out_velocities, double* out_rhs, double* out_residuals, double* out_defects,
                     uint64_t* out_gauges, fl_pressure_diagnostics* out_diagnostics, fl_measurement* measurement) {
  FL_TRY(check_options(options));
  if (!ctx->pressure) ctx->pressure = new Pressure();
  Pressure& P = *state(ctx);
  P.scale_coarse = options->coarse_scale;
  FL_TRY(prepare_system(ctx, P, geometry, blocks, block_count, dt, cells, cell_count, faces, face_count, true));
  // Multigrid levels requested by the options.
  std::vector<Key3> b0 = P.levels[0].keys;
  FL_TRY(build_levels(ctx, P, b0, options->levels));
  hipEvent_t a = ctx->stage[6], b = ctx->stage[7];
  FL_HIP(hipEventRecord(a, ctx->stream));
  Gates gates{};
  EvalResult e0{};
  fl_pressure_diagnostics diag{};
  int st = freeze_and_assemble(ctx, P, dt, *options, gates, e0, diag);
  if (st == FL_OK) st = solve(ctx, P, *options, gates, e0, diag);
  if (st == FL_OK) {
    k_face_velocity<<<grid_for(P.nf), kThreads, 0, ctx->stream>>>(P.nf, P.face_tail, P.face_head, P.face_kind,
                                                                  P.face_area, P.face_ustar, P.face_scale,
                                                                  P.face_lift, P.p, P.face_u);
  }
  (void)hipEventRecord(b, ctx->stream);
  (void)hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  if (out_diagnostics) *out_diagnostics = diag;
  if (st != FL_OK) return st;
  hipStream_t s = ctx->stream;
  const size_t cb = cell_count * sizeof(double), fb = face_count * sizeof(double);
  if (out_pressures) FL_HIP(hipMemcpyAsync(out_pressures, P.p, cb, hipMemcpyDeviceToHost, s));
  if (out_velocities) FL_HIP(hipMemcpyAsync(out_velocities, P.face_u, fb, hipMemcpyDeviceToHost, s));
  if (out_rhs) FL_HIP(hipMemcpyAsync(out_rhs, P.rhs, cb, hipMemcpyDeviceToHost, s));
  if (out_residuals) FL_HIP(hipMemcpyAsync(out_residuals, P.res, cb, hipMemcpyDeviceToHost, s));
  if (out_defects) FL_HIP(hipMemcpyAsync(out_defects, P.def, cb, hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  if (out_gauges) std::copy(P.h_gauges.begin(), P.h_gauges.end(), out_gauges);
  if (measurement) {
    measurement->kernel_ms = ms;
    measurement->copy_ms = 0;
    measurement->launches = 0;
    std::snprintf(measurement->architecture, sizeof(measurement->architecture), "%s", ctx->arch);
  }
  return FL_OK;
}

// ---------------------------------------------------------------- device-built pressure systems
//
// prepare_system and build_levels on the device, for link records that are already device arrays
// (the typed links of a frozen geometry). Each phase's checks are the host's; their failures carry the
// phase in the error word above the ordinal (cells, incidence, faces, the cells' and then the faces'
// right-hand-side representability), so the first failure keeps the host's phase order. Connected
// components over the positive internal links come from a lock-free union-find that always hooks the
// larger root under the smaller, so every component's root is its least cell, as on the host; when a
// component has no anchor (no compliance and no Dirichlet link), its compatibility sum and gauge are
// left to the host's code on downloaded records. The hierarchy is built level by level from per-block
// masks of unknown cells: the coarse blocks of a level are marked in a lookup over its block box and
// numbered by a scan in (z,y,x) order, and coarsening stops where the host's would (no fewer cells,
// the pool, or the requested depth).
namespace {

constexpr uint64_t kSysPhase = uint64_t(1) << 40;

__device__ inline void sys_error(uint64_t* err, int phase, uint64_t ordinal, int code) {
  perr_report(err, uint64_t(phase) * kSysPhase + ordinal, code);
}
__device__ inline bool rep_dev(double a, double b, double p) {
  if (!finite(p)) return false;
  if (a == 0.0 || b == 0.0 || fabs(p) >= 2.2250738585072014e-308) return true;
  if (p == 0.0) return false;
  int ea = 0, eb = 0;
  const double ma = frexp(a, &ea), mb = frexp(b, &eb);
  return fma(ma, mb, -scalbn(p, -(ea + eb))) == 0.0;
}
__device__ inline bool fpos(double v) { return finite(v) && v > 0.0; }

struct SysGrid {
  const int32_t* lut;  // grid block ordinal over box, -1 inactive
  Box box;
  const uint64_t* scan0;  // level-0 index of a marked grid block
  const uint64_t* mark0;  // 1 when the grid block holds a cell
  __device__ int64_t ordinal(int64_t x, int64_t y, int64_t z) const {
    if (x < kI32Min || x > kI32Max || y < kI32Min || y > kI32Max || z < kI32Min || z > kI32Max) return -1;
    const int64_t i = box.index(int32_t(x), int32_t(y), int32_t(z));
    return i < 0 ? -1 : lut[i];
  }
  // Input ordinal of the cell at lattice index (x, y, z), or -1.
  __device__ int32_t cell(const int32_t* fine_cell, int64_t x, int64_t y, int64_t z) const {
    const int64_t bx = floor_div4(x), by = floor_div4(y), bz = floor_div4(z);
    const int64_t g = ordinal(bx, by, bz);
    if (g < 0 || !mark0[g]) return -1;
    const int l = int((z - 4 * bz) * 16 + (y - 4 * by) * 4 + (x - 4 * bx));
    return fine_cell[int64_t(scan0[g]) * 64 + cslot(l)];
  }
};

__global__ void k_sys_cells(const fl_pressure_cell* cells, uint64_t nc, SysGrid G, double volume, double dt,
                            int check_values, double* V, double* S, double* K, double* prev, double* inw,
                            double* comp, double* target, uint64_t* mark, uint64_t* words, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nc) return;
  const fl_pressure_cell c = cells[i];
  int code = 0;
  if (i) {
    const fl_int3 a = cells[i - 1].cell_index, b = c.cell_index;
    const bool less = a.z != b.z ? a.z < b.z : (a.y != b.y ? a.y < b.y : a.x < b.x);
    if (!less) code = FL_INVALID_ARGUMENT;
  }
  double cmp = 0.0, tf = 0.0;
  if (!code && (!fpos(c.fluid_volume) || c.fluid_volume > volume)) code = FL_INVALID_ARGUMENT;
  if (!code && (!finite(c.target_divergence) || !finite(c.previous_pressure) || !finite(c.integrated_inward_wall_flux)))
    code = FL_INVALID_ARGUMENT;
  if (!code && c.bulk_modulus != __builtin_inf()) {
    if (!fpos(c.bulk_modulus)) {
      code = FL_INVALID_ARGUMENT;
    } else {
      const double kdt = c.bulk_modulus * dt;
      if (!fpos(kdt)) code = FL_OUT_OF_RANGE;
      if (!code) {
        cmp = c.fluid_volume / kdt;
        if (!fpos(cmp)) code = FL_OUT_OF_RANGE;
      }
    }
  }
  if (!code) {
    tf = c.fluid_volume * c.target_divergence;
    if (!rep_dev(c.fluid_volume, c.target_divergence, tf)) code = FL_OUT_OF_RANGE;
  }
  const int64_t xyz[3] = {c.cell_index.x, c.cell_index.y, c.cell_index.z};
  for (int a = 0; a < 3 && !code; ++a) {
    if (xyz[a] == kI32Max) {
      code = FL_OUT_OF_RANGE;
      break;
    }
    for (int side = 0; side < 2; ++side) {
      int64_t l[3] = {xyz[0], xyz[1], xyz[2]};
      l[a] += side;
      if (G.ordinal(floor_div4(l[0]), floor_div4(l[1]), floor_div4(l[2])) < 0) {
        code = FL_FAILED_PRECONDITION;
        break;
      }
    }
  }
  if (code) {
    sys_error(err, 0, i, code);
    return;
  }
  V[i] = c.fluid_volume;
  S[i] = c.target_divergence;
  K[i] = c.bulk_modulus;
  prev[i] = c.previous_pressure;
  inw[i] = c.integrated_inward_wall_flux;
  comp[i] = cmp;
  target[i] = tf;
  mark[G.ordinal(floor_div4(xyz[0]), floor_div4(xyz[1]), floor_div4(xyz[2]))] = 1;
  if (c.previous_pressure != 0.0) atomicOr(reinterpret_cast<unsigned long long*>(words + 1), 1ull);
  if (check_values && !rep_dev(cmp, c.previous_pressure, cmp * c.previous_pressure)) sys_error(err, 3, i, FL_OUT_OF_RANGE);
}

__global__ void k_sys_keys0(const uint64_t* mark, const uint64_t* scan, uint64_t nb, const int32_t* blocks,
                            int32_t* keys) {
  const uint64_t g = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (g >= nb || !mark[g]) return;
  const uint64_t b = scan[g];
  for (int a = 0; a < 3; ++a) keys[3 * b + a] = blocks[3 * g + a];
}

__global__ void k_sys_fine(const fl_pressure_cell* cells, uint64_t nc, SysGrid G, int32_t* cell_fine,
                           int32_t* fine_cell) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nc) return;
  const fl_int3 ci = cells[i].cell_index;
  const int64_t bx = floor_div4(ci.x), by = floor_div4(ci.y), bz = floor_div4(ci.z);
  const int64_t g = G.ordinal(bx, by, bz);
  const int l = int((ci.z - 4 * bz) * 16 + (ci.y - 4 * by) * 4 + (ci.x - 4 * bx));
  const int64_t fi = int64_t(G.scan0[g]) * 64 + cslot(l);
  cell_fine[i] = int32_t(fi);
  fine_cell[fi] = int32_t(i);
}

// FreezeLinkFromIncidence (mac_pressure_3d_arithmetic.h:174-245) per face, with the face's slot, its
// endpoints and the exact-incidence checks; cond goes to kcond for the components.
__global__ void k_sys_faces(const fl_pressure_face* faces, uint64_t nf, SysGrid G, const int32_t* fine_cell,
                            double area, double dt, int check_values, int32_t* tail_out, int32_t* head_out,
                            int32_t* kind_out, uint64_t* slot_out, double* fa, double* fb, double* fd, double* fu,
                            double* fw, double* ff, int32_t* fh, double* fl, double* kcond, int32_t* slot_face,
                            uint64_t* words, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nf) return;
  const fl_pressure_face f = faces[i];
  auto slot_of = [&](const fl_pressure_face& q, uint64_t& slot) -> int {
    if (q.axis < 0 || q.axis > 2) return FL_INVALID_ARGUMENT;
    const int64_t x = q.lattice.x, y = q.lattice.y, z = q.lattice.z;
    const int64_t bx = floor_div4(x), by = floor_div4(y), bz = floor_div4(z);
    const int64_t g = G.ordinal(bx, by, bz);
    if (g < 0) return FL_FAILED_PRECONDITION;
    slot = uint64_t(g) * 192 + uint64_t(q.axis) * 64 + uint64_t(((z - 4 * bz) * 4 + (y - 4 * by)) * 4 + (x - 4 * bx));
    return 0;
  };
  uint64_t slot = 0;
  int code = slot_of(f, slot);
  if (code) {
    sys_error(err, 1, i, code);
    return;
  }
  if (i) {
    uint64_t before = 0;
    if (slot_of(faces[i - 1], before) || before >= slot) {
      sys_error(err, 1, i, FL_INVALID_ARGUMENT);
      return;
    }
  }
  int64_t t[3] = {f.lattice.x, f.lattice.y, f.lattice.z};
  t[f.axis] -= 1;
  const int32_t to = G.cell(fine_cell, t[0], t[1], t[2]);
  const int32_t ho = G.cell(fine_cell, f.lattice.x, f.lattice.y, f.lattice.z);
  if (to < 0 && ho < 0) {  // not incident to any cell: the face list is not the exact incident union
    sys_error(err, 1, i, FL_INVALID_ARGUMENT);
    return;
  }
  if (!finite(f.open_area) || f.open_area < 0.0 || f.open_area > area) code = FL_INVALID_ARGUMENT;
  if (!code && (!finite(f.inverse_density) || !finite(f.pressure_gradient_distance) ||
                !finite(f.predicted_normal_velocity) || !finite(f.wall_normal_velocity)))
    code = FL_INVALID_ARGUMENT;
  const bool both = to >= 0 && ho >= 0, one = (to >= 0) != (ho >= 0);
  if (!code) {
    if (f.link_kind == 1) {
      if (!both || !finite(f.lift)) code = FL_INVALID_ARGUMENT;
    } else if (f.link_kind == 2) {
      if (!one || f.open_area == 0.0 || !finite(f.lift)) code = FL_INVALID_ARGUMENT;
    } else if (f.link_kind == 0) {
      if (f.open_area != 0.0) code = FL_INVALID_ARGUMENT;
    } else {
      code = FL_INVALID_ARGUMENT;
    }
  }
  double scale = 0.0, cond = 0.0, wf = 0.0;
  if (!code && f.open_area > 0.0) {
    if (f.inverse_density <= 0.0 || f.pressure_gradient_distance <= 0.0) {
      code = FL_INVALID_ARGUMENT;
    } else {
      scale = (dt * f.inverse_density) / f.pressure_gradient_distance;
      cond = f.open_area * scale;
      if (!fpos(scale) || !fpos(cond)) code = FL_OUT_OF_RANGE;
    }
  }
  if (!code) {
    const double blocked = area - f.open_area;
    if (f.has_prescribed_wall_flux) {
      wf = f.prescribed_wall_flux;
      if (!finite(wf) || (blocked == 0 && wf != 0)) code = FL_INVALID_ARGUMENT;
    } else {
      wf = blocked * f.wall_normal_velocity;
      if (!rep_dev(blocked, f.wall_normal_velocity, wf)) code = FL_OUT_OF_RANGE;
    }
  }
  if (code) {
    sys_error(err, 2, i, code);
    return;
  }
  tail_out[i] = to;
  head_out[i] = ho;
  kind_out[i] = f.link_kind;
  slot_out[i] = slot;
  fa[i] = f.open_area;
  fb[i] = f.inverse_density;
  fd[i] = f.pressure_gradient_distance;
  fu[i] = f.predicted_normal_velocity;
  fw[i] = f.wall_normal_velocity;
  ff[i] = f.prescribed_wall_flux;
  fh[i] = f.has_prescribed_wall_flux;
  fl[i] = f.link_kind == 0 ? 0.0 : f.lift;
  kcond[i] = cond;
  slot_face[slot] = int32_t(i);
  if (cond > 0.0 && f.link_kind == 1 && f.lift != 0.0) atomicOr(reinterpret_cast<unsigned long long*>(words), 1ull);
  if (check_values) {
    if (!rep_dev(f.open_area, f.predicted_normal_velocity, f.open_area * f.predicted_normal_velocity) ||
        ((f.link_kind == 1 || f.link_kind == 2) && !rep_dev(cond, f.lift, cond * f.lift)))
      sys_error(err, 4, i, FL_OUT_OF_RANGE);
  }
}

// The six incident faces of each cell (in the host's order: -x, +x, -y, +y, -z, +z).
__global__ void k_sys_cell_faces(const fl_pressure_cell* cells, uint64_t nc, SysGrid G, const int32_t* slot_face,
                                 int32_t* cell_faces, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nc) return;
  const fl_int3 ci = cells[i].cell_index;
  const int64_t xyz[3] = {ci.x, ci.y, ci.z};
  for (int a = 0; a < 3; ++a)
    for (int side = 0; side < 2; ++side) {
      int64_t l[3] = {xyz[0], xyz[1], xyz[2]};
      l[a] += side;
      const int64_t bx = floor_div4(l[0]), by = floor_div4(l[1]), bz = floor_div4(l[2]);
      const int64_t g = G.ordinal(bx, by, bz);
      const uint64_t slot = uint64_t(g) * 192 + uint64_t(a) * 64 +
                            uint64_t(((l[2] - 4 * bz) * 4 + (l[1] - 4 * by)) * 4 + (l[0] - 4 * bx));
      const int32_t f = slot_face[slot];
      if (f < 0) {
        sys_error(err, 1, i, FL_INVALID_ARGUMENT);
        return;
      }
      cell_faces[6 * i + 2 * a + side] = f;
    }
}

// Root of x with path halving. Links only ever point from a root to a smaller root, so a node's
// grandparent is always one of its ancestors and the halving store needs no atomic: a racing store
// can only leave a longer, still valid path.
__device__ inline int32_t uf_find(int32_t* parent, int32_t x) {
  volatile int32_t* p = parent;
  while (true) {
    const int32_t up = p[x];
    if (up == x) return x;
    const int32_t next = p[up];
    if (next != up) p[x] = next;
    x = next;
  }
}

__global__ void k_uf_init(int32_t* parent, uint64_t nc) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < nc) parent[i] = int32_t(i);
}

// Unites the endpoints of every positive internal link: the larger root is hooked under the smaller.
__global__ void k_uf_union(const double* kcond, const int32_t* tail, const int32_t* head, uint64_t nf,
                           int32_t* parent) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f >= nf || !(kcond[f] > 0.0) || tail[f] < 0 || head[f] < 0) return;
  int32_t a = tail[f], b = head[f];
  while (true) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    if (a < b) {
      const int32_t t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(&parent[a], a, b) == a) return;
  }
}

__global__ void k_uf_roots(int32_t* parent, uint64_t nc, const double* comp, int32_t* root, uint8_t* anchor) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nc) return;
  const int32_t r = uf_find(parent, int32_t(i));
  root[i] = r;
  if (comp[i] > 0.0) anchor[r] = 1;
}

__global__ void k_uf_dirichlet(const int32_t* kind, const int32_t* tail, const int32_t* head, uint64_t nf,
                               const int32_t* root, uint8_t* anchor) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f >= nf || kind[f] != 2) return;
  if (tail[f] >= 0) anchor[root[tail[f]]] = 1;
  if (head[f] >= 0) anchor[root[head[f]]] = 1;
}

// Unanchored components into words[2]; every cell an unknown, the warm start its previous pressure.
__global__ void k_uf_finish(const int32_t* root, const uint8_t* anchor, uint64_t nc, const double* prev,
                            uint8_t* flag, double* p_init, uint64_t* words) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nc) return;
  if (root[i] == int32_t(i) && !anchor[i]) atomicAdd(reinterpret_cast<unsigned long long*>(words + 2), 1ull);
  flag[i] = 1;
  p_init[i] = prev[i];
}

// Unknown-cell mask of each level-0 block, bit l = z * 16 + y * 4 + x.
__global__ void k_occ0(const int32_t* fine_cell, const uint8_t* flag, uint64_t nb, uint64_t* occ) {
  const uint64_t b = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (b >= nb) return;
  uint64_t m = 0;
  for (int s = 0; s < 64; ++s) {
    const int32_t c = fine_cell[b * 64 + s];
    if (c >= 0 && flag[c] == 1) m |= uint64_t(1) << cnat(s);
  }
  occ[b] = m;
}

__host__ __device__ inline int32_t half_floor(int32_t v) { return v >= 0 ? v / 2 : -((-(v + 1)) / 2) - 1; }

// Marks the coarse blocks of the blocks that hold unknown cells (mark[box index] = 1).
__global__ void k_coarse_mark(const int32_t* keys, const uint64_t* occ, uint64_t nb, Box box, uint64_t* mark) {
  const uint64_t b = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (b >= nb || !occ[b]) return;
  mark[box.index(half_floor(keys[3 * b]), half_floor(keys[3 * b + 1]), half_floor(keys[3 * b + 2]))] = 1;
}

// Coarse lookup and keys from the marks and their scan.
__global__ void k_coarse_lut(const uint64_t* mark, const uint64_t* scan, Box box, int32_t* lut, int32_t* keys) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= box.volume()) return;
  if (!mark[i]) {
    lut[i] = -1;
    return;
  }
  const uint64_t c = scan[i];
  lut[i] = int32_t(c);
  const uint64_t x = i % box.dim[0], y = (i / box.dim[0]) % box.dim[1], z = i / (uint64_t(box.dim[0]) * box.dim[1]);
  keys[3 * c] = box.lo[0] + int32_t(x);
  keys[3 * c + 1] = box.lo[1] + int32_t(y);
  keys[3 * c + 2] = box.lo[2] + int32_t(z);
}

// Parent (coarse block, octant) of every fine block whose coarse key exists, and the coarse masks.
__global__ void k_coarse_link(const int32_t* keys, const uint64_t* occ, uint64_t nb, Box box, const int32_t* lut,
                              int32_t* parent, uint64_t* occ_c) {
  const uint64_t b = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (b >= nb) return;
  const int32_t k[3] = {keys[3 * b], keys[3 * b + 1], keys[3 * b + 2]};
  const int32_t pk[3] = {half_floor(k[0]), half_floor(k[1]), half_floor(k[2])};
  const int64_t bi = box.index(pk[0], pk[1], pk[2]);
  const int32_t c = bi < 0 ? -1 : lut[bi];
  const int o[3] = {k[0] - 2 * pk[0], k[1] - 2 * pk[1], k[2] - 2 * pk[2]};
  if (parent) {
    parent[2 * b] = c;
    parent[2 * b + 1] = c < 0 ? -1 : (o[2] * 2 + o[1]) * 2 + o[0];
  }
  uint64_t m = occ[b];
  if (c < 0 || !m) return;
  uint64_t out = 0;
  while (m) {
    const int l = __ffsll(static_cast<long long>(m)) - 1;
    m &= m - 1;
    const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
    const int px = (o[0] * 4 + lx) >> 1, py = (o[1] * 4 + ly) >> 1, pz = (o[2] * 4 + lz) >> 1;
    out |= uint64_t(1) << (pz * 16 + py * 4 + px);
  }
  atomicOr(reinterpret_cast<unsigned long long*>(occ_c + c), static_cast<unsigned long long>(out));
}

__global__ void k_popcount_sum(const uint64_t* occ, uint64_t nb, uint64_t* out) {
  uint64_t s = 0;
  for (uint64_t b = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; b < nb; b += uint64_t(gridDim.x) * blockDim.x)
    s += __popcll(occ[b]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(reinterpret_cast<unsigned long long*>(out), static_cast<unsigned long long>(s));
}

// Six-neighbour table of a level from its lookup.
__global__ void k_level_nbr(const int32_t* keys, uint64_t nb, Box box, const int32_t* lut, int32_t* nbr) {
  const uint64_t b = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (b >= nb) return;
  const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  for (int dir = 0; dir < 6; ++dir) {
    const int64_t i = box.index(keys[3 * b] + d[dir][0], keys[3 * b + 1] + d[dir][1], keys[3 * b + 2] + d[dir][2]);
    nbr[6 * b + dir] = i < 0 ? -1 : lut[i];
  }
}

// Level-0 lookup over the grid's box: the level-0 index of each marked grid block.
__global__ void k_lut0(const int32_t* grid_lut, uint64_t vol, const uint64_t* mark, const uint64_t* scan, int32_t* lut) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= vol) return;
  const int32_t g = grid_lut[i];
  lut[i] = g >= 0 && mark[g] ? int32_t(scan[g]) : -1;
}

__global__ void k_children(const int32_t* parent, uint64_t nb, int32_t* child) {
  const uint64_t b = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (b >= nb || parent[2 * b] < 0) return;
  child[8 * parent[2 * b] + parent[2 * b + 1]] = int32_t(b);
}

// Size of a marked set: last scan entry plus its mark.
__global__ void k_totals_mark(const uint64_t* mark, const uint64_t* scan, uint64_t n, uint64_t* out) {
  if (threadIdx.x == 0) out[0] = n ? scan[n - 1] + (mark[n - 1] ? 1 : 0) : 0;
}

__global__ void k_fill_words(uint64_t* p, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = 0;
}
__global__ void k_fill_int(int32_t* p, uint64_t n, int32_t v) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_byte(uint8_t* p, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = 0;
}

inline unsigned capped(uint64_t n) { return static_cast<unsigned>(std::min<uint64_t>(grid_for(n), 65536)); }

int ensure_device_scratch(fl_ctx* ctx, Pressure& P) {
  const uint64_t grid = std::max<uint64_t>(ctx->cap.blocks, 1), cells = std::max<uint64_t>(ctx->cap.pressure_cells, 1);
  const uint64_t boxes = std::max<uint64_t>({ctx->cap.lookup_blocks, grid, 64});
  if (P.dv_mark && grid <= P.cap_dv_grid && cells <= P.cap_dv_cells && boxes <= P.cap_dv_boxes) return FL_OK;
  void* dv[] = {P.dv_mark, P.dv_slot_face, P.dv_parent, P.dv_root, P.dv_anchor, P.dv_words, P.dv_lut,
                P.dv_occ[0], P.dv_occ[1], P.dv_keys};
  for (void* p : dv)
    if (p) (void)hipFree(p);
  FL_TRY(dal(&P.dv_mark, 2 * boxes));
  FL_TRY(dal(&P.dv_slot_face, 192 * grid));
  FL_TRY(dal(&P.dv_parent, cells));
  FL_TRY(dal(&P.dv_root, cells));
  FL_TRY(dal(&P.dv_anchor, cells));
  FL_TRY(dal(&P.dv_words, 8));
  FL_TRY(dal(&P.dv_lut, boxes));
  FL_TRY(dal(&P.dv_occ[0], grid));
  FL_TRY(dal(&P.dv_occ[1], grid));
  FL_TRY(dal(&P.dv_keys, 6 * grid));
  P.cap_dv_grid = grid;
  P.cap_dv_cells = cells;
  P.cap_dv_boxes = boxes;
  return FL_OK;
}

// Reads n device words into ctx->host_words, with the error word after them.
int read_words(fl_ctx* ctx, const uint64_t* d, int n) {
  FL_HIP(hipMemcpyAsync(ctx->host_words, d, n * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipMemcpyAsync(ctx->host_words + n, ctx->err, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  return FL_OK;
}

// build_levels on the device: level 0 is the nb0 blocks whose keys are in P.dv_keys.
int build_levels_device(fl_ctx* ctx, Pressure& P, uint64_t nb0, int levels_wanted) {
  P.coeff_valid = false;
  P.bf16_valid = false;
  hipStream_t s = ctx->stream;
  free_levels(P);
  P.levels_requested = levels_wanted;
  uint64_t planned = level_bytes(nb0);
  if (planned > P.pool_bytes) return FL_RESOURCE_EXHAUSTED;
  // Level 0: lookup over the grid box, masks of unknown cells.
  Box box = ctx->block_box;
  const uint64_t vol0 = box.volume();
  if (vol0 > P.cap_dv_boxes) return FL_RESOURCE_EXHAUSTED;
  k_lut0<<<grid_for(vol0), kThreads, 0, s>>>(ctx->lut, vol0, P.dv_mark, P.dv_mark + P.cap_dv_boxes, P.dv_lut);
  k_occ0<<<grid_for(nb0), kThreads, 0, s>>>(P.fine_cell, P.cell_flag, nb0, P.dv_occ[0]);
  FL_HIP(hipGetLastError());
  auto add_level = [&](uint64_t nb, const int32_t* keys) -> int {
    Level L;
    L.nb = nb;
    L.nbr = carve<int32_t>(P, 6 * nb);
    L.child = P.levels.empty() ? nullptr : carve<int32_t>(P, 8 * nb);
    L.parent = carve<int32_t>(P, 2 * nb);
    const uint64_t n = nb * 64;
    L.x = carve<float>(P, n);
    L.r = carve<float>(P, n);
    L.diagb = carve<float>(P, n);
    L.b = carve<float>(P, n);
    L.anchor = carve<float>(P, n);
    L.diag = carve<float>(P, n);
    L.kb = carve<uint16_t>(P, 3 * n);
    L.k = carve<float>(P, 3 * n);
    if (P.pool_used > P.pool_bytes) return FL_INTERNAL;
    k_level_nbr<<<grid_for(nb), kThreads, 0, s>>>(keys, nb, box, P.dv_lut, L.nbr);
    FL_HIP(hipGetLastError());
    P.levels.push_back(L);
    return FL_OK;
  };
  const int32_t* keys = P.dv_keys;
  FL_TRY(add_level(nb0, keys));
  uint64_t cells = 0;
  {
    k_fill_words<<<1, 64, 0, s>>>(P.dv_words + 4, 1);
    k_popcount_sum<<<capped(nb0), kThreads, 0, s>>>(P.dv_occ[0], nb0, P.dv_words + 4);
    FL_TRY(read_words(ctx, P.dv_words + 4, 1));
    cells = ctx->host_words[0];
  }
  int cur = 0;
  uint64_t nb = nb0;
  int32_t* next_keys = P.dv_keys + 3 * nb0;
  while (static_cast<int>(P.levels.size()) < levels_wanted && cells > 0) {
    Box cbox;
    for (int a = 0; a < 3; ++a) {
      const int32_t lo = half_floor(box.lo[a]), hi = half_floor(box.lo[a] + box.dim[a] - 1);
      cbox.lo[a] = lo;
      cbox.dim[a] = hi - lo + 1;
    }
    const uint64_t cvol = cbox.volume();
    if (next_keys + 3 * nb > P.dv_keys + 6 * P.cap_dv_grid) break;  // room for at most nb coarse keys
    uint64_t* mark = P.dv_mark;
    uint64_t* scan = P.dv_mark + P.cap_dv_boxes;
    k_fill_words<<<capped(cvol), kThreads, 0, s>>>(mark, cvol);
    k_coarse_mark<<<grid_for(nb), kThreads, 0, s>>>(keys, P.dv_occ[cur], nb, cbox, mark);
    FL_HIP(hipGetLastError());
    FL_TRY(exclusive_scan_u64(ctx, mark, scan, cvol));
    // Coarse lookup (replaces the fine one: the fine level's neighbours are already built).
    k_coarse_lut<<<grid_for(cvol), kThreads, 0, s>>>(mark, scan, cbox, P.dv_lut, next_keys);
    k_fill_words<<<capped(nb), kThreads, 0, s>>>(P.dv_occ[1 - cur], std::max<uint64_t>(nb, 1));
    k_fill_words<<<1, 64, 0, s>>>(P.dv_words + 4, 2);
    k_totals_mark<<<1, 64, 0, s>>>(mark, scan, cvol, P.dv_words + 5);
    k_coarse_link<<<grid_for(nb), kThreads, 0, s>>>(keys, P.dv_occ[cur], nb, cbox, P.dv_lut,
                                                    P.levels.back().parent, P.dv_occ[1 - cur]);
    FL_HIP(hipGetLastError());
    // The coarse block count first, then its cell count.
    FL_TRY(read_words(ctx, P.dv_words + 5, 1));
    const uint64_t nbc = ctx->host_words[0];
    k_popcount_sum<<<capped(nbc), kThreads, 0, s>>>(P.dv_occ[1 - cur], nbc, P.dv_words + 4);
    FL_TRY(read_words(ctx, P.dv_words + 4, 1));
    const uint64_t ccells = ctx->host_words[0];
    if (ccells == cells) break;
    if (planned + level_bytes(nbc) > P.pool_bytes) break;  // the pool bounds the depth
    planned += level_bytes(nbc);
    box = cbox;
    FL_TRY(add_level(nbc, next_keys));
    keys = next_keys;
    next_keys = const_cast<int32_t*>(keys) + 3 * nbc;
    nb = nbc;
    cells = ccells;
    cur = 1 - cur;
  }
  // The coarsest level has no parent; every coarser level's children from the parents below it.
  P.levels.back().parent = nullptr;
  for (size_t l = 1; l < P.levels.size(); ++l) {
    Level& C = P.levels[l];
    k_fill_int<<<capped(8 * C.nb), kThreads, 0, s>>>(C.child, 8 * C.nb, -1);
    k_children<<<grid_for(P.levels[l - 1].nb), kThreads, 0, s>>>(P.levels[l - 1].parent, P.levels[l - 1].nb, C.child);
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

}  // namespace

// prepare_system on device link records: the grid is the context's (ctx->blocks and its lookup).
// Returns the system's status; a component without an anchor sends the records to the host's
// prepare_system.
int prepare_system_device(fl_ctx* ctx, Pressure& P, const fl_pressure_cell* cells, uint64_t nc,
                          const fl_pressure_face* faces, uint64_t nf, double dt, int levels, bool* host_fallback) {
  *host_fallback = false;
  P.coeff_valid = false;
  P.bf16_valid = false;
  hipStream_t s = ctx->stream;
  if (nc == 0 || nf == 0 || !finite_positive(dt)) return FL_INVALID_ARGUMENT;
  const double h = ctx->geometry.h, area = h * h, volume = area * h;
  if (!finite_positive(area) || !finite_positive(volume)) return FL_OUT_OF_RANGE;
  if (nc > ctx->cap.pressure_cells || nf > ctx->cap.pressure_faces) return FL_RESOURCE_EXHAUSTED;
  if (nc > 0x7fffffffULL || nf > 0x7fffffffULL) return FL_RESOURCE_EXHAUSTED;
  FL_TRY(ensure_capacity(ctx, nc, nf, 64 * ctx->nb));
  FL_TRY(ensure_device_scratch(ctx, P));
  const uint64_t ngrid = ctx->nb;
  if (ngrid > P.cap_dv_grid) return FL_RESOURCE_EXHAUSTED;
  P.full_area = area;
  P.full_volume = volume;
  uint64_t* mark = P.dv_mark;
  uint64_t* scan = P.dv_mark + P.cap_dv_boxes;
  FL_TRY(reset_error(ctx));
  k_fill_words<<<capped(ngrid), kThreads, 0, s>>>(mark, ngrid);
  k_fill_words<<<1, 64, 0, s>>>(P.dv_words, 8);
  SysGrid G{ctx->lut, ctx->block_box, scan, mark};
  k_sys_cells<<<grid_for(nc), kThreads, 0, s>>>(cells, nc, G, volume, dt, 1, P.cell_volume, P.cell_s, P.cell_K,
                                                P.cell_prev, P.cell_inwall, P.cell_comp, P.cell_target, mark,
                                                P.dv_words, ctx->err);
  FL_HIP(hipGetLastError());
  FL_TRY(exclusive_scan_u64(ctx, mark, scan, ngrid));
  k_totals_mark<<<1, 64, 0, s>>>(mark, scan, ngrid, P.dv_words + 3);
  FL_TRY(read_words(ctx, P.dv_words + 3, 1));
  const uint64_t nb0 = ctx->host_words[0];
  if (ctx->host_words[1] != ~uint64_t(0) && (ctx->host_words[1] >> 8) < kSysPhase) return int(ctx->host_words[1] & 0xff);
  const uint64_t nfine = nb0 * 64;
  if (nfine > P.cap_fine) return FL_RESOURCE_EXHAUSTED;
  k_sys_keys0<<<grid_for(ngrid), kThreads, 0, s>>>(mark, scan, ngrid, ctx->blocks, P.dv_keys);
  k_fill_int<<<capped(nfine), kThreads, 0, s>>>(P.fine_cell, nfine, -1);
  k_fill_int<<<capped(192 * ngrid), kThreads, 0, s>>>(P.dv_slot_face, 192 * ngrid, -1);
  k_sys_fine<<<grid_for(nc), kThreads, 0, s>>>(cells, nc, G, P.cell_fine, P.fine_cell);
  k_sys_faces<<<grid_for(nf), kThreads, 0, s>>>(faces, nf, G, P.fine_cell, area, dt, 1, P.face_tail, P.face_head,
                                                P.face_kind, P.face_slot, P.face_area, P.face_beta, P.face_dist,
                                                P.face_ustar, P.face_uwall, P.face_flux, P.face_has_flux, P.face_lift,
                                                P.face_k, P.dv_slot_face, P.dv_words, ctx->err);
  k_sys_cell_faces<<<grid_for(nc), kThreads, 0, s>>>(cells, nc, G, P.dv_slot_face, P.cell_faces, ctx->err);
  k_uf_init<<<grid_for(nc), kThreads, 0, s>>>(P.dv_parent, nc);
  k_fill_byte<<<capped(nc), kThreads, 0, s>>>(P.dv_anchor, nc);
  k_uf_union<<<grid_for(nf), kThreads, 0, s>>>(P.face_k, P.face_tail, P.face_head, nf, P.dv_parent);
  k_uf_roots<<<grid_for(nc), kThreads, 0, s>>>(P.dv_parent, nc, P.cell_comp, P.dv_root, P.dv_anchor);
  k_uf_dirichlet<<<grid_for(nf), kThreads, 0, s>>>(P.face_kind, P.face_tail, P.face_head, nf, P.dv_root, P.dv_anchor);
  k_uf_finish<<<grid_for(nc), kThreads, 0, s>>>(P.dv_root, P.dv_anchor, nc, P.cell_prev, P.cell_flag, P.p_init,
                                                P.dv_words);
  FL_HIP(hipGetLastError());
  FL_TRY(read_words(ctx, P.dv_words, 3));
  if (ctx->host_words[3] != ~uint64_t(0)) return int(ctx->host_words[3] & 0xff);
  P.has_jump = ctx->host_words[0] != 0;
  P.p_init_zero = ctx->host_words[1] == 0;
  P.nc = nc;
  P.nf = nf;
  P.h_gauges.clear();
  P.h_cell_comp_root.clear();
  P.h_unknown_cells.clear();
  P.ntiles = 0;  // the learned preconditioner's tiles are not built for device systems
  if (ctx->host_words[2] != 0) {
    *host_fallback = true;
    return FL_OK;
  }
  return build_levels_device(ctx, P, nb0, levels);
}

// The projection of device link records on the context's grid (the typed links of a frozen geometry):
// the system and its hierarchy built on the device, frozen for dt, solved, and the corrected face
// velocities written into unew_slots at the faces' owned slots (other slots untouched).
int project_links_device(fl_ctx* ctx, const fl_pressure_cell* cells, uint64_t nc, const fl_pressure_face* faces,
                         uint64_t nf, double dt, const fl_pressure_options& opt, double* unew_slots,
                         fl_pressure_diagnostics* out_diag) {
  FL_TRY(check_options(&opt));
  if (opt.preconditioner != 0) return FL_FAILED_PRECONDITION;
  if (!ctx->pressure) ctx->pressure = new Pressure();
  Pressure& P = *state(ctx);
  P.scale_coarse = opt.coarse_scale;
  bool fallback = false;
  FL_TRY(prepare_system_device(ctx, P, cells, nc, faces, nf, dt, opt.levels, &fallback));
  if (fallback) {
    // A component without an anchor: its compatibility test and gauge on the host, over the records.
    hipStream_t s = ctx->stream;
    std::vector<fl_pressure_cell> hc(nc);
    std::vector<fl_pressure_face> hf(nf);
    std::vector<fl_int3> hb(ctx->nb);
    FL_HIP(hipMemcpyAsync(hc.data(), cells, nc * sizeof(fl_pressure_cell), hipMemcpyDeviceToHost, s));
    FL_HIP(hipMemcpyAsync(hf.data(), faces, nf * sizeof(fl_pressure_face), hipMemcpyDeviceToHost, s));
    if (ctx->nb) FL_HIP(hipMemcpyAsync(hb.data(), ctx->blocks, ctx->nb * sizeof(fl_int3), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    const fl_geometry g{{ctx->geometry.ox, ctx->geometry.oy, ctx->geometry.oz}, ctx->geometry.h};
    FL_TRY(prepare_system(ctx, P, &g, hb.data(), hb.size(), dt, hc.data(), nc, hf.data(), nf, true));
    std::vector<Key3> b0 = P.levels[0].keys;
    FL_TRY(build_levels(ctx, P, b0, opt.levels));
  }
  Gates gates{};
  EvalResult e0{};
  fl_pressure_diagnostics diag{};
  int st = freeze_and_assemble(ctx, P, dt, opt, gates, e0, diag);
  if (st == FL_OK) st = solve(ctx, P, opt, gates, e0, diag);
  if (out_diag) *out_diag = diag;
  if (st != FL_OK) return st;
  hipStream_t s = ctx->stream;
  k_face_velocity<<<grid_for(P.nf), kThreads, 0, s>>>(P.nf, P.face_tail, P.face_head, P.face_kind, P.face_area,
                                                      P.face_ustar, P.face_scale, P.face_lift, P.p, P.face_u);
  k_scatter_slots<<<grid_for(P.nf), kThreads, 0, s>>>(P.nf, P.face_slot, P.face_u, unew_slots);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// Pressures of the last device projection, in cell order (device array and count).
int last_pressures(fl_ctx* ctx, const double** p, uint64_t* n) {
  Pressure* P = state(ctx);
  if (!P) return FL_FAILED_PRECONDITION;
  *p = P->p;
  *n = P->nc;
  return FL_OK;
}

// ---------------------------------------------------------------- coupled viscous pressure support

namespace {
// z = the level-0 V-cycle's output on the unknown cells, zero on gauge cells.
__global__ void k_store_cells(uint64_t nc, const int32_t* cell_fine, const uint8_t* flag, const float* x, double* z) {
  const uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (c >= nc) return;
  z[c] = flag[c] == 1 ? static_cast<double>(x[cell_fine[c]]) : 0.0;
}
}  // namespace

void coupled_view(Pressure& P, CoupledSystem* out);

int coupled_prepare(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count, double dt,
                    const fl_pressure_cell* cells, size_t cell_count, const fl_pressure_face* faces,
                    size_t face_count, const fl_pressure_options* options, CoupledSystem* out) {
  FL_TRY(check_options(options));
  if (!ctx->pressure) ctx->pressure = new Pressure();
  Pressure& P = *state(ctx);
  P.scale_coarse = options->coarse_scale;
  FL_TRY(prepare_system(ctx, P, geometry, blocks, block_count, dt, cells, cell_count, faces, face_count, true));
  std::vector<Key3> b0 = P.levels[0].keys;
  FL_TRY(build_levels(ctx, P, b0, options->levels));
  Gates gates{};
  EvalResult e0{};
  fl_pressure_diagnostics diag{};
  FL_TRY(freeze_and_assemble(ctx, P, dt, *options, gates, e0, diag));
  coupled_view(P, out);
  return FL_OK;
}

int coupled_prepare_device(fl_ctx* ctx, const fl_pressure_cell* cells, uint64_t nc, const fl_pressure_face* faces,
                           uint64_t nf, double dt, const fl_pressure_options& opt, CoupledSystem* out) {
  FL_TRY(check_options(&opt));
  if (!ctx->pressure) ctx->pressure = new Pressure();
  Pressure& P = *state(ctx);
  P.scale_coarse = opt.coarse_scale;
  bool fallback = false;
  FL_TRY(prepare_system_device(ctx, P, cells, nc, faces, nf, dt, opt.levels, &fallback));
  if (fallback) {  // a component without an anchor: its compatibility test and gauge on the host
    hipStream_t s = ctx->stream;
    std::vector<fl_pressure_cell> hc(nc);
    std::vector<fl_pressure_face> hf(nf);
    std::vector<fl_int3> hb(ctx->nb);
    FL_HIP(hipMemcpyAsync(hc.data(), cells, nc * sizeof(fl_pressure_cell), hipMemcpyDeviceToHost, s));
    FL_HIP(hipMemcpyAsync(hf.data(), faces, nf * sizeof(fl_pressure_face), hipMemcpyDeviceToHost, s));
    if (ctx->nb) FL_HIP(hipMemcpyAsync(hb.data(), ctx->blocks, ctx->nb * sizeof(fl_int3), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    const fl_geometry g{{ctx->geometry.ox, ctx->geometry.oy, ctx->geometry.oz}, ctx->geometry.h};
    FL_TRY(prepare_system(ctx, P, &g, hb.data(), hb.size(), dt, hc.data(), nc, hf.data(), nf, true));
    std::vector<Key3> b0 = P.levels[0].keys;
    FL_TRY(build_levels(ctx, P, b0, opt.levels));
  }
  Gates gates{};
  EvalResult e0{};
  fl_pressure_diagnostics diag{};
  FL_TRY(freeze_and_assemble(ctx, P, dt, opt, gates, e0, diag));
  coupled_view(P, out);
  return FL_OK;
}

void coupled_view(Pressure& P, CoupledSystem* out) {
  CoupledSystem S{};
  S.nc = P.nc;
  S.nf = P.nf;
  S.tail = P.face_tail;
  S.head = P.face_head;
  S.kind = P.face_kind;
  S.slot = P.face_slot;
  S.area = P.face_area;
  S.scale = P.face_scale;
  S.conductance = P.face_k;
  S.ustar = P.face_ustar;
  S.wall = P.face_wall;
  S.lift = P.face_lift;
  S.uwall = P.face_uwall;
  S.has_flux = P.face_has_flux;
  S.cell_faces = P.cell_faces;
  S.flag = P.cell_flag;
  S.volume = P.cell_volume;
  S.target = P.cell_target;
  S.comp = P.cell_comp;
  S.prev = P.cell_prev;
  S.p_init = P.p_init;
  S.p_init_zero = P.p_init_zero;
  *out = S;
}

int coupled_coefficients(fl_ctx* ctx, const double* conductance) {
  Pressure* P = state(ctx);
  if (!P || P->levels.empty()) return FL_FAILED_PRECONDITION;
  return setup_coefficients(ctx, *P, conductance);
}

int coupled_vcycle(fl_ctx* ctx, const double* r, double* z, const fl_pressure_options& o) {
  Pressure* P = state(ctx);
  if (!P || P->levels.empty()) return FL_FAILED_PRECONDITION;
  hipStream_t s = ctx->stream;
  const uint64_t n = P->levels[0].nb * 64;
  k_load_residual<<<grid_for(n), kThreads, 0, s>>>(n, P->fine_cell, P->cell_flag, r, P->r32, P->x32, P->perr);
  FL_HIP(hipGetLastError());
  FL_TRY(vcycle<float>(ctx, *P, 0, o.smoothing_sweeps, o.coarse_sweeps, o.smoother));
  k_store_cells<<<grid_for(P->nc), kThreads, 0, s>>>(P->nc, P->cell_fine, P->cell_flag, P->levels[0].x, z);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int coupled_error(fl_ctx* ctx) {
  Pressure* P = state(ctx);
  if (!P) return FL_OK;
  FL_HIP(hipMemcpyAsync(ctx->host_words, P->perr, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < 2; ++i)
    if (ctx->host_words[i] != ~uint64_t(0)) return static_cast<int>(ctx->host_words[i] & 0xff);
  return FL_OK;
}

}  // namespace fl

// ---------------------------------------------------------------- device-resident liquid step

namespace {
using namespace fl;

__global__ void k_split(const double* in, uint64_t n, int stride, int first, int columns, double* base, uint64_t cap) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < columns; ++c) base[(first + c) * cap + i] = in[i * stride + c];
}
__global__ void k_fill_f64(double* p, uint64_t n, double v) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_pack_gather_rows(const double* cols, uint64_t cap, const uint32_t* perm, uint64_t n,
                                   fl_gather_result* out) {
  const uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (k >= n) return;
  double* r = reinterpret_cast<double*>(out + perm[k]);
  for (int c = 0; c < kGatherDoubles; ++c) r[c] = cols[c * cap + k];
}
__global__ void k_pack_quanta_rows(const int64_t* qm, const int64_t* qp, uint64_t n, fl_face_quanta* out) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  out[i].mass_quanta_sum = qm[i];
  out[i].pure_momentum_quanta_sum = qp[i];
  out[i].stress_force_quanta_sum = 0;
}

int first_error(fl_ctx* ctx) {
  uint64_t w;
  FL_TRY(read_error(ctx, &w));
  return w == ~uint64_t(0) ? FL_OK : static_cast<int>(w & 0xff);
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

}  // namespace

extern "C" {

int fl_scene_upload(fl_ctx* ctx, const fl_geometry* geometry, size_t n, const double* positions_xyz,
                    const double* velocities_xyz, const double* slopes_rowmajor, double particle_mass) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (!geometry || !(geometry->cell_width > 0.0 && std::isnormal(geometry->cell_width))) return FL_INVALID_ARGUMENT;
  if (n > ctx->cap.particles) return FL_RESOURCE_EXHAUSTED;
  (void)hipSetDevice(ctx->device);
  ctx->geometry = make_geometry(geometry->world_origin.x, geometry->world_origin.y, geometry->world_origin.z,
                                geometry->cell_width);
  hipStream_t s = ctx->stream;
  double* stage = reinterpret_cast<double*>(ctx->staging_aos);  // >= 19 doubles per particle
  const struct {
    const double* src;
    int stride, first;
  } parts[3] = {{positions_xyz, 3, 0}, {velocities_xyz, 3, 7}, {slopes_rowmajor, 9, 10}};
  for (const auto& pa : parts) {
    if (!n) break;
    FL_HIP(hipMemcpyAsync(stage, pa.src, n * pa.stride * sizeof(double), hipMemcpyHostToDevice, s));
    k_split<<<grid_for(n), kThreads, 0, s>>>(stage, n, pa.stride, pa.first, pa.stride, ctx->raw.base, ctx->raw.capacity);
    FL_HIP(hipGetLastError());
  }
  if (n) k_fill_f64<<<grid_for(n), kThreads, 0, s>>>(ctx->raw.base + 6 * ctx->raw.capacity, n, particle_mass);
  FL_HIP(hipGetLastError());
  FL_HIP(hipStreamSynchronize(s));
  ctx->raw.trial_aliases_sorted = true;
  ctx->scene_n = n;
  substep_scene_uploaded(ctx);
  ctx->scene_mass = particle_mass;
  ctx->scene_loaded = true;
  ctx->scene_ready = false;
  return FL_OK;
}

// Sort, activation, and the pressure structure of the scene. Faces of the liquid cells: both sides
// liquid -> internal (jump 0); liquid/air inside the domain -> Dirichlet p = 0 at distance h;
// on the domain boundary (lattice 0 or domain_cells along the normal) -> sealed, zero wall velocity.
int fl_scene_prepare(fl_ctx* ctx, const fl_pressure_cell* cells, size_t cell_count, int32_t domain_cells,
                     double density) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (!ctx->scene_loaded) return FL_FAILED_PRECONDITION;
  if (!(density > 0.0) || domain_cells <= 0) return FL_INVALID_ARGUMENT;
  if (cell_count > ctx->cap.pressure_cells) return FL_RESOURCE_EXHAUSTED;
  (void)hipSetDevice(ctx->device);
  const uint64_t n = ctx->scene_n;
  FL_TRY(bin_particles(ctx, ctx->raw.view(), n, true, false));
  FL_TRY(first_error(ctx));
  if (ctx->bin_box.volume() > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  FL_TRY(sort_particles(ctx, n, ctx->raw.view(), true));
  // activation box: particle windows plus cell owners
  int32_t lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = ctx->bin_box.lo[a];
    hi[a] = ctx->bin_box.lo[a] + ctx->bin_box.dim[a];
  }
  std::vector<int32_t> cxyz(3 * cell_count);
  for (size_t i = 0; i < cell_count; ++i) {
    const fl_int3 c = cells[i].cell_index;
    cxyz[3 * i] = c.x;
    cxyz[3 * i + 1] = c.y;
    cxyz[3 * i + 2] = c.z;
    const int32_t o[3] = {int32_t(floor_div4(c.x)), int32_t(floor_div4(c.y)), int32_t(floor_div4(c.z))};
    const int32_t p[3] = {int32_t(floor_div4(int64_t(c.x) + 1)), int32_t(floor_div4(int64_t(c.y) + 1)),
                          int32_t(floor_div4(int64_t(c.z) + 1))};
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], o[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
  Box box;
  for (int a = 0; a < 3; ++a) {
    box.lo[a] = lo[a];
    box.dim[a] = hi[a] - lo[a] + 1;
  }
  if (box.volume() > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  ctx->block_box = box;
  int32_t* d_cells = reinterpret_cast<int32_t*>(ctx->staging_aos);
  if (cell_count * 12 > 152 * ctx->raw.capacity) return FL_RESOURCE_EXHAUSTED;
  if (cell_count)
    FL_HIP(hipMemcpyAsync(d_cells, cxyz.data(), cxyz.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
  FL_TRY(activate_blocks(ctx, ctx->nvalid, d_cells, cell_count));
  FL_TRY(first_error(ctx));
  std::vector<fl_int3> blocks(ctx->nb);
  if (ctx->nb)
    FL_HIP(hipMemcpy(blocks.data(), ctx->blocks, ctx->nb * sizeof(fl_int3), hipMemcpyDeviceToHost));
  // faces of the liquid cells in owned-slot order
  std::unordered_map<Key3, int32_t, Key3Hash> grid;
  grid.reserve(2 * ctx->nb);
  for (size_t i = 0; i < ctx->nb; ++i) grid[{blocks[i].x, blocks[i].y, blocks[i].z}] = int32_t(i);
  const int64_t D = domain_cells;
  std::vector<uint8_t> liquid(size_t(D) * D * D, 0);
  for (size_t i = 0; i < cell_count; ++i) {
    const fl_int3 c = cells[i].cell_index;
    if (c.x < 0 || c.y < 0 || c.z < 0 || c.x >= D || c.y >= D || c.z >= D) return FL_INVALID_ARGUMENT;
    liquid[(size_t(c.z) * D + c.y) * D + c.x] = 1;
  }
  auto is_liquid = [&](int64_t x, int64_t y, int64_t z) {
    return x >= 0 && y >= 0 && z >= 0 && x < D && y < D && z < D && liquid[(size_t(z) * D + y) * D + x];
  };
  std::vector<std::pair<uint64_t, fl_pressure_face>> recs;
  recs.reserve(4 * cell_count);
  const double h = ctx->geometry.h;
  for (size_t i = 0; i < cell_count; ++i) {
    const fl_int3 c = cells[i].cell_index;
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side) {
        int64_t l[3] = {c.x, c.y, c.z};
        l[a] += side;
        // Each face is emitted once: by its head cell, or by its tail when the head is not liquid.
        int64_t t[3] = {l[0], l[1], l[2]};
        t[a] -= 1;
        const bool head_liquid = is_liquid(l[0], l[1], l[2]);
        if (side == 1 && head_liquid) continue;
        const Key3 o{int32_t(floor_div4(l[0])), int32_t(floor_div4(l[1])), int32_t(floor_div4(l[2]))};
        auto it = grid.find(o);
        if (it == grid.end()) return FL_INTERNAL;
        const uint64_t slot = uint64_t(it->second) * 192 + uint64_t(a) * 64 +
                              uint64_t(((l[2] - 4 * o.z) * 4 + (l[1] - 4 * o.y)) * 4 + (l[0] - 4 * o.x));
        fl_pressure_face f{};
        f.axis = a;
        f.lattice = {int32_t(l[0]), int32_t(l[1]), int32_t(l[2])};
        f.inverse_density = 1.0 / density;
        f.pressure_gradient_distance = h;
        const bool wall = l[a] == 0 || l[a] == D;
        const bool both = head_liquid && is_liquid(t[0], t[1], t[2]);
        f.link_kind = wall ? 0 : (both ? 1 : 2);
        f.open_area = wall ? 0.0 : h * h;
        recs.push_back({slot, f});
      }
  }
  std::sort(recs.begin(), recs.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
  std::vector<fl_pressure_face> faces(recs.size());
  for (size_t i = 0; i < recs.size(); ++i) faces[i] = recs[i].second;
  if (!ctx->pressure) ctx->pressure = new Pressure();
  Pressure& P = *static_cast<Pressure*>(ctx->pressure);
  fl_geometry g{{ctx->geometry.ox, ctx->geometry.oy, ctx->geometry.oz}, h};
  // dt only scales conductances here; a step freezes with its own dt (a repeated dt reuses the coefficients).
  FL_TRY(prepare_system(ctx, P, &g, blocks.data(), blocks.size(), 1.0, cells, cell_count, faces.data(), faces.size(),
                        false));
  ctx->scene_ready = true;
  return FL_OK;
}

int fl_time_projection_kernels(fl_ctx* ctx, int reps, fl_kernel_time* out, size_t capacity, size_t* count) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (reps < 1 || !count || (capacity && !out)) return FL_INVALID_ARGUMENT;
  return fl::time_projection_kernels(ctx, reps, out, capacity, count);
}

int fl_step(fl_ctx* ctx, double dt, fl_vec3 gravity, fl_quantization quantization, const fl_pressure_options* options,
            fl_step_report* report) {
  if (!ctx || ctx->poisoned || !ctx->scene_ready || !ctx->pressure) return FL_FAILED_PRECONDITION;
  if (!report) return FL_INVALID_ARGUMENT;
  FL_TRY(check_options(options));
  if (!(dt > 0.0) || !std::isfinite(dt)) return FL_INVALID_ARGUMENT;
  int e;
  if (!std::isnormal(quantization.mass_per_quantum) || std::frexp(quantization.mass_per_quantum, &e) != 0.5 ||
      !std::isnormal(quantization.pure_momentum_per_quantum) ||
      std::frexp(quantization.pure_momentum_per_quantum, &e) != 0.5)
    return FL_INVALID_ARGUMENT;
  if (quantization.declared_particle_capacity > ctx->cap.particles || ctx->scene_n > quantization.declared_particle_capacity)
    return FL_RESOURCE_EXHAUSTED;
  (void)hipSetDevice(ctx->device);
  Pressure& P = *static_cast<Pressure*>(ctx->pressure);
  if (P.levels_requested != options->levels || P.scale_coarse != options->coarse_scale) {
    P.scale_coarse = options->coarse_scale;
    std::vector<Key3> b0 = P.levels[0].keys;
    FL_TRY(build_levels(ctx, P, b0, options->levels));
  }
  hipStream_t s = ctx->stream;
  hipEvent_t* ev = ctx->stage;
  *report = fl_step_report{};
  const uint64_t n = ctx->scene_n;
  const uint64_t cap = std::max<uint64_t>(1, quantization.declared_particle_capacity);
  int bits = 0;
  for (uint64_t v = cap - 1; v; v >>= 1) ++bits;
  const double limit = std::ldexp(1.0, 62 - bits);
  FL_HIP(hipEventRecord(ev[0], s));
  FL_TRY(bin_particles(ctx, ctx->raw.view(), n, true, false));
  FL_TRY(sort_particles(ctx, n, ctx->raw.view(), true));
  FL_HIP(hipEventRecord(ev[1], s));
  FL_TRY(p2g_quanta(ctx, ctx->nvalid, 1.0 / quantization.mass_per_quantum, 1.0 / quantization.pure_momentum_per_quantum,
                    limit, ev[6], ev[7]));
  FL_HIP(hipEventRecord(ev[2], s));
  const double g3[3] = {gravity.x, gravity.y, gravity.z};
  const uint64_t nf_grid = ctx->nb * 192;
  FL_TRY(normalize_forces(ctx, nf_grid, quantization.mass_per_quantum, quantization.pure_momentum_per_quantum, 0.0, g3,
                          dt));
  FL_HIP(hipEventRecord(ev[3], s));
  FL_TRY(first_error(ctx));
  // projection: u* at the pressure faces, freeze for dt, solve, corrected velocities into u_new
  k_gather_slots<<<grid_for(P.nf), kThreads, 0, s>>>(P.nf, P.face_slot, ctx->ustar, P.face_ustar);
  FL_HIP(hipGetLastError());
  Gates gates{};
  EvalResult e0{};
  fl_pressure_diagnostics diag{};
  int st = freeze_and_assemble(ctx, P, dt, *options, gates, e0, diag, true);
  if (st == FL_OK) st = solve(ctx, P, *options, gates, e0, diag);
  report->pressure = diag;
  if (st != FL_OK) {
    report->status = st;
    return st;
  }
  k_face_velocity<<<grid_for(P.nf), kThreads, 0, s>>>(P.nf, P.face_tail, P.face_head, P.face_kind, P.face_area,
                                                      P.face_ustar, P.face_scale, P.face_lift, P.p, P.face_u);
  k_copy_f64<<<grid_for(nf_grid), kThreads, 0, s>>>(nf_grid, ctx->ustar, ctx->unew);
  k_scatter_slots<<<grid_for(P.nf), kThreads, 0, s>>>(P.nf, P.face_slot, P.face_u, ctx->unew);
  FL_HIP(hipGetLastError());
  FL_HIP(hipEventRecord(ev[4], s));
  const float p2g_kernel = elapsed(ev[6], ev[7]);
  FL_TRY(g2p(ctx, ctx->nvalid, ctx->u0, ctx->unew, ev[6], ev[7]));
  FL_HIP(hipEventRecord(ev[5], s));
  FL_HIP(hipEventSynchronize(ev[5]));
  st = first_error(ctx);
  report->p2g_kernel_ms = p2g_kernel;
  report->g2p_kernel_ms = elapsed(ev[6], ev[7]);
  report->sort_ms = elapsed(ev[0], ev[1]);
  report->p2g_ms = elapsed(ev[1], ev[2]);
  report->normalize_ms = elapsed(ev[2], ev[3]);
  report->projection_ms = elapsed(ev[3], ev[4]);
  report->g2p_ms = elapsed(ev[4], ev[5]);
  report->total_ms = elapsed(ev[0], ev[5]);
  report->particles = n;
  report->blocks = ctx->nb;
  report->faces = P.nf;
  report->pressure_cells = P.nc;
  report->nonempty_bins = ctx->nbins;
  report->status = st;
  return st;
}

int fl_scene_download(fl_ctx* ctx, fl_int3* blocks, size_t block_capacity, fl_face_quanta* quanta, double* u0,
                      double* ustar, double* unew, fl_sort_entry* order, fl_gather_result* gather, double* pressures) {
  if (!ctx || ctx->poisoned || !ctx->scene_ready) return FL_FAILED_PRECONDITION;
  (void)hipSetDevice(ctx->device);
  hipStream_t s = ctx->stream;
  const uint64_t nb = ctx->nb, nf = nb * 192, n = ctx->nvalid;
  if (blocks) {
    if (block_capacity < nb) return FL_RESOURCE_EXHAUSTED;
    FL_HIP(hipMemcpyAsync(blocks, ctx->blocks, nb * sizeof(fl_int3), hipMemcpyDeviceToHost, s));
  }
  if (quanta) {
    k_pack_quanta_rows<<<grid_for(nf), kThreads, 0, s>>>(ctx->qm, ctx->qp, nf, ctx->quanta_aos);
    FL_HIP(hipMemcpyAsync(quanta, ctx->quanta_aos, nf * sizeof(fl_face_quanta), hipMemcpyDeviceToHost, s));
  }
  if (u0) FL_HIP(hipMemcpyAsync(u0, ctx->u0, nf * sizeof(double), hipMemcpyDeviceToHost, s));
  if (ustar) FL_HIP(hipMemcpyAsync(ustar, ctx->ustar, nf * sizeof(double), hipMemcpyDeviceToHost, s));
  if (unew) FL_HIP(hipMemcpyAsync(unew, ctx->unew, nf * sizeof(double), hipMemcpyDeviceToHost, s));
  if (order) {
    auto* d = reinterpret_cast<fl_sort_entry*>(ctx->staging_aos);
    FL_TRY(write_sort_entries(ctx, n, d));
    FL_HIP(hipMemcpyAsync(order, d, n * sizeof(fl_sort_entry), hipMemcpyDeviceToHost, s));
  }
  if (gather) {
    k_pack_gather_rows<<<grid_for(n), kThreads, 0, s>>>(ctx->gather, ctx->sorted.capacity, ctx->perm, n,
                                                       ctx->gather_aos);
    FL_HIP(hipMemcpyAsync(gather, ctx->gather_aos, n * sizeof(fl_gather_result), hipMemcpyDeviceToHost, s));
  }
  if (pressures && ctx->pressure) {
    Pressure& P = *static_cast<Pressure*>(ctx->pressure);
    FL_HIP(hipMemcpyAsync(pressures, P.p, P.nc * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  FL_HIP(hipGetLastError());
  FL_HIP(hipStreamSynchronize(s));
  return FL_OK;
}

}  // extern "C"
