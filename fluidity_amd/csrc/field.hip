// The particle-union liquid field of one substep on the device: PrepareField's candidate cells and
// sample lattice (native/src/liquid_substep_3d_staging.cc:759-916) and the signed distances of
// LiquidParticleSphereUnion3D at every sample (FieldDistance, liquid_substep_3d_staging.cc:210-229,
// sampled through SampleWorldBatch, native/src/liquid_substep_3d_cpu.cc:77-120).
//
// Candidates are the cells of every particle's box floor((x -/+ r - o) / h) per axis; centres are
// the 27 cells around each candidate and corners the 8 lattice corners of each candidate, each set
// unique and (z,y,x) sorted. All three are bit sets over one cell box whose rows are padded to whole
// 64-bit words, so a row's x run is a run of words: the candidate bits are set with one atomic OR per
// box row a particle touches, the centre and corner sets are word-wise dilations of the candidate
// words, and a popcount scan of the words writes each list in (z,y,x) order.
//
// A sample's distance is min over all particles of hypot(sample - particle) - r. The particles of
// each sort bin are regrouped by unit sub-cell (floor(X - 1.5) - 4 B, the bin's own 4^3 cells of the
// lattice shifted by 1.5), and a sample visits the sub-cells within rho cells of it, doubling rho until
// the nearest squared distance found is certainly below every unvisited particle's. The minimum of
// the libstdc++ three-argument hypot is then taken over the visited particles whose squared distance
// is within a relative 1e-10 of the least one: hypot is within an ulp of the exact distance, so no
// other particle can hold the minimum, and the result equals the reference's scan over all
// particles bit for bit (min is exact and order-free).
#include "context.h"
#include "stdmath.h"

namespace fl {

struct FieldBuffers {
  uint64_t* cand = nullptr;     // candidate bits
  uint64_t* centre = nullptr;   // centre-sample bits
  uint64_t* corner = nullptr;   // corner-sample bits
  uint64_t* counts = nullptr;   // 3 * words popcounts
  uint64_t* scans = nullptr;    // 3 * words exclusive scans
  uint64_t* tot = nullptr;      // the three set sizes
  uint64_t cap_words = 0;
  fl_int3* cells = nullptr;     // candidates
  fl_center_distance* centres = nullptr;
  fl_center_distance* corners = nullptr;
  uint64_t cap_cells = 0, cap_centres = 0, cap_corners = 0;
  uint2* sub_range = nullptr;    // 64 per bin of the bin box: [first, last) particles of each sub-cell
  uint64_t cap_sub = 0;
  double* sub_pos = nullptr;    // 4 doubles per particle: x, y, z of the particles in sub-cell order
  uint64_t cap_pos = 0;
  void release() {
    void* ptrs[] = {cand, centre, corner, counts, scans, tot, cells, centres, corners, sub_range, sub_pos};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    *this = FieldBuffers{};
  }
};

namespace {

template <typename T>
int grow_buf(T** p, uint64_t* cap, uint64_t need) {
  if (need <= *cap && *p) return FL_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  const uint64_t n = need < 1 ? 1 : need;
  if (hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return FL_RESOURCE_EXHAUSTED;
  }
  *cap = n;
  return FL_OK;
}

// The bit-set box: cells lo + (x, y, z), rows of `row_words` words.
struct BitBox {
  int32_t lo[3];
  int32_t dim[3];
  int32_t row_words;
  __host__ __device__ uint64_t words() const { return uint64_t(dim[2]) * dim[1] * row_words; }
};

__global__ void k_zero_words(uint64_t* w, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    w[i] = 0;
}

// add_sphere of PrepareField (liquid_substep_3d_staging.cc:785-824) for every particle: the cells of
// its box per axis, floor((c - r - o) / h) .. floor((c + r - o) / h), with the reference's range
// check (OutOfRange, reported at the particle's ordinal in the caller's order).
__global__ void k_mark_candidates(const double* sx, const double* sy, const double* sz, const uint32_t* perm,
                                  uint64_t n, double ox, double oy, double oz, double h, double r, BitBox box,
                                  uint64_t* bits, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const double c[3] = {sx[i], sy[i], sz[i]}, o[3] = {ox, oy, oz};
  int32_t lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    const double low = floor(((c[a] - r) - o[a]) / h);
    const double high = floor(((c[a] + r) - o[a]) / h);
    if (!finite(low) || !finite(high) || low < double(kI32Min) + 1 || high > double(kI32Max) - 1) {
      atomicMin(reinterpret_cast<unsigned long long*>(err),
                static_cast<unsigned long long>(error_word(perm ? perm[i] : i, kOutOfRange)));
      return;
    }
    lo[a] = static_cast<int32_t>(low) - box.lo[a];
    hi[a] = static_cast<int32_t>(high) - box.lo[a];
    if (lo[a] < 0 || hi[a] >= box.dim[a]) {  // the box is built to hold every particle's cells
      atomicMin(reinterpret_cast<unsigned long long*>(err),
                static_cast<unsigned long long>(error_word(perm ? perm[i] : i, kInternal)));
      return;
    }
  }
  for (int32_t z = lo[2]; z <= hi[2]; ++z)
    for (int32_t y = lo[1]; y <= hi[1]; ++y) {
      uint64_t* row = bits + (uint64_t(z) * box.dim[1] + y) * box.row_words;
      int32_t x = lo[0];
      while (x <= hi[0]) {
        const int32_t w = x >> 6, b = x & 63;
        const int32_t last = min(hi[0], (w << 6) + 63);
        const int32_t len = last - x + 1;
        const uint64_t mask = (len == 64 ? ~uint64_t(0) : ((uint64_t(1) << len) - 1)) << b;
        atomicOr(reinterpret_cast<unsigned long long*>(row + w), static_cast<unsigned long long>(mask));
        x = last + 1;
      }
    }
}

// The same marks, one workgroup per non-empty bin: the bin's particles set their cells in an LDS bit
// window of W^3 cells (a bin holds lattice coordinates [4B + 1.5, 4B + 5.5), so with R = ceil(r / h)
// every box lies in [4B - R, 4B + 6 + R] on each axis, one cell of slack included), then each window
// row is ORed into the global bits with at most two atomics. A particle whose box leaves the window
// (never, short of rounding beyond the slack) marks its cells directly.
constexpr int kMaxWindow = 15;
__global__ __launch_bounds__(256) void k_mark_candidates_bins(const double* sx, const double* sy, const double* sz,
                                                              const uint32_t* perm, const uint32_t* bkey,
                                                              const uint32_t* boff, const uint32_t* bn, Box bins,
                                                              int32_t reach, double ox, double oy, double oz, double h,
                                                              double r, BitBox box, uint64_t* bits, uint64_t* err) {
  __shared__ unsigned int win[(kMaxWindow * kMaxWindow * kMaxWindow + 31) / 32];
  const int W = 7 + 2 * reach;
  const int nwords = (W * W * W + 31) / 32;
  const uint32_t key = bkey[blockIdx.x], off = boff[blockIdx.x], cnt = bn[blockIdx.x];
  const int32_t B[3] = {bins.lo[0] + static_cast<int32_t>(key % bins.dim[0]),
                        bins.lo[1] + static_cast<int32_t>((key / bins.dim[0]) % bins.dim[1]),
                        bins.lo[2] + static_cast<int32_t>(key / (uint32_t(bins.dim[0]) * bins.dim[1]))};
  const int32_t w0[3] = {4 * B[0] - reach, 4 * B[1] - reach, 4 * B[2] - reach};
  for (int k = threadIdx.x; k < nwords; k += blockDim.x) win[k] = 0;
  __syncthreads();
  const double o[3] = {ox, oy, oz};
  for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) {
    const uint32_t i = off + k;
    const double c[3] = {sx[i], sy[i], sz[i]};
    int32_t lo[3], hi[3];
    bool ok = true, inside = true;
    for (int a = 0; a < 3; ++a) {
      const double low = floor(((c[a] - r) - o[a]) / h);
      const double high = floor(((c[a] + r) - o[a]) / h);
      if (!finite(low) || !finite(high) || low < double(kI32Min) + 1 || high > double(kI32Max) - 1) {
        atomicMin(reinterpret_cast<unsigned long long*>(err),
                  static_cast<unsigned long long>(error_word(perm[i], kOutOfRange)));
        ok = false;
        break;
      }
      lo[a] = static_cast<int32_t>(low);
      hi[a] = static_cast<int32_t>(high);
      inside = inside && lo[a] >= w0[a] && hi[a] < w0[a] + W;
      if (lo[a] - box.lo[a] < 0 || hi[a] - box.lo[a] >= box.dim[a]) {
        atomicMin(reinterpret_cast<unsigned long long*>(err),
                  static_cast<unsigned long long>(error_word(perm[i], kInternal)));
        ok = false;
        break;
      }
    }
    if (!ok) continue;
    if (inside) {
      for (int32_t z = lo[2]; z <= hi[2]; ++z)
        for (int32_t y = lo[1]; y <= hi[1]; ++y)
          for (int32_t x = lo[0]; x <= hi[0]; ++x) {
            const int b = ((z - w0[2]) * W + (y - w0[1])) * W + (x - w0[0]);
            atomicOr(&win[b >> 5], 1u << (b & 31));
          }
    } else {
      for (int32_t z = lo[2]; z <= hi[2]; ++z)
        for (int32_t y = lo[1]; y <= hi[1]; ++y) {
          uint64_t* row = bits + (uint64_t(z - box.lo[2]) * box.dim[1] + (y - box.lo[1])) * box.row_words;
          for (int32_t x = lo[0]; x <= hi[0]; ++x) {
            const int32_t bx = x - box.lo[0];
            atomicOr(reinterpret_cast<unsigned long long*>(row + (bx >> 6)),
                     static_cast<unsigned long long>(uint64_t(1) << (bx & 63)));
          }
        }
    }
  }
  __syncthreads();
  for (int row = threadIdx.x; row < W * W; row += blockDim.x) {
    uint32_t m = 0;  // W <= 15 bits of this window row
    const int first = row * W;
    for (int x = 0; x < W; ++x) {
      const int b = first + x;
      m |= ((win[b >> 5] >> (b & 31)) & 1u) << x;
    }
    if (!m) continue;
    const int32_t gz = w0[2] + row / W, gy = w0[1] + row % W;
    if (gz - box.lo[2] < 0 || gz - box.lo[2] >= box.dim[2] || gy - box.lo[1] < 0 || gy - box.lo[1] >= box.dim[1])
      continue;  // only empty rows can fall outside (every set bit came from a checked box)
    uint64_t* grow_ = bits + (uint64_t(gz - box.lo[2]) * box.dim[1] + (gy - box.lo[1])) * box.row_words;
    const int32_t bx = w0[0] - box.lo[0];  // >= 0 for any row with a set bit
    const uint64_t wide = uint64_t(m);
    const int32_t word = bx >> 6, shift = bx & 63;
    atomicOr(reinterpret_cast<unsigned long long*>(grow_ + word), static_cast<unsigned long long>(wide << shift));
    if (shift && (wide >> (64 - shift)))
      atomicOr(reinterpret_cast<unsigned long long*>(grow_ + word + 1),
               static_cast<unsigned long long>(wide >> (64 - shift)));
  }
}

// Centre and corner sets from the candidate bits, one output word per lane. Centres: a cell within
// one cell of a candidate on every axis; corners: corner (x, y, z) of candidate (x - i, y - j, z - k),
// i, j, k in {0, 1}. The box keeps two empty cells around every candidate, so nothing spills out.
__global__ void k_dilate(const uint64_t* cand, BitBox box, uint64_t* centre, uint64_t* corner) {
  const uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (t >= box.words()) return;
  const int32_t wx = static_cast<int32_t>(t % box.row_words);
  const uint64_t row = t / box.row_words;
  const int32_t y = static_cast<int32_t>(row % box.dim[1]), z = static_cast<int32_t>(row / box.dim[1]);
  uint64_t c = 0, k = 0;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy) {
      const int32_t yy = y + dy, zz = z + dz;
      if (yy < 0 || zz < 0 || yy >= box.dim[1] || zz >= box.dim[2]) continue;
      const uint64_t* r = cand + (uint64_t(zz) * box.dim[1] + yy) * box.row_words;
      const uint64_t w = r[wx];
      const uint64_t prev = wx > 0 ? r[wx - 1] : 0, next = wx + 1 < box.row_words ? r[wx + 1] : 0;
      const uint64_t up = (w << 1) | (prev >> 63);    // bit x set when x - 1 is a candidate
      const uint64_t down = (w >> 1) | (next << 63);  // bit x set when x + 1 is a candidate
      c |= w | up | down;
      if (dz <= 0 && dy <= 0) k |= w | up;
    }
  centre[t] = c;
  corner[t] = k;
}

__global__ void k_popcounts(const uint64_t* a, const uint64_t* b, const uint64_t* c, uint64_t n, uint64_t* out) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  out[i] = __popcll(a[i]);
  out[n + i] = __popcll(b[i]);
  out[2 * n + i] = __popcll(c[i]);
}

// Totals of the three scans into tot[0..2].
__global__ void k_totals(const uint64_t* counts, const uint64_t* scans, uint64_t n, uint64_t* tot) {
  if (threadIdx.x < 3) tot[threadIdx.x] = n ? scans[threadIdx.x * n + n - 1] + counts[threadIdx.x * n + n - 1] : 0;
}

// Writes the set bits of each word in (z,y,x) order: candidates as cells, samples as records with
// the distance still unset.
__global__ void k_compact(const uint64_t* cand, const uint64_t* centre, const uint64_t* corner, const uint64_t* scans,
                          BitBox box, fl_int3* cells, fl_center_distance* centres, fl_center_distance* corners) {
  const uint64_t n = box.words();
  const uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  const int32_t wx = static_cast<int32_t>(t % box.row_words);
  const uint64_t row = t / box.row_words;
  const int32_t y = static_cast<int32_t>(row % box.dim[1]) + box.lo[1];
  const int32_t z = static_cast<int32_t>(row / box.dim[1]) + box.lo[2];
  const uint64_t words[3] = {cand[t], centre[t], corner[t]};
  for (int s = 0; s < 3; ++s) {
    uint64_t w = words[s];
    uint64_t pos = scans[s * n + t];
    while (w) {
      const int b = __ffsll(static_cast<long long>(w)) - 1;
      w &= w - 1;
      const int32_t x = box.lo[0] + (wx << 6) + b;
      if (s == 0) {
        cells[pos] = fl_int3{x, y, z};
      } else {
        fl_center_distance& d = (s == 1 ? centres : corners)[pos];
        d.cell_index = fl_int3{x, y, z};
        d.pad = 0;
        d.signed_distance = 0;
      }
      ++pos;
    }
  }
}

__device__ inline int32_t floor_i32(double v) { return static_cast<int32_t>(floor(v)); }

// Regroups the particles of every non-empty bin by unit sub-cell: sub-cell u = floor(X - 1.5) - 4 B
// on each axis with X = (x - o) / h (clamped into the bin; the search below allows for the rounding),
// and writes each sub-cell's particle range into the dense table over the bin box.
__global__ __launch_bounds__(256) void k_subcells(const double* sx, const double* sy, const double* sz,
                                                  const uint32_t* bkey, const uint32_t* boff, const uint32_t* bn,
                                                  Box bins, double ox, double oy, double oz, double h,
                                                  uint2* sub_range, double* sub_pos) {
  __shared__ uint32_t count[64], start[65], fill[64];
  const uint32_t b = blockIdx.x;
  const uint32_t key = bkey[b], off = boff[b], cnt = bn[b];
  const int32_t bx = bins.lo[0] + static_cast<int32_t>(key % bins.dim[0]);
  const int32_t by = bins.lo[1] + static_cast<int32_t>((key / bins.dim[0]) % bins.dim[1]);
  const int32_t bz = bins.lo[2] + static_cast<int32_t>(key / (uint32_t(bins.dim[0]) * bins.dim[1]));
  if (threadIdx.x < 64) {
    count[threadIdx.x] = 0;
    fill[threadIdx.x] = 0;
  }
  __syncthreads();
  auto sub = [&](uint32_t i) {
    int u[3];
    const double X[3] = {(sx[i] - ox) / h, (sy[i] - oy) / h, (sz[i] - oz) / h};
    const int32_t B[3] = {bx, by, bz};
    for (int a = 0; a < 3; ++a) {
      const int32_t v = floor_i32(X[a] - 1.5) - 4 * B[a];
      u[a] = v < 0 ? 0 : (v > 3 ? 3 : v);
    }
    return (u[2] * 4 + u[1]) * 4 + u[0];
  };
  for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) atomicAdd(&count[sub(off + k)], 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int q = 0; q < 64; ++q) {
      start[q] = run;
      run += count[q];
    }
    start[64] = run;
  }
  __syncthreads();
  if (threadIdx.x < 64)
    sub_range[uint64_t(key) * 64 + threadIdx.x] = make_uint2(off + start[threadIdx.x], off + start[threadIdx.x + 1]);
  for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) {
    const uint32_t i = off + k;
    const int q = sub(i);
    const uint32_t pos = off + start[q] + atomicAdd(&fill[q], 1u);
    sub_pos[4 * uint64_t(pos)] = sx[i];
    sub_pos[4 * uint64_t(pos) + 1] = sy[i];
    sub_pos[4 * uint64_t(pos) + 2] = sz[i];
  }
}

__global__ void k_zero_ranges(uint2* r, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    r[i] = make_uint2(0, 0);
}

struct SubGrid {
  const uint2* range;  // 64 per bin of the bin box
  const double* pos;
  Box bins;
};

// FieldDistance of LiquidParticleSphereUnion3D at every sample (centres at index + 0.5, corners at
// the index), world position origin + lattice * h as the reference rounds it. One pass keeps the least
// squared distance d2 and the least hypot among the particles whose d2 is within a relative 1e-10 of
// it: a particle more than that above the least d2 is farther than the hypot minimum by far more than
// hypot's rounding. Sub-cells are visited in rings of growing radius rho (cells); within a ring a
// sub-cell whose box is farther than the current least distance is skipped, and the search ends when
// every unvisited sub-cell lies beyond the least distance.
__global__ void k_field_distance(fl_center_distance* samples, uint64_t n, int corner, SubGrid G, double ox, double oy,
                                 double oz, double h, double r, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const fl_int3 c = samples[i].cell_index;
  const double half = corner ? 0.0 : 0.5;
  const double L[3] = {c.x + half, c.y + half, c.z + half};
  const double w[3] = {ox + L[0] * h, oy + L[1] * h, oz + L[2] * h};
  const double h2 = h * h;
  double best = __builtin_inf(), least = __builtin_inf();
  int32_t blo[3], bhi[3];
  for (int a = 0; a < 3; ++a) {
    blo[a] = 4 * G.bins.lo[a];
    bhi[a] = 4 * (G.bins.lo[a] + G.bins.dim[a]) - 1;
  }
  // Lattice gap between the sample and sub-cell u on one axis (u covers [u + 1.5, u + 2.5)), less a
  // slack for particles whose computed sub-cell rounding moved across a face.
  auto gap = [&](int a, int32_t u) {
    const double g = fmax(fmax((u + 1.5) - L[a], L[a] - (u + 2.5)), 0.0);
    return fmax(g - 1e-6, 0.0);
  };
  auto visit = [&](int32_t ux, int32_t uy, int32_t uz) {
    const int32_t b[3] = {static_cast<int32_t>(floor_div4(ux)), static_cast<int32_t>(floor_div4(uy)),
                          static_cast<int32_t>(floor_div4(uz))};
    const int64_t bi = G.bins.index(b[0], b[1], b[2]);
    if (bi < 0) return;
    const int q = ((uz - 4 * b[2]) * 4 + (uy - 4 * b[1])) * 4 + (ux - 4 * b[0]);
    const uint2 rg = G.range[uint64_t(bi) * 64 + q];
    for (uint32_t k = rg.x; k < rg.y; ++k) {
      const double* p = G.pos + 4 * uint64_t(k);
      const double dx = w[0] - p[0], dy = w[1] - p[1], dz = w[2] - p[2];
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < best * (1 - 1e-10)) {
        best = d2;
        least = hypot3(dx, dy, dz);
      } else if (d2 <= best * (1 + 1e-10)) {
        best = fmin(best, d2);
        least = fmin(least, hypot3(dx, dy, dz));
      }
    }
  };
  int32_t plo[3] = {0, 0, 0}, phi[3] = {-1, -1, -1};  // the previous ring's cube (empty at first)
  for (double rho = 1;; rho *= 2) {
    int32_t lo[3], hi[3];
    bool all = true;
    for (int a = 0; a < 3; ++a) {
      const double l = floor(L[a] - rho - 1.5), u = floor(L[a] + rho - 1.5);
      lo[a] = l <= blo[a] ? blo[a] : static_cast<int32_t>(fmin(l, double(bhi[a]) + 1));
      hi[a] = u >= bhi[a] ? bhi[a] : static_cast<int32_t>(fmax(u, double(blo[a]) - 1));
      all = all && l <= blo[a] && u >= bhi[a];
    }
    // Sub-cells touching the sample first (they hold the nearest particle of a liquid sample), then
    // the rest of the ring, each skipped when its box is farther than the least distance so far.
    for (int pass = 0; pass < 2; ++pass)
      for (int32_t uz = lo[2]; uz <= hi[2]; ++uz) {
        const double gz = gap(2, uz);
        for (int32_t uy = lo[1]; uy <= hi[1]; ++uy) {
          const double gy = gap(1, uy);
          for (int32_t ux = lo[0]; ux <= hi[0]; ++ux) {
            const bool old = ux >= plo[0] && ux <= phi[0] && uy >= plo[1] && uy <= phi[1] && uz >= plo[2] &&
                             uz <= phi[2];
            if (old) continue;
            const double gx = gap(0, ux);
            const double g2 = gx * gx + gy * gy + gz * gz;
            if ((g2 == 0) != (pass == 0)) continue;
            if (g2 * h2 > best * (1 + 1e-9)) continue;
            visit(ux, uy, uz);
          }
        }
      }
    for (int a = 0; a < 3; ++a) {
      plo[a] = lo[a];
      phi[a] = hi[a];
    }
    // Every unvisited particle lies at least (rho - rounding) cells from the sample.
    if (finite(best) && sqrt(best) * (1 + 1e-12) < (rho - 1e-6) * h) break;
    if (all) break;
  }
  const double d = least - r;
  if (!finite(d)) {
    atomicMin(reinterpret_cast<unsigned long long*>(err), static_cast<unsigned long long>(error_word(i, kOutOfRange)));
    return;
  }
  samples[i].signed_distance = d;
}

}  // namespace

void release_field(fl_ctx* ctx) {
  auto* B = static_cast<FieldBuffers*>(ctx->field);
  if (!B) return;
  B->release();
  delete B;
  ctx->field = nullptr;
}

// The particle-union field of the sorted particles (ctx->sorted, ctx->nvalid of them, binned and
// sorted into ctx's bin list): candidates, centre and corner samples with their distances, on the
// device. Refuses more candidates than max_candidates (ResourceExhausted).
int sample_particle_field(fl_ctx* ctx, double radius, uint64_t max_candidates, FieldSamples* out) {
  if (!(radius > 0) || !finite(radius)) return FL_INVALID_ARGUMENT;
  if (!ctx->field) ctx->field = new FieldBuffers();
  FieldBuffers& B = *static_cast<FieldBuffers*>(ctx->field);
  hipStream_t s = ctx->stream;
  const uint64_t n = ctx->nvalid;
  *out = FieldSamples{};
  if (n == 0) return FL_OK;
  const Geometry g = ctx->geometry;
  const double h = g.h;
  // Bin B holds lattice coordinates [4B + 1.5, 4B + 5.5); a particle's cells reach r / h beyond, and
  // the sample sets one cell further; two more cells keep the dilation inside the box.
  const double reach = std::ceil(radius / h);
  if (!(reach < 1e6)) return FL_OUT_OF_RANGE;
  const int32_t m = static_cast<int32_t>(reach) + 4;
  BitBox box{};
  for (int a = 0; a < 3; ++a) {
    const int64_t lo = 4 * int64_t(ctx->bin_box.lo[a]) + 1 - m;
    const int64_t hi = 4 * (int64_t(ctx->bin_box.lo[a]) + ctx->bin_box.dim[a] - 1) + 6 + m;
    if (lo < kI32Min + 2 || hi > kI32Max - 2) return FL_OUT_OF_RANGE;
    box.lo[a] = static_cast<int32_t>(lo);
    box.dim[a] = static_cast<int32_t>(hi - lo + 1);
  }
  box.row_words = (box.dim[0] + 63) / 64;
  const uint64_t nw = box.words();
  if (nw > 64 * std::max<uint64_t>(ctx->cap.lookup_blocks, 64)) return FL_RESOURCE_EXHAUSTED;
  if (nw > B.cap_words || !B.cand) {
    // The three bit sets, their popcounts and scans grow together.
    for (uint64_t** p : {&B.cand, &B.centre, &B.corner, &B.counts, &B.scans})
      if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
      }
    B.cap_words = 0;
    for (uint64_t** p : {&B.cand, &B.centre, &B.corner})
      if (hipMalloc(reinterpret_cast<void**>(p), nw * sizeof(uint64_t)) != hipSuccess) return FL_RESOURCE_EXHAUSTED;
    for (uint64_t** p : {&B.counts, &B.scans})
      if (hipMalloc(reinterpret_cast<void**>(p), 3 * nw * sizeof(uint64_t)) != hipSuccess) return FL_RESOURCE_EXHAUSTED;
    B.cap_words = nw;
  }
  if (!B.tot && hipMalloc(reinterpret_cast<void**>(&B.tot), 4 * sizeof(uint64_t)) != hipSuccess)
    return FL_RESOURCE_EXHAUSTED;
  const ParticleView v = ctx->sorted.view();
  k_zero_words<<<grid_for(nw), kThreads, 0, s>>>(B.cand, nw);
  FL_TRY(reset_error(ctx));
  const int32_t creach = static_cast<int32_t>(reach);
  if (7 + 2 * creach <= kMaxWindow) {
    k_mark_candidates_bins<<<ctx->nbins, 256, 0, s>>>(v.s[0], v.s[1], v.s[2], ctx->perm, ctx->bin_key, ctx->bin_offset,
                                                      ctx->bin_n, ctx->bin_box, creach, g.ox, g.oy, g.oz, h, radius,
                                                      box, B.cand, ctx->err);
  } else {
    k_mark_candidates<<<grid_for(n), kThreads, 0, s>>>(v.s[0], v.s[1], v.s[2], ctx->perm, n, g.ox, g.oy, g.oz, h,
                                                       radius, box, B.cand, ctx->err);
  }
  k_dilate<<<grid_for(nw), kThreads, 0, s>>>(B.cand, box, B.centre, B.corner);
  k_popcounts<<<grid_for(nw), kThreads, 0, s>>>(B.cand, B.centre, B.corner, nw, B.counts);
  FL_HIP(hipGetLastError());
  for (int q = 0; q < 3; ++q) FL_TRY(exclusive_scan_u64(ctx, B.counts + q * nw, B.scans + q * nw, nw));
  k_totals<<<1, 64, 0, s>>>(B.counts, B.scans, nw, B.tot);
  FL_HIP(hipMemcpyAsync(ctx->host_words, B.tot, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FL_HIP(hipMemcpyAsync(ctx->host_words + 3, ctx->err, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  if (ctx->host_words[3] != ~uint64_t(0)) return static_cast<int>(ctx->host_words[3] & 0xff);
  const uint64_t nc = ctx->host_words[0], ncen = ctx->host_words[1], ncor = ctx->host_words[2];
  if (max_candidates && nc > max_candidates) return FL_RESOURCE_EXHAUSTED;
  FL_TRY(grow_buf(&B.cells, &B.cap_cells, nc));
  FL_TRY(grow_buf(&B.centres, &B.cap_centres, ncen));
  FL_TRY(grow_buf(&B.corners, &B.cap_corners, ncor));
  k_compact<<<grid_for(nw), kThreads, 0, s>>>(B.cand, B.centre, B.corner, B.scans, box, B.cells, B.centres, B.corners);
  // Sub-cell regrouping of the sorted particles for the nearest-particle search.
  const uint64_t nkeys = ctx->bin_box.volume();
  FL_TRY(grow_buf(&B.sub_range, &B.cap_sub, 64 * nkeys));
  FL_TRY(grow_buf(&B.sub_pos, &B.cap_pos, 4 * n));
  k_zero_ranges<<<std::min<uint64_t>(grid_for(64 * nkeys), 65536), kThreads, 0, s>>>(B.sub_range, 64 * nkeys);
  k_subcells<<<ctx->nbins, 256, 0, s>>>(v.s[0], v.s[1], v.s[2], ctx->bin_key, ctx->bin_offset, ctx->bin_n,
                                        ctx->bin_box, g.ox, g.oy, g.oz, h, B.sub_range, B.sub_pos);
  FL_HIP(hipGetLastError());
  const SubGrid G{B.sub_range, B.sub_pos, ctx->bin_box};
  FL_TRY(reset_error(ctx));
  if (ncen) k_field_distance<<<grid_for(ncen), kThreads, 0, s>>>(B.centres, ncen, 0, G, g.ox, g.oy, g.oz, h, radius, ctx->err);
  if (ncor) k_field_distance<<<grid_for(ncor), kThreads, 0, s>>>(B.corners, ncor, 1, G, g.ox, g.oy, g.oz, h, radius, ctx->err);
  FL_HIP(hipGetLastError());
  uint64_t word = 0;
  FL_TRY(read_error(ctx, &word));
  if (word != ~uint64_t(0)) return static_cast<int>(word & 0xff);
  out->cells = B.cells;
  out->cell_count = nc;
  out->centres = B.centres;
  out->centre_count = ncen;
  out->corners = B.corners;
  out->corner_count = ncor;
  out->cell_box_lo = fl_int3{box.lo[0], box.lo[1], box.lo[2]};
  out->cell_box_hi = fl_int3{box.lo[0] + box.dim[0] - 1, box.lo[1] + box.dim[1] - 1, box.lo[2] + box.dim[2] - 1};
  return FL_OK;
}

}  // namespace fl
