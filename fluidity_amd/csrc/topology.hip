// Topology kernels: particle bins, the stable (z,y,x)-bin sort, block activation and the dense
// block lookup.
//
// The sort reproduces LiquidStepWorkspace3D's order exactly (native/src/liquid_step_workspace_3d.cc:245-268:
// signed (z,y,x) bin, then original ordinal). It is a counting sort: an atomic histogram over the
// bin box hands out provisional slots, an exclusive scan places every bin, and each bin's segment is
// then sorted by original ordinal in LDS, so the result is unique and independent of atomic order.
// Activation marks the owners of all 81 stencil faces of every particle and of the six faces of
// every pressure cell in a dense bitmap over the block box; compacting the bitmap in (z,y,x) order
// yields the sorted unique block list of native/src/liquid_step_workspace_3d.cc:208-244, :272-276.
#include "context.h"

namespace fl {
namespace {

__global__ void k_fill_i64(int64_t* p, uint64_t n, int64_t v) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_i32(int32_t* p, uint64_t n, int32_t v) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u8(uint8_t* p, uint64_t n, uint8_t v) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void k_init_scalars(uint64_t* err, int32_t* red) {
  if (threadIdx.x == 0) {
    for (int k = 0; k < 4; ++k) err[k] = ~uint64_t(0);
    for (int a = 0; a < 3; ++a) {
      red[a] = 0x7fffffff;
      red[3 + a] = -0x7fffffff - 1;
    }
    red[6] = 0;
  }
}

inline unsigned capped_grid(uint64_t n) {
  const uint64_t g = (n + kThreads - 1) / kThreads;
  return static_cast<unsigned>(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

// First failure of one particle up to and including its bin, in the reference's order:
// deposit-input validation (mac_deposit_3d_internal.h:25-37), sorted and trial normalisation,
// then the bin range (sparse_mac_grid_3d.cc:463-474). Gathers skip the mass/velocity checks and
// plain bin queries skip the trial position (sparse_mac_grid_3d.cc:438-449). Scalar code only, so
// nothing spills to scratch.
template <int kMode, bool kPow2>
__global__ void k_bin(ParticleView p, uint64_t n, Geometry g, int32_t* bins, uint64_t* err, int32_t* partial) {
  __shared__ int32_t smin[3][kThreads / 64], smax[3][kThreads / 64];
  int32_t mn0 = 0x7fffffff, mn1 = 0x7fffffff, mn2 = 0x7fffffff;
  int32_t mx0 = -0x7fffffff - 1, mx1 = -0x7fffffff - 1, mx2 = -0x7fffffff - 1;
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    int32_t b0 = 0x7fffffff, b1 = 0x7fffffff, b2 = 0x7fffffff;
    int st = kOk;
    if (kMode == 4) {  // boundary gather: the retained velocity must be finite
      if (!(finite(p.v[0][i]) && finite(p.v[1][i]) && finite(p.v[2][i]))) st = kInvalidArgument;
    }
    if (kMode == 1 || kMode == 3) {
      bool good = finite(p.m[i]) && p.m[i] >= 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) good = good && finite(p.v[a][i]);
      if (kMode == 1) {
#pragma unroll
        for (int k = 0; k < 9; ++k) good = good && finite(p.c[k][i]);
      }
      if (!good) st = kInvalidArgument;
    }
    const double s0 = p.s[0][i], s1 = p.s[1][i], s2 = p.s[2][i];
    if (st == kOk && !(finite(s0) && finite(s1) && finite(s2))) st = kInvalidArgument;
    const auto c0 = normalize<kPow2>(s0, g.ox, g.h, g.inv_h);
    const auto c1 = normalize<kPow2>(s1, g.oy, g.h, g.inv_h);
    const auto c2 = normalize<kPow2>(s2, g.oz, g.h, g.inv_h);
    if (st == kOk && !(finite(c0.X) && finite(c1.X) && finite(c2.X))) st = kOutOfRange;
    if (kMode != 0 && p.t[0] != p.s[0]) {
      const double t0 = p.t[0][i], t1 = p.t[1][i], t2 = p.t[2][i];
      if (st == kOk && !(finite(t0) && finite(t1) && finite(t2))) st = kInvalidArgument;
      if (st == kOk) {
        const bool fin = finite(normalize<kPow2>(t0, g.ox, g.h, g.inv_h).X) &&
                         finite(normalize<kPow2>(t1, g.oy, g.h, g.inv_h).X) &&
                         finite(normalize<kPow2>(t2, g.oz, g.h, g.inv_h).X);
        if (!fin) st = kOutOfRange;
      }
    }
    if (st == kOk) st = checked_bin(c0, b0);
    if (st == kOk) st = checked_bin(c1, b1);
    if (st == kOk) st = checked_bin(c2, b2);
    if (st == kOk) {
      bins[3 * i + 0] = b0;
      bins[3 * i + 1] = b1;
      bins[3 * i + 2] = b2;
      mn0 = min(mn0, b0);
      mn1 = min(mn1, b1);
      mn2 = min(mn2, b2);
      mx0 = max(mx0, b0);
      mx1 = max(mx1, b1);
      mx2 = max(mx2, b2);
    } else {
      atomicMin(reinterpret_cast<unsigned long long*>(err), static_cast<unsigned long long>(error_word(i, st)));
      bins[3 * i + 0] = bins[3 * i + 1] = bins[3 * i + 2] = 0x7fffffff;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    mn0 = min(mn0, __shfl_xor(mn0, o));
    mn1 = min(mn1, __shfl_xor(mn1, o));
    mn2 = min(mn2, __shfl_xor(mn2, o));
    mx0 = max(mx0, __shfl_xor(mx0, o));
    mx1 = max(mx1, __shfl_xor(mx1, o));
    mx2 = max(mx2, __shfl_xor(mx2, o));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    smin[0][wave] = mn0;
    smin[1][wave] = mn1;
    smin[2][wave] = mn2;
    smax[0][wave] = mx0;
    smax[1][wave] = mx1;
    smax[2][wave] = mx2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int a = threadIdx.x;
    int32_t lo = smin[a][0], hi = smax[a][0];
    for (int w = 1; w < kThreads / 64; ++w) {
      lo = min(lo, smin[a][w]);
      hi = max(hi, smax[a][w]);
    }
    partial[6 * blockIdx.x + a] = lo;
    partial[6 * blockIdx.x + 3 + a] = hi;
  }
}

// Folds the per-workgroup bin boxes in a fixed order into red[0..5].
__global__ void k_box_finalize(const int32_t* partial, unsigned groups, int32_t* red) {
  __shared__ int32_t s[6][kThreads];
  int32_t v[6] = {0x7fffffff, 0x7fffffff, 0x7fffffff, -0x7fffffff - 1, -0x7fffffff - 1, -0x7fffffff - 1};
  for (unsigned g = threadIdx.x; g < groups; g += blockDim.x)
    for (int a = 0; a < 3; ++a) {
      v[a] = min(v[a], partial[6 * g + a]);
      v[3 + a] = max(v[3 + a], partial[6 * g + 3 + a]);
    }
  for (int q = 0; q < 6; ++q) s[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int st = kThreads / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = min(s[a][threadIdx.x], s[a][threadIdx.x + st]);
        s[3 + a][threadIdx.x] = max(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) red[threadIdx.x] = s[threadIdx.x][0];
}

// Histogram over bin keys with one atomic per distinct key in each wavefront: lanes sharing a key
// (neighbouring particles usually do) take consecutive provisional slots from a single atomicAdd.
__global__ void k_count(const int32_t* bins, uint64_t n, Box box, uint32_t* key, uint32_t* rank, uint32_t* count) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t k = 0xffffffffu;
  bool valid = false;
  if (i < n) {
    const int32_t bx = bins[3 * i], by = bins[3 * i + 1], bz = bins[3 * i + 2];
    if (!(bx == 0x7fffffff && by == 0x7fffffff && bz == 0x7fffffff)) {
      k = static_cast<uint32_t>(box.index(bx, by, bz));
      valid = true;
    }
    key[i] = k;
  }
  uint64_t pending = __ballot(valid);
  const uint64_t below = (uint64_t(1) << lane) - 1;
  while (pending) {
    const int leader = __ffsll(static_cast<long long>(pending)) - 1;
    const uint32_t lk = __shfl(k, leader);
    const uint64_t same = __ballot(valid && k == lk);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&count[lk], static_cast<uint32_t>(__popcll(same)));
    base = __shfl(base, leader);
    if (valid && k == lk) rank[i] = base + static_cast<uint32_t>(__popcll(same & below));
    pending &= ~same;
  }
}

// Packs (count, nonempty) so one scan yields both the particle offset and the bin-list position.
__global__ void k_pack_counts(const uint32_t* count, uint64_t n, uint64_t* out) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = count[i];
  out[i] = uint64_t(c) | (c ? (uint64_t(1) << 32) : 0);
}

constexpr unsigned kBinGroups = 2048;  // workgroups of the bin kernel (grid-stride)
constexpr int kScanItems = 4;
constexpr int kScanTile = kThreads * kScanItems;

__global__ void k_scan_tile(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* partial) {
  __shared__ uint64_t wave_sum[kThreads / 64];
  const uint64_t base = blockIdx.x * uint64_t(kScanTile) + threadIdx.x * kScanItems;
  uint64_t v[kScanItems];
  uint64_t sum = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    sum += v[k];
  }
  // inclusive scan of per-thread sums within the wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t inc = sum;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t t = __shfl_up(inc, o);
    if (lane >= o) inc += t;
  }
  if (lane == 63) wave_sum[wave] = inc;
  __syncthreads();
  uint64_t wave_off = 0;
  for (int w = 0; w < wave; ++w) wave_off += wave_sum[w];
  uint64_t run = wave_off + inc - sum;
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == kThreads - 1) partial[blockIdx.x] = wave_off + inc;
}

__global__ void k_scan_add(uint64_t* out, uint64_t n, const uint64_t* partial_scanned) {
  const uint64_t base = blockIdx.x * uint64_t(kScanTile) + threadIdx.x * kScanItems;
  const uint64_t add = partial_scanned[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) out[base + k] += add;
}

__global__ void k_make_bin_list(const uint32_t* count, const uint64_t* scan, uint64_t nkeys, uint32_t* bkey,
                                uint32_t* boff, uint32_t* bn, int32_t* red) {
  const uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (k >= nkeys) return;
  const uint32_t c = count[k];
  if (c) {
    const uint32_t pos = static_cast<uint32_t>(scan[k] >> 32);
    bkey[pos] = static_cast<uint32_t>(k);
    boff[pos] = static_cast<uint32_t>(scan[k] & 0xffffffffu);
    bn[pos] = c;
  }
  if (k == nkeys - 1) {
    red[6] = static_cast<int32_t>((scan[k] >> 32) + (c ? 1 : 0));
    red[8] = static_cast<int32_t>((scan[k] & 0xffffffffu) + c);
  }
}

__global__ void k_scatter(const uint32_t* key, const uint32_t* rank, uint64_t n, const uint64_t* scan, uint32_t* perm) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n || key[i] == 0xffffffffu) return;
  perm[(scan[key[i]] & 0xffffffffu) + rank[i]] = static_cast<uint32_t>(i);
}

constexpr int kSortCap = 4096;

__device__ void bitonic_lds(uint32_t* s, int m) {
  for (int size = 2; size <= m; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < m / 2; t += blockDim.x) {
        const int lo = 2 * t - (t & (stride - 1));
        const int hi = lo + stride;
        const bool up = ((lo & size) == 0);
        const uint32_t a = s[lo], b = s[hi];
        if ((a > b) == up) {
          s[lo] = b;
          s[hi] = a;
        }
      }
      __syncthreads();
    }
}

// Sorts every bin segment of <= kSortCap entries by original ordinal; longer segments are sorted
// in kSortCap chunks here and merged by k_merge_segment.
__global__ void k_segment_sort(uint32_t* perm, const uint32_t* boff, const uint32_t* bn, uint32_t nbins) {
  __shared__ uint32_t s[kSortCap];
  const uint32_t b = blockIdx.x;
  if (b >= nbins) return;
  const uint32_t off = boff[b], n = bn[b];
  for (uint32_t c0 = 0; c0 < n; c0 += kSortCap) {
    const uint32_t cn = min(n - c0, uint32_t(kSortCap));
    int m = 1;
    while (m < static_cast<int>(cn)) m <<= 1;
    for (int t = threadIdx.x; t < m; t += blockDim.x) s[t] = t < static_cast<int>(cn) ? perm[off + c0 + t] : 0xffffffffu;
    __syncthreads();
    bitonic_lds(s, m);
    for (int t = threadIdx.x; t < static_cast<int>(cn); t += blockDim.x) perm[off + c0 + t] = s[t];
    __syncthreads();
  }
}

__device__ inline uint32_t lower_bound_u32(const uint32_t* a, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Bottom-up merge of the sorted kSortCap chunks of one long segment (values are unique ordinals,
// so each element's output index is its own index plus its rank in the partner run).
__global__ void k_merge_segment(uint32_t* perm, uint32_t* tmp, const uint32_t* boff, const uint32_t* bn,
                                uint32_t nbins) {
  const uint32_t b = blockIdx.x;
  if (b >= nbins) return;
  const uint32_t off = boff[b], n = bn[b];
  if (n <= kSortCap) return;
  uint32_t* src = perm + off;
  uint32_t* dst = tmp + off;
  for (uint32_t width = kSortCap; width < n; width <<= 1) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      const uint32_t run = i / width, pair = run >> 1;
      const uint32_t a0 = pair * 2 * width;
      const uint32_t b0 = min(a0 + width, n), b1 = min(a0 + 2 * width, n);
      uint32_t out;
      if ((run & 1) == 0)
        out = a0 + (i - a0) + lower_bound_u32(src + b0, b1 - b0, src[i]);
      else
        out = a0 + (i - b0) + lower_bound_u32(src + a0, b0 - a0, src[i]);
      dst[out] = src[i];
    }
    __threadfence();
    __syncthreads();
    uint32_t* t = src;
    src = dst;
    dst = t;
  }
  if (src != perm + off) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) perm[off + i] = src[i];
  }
}

__global__ void k_permute(const double* src, double* dst, uint64_t cap_src, uint64_t cap_dst, uint32_t column_mask,
                          const uint32_t* perm, uint64_t n) {
  const uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = perm[k];
  for (int c = 0; c < kParticleDoubles; ++c)
    if (column_mask & (1u << c)) dst[c * cap_dst + k] = src[c * cap_src + i];
}

__global__ void k_sort_entries(const int32_t* bins, const uint32_t* perm, uint64_t n, fl_sort_entry* out) {
  const uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = perm[k];
  fl_sort_entry e;
  e.bin_index.x = bins[3 * i];
  e.bin_index.y = bins[3 * i + 1];
  e.bin_index.z = bins[3 * i + 2];
  e.pad = 0;
  e.original_ordinal = i;
  out[k] = e;
}

// Owners of the geometry stencils of one particle (sorted = trial position), per face axis.
template <bool kPow2>
__global__ void k_activate_particles(const double* sx, const double* sy, const double* sz, const uint32_t* perm,
                                     uint64_t n, Geometry g, Box box, uint8_t* mark, uint64_t* err) {
  const uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (k >= n) return;
  const double s[3] = {sx[k], sy[k], sz[k]};
  CoordT<kPow2> c[3];
  int32_t bin[3];
  for (int a = 0; a < 3; ++a) {
    c[a] = normalize<kPow2>(s[a], comp(g, a), g.h, g.inv_h);
    checked_bin(c[a], bin[a]);
  }
  int32_t lo[2][3], hi[2][3];  // [normal?][coordinate]
  for (int a = 0; a < 3; ++a)
    for (int kind = 0; kind < 2; ++kind) {
      AxisStencil st;
      const int code = axis_stencil(c[a], kind == 1 ? 0.0 : 0.5, bin[a], g.h, st);
      if (code) {
        atomicMin(reinterpret_cast<unsigned long long*>(err),
                  static_cast<unsigned long long>(error_word(perm ? perm[k] : k, code)));
        return;
      }
      lo[kind][a] = static_cast<int32_t>(floor_div4(st.first));
      hi[kind][a] = static_cast<int32_t>(floor_div4(int64_t(st.first) + 2));
    }
  for (int axis = 0; axis < 3; ++axis) {
    const int kx = axis == 0, ky = axis == 1, kz = axis == 2;
    for (int z = lo[kz][2]; z <= hi[kz][2]; ++z)
      for (int y = lo[ky][1]; y <= hi[ky][1]; ++y)
        for (int x = lo[kx][0]; x <= hi[kx][0]; ++x) {
          const int64_t i = box.index(x, y, z);
          if (i >= 0) mark[i] = 1;
        }
  }
}

__global__ void k_activate_cells(const int32_t* cells, uint64_t nc, Box box, uint8_t* mark) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nc) return;
  const int32_t x = cells[3 * i], y = cells[3 * i + 1], z = cells[3 * i + 2];
  const int32_t ox = int32_t(floor_div4(x)), oy = int32_t(floor_div4(y)), oz = int32_t(floor_div4(z));
  const int64_t idx[4] = {box.index(ox, oy, oz), box.index(int32_t(floor_div4(int64_t(x) + 1)), oy, oz),
                          box.index(ox, int32_t(floor_div4(int64_t(y) + 1)), oz),
                          box.index(ox, oy, int32_t(floor_div4(int64_t(z) + 1)))};
  for (int k = 0; k < 4; ++k)
    if (idx[k] >= 0) mark[idx[k]] = 1;
}

__global__ void k_pack_marks(const uint8_t* mark, uint64_t n, uint64_t* out) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = mark[i];
}

__global__ void k_compact_marks(const uint8_t* mark, const uint64_t* scan, uint64_t n, Box box, int32_t* blocks,
                                uint64_t block_capacity, int32_t* red) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  if (mark[i] && scan[i] < block_capacity) {
    const uint64_t pos = scan[i];
    const int32_t x = box.lo[0] + int32_t(i % box.dim[0]);
    const int32_t y = box.lo[1] + int32_t((i / box.dim[0]) % box.dim[1]);
    const int32_t z = box.lo[2] + int32_t(i / (uint64_t(box.dim[0]) * box.dim[1]));
    blocks[3 * pos] = x;
    blocks[3 * pos + 1] = y;
    blocks[3 * pos + 2] = z;
  }
  if (i == n - 1) red[7] = int32_t(scan[i] + mark[i]);
}

__global__ void k_scatter_lut(const int32_t* blocks, uint64_t nb, Box box, int32_t* lut) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nb) return;
  lut[box.index(blocks[3 * i], blocks[3 * i + 1], blocks[3 * i + 2])] = static_cast<int32_t>(i);
}

}  // namespace

int fill_i64(fl_ctx* ctx, int64_t* p, uint64_t n, int64_t v) {
  if (n == 0) return FL_OK;
  k_fill_i64<<<capped_grid(n), kThreads, 0, ctx->stream>>>(p, n, v);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int fill_i32(fl_ctx* ctx, int32_t* p, uint64_t n, int32_t v) {
  if (n == 0) return FL_OK;
  k_fill_i32<<<capped_grid(n), kThreads, 0, ctx->stream>>>(p, n, v);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int reset_error(fl_ctx* ctx) {
  k_init_scalars<<<1, 64, 0, ctx->stream>>>(ctx->err, ctx->reduce_i32);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int read_error(fl_ctx* ctx, uint64_t* word) {
  FL_HIP(hipMemcpyAsync(ctx->host_words, ctx->err, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  *word = ctx->host_words[0];
  return FL_OK;
}

int exclusive_scan_u64(fl_ctx* ctx, const uint64_t* in, uint64_t* out, uint64_t n) {
  if (n == 0) return FL_OK;
  const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
  // Partial sums live at the front of ctx->scan_partial; deeper levels follow.
  static thread_local uint64_t depth_offset = 0;
  uint64_t* partial = ctx->scan_partial + depth_offset;
  k_scan_tile<<<static_cast<unsigned>(tiles), kThreads, 0, ctx->stream>>>(in, out, n, partial);
  FL_HIP(hipGetLastError());
  if (tiles > 1) {
    const uint64_t saved = depth_offset;
    depth_offset += tiles + 1;
    uint64_t* scanned = ctx->scan_partial + depth_offset;
    depth_offset += tiles + 1;
    const int s = exclusive_scan_u64(ctx, partial, scanned, tiles);
    depth_offset = saved;
    if (s) return s;
    k_scan_add<<<static_cast<unsigned>(tiles), kThreads, 0, ctx->stream>>>(out, n, scanned);
    FL_HIP(hipGetLastError());
  }
  return FL_OK;
}

// Bins every particle and reduces the bin box. mode: deposit checks (1), gather checks (2) or the
// plain sort bin (0). Returns the first failing particle's status when mode is 0.
int bin_particles(fl_ctx* ctx, const ParticleView& p, uint64_t n, bool deposit_checks, bool gather_checks) {
  return bin_particles_mode(ctx, p, n, deposit_checks ? kBinDeposit : (gather_checks ? kBinGather : kBinPlain));
}

int bin_particles_mode(fl_ctx* ctx, const ParticleView& p, uint64_t n, int mode) {
  FL_TRY(reset_error(ctx));
  if (mode < kBinPlain || mode > kBinBoundaryGather) return FL_INTERNAL;
  if (n) {
    const bool pow2 = ctx->geometry.inv_h != 0.0;
    using K = decltype(&k_bin<0, true>);
    static const K table[2][5] = {{k_bin<0, false>, k_bin<1, false>, k_bin<2, false>, k_bin<3, false>, k_bin<4, false>},
                                  {k_bin<0, true>, k_bin<1, true>, k_bin<2, true>, k_bin<3, true>, k_bin<4, true>}};
    auto* kernel = table[pow2 ? 1 : 0][mode];
    const unsigned groups = std::min<unsigned>(grid_for(n), kBinGroups);
    int32_t* partial = reinterpret_cast<int32_t*>(ctx->scan_partial);  // 6 * kBinGroups ints
    kernel<<<groups, kThreads, 0, ctx->stream>>>(p, n, ctx->geometry, ctx->bin, ctx->err, partial);
    k_box_finalize<<<1, kThreads, 0, ctx->stream>>>(partial, groups, ctx->reduce_i32);
    FL_HIP(hipGetLastError());
  }
  FL_HIP(hipMemcpyAsync(ctx->host_words, ctx->reduce_i32, 8 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  const int32_t* r = reinterpret_cast<const int32_t*>(ctx->host_words);
  Box box;
  if (r[0] <= r[3]) {
    for (int a = 0; a < 3; ++a) {
      box.lo[a] = r[a];
      box.dim[a] = r[3 + a] - r[a] + 1;
    }
  }
  ctx->bin_box = box;
  return FL_OK;
}

// Counting sort by bin key, then deterministic in-segment ordering by original ordinal; permutes
// the particle columns into ctx->sorted when requested.
int sort_particles(fl_ctx* ctx, uint64_t n, const ParticleView& src, bool permute, uint32_t permute_columns) {
  const Box& box = ctx->bin_box;
  const uint64_t nkeys = box.volume();
  ctx->nbins = 0;
  ctx->nvalid = 0;
  if (n == 0 || nkeys == 0) return FL_OK;
  if (nkeys > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  k_fill_i32<<<capped_grid(nkeys), kThreads, 0, ctx->stream>>>(reinterpret_cast<int32_t*>(ctx->bin_count), nkeys, 0);
  k_count<<<grid_for(n), kThreads, 0, ctx->stream>>>(ctx->bin, n, box, ctx->key, ctx->rank, ctx->bin_count);
  k_pack_counts<<<grid_for(nkeys), kThreads, 0, ctx->stream>>>(ctx->bin_count, nkeys, ctx->bin_scan);
  FL_HIP(hipGetLastError());
  uint64_t* scan = ctx->bin_scan + nkeys;  // second half holds the exclusive scan
  FL_TRY(exclusive_scan_u64(ctx, ctx->bin_scan, scan, nkeys));
  k_make_bin_list<<<grid_for(nkeys), kThreads, 0, ctx->stream>>>(ctx->bin_count, scan, nkeys, ctx->bin_key,
                                                                  ctx->bin_offset, ctx->bin_n, ctx->reduce_i32);
  k_scatter<<<grid_for(n), kThreads, 0, ctx->stream>>>(ctx->key, ctx->rank, n, scan, ctx->perm);
  FL_HIP(hipGetLastError());
  FL_HIP(hipMemcpyAsync(ctx->host_words, ctx->reduce_i32 + 6, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  ctx->nbins = static_cast<uint32_t>(reinterpret_cast<const int32_t*>(ctx->host_words)[0]);
  ctx->nvalid = static_cast<uint64_t>(reinterpret_cast<const int32_t*>(ctx->host_words)[2]);
  if (ctx->nbins) {
    k_segment_sort<<<ctx->nbins, kThreads, 0, ctx->stream>>>(ctx->perm, ctx->bin_offset, ctx->bin_n, ctx->nbins);
    k_merge_segment<<<ctx->nbins, kThreads, 0, ctx->stream>>>(ctx->perm, ctx->perm_tmp, ctx->bin_offset, ctx->bin_n,
                                                              ctx->nbins);
    FL_HIP(hipGetLastError());
  }
  if (permute) {
    // The sorted copy holds the valid particles in bin order; columns follow ParticleBuffers. The
    // trial columns are skipped when they alias the sorted positions.
    const uint64_t valid = ctx->nvalid;
    const uint32_t mask = ctx->raw.trial_aliases_sorted ? (((1u << kParticleDoubles) - 1) & ~0x38u)
                                                        : ((1u << kParticleDoubles) - 1);
    if (valid)
      k_permute<<<grid_for(valid), kThreads, 0, ctx->stream>>>(src.s[0], ctx->sorted.base, ctx->raw.capacity,
                                                               ctx->sorted.capacity, permute_columns & mask, ctx->perm,
                                                               valid);
    FL_HIP(hipGetLastError());
    ctx->sorted.trial_aliases_sorted = ctx->raw.trial_aliases_sorted;
  }
  return FL_OK;
}

int write_sort_entries(fl_ctx* ctx, uint64_t n, fl_sort_entry* d_out) {
  if (!n) return FL_OK;
  k_sort_entries<<<grid_for(n), kThreads, 0, ctx->stream>>>(ctx->bin, ctx->perm, n, d_out);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// Activation over the sorted particles (ctx->sorted) and device pressure cells.
int activate_blocks(fl_ctx* ctx, uint64_t n, const int32_t* d_cells, uint64_t cell_count) {
  Box box = ctx->block_box;  // set by the caller
  const uint64_t vol = box.volume();
  if (vol > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  if (vol == 0) {  // nothing to activate
    ctx->nb = 0;
    return FL_OK;
  }
  k_fill_u8<<<capped_grid(vol), kThreads, 0, ctx->stream>>>(ctx->mark, vol, 0);
  FL_HIP(hipGetLastError());
  if (n) {
    const ParticleView v = ctx->sorted.view();
    auto* kernel = ctx->geometry.inv_h != 0.0 ? k_activate_particles<true> : k_activate_particles<false>;
    kernel<<<grid_for(n), kThreads, 0, ctx->stream>>>(v.s[0], v.s[1], v.s[2], ctx->perm, n, ctx->geometry, box,
                                                      ctx->mark, ctx->err);
    FL_HIP(hipGetLastError());
  }
  if (cell_count) {
    k_activate_cells<<<grid_for(cell_count), kThreads, 0, ctx->stream>>>(d_cells, cell_count, box, ctx->mark);
    FL_HIP(hipGetLastError());
  }
  k_pack_marks<<<grid_for(vol), kThreads, 0, ctx->stream>>>(ctx->mark, vol, ctx->bin_scan);
  FL_TRY(exclusive_scan_u64(ctx, ctx->bin_scan, ctx->bin_scan + vol, vol));
  // Compaction writes the block list into ctx->blocks; the count goes to reduce_i32[7].
  k_compact_marks<<<grid_for(vol), kThreads, 0, ctx->stream>>>(ctx->mark, ctx->bin_scan + vol, vol, box,
                                                                ctx->blocks, ctx->cap.blocks, ctx->reduce_i32);
  FL_HIP(hipGetLastError());
  FL_HIP(hipMemcpyAsync(ctx->host_words, ctx->reduce_i32 + 7, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  const uint64_t nb = static_cast<uint64_t>(reinterpret_cast<const int32_t*>(ctx->host_words)[0]);
  if (nb > ctx->cap.blocks) return FL_RESOURCE_EXHAUSTED;
  ctx->nb = nb;
  k_fill_i32<<<capped_grid(vol), kThreads, 0, ctx->stream>>>(ctx->lut, vol, -1);
  if (nb) k_scatter_lut<<<grid_for(nb), kThreads, 0, ctx->stream>>>(ctx->blocks, nb, box, ctx->lut);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// Canonical host block list -> device blocks + lookup table over their box.
int build_lookup(fl_ctx* ctx, const int32_t* host_blocks, uint64_t nb) {
  if (nb > ctx->cap.blocks) return FL_RESOURCE_EXHAUSTED;
  Box box;
  if (nb) {
    int32_t lo[3], hi[3];
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = host_blocks[a];
    for (uint64_t i = 1; i < nb; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], host_blocks[3 * i + a]);
        hi[a] = std::max(hi[a], host_blocks[3 * i + a]);
      }
    for (int a = 0; a < 3; ++a) {
      box.lo[a] = lo[a];
      box.dim[a] = hi[a] - lo[a] + 1;
    }
  }
  if (box.volume() > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  ctx->block_box = box;
  ctx->nb = nb;
  if (nb) {
    FL_HIP(hipMemcpyAsync(ctx->blocks, host_blocks, 3 * sizeof(int32_t) * nb, hipMemcpyHostToDevice, ctx->stream));
    k_fill_i32<<<capped_grid(box.volume()), kThreads, 0, ctx->stream>>>(ctx->lut, box.volume(), -1);
    k_scatter_lut<<<grid_for(nb), kThreads, 0, ctx->stream>>>(ctx->blocks, nb, box, ctx->lut);
    FL_HIP(hipGetLastError());
  }
  return FL_OK;
}

GridView grid_view(fl_ctx* ctx) {
  GridView v;
  v.g = ctx->geometry;
  v.blocks = ctx->blocks;
  v.lut = ctx->lut;
  v.box = ctx->block_box;
  v.nb = ctx->nb;
  return v;
}

BinList bin_list(fl_ctx* ctx) {
  BinList b;
  b.key = ctx->bin_key;
  b.offset = ctx->bin_offset;
  b.count = ctx->bin_n;
  b.n = ctx->nbins;
  b.box = ctx->bin_box;
  return b;
}

}  // namespace fl
