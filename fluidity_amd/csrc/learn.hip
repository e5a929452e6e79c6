// Learned preconditioner on gfx950: a 3-level 3D U-Net on 8^3 tiles of the level-0 pressure grid
// (channels 8/16/32, 3x3x3 convolutions with zero padding, ReLU, 2x2x2 average pooling, nearest
// upsampling, skips by channel concatenation) mapping (normalised residual, liquid, air, solid) to a
// pressure correction; its masked mean-squared loss, exact backward pass and Adam training.
//
// The reference has no learned component. The network and its gradient are checked against the FP64
// oracle (oracle/unet.cc), which is pinned by finite differences.
//
// Layout: a batch of tiles keeps every activation of a tile in one contiguous record of kActFloats
// floats ([channel][voxel], voxel = (z*s + y)*s + x), so one lane per (tile, voxel) reads and writes
// consecutive addresses. Weights are read with wave-uniform (scalar) loads. Accumulation is FP32.
#include "learn.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace fl {
namespace {

struct LayerSpec {
  int side, a, ca, b, cb, cout, y;  // input offsets a (ca channels) and b (cb channels), output offset y
  bool relu;
};
constexpr LayerSpec kLayers[kNetLayers] = {
    {8, kActIn, 4, 0, 0, 8, kActA0, true},     // encoder 8^3
    {8, kActA0, 8, 0, 0, 8, kActA1, true},     // -> skip s1
    {4, kActP1, 8, 0, 0, 16, kActA2, true},    // encoder 4^3
    {4, kActA2, 16, 0, 0, 16, kActA3, true},   // -> skip s2
    {2, kActP2, 16, 0, 0, 32, kActA4, true},   // bottleneck 2^3
    {2, kActA4, 32, 0, 0, 32, kActA5, true},   //
    {4, kActU5, 32, kActA3, 16, 16, kActA6, true},  // up(x5) ++ s2
    {4, kActA6, 16, 0, 0, 16, kActA7, true},   //
    {8, kActU7, 16, kActA1, 8, 8, kActA8, true},    // up(x7) ++ s1
    {8, kActA8, 8, 0, 0, 1, kActOut, false},   // output
};

uint64_t layer_weights(int l) { return uint64_t(kLayers[l].cout) * (kLayers[l].ca + kLayers[l].cb) * 27; }

uint64_t layer_offset(int l) {
  uint64_t o = 0;
  for (int i = 0; i < l; ++i) o += layer_weights(i) + kLayers[i].cout;
  return o;
}

__device__ inline void tap(int k, int& dz, int& dy, int& dx) {
  dz = k / 9 - 1;
  dy = (k / 3) % 3 - 1;
  dx = k % 3 - 1;
}

// y[co][v] = relu?(b[co] + sum_{ci,k} w[co][ci][k] x[ci][v + off(k)]), x = a-channels ++ b-channels.
template <int COUT>
__global__ void k_conv_fwd(float* act, uint64_t ntiles, LayerSpec L, const float* w, const float* bias) {
  const int s = L.side, n = s * s * s;
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (gid >= ntiles * uint64_t(n)) return;
  const uint64_t t = gid / uint64_t(n);
  const int v = static_cast<int>(gid % uint64_t(n));
  const int z = v / (s * s), y = (v / s) % s, x = v % s;
  float* base = act + t * kActFloats;
  float acc[COUT];
#pragma unroll
  for (int co = 0; co < COUT; ++co) acc[co] = bias[co];
  const int cin = L.ca + L.cb;
  for (int ci = 0; ci < cin; ++ci) {
    const float* src = ci < L.ca ? base + L.a + ci * n : base + L.b + (ci - L.ca) * n;
    for (int k = 0; k < 27; ++k) {
      int dz, dy, dx;
      tap(k, dz, dy, dx);
      const int zz = z + dz, yy = y + dy, xx = x + dx;
      if (zz < 0 || zz >= s || yy < 0 || yy >= s || xx < 0 || xx >= s) continue;
      const float xv = src[(zz * s + yy) * s + xx];
      const float* wk = w + ci * 27 + k;
#pragma unroll
      for (int co = 0; co < COUT; ++co) acc[co] = fmaf(wk[co * cin * 27], xv, acc[co]);
    }
  }
  float* dst = base + L.y;
#pragma unroll
  for (int co = 0; co < COUT; ++co) dst[co * n + v] = (L.relu && acc[co] < 0.f) ? 0.f : acc[co];
}

// dL/dx for one convolution: dx[ci][v] (+)= sum_{co,k} w[co][ci][k] g[co][v - off(k)], where g (at the
// layer's output offset in `grad`) already includes the ReLU derivative. a-channels and b-channels of
// the input gradient are added to what is there (both gradients of a skip tensor meet by addition).
template <int CIN>
__global__ void k_conv_bwd_data(float* grad, uint64_t ntiles, LayerSpec L, const float* w) {
  const int s = L.side, n = s * s * s;
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (gid >= ntiles * uint64_t(n)) return;
  const uint64_t t = gid / uint64_t(n);
  const int v = static_cast<int>(gid % uint64_t(n));
  const int z = v / (s * s), y = (v / s) % s, x = v % s;
  float* base = grad + t * kActFloats;
  float acc[CIN];
#pragma unroll
  for (int ci = 0; ci < CIN; ++ci) acc[ci] = 0.f;
  for (int co = 0; co < L.cout; ++co) {
    const float* gy = base + L.y + co * n;
    for (int k = 0; k < 27; ++k) {
      int dz, dy, dx;
      tap(k, dz, dy, dx);
      const int zz = z - dz, yy = y - dy, xx = x - dx;
      if (zz < 0 || zz >= s || yy < 0 || yy >= s || xx < 0 || xx >= s) continue;
      const float gv = gy[(zz * s + yy) * s + xx];
      const float* wk = w + co * CIN * 27 + k;
#pragma unroll
      for (int ci = 0; ci < CIN; ++ci) acc[ci] = fmaf(wk[ci * 27], gv, acc[ci]);
    }
  }
#pragma unroll
  for (int ci = 0; ci < CIN; ++ci) {
    float* d = ci < L.ca ? base + L.a + ci * n : base + L.b + (ci - L.ca) * n;
    d[v] += acc[ci];
  }
}

// Per-chunk partial gradients of one layer's weights and biases: lane (chunk, j), j < weights is
// w[co][ci][k], j >= weights the bias of co = j - weights. Tiles of a chunk are summed in order.
__global__ void k_conv_bwd_w(const float* act, const float* grad, uint64_t ntiles, uint64_t chunk, LayerSpec L,
                             uint64_t nw, float* partial) {
  const uint64_t per = nw + L.cout;
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  const uint64_t nchunks = (ntiles + chunk - 1) / chunk;
  if (gid >= per * nchunks) return;
  const uint64_t c = gid / per, j = gid % per;
  const int s = L.side, n = s * s * s, cin = L.ca + L.cb;
  const uint64_t t0 = c * chunk, t1 = std::min<uint64_t>(ntiles, t0 + chunk);
  float acc = 0.f;
  if (j >= nw) {
    const int co = static_cast<int>(j - nw);
    for (uint64_t t = t0; t < t1; ++t) {
      const float* g = grad + t * kActFloats + L.y + co * n;
      for (int v = 0; v < n; ++v) acc += g[v];
    }
  } else {
    const int co = static_cast<int>(j / (uint64_t(cin) * 27));
    const int ci = static_cast<int>((j / 27) % cin);
    const int k = static_cast<int>(j % 27);
    int dz, dy, dx;
    tap(k, dz, dy, dx);
    for (uint64_t t = t0; t < t1; ++t) {
      const float* g = grad + t * kActFloats + L.y + co * n;
      const float* xs = act + t * kActFloats + (ci < L.ca ? L.a + ci * n : L.b + (ci - L.ca) * n);
      for (int z = 0; z < s; ++z) {
        const int zz = z + dz;
        if (zz < 0 || zz >= s) continue;
        for (int y = 0; y < s; ++y) {
          const int yy = y + dy;
          if (yy < 0 || yy >= s) continue;
          for (int x = 0; x < s; ++x) {
            const int xx = x + dx;
            if (xx < 0 || xx >= s) continue;
            acc = fmaf(g[(z * s + y) * s + x], xs[(zz * s + yy) * s + xx], acc);
          }
        }
      }
    }
  }
  partial[c * per + j] = acc;
}

// out[j] = sum over chunks in order (deterministic).
__global__ void k_reduce_chunks(const float* partial, uint64_t nchunks, uint64_t per, float* out) {
  const uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (j >= per) return;
  float a = 0.f;
  for (uint64_t c = 0; c < nchunks; ++c) a += partial[c * per + j];
  out[j] = a;
}

__global__ void k_relu_bwd(const float* act, float* grad, uint64_t ntiles, int off, int count) {
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (gid >= ntiles * uint64_t(count)) return;
  const uint64_t t = gid / uint64_t(count);
  const uint64_t i = t * kActFloats + off + gid % uint64_t(count);
  if (!(act[i] > 0.f)) grad[i] = 0.f;
}

// 2x2x2 average pooling from side 2h to side h (fwd), and its adjoint added into the input gradient.
__global__ void k_pool_fwd(float* act, uint64_t ntiles, int ch, int h, int off_in, int off_out) {
  const int n = h * h * h;
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (gid >= ntiles * uint64_t(ch * n)) return;
  const uint64_t t = gid / uint64_t(ch * n);
  const int r = static_cast<int>(gid % uint64_t(ch * n)), c = r / n, v = r % n;
  const int z = v / (h * h), y = (v / h) % h, x = v % h, s2 = 2 * h;
  const float* in = act + t * kActFloats + off_in + c * s2 * s2 * s2;
  float a = 0.f;
  for (int o = 0; o < 8; ++o) a += in[((2 * z + (o >> 2)) * s2 + 2 * y + ((o >> 1) & 1)) * s2 + 2 * x + (o & 1)];
  act[t * kActFloats + off_out + c * n + v] = 0.125f * a;
}
__global__ void k_pool_bwd(float* grad, uint64_t ntiles, int ch, int h, int off_in, int off_out) {
  const int s2 = 2 * h, n2 = s2 * s2 * s2;
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (gid >= ntiles * uint64_t(ch * n2)) return;
  const uint64_t t = gid / uint64_t(ch * n2);
  const int r = static_cast<int>(gid % uint64_t(ch * n2)), c = r / n2, v = r % n2;
  const int z = v / (s2 * s2), y = (v / s2) % s2, x = v % s2;
  grad[t * kActFloats + off_in + c * n2 + v] +=
      0.125f * grad[t * kActFloats + off_out + c * h * h * h + ((z >> 1) * h + (y >> 1)) * h + (x >> 1)];
}

// Nearest upsampling from side h to 2h (fwd) and its adjoint (sum of the 8 children, in order).
__global__ void k_up_fwd(float* act, uint64_t ntiles, int ch, int h, int off_in, int off_out) {
  const int s2 = 2 * h, n2 = s2 * s2 * s2;
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (gid >= ntiles * uint64_t(ch * n2)) return;
  const uint64_t t = gid / uint64_t(ch * n2);
  const int r = static_cast<int>(gid % uint64_t(ch * n2)), c = r / n2, v = r % n2;
  const int z = v / (s2 * s2), y = (v / s2) % s2, x = v % s2;
  act[t * kActFloats + off_out + c * n2 + v] =
      act[t * kActFloats + off_in + c * h * h * h + ((z >> 1) * h + (y >> 1)) * h + (x >> 1)];
}
__global__ void k_up_bwd(float* grad, uint64_t ntiles, int ch, int h, int off_in, int off_out) {
  const int n = h * h * h, s2 = 2 * h;
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (gid >= ntiles * uint64_t(ch * n)) return;
  const uint64_t t = gid / uint64_t(ch * n);
  const int r = static_cast<int>(gid % uint64_t(ch * n)), c = r / n, v = r % n;
  const int z = v / (h * h), y = (v / h) % h, x = v % h;
  const float* g = grad + t * kActFloats + off_out + c * s2 * s2 * s2;
  float a = 0.f;
  for (int o = 0; o < 8; ++o) a += g[((2 * z + (o >> 2)) * s2 + 2 * y + ((o >> 1) & 1)) * s2 + 2 * x + (o & 1)];
  grad[t * kActFloats + off_in + c * n + v] += a;
}

// Fixed-order block sum into partial[blockIdx.x].
__device__ inline void block_sum(float v, float* partial) {
  __shared__ float red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_mask_sum(const float* act, uint64_t ntiles, float* partial) {
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  float m = 0.f;
  if (gid < ntiles * 512) m = act[(gid / 512) * kActFloats + kActIn + 512 + gid % 512];
  block_sum(m, partial);
}

// Loss 0.5 m (out - target)^2 / sum(m) per lane into block partials; dL/dout into the gradient record.
__global__ void k_loss(const float* act, const float* target, uint64_t ntiles, const float* msum, float* grad,
                       float* partial) {
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  float l = 0.f;
  if (gid < ntiles * 512) {
    const uint64_t t = gid / 512;
    const int v = static_cast<int>(gid % 512);
    const float inv = msum[0] > 0.f ? 1.f / msum[0] : 0.f;
    const float m = act[t * kActFloats + kActIn + 512 + v];
    const float d = act[t * kActFloats + kActOut + v] - target[t * 512 + v];
    l = 0.5f * m * d * d * inv;
    grad[t * kActFloats + kActOut + v] = m * d * inv;
  }
  block_sum(l, partial);
}

__global__ void k_sum_partials(const float* partial, unsigned n, float* out) {
  __shared__ float red[256];
  float a = 0.f;
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) a += partial[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void k_adam(float* w, const float* g, float* m1, float* m2, uint64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ void k_fill_f32(float* p, uint64_t n, float v) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// BF16 inference path: every convolution but the 1-channel output as an implicit GEMM on the matrix
// cores, Y[m][co] = sum_k X[m][k] W[k][co] with m a voxel of one of TPW tiles and k = ci*27 + tap,
// products in BF16, accumulation in FP32 (v_mfma_f32_16x16x16bf16_1k). The tiles' input channels are
// staged zero-padded in LDS as BF16 ((SIDE+2)^3 per channel), the weights transposed as [co][k] (k
// padded to 16, co to 16), and the im2col offset of every k in a table, so an A fragment is four LDS
// loads at base(m) + off(k) with no index arithmetic. Each wave owns a run of 16-row blocks and
// reuses every B fragment across them.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t bf16_bits(float v) {
  const uint32_t u = __float_as_uint(v);
  return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // nearest even (finite v)
}

template <int SIDE, int CIN, int COUT, int TPW>
__global__ __launch_bounds__(256) void k_conv_mfma(float* act, uint64_t ntiles, LayerSpec L, const float* w,
                                                   const float* bias) {
  constexpr int P = SIDE + 2, P3 = P * P * P, V = SIDE * SIDE * SIDE;
  constexpr int K = CIN * 27, KP = (K + 15) / 16 * 16;
  constexpr int NP = COUT < 16 ? 16 : COUT, NB = NP / 16;
  constexpr int M = TPW * V, RBW = M / 64;  // 16-row blocks per wave
  static_assert(M % 64 == 0, "four waves of whole row blocks");
  __shared__ uint16_t xs[TPW * CIN * P3];
  __shared__ uint16_t wt[NP * KP];
  __shared__ int32_t koff[KP];
  const uint64_t t0 = blockIdx.x * uint64_t(TPW);
  for (int idx = threadIdx.x; idx < NP * KP; idx += 256) {
    const int co = idx / KP, k = idx % KP;
    float v = 0.f;
    if (k < K && co < COUT) v = w[(co * CIN + k / 27) * 27 + k % 27];
    wt[idx] = bf16_bits(v);
  }
  for (int k = threadIdx.x; k < KP; k += 256) {
    int o = 0;
    if (k < K) {
      const int ci = k / 27, tp = k % 27;
      o = ci * P3 + ((tp / 9 - 1) * P + ((tp / 3) % 3 - 1)) * P + (tp % 3 - 1);
    }
    koff[k] = o;
  }
  for (int idx = threadIdx.x; idx < TPW * CIN * P3; idx += 256) {
    const int tt = idx / (CIN * P3), rem = idx % (CIN * P3), ci = rem / P3, pp = rem % P3;
    const int pz = pp / (P * P), py = (pp / P) % P, px = pp % P;
    float v = 0.f;
    if (t0 + tt < ntiles && pz >= 1 && pz <= SIDE && py >= 1 && py <= SIDE && px >= 1 && px <= SIDE) {
      const float* base = act + (t0 + tt) * kActFloats;
      const float* src = ci < L.ca ? base + L.a + ci * V : base + L.b + (ci - L.ca) * V;
      v = src[((pz - 1) * SIDE + (py - 1)) * SIDE + (px - 1)];
    }
    xs[idx] = bf16_bits(v);
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, kq = lane >> 4;
  int base_m[RBW];
#pragma unroll
  for (int j = 0; j < RBW; ++j) {
    const int m = (wave * RBW + j) * 16 + r;
    const int tt = m / V, v = m % V;
    const int z = v / (SIDE * SIDE), y = (v / SIDE) % SIDE, x = v % SIDE;
    base_m[j] = tt * CIN * P3 + ((z + 1) * P + (y + 1)) * P + (x + 1);
  }
  f32x4 acc[RBW][NB];
#pragma unroll
  for (int j = 0; j < RBW; ++j)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[j][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < KP; k0 += 16) {
    const int kb = k0 + kq * 4;
    const int o0 = koff[kb], o1 = koff[kb + 1], o2 = koff[kb + 2], o3 = koff[kb + 3];
    bf16x4 bf[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const uint16_t* col = wt + (nb * 16 + r) * KP + kb;
      bf[nb] = bf16x4{static_cast<short>(col[0]), static_cast<short>(col[1]), static_cast<short>(col[2]),
                      static_cast<short>(col[3])};
    }
#pragma unroll
    for (int j = 0; j < RBW; ++j) {
      const int bm = base_m[j];
      const bf16x4 a = {static_cast<short>(xs[bm + o0]), static_cast<short>(xs[bm + o1]),
                        static_cast<short>(xs[bm + o2]), static_cast<short>(xs[bm + o3])};
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[j][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bf[nb], acc[j][nb], 0, 0, 0);
    }
  }
  // D[i][j] sits in lane j + 16 (i / 4), element i % 4: column r is the output channel.
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int co = nb * 16 + r;
    if (co >= COUT) continue;
    const float bco = bias[co];
#pragma unroll
    for (int j = 0; j < RBW; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = (wave * RBW + j) * 16 + kq * 4 + e;
        const int tt = m / V, v = m % V;
        if (t0 + tt >= ntiles) continue;
        float val = acc[j][nb][e] + bco;
        if (L.relu && val < 0.f) val = 0.f;
        act[(t0 + tt) * kActFloats + L.y + co * V + v] = val;
      }
  }
}

template <int SIDE, int CIN, int COUT, int TPW>
void launch_conv_mfma(hipStream_t s, float* act, uint64_t nt, const LayerSpec& L, const float* w, const float* b) {
  const unsigned g = static_cast<unsigned>((nt + TPW - 1) / TPW);
  k_conv_mfma<SIDE, CIN, COUT, TPW><<<g, 256, 0, s>>>(act, nt, L, w, b);
}

template <int N>
void conv_fwd_n(hipStream_t s, float* act, uint64_t nt, const LayerSpec& L, const float* w, const float* b) {
  const uint64_t lanes = nt * uint64_t(L.side * L.side * L.side);
  k_conv_fwd<N><<<grid_for(lanes), kThreads, 0, s>>>(act, nt, L, w, b);
}
template <int N>
void conv_bwd_n(hipStream_t s, float* grad, uint64_t nt, const LayerSpec& L, const float* w) {
  const uint64_t lanes = nt * uint64_t(L.side * L.side * L.side);
  k_conv_bwd_data<N><<<grid_for(lanes), kThreads, 0, s>>>(grad, nt, L, w);
}

int conv_fwd(hipStream_t s, float* act, uint64_t nt, int l, const float* params, bool mfma = false) {
  const LayerSpec& L = kLayers[l];
  const float* w = params + layer_offset(l);
  const float* b = w + layer_weights(l);
  if (mfma && L.cout > 1) {
    switch (l) {
      case 0: launch_conv_mfma<8, 4, 8, 1>(s, act, nt, L, w, b); break;
      case 1: launch_conv_mfma<8, 8, 8, 1>(s, act, nt, L, w, b); break;
      case 2: launch_conv_mfma<4, 8, 16, 8>(s, act, nt, L, w, b); break;
      case 3: launch_conv_mfma<4, 16, 16, 8>(s, act, nt, L, w, b); break;
      case 4: launch_conv_mfma<2, 16, 32, 16>(s, act, nt, L, w, b); break;
      case 5: launch_conv_mfma<2, 32, 32, 16>(s, act, nt, L, w, b); break;
      case 6: launch_conv_mfma<4, 48, 16, 4>(s, act, nt, L, w, b); break;
      case 7: launch_conv_mfma<4, 16, 16, 8>(s, act, nt, L, w, b); break;
      case 8: launch_conv_mfma<8, 24, 8, 1>(s, act, nt, L, w, b); break;
      default: return FL_INTERNAL;
    }
    FL_HIP(hipGetLastError());
    return FL_OK;
  }
  switch (L.cout) {
    case 1: conv_fwd_n<1>(s, act, nt, L, w, b); break;
    case 8: conv_fwd_n<8>(s, act, nt, L, w, b); break;
    case 16: conv_fwd_n<16>(s, act, nt, L, w, b); break;
    case 32: conv_fwd_n<32>(s, act, nt, L, w, b); break;
    default: return FL_INTERNAL;
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int conv_bwd_data(hipStream_t s, float* grad, uint64_t nt, int l, const float* params) {
  const LayerSpec& L = kLayers[l];
  const float* w = params + layer_offset(l);
  switch (L.ca + L.cb) {
    case 4: conv_bwd_n<4>(s, grad, nt, L, w); break;
    case 8: conv_bwd_n<8>(s, grad, nt, L, w); break;
    case 16: conv_bwd_n<16>(s, grad, nt, L, w); break;
    case 24: conv_bwd_n<24>(s, grad, nt, L, w); break;
    case 32: conv_bwd_n<32>(s, grad, nt, L, w); break;
    case 48: conv_bwd_n<48>(s, grad, nt, L, w); break;
    default: return FL_INTERNAL;
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

}  // namespace

uint64_t net_param_count() { return layer_offset(kNetLayers); }

int net_forward(hipStream_t s, const float* params, float* act, uint64_t nt, bool mfma) {
  if (!nt) return FL_OK;
  FL_TRY(conv_fwd(s, act, nt, 0, params, mfma));
  FL_TRY(conv_fwd(s, act, nt, 1, params, mfma));
  k_pool_fwd<<<grid_for(nt * 8 * 64), kThreads, 0, s>>>(act, nt, 8, 4, kActA1, kActP1);
  FL_TRY(conv_fwd(s, act, nt, 2, params, mfma));
  FL_TRY(conv_fwd(s, act, nt, 3, params, mfma));
  k_pool_fwd<<<grid_for(nt * 16 * 8), kThreads, 0, s>>>(act, nt, 16, 2, kActA3, kActP2);
  FL_TRY(conv_fwd(s, act, nt, 4, params, mfma));
  FL_TRY(conv_fwd(s, act, nt, 5, params, mfma));
  k_up_fwd<<<grid_for(nt * 32 * 64), kThreads, 0, s>>>(act, nt, 32, 2, kActA5, kActU5);
  FL_TRY(conv_fwd(s, act, nt, 6, params, mfma));
  FL_TRY(conv_fwd(s, act, nt, 7, params, mfma));
  k_up_fwd<<<grid_for(nt * 16 * 512), kThreads, 0, s>>>(act, nt, 16, 4, kActA7, kActU7);
  FL_TRY(conv_fwd(s, act, nt, 8, params, mfma));
  FL_TRY(conv_fwd(s, act, nt, 9, params));
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int net_backward(hipStream_t s, const float* params, const float* act, float* grad, uint64_t nt, float* partial,
                 uint64_t partial_floats, float* grad_params) {
  if (!nt) return FL_OK;
  const uint64_t chunk = 8, nchunks = (nt + chunk - 1) / chunk;
  auto weights = [&](int l) -> int {
    const LayerSpec& L = kLayers[l];
    const uint64_t nw = layer_weights(l), per = nw + L.cout;
    if (per * nchunks > partial_floats) return FL_RESOURCE_EXHAUSTED;
    k_conv_bwd_w<<<grid_for(per * nchunks), kThreads, 0, s>>>(act, grad, nt, chunk, L, nw, partial);
    k_reduce_chunks<<<grid_for(per), kThreads, 0, s>>>(partial, nchunks, per, grad_params + layer_offset(l));
    FL_HIP(hipGetLastError());
    return FL_OK;
  };
  auto relu = [&](int l) {
    const LayerSpec& L = kLayers[l];
    const int count = L.cout * L.side * L.side * L.side;
    k_relu_bwd<<<grid_for(nt * count), kThreads, 0, s>>>(act, grad, nt, L.y, count);
  };
  // Layer 9 (linear), 8, up7, 7, 6, up5, 5, 4, pool2, 3, 2, pool1, 1, 0.
  FL_TRY(weights(9));
  FL_TRY(conv_bwd_data(s, grad, nt, 9, params));
  relu(8);
  FL_TRY(weights(8));
  FL_TRY(conv_bwd_data(s, grad, nt, 8, params));
  k_up_bwd<<<grid_for(nt * 16 * 64), kThreads, 0, s>>>(grad, nt, 16, 4, kActA7, kActU7);
  relu(7);
  FL_TRY(weights(7));
  FL_TRY(conv_bwd_data(s, grad, nt, 7, params));
  relu(6);
  FL_TRY(weights(6));
  FL_TRY(conv_bwd_data(s, grad, nt, 6, params));
  k_up_bwd<<<grid_for(nt * 32 * 8), kThreads, 0, s>>>(grad, nt, 32, 2, kActA5, kActU5);
  relu(5);
  FL_TRY(weights(5));
  FL_TRY(conv_bwd_data(s, grad, nt, 5, params));
  relu(4);
  FL_TRY(weights(4));
  FL_TRY(conv_bwd_data(s, grad, nt, 4, params));
  k_pool_bwd<<<grid_for(nt * 16 * 64), kThreads, 0, s>>>(grad, nt, 16, 2, kActA3, kActP2);
  relu(3);
  FL_TRY(weights(3));
  FL_TRY(conv_bwd_data(s, grad, nt, 3, params));
  relu(2);
  FL_TRY(weights(2));
  FL_TRY(conv_bwd_data(s, grad, nt, 2, params));
  k_pool_bwd<<<grid_for(nt * 8 * 512), kThreads, 0, s>>>(grad, nt, 8, 4, kActA1, kActP1);
  relu(1);
  FL_TRY(weights(1));
  FL_TRY(conv_bwd_data(s, grad, nt, 1, params));
  relu(0);
  FL_TRY(weights(0));
  FL_TRY(conv_bwd_data(s, grad, nt, 0, params));  // input gradient, for checks
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int net_loss(hipStream_t s, const float* act, const float* target, float* grad, uint64_t nt, float* partial,
             float* scalars) {
  const unsigned g = grid_for(nt * 512);
  k_mask_sum<<<g, kThreads, 0, s>>>(act, nt, partial);
  k_sum_partials<<<1, kThreads, 0, s>>>(partial, g, scalars);
  k_loss<<<g, kThreads, 0, s>>>(act, target, nt, scalars, grad, partial);
  k_sum_partials<<<1, kThreads, 0, s>>>(partial, g, scalars + 1);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int net_clear(hipStream_t s, float* p, uint64_t n) {
  k_fill_f32<<<std::max<unsigned>(1, std::min<unsigned>(grid_for(n), 4096)), kThreads, 0, s>>>(p, n, 0.f);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int net_adam(hipStream_t s, float* w, const float* g, float* m1, float* m2, uint64_t n, float lr, int step) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, float(step)), c2 = 1.f - std::pow(b2, float(step));
  k_adam<<<grid_for(n), kThreads, 0, s>>>(w, g, m1, m2, n, lr, b1, b2, 1e-8f, c1, c2);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

}  // namespace fl

// ------------------------------------------------------------------ state, training, C-ABI
namespace fl {
namespace {

// Copies the selected training pairs into the batch records (inputs) and the target rows.
__global__ void k_load_batch(const float* data_in, const float* data_target, const int32_t* index, uint64_t nt,
                             float* act, float* target) {
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (gid >= nt * 2560) return;
  const uint64_t t = gid / 2560;
  const int j = static_cast<int>(gid % 2560);
  const uint64_t src = static_cast<uint64_t>(index[t]);
  if (j < 2048)
    act[t * kActFloats + kActIn + j] = data_in[src * 2048 + j];
  else
    target[t * 512 + (j - 2048)] = data_target[src * 512 + (j - 2048)];
}

// Host tile-major arrays [n][2048] <-> the input slots of the activation records.
__global__ void k_put_inputs(const float* in, uint64_t nt, float* act) {
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (gid < nt * 2048) act[(gid / 2048) * kActFloats + kActIn + gid % 2048] = in[gid];
}
__global__ void k_get_slot(const float* act, uint64_t nt, int off, int count, float* out) {
  const uint64_t gid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (gid < nt * uint64_t(count)) out[gid] = act[(gid / count) * kActFloats + off + gid % count];
}

LearnState* state(fl_ctx* ctx) { return static_cast<LearnState*>(ctx->learn); }

template <typename T>
int lalloc(T** p, uint64_t n) {
  *p = nullptr;
  if (!n) return FL_OK;
  FL_HIP(hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)));
  return FL_OK;
}
template <typename T>
void lfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

void destroy_learn(fl_ctx* ctx) {
  LearnState* L = state(ctx);
  if (!L) return;
  lfree(L->params);
  lfree(L->gparams);
  lfree(L->m1);
  lfree(L->m2);
  lfree(L->act);
  lfree(L->grad);
  lfree(L->partial);
  lfree(L->scalars);
  lfree(L->scale);
  lfree(L->batch_target);
  lfree(L->batch_index);
  lfree(L->stage);
  lfree(L->data_in);
  lfree(L->data_target);
  delete L;
  ctx->learn = nullptr;
}

// One epoch of Adam over the training set in a seeded random order; returns the mean batch loss.
int train_epoch(fl_ctx* ctx, uint64_t batch, float lr, uint64_t seed, float* mean_loss) {
  LearnState* L = state(ctx);
  hipStream_t s = ctx->stream;
  std::vector<int32_t> order(L->ndata);
  for (uint64_t i = 0; i < L->ndata; ++i) order[i] = static_cast<int32_t>(i);
  std::mt19937_64 rng(seed);
  std::shuffle(order.begin(), order.end(), rng);
  double total = 0.0;
  int batches = 0;
  for (uint64_t b0 = 0; b0 < L->ndata; b0 += batch) {
    const uint64_t nt = std::min<uint64_t>(batch, L->ndata - b0);
    FL_HIP(hipMemcpyAsync(L->batch_index, order.data() + b0, nt * sizeof(int32_t), hipMemcpyHostToDevice, s));
    k_load_batch<<<grid_for(nt * 2560), kThreads, 0, s>>>(L->data_in, L->data_target, L->batch_index, nt, L->act,
                                                          L->batch_target);
    FL_TRY(net_forward(s, L->params, L->act, nt));
    FL_TRY(net_clear(s, L->grad, nt * kActFloats));
    FL_TRY(net_loss(s, L->act, L->batch_target, L->grad, nt, L->partial, L->scalars));
    FL_TRY(net_backward(s, L->params, L->act, L->grad, nt, L->partial, L->partial_floats, L->gparams));
    FL_TRY(net_adam(s, L->params, L->gparams, L->m1, L->m2, L->np, lr, ++L->step));
    float h[2];
    FL_HIP(hipMemcpyAsync(h, L->scalars, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    total += h[1];
    ++batches;
  }
  *mean_loss = batches ? static_cast<float>(total / batches) : 0.f;
  return FL_OK;
}

// Mean loss over the training set, forward only (FP32), in order.
int eval_pairs(fl_ctx* ctx, uint64_t batch, float* mean_loss) {
  LearnState* L = state(ctx);
  hipStream_t s = ctx->stream;
  std::vector<int32_t> order(L->ndata);
  for (uint64_t i = 0; i < L->ndata; ++i) order[i] = static_cast<int32_t>(i);
  double total = 0.0;
  uint64_t count = 0;
  for (uint64_t b0 = 0; b0 < L->ndata; b0 += batch) {
    const uint64_t nt = std::min<uint64_t>(batch, L->ndata - b0);
    FL_HIP(hipMemcpyAsync(L->batch_index, order.data() + b0, nt * sizeof(int32_t), hipMemcpyHostToDevice, s));
    k_load_batch<<<grid_for(nt * 2560), kThreads, 0, s>>>(L->data_in, L->data_target, L->batch_index, nt, L->act,
                                                          L->batch_target);
    FL_TRY(net_forward(s, L->params, L->act, nt));
    FL_TRY(net_clear(s, L->grad, nt * kActFloats));
    FL_TRY(net_loss(s, L->act, L->batch_target, L->grad, nt, L->partial, L->scalars));
    float h[2];
    FL_HIP(hipMemcpyAsync(h, L->scalars, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    total += double(h[1]) * double(nt);
    count += nt;
  }
  *mean_loss = count ? static_cast<float>(total / double(count)) : 0.f;
  return FL_OK;
}

}  // namespace fl

using fl::LearnState;

extern "C" {

size_t fl_learn_param_count(void) { return fl::net_param_count(); }

int fl_learn_setup(fl_ctx* ctx, uint64_t tile_capacity, uint64_t pair_capacity) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (tile_capacity == 0) return FL_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  fl::destroy_learn(ctx);
  LearnState* L = new LearnState();
  ctx->learn = L;
  L->np = fl::net_param_count();
  L->tile_cap = tile_capacity;
  L->data_cap = pair_capacity;
  const uint64_t train_tiles = std::min<uint64_t>(tile_capacity, 1024);
  L->partial_floats = std::max<uint64_t>(27680 * ((train_tiles + 7) / 8), 2 * tile_capacity + 1024);
  int st = FL_OK;
  auto A = [&](auto** p, uint64_t n) {
    if (st == FL_OK) st = fl::lalloc(p, n);
  };
  A(&L->params, L->np);
  A(&L->gparams, L->np);
  A(&L->m1, L->np);
  A(&L->m2, L->np);
  A(&L->act, tile_capacity * fl::kActFloats);
  A(&L->grad, std::min<uint64_t>(tile_capacity, 1024) * fl::kActFloats);
  A(&L->partial, L->partial_floats);
  A(&L->scalars, 8);
  A(&L->scale, tile_capacity);
  A(&L->batch_target, tile_capacity * 512);
  A(&L->batch_index, tile_capacity);
  A(&L->stage, std::min<uint64_t>(tile_capacity, 1024) * 2048);
  A(&L->data_in, pair_capacity * 2048);
  A(&L->data_target, pair_capacity * 512);
  if (st != FL_OK) {
    fl::destroy_learn(ctx);
    return FL_RESOURCE_EXHAUSTED;
  }
  FL_TRY(fl::net_clear(ctx->stream, L->params, L->np));
  FL_TRY(fl::net_clear(ctx->stream, L->m1, L->np));
  FL_TRY(fl::net_clear(ctx->stream, L->m2, L->np));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  return FL_OK;
}

int fl_learn_set_params(fl_ctx* ctx, const float* params, size_t count) {
  if (!ctx || !ctx->learn) return FL_FAILED_PRECONDITION;
  LearnState* L = static_cast<LearnState*>(ctx->learn);
  if (count != L->np || !params) return FL_INVALID_ARGUMENT;
  FL_HIP(hipMemcpy(L->params, params, count * sizeof(float), hipMemcpyHostToDevice));
  FL_TRY(fl::net_clear(ctx->stream, L->m1, L->np));
  FL_TRY(fl::net_clear(ctx->stream, L->m2, L->np));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  L->step = 0;
  return FL_OK;
}

int fl_learn_get_params(fl_ctx* ctx, float* out, size_t count) {
  if (!ctx || !ctx->learn) return FL_FAILED_PRECONDITION;
  LearnState* L = static_cast<LearnState*>(ctx->learn);
  if (count != L->np || !out) return FL_INVALID_ARGUMENT;
  FL_HIP(hipMemcpy(out, L->params, count * sizeof(float), hipMemcpyDeviceToHost));
  return FL_OK;
}

static int learn_forward(fl_ctx* ctx, const float* inputs, size_t n, float* outputs, bool mfma) {
  if (!ctx || !ctx->learn) return FL_FAILED_PRECONDITION;
  LearnState* L = static_cast<LearnState*>(ctx->learn);
  if (n > std::min<uint64_t>(L->tile_cap, 1024)) return FL_RESOURCE_EXHAUSTED;
  hipStream_t s = ctx->stream;
  if (!n) return FL_OK;
  FL_HIP(hipMemcpyAsync(L->stage, inputs, n * 2048 * sizeof(float), hipMemcpyHostToDevice, s));
  fl::k_put_inputs<<<fl::grid_for(n * 2048), fl::kThreads, 0, s>>>(L->stage, n, L->act);
  FL_TRY(fl::net_forward(s, L->params, L->act, n, mfma));
  fl::k_get_slot<<<fl::grid_for(n * 512), fl::kThreads, 0, s>>>(L->act, n, fl::kActOut, 512, L->stage);
  FL_HIP(hipMemcpyAsync(outputs, L->stage, n * 512 * sizeof(float), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  return FL_OK;
}

int fl_learn_forward(fl_ctx* ctx, const float* inputs, size_t n, float* outputs) {
  return learn_forward(ctx, inputs, n, outputs, false);
}

int fl_learn_infer(fl_ctx* ctx, const float* inputs, size_t n, float* outputs) {
  return learn_forward(ctx, inputs, n, outputs, true);
}

int fl_learn_loss_grad(fl_ctx* ctx, const float* inputs, const float* targets, size_t n, float* loss,
                       float* grad_params, float* grad_inputs) {
  if (!ctx || !ctx->learn) return FL_FAILED_PRECONDITION;
  LearnState* L = static_cast<LearnState*>(ctx->learn);
  if (n > std::min<uint64_t>(L->tile_cap, 1024)) return FL_RESOURCE_EXHAUSTED;
  if (!n || !loss) return FL_INVALID_ARGUMENT;
  hipStream_t s = ctx->stream;
  FL_HIP(hipMemcpyAsync(L->stage, inputs, n * 2048 * sizeof(float), hipMemcpyHostToDevice, s));
  FL_HIP(hipMemcpyAsync(L->batch_target, targets, n * 512 * sizeof(float), hipMemcpyHostToDevice, s));
  fl::k_put_inputs<<<fl::grid_for(n * 2048), fl::kThreads, 0, s>>>(L->stage, n, L->act);
  FL_TRY(fl::net_forward(s, L->params, L->act, n));
  FL_TRY(fl::net_clear(s, L->grad, n * fl::kActFloats));
  FL_TRY(fl::net_loss(s, L->act, L->batch_target, L->grad, n, L->partial, L->scalars));
  FL_TRY(fl::net_backward(s, L->params, L->act, L->grad, n, L->partial, L->partial_floats, L->gparams));
  float h[2];
  FL_HIP(hipMemcpyAsync(h, L->scalars, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
  if (grad_params) FL_HIP(hipMemcpyAsync(grad_params, L->gparams, L->np * sizeof(float), hipMemcpyDeviceToHost, s));
  if (grad_inputs) {
    fl::k_get_slot<<<fl::grid_for(n * 2048), fl::kThreads, 0, s>>>(L->grad, n, fl::kActIn, 2048, L->stage);
    FL_HIP(hipMemcpyAsync(grad_inputs, L->stage, n * 2048 * sizeof(float), hipMemcpyDeviceToHost, s));
  }
  FL_HIP(hipStreamSynchronize(s));
  *loss = h[1];
  return FL_OK;
}

int fl_learn_add_pairs(fl_ctx* ctx, const float* inputs, const float* targets, size_t n) {
  if (!ctx || !ctx->learn) return FL_FAILED_PRECONDITION;
  LearnState* L = static_cast<LearnState*>(ctx->learn);
  if (L->ndata + n > L->data_cap) return FL_RESOURCE_EXHAUSTED;
  hipStream_t s = ctx->stream;
  FL_HIP(hipMemcpyAsync(L->data_in + L->ndata * 2048, inputs, n * 2048 * sizeof(float), hipMemcpyHostToDevice, s));
  FL_HIP(hipMemcpyAsync(L->data_target + L->ndata * 512, targets, n * 512 * sizeof(float), hipMemcpyHostToDevice, s));
  FL_HIP(hipStreamSynchronize(s));
  L->ndata += n;
  return FL_OK;
}

int fl_learn_pairs(fl_ctx* ctx, size_t* count) {
  if (!ctx || !ctx->learn || !count) return FL_FAILED_PRECONDITION;
  *count = static_cast<LearnState*>(ctx->learn)->ndata;
  return FL_OK;
}

int fl_learn_collect(fl_ctx* ctx, int every, int clear) {
  if (!ctx || !ctx->learn) return FL_FAILED_PRECONDITION;
  if (every < 0) return FL_INVALID_ARGUMENT;
  LearnState* L = static_cast<LearnState*>(ctx->learn);
  L->collect_every = every;
  if (clear) L->ndata = 0;
  return FL_OK;
}

int fl_learn_train(fl_ctx* ctx, int epochs, size_t batch, float learning_rate, uint64_t seed, float* epoch_losses) {
  if (!ctx || !ctx->learn) return FL_FAILED_PRECONDITION;
  LearnState* L = static_cast<LearnState*>(ctx->learn);
  if (epochs < 1 || batch < 1 || !(learning_rate > 0.f)) return FL_INVALID_ARGUMENT;
  if (batch > std::min<uint64_t>(L->tile_cap, 1024)) return FL_RESOURCE_EXHAUSTED;
  if (L->ndata == 0) return FL_FAILED_PRECONDITION;
  (void)hipSetDevice(ctx->device);
  for (int e = 0; e < epochs; ++e) {
    float loss = 0.f;
    FL_TRY(fl::train_epoch(ctx, batch, learning_rate, seed + uint64_t(e), &loss));
    if (epoch_losses) epoch_losses[e] = loss;
  }
  return FL_OK;
}

int fl_learn_eval(fl_ctx* ctx, size_t batch, float* mean_loss) {
  if (!ctx || !ctx->learn || !mean_loss) return FL_FAILED_PRECONDITION;
  LearnState* L = static_cast<LearnState*>(ctx->learn);
  if (batch < 1) return FL_INVALID_ARGUMENT;
  if (batch > std::min<uint64_t>(L->tile_cap, 1024)) return FL_RESOURCE_EXHAUSTED;
  if (L->ndata == 0) return FL_FAILED_PRECONDITION;
  return fl::eval_pairs(ctx, batch, mean_loss);
}

}  // extern "C"
