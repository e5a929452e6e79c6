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

// BF16 inference: the whole network on one 8^3 tile per workgroup (4 waves), every activation kept in
// LDS as BF16, so a tile reads its 4 x 512 inputs and writes its 512 outputs and nothing else goes
// to HBM. Activations are stored channel-innermost with a one-voxel zero halo ([(S+2)^3][CS], CS the
// channel stride), so the four channels at one tap are one 8-byte LDS read. Every convolution, the
// 1-channel output included, is an implicit GEMM on the matrix cores with the output channels as
// rows: Y^T[co][m] = sum_k W[co][k] X^T[k][m], k = tap * CIN + ci, m a voxel, products in BF16 and
// accumulation in FP32 (v_mfma_f32_16x16x16bf16_1k). The weights are packed once per call into the
// A-operand fragment order of each (16-row output block, 16-wide k step), so a wave's weight load is
// one coalesced 512-byte read; the B operand is one LDS read per lane; each lane of the result holds
// four consecutive output channels of one voxel, stored as one 8-byte LDS write after bias and ReLU.
// Pooling (FP32 sum of the 8 BF16 children) and nearest upsampling run between the convolutions.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t bf16_bits(float v) {
  const uint32_t u = __float_as_uint(v);
  return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // nearest even (finite v)
}
__device__ inline float bf16_value(short h) { return __uint_as_float(uint32_t(uint16_t(h)) << 16); }
__device__ inline bf16x4 bf16_pack(float a, float b, float c, float d) {
  return bf16x4{static_cast<short>(bf16_bits(a)), static_cast<short>(bf16_bits(b)), static_cast<short>(bf16_bits(c)),
                static_cast<short>(bf16_bits(d))};
}

constexpr int frag_steps(int l) { return (27 * (kLayers[l].ca + kLayers[l].cb) + 15) / 16; }
constexpr int frag_blocks(int l) { return (kLayers[l].cout + 15) / 16; }
constexpr int frag_offset(int l) {
  int o = 0;
  for (int i = 0; i < l; ++i) o += frag_steps(i) * frag_blocks(i) * 64;
  return o;
}
constexpr int param_offset(int l) {
  int o = 0;
  for (int i = 0; i < l; ++i) o += kLayers[i].cout * (kLayers[i].ca + kLayers[i].cb) * 27 + kLayers[i].cout;
  return o;
}

struct FragSpec {
  int cin[kNetLayers], cout[kNetLayers], steps[kNetLayers], first[kNetLayers + 1], param[kNetLayers];
};

// Fragment f of layer l (f = (block * steps + step) * 64 + lane): W[co][k] for co = 16 block + lane % 16
// and k = 16 step + 4 (lane / 16) + e, e = 0..3, zero outside the layer's channels and taps.
__global__ void k_pack_fragments(const float* params, FragSpec spec, bf16x4* frags) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= spec.first[kNetLayers]) return;
  int l = 0;
  while (f >= spec.first[l + 1]) ++l;
  const int local = f - spec.first[l], lane = local & 63, step = (local >> 6) % spec.steps[l];
  const int block = (local >> 6) / spec.steps[l];
  const int cin = spec.cin[l], co = block * 16 + (lane & 15);
  const float* w = params + spec.param[l];
  float v[4];
  for (int e = 0; e < 4; ++e) {
    const int k = 16 * step + 4 * (lane >> 4) + e;
    v[e] = (co < spec.cout[l] && k < 27 * cin) ? w[(co * cin + k % cin) * 27 + k / cin] : 0.f;
  }
  frags[f] = bf16_pack(v[0], v[1], v[2], v[3]);
}

// Padded LDS index of voxel (z, y, x) of a side-S grid.
template <int S>
__device__ constexpr int padded(int z, int y, int x) {
  return ((z + 1) * (S + 2) + (y + 1)) * (S + 2) + (x + 1);
}

// Zeroes every channel slot of the halo voxels of a side-S buffer.
template <int S, int CS>
__device__ void zero_halo(uint16_t* buf) {
  constexpr int P = S + 2;
  for (int i = threadIdx.x; i < P * P * P; i += 256) {
    const int pz = i / (P * P), py = (i / P) % P, px = i % P;
    if (pz == 0 || pz == P - 1 || py == 0 || py == P - 1 || px == 0 || px == P - 1) {
      bf16x4* d = reinterpret_cast<bf16x4*>(buf + i * CS);
#pragma unroll
      for (int c = 0; c < CS / 4; ++c) d[c] = bf16x4{0, 0, 0, 0};
    }
  }
}

// Weight fragments of k-steps s0 .. s0 + U - 1 (clamped to the last step) of CPW output blocks.
template <int U, int CPW, int KS>
__device__ inline void load_fragments(bf16x4 (&w)[U][CPW], const bf16x4* wf, int cb0, int s0) {
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int c = 0; c < CPW; ++c) w[u][c] = wf[((cb0 + c) * KS + min(s0 + u, KS - 1)) * 64];
}

// k-steps s0 .. s0 + U - 1 (those below KS) of the implicit GEMM: the lane's four k of a step are
// four channels at one tap, one 8-byte LDS read per column block.
template <int U, int BPW, int CPW, int KS, int CIN, int P, int XCS>
__device__ inline void conv_steps(const bf16x4 (&w)[U][CPW], const uint16_t* x, const int (&vb)[BPW], int q, int s0,
                                  f32x4 (&acc)[BPW][CPW]) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int s = s0 + u;
    if (s >= KS) break;
    const int k0 = 16 * s + 4 * q;
    int off = 0;  // padding k: weights are zero, any in-range voxel will do
    if (k0 < 27 * CIN) {
      const int tp = k0 / CIN, ci = k0 % CIN;
      off = (((tp / 9 - 1) * P + ((tp / 3) % 3 - 1)) * P + (tp % 3 - 1)) * XCS + ci;
    }
#pragma unroll
    for (int b = 0; b < BPW; ++b) {
      const bf16x4 xb = *reinterpret_cast<const bf16x4*>(x + vb[b] + off);
#pragma unroll
      for (int c = 0; c < CPW; ++c) acc[b][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w[u][c], xb, acc[b][c], 0, 0, 0);
    }
  }
}

// Convolution l from the LDS buffer at XOFF (channel stride XCS) into channels YCH0.. of the buffer at
// YOFF (stride YCS), or, for the output layer, into out[512] in FP32. A side-8 layer gives each wave 8
// column blocks of 16 voxels, a side-4 layer one; a side-2 layer (one half-used block) gives each of
// its 16-channel output blocks to one wave.
template <int l, int XOFF, int XCS, int YOFF, int YCS, int YCH0>
__device__ void conv_lds(uint16_t* lds, const bf16x4* frags, const float* params, float* out) {
  constexpr LayerSpec L = kLayers[l];
  constexpr int S = L.side, P = S + 2, V = S * S * S, CIN = L.ca + L.cb, COUT = L.cout;
  constexpr int KS = frag_steps(l), NCB = frag_blocks(l), NBLK = (V + 15) / 16;
  constexpr int BPW = NBLK >= 4 ? NBLK / 4 : 1, CPW = NBLK >= 4 ? NCB : 1;
  static_assert(NBLK >= 4 ? NBLK % 4 == 0 : NCB <= 4, "work split");
  static_assert(CIN % 4 == 0 && XCS >= CIN && XCS % 4 == 0 && YCS % 4 == 0, "8-byte channel groups");
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, j = lane & 15, q = lane >> 4;
  int blk0 = 0, cb0 = 0;
  if (NBLK >= 4) {
    blk0 = wave * BPW;
  } else {
    if (wave >= NCB) return;
    cb0 = wave;
  }
  const uint16_t* x = lds + XOFF;
  int vb[BPW];
#pragma unroll
  for (int b = 0; b < BPW; ++b) {
    const int m = ((blk0 + b) * 16 + j) % V;  // the unused half of a side-2 block repeats real voxels
    vb[b] = padded<S>(m / (S * S), (m / S) % S, m % S) * XCS;
  }
  f32x4 acc[BPW][CPW];
#pragma unroll
  for (int b = 0; b < BPW; ++b)
#pragma unroll
    for (int c = 0; c < CPW; ++c) acc[b][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  // Weight fragments are loaded one chunk of U k-steps ahead of their use (two register sets in
  // turn), so the L2 latency of a load overlaps the matrix work of the chunk before.
  constexpr int U = BPW >= 4 ? 4 : 16;
  const bf16x4* wf = frags + frag_offset(l) + lane;
  bf16x4 w0[U][CPW], w1[U][CPW];
  load_fragments<U, CPW, KS>(w0, wf, cb0, 0);
  for (int s0 = 0; s0 < KS; s0 += 2 * U) {
    load_fragments<U, CPW, KS>(w1, wf, cb0, s0 + U);
    conv_steps<U, BPW, CPW, KS, CIN, P, XCS>(w0, x, vb, q, s0, acc);
    load_fragments<U, CPW, KS>(w0, wf, cb0, s0 + 2 * U);
    conv_steps<U, BPW, CPW, KS, CIN, P, XCS>(w1, x, vb, q, s0 + U, acc);
  }
  // D[co][m] sits in lane m + 16 (co / 4), element co % 4.
  constexpr int BOFF = param_offset(l) + COUT * CIN * 27;
#pragma unroll
  for (int c = 0; c < CPW; ++c) {
    const int co0 = (cb0 + c) * 16 + 4 * q;
    if (co0 >= COUT) continue;
#pragma unroll
    for (int b = 0; b < BPW; ++b) {
      const int m = (blk0 + b) * 16 + j;
      if (m >= V) continue;
      if constexpr (COUT == 1) {
        out[m] = acc[b][c][0] + params[BOFF];
      } else {
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          v[e] = acc[b][c][e] + params[BOFF + co0 + e];
          if (L.relu && v[e] < 0.f) v[e] = 0.f;
        }
        *reinterpret_cast<bf16x4*>(lds + YOFF + padded<S>(m / (S * S), (m / S) % S, m % S) * YCS + YCH0 + co0) =
            bf16_pack(v[0], v[1], v[2], v[3]);
      }
    }
  }
}

// 2x2x2 average pooling of C channels from side S to S/2 (FP32 sum of the 8 children, in z, y, x order).
template <int S, int C, int XOFF, int XCS, int XCH0, int YOFF, int YCS>
__device__ void pool_lds(uint16_t* lds) {
  constexpr int H = S / 2, G = C / 4;
  for (int t = threadIdx.x; t < H * H * H * G; t += 256) {
    const int v = t / G, g = t % G, z = v / (H * H), y = (v / H) % H, x = v % H;
    float a[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const bf16x4 h = *reinterpret_cast<const bf16x4*>(
          lds + XOFF + padded<S>(2 * z + (o >> 2), 2 * y + ((o >> 1) & 1), 2 * x + (o & 1)) * XCS + XCH0 + 4 * g);
#pragma unroll
      for (int e = 0; e < 4; ++e) a[e] += bf16_value(h[e]);
    }
    *reinterpret_cast<bf16x4*>(lds + YOFF + padded<H>(z, y, x) * YCS + 4 * g) =
        bf16_pack(0.125f * a[0], 0.125f * a[1], 0.125f * a[2], 0.125f * a[3]);
  }
}

// Nearest upsampling of C channels from side S to 2S into channels YCH0.. of the target.
template <int S, int C, int XOFF, int XCS, int YOFF, int YCS, int YCH0>
__device__ void up_lds(uint16_t* lds) {
  constexpr int D = 2 * S, G = C / 4;
  for (int t = threadIdx.x; t < D * D * D * G; t += 256) {
    const int v = t / G, g = t % G, z = v / (D * D), y = (v / D) % D, x = v % D;
    *reinterpret_cast<bf16x4*>(lds + YOFF + padded<D>(z, y, x) * YCS + YCH0 + 4 * g) =
        *reinterpret_cast<const bf16x4*>(lds + XOFF + padded<S>(z >> 1, y >> 1, x >> 1) * XCS + 4 * g);
  }
}

// LDS plan, in BF16 elements. The concatenation (up(x7) ++ s1, 24 channels at side 8) has its own
// region for the whole pass, since s1 is written by layer 1; everything else shares a scratch region
// whose buffers are placed so that no phase reads what it writes. Channel strides of 16/48 channels
// are padded by 4 so that 16 consecutive voxels fall on distinct LDS banks.
namespace plan {
constexpr int kC8 = 0, kC8Stride = 24;              // [10^3][24]: up(x7) 0..15, s1 16..23
constexpr int kBase = 1000 * kC8Stride;
constexpr int kX0 = kBase, kA0 = kBase + 4000;      // [10^3][4], [10^3][8]
constexpr int kP1 = kBase + 4320, kA2 = kBase;      // [6^3][8], [6^3][20]
constexpr int kC6 = kBase + 4320, kC6Stride = 52;   // [6^3][52]: up(x5) 0..31, s2 32..47
constexpr int kP2 = kBase + 2048, kA4 = kBase, kA5 = kBase + 2048;  // [4^3][16], [4^3][32] x 2
constexpr int kA6 = kBase, kA7 = kBase + 4320;      // [6^3][20] x 2
constexpr int kA8 = kBase;                          // [10^3][8]
constexpr int kElements = kC6 + 216 * kC6Stride;
static_assert(kA0 + 8000 <= kElements && kA7 + 4320 <= kElements, "plan");
}  // namespace plan

__global__ __launch_bounds__(256) void k_unet_infer(float* act, const bf16x4* frags, const float* params) {
  using namespace plan;
  __shared__ __align__(16) uint16_t lds[kElements];
  float* rec = act + blockIdx.x * uint64_t(kActFloats);
  for (int v = threadIdx.x; v < 512; v += 256) {
    const float* in = rec + kActIn + v;
    *reinterpret_cast<bf16x4*>(lds + kX0 + padded<8>(v >> 6, (v >> 3) & 7, v & 7) * 4) =
        bf16_pack(in[0], in[512], in[1024], in[1536]);
  }
  zero_halo<8, 4>(lds + kX0);
  zero_halo<8, 8>(lds + kA0);
  zero_halo<8, kC8Stride>(lds + kC8);
  __syncthreads();
  conv_lds<0, kX0, 4, kA0, 8, 0>(lds, frags, params, nullptr);
  __syncthreads();
  conv_lds<1, kA0, 8, kC8, kC8Stride, 16>(lds, frags, params, nullptr);
  __syncthreads();
  pool_lds<8, 8, kC8, kC8Stride, 16, kP1, 8>(lds);
  zero_halo<4, 8>(lds + kP1);
  __syncthreads();
  conv_lds<2, kP1, 8, kA2, 20, 0>(lds, frags, params, nullptr);
  zero_halo<4, 20>(lds + kA2);
  __syncthreads();
  conv_lds<3, kA2, 20, kC6, kC6Stride, 32>(lds, frags, params, nullptr);
  zero_halo<4, kC6Stride>(lds + kC6);
  __syncthreads();
  pool_lds<4, 16, kC6, kC6Stride, 32, kP2, 16>(lds);
  zero_halo<2, 16>(lds + kP2);
  __syncthreads();
  conv_lds<4, kP2, 16, kA4, 32, 0>(lds, frags, params, nullptr);
  zero_halo<2, 32>(lds + kA4);
  __syncthreads();
  conv_lds<5, kA4, 32, kA5, 32, 0>(lds, frags, params, nullptr);
  __syncthreads();
  up_lds<2, 32, kA5, 32, kC6, kC6Stride, 0>(lds);
  __syncthreads();
  conv_lds<6, kC6, kC6Stride, kA6, 20, 0>(lds, frags, params, nullptr);
  zero_halo<4, 20>(lds + kA6);
  __syncthreads();
  conv_lds<7, kA6, 20, kA7, 20, 0>(lds, frags, params, nullptr);
  __syncthreads();
  up_lds<4, 16, kA7, 20, kC8, kC8Stride, 0>(lds);
  __syncthreads();
  conv_lds<8, kC8, kC8Stride, kA8, 8, 0>(lds, frags, params, nullptr);
  zero_halo<8, 8>(lds + kA8);
  __syncthreads();
  conv_lds<9, kA8, 8, 0, 4, 0>(lds, frags, params, rec + kActOut);
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

int conv_fwd(hipStream_t s, float* act, uint64_t nt, int l, const float* params) {
  const LayerSpec& L = kLayers[l];
  const float* w = params + layer_offset(l);
  const float* b = w + layer_weights(l);
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

int net_forward(hipStream_t s, const float* params, float* act, uint64_t nt) {
  if (!nt) return FL_OK;
  FL_TRY(conv_fwd(s, act, nt, 0, params));
  FL_TRY(conv_fwd(s, act, nt, 1, params));
  k_pool_fwd<<<grid_for(nt * 8 * 64), kThreads, 0, s>>>(act, nt, 8, 4, kActA1, kActP1);
  FL_TRY(conv_fwd(s, act, nt, 2, params));
  FL_TRY(conv_fwd(s, act, nt, 3, params));
  k_pool_fwd<<<grid_for(nt * 16 * 8), kThreads, 0, s>>>(act, nt, 16, 2, kActA3, kActP2);
  FL_TRY(conv_fwd(s, act, nt, 4, params));
  FL_TRY(conv_fwd(s, act, nt, 5, params));
  k_up_fwd<<<grid_for(nt * 32 * 64), kThreads, 0, s>>>(act, nt, 32, 2, kActA5, kActU5);
  FL_TRY(conv_fwd(s, act, nt, 6, params));
  FL_TRY(conv_fwd(s, act, nt, 7, params));
  k_up_fwd<<<grid_for(nt * 16 * 512), kThreads, 0, s>>>(act, nt, 16, 4, kActA7, kActU7);
  FL_TRY(conv_fwd(s, act, nt, 8, params));
  FL_TRY(conv_fwd(s, act, nt, 9, params));
  FL_HIP(hipGetLastError());
  return FL_OK;
}

uint64_t net_fragment_bytes() { return uint64_t(frag_offset(kNetLayers)) * sizeof(bf16x4); }

int net_infer(hipStream_t s, const float* params, void* frags, float* act, uint64_t nt) {
  if (!nt) return FL_OK;
  if (nt > 0x7fffffffu) return FL_RESOURCE_EXHAUSTED;
  FragSpec spec{};
  for (int l = 0; l < kNetLayers; ++l) {
    spec.cin[l] = kLayers[l].ca + kLayers[l].cb;
    spec.cout[l] = kLayers[l].cout;
    spec.steps[l] = frag_steps(l);
    spec.first[l] = frag_offset(l);
    spec.param[l] = param_offset(l);
  }
  spec.first[kNetLayers] = frag_offset(kNetLayers);
  bf16x4* f = static_cast<bf16x4*>(frags);
  k_pack_fragments<<<(spec.first[kNetLayers] + 255) / 256, 256, 0, s>>>(params, spec, f);
  k_unet_infer<<<static_cast<unsigned>(nt), 256, 0, s>>>(act, f, params);
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
  lfree(L->frags);
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
  A(&L->frags, fl::net_fragment_bytes());
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
  if (mfma) {
    FL_TRY(fl::net_infer(s, L->params, L->frags, L->act, n));
  } else {
    FL_TRY(fl::net_forward(s, L->params, L->act, n));
  }
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
