// FP64 restatement of the learned preconditioner's network: a 3-level 3D U-Net on 8^3 tiles
// (channels 8/16/32, 3x3x3 convolutions with zero padding, ReLU, 2x2x2 average pooling, nearest
// upsampling, skip connections by channel concatenation), its masked mean-squared loss and the exact
// gradient by reverse-mode differentiation.
//
// The reference has no learned component, so nothing here follows a reference file; the forward and
// backward passes are pinned by finite-difference gradient checks (tests/test_oracle_unet.py) and
// serve as the checker of the device kernels. Test infrastructure only.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <vector>

#include "oracle.h"

namespace {

// Layer table: input channels (a + b when the input is a concatenation), output channels, cube side.
struct Conv {
  int cin, cout, side;
  bool relu;
};
constexpr Conv kConvs[10] = {
    {4, 8, 8, true},    // 0  encoder 8^3
    {8, 8, 8, true},    // 1  -> skip s1
    {8, 16, 4, true},   // 2  encoder 4^3 (after pooling)
    {16, 16, 4, true},  // 3  -> skip s2
    {16, 32, 2, true},  // 4  bottleneck 2^3 (after pooling)
    {32, 32, 2, true},  // 5
    {48, 16, 4, true},  // 6  up(x5) ++ s2
    {16, 16, 4, true},  // 7
    {24, 8, 8, true},   // 8  up(x7) ++ s1
    {8, 1, 8, false},   // 9  output
};

size_t ParamCount() {
  size_t n = 0;
  for (const Conv& c : kConvs) n += static_cast<size_t>(c.cout) * c.cin * 27 + c.cout;
  return n;
}

struct Params {
  const double* w[10];
  const double* b[10];
  explicit Params(const double* p) {
    for (int l = 0; l < 10; ++l) {
      w[l] = p;
      p += static_cast<size_t>(kConvs[l].cout) * kConvs[l].cin * 27;
      b[l] = p;
      p += kConvs[l].cout;
    }
  }
};

using Vec = std::vector<double>;

// y[co][v] = b[co] + sum_{ci,k} w[co][ci][k] x[ci][v + off(k)], zero outside the cube.
void ConvFwd(const Conv& c, const double* w, const double* b, const double* x, double* y) {
  const int s = c.side, n = s * s * s;
  for (int co = 0; co < c.cout; ++co)
    for (int z = 0; z < s; ++z)
      for (int yy = 0; yy < s; ++yy)
        for (int xx = 0; xx < s; ++xx) {
          double acc = b[co];
          for (int ci = 0; ci < c.cin; ++ci)
            for (int k = 0; k < 27; ++k) {
              const int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
              const int zz = z + dz, y2 = yy + dy, x2 = xx + dx;
              if (zz < 0 || zz >= s || y2 < 0 || y2 >= s || x2 < 0 || x2 >= s) continue;
              acc += w[(co * c.cin + ci) * 27 + k] * x[ci * n + (zz * s + y2) * s + x2];
            }
          y[co * n + (z * s + yy) * s + xx] = c.relu && acc < 0.0 ? 0.0 : acc;
        }
}

// Given g = dL/dy (already multiplied by the ReLU derivative), accumulates dL/dw, dL/db and dL/dx.
void ConvBwd(const Conv& c, const double* w, const double* x, const double* g, double* gw, double* gb, double* gx) {
  const int s = c.side, n = s * s * s;
  for (int co = 0; co < c.cout; ++co)
    for (int z = 0; z < s; ++z)
      for (int yy = 0; yy < s; ++yy)
        for (int xx = 0; xx < s; ++xx) {
          const double gv = g[co * n + (z * s + yy) * s + xx];
          if (gv == 0.0) continue;
          gb[co] += gv;
          for (int ci = 0; ci < c.cin; ++ci)
            for (int k = 0; k < 27; ++k) {
              const int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
              const int zz = z + dz, y2 = yy + dy, x2 = xx + dx;
              if (zz < 0 || zz >= s || y2 < 0 || y2 >= s || x2 < 0 || x2 >= s) continue;
              const int xi = ci * n + (zz * s + y2) * s + x2;
              gw[(co * c.cin + ci) * 27 + k] += gv * x[xi];
              if (gx) gx[xi] += gv * w[(co * c.cin + ci) * 27 + k];
            }
        }
}

void Pool(int ch, int side, const double* x, double* y) {  // side = output side
  const int s2 = 2 * side;
  for (int c = 0; c < ch; ++c)
    for (int z = 0; z < side; ++z)
      for (int yy = 0; yy < side; ++yy)
        for (int xx = 0; xx < side; ++xx) {
          double acc = 0.0;
          for (int o = 0; o < 8; ++o)
            acc += x[c * s2 * s2 * s2 + ((2 * z + (o >> 2)) * s2 + 2 * yy + ((o >> 1) & 1)) * s2 + 2 * xx + (o & 1)];
          y[c * side * side * side + (z * side + yy) * side + xx] = acc * 0.125;
        }
}

void PoolBwd(int ch, int side, const double* g, double* gx) {
  const int s2 = 2 * side;
  for (int c = 0; c < ch; ++c)
    for (int z = 0; z < s2; ++z)
      for (int yy = 0; yy < s2; ++yy)
        for (int xx = 0; xx < s2; ++xx)
          gx[c * s2 * s2 * s2 + (z * s2 + yy) * s2 + xx] +=
              0.125 * g[c * side * side * side + ((z >> 1) * side + (yy >> 1)) * side + (xx >> 1)];
}

void Up(int ch, int side, const double* x, double* y) {  // side = output side
  const int h = side / 2;
  for (int c = 0; c < ch; ++c)
    for (int z = 0; z < side; ++z)
      for (int yy = 0; yy < side; ++yy)
        for (int xx = 0; xx < side; ++xx)
          y[c * side * side * side + (z * side + yy) * side + xx] = x[c * h * h * h + ((z >> 1) * h + (yy >> 1)) * h + (xx >> 1)];
}

void UpBwd(int ch, int side, const double* g, double* gx) {
  const int h = side / 2;
  for (int c = 0; c < ch; ++c)
    for (int z = 0; z < side; ++z)
      for (int yy = 0; yy < side; ++yy)
        for (int xx = 0; xx < side; ++xx)
          gx[c * h * h * h + ((z >> 1) * h + (yy >> 1)) * h + (xx >> 1)] += g[c * side * side * side + (z * side + yy) * side + xx];
}

void ReluMask(const Vec& y, Vec& g) {
  for (size_t i = 0; i < y.size(); ++i)
    if (!(y[i] > 0.0)) g[i] = 0.0;
}

// All activations of one tile.
struct Acts {
  Vec a0, a1, p1, a2, a3, p2, a4, a5, u5, c6, a6, a7, u7, c8, a8, out;
  Acts()
      : a0(8 * 512), a1(8 * 512), p1(8 * 64), a2(16 * 64), a3(16 * 64), p2(16 * 8), a4(32 * 8), a5(32 * 8),
        u5(32 * 64), c6(48 * 64), a6(16 * 64), a7(16 * 64), u7(16 * 512), c8(24 * 512), a8(8 * 512), out(512) {}
};

void Forward(const Params& P, const double* in, Acts& A) {
  ConvFwd(kConvs[0], P.w[0], P.b[0], in, A.a0.data());
  ConvFwd(kConvs[1], P.w[1], P.b[1], A.a0.data(), A.a1.data());
  Pool(8, 4, A.a1.data(), A.p1.data());
  ConvFwd(kConvs[2], P.w[2], P.b[2], A.p1.data(), A.a2.data());
  ConvFwd(kConvs[3], P.w[3], P.b[3], A.a2.data(), A.a3.data());
  Pool(16, 2, A.a3.data(), A.p2.data());
  ConvFwd(kConvs[4], P.w[4], P.b[4], A.p2.data(), A.a4.data());
  ConvFwd(kConvs[5], P.w[5], P.b[5], A.a4.data(), A.a5.data());
  Up(32, 4, A.a5.data(), A.u5.data());
  std::copy(A.u5.begin(), A.u5.end(), A.c6.begin());  // channels 0..31 upsampled, 32..47 skip
  std::copy(A.a3.begin(), A.a3.end(), A.c6.begin() + 32 * 64);
  ConvFwd(kConvs[6], P.w[6], P.b[6], A.c6.data(), A.a6.data());
  ConvFwd(kConvs[7], P.w[7], P.b[7], A.a6.data(), A.a7.data());
  Up(16, 8, A.a7.data(), A.u7.data());
  std::copy(A.u7.begin(), A.u7.end(), A.c8.begin());  // channels 0..15 upsampled, 16..23 skip
  std::copy(A.a1.begin(), A.a1.end(), A.c8.begin() + 16 * 512);
  ConvFwd(kConvs[8], P.w[8], P.b[8], A.c8.data(), A.a8.data());
  ConvFwd(kConvs[9], P.w[9], P.b[9], A.a8.data(), A.out.data());
}

}  // namespace

extern "C" {

size_t flo_unet_param_count(void) { return ParamCount(); }

int flo_unet_forward(const double* params, size_t n, const double* in, double* out) {
  const Params P(params);
  Acts A;
  for (size_t t = 0; t < n; ++t) {
    Forward(P, in + t * 4 * 512, A);
    std::copy(A.out.begin(), A.out.end(), out + t * 512);
  }
  return FLO_OK;
}

// loss = sum_t sum_v m (out - target)^2 / (2 sum m), m = the liquid channel (input channel 1).
int flo_unet_loss_grad(const double* params, size_t n, const double* in, const double* target, double* loss,
                       double* grad_params, double* grad_in) {
  const Params P(params);
  const size_t np = ParamCount();
  std::vector<double> gp(np, 0.0);
  double msum = 0.0;
  for (size_t t = 0; t < n; ++t)
    for (int v = 0; v < 512; ++v) msum += in[t * 2048 + 512 + v];
  const double inv = msum > 0.0 ? 1.0 / msum : 0.0;
  double L = 0.0;
  // Gradient slices of each layer inside gp.
  double* gw[10];
  double* gb[10];
  {
    double* q = gp.data();
    for (int l = 0; l < 10; ++l) {
      gw[l] = q;
      q += static_cast<size_t>(kConvs[l].cout) * kConvs[l].cin * 27;
      gb[l] = q;
      q += kConvs[l].cout;
    }
  }
  Acts A;
  for (size_t t = 0; t < n; ++t) {
    const double* x = in + t * 2048;
    Forward(P, x, A);
    Vec g_out(512);
    for (int v = 0; v < 512; ++v) {
      const double m = x[512 + v], d = A.out[v] - target[t * 512 + v];
      L += 0.5 * m * d * d * inv;
      g_out[v] = m * d * inv;
    }
    Vec g_a8(8 * 512, 0.0), g_c8(24 * 512, 0.0), g_a7(16 * 64, 0.0), g_a6(16 * 64, 0.0), g_c6(48 * 64, 0.0),
        g_a5(32 * 8, 0.0), g_a4(32 * 8, 0.0), g_p2(16 * 8, 0.0), g_a3(16 * 64, 0.0), g_a2(16 * 64, 0.0),
        g_p1(8 * 64, 0.0), g_a1(8 * 512, 0.0), g_a0(8 * 512, 0.0), g_x(4 * 512, 0.0);
    ConvBwd(kConvs[9], P.w[9], A.a8.data(), g_out.data(), gw[9], gb[9], g_a8.data());
    ReluMask(A.a8, g_a8);
    ConvBwd(kConvs[8], P.w[8], A.c8.data(), g_a8.data(), gw[8], gb[8], g_c8.data());
    Vec g_u7(g_c8.begin(), g_c8.begin() + 16 * 512);
    for (int i = 0; i < 8 * 512; ++i) g_a1[i] += g_c8[16 * 512 + i];
    UpBwd(16, 8, g_u7.data(), g_a7.data());
    ReluMask(A.a7, g_a7);
    ConvBwd(kConvs[7], P.w[7], A.a6.data(), g_a7.data(), gw[7], gb[7], g_a6.data());
    ReluMask(A.a6, g_a6);
    ConvBwd(kConvs[6], P.w[6], A.c6.data(), g_a6.data(), gw[6], gb[6], g_c6.data());
    Vec g_u5(g_c6.begin(), g_c6.begin() + 32 * 64);
    for (int i = 0; i < 16 * 64; ++i) g_a3[i] += g_c6[32 * 64 + i];
    UpBwd(32, 4, g_u5.data(), g_a5.data());
    ReluMask(A.a5, g_a5);
    ConvBwd(kConvs[5], P.w[5], A.a4.data(), g_a5.data(), gw[5], gb[5], g_a4.data());
    ReluMask(A.a4, g_a4);
    ConvBwd(kConvs[4], P.w[4], A.p2.data(), g_a4.data(), gw[4], gb[4], g_p2.data());
    PoolBwd(16, 2, g_p2.data(), g_a3.data());
    ReluMask(A.a3, g_a3);
    ConvBwd(kConvs[3], P.w[3], A.a2.data(), g_a3.data(), gw[3], gb[3], g_a2.data());
    ReluMask(A.a2, g_a2);
    ConvBwd(kConvs[2], P.w[2], A.p1.data(), g_a2.data(), gw[2], gb[2], g_p1.data());
    PoolBwd(8, 4, g_p1.data(), g_a1.data());
    ReluMask(A.a1, g_a1);
    ConvBwd(kConvs[1], P.w[1], A.a0.data(), g_a1.data(), gw[1], gb[1], g_a0.data());
    ReluMask(A.a0, g_a0);
    ConvBwd(kConvs[0], P.w[0], x, g_a0.data(), gw[0], gb[0], grad_in ? g_x.data() : nullptr);
    if (grad_in) std::copy(g_x.begin(), g_x.end(), grad_in + t * 2048);
  }
  *loss = L;
  if (grad_params) std::copy(gp.begin(), gp.end(), grad_params);
  return FLO_OK;
}

}  // extern "C"
