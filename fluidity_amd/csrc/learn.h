// Learned preconditioner network: activation layout of one 8^3 tile and the device entry points.
#pragma once

#include "context.h"

namespace fl {

constexpr int kNetLayers = 10;
// Float offsets of each activation inside a tile's record ([channel][voxel]).
constexpr int kActIn = 0;                  // 4 x 8^3: normalised residual, liquid, air, solid
constexpr int kActA0 = kActIn + 4 * 512;   // 8 x 8^3
constexpr int kActA1 = kActA0 + 8 * 512;   // 8 x 8^3 (skip s1)
constexpr int kActP1 = kActA1 + 8 * 512;   // 8 x 4^3
constexpr int kActA2 = kActP1 + 8 * 64;    // 16 x 4^3
constexpr int kActA3 = kActA2 + 16 * 64;   // 16 x 4^3 (skip s2)
constexpr int kActP2 = kActA3 + 16 * 64;   // 16 x 2^3
constexpr int kActA4 = kActP2 + 16 * 8;    // 32 x 2^3
constexpr int kActA5 = kActA4 + 32 * 8;    // 32 x 2^3
constexpr int kActU5 = kActA5 + 32 * 8;    // 32 x 4^3
constexpr int kActA6 = kActU5 + 32 * 64;   // 16 x 4^3
constexpr int kActA7 = kActA6 + 16 * 64;   // 16 x 4^3
constexpr int kActU7 = kActA7 + 16 * 64;   // 16 x 8^3
constexpr int kActA8 = kActU7 + 16 * 512;  // 8 x 8^3
constexpr int kActOut = kActA8 + 8 * 512;  // 1 x 8^3
constexpr int kActFloats = kActOut + 512;  // 30,336 floats per tile

// Network state of a context: parameters and Adam moments, activation and gradient records for up to
// tile_cap tiles, reduction scratch, and the training set of (input, target) tile pairs.
struct LearnState {
  float* params = nullptr;
  float* gparams = nullptr;
  float* m1 = nullptr;
  float* m2 = nullptr;
  uint64_t np = 0;
  int step = 0;
  float* act = nullptr;
  float* grad = nullptr;
  uint64_t tile_cap = 0;
  float* partial = nullptr;
  uint64_t partial_floats = 0;
  float* scalars = nullptr;  // [8]
  float* scale = nullptr;    // [tile_cap] residual scale of each tile
  float* batch_target = nullptr;  // [tile_cap][512]
  int32_t* batch_index = nullptr;  // [tile_cap]
  float* stage = nullptr;          // [min(tile_cap, 1024)][2048] host staging of the span-based calls
  float* data_in = nullptr;        // [data_cap][4][512]
  float* data_target = nullptr;    // [data_cap][512]
  uint64_t data_cap = 0, ndata = 0;
  int collect_every = 0;  // > 0: the MGPCG solver appends (residual, correction) pairs every k-th iteration
  uint8_t* frags = nullptr;  // [net_fragment_bytes()] BF16 weights in matrix-core operand order
};

uint64_t net_param_count();
// FP32 forward pass over nt tiles whose inputs are at kActIn of each record (training, and the FP32
// reference of the inference path); every activation is kept in the records for the backward pass.
int net_forward(hipStream_t s, const float* params, float* act, uint64_t nt);
// BF16 inference over nt tiles: reads the inputs at kActIn and writes the output at kActOut of each
// record, with the whole network of a tile in one workgroup on the matrix cores. frags holds
// net_fragment_bytes() bytes of device memory; the weights are repacked into it on every call.
uint64_t net_fragment_bytes();
int net_infer(hipStream_t s, const float* params, void* frags, float* act, uint64_t nt);
// Masked loss: scalars[0] = sum of the liquid mask, scalars[1] = loss; writes dL/dout into grad.
int net_loss(hipStream_t s, const float* act, const float* target, float* grad, uint64_t nt, float* partial,
             float* scalars);
// Backward pass: grad must be zero except dL/dout at kActOut; writes every parameter gradient and
// leaves dL/dinput at kActIn of each gradient record.
int net_backward(hipStream_t s, const float* params, const float* act, float* grad, uint64_t nt, float* partial,
                 uint64_t partial_floats, float* grad_params);
int net_clear(hipStream_t s, float* p, uint64_t n);
int net_adam(hipStream_t s, float* w, const float* g, float* m1, float* m2, uint64_t n, float lr, int step);

}  // namespace fl
