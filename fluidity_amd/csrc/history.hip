// Topology preparation on the device: the checks of a previous state's topology records, the
// particle and pressure-cell input checks, canonical unique pressure cells, and the remap of block,
// cell and face history by identity (native/src/liquid_step_workspace_3d.cc:171-341; the record
// checks of ValidateLiquidState3D, native/src/liquid_state_3d.cc:559-604).
//
// Every check writes min(phase, ordinal, code) into the context's error word, so the first failing
// record of the first failing phase is reported, as the reference's sequential loops do. Lookups in
// the previous state are binary searches over its sorted records; pressure cells are made unique
// with a bitmap over the active blocks' box (one bit per cell, 64 cells per word, in (z,y,x) order),
// whose word population counts are scanned and compacted.
#include "context.h"

namespace fl {
namespace {

__device__ inline bool key_less(fl_int3 a, fl_int3 b) {
  if (a.z != b.z) return a.z < b.z;
  if (a.y != b.y) return a.y < b.y;
  return a.x < b.x;
}
__device__ inline bool key_eq(fl_int3 a, fl_int3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
__device__ inline fl_int3 owner_of(fl_int3 l) {
  return fl_int3{int32_t(floor_div4(l.x)), int32_t(floor_div4(l.y)), int32_t(floor_div4(l.z))};
}

// Ordinal of the block with key k in the sorted records, or n.
__device__ inline uint64_t find_block(const fl_block_state* b, uint64_t n, fl_int3 k) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (key_less(b[mid].block_index, k)) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && key_eq(b[lo].block_index, k)) ? lo : n;
}

constexpr int kPhaseShift = 40;  // ordinals below 2^40; the phase orders the record kinds
__device__ inline void fail(uint64_t* err, uint64_t phase, uint64_t ordinal, int code) {
  atomicMin(reinterpret_cast<unsigned long long*>(err),
            static_cast<unsigned long long>(error_word((phase << kPhaseShift) | ordinal, code)));
}

// Previous blocks: coordinates inside the int32 lattice, strictly increasing (z,y,x) keys.
__global__ void k_check_blocks(const fl_block_state* b, uint64_t nb, uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nb) return;
  const fl_int3 k = b[i].block_index;
  const int32_t c[3] = {k.x, k.y, k.z};
  for (int a = 0; a < 3; ++a)
    if (c[a] < kI32Min / 4 || c[a] > (kI32Max - 3) / 4) {
      fail(err, 0, i, kOutOfRange);
      return;
    }
  if (i > 0 && !key_less(b[i - 1].block_index, k)) fail(err, 0, i, kInvalidArgument);
}

// Previous cells: strictly increasing keys, an active owner, an occupancy and a finite pressure.
__global__ void k_check_cells(const fl_cell_state* c, uint64_t nc, const fl_block_state* b, uint64_t nb,
                              uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nc) return;
  const fl_cell_state r = c[i];
  if ((i > 0 && !key_less(c[i - 1].cell_index, r.cell_index)) || find_block(b, nb, owner_of(r.cell_index)) == nb ||
      r.occupancy > 2 || !finite(r.previous_pressure))
    fail(err, 1, i, kInvalidArgument);
}

// Previous faces: axis, validity and velocity; an active owner; the owned slot of the face's index,
// strictly increasing.
__global__ void k_check_faces(const fl_face_state* f, uint64_t nf, const fl_block_state* b, uint64_t nb,
                              uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nf) return;
  const fl_face_state r = f[i];
  if (r.axis > 2 || r.validity > 3 || !finite(r.carrier_velocity)) {
    fail(err, 2, i, kInvalidArgument);
    return;
  }
  const uint64_t o = find_block(b, nb, owner_of(r.lattice));
  if (o == nb) {
    fail(err, 2, i, kInvalidArgument);
    return;
  }
  const fl_int3 k = b[o].block_index;
  const uint64_t lx = uint64_t(int64_t(r.lattice.x) - int64_t(k.x) * 4);
  const uint64_t ly = uint64_t(int64_t(r.lattice.y) - int64_t(k.y) * 4);
  const uint64_t lz = uint64_t(int64_t(r.lattice.z) - int64_t(k.z) * 4);
  const uint64_t expected = o * 192 + uint64_t(r.axis) * 64 + lz * 16 + ly * 4 + lx;
  if (r.face_slot != expected || (i > 0 && f[i - 1].face_slot >= expected)) fail(err, 2, i, kInvalidArgument);
}

// Particle checks of PrepareTopology in particle order: finite position, representable bin, then the
// geometry stencils of the three face axes (positions are both sorted and trial positions).
template <bool kPow2>
__global__ void k_check_particles(const double* sx, const double* sy, const double* sz, uint64_t n, Geometry g,
                                  uint64_t* err) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const double s[3] = {sx[i], sy[i], sz[i]};
  if (!(finite(s[0]) && finite(s[1]) && finite(s[2]))) {
    fail(err, 3, i, kInvalidArgument);
    return;
  }
  CoordT<kPow2> c[3];
  for (int a = 0; a < 3; ++a) {
    c[a] = normalize<kPow2>(s[a], comp(g, a), g.h, g.inv_h);
    if (!finite(c[a].X)) {
      fail(err, 3, i, kOutOfRange);
      return;
    }
  }
  int32_t bin[3];
  for (int a = 0; a < 3; ++a) {
    const int code = checked_bin(c[a], bin[a]);
    if (code) {
      fail(err, 3, i, code);
      return;
    }
  }
  for (int a = 0; a < 3; ++a)
    for (int kind = 0; kind < 2; ++kind) {
      AxisStencil st;
      const int code = axis_stencil(c[a], kind == 1 ? 0.0 : 0.5, bin[a], g.h, st);
      if (code) {
        fail(err, 3, i, code);
        return;
      }
    }
}

// Pressure cells: no coordinate may be INT32_MAX (liquid_step_workspace_3d.cc:225-231).
__global__ void k_check_cell_input(const int32_t* cells, uint64_t nc, uint64_t* err) {
  const uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (j >= nc) return;
  if (cells[3 * j] == kI32Max || cells[3 * j + 1] == kI32Max || cells[3 * j + 2] == kI32Max)
    fail(err, 4, j, kOutOfRange);
}

// Owner-block range of the pressure cells' six faces: min of the cell's owner, max of its +1 owner.
__global__ void k_cell_owner_box(const int32_t* cells, uint64_t nc, int32_t* red) {
  const uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (j >= nc) return;
  for (int a = 0; a < 3; ++a) {
    const int64_t c = cells[3 * j + a];
    atomicMin(&red[a], int32_t(floor_div4(c)));
    atomicMax(&red[3 + a], int32_t(floor_div4(c + 1)));
  }
}

__global__ void k_box_init(int32_t* red) {
  if (threadIdx.x < 3) {
    red[threadIdx.x] = 0x7fffffff;
    red[3 + threadIdx.x] = -0x7fffffff - 1;
  }
}

__global__ void k_clear_words(unsigned long long* w, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    w[i] = 0ull;
}

__global__ void k_mark_cells(const int32_t* cells, uint64_t nc, Box box, unsigned long long* words) {
  const uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (j >= nc) return;
  const int64_t idx = box.index(cells[3 * j], cells[3 * j + 1], cells[3 * j + 2]);
  if (idx >= 0) atomicOr(&words[idx >> 6], 1ull << (idx & 63));
}

__global__ void k_word_counts(const unsigned long long* words, uint64_t nw, uint64_t* counts) {
  const uint64_t w = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (w < nw) counts[w] = static_cast<uint64_t>(__popcll(words[w]));
}

// Cells of every set bit, in bit order, at the word's scanned position; the total into *count.
__global__ void k_compact_cells(const unsigned long long* words, const uint64_t* scan, uint64_t nw, Box box,
                                int32_t* out, uint64_t* count) {
  const uint64_t w = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (w >= nw) return;
  unsigned long long bits = words[w];
  uint64_t pos = scan[w];
  if (w == nw - 1) *count = pos + static_cast<uint64_t>(__popcll(bits));
  const uint64_t plane = uint64_t(box.dim[0]) * box.dim[1];
  while (bits) {
    const int b = __ffsll(bits) - 1;
    bits &= bits - 1;
    const uint64_t idx = w * 64 + uint64_t(b);
    out[3 * pos] = box.lo[0] + int32_t(idx % box.dim[0]);
    out[3 * pos + 1] = box.lo[1] + int32_t((idx / box.dim[0]) % box.dim[1]);
    out[3 * pos + 2] = box.lo[2] + int32_t(idx / plane);
    ++pos;
  }
}

// Blocks: the previous record with its inactive age cleared, or a new record of this generation;
// old[i] keeps the previous ordinal (-1 for a new block) for the face remap.
__global__ void k_remap_blocks(const int32_t* blocks, uint64_t nb, const fl_block_state* prev, uint64_t prev_nb,
                               uint64_t generation, fl_block_state* out, int64_t* old) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nb) return;
  const fl_int3 k{blocks[3 * i], blocks[3 * i + 1], blocks[3 * i + 2]};
  const uint64_t o = find_block(prev, prev_nb, k);
  fl_block_state r{};
  r.block_index = k;
  r.pad = 0;
  r.generation = generation;
  r.activation_age = 0;
  r.inactive_age = 0;
  if (o < prev_nb) {
    r.generation = prev[o].generation;
    r.activation_age = prev[o].activation_age;
  }
  out[i] = r;
  old[i] = o < prev_nb ? int64_t(o) : -1;
}

__global__ void k_remap_cells(const int32_t* cells, uint64_t nc, const fl_cell_state* prev, uint64_t prev_nc,
                              fl_cell_state* out) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= nc) return;
  const fl_int3 k{cells[3 * i], cells[3 * i + 1], cells[3 * i + 2]};
  uint64_t lo = 0, hi = prev_nc;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (key_less(prev[mid].cell_index, k)) lo = mid + 1;
    else hi = mid;
  }
  fl_cell_state r{};
  r.cell_index = k;
  r.occupancy = 1;
  r.pad[0] = r.pad[1] = r.pad[2] = 0;
  r.previous_pressure = 0.0;
  if (lo < prev_nc && key_eq(prev[lo].cell_index, k)) {
    r.occupancy = prev[lo].occupancy;
    r.previous_pressure = prev[lo].previous_pressure;
  }
  out[i] = r;
}

// Faces: slot s of block i is face (axis, 4 * block + local). When block i existed, its previous
// slot is old[i] * 192 + s % 192, and the previous record with that slot (the checks above tie every
// previous record's slot to its identity) is copied with the new slot.
__global__ void k_remap_faces(const int32_t* blocks, uint64_t nb, const int64_t* old, const fl_face_state* prev,
                              uint64_t prev_nf, fl_face_state* out) {
  const uint64_t s = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (s >= nb * 192) return;
  const uint64_t i = s / 192, within = s % 192, local = within % 64;
  fl_face_state r{};
  r.axis = static_cast<uint8_t>(within / 64);
  r.pad0[0] = r.pad0[1] = r.pad0[2] = 0;
  r.lattice = fl_int3{blocks[3 * i] * 4 + int32_t(local % 4), blocks[3 * i + 1] * 4 + int32_t((local / 4) % 4),
                      blocks[3 * i + 2] * 4 + int32_t(local / 16)};
  r.face_slot = s;
  r.carrier_velocity = 0.0;
  r.validity = 0;
  r.side_tag = 0;
  r.structurally_eligible = 1;
  for (int k = 0; k < 7; ++k) r.pad1[k] = r.pad2[k] = 0;
  if (old[i] >= 0) {
    const uint64_t want = uint64_t(old[i]) * 192 + within;
    uint64_t lo = 0, hi = prev_nf;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (prev[mid].face_slot < want) lo = mid + 1;
      else hi = mid;
    }
    if (lo < prev_nf && prev[lo].face_slot == want) {
      r.carrier_velocity = prev[lo].carrier_velocity;
      r.validity = prev[lo].validity;
      r.side_tag = prev[lo].side_tag;
      r.structurally_eligible = prev[lo].structurally_eligible;
    }
  }
  out[s] = r;
}

}  // namespace

int check_history(fl_ctx* ctx, const fl_block_state* b, uint64_t nb, const fl_cell_state* c, uint64_t nc,
                  const fl_face_state* f, uint64_t nf) {
  hipStream_t s = ctx->stream;
  if (nb) k_check_blocks<<<grid_for(nb), kThreads, 0, s>>>(b, nb, ctx->err);
  if (nc) k_check_cells<<<grid_for(nc), kThreads, 0, s>>>(c, nc, b, nb, ctx->err);
  if (nf) k_check_faces<<<grid_for(nf), kThreads, 0, s>>>(f, nf, b, nb, ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int check_topology_inputs(fl_ctx* ctx, uint64_t n, const int32_t* cells, uint64_t nc) {
  hipStream_t s = ctx->stream;
  if (n) {
    const ParticleView v = ctx->raw.view();
    auto* kernel = ctx->geometry.inv_h != 0.0 ? k_check_particles<true> : k_check_particles<false>;
    kernel<<<grid_for(n), kThreads, 0, s>>>(v.s[0], v.s[1], v.s[2], n, ctx->geometry, ctx->err);
  }
  if (nc) k_check_cell_input<<<grid_for(nc), kThreads, 0, s>>>(cells, nc, ctx->err);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int cell_owner_box(fl_ctx* ctx, const int32_t* cells, uint64_t nc, int32_t* d_red, int32_t lo[3], int32_t hi[3]) {
  if (!nc) return FL_OK;
  hipStream_t s = ctx->stream;
  k_box_init<<<1, 64, 0, s>>>(d_red);
  k_cell_owner_box<<<grid_for(nc), kThreads, 0, s>>>(cells, nc, d_red);
  FL_HIP(hipGetLastError());
  int32_t h[6];
  FL_HIP(hipMemcpyAsync(h, d_red, sizeof h, hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  for (int a = 0; a < 3; ++a) {
    lo[a] = h[a];
    hi[a] = h[3 + a];
  }
  return FL_OK;
}

int unique_cells(fl_ctx* ctx, const int32_t* cells, uint64_t nc, void* words_buffer, int32_t* out, uint64_t* d_count,
                 uint64_t* count) {
  *count = 0;
  if (!nc) return FL_OK;
  // Every cell's owner block is active, so the cells lie in four times the blocks' box.
  Box box;
  for (int a = 0; a < 3; ++a) {
    if (ctx->block_box.dim[a] > (1 << 28)) return FL_RESOURCE_EXHAUSTED;
    box.lo[a] = ctx->block_box.lo[a] * 4;
    box.dim[a] = ctx->block_box.dim[a] * 4;
  }
  const uint64_t nw = (box.volume() + 63) / 64;
  if (nw > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  hipStream_t s = ctx->stream;
  auto* words = static_cast<unsigned long long*>(words_buffer);
  k_clear_words<<<grid_for(nw), kThreads, 0, s>>>(words, nw);
  k_mark_cells<<<grid_for(nc), kThreads, 0, s>>>(cells, nc, box, words);
  k_word_counts<<<grid_for(nw), kThreads, 0, s>>>(words, nw, ctx->bin_scan);
  FL_HIP(hipGetLastError());
  FL_TRY(exclusive_scan_u64(ctx, ctx->bin_scan, ctx->bin_scan + nw, nw));
  k_compact_cells<<<grid_for(nw), kThreads, 0, s>>>(words, ctx->bin_scan + nw, nw, box, out, d_count);
  FL_HIP(hipGetLastError());
  FL_HIP(hipMemcpyAsync(ctx->host_words, d_count, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  *count = ctx->host_words[0];
  return FL_OK;
}

int remap_topology(fl_ctx* ctx, const fl_block_state* pb, uint64_t pnb, const fl_cell_state* pc, uint64_t pnc,
                   const fl_face_state* pf, uint64_t pnf, uint64_t generation, const int32_t* cells, uint64_t nc,
                   int64_t* old, fl_block_state* ob, fl_cell_state* oc, fl_face_state* of) {
  hipStream_t s = ctx->stream;
  const uint64_t nb = ctx->nb;
  if (nb) {
    k_remap_blocks<<<grid_for(nb), kThreads, 0, s>>>(ctx->blocks, nb, pb, pnb, generation, ob, old);
    k_remap_faces<<<grid_for(nb * 192), kThreads, 0, s>>>(ctx->blocks, nb, old, pf, pnf, of);
  }
  if (nc) k_remap_cells<<<grid_for(nc), kThreads, 0, s>>>(cells, nc, pc, pnc, oc);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

}  // namespace fl
