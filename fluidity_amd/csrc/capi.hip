// C-ABI entry points (include/fluidity_hip.h): context lifecycle and the span-based drop-ins for
// the reference operators. Argument checks run in the reference's order so that every call
// returns the same status code the CPU operator would; outputs are written only on success.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include <limits>

#include "context.h"

namespace fl {
int project_pressure(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count, double dt,
                     const fl_pressure_cell* cells, size_t cell_count, const fl_pressure_face* faces,
                     size_t face_count, const fl_pressure_options* options, double* out_pressures,
                     double* out_velocities, double* out_rhs, double* out_residuals, double* out_defects,
                     uint64_t* out_gauges, fl_pressure_diagnostics* out_diagnostics, fl_measurement* measurement);
int normalize_forces_stress(fl_ctx* ctx, const int64_t* qs, uint64_t nf, double qm, double qp, double min_mass,
                            const double g[3], double dt);
void destroy_pressure(fl_ctx* ctx);
void destroy_learn(fl_ctx* ctx);
int reserve_pressure(fl_ctx* ctx);
}  // namespace fl

namespace {

using namespace fl;

__global__ void k_split_columns(const double* in, uint64_t n, int stride, int columns, double* base, uint64_t cap) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < columns; ++c) base[c * cap + i] = in[i * stride + c];
}

// Compatible-deposit particle records: the 19 transfer doubles into the columns, then the side tag
// and the affine switch into their own caller-order arrays.
__global__ void k_split_boundary(const fl_boundary_particle* in, uint64_t n, double* base, uint64_t cap,
                                 int64_t* side, uint8_t* affine) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const double* r = reinterpret_cast<const double*>(&in[i].transfer);
  for (int c = 0; c < kParticleDoubles; ++c) base[c * cap + i] = r[c];
  side[i] = in[i].side_tag;
  affine[i] = in[i].affine_enabled;
}

__global__ void k_split_tags(const fl_face_tag* in, uint64_t n, int64_t* side, uint8_t* eligible) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  side[i] = in[i].side_tag;
  eligible[i] = in[i].structurally_eligible;
}

__global__ void k_split_quanta(const fl_face_quanta* in, uint64_t n, int64_t* qm, int64_t* qp, int64_t* qs) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  qm[i] = in[i].mass_quanta_sum;
  qp[i] = in[i].pure_momentum_quanta_sum;
  qs[i] = in[i].stress_force_quanta_sum;
}

__global__ void k_pack_quanta(const int64_t* qm, const int64_t* qp, uint64_t n, fl_face_quanta* out) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  fl_face_quanta q;
  q.mass_quanta_sum = qm[i];
  q.pure_momentum_quanta_sum = qp[i];
  q.stress_force_quanta_sum = 0;
  out[i] = q;
}

__global__ void k_pack_gather(const double* cols, uint64_t cap, const uint32_t* perm, uint64_t n,
                              fl_gather_result* out) {
  const uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (k >= n) return;
  double* r = reinterpret_cast<double*>(out + perm[k]);
  for (int c = 0; c < kGatherDoubles; ++c) r[c] = cols[c * cap + k];
}

bool powers_of_two_ok(const fl_quantization& q) {
  int e;
  return std::isnormal(q.mass_per_quantum) && std::frexp(q.mass_per_quantum, &e) == 0.5 &&
         std::isnormal(q.pure_momentum_per_quantum) && std::frexp(q.pure_momentum_per_quantum, &e) == 0.5;
}

int check_geometry(const fl_geometry* g) {
  if (!g) return FL_INVALID_ARGUMENT;
  if (!std::isfinite(g->world_origin.x) || !std::isfinite(g->world_origin.y) || !std::isfinite(g->world_origin.z))
    return FL_INVALID_ARGUMENT;
  if (!(g->cell_width > 0.0 && std::isnormal(g->cell_width))) return FL_INVALID_ARGUMENT;
  return FL_OK;
}

// SparseMacGridView3D::Create rules (native/src/sparse_mac_grid_3d.cc:557-577).
int check_blocks(const fl_int3* b, size_t n) {
  const int64_t lo = -2147483648LL / 4, hi = 2147483647LL / 4;
  for (size_t i = 0; i < n; ++i) {
    for (int32_t c : {b[i].x, b[i].y, b[i].z})
      if (c < lo || c > hi) return FL_OUT_OF_RANGE;
    if (i) {
      const fl_int3 p = b[i - 1], q = b[i];
      const bool less = p.z != q.z ? p.z < q.z : (p.y != q.y ? p.y < q.y : p.x < q.x);
      if (!less) return FL_INVALID_ARGUMENT;
    }
  }
  return FL_OK;
}

bool overlaps(const void* a, size_t an, const void* b, size_t bn) {
  if (!a || !b || !an || !bn) return false;
  const auto* pa = static_cast<const unsigned char*>(a);
  const auto* pb = static_cast<const unsigned char*>(b);
  return pa < pb + bn && pb < pa + an;
}

// Every span-based call starts here. Those calls reuse the particle, block and face buffers that
// also hold a device-resident scene, so the scene must be uploaded and prepared again afterwards.
void set_geometry(fl_ctx* ctx, const fl_geometry* g) {
  ctx->geometry = make_geometry(g->world_origin.x, g->world_origin.y, g->world_origin.z, g->cell_width);
  ctx->scene_loaded = false;
  ctx->scene_ready = false;
}

// hipEvent interval on the context stream; each timer owns its events so timers may nest.
struct Timer {
  fl_ctx* ctx;
  hipEvent_t a = nullptr, b = nullptr;
  explicit Timer(fl_ctx* c) : ctx(c) {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, ctx->stream);
  }
  ~Timer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  double stop() {
    (void)hipEventRecord(b, ctx->stream);
    (void)hipEventSynchronize(b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
  }
};

void fill_measurement(fl_ctx* ctx, fl_measurement* m, double kernel_ms, double copy_ms, uint32_t launches) {
  if (!m) return;
  m->kernel_ms = kernel_ms;
  m->copy_ms = copy_ms;
  m->launches = launches;
  std::snprintf(m->architecture, sizeof(m->architecture), "%s", ctx->arch);
}

// Device copies of the two velocity histories of a boundary gather, the particle side tags and affine
// switches, and the continuation list; allocated on the first boundary gather and grown with it.
struct BoundaryBuffers {
  uint64_t faces = 0, particles = 0;
  double* u[2] = {};
  uint8_t* validity[2] = {};
  int64_t* side[2] = {};
  uint8_t* eligible[2] = {};
  fl_wall_sample* walls[2] = {};
  int64_t* pside = nullptr;
  uint8_t* paffine = nullptr;
  uint32_t* slow = nullptr;
  uint32_t* slow_count = nullptr;
  void release() {
    void* ptrs[] = {u[0], u[1], validity[0], validity[1], side[0], side[1], eligible[0], eligible[1], walls[0], walls[1],
                    pside, paffine, slow, slow_count};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    *this = BoundaryBuffers{};
  }
};

void destroy_boundary(fl_ctx* ctx) {
  auto* b = static_cast<BoundaryBuffers*>(ctx->boundary);
  if (!b) return;
  b->release();
  delete b;
  ctx->boundary = nullptr;
}

__global__ void k_split_history(const fl_face_tag* tags, uint64_t n, int64_t* side, uint8_t* eligible) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  side[i] = tags[i].side_tag;
  eligible[i] = tags[i].structurally_eligible;
}

__global__ void k_pack_boundary_gather(const double* cols, uint64_t cap, const double* extra, uint64_t extra_cap,
                                       const uint32_t* perm, uint64_t n, fl_boundary_gather* out) {
  const uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (k >= n) return;
  fl_boundary_gather& r = out[perm[k]];
  double* d = reinterpret_cast<double*>(&r);
  // new PIC (0-2), increment (3-5), then the fitted slopes and the gradient after the retention (6-8).
  for (int c = 0; c < 6; ++c) d[c] = cols[c * cap + k];
  for (int c = 0; c < 3; ++c) d[6 + c] = extra[c * extra_cap + k];
  for (int c = 6; c < 24; ++c) d[3 + c] = cols[c * cap + k];
  const int flags = static_cast<int>(extra[3 * extra_cap + k]);
  r.affine_enabled = flags & 1;
  r.old_pic_fallback = (flags >> 1) & 1;
  r.new_pic_fallback = (flags >> 2) & 1;
  r.low_mass_fallback = (flags >> 3) & 1;
  for (int c = 0; c < 4; ++c) r.pad[c] = 0;
}

// Device copies of a previous state's topology records, the pressure cells and the prepared records;
// allocated on the first topology preparation and grown with it.
struct TopologyBuffers {
  fl_block_state *prev_blocks = nullptr, *blocks = nullptr;
  fl_cell_state *prev_cells = nullptr, *cells = nullptr;
  fl_face_state *prev_faces = nullptr, *faces = nullptr;
  int32_t *cells_in = nullptr, *cells_unique = nullptr, *box = nullptr;
  int64_t* old = nullptr;
  uint64_t* words = nullptr;
  uint64_t* count = nullptr;
  uint64_t cap_prev_blocks = 0, cap_blocks = 0, cap_prev_cells = 0, cap_cells = 0, cap_prev_faces = 0, cap_faces = 0,
           cap_cells_in = 0, cap_cells_unique = 0, cap_old = 0, cap_words = 0;
  void release() {
    void* ptrs[] = {prev_blocks, blocks, prev_cells, cells, prev_faces, faces, cells_in, cells_unique, box, old, words,
                    count};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    *this = TopologyBuffers{};
  }
};

// Device copies of one particle update's records; allocated on first use and grown with it.
struct UpdateBuffers {
  fl_liquid_particle* particles = nullptr;
  fl_boundary_gather* gathered = nullptr;
  double* distances = nullptr;
  uint64_t cap_particles = 0, cap_gathered = 0, cap_distances = 0;
  void release() {
    void* ptrs[] = {particles, gathered, distances};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    *this = UpdateBuffers{};
  }
};

void destroy_update(fl_ctx* ctx) {
  auto* b = static_cast<UpdateBuffers*>(ctx->update);
  if (!b) return;
  b->release();
  delete b;
  ctx->update = nullptr;
}

// Device copies of one link build's view, fields and outputs; allocated on first use and grown with it.
struct LinkBuffers {
  fl_int3* blocks = nullptr;
  fl_center_distance* centers = nullptr;
  fl_frozen_cell* cells = nullptr;
  fl_frozen_face* faces = nullptr;
  fl_cell_fields* fields = nullptr;
  double* predicted = nullptr;
  fl_optional_flux* wetted = nullptr;
  fl_air_pressure* air = nullptr;
  fl_internal_interface* interfaces = nullptr;
  fl_pressure_cell* out_cells = nullptr;
  fl_pressure_face* out_faces = nullptr;
  uint64_t* lower = nullptr;
  uint64_t cap[12] = {};
  void release() {
    void* ptrs[] = {blocks, centers, cells, faces, fields, predicted, wetted, air, interfaces, out_cells, out_faces,
                    lower};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    *this = LinkBuffers{};
  }
};

struct FreezeBuffers {
  fl_int3* blocks = nullptr;
  fl_int3* candidates = nullptr;
  fl_center_distance* centers = nullptr;
  fl_center_distance* corners = nullptr;
  uint64_t* keep = nullptr;
  fl_frozen_cell* staged = nullptr;
  uint64_t* slots = nullptr;
  unsigned long long* words = nullptr;
  uint64_t* word_counts = nullptr;
  uint64_t* face_slots = nullptr;
  fl_frozen_cell* cells = nullptr;
  fl_frozen_face* faces = nullptr;
  uint8_t* needs = nullptr;
  uint64_t* counts = nullptr;
  double* phi = nullptr;
  uint64_t* work = nullptr;
  double* uncut = nullptr;
  fl_solid_corner* solids = nullptr;
  double* face_velocities = nullptr;
  double* solid_phi = nullptr;
  uint64_t cap[20] = {};
  void release() {
    void* ptrs[] = {blocks, candidates, centers, corners, keep,  staged, slots, words,  word_counts,     face_slots,
                    cells,  faces,      needs,   counts,  phi,   work,   uncut, solids, face_velocities, solid_phi};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    *this = FreezeBuffers{};
  }
};

// Device arrays of the viscosity operators; allocated on first use (or by fl_reserve) and grown with them.
struct ViscosityBuffers {
  fl_viscosity_cell* cells = nullptr;
  fl_strain_row* rows = nullptr;
  uint64_t *counts = nullptr, *offsets = nullptr, *entries = nullptr, *row_nnz = nullptr, *row_ptr = nullptr,
           *inc_ptr = nullptr;
  double *inertia = nullptr, *rhs = nullptr, *v = nullptr, *r = nullptr, *z = nullptr, *d = nullptr, *q = nullptr,
         *diag = nullptr, *strain = nullptr, *partial = nullptr, *scal = nullptr, *term_coef = nullptr,
         *inc_coef = nullptr;
  uint32_t *flag = nullptr, *inc_count = nullptr, *cursor = nullptr, *term_slot = nullptr, *inc_row = nullptr;
  uint64_t cap[28] = {};
  void release() {
    void* ptrs[] = {cells, rows,   counts, offsets, entries, row_nnz,   row_ptr,  inc_ptr, inertia,
                    rhs,   v,      r,      z,       d,       q,         diag,     strain,  partial,
                    scal,  term_coef, inc_coef, flag, inc_count, cursor, term_slot, inc_row};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    *this = ViscosityBuffers{};
  }
};

void destroy_viscosity(fl_ctx* ctx) {
  auto* b = static_cast<ViscosityBuffers*>(ctx->viscosity);
  if (!b) return;
  b->release();
  delete b;
  ctx->viscosity = nullptr;
}

void destroy_freeze(fl_ctx* ctx) {
  auto* b = static_cast<FreezeBuffers*>(ctx->freeze);
  if (!b) return;
  b->release();
  delete b;
  ctx->freeze = nullptr;
}

void destroy_links(fl_ctx* ctx) {
  auto* b = static_cast<LinkBuffers*>(ctx->links);
  if (!b) return;
  b->release();
  delete b;
  ctx->links = nullptr;
}

void destroy_topology(fl_ctx* ctx) {
  auto* b = static_cast<TopologyBuffers*>(ctx->topology);
  if (!b) return;
  b->release();
  delete b;
  ctx->topology = nullptr;
}

// (Re)allocates *p for at least need elements (capacity in *cap).
template <typename T>
int grow(T** p, uint64_t* cap, uint64_t need) {
  if (need <= *cap && *p) return FL_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  const uint64_t n = std::max<uint64_t>(need, 1);
  if (hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return FL_RESOURCE_EXHAUSTED;
  }
  *cap = n;
  return FL_OK;
}

// Viscosity arrays for n cells, nr strain rows and nf faces.
int grow_viscosity(fl_ctx* ctx, uint64_t n, uint64_t nr, uint64_t nf) {
  if (!ctx->viscosity) ctx->viscosity = new (std::nothrow) ViscosityBuffers();
  auto* B = static_cast<ViscosityBuffers*>(ctx->viscosity);
  if (!B) return FL_RESOURCE_EXHAUSTED;
  FL_TRY(grow(&B->cells, &B->cap[0], n));
  FL_TRY(grow(&B->rows, &B->cap[1], nr));
  FL_TRY(grow(&B->counts, &B->cap[2], std::max(n, nf)));
  FL_TRY(grow(&B->offsets, &B->cap[3], n));
  FL_TRY(grow(&B->row_nnz, &B->cap[4], nr));
  FL_TRY(grow(&B->row_ptr, &B->cap[5], nr + 1));
  FL_TRY(grow(&B->inc_ptr, &B->cap[6], nf + 1));
  double** fv[] = {&B->inertia, &B->rhs, &B->v, &B->r, &B->z, &B->d, &B->q, &B->diag};
  for (int i = 0; i < 8; ++i) FL_TRY(grow(fv[i], &B->cap[7 + i], nf));
  FL_TRY(grow(&B->strain, &B->cap[15], nr));
  FL_TRY(grow(&B->partial, &B->cap[16], 2048));  // two partial sets of up to 1024 workgroups
  FL_TRY(grow(&B->scal, &B->cap[17], 8));
  FL_TRY(grow(&B->flag, &B->cap[18], 2));
  FL_TRY(grow(&B->inc_count, &B->cap[19], nf));
  FL_TRY(grow(&B->cursor, &B->cap[20], nf));
  return FL_OK;
}

// The packed operator's arrays for nt terms (contents not kept).
int grow_viscosity_terms(fl_ctx* ctx, uint64_t nt) {
  auto* B = static_cast<ViscosityBuffers*>(ctx->viscosity);
  FL_TRY(grow(&B->entries, &B->cap[21], nt));
  FL_TRY(grow(&B->term_slot, &B->cap[22], nt));
  FL_TRY(grow(&B->term_coef, &B->cap[23], nt));
  FL_TRY(grow(&B->inc_row, &B->cap[24], nt));
  FL_TRY(grow(&B->inc_coef, &B->cap[25], nt));
  return FL_OK;
}

// Whether the context's scan partials hold an exclusive scan of n values (the tile sums of two levels).
bool scan_fits(fl_ctx* ctx, uint64_t n) {
  const uint64_t partials = std::max<uint64_t>(std::max<uint64_t>(64, ctx->cap.lookup_blocks) / 128 + 4096, 8192);
  return 2 * (n / 1024 + 2) + 16 <= partials;
}

ViscosityDevice viscosity_device(fl_ctx* ctx) {
  auto& B = *static_cast<ViscosityBuffers*>(ctx->viscosity);
  ViscosityDevice d{};
  d.cells = B.cells;
  d.rows = B.rows;
  d.row_capacity = B.cap[1];
  d.counts = B.counts;
  d.offsets = B.offsets;
  d.inertia = B.inertia;
  d.rhs = B.rhs;
  d.v = B.v, d.r = B.r, d.z = B.z, d.d = B.d, d.q = B.q, d.diag = B.diag, d.strain = B.strain;
  d.partial = B.partial, d.partial_capacity = B.cap[16], d.scal = B.scal, d.flag = B.flag;
  d.inc_count = B.inc_count, d.cursor = B.cursor;
  d.row_nnz = B.row_nnz, d.row_ptr = B.row_ptr, d.inc_ptr = B.inc_ptr;
  d.term_capacity = std::min({B.cap[21], B.cap[22], B.cap[23], B.cap[24], B.cap[25]});
  d.term_slot = B.term_slot, d.inc_row = B.inc_row, d.term_coef = B.term_coef, d.inc_coef = B.inc_coef;
  d.entries = B.entries;
  return d;
}

int first_error_code(fl_ctx* ctx) {
  uint64_t w;
  FL_TRY(read_error(ctx, &w));
  return w == ~uint64_t(0) ? FL_OK : static_cast<int>(w & 0xff);
}

// Caller records (the 224 B boundary-gather results are the largest), face quanta, or one particle's
// 81 compatible-stencil records, whichever is largest.
uint64_t staging_bytes(const fl_ctx* c) {
  return std::max<uint64_t>({sizeof(fl_boundary_gather) * std::max<uint64_t>(1, c->cap.particles),
                             24 * 192 * std::max<uint64_t>(1, c->cap.blocks), 81 * sizeof(fl_compatible_sample)});
}

template <typename T>
int dalloc(T** p, uint64_t count) {
  *p = nullptr;
  if (!count) return FL_OK;
  FL_HIP(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
  return FL_OK;
}

int allocate(fl_ctx* c) {
  const uint64_t np = std::max<uint64_t>(1, c->cap.particles);
  const uint64_t nb = std::max<uint64_t>(1, c->cap.blocks);
  const uint64_t lk = std::max<uint64_t>(64, c->cap.lookup_blocks);
  const uint64_t nf = 192 * nb;
  c->raw.capacity = c->sorted.capacity = np;
  FL_TRY(dalloc(&c->raw.base, kParticleDoubles * np));
  FL_TRY(dalloc(&c->sorted.base, kParticleDoubles * np));
  FL_TRY(dalloc(reinterpret_cast<char**>(&c->staging_aos), staging_bytes(c)));
  FL_TRY(dalloc(&c->bin, 3 * np));
  FL_TRY(dalloc(&c->key, np));
  FL_TRY(dalloc(&c->rank, np));
  FL_TRY(dalloc(&c->perm, np));
  FL_TRY(dalloc(&c->perm_tmp, np));
  FL_TRY(dalloc(&c->bin_count, lk));
  FL_TRY(dalloc(&c->bin_scan, 2 * lk));
  FL_TRY(dalloc(&c->bin_key, lk));
  FL_TRY(dalloc(&c->bin_offset, lk));
  FL_TRY(dalloc(&c->bin_n, lk));
  FL_TRY(dalloc(&c->scan_partial, std::max<uint64_t>(lk / 128 + 4096, 8192)));  // also the bin-box partials
  FL_TRY(dalloc(&c->blocks, 3 * nb));
  FL_TRY(dalloc(&c->lut, lk));
  FL_TRY(dalloc(&c->mark, lk));
  FL_TRY(dalloc(&c->qm, nf));
  FL_TRY(dalloc(&c->qp, nf));
  FL_TRY(dalloc(&c->face_mass, nf));
  FL_TRY(dalloc(&c->u0, nf));
  FL_TRY(dalloc(&c->ustar, nf));
  FL_TRY(dalloc(&c->unew, nf));
  FL_TRY(dalloc(&c->validity, nf));
  FL_TRY(dalloc(&c->quanta_aos, nf));
  FL_TRY(dalloc(&c->gather, kGatherDoubles * np));
  FL_TRY(dalloc(&c->gather_aos, np));
  FL_TRY(dalloc(&c->err, 4));
  FL_TRY(dalloc(&c->reduce_i32, 16));
  FL_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->host_words), 64 * sizeof(uint64_t), hipHostMallocDefault));
  return FL_OK;
}

void release(fl_ctx* c) {
  void* ptrs[] = {c->raw.base, c->sorted.base, c->staging_aos, c->bin, c->key, c->rank, c->perm, c->perm_tmp,
                  c->bin_count, c->bin_scan, c->bin_key, c->bin_offset, c->bin_n, c->scan_partial, c->blocks,
                  c->lut, c->mark, c->qm, c->qp, c->face_mass, c->u0, c->ustar, c->unew, c->validity,
                  c->quanta_aos, c->gather, c->gather_aos, c->err, c->reduce_i32};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->host_words) (void)hipHostFree(c->host_words);
}

// Uploads caller particle records, bins them (deposit or gather checks) and sorts them by bin.
int stage_particles(fl_ctx* ctx, const fl_particle* particles, size_t n, bool deposit, double* copy_ms) {
  Timer t(ctx);
  if (n) FL_HIP(hipMemcpyAsync(ctx->staging_aos, particles, n * sizeof(fl_particle), hipMemcpyHostToDevice, ctx->stream));
  *copy_ms += t.stop();
  FL_TRY(unpack_particles(ctx, n));
  FL_TRY(bin_particles(ctx, ctx->raw.view(), n, deposit, !deposit));
  if (ctx->bin_box.volume() > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  return sort_particles(ctx, n, ctx->raw.view(), true);
}

}  // namespace

namespace fl {
int viscosity_reserve(fl_ctx* ctx, uint64_t n, uint64_t nr, uint64_t nf, ViscosityDevice* out) {
  FL_TRY(grow_viscosity(ctx, std::max<uint64_t>(n, 1), std::max<uint64_t>(nr, 1), std::max<uint64_t>(nf, 1)));
  *out = viscosity_device(ctx);
  return FL_OK;
}
int viscosity_reserve_terms(fl_ctx* ctx, uint64_t nt, ViscosityDevice* out) {
  FL_TRY(grow_viscosity_terms(ctx, nt));
  *out = viscosity_device(ctx);
  return FL_OK;
}
}  // namespace fl

extern "C" {

int fl_create(const fl_capacities* capacities, int device, fl_ctx** out) {
  if (!capacities || !out) return FL_INVALID_ARGUMENT;
  if (capacities->particles > 0xffffffffULL || capacities->lookup_blocks > 0x7fffffffULL) return FL_RESOURCE_EXHAUSTED;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return FL_FAILED_PRECONDITION;
  auto* c = new (std::nothrow) fl_ctx();
  if (!c) return FL_RESOURCE_EXHAUSTED;
  c->device = device;
  c->cap = *capacities;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return FL_FAILED_PRECONDITION;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess)
    std::snprintf(c->arch, sizeof(c->arch), "%s", prop.gcnArchName);
  int s = FL_OK;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    s = FL_INTERNAL;
  for (int i = 0; i < 8 && s == FL_OK; ++i)
    if (hipEventCreate(&c->stage[i]) != hipSuccess) s = FL_INTERNAL;
  if (s == FL_OK) s = allocate(c);
  if (s == FL_OK) s = fl::reserve_pressure(c);
  if (s != FL_OK) {
    fl_destroy(c);
    return s == FL_INTERNAL ? FL_RESOURCE_EXHAUSTED : s;
  }
  *out = c;
  return FL_OK;
}

void fl_destroy(fl_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  fl::destroy_pressure(ctx);
  fl::destroy_learn(ctx);
  destroy_boundary(ctx);
  destroy_topology(ctx);
  destroy_update(ctx);
  destroy_links(ctx);
  destroy_freeze(ctx);
  destroy_viscosity(ctx);
  fl::release_field(ctx);
  fl::release_substep(ctx);
  fl::release_coupled(ctx);
  release(ctx);
  for (hipEvent_t e : ctx->stage)
    if (e) (void)hipEventDestroy(e);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int fl_device_name(fl_ctx* ctx, char* out, size_t size) {
  if (!ctx || !out || !size) return FL_INVALID_ARGUMENT;
  std::snprintf(out, size, "%s", ctx->arch);
  return FL_OK;
}

int fl_reserve(fl_ctx* ctx) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  (void)hipSetDevice(ctx->device);
  const fl_capacities& c = ctx->cap;
  const uint64_t nb = c.blocks, nc = c.pressure_cells, nf = 192 * c.blocks, np = c.particles;
  // Topology preparation: previous and prepared records, pressure cells, the cell bitmap.
  if (!ctx->topology) ctx->topology = new (std::nothrow) TopologyBuffers();
  if (!ctx->topology) return FL_RESOURCE_EXHAUSTED;
  {
    TopologyBuffers& B = *static_cast<TopologyBuffers*>(ctx->topology);
    FL_TRY(grow(&B.prev_blocks, &B.cap_prev_blocks, nb));
    FL_TRY(grow(&B.prev_cells, &B.cap_prev_cells, nc));
    FL_TRY(grow(&B.prev_faces, &B.cap_prev_faces, nf));
    FL_TRY(grow(&B.cells_in, &B.cap_cells_in, 3 * nc));
    FL_TRY(grow(&B.cells_unique, &B.cap_cells_unique, 3 * nc));
    if (!B.box) FL_HIP(hipMalloc(reinterpret_cast<void**>(&B.box), 8 * sizeof(int32_t)));
    if (!B.count) FL_HIP(hipMalloc(reinterpret_cast<void**>(&B.count), sizeof(uint64_t)));
    FL_TRY(grow(&B.words, &B.cap_words, c.lookup_blocks));
    FL_TRY(grow(&B.blocks, &B.cap_blocks, nb));
    FL_TRY(grow(&B.old, &B.cap_old, nb));
    FL_TRY(grow(&B.cells, &B.cap_cells, nc));
    FL_TRY(grow(&B.faces, &B.cap_faces, nf));
  }
  // Particle update records.
  if (!ctx->update) ctx->update = new (std::nothrow) UpdateBuffers();
  if (!ctx->update) return FL_RESOURCE_EXHAUSTED;
  {
    UpdateBuffers& B = *static_cast<UpdateBuffers*>(ctx->update);
    FL_TRY(grow(&B.particles, &B.cap_particles, np));
    FL_TRY(grow(&B.gathered, &B.cap_gathered, np));
    FL_TRY(grow(&B.distances, &B.cap_distances, np));
  }
  // Boundary gather histories (two banks of every owned slot) and particle attributes.
  if (!ctx->boundary) ctx->boundary = new (std::nothrow) BoundaryBuffers();
  if (!ctx->boundary) return FL_RESOURCE_EXHAUSTED;
  {
    BoundaryBuffers& B = *static_cast<BoundaryBuffers*>(ctx->boundary);
    const uint64_t need_f = std::max<uint64_t>(nf, 1), need_p = std::max<uint64_t>(np, 1);
    if (B.faces < need_f || B.particles < need_p) {
      B.release();
      for (int k = 0; k < 2; ++k) {
        FL_TRY(dalloc(&B.u[k], need_f));
        FL_TRY(dalloc(&B.validity[k], need_f));
        FL_TRY(dalloc(&B.side[k], need_f));
        FL_TRY(dalloc(&B.eligible[k], need_f));
        FL_TRY(dalloc(&B.walls[k], need_f));
      }
      FL_TRY(dalloc(&B.pside, need_p));
      FL_TRY(dalloc(&B.paffine, need_p));
      FL_TRY(dalloc(&B.slow, need_p));
      FL_TRY(dalloc(&B.slow_count, 1));
      B.faces = need_f;
      B.particles = need_p;
    }
  }
  // Frozen geometry of up to pressure_cells candidates, 27 centre and 8 corner samples each.
  if (!ctx->freeze) ctx->freeze = new (std::nothrow) FreezeBuffers();
  if (!ctx->freeze) return FL_RESOURCE_EXHAUSTED;
  {
    FreezeBuffers& B = *static_cast<FreezeBuffers*>(ctx->freeze);
    const uint64_t nwords = 3 * nb;
    FL_TRY(grow(&B.blocks, &B.cap[0], nb));
    FL_TRY(grow(&B.candidates, &B.cap[1], nc));
    FL_TRY(grow(&B.centers, &B.cap[2], 27 * nc));
    FL_TRY(grow(&B.corners, &B.cap[3], 8 * nc));
    FL_TRY(grow(&B.keep, &B.cap[4], 2 * nc));
    FL_TRY(grow(&B.staged, &B.cap[5], nc));
    FL_TRY(grow(&B.slots, &B.cap[6], 6 * nc));
    FL_TRY(grow(&B.words, &B.cap[7], nwords));
    FL_TRY(grow(&B.word_counts, &B.cap[8], 2 * nwords));
    FL_TRY(grow(&B.face_slots, &B.cap[9], 6 * nc));
    FL_TRY(grow(&B.cells, &B.cap[10], nc));
    FL_TRY(grow(&B.faces, &B.cap[11], 6 * nc));
    FL_TRY(grow(&B.needs, &B.cap[12], nc));
    FL_TRY(grow(&B.counts, &B.cap[13], 4));
    FL_TRY(grow(&B.phi, &B.cap[14], 9 * nc));
    FL_TRY(grow(&B.work, &B.cap[15], nc));
    FL_TRY(grow(&B.uncut, &B.cap[16], 2));
    // A sampled solid: 8 corner samples per candidate, one face velocity per owned slot.
    FL_TRY(grow(&B.solids, &B.cap[17], 8 * nc));
    FL_TRY(grow(&B.face_velocities, &B.cap[18], 192 * nb));
    FL_TRY(grow(&B.solid_phi, &B.cap[19], 8 * nc));
  }
  // Pressure links of a frozen view: blocks, 27 centre samples per cell, cells and faces, the fields.
  if (!ctx->links) ctx->links = new (std::nothrow) LinkBuffers();
  if (!ctx->links) return FL_RESOURCE_EXHAUSTED;
  {
    LinkBuffers& B = *static_cast<LinkBuffers*>(ctx->links);
    const uint64_t pf = c.pressure_faces;
    FL_TRY(grow(&B.blocks, &B.cap[0], nb));
    FL_TRY(grow(&B.centers, &B.cap[1], 27 * nc));
    FL_TRY(grow(&B.cells, &B.cap[2], nc));
    FL_TRY(grow(&B.faces, &B.cap[3], pf));
    FL_TRY(grow(&B.fields, &B.cap[4], nc));
    FL_TRY(grow(&B.predicted, &B.cap[5], nf));
    FL_TRY(grow(&B.wetted, &B.cap[6], pf));
    FL_TRY(grow(&B.air, &B.cap[7], nc));
    FL_TRY(grow(&B.interfaces, &B.cap[8], pf));
    FL_TRY(grow(&B.out_cells, &B.cap[9], nc));
    FL_TRY(grow(&B.out_faces, &B.cap[10], pf));
    FL_TRY(grow(&B.lower, &B.cap[11], 1));
  }
  // Viscosity of up to pressure_cells cells (six strain rows each, at most 18 terms per cell) on the
  // owned face slots.
  FL_TRY(grow_viscosity(ctx, nc, 6 * nc, nf));
  FL_TRY(grow_viscosity_terms(ctx, 18 * nc));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  return FL_OK;
}

int fl_bin_sort(fl_ctx* ctx, const fl_geometry* geometry, const fl_vec3* positions, size_t n, fl_sort_entry* out,
                fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  FL_TRY(check_geometry(geometry));
  if (n > ctx->cap.particles) return FL_RESOURCE_EXHAUSTED;
  if (n && (!positions || !out)) return FL_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  set_geometry(ctx, geometry);
  double copy_ms = 0;
  {
    Timer t(ctx);
    if (n) FL_HIP(hipMemcpyAsync(ctx->staging_aos, positions, n * sizeof(fl_vec3), hipMemcpyHostToDevice, ctx->stream));
    copy_ms += t.stop();
  }
  Timer t(ctx);
  if (n) {
    k_split_columns<<<grid_for(n), kThreads, 0, ctx->stream>>>(reinterpret_cast<const double*>(ctx->staging_aos), n, 3,
                                                              3, ctx->raw.base, ctx->raw.capacity);
    FL_HIP(hipGetLastError());
  }
  ctx->raw.trial_aliases_sorted = true;
  FL_TRY(bin_particles(ctx, ctx->raw.view(), n, false, false));
  FL_TRY(first_error_code(ctx));
  if (ctx->bin_box.volume() > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  FL_TRY(sort_particles(ctx, n, ctx->raw.view(), false));
  auto* d_out = reinterpret_cast<fl_sort_entry*>(ctx->staging_aos);
  FL_TRY(write_sort_entries(ctx, n, d_out));
  const double kernel_ms = t.stop();
  Timer c(ctx);
  if (n) FL_HIP(hipMemcpyAsync(out, d_out, n * sizeof(fl_sort_entry), hipMemcpyDeviceToHost, ctx->stream));
  copy_ms += c.stop();
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 8);
  return FL_OK;
}

int fl_activate(fl_ctx* ctx, const fl_geometry* geometry, const fl_vec3* positions, size_t n,
                const fl_int3* pressure_cells, size_t cell_count, fl_int3* out_blocks, size_t* io_count,
                fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  FL_TRY(check_geometry(geometry));
  if (!io_count) return FL_INVALID_ARGUMENT;
  if (n > ctx->cap.particles || cell_count > ctx->cap.pressure_cells) return FL_RESOURCE_EXHAUSTED;
  if (cell_count * sizeof(fl_int3) > staging_bytes(ctx)) return FL_RESOURCE_EXHAUSTED;
  (void)hipSetDevice(ctx->device);
  set_geometry(ctx, geometry);
  double copy_ms = 0;
  {
    Timer t(ctx);
    if (n) FL_HIP(hipMemcpyAsync(ctx->staging_aos, positions, n * sizeof(fl_vec3), hipMemcpyHostToDevice, ctx->stream));
    copy_ms += t.stop();
  }
  Timer t(ctx);
  if (n) {
    k_split_columns<<<grid_for(n), kThreads, 0, ctx->stream>>>(reinterpret_cast<const double*>(ctx->staging_aos), n, 3,
                                                              3, ctx->raw.base, ctx->raw.capacity);
    FL_HIP(hipGetLastError());
  }
  ctx->raw.trial_aliases_sorted = true;
  FL_TRY(bin_particles(ctx, ctx->raw.view(), n, false, false));
  FL_TRY(first_error_code(ctx));
  // Cells: INT32_MAX coordinates are unrepresentable (liquid_step_workspace_3d.cc:225-231).
  int32_t clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
  for (size_t i = 0; i < cell_count; ++i) {
    const fl_int3 c = pressure_cells[i];
    if (c.x == 2147483647 || c.y == 2147483647 || c.z == 2147483647) return FL_OUT_OF_RANGE;
    const int32_t o[3] = {int32_t(floor_div4(c.x)), int32_t(floor_div4(c.y)), int32_t(floor_div4(c.z))};
    const int32_t p[3] = {int32_t(floor_div4(int64_t(c.x) + 1)), int32_t(floor_div4(int64_t(c.y) + 1)),
                          int32_t(floor_div4(int64_t(c.z) + 1))};
    for (int a = 0; a < 3; ++a) {
      if (i == 0 || o[a] < clo[a]) clo[a] = o[a];
      if (i == 0 || p[a] > chi[a]) chi[a] = p[a];
    }
  }
  if (ctx->bin_box.volume() > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  FL_TRY(sort_particles(ctx, n, ctx->raw.view(), true, 0x7u));
  Box box;
  bool any = false;
  int32_t lo[3], hi[3];
  if (n && ctx->bin_box.volume()) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = ctx->bin_box.lo[a];
      hi[a] = ctx->bin_box.lo[a] + ctx->bin_box.dim[a];  // bins B..B+1 own the stencils
    }
    any = true;
  }
  if (cell_count) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = any ? std::min(lo[a], clo[a]) : clo[a];
      hi[a] = any ? std::max(hi[a], chi[a]) : chi[a];
    }
    any = true;
  }
  if (any)
    for (int a = 0; a < 3; ++a) {
      box.lo[a] = lo[a];
      box.dim[a] = hi[a] - lo[a] + 1;
    }
  if (box.volume() > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  ctx->block_box = box;
  int32_t* d_cells = reinterpret_cast<int32_t*>(ctx->staging_aos);
  if (cell_count)
    FL_HIP(hipMemcpyAsync(d_cells, pressure_cells, cell_count * sizeof(fl_int3), hipMemcpyHostToDevice, ctx->stream));
  FL_TRY(activate_blocks(ctx, ctx->nvalid, d_cells, cell_count));
  FL_TRY(first_error_code(ctx));
  const double kernel_ms = t.stop();
  if (ctx->nb > *io_count) return FL_RESOURCE_EXHAUSTED;
  Timer c(ctx);
  if (ctx->nb) FL_HIP(hipMemcpyAsync(out_blocks, ctx->blocks, ctx->nb * sizeof(fl_int3), hipMemcpyDeviceToHost, ctx->stream));
  copy_ms += c.stop();
  *io_count = ctx->nb;
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 12);
  return FL_OK;
}

int fl_prepare_topology(fl_ctx* ctx, const fl_geometry* geometry, const fl_topology_capacities* capacities,
                        const fl_vec3* positions, size_t n, const fl_int3* pressure_cells, size_t cell_count,
                        const fl_topology_history* previous, fl_sort_entry* out_order, fl_block_state* out_blocks,
                        uint64_t* out_block_count, fl_cell_state* out_cells, uint64_t* out_cell_count,
                        fl_face_state* out_faces, uint64_t* out_face_count, fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  FL_TRY(check_geometry(geometry));
  if (!capacities || !out_block_count || !out_cell_count || !out_face_count) return FL_INVALID_ARGUMENT;
  if ((n && (!positions || !out_order)) || (cell_count && !pressure_cells)) return FL_INVALID_ARGUMENT;
  if (!out_blocks || !out_cells || !out_faces) return FL_INVALID_ARGUMENT;
  if (previous && ((previous->block_count && !previous->blocks) || (previous->cell_count && !previous->cells) ||
                   (previous->face_count && !previous->faces)))
    return FL_INVALID_ARGUMENT;
  // The workspace's capacities (liquid_step_workspace_3d.cc:182-184), then the context's.
  if (n > capacities->particles || cell_count > capacities->cells) return FL_RESOURCE_EXHAUSTED;
  if (n > ctx->cap.particles || cell_count > ctx->cap.pressure_cells) return FL_RESOURCE_EXHAUSTED;
  if (previous && (previous->block_count > ctx->cap.blocks || previous->cell_count > ctx->cap.pressure_cells ||
                   previous->face_count > 192 * ctx->cap.blocks))
    return FL_RESOURCE_EXHAUSTED;
  (void)hipSetDevice(ctx->device);
  if (!ctx->topology) ctx->topology = new TopologyBuffers();
  TopologyBuffers& B = *static_cast<TopologyBuffers*>(ctx->topology);
  hipStream_t s = ctx->stream;
  set_geometry(ctx, geometry);
  double copy_ms = 0, kernel_ms = 0;
  uint64_t generation = 1;
  const uint64_t pnb = previous ? previous->block_count : 0, pnc = previous ? previous->cell_count : 0,
                 pnf = previous ? previous->face_count : 0;
  FL_TRY(grow(&B.prev_blocks, &B.cap_prev_blocks, pnb));
  FL_TRY(grow(&B.prev_cells, &B.cap_prev_cells, pnc));
  FL_TRY(grow(&B.prev_faces, &B.cap_prev_faces, pnf));
  FL_TRY(grow(&B.cells_in, &B.cap_cells_in, 3 * cell_count));
  FL_TRY(grow(&B.cells_unique, &B.cap_cells_unique, 3 * cell_count));
  if (!B.box) FL_HIP(hipMalloc(reinterpret_cast<void**>(&B.box), 8 * sizeof(int32_t)));
  if (!B.count) FL_HIP(hipMalloc(reinterpret_cast<void**>(&B.count), sizeof(uint64_t)));
  FL_TRY(grow(&B.words, &B.cap_words, ctx->cap.lookup_blocks));
  {
    Timer c(ctx);
    if (n) FL_HIP(hipMemcpyAsync(ctx->staging_aos, positions, n * sizeof(fl_vec3), hipMemcpyHostToDevice, s));
    if (cell_count)
      FL_HIP(hipMemcpyAsync(B.cells_in, pressure_cells, cell_count * sizeof(fl_int3), hipMemcpyHostToDevice, s));
    if (pnb) FL_HIP(hipMemcpyAsync(B.prev_blocks, previous->blocks, pnb * sizeof(fl_block_state), hipMemcpyHostToDevice, s));
    if (pnc) FL_HIP(hipMemcpyAsync(B.prev_cells, previous->cells, pnc * sizeof(fl_cell_state), hipMemcpyHostToDevice, s));
    if (pnf) FL_HIP(hipMemcpyAsync(B.prev_faces, previous->faces, pnf * sizeof(fl_face_state), hipMemcpyHostToDevice, s));
    copy_ms += c.stop();
  }
  Timer t(ctx);
  if (previous) {
    const fl_geometry& pg = previous->geometry;
    if (!(pg.world_origin.x == geometry->world_origin.x && pg.world_origin.y == geometry->world_origin.y &&
          pg.world_origin.z == geometry->world_origin.z && pg.cell_width == geometry->cell_width))
      return FL_INVALID_ARGUMENT;
    if (pnb > UINT64_MAX / 192) return FL_OUT_OF_RANGE;
    FL_TRY(reset_error(ctx));
    FL_TRY(check_history(ctx, B.prev_blocks, pnb, B.prev_cells, pnc, B.prev_faces, pnf));
    FL_TRY(first_error_code(ctx));
    if (previous->epoch == UINT64_MAX) return FL_OUT_OF_RANGE;
    generation = previous->epoch + 1;
  }
  if (n) {
    k_split_columns<<<grid_for(n), kThreads, 0, s>>>(reinterpret_cast<const double*>(ctx->staging_aos), n, 3, 3,
                                                    ctx->raw.base, ctx->raw.capacity);
    FL_HIP(hipGetLastError());
  }
  ctx->raw.trial_aliases_sorted = true;
  FL_TRY(reset_error(ctx));
  FL_TRY(check_topology_inputs(ctx, n, B.cells_in, cell_count));
  FL_TRY(first_error_code(ctx));
  FL_TRY(bin_particles(ctx, ctx->raw.view(), n, false, false));
  FL_TRY(first_error_code(ctx));
  if (ctx->bin_box.volume() > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  FL_TRY(sort_particles(ctx, n, ctx->raw.view(), true, 0x7u));
  int32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  bool any = false;
  if (n && ctx->bin_box.volume()) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = ctx->bin_box.lo[a];
      hi[a] = ctx->bin_box.lo[a] + ctx->bin_box.dim[a];  // bins B..B+1 own the stencils
    }
    any = true;
  }
  if (cell_count) {
    int32_t clo[3], chi[3];
    FL_TRY(cell_owner_box(ctx, B.cells_in, cell_count, B.box, clo, chi));
    for (int a = 0; a < 3; ++a) {
      lo[a] = any ? std::min(lo[a], clo[a]) : clo[a];
      hi[a] = any ? std::max(hi[a], chi[a]) : chi[a];
    }
    any = true;
  }
  Box box;
  if (any)
    for (int a = 0; a < 3; ++a) {
      box.lo[a] = lo[a];
      box.dim[a] = hi[a] - lo[a] + 1;
    }
  if (box.volume() > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  ctx->block_box = box;
  FL_TRY(activate_blocks(ctx, ctx->nvalid, B.cells_in, cell_count));
  FL_TRY(first_error_code(ctx));
  const uint64_t nb = ctx->nb;
  if (nb > capacities->blocks || nb > capacities->faces / 192) return FL_RESOURCE_EXHAUSTED;
  uint64_t nc = 0;
  FL_TRY(unique_cells(ctx, B.cells_in, cell_count, B.words, B.cells_unique, B.count, &nc));
  FL_TRY(grow(&B.blocks, &B.cap_blocks, nb));
  FL_TRY(grow(&B.old, &B.cap_old, nb));
  FL_TRY(grow(&B.cells, &B.cap_cells, nc));
  FL_TRY(grow(&B.faces, &B.cap_faces, 192 * nb));
  FL_TRY(remap_topology(ctx, B.prev_blocks, pnb, B.prev_cells, pnc, B.prev_faces, pnf, generation, B.cells_unique, nc,
                        B.old, B.blocks, B.cells, B.faces));
  auto* d_order = reinterpret_cast<fl_sort_entry*>(ctx->staging_aos);
  FL_TRY(write_sort_entries(ctx, n, d_order));
  kernel_ms = t.stop();
  Timer c(ctx);
  if (n) FL_HIP(hipMemcpyAsync(out_order, d_order, n * sizeof(fl_sort_entry), hipMemcpyDeviceToHost, s));
  if (nb) FL_HIP(hipMemcpyAsync(out_blocks, B.blocks, nb * sizeof(fl_block_state), hipMemcpyDeviceToHost, s));
  if (nc) FL_HIP(hipMemcpyAsync(out_cells, B.cells, nc * sizeof(fl_cell_state), hipMemcpyDeviceToHost, s));
  if (nb) FL_HIP(hipMemcpyAsync(out_faces, B.faces, 192 * nb * sizeof(fl_face_state), hipMemcpyDeviceToHost, s));
  copy_ms += c.stop();
  *out_block_count = nb;
  *out_cell_count = nc;
  *out_face_count = 192 * nb;
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 16);
  return FL_OK;
}

int fl_update_particles(fl_ctx* ctx, const fl_particle_update_options* options, const fl_boundary_gather* gathered,
                        const double* liquid_distances, size_t n, fl_liquid_particle* particles,
                        fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (!options || (n && (!gathered || !liquid_distances || !particles))) return FL_INVALID_ARGUMENT;
  if (n && (overlaps(particles, n * sizeof(fl_liquid_particle), gathered, n * sizeof(fl_boundary_gather)) ||
            overlaps(particles, n * sizeof(fl_liquid_particle), liquid_distances, n * sizeof(double))))
    return FL_INVALID_ARGUMENT;
  // An unknown obstacle kind fails the first particle's contact as InvalidArgument, like its other
  // contact input checks.
  if (n && options->obstacle.kind > FL_OBSTACLE_TANK) return FL_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  if (!ctx->update) ctx->update = new UpdateBuffers();
  UpdateBuffers& B = *static_cast<UpdateBuffers*>(ctx->update);
  FL_TRY(grow(&B.particles, &B.cap_particles, n));
  FL_TRY(grow(&B.gathered, &B.cap_gathered, n));
  FL_TRY(grow(&B.distances, &B.cap_distances, n));
  hipStream_t s = ctx->stream;
  double copy_ms = 0;
  {
    Timer c(ctx);
    if (n) {
      FL_HIP(hipMemcpyAsync(B.particles, particles, n * sizeof(fl_liquid_particle), hipMemcpyHostToDevice, s));
      FL_HIP(hipMemcpyAsync(B.gathered, gathered, n * sizeof(fl_boundary_gather), hipMemcpyHostToDevice, s));
      FL_HIP(hipMemcpyAsync(B.distances, liquid_distances, n * sizeof(double), hipMemcpyHostToDevice, s));
    }
    copy_ms += c.stop();
  }
  Timer t(ctx);
  FL_TRY(reset_error(ctx));
  FL_TRY(update_particles(ctx, *options, B.gathered, B.distances, n, B.particles));
  FL_TRY(first_error_code(ctx));
  const double kernel_ms = t.stop();
  Timer c(ctx);
  if (n) FL_HIP(hipMemcpyAsync(particles, B.particles, n * sizeof(fl_liquid_particle), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  copy_ms += c.stop();
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 1);
  return FL_OK;
}

int fl_build_pressure_links(fl_ctx* ctx, const fl_frozen_view* view, const fl_pressure_fields* fields,
                            fl_pressure_cell* out_cells, fl_pressure_face* out_faces, uint64_t* out_lower_theta_count,
                            fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (!view || !fields || !out_lower_theta_count) return FL_INVALID_ARGUMENT;
  const fl_frozen_view& g = *view;
  const fl_pressure_fields& F = *fields;
  if ((g.block_count && !g.blocks) || (g.center_count && !g.center_distances) || (g.cell_count && !g.cells) ||
      (g.face_count && !g.faces) || (g.cell_count && !out_cells) || (g.face_count && !out_faces) ||
      (F.cell_count && !F.cells) || (F.predicted_count && !F.predicted_face_velocities) ||
      (F.wetted_count && !F.wetted_piston_face_fluxes) || (F.air_count && !F.air_cell_pressures) ||
      (F.interface_count && !F.internal_interfaces))
    return FL_INVALID_ARGUMENT;
  // SparseMacGridView3D::Create of the view's keys, then ValidateFrozenView's scalar checks.
  FL_TRY(check_geometry(&g.geometry));
  FL_TRY(check_blocks(g.blocks, g.block_count));
  const auto positive_finite = [](double v) { return std::isfinite(v) && v > 0; };
  if (!positive_finite(g.step_duration) || !positive_finite(g.liquid_density) || !std::isfinite(g.surface_tension) ||
      g.surface_tension < 0)
    return FL_INVALID_ARGUMENT;
  const double h = g.geometry.cell_width, full_area = h * h, full_volume = full_area * h;
  if (!std::isnormal(full_area) || !std::isnormal(full_volume) || !std::isnormal(1 / g.liquid_density))
    return FL_OUT_OF_RANGE;
  (void)hipSetDevice(ctx->device);
  if (!ctx->links) ctx->links = new LinkBuffers();
  LinkBuffers& B = *static_cast<LinkBuffers*>(ctx->links);
  FL_TRY(grow(&B.blocks, &B.cap[0], g.block_count));
  FL_TRY(grow(&B.centers, &B.cap[1], g.center_count));
  FL_TRY(grow(&B.cells, &B.cap[2], g.cell_count));
  FL_TRY(grow(&B.faces, &B.cap[3], g.face_count));
  FL_TRY(grow(&B.fields, &B.cap[4], F.cell_count));
  FL_TRY(grow(&B.predicted, &B.cap[5], F.predicted_count));
  FL_TRY(grow(&B.wetted, &B.cap[6], F.wetted_count));
  FL_TRY(grow(&B.air, &B.cap[7], F.air_count));
  FL_TRY(grow(&B.interfaces, &B.cap[8], F.interface_count));
  FL_TRY(grow(&B.out_cells, &B.cap[9], g.cell_count));
  FL_TRY(grow(&B.out_faces, &B.cap[10], g.face_count));
  FL_TRY(grow(&B.lower, &B.cap[11], 1));
  hipStream_t s = ctx->stream;
  double copy_ms = 0;
  const auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    if (bytes) FL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return FL_OK;
  };
  {
    Timer c(ctx);
    FL_TRY(up(B.blocks, g.blocks, g.block_count * sizeof(fl_int3)));
    FL_TRY(up(B.centers, g.center_distances, g.center_count * sizeof(fl_center_distance)));
    FL_TRY(up(B.cells, g.cells, g.cell_count * sizeof(fl_frozen_cell)));
    FL_TRY(up(B.faces, g.faces, g.face_count * sizeof(fl_frozen_face)));
    FL_TRY(up(B.fields, F.cells, F.cell_count * sizeof(fl_cell_fields)));
    FL_TRY(up(B.predicted, F.predicted_face_velocities, F.predicted_count * sizeof(double)));
    FL_TRY(up(B.wetted, F.wetted_piston_face_fluxes, F.wetted_count * sizeof(fl_optional_flux)));
    FL_TRY(up(B.air, F.air_cell_pressures, F.air_count * sizeof(fl_air_pressure)));
    FL_TRY(up(B.interfaces, F.internal_interfaces, F.interface_count * sizeof(fl_internal_interface)));
    copy_ms += c.stop();
  }
  Timer t(ctx);
  const LinkDevice d{B.blocks, g.block_count, B.centers, g.center_count, B.cells, g.cell_count, B.faces, g.face_count,
                     h, g.step_duration, g.liquid_density, g.surface_tension, B.fields, B.predicted,
                     F.ambient_pressure, B.wetted, F.wetted_count, B.air, F.air_count, B.interfaces,
                     F.interface_count};
  FL_TRY(reset_error(ctx));
  FL_TRY(build_links_device(ctx, d, nullptr, nullptr, nullptr, true));
  FL_TRY(first_error_code(ctx));
  // LinksCore's extents (liquid_boundary_geometry_3d.cc:1085-1094).
  if (F.cell_count != g.cell_count || F.predicted_count != g.face_count ||
      (F.wetted_count && F.wetted_count != g.face_count) || F.air_count > g.center_count ||
      F.interface_count > g.face_count || !std::isfinite(F.ambient_pressure))
    return FL_INVALID_ARGUMENT;
  FL_HIP(hipMemsetAsync(B.lower, 0, sizeof(uint64_t), s));
  FL_TRY(build_links_device(ctx, d, B.out_cells, B.out_faces, B.lower, false));
  FL_TRY(first_error_code(ctx));
  const double kernel_ms = t.stop();
  Timer c(ctx);
  if (g.cell_count)
    FL_HIP(hipMemcpyAsync(out_cells, B.out_cells, g.cell_count * sizeof(fl_pressure_cell), hipMemcpyDeviceToHost, s));
  if (g.face_count)
    FL_HIP(hipMemcpyAsync(out_faces, B.out_faces, g.face_count * sizeof(fl_pressure_face), hipMemcpyDeviceToHost, s));
  FL_HIP(hipMemcpyAsync(ctx->host_words, B.lower, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  copy_ms += c.stop();
  *out_lower_theta_count = ctx->host_words[0];
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 8);
  return FL_OK;
}

int fl_freeze_geometry(fl_ctx* ctx, const fl_freeze_inputs* inputs, fl_frozen_cell* out_cells, uint64_t* out_cell_count,
                       fl_frozen_face* out_faces, uint64_t* out_face_count, double* out_capillary_step_bound,
                       fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (!inputs || !out_cell_count || !out_face_count || !out_capillary_step_bound) return FL_INVALID_ARGUMENT;
  const fl_freeze_inputs& I = *inputs;
  if ((I.block_count && !I.blocks) || (I.cell_count && (!I.cell_indices || !out_cells || !out_faces)) ||
      (I.center_count && !I.center_distances) || (I.corner_count && !I.corner_distances))
    return FL_INVALID_ARGUMENT;
  FL_TRY(check_geometry(&I.geometry));
  FL_TRY(check_blocks(I.blocks, I.block_count));
  // FreezeCore's scalar checks in its order (liquid_boundary_geometry_3d.cc:772-796), with ValidateObstacle
  // (:132-175).
  const auto positive_finite = [](double v) { return std::isfinite(v) && v > 0; };
  if (!positive_finite(I.step_duration) || !positive_finite(I.liquid_density) || !std::isfinite(I.surface_tension) ||
      I.surface_tension < 0)
    return FL_INVALID_ARGUMENT;
  if (I.has_exact_curvature && !std::isfinite(I.exact_curvature)) return FL_INVALID_ARGUMENT;
  if (I.max_cells && I.cell_count > I.max_cells) return FL_RESOURCE_EXHAUSTED;
  const fl_obstacle& o = I.obstacle;
  if (o.kind == FL_OBSTACLE_HALFSPACE || o.kind == FL_OBSTACLE_SLAB) {
    double low = o.coordinate, high = o.coordinate;
    if (o.kind == FL_OBSTACLE_SLAB) {
      if (!positive_finite(o.half_width)) return FL_INVALID_ARGUMENT;
      low = o.coordinate - o.half_width;
      high = o.coordinate + o.half_width;
      if (!std::isfinite(o.coordinate)) return FL_INVALID_ARGUMENT;
      if (!std::isfinite(low) || !std::isfinite(high) || low >= high) return FL_OUT_OF_RANGE;
    }
    const double n[3] = {o.unit_normal.x, o.unit_normal.y, o.unit_normal.z};
    const double v[3] = {o.wall_velocity.x, o.wall_velocity.y, o.wall_velocity.z};
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(n[a]) || !std::isfinite(v[a])) return FL_INVALID_ARGUMENT;
    if (!std::isfinite(low) || std::abs(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] - 1.0) > 1e-12)
      return FL_INVALID_ARGUMENT;
  } else if (o.kind == FL_OBSTACLE_TANK) {
    // CheckTankStep (base/liquid_boundary_math.h:252-276): finite walls and piston, ordered x and z walls,
    // a representable piston pose strictly inside the high-x wall at the step's end.
    const fl_tank& t = o.tank;
    if (!std::isfinite(t.start_low_x) || !std::isfinite(t.high_x) || !std::isfinite(t.floor_y) ||
        !std::isfinite(t.low_z) || !std::isfinite(t.high_z) || !std::isfinite(t.piston_velocity_x))
      return FL_INVALID_ARGUMENT;
    if (t.start_low_x >= t.high_x || t.low_z >= t.high_z) return FL_INVALID_ARGUMENT;
    const double end_x = std::fma(t.piston_velocity_x, I.step_duration, t.start_low_x);
    if (!std::isfinite(t.piston_velocity_x * I.step_duration) || !std::isfinite(end_x) ||
        !std::isfinite(t.high_x - t.start_low_x) || !std::isfinite(t.high_x - end_x) ||
        !std::isfinite(t.high_z - t.low_z))
      return FL_OUT_OF_RANGE;
    if (end_x >= t.high_x) return FL_FAILED_PRECONDITION;
  } else if (o.kind != FL_OBSTACLE_ABSENT) {
    return FL_INVALID_ARGUMENT;
  }
  const double h = I.geometry.cell_width, full_area = h * h, full_volume = full_area * h;
  if (!std::isnormal(full_area) || !std::isnormal(full_volume) || !std::isnormal(1 / I.liquid_density))
    return FL_OUT_OF_RANGE;
  // The sampled solid's checks (:803-818): a half-space or no obstacle, dense face velocities of the
  // grid's size and finite. The reference checks the centre and corner samples first; a failure of
  // either is InvalidArgument too, so the order does not change the status.
  const bool sampled = I.has_sampled_solid != 0;
  if (I.has_sampled_solid > 3) return FL_INVALID_ARGUMENT;
  if (sampled) {
    if ((I.solid_corner_count && !I.solid_corners) || (I.face_velocity_count && !I.face_normal_velocities))
      return FL_INVALID_ARGUMENT;
    bool bad = (o.kind != FL_OBSTACLE_HALFSPACE && o.kind != FL_OBSTACLE_ABSENT) ||
               (I.face_velocity_count != 0 && I.face_velocity_count != 192 * I.block_count);
    for (uint64_t f = 0; !bad && f < I.face_velocity_count; ++f) bad = !std::isfinite(I.face_normal_velocities[f]);
    if (bad) return FL_INVALID_ARGUMENT;
  }
  // The device scans (candidates and bitmap words) share the context's scan partials.
  const uint64_t nwords = 3 * I.block_count;
  const uint64_t partials = std::max<uint64_t>(std::max<uint64_t>(64, ctx->cap.lookup_blocks) / 128 + 4096, 8192);
  for (uint64_t n : {I.cell_count, nwords})
    if (2 * (n / 1024 + 2) + 16 > partials) return FL_RESOURCE_EXHAUSTED;
  (void)hipSetDevice(ctx->device);
  if (!ctx->freeze) ctx->freeze = new FreezeBuffers();
  FreezeBuffers& B = *static_cast<FreezeBuffers*>(ctx->freeze);
  const uint64_t nc_in = I.cell_count;
  FL_TRY(grow(&B.blocks, &B.cap[0], I.block_count));
  FL_TRY(grow(&B.candidates, &B.cap[1], nc_in));
  FL_TRY(grow(&B.centers, &B.cap[2], I.center_count));
  FL_TRY(grow(&B.corners, &B.cap[3], I.corner_count));
  FL_TRY(grow(&B.keep, &B.cap[4], 2 * nc_in));
  FL_TRY(grow(&B.staged, &B.cap[5], nc_in));
  FL_TRY(grow(&B.slots, &B.cap[6], 6 * nc_in));
  FL_TRY(grow(&B.words, &B.cap[7], nwords));
  FL_TRY(grow(&B.word_counts, &B.cap[8], 2 * nwords));
  FL_TRY(grow(&B.face_slots, &B.cap[9], 6 * nc_in));
  FL_TRY(grow(&B.cells, &B.cap[10], nc_in));
  FL_TRY(grow(&B.faces, &B.cap[11], 6 * nc_in));
  FL_TRY(grow(&B.needs, &B.cap[12], nc_in));
  FL_TRY(grow(&B.counts, &B.cap[13], 4));
  FL_TRY(grow(&B.phi, &B.cap[14], 9 * nc_in));
  FL_TRY(grow(&B.work, &B.cap[15], nc_in));
  FL_TRY(grow(&B.uncut, &B.cap[16], 2));
  if (sampled) {
    FL_TRY(grow(&B.solids, &B.cap[17], I.solid_corner_count));
    FL_TRY(grow(&B.face_velocities, &B.cap[18], I.face_velocity_count));
    FL_TRY(grow(&B.solid_phi, &B.cap[19], 8 * nc_in));
  }
  hipStream_t s = ctx->stream;
  double copy_ms = 0;
  const auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    if (bytes) FL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return FL_OK;
  };
  {
    Timer c(ctx);
    FL_TRY(up(B.blocks, I.blocks, I.block_count * sizeof(fl_int3)));
    FL_TRY(up(B.candidates, I.cell_indices, nc_in * sizeof(fl_int3)));
    FL_TRY(up(B.centers, I.center_distances, I.center_count * sizeof(fl_center_distance)));
    FL_TRY(up(B.corners, I.corner_distances, I.corner_count * sizeof(fl_center_distance)));
    if (sampled) {
      FL_TRY(up(B.solids, I.solid_corners, I.solid_corner_count * sizeof(fl_solid_corner)));
      FL_TRY(up(B.face_velocities, I.face_normal_velocities, I.face_velocity_count * sizeof(double)));
    }
    copy_ms += c.stop();
  }
  Timer t(ctx);
  const FreezeDevice d{B.blocks,     I.block_count,
                       I.geometry.world_origin, h,
                       full_volume,  B.candidates,
                       nc_in,        B.centers,
                       I.center_count, B.corners,
                       I.corner_count, o,
                       B.keep,       B.keep + nc_in,
                       B.staged,     B.slots,
                       B.words,      nwords,
                       B.word_counts, B.word_counts + nwords,
                       B.face_slots, B.cells,
                       B.faces,      B.needs,
                       B.counts,     B.phi,
                       B.work,       B.uncut,
                       sampled ? static_cast<int>(I.has_sampled_solid) : 0, sampled ? B.solids : nullptr,
                       sampled ? I.solid_corner_count : 0, sampled ? B.face_velocities : nullptr,
                       sampled ? I.face_velocity_count : 0, sampled ? B.solid_phi : nullptr};
  const auto read_counts = [&]() -> int {
    FL_HIP(hipMemcpyAsync(ctx->host_words, B.counts, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    return FL_OK;
  };
  FL_TRY(reset_error(ctx));
  FL_TRY(freeze_cells_device(ctx, d));
  FL_TRY(first_error_code(ctx));
  FL_TRY(read_counts());
  const uint64_t nc = nc_in ? ctx->host_words[0] : 0, nf = nc_in ? ctx->host_words[1] : 0;
  if (I.max_faces && nf > I.max_faces) return FL_RESOURCE_EXHAUSTED;
  FL_TRY(freeze_faces_device(ctx, d, nc, nf, I.surface_tension != 0));
  FL_TRY(first_error_code(ctx));
  FL_TRY(read_counts());
  const bool capillary = ctx->host_words[2] != 0;
  if (I.has_exact_curvature) {
    if (capillary && I.exact_curvature != 0 && !std::isnormal(I.surface_tension * I.exact_curvature))
      return FL_OUT_OF_RANGE;
    FL_TRY(freeze_curvature_device(ctx, d, nc, true, I.exact_curvature));
  } else {
    FL_TRY(freeze_curvature_device(ctx, d, nc, false, 0.0));
    FL_TRY(first_error_code(ctx));
  }
  double bound = std::numeric_limits<double>::infinity();
  if (capillary) {
    constexpr double kPi = 3.141592653589793238462643383279502884;
    const double square = I.liquid_density * full_volume / (4 * kPi * I.surface_tension);
    if (!std::isnormal(square)) return FL_OUT_OF_RANGE;
    bound = std::sqrt(square);
    if (I.step_duration > bound) return FL_FAILED_PRECONDITION;
  }
  const double kernel_ms = t.stop();
  Timer c(ctx);
  if (nc) FL_HIP(hipMemcpyAsync(out_cells, B.cells, nc * sizeof(fl_frozen_cell), hipMemcpyDeviceToHost, s));
  if (nf) FL_HIP(hipMemcpyAsync(out_faces, B.faces, nf * sizeof(fl_frozen_face), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  copy_ms += c.stop();
  *out_cell_count = nc;
  *out_face_count = nf;
  *out_capillary_step_bound = bound;
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 13);
  return FL_OK;
}

int fl_p2g_quanta(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
                  const fl_particle* particles, size_t n, fl_quantization quantization, fl_face_quanta* out,
                  size_t face_count, fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  FL_TRY(check_geometry(geometry));
  FL_TRY(check_blocks(blocks, block_count));
  if (!powers_of_two_ok(quantization)) return FL_INVALID_ARGUMENT;
  if (quantization.declared_particle_capacity > ctx->cap.particles) return FL_RESOURCE_EXHAUSTED;
  if (n > quantization.declared_particle_capacity) return FL_RESOURCE_EXHAUSTED;
  if (face_count != 192 * block_count) return FL_INVALID_ARGUMENT;
  if (overlaps(out, face_count * sizeof(fl_face_quanta), particles, n * sizeof(fl_particle))) return FL_INVALID_ARGUMENT;
  if (block_count > ctx->cap.blocks) return FL_RESOURCE_EXHAUSTED;
  (void)hipSetDevice(ctx->device);
  set_geometry(ctx, geometry);
  // Contribution limit 2^(62 - bit_width(cap - 1)) (transfers_3d.cc:164-168).
  const uint64_t cap = std::max<uint64_t>(1, quantization.declared_particle_capacity);
  int bits = 0;
  for (uint64_t v = cap - 1; v; v >>= 1) ++bits;
  const double limit = std::ldexp(1.0, 62 - bits);
  double copy_ms = 0;
  FL_TRY(build_lookup(ctx, reinterpret_cast<const int32_t*>(blocks), block_count));
  Timer t(ctx);
  FL_TRY(stage_particles(ctx, particles, n, true, &copy_ms));
  FL_TRY(p2g_quanta(ctx, ctx->nvalid, 1.0 / quantization.mass_per_quantum,
                    1.0 / quantization.pure_momentum_per_quantum, limit));
  FL_TRY(first_error_code(ctx));
  const uint64_t nf = face_count;
  if (nf) {
    k_pack_quanta<<<grid_for(nf), kThreads, 0, ctx->stream>>>(ctx->qm, ctx->qp, nf, ctx->quanta_aos);
    FL_HIP(hipGetLastError());
  }
  const double kernel_ms = t.stop() - copy_ms;
  Timer c(ctx);
  if (nf) FL_HIP(hipMemcpyAsync(out, ctx->quanta_aos, nf * sizeof(fl_face_quanta), hipMemcpyDeviceToHost, ctx->stream));
  copy_ms += c.stop();
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 10);
  return FL_OK;
}

int fl_p2g_compatible(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
                      const fl_boundary_particle* particles, size_t n, const fl_face_tag* tags, size_t tag_count,
                      fl_quantization quantization, fl_face_quanta* out, size_t face_count,
                      fl_compatible_sample* stencils, fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  FL_TRY(check_geometry(geometry));
  FL_TRY(check_blocks(blocks, block_count));
  // Capacity, then quantization, then tag coverage (liquid_boundary_transfer_3d.cc:330-341).
  if (n > ctx->cap.particles || block_count > ctx->cap.blocks) return FL_RESOURCE_EXHAUSTED;
  if (!powers_of_two_ok(quantization)) return FL_INVALID_ARGUMENT;
  if (quantization.declared_particle_capacity > ctx->cap.particles) return FL_RESOURCE_EXHAUSTED;
  if (n > quantization.declared_particle_capacity) return FL_RESOURCE_EXHAUSTED;
  if (face_count != 192 * block_count || tag_count != face_count) return FL_INVALID_ARGUMENT;
  if ((n && !particles) || (face_count && (!out || !tags))) return FL_INVALID_ARGUMENT;
  const size_t pb = n * sizeof(fl_boundary_particle), ob = face_count * sizeof(fl_face_quanta),
               tb = tag_count * sizeof(fl_face_tag), sb = stencils ? n * 81 * sizeof(fl_compatible_sample) : 0;
  if (overlaps(out, ob, particles, pb) || overlaps(out, ob, tags, tb) || overlaps(stencils, sb, particles, pb) ||
      overlaps(stencils, sb, tags, tb) || overlaps(stencils, sb, out, ob))
    return FL_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  set_geometry(ctx, geometry);
  const uint64_t cap = std::max<uint64_t>(1, quantization.declared_particle_capacity);
  int bits = 0;
  for (uint64_t v = cap - 1; v; v >>= 1) ++bits;
  const double limit = std::ldexp(1.0, 62 - bits);
  double copy_ms = 0;
  FL_TRY(build_lookup(ctx, reinterpret_cast<const int32_t*>(blocks), block_count));
  // Caller-order side tags and affine switches live in the gather buffer, slot tags in u0 / validity;
  // neither is in use during a deposit.
  int64_t* side = reinterpret_cast<int64_t*>(ctx->gather);
  uint8_t* affine = reinterpret_cast<uint8_t*>(ctx->gather + ctx->cap.particles);
  int64_t* slot_side = reinterpret_cast<int64_t*>(ctx->u0);
  uint8_t* slot_ok = ctx->validity;
  auto* staged = reinterpret_cast<fl_boundary_particle*>(ctx->staging_aos);
  Timer t(ctx);
  {
    Timer c(ctx);
    if (n) FL_HIP(hipMemcpyAsync(staged, particles, pb, hipMemcpyHostToDevice, ctx->stream));
    copy_ms += c.stop();
  }
  if (n) {
    k_split_boundary<<<grid_for(n), kThreads, 0, ctx->stream>>>(staged, n, ctx->raw.base, ctx->raw.capacity, side,
                                                                affine);
    FL_HIP(hipGetLastError());
  }
  ctx->raw.trial_aliases_sorted = false;
  FL_TRY(bin_particles_mode(ctx, ctx->raw.view(), n, kBinCompatible));
  if (ctx->bin_box.volume() > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  FL_TRY(sort_particles(ctx, n, ctx->raw.view(), true));
  {
    Timer c(ctx);
    if (face_count) FL_HIP(hipMemcpyAsync(ctx->staging_aos, tags, tb, hipMemcpyHostToDevice, ctx->stream));
    copy_ms += c.stop();
  }
  if (face_count) {
    k_split_tags<<<grid_for(face_count), kThreads, 0, ctx->stream>>>(
        reinterpret_cast<const fl_face_tag*>(ctx->staging_aos), face_count, slot_side, slot_ok);
    FL_HIP(hipGetLastError());
  }
  FL_TRY(p2g_compatible(ctx, ctx->nvalid, slot_side, slot_ok, side, affine, 1.0 / quantization.mass_per_quantum,
                        1.0 / quantization.pure_momentum_per_quantum, limit));
  FL_TRY(first_error_code(ctx));
  const uint64_t nf = face_count;
  if (nf) {
    k_pack_quanta<<<grid_for(nf), kThreads, 0, ctx->stream>>>(ctx->qm, ctx->qp, nf, ctx->quanta_aos);
    FL_HIP(hipGetLastError());
  }
  const double kernel_ms = t.stop() - copy_ms;
  std::vector<fl_face_quanta> quanta_host;
  Timer c(ctx);
  // Nothing reaches the caller until every stage has succeeded: the quanta wait in host memory while
  // the stencils are formed chunk by chunk in the staging buffer.
  if (nf && stencils) {
    quanta_host.resize(nf);
    FL_HIP(hipMemcpyAsync(quanta_host.data(), ctx->quanta_aos, nf * sizeof(fl_face_quanta), hipMemcpyDeviceToHost,
                          ctx->stream));
  } else if (nf) {
    FL_HIP(hipMemcpyAsync(out, ctx->quanta_aos, nf * sizeof(fl_face_quanta), hipMemcpyDeviceToHost, ctx->stream));
  }
  if (stencils && n) {
    auto* d = reinterpret_cast<fl_compatible_sample*>(ctx->staging_aos);
    const uint64_t chunk = staging_bytes(ctx) / (81 * sizeof(fl_compatible_sample));
    for (uint64_t lo = 0; lo < n; lo += chunk) {
      const uint64_t cnt = std::min<uint64_t>(chunk, n - lo);
      FL_TRY(compatible_stencils(ctx, lo, cnt, slot_side, slot_ok, side, affine, d));
      FL_HIP(hipMemcpyAsync(stencils + 81 * lo, d, cnt * 81 * sizeof(fl_compatible_sample), hipMemcpyDeviceToHost,
                            ctx->stream));
    }
  }
  FL_HIP(hipStreamSynchronize(ctx->stream));
  if (nf && stencils) std::copy(quanta_host.begin(), quanta_host.end(), out);
  copy_ms += c.stop();
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 11);
  return FL_OK;
}

// ValidateHistory (liquid_boundary_transfer_3d.cc:97-145), in the reference's order.
static int validate_history(const fl_velocity_history* h, size_t face_count, double stage_time) {
  if (h->velocity_sample_time != stage_time || h->wall_sample_time != stage_time) return FL_INVALID_ARGUMENT;
  if (h->count != face_count) return FL_INVALID_ARGUMENT;
  if (face_count && (!h->face_velocities || !h->validity || !h->face_tags || !h->wall_samples))
    return FL_INVALID_ARGUMENT;
  for (size_t s = 0; s < face_count; ++s) {
    const uint8_t v = h->validity[s];
    const bool usable = v == 1 || v == 2;
    if (v != 0 && v != 3 && !usable) return FL_INVALID_ARGUMENT;
    if (usable && !std::isfinite(h->face_velocities[s])) return FL_INVALID_ARGUMENT;
    const fl_wall_sample& w = h->wall_samples[s];
    if (w.available) {
      const fl_vec3 n = w.unit_normal, u = w.wall_velocity;
      if (!std::isfinite(u.x) || !std::isfinite(u.y) || !std::isfinite(u.z) || !std::isfinite(n.x) ||
          !std::isfinite(n.y) || !std::isfinite(n.z))
        return FL_INVALID_ARGUMENT;
      if (std::fabs(((n.x * n.x + n.y * n.y) + n.z * n.z) - 1.0) > 1e-12) return FL_INVALID_ARGUMENT;
    }
  }
  return FL_OK;
}

int fl_g2p_boundary(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
                    const fl_boundary_particle* particles, size_t n, double old_time, double new_time,
                    const fl_velocity_history* oh, const fl_velocity_history* nh, fl_boundary_gather* out,
                    uint64_t* out_counts, fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  FL_TRY(check_geometry(geometry));
  FL_TRY(check_blocks(blocks, block_count));
  if (n > ctx->cap.particles || block_count > ctx->cap.blocks) return FL_RESOURCE_EXHAUSTED;
  if (!oh || !nh || (n && (!particles || !out))) return FL_INVALID_ARGUMENT;
  const uint64_t nf = 192 * block_count;
  // Stage times, then each history (liquid_boundary_transfer_3d.cc:528-545).
  if (!std::isfinite(old_time) || !std::isfinite(new_time) || new_time <= old_time) return FL_INVALID_ARGUMENT;
  FL_TRY(validate_history(oh, nf, old_time));
  FL_TRY(validate_history(nh, nf, new_time));
  const size_t ob = n * sizeof(fl_boundary_gather);
  if (overlaps(out, ob, particles, n * sizeof(fl_boundary_particle))) return FL_INVALID_ARGUMENT;
  for (const fl_velocity_history* h : {oh, nh})
    if (overlaps(out, ob, h->face_velocities, nf * sizeof(double)) || overlaps(out, ob, h->validity, nf) ||
        overlaps(out, ob, h->face_tags, nf * sizeof(fl_face_tag)) ||
        overlaps(out, ob, h->wall_samples, nf * sizeof(fl_wall_sample)))
      return FL_INVALID_ARGUMENT;
  const bool same = oh->face_velocities == nh->face_velocities && oh->validity == nh->validity &&
                    oh->face_tags == nh->face_tags && oh->wall_samples == nh->wall_samples;
  (void)hipSetDevice(ctx->device);
  set_geometry(ctx, geometry);
  // History and particle-attribute buffers: allocated on the first boundary gather, grown with it.
  if (!ctx->boundary) ctx->boundary = new (std::nothrow) BoundaryBuffers();
  auto* B = static_cast<BoundaryBuffers*>(ctx->boundary);
  if (!B) return FL_RESOURCE_EXHAUSTED;
  const uint64_t need_f = std::max<uint64_t>(nf, 1), need_p = std::max<uint64_t>(ctx->cap.particles, 1);
  if (B->faces < need_f || B->particles < need_p) {
    B->release();
    for (int k = 0; k < 2; ++k) {
      FL_TRY(dalloc(&B->u[k], need_f));
      FL_TRY(dalloc(&B->validity[k], need_f));
      FL_TRY(dalloc(&B->side[k], need_f));
      FL_TRY(dalloc(&B->eligible[k], need_f));
      FL_TRY(dalloc(&B->walls[k], need_f));
    }
    FL_TRY(dalloc(&B->pside, need_p));
    FL_TRY(dalloc(&B->paffine, need_p));
    FL_TRY(dalloc(&B->slow, need_p));
    FL_TRY(dalloc(&B->slow_count, 1));
    B->faces = need_f;
    B->particles = need_p;
  }
  double copy_ms = 0;
  FL_TRY(build_lookup(ctx, reinterpret_cast<const int32_t*>(blocks), block_count));
  {
    Timer c(ctx);
    const fl_velocity_history* hs[2] = {oh, nh};
    for (int k = 0; k < (same ? 1 : 2) && nf; ++k) {
      FL_HIP(hipMemcpyAsync(B->u[k], hs[k]->face_velocities, nf * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      FL_HIP(hipMemcpyAsync(B->validity[k], hs[k]->validity, nf, hipMemcpyHostToDevice, ctx->stream));
      FL_HIP(hipMemcpyAsync(B->walls[k], hs[k]->wall_samples, nf * sizeof(fl_wall_sample), hipMemcpyHostToDevice,
                            ctx->stream));
      // Tags are staged in the face-quanta buffer (24 B per slot), then split into side and eligibility.
      FL_HIP(hipMemcpyAsync(ctx->quanta_aos, hs[k]->face_tags, nf * sizeof(fl_face_tag), hipMemcpyHostToDevice,
                            ctx->stream));
      k_split_history<<<grid_for(nf), kThreads, 0, ctx->stream>>>(reinterpret_cast<const fl_face_tag*>(ctx->quanta_aos),
                                                                  nf, B->side[k], B->eligible[k]);
      FL_HIP(hipGetLastError());
    }
    copy_ms += c.stop();
  }
  Timer t(ctx);
  auto* staged = reinterpret_cast<fl_boundary_particle*>(ctx->staging_aos);
  {
    Timer c(ctx);
    if (n) FL_HIP(hipMemcpyAsync(staged, particles, n * sizeof(fl_boundary_particle), hipMemcpyHostToDevice, ctx->stream));
    copy_ms += c.stop();
  }
  if (n) {
    k_split_boundary<<<grid_for(n), kThreads, 0, ctx->stream>>>(staged, n, ctx->raw.base, ctx->raw.capacity, B->pside,
                                                                B->paffine);
    FL_HIP(hipGetLastError());
  }
  ctx->raw.trial_aliases_sorted = false;
  FL_TRY(bin_particles_mode(ctx, ctx->raw.view(), n, kBinBoundaryGather));
  if (ctx->bin_box.volume() > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  FL_TRY(sort_particles(ctx, n, ctx->raw.view(), true));
  BoundaryDevice d;
  for (int k = 0; k < 2; ++k) {
    const int src = same ? 0 : k;
    d.u[k] = B->u[src];
    d.validity[k] = B->validity[src];
    d.side[k] = B->side[src];
    d.eligible[k] = B->eligible[src];
    d.walls[k] = B->walls[src];
  }
  d.same = same ? 1 : 0;
  d.pside = B->pside;
  d.paffine = B->paffine;
  d.extra = ctx->raw.base;  // the caller-order columns are free once the particles are sorted
  d.extra_cap = ctx->raw.capacity;
  d.slow = B->slow;
  d.slow_count = B->slow_count;
  uint32_t slow = 0;
  FL_TRY(g2p_boundary(ctx, ctx->nvalid, d, &slow));
  FL_TRY(first_error_code(ctx));
  auto* packed = reinterpret_cast<fl_boundary_gather*>(ctx->staging_aos);
  if (ctx->nvalid) {
    k_pack_boundary_gather<<<grid_for(ctx->nvalid), kThreads, 0, ctx->stream>>>(
        ctx->gather, ctx->sorted.capacity, ctx->raw.base, ctx->raw.capacity, ctx->perm, ctx->nvalid, packed);
    FL_HIP(hipGetLastError());
  }
  const double kernel_ms = t.stop() - copy_ms;
  Timer c(ctx);
  if (n) FL_HIP(hipMemcpyAsync(out, packed, ob, hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  copy_ms += c.stop();
  if (out_counts) {
    out_counts[0] = out_counts[1] = out_counts[2] = 0;
    for (size_t i = 0; i < n; ++i) {
      out_counts[0] += out[i].old_pic_fallback;
      out_counts[1] += out[i].new_pic_fallback;
      out_counts[2] += out[i].low_mass_fallback;
    }
  }
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 12);
  (void)slow;
  return FL_OK;
}

__global__ void k_pack_deposit(const double* m, const double* q, uint64_t n, fl_face_deposit* out) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = fl_face_deposit{m[i], q[i]};
}

int fl_p2g(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
           const fl_particle* particles, size_t n, fl_face_deposit* out, size_t face_count,
           fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  FL_TRY(check_geometry(geometry));
  FL_TRY(check_blocks(blocks, block_count));
  if (face_count != 192 * block_count) return FL_INVALID_ARGUMENT;
  if (overlaps(out, face_count * sizeof(fl_face_deposit), particles, n * sizeof(fl_particle)))
    return FL_INVALID_ARGUMENT;
  if (n > ctx->cap.particles || block_count > ctx->cap.blocks) return FL_RESOURCE_EXHAUSTED;
  (void)hipSetDevice(ctx->device);
  set_geometry(ctx, geometry);
  double copy_ms = 0;
  FL_TRY(build_lookup(ctx, reinterpret_cast<const int32_t*>(blocks), block_count));
  Timer t(ctx);
  FL_TRY(stage_particles(ctx, particles, n, true, &copy_ms));
  FL_TRY(p2g_quanta(ctx, ctx->nvalid, 1.0, 1.0, std::numeric_limits<double>::max(), nullptr, nullptr, true));
  FL_TRY(first_error_code(ctx));
  const uint64_t nf = face_count;
  fl_face_deposit* staged = reinterpret_cast<fl_face_deposit*>(ctx->quanta_aos);
  if (nf) {
    k_pack_deposit<<<grid_for(nf), kThreads, 0, ctx->stream>>>(reinterpret_cast<const double*>(ctx->qm),
                                                               reinterpret_cast<const double*>(ctx->qp), nf, staged);
    FL_HIP(hipGetLastError());
  }
  const double kernel_ms = t.stop() - copy_ms;
  Timer c(ctx);
  if (nf) FL_HIP(hipMemcpyAsync(out, staged, nf * sizeof(fl_face_deposit), hipMemcpyDeviceToHost, ctx->stream));
  copy_ms += c.stop();
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 10);
  return FL_OK;
}

int fl_assemble_full_cell_strain(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
                                 const fl_viscosity_cell* cells, size_t cell_count, fl_strain_row* out,
                                 size_t* io_count, fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (!io_count || (cell_count && (!cells || !out)) || (block_count && !blocks)) return FL_INVALID_ARGUMENT;
  FL_TRY(check_geometry(geometry));
  FL_TRY(check_blocks(blocks, block_count));
  if (cell_count > ctx->cap.pressure_cells || !scan_fits(ctx, cell_count)) return FL_RESOURCE_EXHAUSTED;
  (void)hipSetDevice(ctx->device);
  set_geometry(ctx, geometry);
  const uint64_t n = cell_count, capacity = 6 * n;
  FL_TRY(grow_viscosity(ctx, n, capacity, 1));
  double copy_ms = 0;
  {
    Timer c(ctx);
    FL_TRY(build_lookup(ctx, reinterpret_cast<const int32_t*>(blocks), block_count));
    auto& B = *static_cast<ViscosityBuffers*>(ctx->viscosity);
    if (n) FL_HIP(hipMemcpyAsync(B.cells, cells, n * sizeof(fl_viscosity_cell), hipMemcpyHostToDevice, ctx->stream));
    copy_ms += c.stop();
  }
  Timer t(ctx);
  uint64_t total = 0;
  FL_TRY(fl::assemble_strain(ctx, viscosity_device(ctx), n, geometry->cell_width, &total));
  const double kernel_ms = t.stop();
  if (total > *io_count) return FL_RESOURCE_EXHAUSTED;
  Timer c(ctx);
  auto& B = *static_cast<ViscosityBuffers*>(ctx->viscosity);
  if (total) FL_HIP(hipMemcpyAsync(out, B.rows, total * sizeof(fl_strain_row), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  copy_ms += c.stop();
  *io_count = total;
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 5);
  return FL_OK;
}

int fl_solve_viscous_block(fl_ctx* ctx, const fl_strain_row* rows, size_t row_count, const double* inertia,
                           size_t face_count, double dt, const double* rhs, uint64_t max_iterations,
                           double relative_tolerance, double* out_velocity, uint64_t* out_iterations,
                           double* out_max_residual, fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (!out_iterations || !out_max_residual || (row_count && !rows) ||
      (face_count && (!inertia || !rhs || !out_velocity)))
    return FL_INVALID_ARGUMENT;
  // ValidateVelocityBlockInputs (viscosity.cc:1331-1358), in its order.
  if (!std::isfinite(dt) || dt < 0.0) return FL_INVALID_ARGUMENT;
  if (max_iterations == 0 || !std::isfinite(relative_tolerance) || relative_tolerance <= 0.0 ||
      relative_tolerance >= 1.0)
    return FL_INVALID_ARGUMENT;
  double max_rhs = 0.0;
  for (size_t f = 0; f < face_count; ++f) {
    if (!std::isfinite(inertia[f]) || inertia[f] < 0.0 || !std::isfinite(rhs[f])) return FL_INVALID_ARGUMENT;
    if (inertia[f] == 0.0 && rhs[f] != 0.0) return FL_INVALID_ARGUMENT;
    max_rhs = std::max(max_rhs, std::fabs(rhs[f]));
  }
  // Packed indices are 32-bit: face slots and row ordinals below 2^32.
  if (face_count > 192 * std::max<uint64_t>(ctx->cap.blocks, 1) ||
      row_count > 6 * std::max<uint64_t>(ctx->cap.pressure_cells, 1) || !scan_fits(ctx, face_count) ||
      !scan_fits(ctx, row_count) || face_count >= (uint64_t(1) << 32) || row_count >= (uint64_t(1) << 32))
    return FL_RESOURCE_EXHAUSTED;
  (void)hipSetDevice(ctx->device);
  FL_TRY(grow_viscosity(ctx, 1, std::max<uint64_t>(row_count, 1), std::max<uint64_t>(face_count, 1)));
  auto& B = *static_cast<ViscosityBuffers*>(ctx->viscosity);
  double copy_ms = 0;
  {
    Timer c(ctx);
    if (row_count)
      FL_HIP(hipMemcpyAsync(B.rows, rows, row_count * sizeof(fl_strain_row), hipMemcpyHostToDevice, ctx->stream));
    if (face_count) {
      FL_HIP(hipMemcpyAsync(B.inertia, inertia, face_count * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      FL_HIP(hipMemcpyAsync(B.rhs, rhs, face_count * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    copy_ms += c.stop();
  }
  Timer t(ctx);
  uint64_t iterations = 0, terms = 0;
  double max_residual = 0.0;
  FL_TRY(fl::weight_viscous_rows(ctx, viscosity_device(ctx), row_count, face_count, dt, &terms));
  FL_TRY(grow_viscosity_terms(ctx, terms));
  FL_TRY(fl::solve_viscous_block(ctx, viscosity_device(ctx), row_count, face_count, terms, dt, max_rhs, max_iterations,
                                 relative_tolerance, &iterations, &max_residual));
  const double kernel_ms = t.stop();
  Timer c(ctx);
  if (face_count)
    FL_HIP(hipMemcpyAsync(out_velocity, B.v, face_count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  copy_ms += c.stop();
  *out_iterations = iterations;
  *out_max_residual = max_residual;
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, static_cast<uint32_t>(8 + 12 * iterations));
  return FL_OK;
}

int fl_project_viscous(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
                       double dt, const fl_pressure_cell* cells, size_t cell_count, const fl_pressure_face* faces,
                       size_t face_count, const fl_strain_row* rows, size_t row_count,
                       const fl_pressure_options* options, double* out_pressures, double* out_velocities,
                       double* out_defects, fl_viscous_diagnostics* out_diagnostics, fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (!options || (row_count && !rows) || !cells || !faces) return FL_INVALID_ARGUMENT;
  if (options->max_iterations == 0 || options->max_iterations > 256) return FL_INVALID_ARGUMENT;
  if (options->preconditioner != 0 && options->preconditioner != 2) return FL_INVALID_ARGUMENT;
  FL_TRY(check_geometry(geometry));
  FL_TRY(check_blocks(blocks, block_count));
  if (face_count >= (uint64_t(1) << 32) || row_count >= (uint64_t(1) << 32) ||
      row_count > 6 * std::max<uint64_t>(ctx->cap.pressure_cells, 1) || !scan_fits(ctx, face_count) ||
      !scan_fits(ctx, row_count))
    return FL_RESOURCE_EXHAUSTED;
  (void)hipSetDevice(ctx->device);
  const bool jacobi = options->preconditioner == 2;
  fl_pressure_options popt = *options;
  popt.preconditioner = 0;
  Timer t(ctx);
  fl::CoupledSystem S{};
  FL_TRY(fl::coupled_prepare(ctx, geometry, blocks, block_count, dt, cells, cell_count, faces, face_count, &popt, &S));
  FL_TRY(grow_viscosity(ctx, 1, std::max<uint64_t>(row_count, 1), std::max<uint64_t>(face_count, 1)));
  auto& B = *static_cast<ViscosityBuffers*>(ctx->viscosity);
  if (row_count)
    FL_HIP(hipMemcpyAsync(B.rows, rows, row_count * sizeof(fl_strain_row), hipMemcpyHostToDevice, ctx->stream));
  FL_TRY(fl::coupled_rows(ctx, S, viscosity_device(ctx), row_count, dt));
  uint64_t terms = 0;
  FL_TRY(fl::weight_viscous_rows(ctx, viscosity_device(ctx), row_count, face_count, dt, &terms));
  FL_TRY(grow_viscosity_terms(ctx, terms));
  fl::CoupledResult r{};
  const int st = fl::coupled_solve(ctx, S, viscosity_device(ctx), row_count, terms, dt, popt, jacobi, &r);
  const double kernel_ms = t.stop();
  if (out_diagnostics) {
    fl_viscous_diagnostics d{};
    d.pressure_iterations = r.iterations;
    d.constraint_residual_l2 = r.residual_l2;
    d.initial_constraint_residual_l2 = r.initial_residual_l2;
    d.max_cell_constraint_defect = r.max_defect;
    d.max_momentum_residual = r.max_momentum_residual;
    d.velocity_solves = r.velocity_solves;
    d.velocity_iterations = r.velocity_iterations;
    // AccumulateInwardExplicitWallFlux (mac_pressure_3d_arithmetic.h:97-108) over the input faces.
    double sum = 0.0, corr = 0.0;
    for (size_t f = 0; f < face_count; ++f) {
      if (!faces[f].has_prescribed_wall_flux) continue;
      const int a = faces[f].axis;
      fl_int3 tail = faces[f].lattice;
      int32_t& along = a == 0 ? tail.x : a == 1 ? tail.y : tail.z;
      const bool tail_exists = along > std::numeric_limits<int32_t>::min();
      if (tail_exists) along -= 1;
      auto find = [&](const fl_int3& k) {
        size_t lo = 0, hi = cell_count;
        while (lo < hi) {
          const size_t mid = (lo + hi) / 2;
          const fl_int3& c = cells[mid].cell_index;
          const bool less = c.z != k.z ? c.z < k.z : c.y != k.y ? c.y < k.y : c.x < k.x;
          if (less) lo = mid + 1;
          else hi = mid;
        }
        return lo < cell_count && cells[lo].cell_index.x == k.x && cells[lo].cell_index.y == k.y &&
               cells[lo].cell_index.z == k.z;
      };
      const bool has_tail = tail_exists && find(tail);
      const bool has_head = find(faces[f].lattice);
      if (has_tail == has_head) continue;
      const double term = (has_tail ? -1.0 : 1.0) * faces[f].prescribed_wall_flux;
      const double n = sum + term;
      corr += std::fabs(sum) >= std::fabs(term) ? (sum - n) + term : (term - n) + sum;
      sum = n;
    }
    d.inward_prescribed_wall_flux = sum + corr;
    *out_diagnostics = d;
  }
  if (st != FL_OK) return st;
  const double *p = nullptr, *u = nullptr, *def = nullptr;
  FL_TRY(fl::coupled_outputs(ctx, &p, &u, &def));
  Timer c(ctx);
  hipStream_t s = ctx->stream;
  if (out_pressures) FL_HIP(hipMemcpyAsync(out_pressures, p, cell_count * sizeof(double), hipMemcpyDeviceToHost, s));
  if (out_velocities) FL_HIP(hipMemcpyAsync(out_velocities, u, face_count * sizeof(double), hipMemcpyDeviceToHost, s));
  if (out_defects) FL_HIP(hipMemcpyAsync(out_defects, def, cell_count * sizeof(double), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  const double copy_ms = c.stop();
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, static_cast<uint32_t>(r.velocity_iterations));
  return FL_OK;
}

int fl_normalize_forces(fl_ctx* ctx, const fl_face_quanta* quanta, size_t face_count, fl_quantization quantization,
                        double minimum_face_mass, fl_vec3 acceleration, double dt, double* out_mass,
                        double* out_velocity, uint8_t* out_validity, double* out_predicted,
                        fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (face_count > 192 * ctx->cap.blocks) return FL_RESOURCE_EXHAUSTED;
  if (!powers_of_two_ok(quantization)) return FL_INVALID_ARGUMENT;
  if (!std::isfinite(minimum_face_mass) || minimum_face_mass < 0.0) return FL_INVALID_ARGUMENT;
  if (face_count % 192 != 0) return FL_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  double copy_ms = 0;
  Timer t0(ctx);
  if (face_count)
    FL_HIP(hipMemcpyAsync(ctx->quanta_aos, quanta, face_count * sizeof(fl_face_quanta), hipMemcpyHostToDevice,
                          ctx->stream));
  copy_ms += t0.stop();
  Timer t(ctx);
  FL_TRY(reset_error(ctx));
  int64_t* qs = reinterpret_cast<int64_t*>(ctx->staging_aos);
  if (face_count) {
    k_split_quanta<<<grid_for(face_count), kThreads, 0, ctx->stream>>>(ctx->quanta_aos, face_count, ctx->qm, ctx->qp,
                                                                       qs);
    FL_HIP(hipGetLastError());
  }
  const double g[3] = {acceleration.x, acceleration.y, acceleration.z};
  FL_TRY(normalize_forces_stress(ctx, qs, face_count, quantization.mass_per_quantum,
                                 quantization.pure_momentum_per_quantum, minimum_face_mass, g, dt));
  FL_HIP(hipMemcpyAsync(ctx->host_words, ctx->err, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  const double kernel_ms = t.stop();
  if (ctx->host_words[0] != ~uint64_t(0)) return static_cast<int>(ctx->host_words[0] & 0xff);
  if (ctx->host_words[1] != ~uint64_t(0)) return static_cast<int>(ctx->host_words[1] & 0xff);
  Timer c(ctx);
  const size_t bytes = face_count * sizeof(double);
  if (face_count) {
    if (out_mass) FL_HIP(hipMemcpyAsync(out_mass, ctx->face_mass, bytes, hipMemcpyDeviceToHost, ctx->stream));
    if (out_velocity) FL_HIP(hipMemcpyAsync(out_velocity, ctx->u0, bytes, hipMemcpyDeviceToHost, ctx->stream));
    if (out_validity) FL_HIP(hipMemcpyAsync(out_validity, ctx->validity, face_count, hipMemcpyDeviceToHost, ctx->stream));
    if (out_predicted) FL_HIP(hipMemcpyAsync(out_predicted, ctx->ustar, bytes, hipMemcpyDeviceToHost, ctx->stream));
  }
  copy_ms += c.stop();
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 3);
  return FL_OK;
}

int fl_apply_forces(fl_ctx* ctx, const double* old_face_velocities, size_t face_count, fl_vec3 acceleration,
                    double dt, double* out_predicted, fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (face_count % 192 != 0) return FL_INVALID_ARGUMENT;
  if (face_count > 192 * ctx->cap.blocks) return FL_RESOURCE_EXHAUSTED;
  if (face_count && (!old_face_velocities || !out_predicted)) return FL_INVALID_ARGUMENT;
  if (overlaps(out_predicted, face_count * sizeof(double), old_face_velocities, face_count * sizeof(double)))
    return FL_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  double copy_ms = 0;
  Timer t0(ctx);
  if (face_count)
    FL_HIP(hipMemcpyAsync(ctx->u0, old_face_velocities, face_count * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
  copy_ms += t0.stop();
  Timer t(ctx);
  FL_TRY(reset_error(ctx));
  const double g[3] = {acceleration.x, acceleration.y, acceleration.z};
  FL_TRY(apply_forces(ctx, ctx->u0, face_count, g, dt, ctx->ustar));
  FL_HIP(hipMemcpyAsync(ctx->host_words, ctx->err, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  FL_HIP(hipStreamSynchronize(ctx->stream));
  const double kernel_ms = t.stop();
  if (ctx->host_words[0] != ~uint64_t(0)) return static_cast<int>(ctx->host_words[0] & 0xff);
  Timer c(ctx);
  if (face_count)
    FL_HIP(hipMemcpyAsync(out_predicted, ctx->ustar, face_count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  copy_ms += c.stop();
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 1);
  return FL_OK;
}

int fl_g2p(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
           const fl_particle* particles, size_t n, const double* old_face_velocities,
           const double* new_face_velocities, size_t face_count, fl_gather_result* out, fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  FL_TRY(check_geometry(geometry));
  FL_TRY(check_blocks(blocks, block_count));
  if (face_count != 192 * block_count) return FL_INVALID_ARGUMENT;
  if (n > ctx->cap.particles || block_count > ctx->cap.blocks) return FL_RESOURCE_EXHAUSTED;
  const size_t out_bytes = n * sizeof(fl_gather_result);
  if (overlaps(out, out_bytes, particles, n * sizeof(fl_particle)) ||
      overlaps(out, out_bytes, old_face_velocities, face_count * sizeof(double)) ||
      overlaps(out, out_bytes, new_face_velocities, face_count * sizeof(double)))
    return FL_INVALID_ARGUMENT;
  for (size_t f = 0; f < face_count; ++f)
    if (!std::isfinite(old_face_velocities[f]) || !std::isfinite(new_face_velocities[f])) return FL_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  set_geometry(ctx, geometry);
  double copy_ms = 0;
  FL_TRY(build_lookup(ctx, reinterpret_cast<const int32_t*>(blocks), block_count));
  {
    Timer c(ctx);
    if (face_count) {
      FL_HIP(hipMemcpyAsync(ctx->u0, old_face_velocities, face_count * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      FL_HIP(hipMemcpyAsync(ctx->unew, new_face_velocities, face_count * sizeof(double), hipMemcpyHostToDevice,
                            ctx->stream));
    }
    copy_ms += c.stop();
  }
  Timer t(ctx);
  FL_TRY(stage_particles(ctx, particles, n, false, &copy_ms));
  FL_TRY(g2p(ctx, ctx->nvalid, ctx->u0, ctx->unew));
  FL_TRY(first_error_code(ctx));
  if (ctx->nvalid) {
    k_pack_gather<<<grid_for(ctx->nvalid), kThreads, 0, ctx->stream>>>(ctx->gather, ctx->sorted.capacity, ctx->perm,
                                                                       ctx->nvalid, ctx->gather_aos);
    FL_HIP(hipGetLastError());
  }
  const double kernel_ms = t.stop() - copy_ms;
  Timer c(ctx);
  if (n) FL_HIP(hipMemcpyAsync(out, ctx->gather_aos, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
  copy_ms += c.stop();
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 10);
  return FL_OK;
}

int fl_project(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count, double dt,
               const fl_pressure_cell* cells, size_t cell_count, const fl_pressure_face* faces, size_t face_count,
               const fl_pressure_options* options, double* out_pressures, double* out_velocities, double* out_rhs,
               double* out_residuals, double* out_defects, uint64_t* out_gauges,
               fl_pressure_diagnostics* out_diagnostics, fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  FL_TRY(check_geometry(geometry));
  FL_TRY(check_blocks(blocks, block_count));
  (void)hipSetDevice(ctx->device);
  set_geometry(ctx, geometry);
  return fl::project_pressure(ctx, geometry, blocks, block_count, dt, cells, cell_count, faces, face_count, options,
                              out_pressures, out_velocities, out_rhs, out_residuals, out_defects, out_gauges,
                              out_diagnostics, measurement);
}

int fl_project_device_built(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
                            double dt, const fl_pressure_cell* cells, size_t cell_count, const fl_pressure_face* faces,
                            size_t face_count, const fl_pressure_options* options, double* out_pressures,
                            double* out_velocities, fl_pressure_diagnostics* out_diagnostics) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  FL_TRY(check_geometry(geometry));
  FL_TRY(check_blocks(blocks, block_count));
  if (!options || (cell_count && !cells) || (face_count && !faces)) return FL_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  set_geometry(ctx, geometry);
  FL_TRY(fl::build_lookup(ctx, reinterpret_cast<const int32_t*>(blocks), block_count));
  fl_pressure_diagnostics diag{};
  const int st = fl::project_records_device(ctx, cells, cell_count, faces, face_count, dt, *options, out_pressures,
                                            out_velocities, &diag);
  if (out_diagnostics) *out_diagnostics = diag;
  return st;
}

int fl_sample_particle_field(fl_ctx* ctx, const fl_geometry* geometry, const fl_vec3* positions, size_t n,
                             double radius, uint64_t candidate_capacity, fl_int3* out_cells, uint64_t* cell_count,
                             fl_center_distance* out_centres, uint64_t centre_capacity, uint64_t* centre_count,
                             fl_center_distance* out_corners, uint64_t corner_capacity, uint64_t* corner_count,
                             fl_measurement* measurement) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  FL_TRY(check_geometry(geometry));
  if (!cell_count || !centre_count || !corner_count || (n && !positions)) return FL_INVALID_ARGUMENT;
  if (!std::isfinite(radius) || radius <= 0) return FL_INVALID_ARGUMENT;
  if (n > ctx->cap.particles) return FL_RESOURCE_EXHAUSTED;
  (void)hipSetDevice(ctx->device);
  set_geometry(ctx, geometry);
  double copy_ms = 0;
  {
    Timer t(ctx);
    if (n) FL_HIP(hipMemcpyAsync(ctx->staging_aos, positions, n * sizeof(fl_vec3), hipMemcpyHostToDevice, ctx->stream));
    copy_ms += t.stop();
  }
  Timer t(ctx);
  if (n) {
    k_split_columns<<<grid_for(n), kThreads, 0, ctx->stream>>>(reinterpret_cast<const double*>(ctx->staging_aos), n, 3,
                                                              3, ctx->raw.base, ctx->raw.capacity);
    FL_HIP(hipGetLastError());
  }
  ctx->raw.trial_aliases_sorted = true;
  // SampleWorldBatch rejects non-finite particles as InvalidArgument before any sample.
  FL_TRY(bin_particles(ctx, ctx->raw.view(), n, false, false));
  uint64_t word = 0;
  FL_TRY(fl::read_error(ctx, &word));
  if (word != ~uint64_t(0)) {
    const int code = static_cast<int>(word & 0xff);
    return code == FL_OUT_OF_RANGE ? FL_OUT_OF_RANGE : FL_INVALID_ARGUMENT;
  }
  if (ctx->bin_box.volume() > ctx->cap.lookup_blocks) return FL_RESOURCE_EXHAUSTED;
  FL_TRY(sort_particles(ctx, n, ctx->raw.view(), true, 0x7u));
  fl::FieldSamples F{};
  FL_TRY(fl::sample_particle_field(ctx, radius, candidate_capacity, &F));
  const double kernel_ms = t.stop();
  *cell_count = F.cell_count;
  *centre_count = F.centre_count;
  *corner_count = F.corner_count;
  if (!out_cells || !out_centres || !out_corners) {
    fill_measurement(ctx, measurement, kernel_ms, copy_ms, 12);
    return FL_OK;
  }
  if (F.cell_count > candidate_capacity || F.centre_count > centre_capacity || F.corner_count > corner_capacity)
    return FL_RESOURCE_EXHAUSTED;
  Timer c(ctx);
  hipStream_t s = ctx->stream;
  if (F.cell_count) FL_HIP(hipMemcpyAsync(out_cells, F.cells, F.cell_count * sizeof(fl_int3), hipMemcpyDeviceToHost, s));
  if (F.centre_count)
    FL_HIP(hipMemcpyAsync(out_centres, F.centres, F.centre_count * sizeof(fl_center_distance), hipMemcpyDeviceToHost, s));
  if (F.corner_count)
    FL_HIP(hipMemcpyAsync(out_corners, F.corners, F.corner_count * sizeof(fl_center_distance), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  copy_ms += c.stop();
  fill_measurement(ctx, measurement, kernel_ms, copy_ms, 12);
  return FL_OK;
}

int fl_set_transfer_precision(fl_ctx* ctx, int bits) {
  if (!ctx || ctx->poisoned) return FL_FAILED_PRECONDITION;
  if (bits != 32 && bits != 64) return FL_INVALID_ARGUMENT;
  ctx->transfer_bits = bits;
  return FL_OK;
}

int fl_substep(fl_ctx* ctx, const fl_substep_options* options, fl_substep_report* report) {
  if (!ctx || ctx->poisoned || !ctx->scene_loaded) return FL_FAILED_PRECONDITION;
  if (!options || !report) return FL_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  fl_substep_report r{};
  const int st = fl::substep(ctx, *options, r);
  r.status = st;
  *report = r;
  return st;
}

int fl_substep_download(fl_ctx* ctx, double* positions_xyz, double* velocities_xyz, double* slopes_rowmajor,
                        double* volume_ratios, uint8_t* flags, size_t n) {
  if (!ctx || ctx->poisoned || !ctx->scene_loaded) return FL_FAILED_PRECONDITION;
  (void)hipSetDevice(ctx->device);
  return fl::substep_download(ctx, positions_xyz, velocities_xyz, slopes_rowmajor, volume_ratios, flags, n);
}

}  // extern "C"

// ------------------------------------------------------------------ measurement utility
namespace {
__global__ void k_copy_probe(const float4* src, float4* dst, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

extern "C" int fl_probe_copy(int device, size_t bytes, int reps, double* gbs) {
  if (bytes < 16 || reps < 1 || !gbs) return FL_INVALID_ARGUMENT;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return FL_FAILED_PRECONDITION;
  FL_HIP(hipSetDevice(device));
  const uint64_t n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess) return FL_RESOURCE_EXHAUSTED;
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipFree(a);
    return FL_RESOURCE_EXHAUSTED;
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipMemset(a, 0, n * 16);
  const unsigned grid = 256 * 32;  // 32 workgroups per CU, grid-stride
  k_copy_probe<<<grid, 256>>>(a, b, n);
  (void)hipEventRecord(e0, nullptr);
  for (int r = 0; r < reps; ++r) k_copy_probe<<<grid, 256>>>(a, b, n);
  (void)hipEventRecord(e1, nullptr);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  *gbs = 2.0 * double(n) * 16.0 * reps / (double(ms) * 1e-3) / 1e9;
  return FL_OK;
}
