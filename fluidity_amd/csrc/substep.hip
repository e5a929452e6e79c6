// One device-resident liquid substep over moving particles, in the stage order of AttemptImpl
// (native/src/liquid_substep_3d.cc:1175-2820): sort the particles into their bins (the trial topology),
// sample the particle-union liquid field (PrepareField), activate the blocks of the particles' stencils
// and of the candidates' faces, deposit and normalize the face quanta and apply gravity, freeze the
// boundary geometry, build the typed pressure links with the predicted face velocities, project on a
// system and hierarchy built on the device, gather the particles' grid velocities, and update the
// particles (ASFLIP, advection, contact, volume ratio). Every stage runs on the context's stream on
// device arrays; the host reads only counts and error words between stages. At the end the sorted
// particle columns become the scene's particles, so the next substep sorts the moved particles.
#include <cmath>
#include <cstring>

#include "context.h"

namespace fl {

struct SubstepBuffers {
  // frozen geometry staging (FreezeDevice)
  uint64_t *keep = nullptr, *slots = nullptr, *word_counts = nullptr, *face_slots = nullptr, *counts = nullptr,
           *work = nullptr;
  unsigned long long* words = nullptr;
  fl_frozen_cell *staged = nullptr, *cells = nullptr;
  fl_frozen_face* faces = nullptr;
  uint8_t* needs = nullptr;
  double *phi = nullptr, *uncut = nullptr;
  // links
  fl_cell_fields* fields = nullptr;
  double* predicted = nullptr;
  fl_pressure_cell* pcells = nullptr;
  fl_pressure_face* pfaces = nullptr;
  uint64_t* lower = nullptr;
  // particle state beyond the scene's columns, in the scene's order and in sort order
  double *ratio = nullptr, *ratio_sorted = nullptr;
  uint8_t *flags = nullptr, *flags_sorted = nullptr;
  uint32_t *contact = nullptr, *contact_count = nullptr;  // particles near a wall for the contact pass
  // sampled solid: the solid's corner samples at the field's corner keys, and the freeze's per-cell copy
  fl_solid_corner* solids = nullptr;
  double* solid_phi = nullptr;
  uint64_t cap[26] = {};
  hipEvent_t ev[12] = {};
  bool fresh = true;  // the scene was uploaded since the last substep: volume ratios 1, PIC unset
  void release() {
    void* ptrs[] = {keep,   slots,  word_counts, face_slots, counts, work,      words,        staged,
                    cells,  faces,  needs,       phi,        uncut,  fields,    predicted,    pcells,
                    pfaces, lower,  ratio,       ratio_sorted, flags, flags_sorted, contact, contact_count,
                    solids, solid_phi};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    *this = SubstepBuffers{};
  }
};

namespace {

template <typename T>
int grow_to(T** p, uint64_t* cap, uint64_t need) {
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

__global__ void k_state_init(double* ratio, uint8_t* flags, uint64_t n) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  ratio[i] = 1.0;
  flags[i] = 0;
}

__global__ void k_state_permute(const double* ratio, const uint8_t* flags, const uint32_t* perm, uint64_t n,
                                double* ratio_out, uint8_t* flags_out) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  ratio_out[i] = ratio[perm[i]];
  flags_out[i] = flags[perm[i]];
}

// LiquidPressureCellFields3D of the frozen cells: an incompressible liquid of one density at rest
// (zero target divergence, infinite bulk modulus, previous pressure zero).
__global__ void k_cell_fields(fl_cell_fields* f, uint64_t n, double density) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  fl_cell_fields c{};
  c.target_divergence = 0.0;
  c.bulk_modulus = __builtin_inf();
  c.previous_pressure = 0.0;
  c.density = density;
  c.has_density = 1;
  f[i] = c;
}

// Predicted normal velocity of every frozen face: u* at its owned slot.
__global__ void k_predicted(const fl_frozen_face* faces, uint64_t n, const double* ustar, double* out) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = ustar[faces[i].face_slot];
}

// The scene's level-set solids at the field's corner keys: signed distance (positive outside) of the
// union of the selected parts and the wall velocity of the nearest part. cos_a, sin_a: the paddle's
// turn about +y (body = R^T (x - pivot)).
struct SolidParams {
  uint32_t parts;
  double sx, sy, sz, sr;
  double px, py, pz, ex, ey, ez, cos_a, sin_a, omega;
  double wlx, wlz, whx, whz;
};

__global__ void k_sample_solid(const fl_center_distance* corners, uint64_t n, double ox, double oy, double oz, double h,
                               SolidParams sp, fl_solid_corner* out) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const fl_int3 k = corners[i].cell_index;
  const double x = ox + k.x * h, y = oy + k.y * h, z = oz + k.z * h;
  double best = __builtin_inf();
  double vx = 0.0, vz = 0.0;
  if (sp.parts & 1u) {
    const double dx = x - sp.sx, dy = y - sp.sy, dz = z - sp.sz;
    best = sqrt(dx * dx + dy * dy + dz * dz) - sp.sr;
  }
  if (sp.parts & 2u) {
    const double rx = x - sp.px, ry = y - sp.py, rz = z - sp.pz;
    const double bx = sp.cos_a * rx - sp.sin_a * rz, bz = sp.sin_a * rx + sp.cos_a * rz;
    const double qx = fabs(bx) - sp.ex, qy = fabs(ry) - sp.ey, qz = fabs(bz) - sp.ez;
    const double mx = fmax(qx, 0.0), my = fmax(qy, 0.0), mz = fmax(qz, 0.0);
    const double d = sqrt(mx * mx + my * my + mz * mz) + fmin(fmax(qx, fmax(qy, qz)), 0.0);
    if (d < best) {
      best = d;
      vx = sp.omega * rz;
      vz = -sp.omega * rx;
    }
  }
  if (sp.parts & 4u) {
    const double d = fmin(fmin(x - sp.wlx, sp.whx - x), fmin(z - sp.wlz, sp.whz - z));
    if (d < best) {
      best = d;
      vx = 0.0;
      vz = 0.0;
    }
  }
  fl_solid_corner c{};
  c.corner_index = k;
  c.signed_distance = best;
  c.wall_velocity = fl_vec3{vx, 0.0, vz};
  out[i] = c;
}

// MacViscosityCell3D of the pressure cells with one dynamic viscosity.
__global__ void k_viscosity_cells(const fl_pressure_cell* cells, uint64_t n, double mu, fl_viscosity_cell* out) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  fl_viscosity_cell c{};
  c.cell_index = cells[i].cell_index;
  c.dynamic_viscosity = mu;
  out[i] = c;
}

// Accepted link velocities into their grid slots.
__global__ void k_scatter_links(const uint64_t* slot, const double* u, uint64_t nf, double* unew) {
  const uint64_t f = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (f < nf) unew[slot[f]] = u[f];
}

__global__ void k_copy_slots(const double* a, uint64_t n, double* b) {
  const uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// The first error word of the last stage: its status, with the failing ordinal kept for the report.
uint64_t g_last_error_ordinal = ~uint64_t(0);
int error_code(fl_ctx* ctx) {
  uint64_t w = 0;
  FL_TRY(read_error(ctx, &w));
  if (w == ~uint64_t(0)) return FL_OK;
  g_last_error_ordinal = w >> 8;
  return static_cast<int>(w & 0xff);
}

float elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

int check_substep_options(const fl_substep_options* o) {
  if (!o) return FL_INVALID_ARGUMENT;
  if (!(o->step_duration > 0.0) || !finite(o->step_duration)) return FL_INVALID_ARGUMENT;
  if (!finite(o->gravity.x) || !finite(o->gravity.y) || !finite(o->gravity.z)) return FL_INVALID_ARGUMENT;
  int e;
  const fl_quantization& q = o->quantization;
  if (!std::isnormal(q.mass_per_quantum) || std::frexp(q.mass_per_quantum, &e) != 0.5 ||
      !std::isnormal(q.pure_momentum_per_quantum) || std::frexp(q.pure_momentum_per_quantum, &e) != 0.5)
    return FL_INVALID_ARGUMENT;
  if (!(o->minimum_face_mass >= 0.0) || !finite(o->minimum_face_mass)) return FL_INVALID_ARGUMENT;
  if (!(o->field_radius > 0.0) || !finite(o->field_radius)) return FL_INVALID_ARGUMENT;
  if (!(o->liquid_density > 0.0) || !finite(o->liquid_density)) return FL_INVALID_ARGUMENT;
  if (!(o->surface_tension >= 0.0) || !finite(o->surface_tension) || !finite(o->ambient_pressure))
    return FL_INVALID_ARGUMENT;
  if (!finite(o->particle_energy_blend) || o->particle_energy_blend < 0 || o->particle_energy_blend > 1)
    return FL_INVALID_ARGUMENT;
  if (o->obstacle.kind > FL_OBSTACLE_TANK) return FL_INVALID_ARGUMENT;
  if (!(o->dynamic_viscosity >= 0.0) || !finite(o->dynamic_viscosity)) return FL_INVALID_ARGUMENT;
  const fl_scene_solid& q2 = o->solid;
  if (q2.parts > 7 || q2.split_rule > 3) return FL_INVALID_ARGUMENT;
  if (q2.parts) {
    // The sampled-solid freeze takes an absent or half-space obstacle.
    if (o->obstacle.kind != FL_OBSTACLE_ABSENT && o->obstacle.kind != FL_OBSTACLE_HALFSPACE) return FL_INVALID_ARGUMENT;
    const double v[] = {q2.sphere_centre.x, q2.sphere_centre.y, q2.sphere_centre.z, q2.sphere_radius,
                        q2.paddle_pivot.x, q2.paddle_pivot.y, q2.paddle_pivot.z, q2.paddle_half_extent.x,
                        q2.paddle_half_extent.y, q2.paddle_half_extent.z, q2.paddle_angle,
                        q2.paddle_angular_velocity, q2.walls_low.x, q2.walls_low.z, q2.walls_high.x,
                        q2.walls_high.z};
    for (double a : v)
      if (!finite(a)) return FL_INVALID_ARGUMENT;
    if ((q2.parts & 1u) && !(q2.sphere_radius > 0.0)) return FL_INVALID_ARGUMENT;
    if ((q2.parts & 2u) && !(q2.paddle_half_extent.x > 0.0 && q2.paddle_half_extent.y > 0.0 &&
                             q2.paddle_half_extent.z > 0.0))
      return FL_INVALID_ARGUMENT;
    if ((q2.parts & 4u) && !(q2.walls_low.x < q2.walls_high.x && q2.walls_low.z < q2.walls_high.z))
      return FL_INVALID_ARGUMENT;
  }
  return FL_OK;
}

}  // namespace

void release_substep(fl_ctx* ctx) {
  auto* B = static_cast<SubstepBuffers*>(ctx->substep);
  if (!B) return;
  B->release();
  delete B;
  ctx->substep = nullptr;
}

void substep_scene_uploaded(fl_ctx* ctx) {
  if (ctx->substep) static_cast<SubstepBuffers*>(ctx->substep)->fresh = true;
}

int substep(fl_ctx* ctx, const fl_substep_options& o, fl_substep_report& rep) {
  FL_TRY(check_substep_options(&o));
  const uint64_t n = ctx->scene_n;
  const fl_quantization& q = o.quantization;
  if (q.declared_particle_capacity > ctx->cap.particles || n > q.declared_particle_capacity)
    return FL_RESOURCE_EXHAUSTED;
  if (!ctx->substep) ctx->substep = new SubstepBuffers();
  SubstepBuffers& B = *static_cast<SubstepBuffers*>(ctx->substep);
  for (hipEvent_t& e : B.ev)
    if (!e) FL_HIP(hipEventCreate(&e));
  hipStream_t s = ctx->stream;
  const uint64_t cap_p = ctx->cap.particles;
  FL_TRY(grow_to(&B.ratio, &B.cap[0], cap_p));
  FL_TRY(grow_to(&B.ratio_sorted, &B.cap[1], cap_p));
  FL_TRY(grow_to(&B.flags, &B.cap[2], cap_p));
  FL_TRY(grow_to(&B.flags_sorted, &B.cap[3], cap_p));
  if (B.fresh) {
    if (n) k_state_init<<<grid_for(n), kThreads, 0, s>>>(B.ratio, B.flags, n);
    FL_HIP(hipGetLastError());
    B.fresh = false;
  }
  rep = fl_substep_report{};
  rep.particles = n;
  const double dt = o.step_duration, h = ctx->geometry.h;
  hipEvent_t* ev = B.ev;
  g_last_error_ordinal = ~uint64_t(0);
  auto fail = [&](int stage, int code) {
    rep.stage = stage;
    rep.status = code;
    rep.error_ordinal = g_last_error_ordinal;
    return code;
  };
#define STAGE(stage, expr)                 \
  do {                                     \
    const int st_ = (expr);                \
    if (st_) return fail((stage), st_);    \
  } while (0)
  // 1. Trial topology: bins with the deposit's input checks and the stable sort.
  FL_HIP(hipEventRecord(ev[0], s));
  STAGE(1, bin_particles(ctx, ctx->raw.view(), n, true, false));
  STAGE(1, error_code(ctx));
  if (ctx->bin_box.volume() > ctx->cap.lookup_blocks) return fail(1, FL_RESOURCE_EXHAUSTED);
  STAGE(1, sort_particles(ctx, n, ctx->raw.view(), true));
  if (ctx->nvalid != n) return fail(1, FL_INTERNAL);
  if (n) k_state_permute<<<grid_for(n), kThreads, 0, s>>>(B.ratio, B.flags, ctx->perm, n, B.ratio_sorted, B.flags_sorted);
  FL_HIP(hipGetLastError());
  // 2. The liquid field: candidates, samples and distances.
  FL_HIP(hipEventRecord(ev[1], s));
  FieldSamples F{};
  STAGE(2, sample_particle_field(ctx, o.field_radius, ctx->cap.pressure_cells, &F));
  rep.candidates = F.cell_count;
  rep.centre_samples = F.centre_count;
  rep.corner_samples = F.corner_count;
  // Activation: the particles' stencil owners and the owners of the candidates' faces.
  {
    Box box;
    for (int a = 0; a < 3; ++a) {
      const int32_t flo[3] = {F.cell_box_lo.x, F.cell_box_lo.y, F.cell_box_lo.z};
      const int32_t fhi[3] = {F.cell_box_hi.x, F.cell_box_hi.y, F.cell_box_hi.z};
      const int32_t lo = std::min<int32_t>(ctx->bin_box.lo[a], static_cast<int32_t>(floor_div4(flo[a])));
      const int32_t hi = std::max<int32_t>(ctx->bin_box.lo[a] + ctx->bin_box.dim[a],
                                           static_cast<int32_t>(floor_div4(int64_t(fhi[a]) + 1)));
      box.lo[a] = lo;
      box.dim[a] = hi - lo + 1;
    }
    if (box.volume() > ctx->cap.lookup_blocks) return fail(2, FL_RESOURCE_EXHAUSTED);
    ctx->block_box = box;
    FL_TRY(reset_error(ctx));
    STAGE(2, activate_blocks(ctx, n, reinterpret_cast<const int32_t*>(F.cells), F.cell_count));
    STAGE(2, error_code(ctx));
  }
  const uint64_t nb = ctx->nb, nslots = 192 * nb;
  rep.blocks = nb;
  // 3. Deposit, normalize and gravity.
  FL_HIP(hipEventRecord(ev[2], s));
  {
    const uint64_t cap = std::max<uint64_t>(1, q.declared_particle_capacity);
    int bits = 0;
    for (uint64_t v = cap - 1; v; v >>= 1) ++bits;
    const double limit = std::ldexp(1.0, 62 - bits);
    FL_TRY(reset_error(ctx));
    STAGE(3, p2g_quanta(ctx, n, 1.0 / q.mass_per_quantum, 1.0 / q.pure_momentum_per_quantum, limit));
    STAGE(3, error_code(ctx));
    const double g3[3] = {o.gravity.x, o.gravity.y, o.gravity.z};
    FL_TRY(reset_error(ctx));
    STAGE(3, normalize_forces(ctx, nslots, q.mass_per_quantum, q.pure_momentum_per_quantum, o.minimum_face_mass, g3,
                              dt));
    STAGE(3, error_code(ctx));
  }
  // 4. Frozen geometry of the candidates.
  FL_HIP(hipEventRecord(ev[3], s));
  const uint64_t ncand = F.cell_count;
  const uint64_t nwords = 3 * nb;
  {
    const uint64_t partials = std::max<uint64_t>(std::max<uint64_t>(64, ctx->cap.lookup_blocks) / 128 + 4096, 8192);
    for (uint64_t m : {ncand, nwords})
      if (2 * (m / 1024 + 2) + 16 > partials) return fail(4, FL_RESOURCE_EXHAUSTED);
  }
  FL_TRY(grow_to(&B.keep, &B.cap[4], 2 * ncand));
  FL_TRY(grow_to(&B.staged, &B.cap[5], ncand));
  FL_TRY(grow_to(&B.slots, &B.cap[6], 6 * ncand));
  FL_TRY(grow_to(&B.words, &B.cap[7], nwords));
  FL_TRY(grow_to(&B.word_counts, &B.cap[8], 2 * nwords));
  FL_TRY(grow_to(&B.face_slots, &B.cap[9], 6 * ncand));
  FL_TRY(grow_to(&B.cells, &B.cap[10], ncand));
  FL_TRY(grow_to(&B.faces, &B.cap[11], 6 * ncand));
  FL_TRY(grow_to(&B.needs, &B.cap[12], ncand));
  FL_TRY(grow_to(&B.counts, &B.cap[13], 4));
  FL_TRY(grow_to(&B.phi, &B.cap[14], 9 * ncand));
  FL_TRY(grow_to(&B.work, &B.cap[15], ncand));
  FL_TRY(grow_to(&B.uncut, &B.cap[16], 2));
  const double full_area = h * h, full_volume = full_area * h;
  const FreezeDevice fd{reinterpret_cast<const fl_int3*>(ctx->blocks),
                        nb,
                        fl_vec3{ctx->geometry.ox, ctx->geometry.oy, ctx->geometry.oz},
                        h,
                        full_volume,
                        F.cells,
                        ncand,
                        F.centres,
                        F.centre_count,
                        F.corners,
                        F.corner_count,
                        o.obstacle,
                        B.keep,
                        B.keep + ncand,
                        B.staged,
                        B.slots,
                        B.words,
                        nwords,
                        B.word_counts,
                        B.word_counts + nwords,
                        B.face_slots,
                        B.cells,
                        B.faces,
                        B.needs,
                        B.counts,
                        B.phi,
                        B.work,
                        B.uncut,
                        0,
                        nullptr,
                        0,
                        nullptr,
                        0,
                        nullptr};
  FreezeDevice fd_run = fd;
  if (o.solid.parts) {
    FL_TRY(grow_to(&B.solids, &B.cap[24], F.corner_count));
    FL_TRY(grow_to(&B.solid_phi, &B.cap[25], 8 * ncand));
    const fl_scene_solid& q2 = o.solid;
    SolidParams sp{};
    sp.parts = q2.parts;
    sp.sx = q2.sphere_centre.x, sp.sy = q2.sphere_centre.y, sp.sz = q2.sphere_centre.z, sp.sr = q2.sphere_radius;
    sp.px = q2.paddle_pivot.x, sp.py = q2.paddle_pivot.y, sp.pz = q2.paddle_pivot.z;
    sp.ex = q2.paddle_half_extent.x, sp.ey = q2.paddle_half_extent.y, sp.ez = q2.paddle_half_extent.z;
    sp.cos_a = std::cos(q2.paddle_angle);
    sp.sin_a = std::sin(q2.paddle_angle);
    sp.omega = q2.paddle_angular_velocity;
    sp.wlx = q2.walls_low.x, sp.wlz = q2.walls_low.z, sp.whx = q2.walls_high.x, sp.whz = q2.walls_high.z;
    if (F.corner_count)
      k_sample_solid<<<grid_for(F.corner_count), kThreads, 0, s>>>(F.corners, F.corner_count, ctx->geometry.ox,
                                                                   ctx->geometry.oy, ctx->geometry.oz, h, sp,
                                                                   B.solids);
    FL_HIP(hipGetLastError());
    fd_run.sampled = q2.split_rule == 0 ? 1 : static_cast<int>(q2.split_rule);
    fd_run.solids = B.solids;
    fd_run.nsolids = F.corner_count;
    fd_run.solid_phi = B.solid_phi;
  }
  auto read_counts = [&]() -> int {
    FL_HIP(hipMemcpyAsync(ctx->host_words, B.counts, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    return FL_OK;
  };
  FL_TRY(reset_error(ctx));
  STAGE(4, freeze_cells_device(ctx, fd_run));
  STAGE(4, error_code(ctx));
  FL_TRY(read_counts());
  const uint64_t nc = ncand ? ctx->host_words[0] : 0, nf = ncand ? ctx->host_words[1] : 0;
  if (nc == 0) return fail(4, FL_FAILED_PRECONDITION);
  STAGE(4, freeze_faces_device(ctx, fd_run, nc, nf, o.surface_tension != 0));
  STAGE(4, error_code(ctx));
  FL_TRY(read_counts());
  const bool capillary = ctx->host_words[2] != 0;
  STAGE(4, freeze_curvature_device(ctx, fd_run, nc, false, 0.0));
  STAGE(4, error_code(ctx));
  rep.capillary_step_bound = __builtin_inf();
  if (capillary) {
    constexpr double kPi = 3.141592653589793238462643383279502884;
    const double square = o.liquid_density * full_volume / (4 * kPi * o.surface_tension);
    if (!std::isnormal(square)) return fail(4, FL_OUT_OF_RANGE);
    rep.capillary_step_bound = std::sqrt(square);
    if (dt > rep.capillary_step_bound) return fail(4, FL_FAILED_PRECONDITION);
  }
  rep.frozen_cells = nc;
  rep.frozen_faces = nf;
  // 5. Typed pressure links with the predicted face velocities.
  FL_HIP(hipEventRecord(ev[4], s));
  if (nc > ctx->cap.pressure_cells || nf > ctx->cap.pressure_faces) return fail(5, FL_RESOURCE_EXHAUSTED);
  FL_TRY(grow_to(&B.fields, &B.cap[17], nc));
  FL_TRY(grow_to(&B.predicted, &B.cap[18], nf));
  FL_TRY(grow_to(&B.pcells, &B.cap[19], nc));
  FL_TRY(grow_to(&B.pfaces, &B.cap[20], nf));
  FL_TRY(grow_to(&B.lower, &B.cap[21], 1));
  k_cell_fields<<<grid_for(nc), kThreads, 0, s>>>(B.fields, nc, o.liquid_density);
  k_predicted<<<grid_for(nf), kThreads, 0, s>>>(B.faces, nf, ctx->ustar, B.predicted);
  FL_HIP(hipGetLastError());
  {
    const LinkDevice ld{reinterpret_cast<const fl_int3*>(ctx->blocks), nb, F.centres, F.centre_count, B.cells, nc,
                        B.faces, nf, h, dt, o.liquid_density, o.surface_tension, B.fields, B.predicted,
                        o.ambient_pressure, nullptr, 0, nullptr, 0, nullptr, 0};
    FL_TRY(reset_error(ctx));
    STAGE(5, build_links_device(ctx, ld, nullptr, nullptr, nullptr, true));
    STAGE(5, error_code(ctx));
    FL_HIP(hipMemsetAsync(B.lower, 0, sizeof(uint64_t), s));
    STAGE(5, build_links_device(ctx, ld, B.pcells, B.pfaces, B.lower, false));
    STAGE(5, error_code(ctx));
    FL_HIP(hipMemcpyAsync(ctx->host_words, B.lower, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    rep.lower_theta_count = ctx->host_words[0];
  }
  // 6. Projection: u_new = u* off the pressure faces, the corrected velocities on them.
  FL_HIP(hipEventRecord(ev[5], s));
  k_copy_slots<<<grid_for(nslots), kThreads, 0, s>>>(ctx->ustar, nslots, ctx->unew);
  FL_HIP(hipGetLastError());
  FL_TRY(reset_error(ctx));
  if (o.dynamic_viscosity > 0.0) {
    // The coupled viscous pressure solve on the full-cell strain rows of the pressure cells.
    // preconditioner 2: the reference's Jacobi diagonal; otherwise the V-cycle of the effective conductances.
    const bool jacobi = o.pressure.preconditioner == 2;
    fl_pressure_options popt = o.pressure;
    popt.preconditioner = 0;
    CoupledSystem S{};
    STAGE(6, coupled_prepare_device(ctx, B.pcells, nc, B.pfaces, nf, dt, popt, &S));
    ViscosityDevice V{};
    STAGE(6, viscosity_reserve(ctx, nc, 6 * nc, nf, &V));
    k_viscosity_cells<<<grid_for(nc), kThreads, 0, s>>>(B.pcells, nc, o.dynamic_viscosity,
                                                        const_cast<fl_viscosity_cell*>(V.cells));
    FL_HIP(hipGetLastError());
    uint64_t nr = 0, nt = 0;
    STAGE(6, assemble_strain(ctx, V, nc, h, &nr));
    rep.strain_rows = nr;
    STAGE(6, coupled_rows(ctx, S, V, nr, dt));
    STAGE(6, weight_viscous_rows(ctx, V, nr, nf, dt, &nt));
    STAGE(6, viscosity_reserve_terms(ctx, nt, &V));
    CoupledResult cr{};
    const int cst = coupled_solve(ctx, S, V, nr, nt, dt, popt, jacobi, &cr);
    rep.viscous.pressure_iterations = cr.iterations;
    rep.viscous.constraint_residual_l2 = cr.residual_l2;
    rep.viscous.initial_constraint_residual_l2 = cr.initial_residual_l2;
    rep.viscous.max_cell_constraint_defect = cr.max_defect;
    rep.viscous.max_momentum_residual = cr.max_momentum_residual;
    rep.viscous.velocity_solves = cr.velocity_solves;
    rep.viscous.velocity_iterations = cr.velocity_iterations;
    STAGE(6, cst);
    const double *cp = nullptr, *cu = nullptr, *cd = nullptr;
    STAGE(6, coupled_outputs(ctx, &cp, &cu, &cd));
    k_scatter_links<<<grid_for(nf), kThreads, 0, s>>>(S.slot, cu, nf, ctx->unew);
    FL_HIP(hipGetLastError());
  } else {
    STAGE(6, project_links_device(ctx, B.pcells, nc, B.pfaces, nf, dt, o.pressure, ctx->unew, &rep.pressure));
  }
  // 7. Gather (GatherMac3D of the old and new face velocities).
  FL_HIP(hipEventRecord(ev[6], s));
  FL_TRY(reset_error(ctx));
  STAGE(7, g2p(ctx, n, ctx->u0, ctx->unew));
  STAGE(7, error_code(ctx));
  // 8. Particle update of the sorted columns; the liquid field at every particle is -r (the union
  // includes the particle's own sphere).
  FL_HIP(hipEventRecord(ev[7], s));
  {
    fl_particle_update_options uo{};
    uo.cell_width = h;
    uo.step_duration = dt;
    uo.particle_energy_blend = o.particle_energy_blend;
    uo.friction_coefficient = o.friction_coefficient;
    uo.penetration_tolerance = o.penetration_tolerance;
    uo.obstacle = o.obstacle;
    ParticleColumns P{};
    for (int a = 0; a < 3; ++a) {
      P.x[a] = ctx->sorted.column(a);
      P.v[a] = ctx->sorted.column(7 + a);
    }
    for (int k = 0; k < 9; ++k) P.c[k] = ctx->sorted.column(10 + k);
    P.volume_ratio = B.ratio_sorted;
    P.flags = B.flags_sorted;
    P.ordinal = ctx->perm;
    FL_TRY(reset_error(ctx));
    FL_HIP(hipMemsetAsync(B.lower, 0, sizeof(uint64_t), s));
    FL_TRY(grow_to(&B.contact, &B.cap[22], n));
    FL_TRY(grow_to(&B.contact_count, &B.cap[23], 1));
    STAGE(8, update_columns(ctx, uo, n, P, ctx->gather, ctx->sorted.capacity, -o.field_radius,
                            reinterpret_cast<unsigned long long*>(B.lower), B.contact, B.contact_count));
    STAGE(8, error_code(ctx));
    FL_HIP(hipMemcpyAsync(ctx->host_words, B.lower, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    std::memcpy(&rep.max_speed, ctx->host_words, sizeof(double));
  }
  FL_HIP(hipEventRecord(ev[8], s));
  FL_HIP(hipEventSynchronize(ev[8]));
#undef STAGE
  // The sorted, updated particles are the scene's particles from here on.
  std::swap(ctx->raw, ctx->sorted);
  ctx->raw.trial_aliases_sorted = true;
  ctx->sorted.trial_aliases_sorted = true;
  std::swap(B.ratio, B.ratio_sorted);
  std::swap(B.flags, B.flags_sorted);
  ctx->scene_ready = false;
  rep.sort_ms = elapsed_ms(ev[0], ev[1]);
  rep.field_ms = elapsed_ms(ev[1], ev[2]);
  rep.deposit_ms = elapsed_ms(ev[2], ev[3]);
  rep.freeze_ms = elapsed_ms(ev[3], ev[4]);
  rep.links_ms = elapsed_ms(ev[4], ev[5]);
  rep.projection_ms = elapsed_ms(ev[5], ev[6]);
  rep.gather_ms = elapsed_ms(ev[6], ev[7]);
  rep.update_ms = elapsed_ms(ev[7], ev[8]);
  rep.total_ms = elapsed_ms(ev[0], ev[8]);
  return FL_OK;
}

// ProjectMacPressure3D of host records with the system and hierarchy built on the device (the
// substep's projection path): the records are copied to the device, the grid is the context's.
int project_records_device(fl_ctx* ctx, const fl_pressure_cell* cells, uint64_t nc, const fl_pressure_face* faces,
                           uint64_t nf, double dt, const fl_pressure_options& opt, double* out_pressures,
                           double* out_velocities, fl_pressure_diagnostics* diag) {
  if (!ctx->substep) ctx->substep = new SubstepBuffers();
  SubstepBuffers& B = *static_cast<SubstepBuffers*>(ctx->substep);
  if (nc > ctx->cap.pressure_cells || nf > ctx->cap.pressure_faces) return FL_RESOURCE_EXHAUSTED;
  FL_TRY(grow_to(&B.pcells, &B.cap[19], nc));
  FL_TRY(grow_to(&B.pfaces, &B.cap[20], nf));
  FL_TRY(grow_to(&B.predicted, &B.cap[18], 192 * std::max<uint64_t>(ctx->nb, 1)));
  hipStream_t s = ctx->stream;
  if (nc) FL_HIP(hipMemcpyAsync(B.pcells, cells, nc * sizeof(fl_pressure_cell), hipMemcpyHostToDevice, s));
  if (nf) FL_HIP(hipMemcpyAsync(B.pfaces, faces, nf * sizeof(fl_pressure_face), hipMemcpyHostToDevice, s));
  FL_TRY(reset_error(ctx));
  FL_TRY(project_links_device(ctx, B.pcells, nc, B.pfaces, nf, dt, opt, B.predicted, diag));
  const double* p = nullptr;
  uint64_t np = 0;
  FL_TRY(last_pressures(ctx, &p, &np));
  if (np != nc) return FL_INTERNAL;
  if (out_pressures && nc) FL_HIP(hipMemcpyAsync(out_pressures, p, nc * sizeof(double), hipMemcpyDeviceToHost, s));
  if (out_velocities && nf) {
    // The corrected velocities went to their owned slots; gather them back in face order.
    std::vector<double> slots(192 * ctx->nb);
    std::vector<int32_t> lut(ctx->block_box.volume());
    FL_HIP(hipMemcpyAsync(slots.data(), B.predicted, slots.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    if (!lut.empty())
      FL_HIP(hipMemcpyAsync(lut.data(), ctx->lut, lut.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    for (uint64_t f = 0; f < nf; ++f) {
      const fl_pressure_face& q = faces[f];
      const int32_t l[3] = {q.lattice.x, q.lattice.y, q.lattice.z};
      const int64_t b[3] = {floor_div4(l[0]), floor_div4(l[1]), floor_div4(l[2])};
      const int64_t bi = ctx->block_box.index(int32_t(b[0]), int32_t(b[1]), int32_t(b[2]));
      const int32_t g = bi >= 0 ? lut[bi] : -1;
      if (g < 0) return FL_INTERNAL;
      const uint64_t slot = uint64_t(g) * 192 + uint64_t(q.axis) * 64 +
                            uint64_t(((l[2] - 4 * b[2]) * 4 + (l[1] - 4 * b[1])) * 4 + (l[0] - 4 * b[0]));
      out_velocities[f] = slots[slot];
    }
  }
  FL_HIP(hipStreamSynchronize(s));
  return FL_OK;
}

// Particle columns and state of the scene in its current order (the last substep's sort order).
int substep_download(fl_ctx* ctx, double* positions, double* velocities, double* slopes, double* ratios,
                     uint8_t* flags, uint64_t n) {
  if (n != ctx->scene_n) return FL_INVALID_ARGUMENT;
  hipStream_t s = ctx->stream;
  SubstepBuffers* B = static_cast<SubstepBuffers*>(ctx->substep);
  std::vector<double> col(n);
  auto pull = [&](int column, double* out, int stride, int offset) -> int {
    if (!n) return FL_OK;
    FL_HIP(hipMemcpyAsync(col.data(), ctx->raw.column(column), n * sizeof(double), hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    for (uint64_t i = 0; i < n; ++i) out[i * stride + offset] = col[i];
    return FL_OK;
  };
  for (int a = 0; a < 3; ++a) {
    if (positions) FL_TRY(pull(a, positions, 3, a));
    if (velocities) FL_TRY(pull(7 + a, velocities, 3, a));
  }
  if (slopes)
    for (int k = 0; k < 9; ++k) FL_TRY(pull(10 + k, slopes, 9, k));
  const bool have = B && !B->fresh;
  if (ratios) {
    if (have && n) {
      FL_HIP(hipMemcpy(ratios, B->ratio, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      for (uint64_t i = 0; i < n; ++i) ratios[i] = 1.0;
    }
  }
  if (flags) {
    if (have && n) {
      FL_HIP(hipMemcpy(flags, B->flags, n, hipMemcpyDeviceToHost));
    } else {
      for (uint64_t i = 0; i < n; ++i) flags[i] = 0;
    }
  }
  return FL_OK;
}

}  // namespace fl
