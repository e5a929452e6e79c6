// Host-side context of the HIP hot path: device buffers, stream, timing and launch helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/fluidity_hip.h"
#include "lattice.h"

#define FL_HIP(expr)                                                                          \
  do {                                                                                        \
    hipError_t fl_e_ = (expr);                                                                \
    if (fl_e_ != hipSuccess) {                                                                \
      std::fprintf(stderr, "fluidity_hip: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(fl_e_), \
                   __FILE__, __LINE__);                                                       \
      return FL_INTERNAL;                                                                     \
    }                                                                                         \
  } while (0)

#define FL_TRY(expr)        \
  do {                      \
    int fl_s_ = (expr);     \
    if (fl_s_) return fl_s_; \
  } while (0)

namespace fl {

constexpr int kThreads = 256;
constexpr int kParticleDoubles = 19;  // sorted xyz, trial xyz, mass, velocity xyz, slopes 3x3 row-major
constexpr int kGatherDoubles = 24;    // MacParticleGatherResult3D

// Structure-of-arrays particle view; trial may alias sorted.
struct ParticleView {
  const double* s[3];
  const double* t[3];
  const double* m;
  const double* v[3];
  const double* c[9];
};

struct ParticleBuffers {
  double* base = nullptr;  // kParticleDoubles * capacity
  uint64_t capacity = 0;
  bool trial_aliases_sorted = false;
  ParticleView view() const {
    ParticleView p;
    for (int a = 0; a < 3; ++a) p.s[a] = base + a * capacity;
    for (int a = 0; a < 3; ++a) p.t[a] = trial_aliases_sorted ? p.s[a] : base + (3 + a) * capacity;
    p.m = base + 6 * capacity;
    for (int a = 0; a < 3; ++a) p.v[a] = base + (7 + a) * capacity;
    for (int k = 0; k < 9; ++k) p.c[k] = base + (10 + k) * capacity;
    return p;
  }
  double* column(int k) const { return base + k * capacity; }
};

// Dense ordinal table over a box of blocks (or bins), (z,y,x) linear order.
struct Box {
  int32_t lo[3] = {0, 0, 0};
  int32_t dim[3] = {0, 0, 0};
  __host__ __device__ uint64_t volume() const { return uint64_t(dim[0]) * dim[1] * dim[2]; }
  __host__ __device__ int64_t index(int32_t x, int32_t y, int32_t z) const {
    const int64_t rx = int64_t(x) - lo[0], ry = int64_t(y) - lo[1], rz = int64_t(z) - lo[2];
    if (rx < 0 || ry < 0 || rz < 0 || rx >= dim[0] || ry >= dim[1] || rz >= dim[2]) return -1;
    return (rz * dim[1] + ry) * dim[0] + rx;
  }
};

// The sparse grid as the kernels see it: canonical blocks plus a dense lookup over their box.
struct GridView {
  Geometry g;
  const int32_t* blocks;  // 3 * nb, xyz interleaved
  const int32_t* lut;     // box.volume() ordinals, -1 if inactive
  Box box;
  uint64_t nb;
  __device__ int32_t ordinal(int32_t x, int32_t y, int32_t z) const {
    const int64_t i = box.index(x, y, z);
    return i < 0 ? -1 : lut[i];
  }
};

// Non-empty bins of the sorted particle order.
struct BinList {
  const uint32_t* key;     // linear key in the bin box
  const uint32_t* offset;  // first sorted particle
  const uint32_t* count;
  uint32_t n;
  Box box;
};

}  // namespace fl

struct fl_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t stage[8] = {};
  fl_capacities cap{};
  char arch[32] = {};
  bool poisoned = false;

  // particles
  fl::ParticleBuffers raw, sorted;
  fl_particle* staging_aos = nullptr;  // device copy of caller records
  int32_t* bin = nullptr;              // 3 * cap.particles
  uint32_t* key = nullptr;
  uint32_t* rank = nullptr;
  uint32_t* perm = nullptr;
  uint32_t* perm_tmp = nullptr;
  // bins (capacity cap.lookup_blocks)
  uint32_t* bin_count = nullptr;
  uint64_t* bin_scan = nullptr;
  uint32_t* bin_key = nullptr;
  uint32_t* bin_offset = nullptr;
  uint32_t* bin_n = nullptr;
  uint64_t* scan_partial = nullptr;  // block sums of the scans
  // grid
  int32_t* blocks = nullptr;  // 3 * cap.blocks
  int32_t* lut = nullptr;     // cap.lookup_blocks
  uint8_t* mark = nullptr;    // activation bitmap, cap.lookup_blocks
  uint64_t nb = 0;
  fl::Box block_box;
  fl::Geometry geometry{};
  // faces (cap.blocks * 192)
  int64_t* qm = nullptr;
  int64_t* qp = nullptr;
  double* face_mass = nullptr;
  double* u0 = nullptr;
  double* ustar = nullptr;
  double* unew = nullptr;
  uint8_t* validity = nullptr;
  fl_face_quanta* quanta_aos = nullptr;
  // gather output, kGatherDoubles * cap.particles (component-major)
  double* gather = nullptr;
  fl_gather_result* gather_aos = nullptr;
  // scalars
  uint64_t* err = nullptr;      // first error word
  int32_t* reduce_i32 = nullptr;  // 16 ints: box min/max etc.
  uint64_t* host_words = nullptr;  // pinned host scratch (64 words)

  // device-resident scene
  // Transfer arithmetic of the device-resident step and substep: 64 (bit-exact FP64, the default) or
  // 32 (FP32 stencils and contributions into the same integer quanta).
  int transfer_bits = 64;
  uint64_t scene_n = 0;
  double scene_mass = 0;
  bool scene_loaded = false;  // fl_scene_upload ran and no span-based call reused its buffers since
  bool scene_ready = false;   // ... and fl_scene_prepare ran on it
  uint32_t nbins = 0;
  uint64_t nvalid = 0;
  fl::Box bin_box;

  // pressure solver state (opaque, owned by pressure.hip)
  void* pressure = nullptr;
  // learned preconditioner state (fl::LearnState, owned by learn.hip)
  void* learn = nullptr;
  // boundary-gather history buffers (owned by capi.hip, allocated on first use)
  void* boundary = nullptr;
  // topology-preparation buffers (owned by capi.hip, allocated on first use)
  void* topology = nullptr;
  // particle-update buffers (owned by capi.hip, allocated on first use)
  void* update = nullptr;
  // pressure-link buffers (owned by capi.hip, allocated on first use)
  void* links = nullptr;
  void* freeze = nullptr;
  void* viscosity = nullptr;  // ViscosityBuffers (capi.hip)
  void* field = nullptr;      // FieldBuffers (field.hip)
  void* substep = nullptr;    // SubstepBuffers (substep.hip)
  void* coupled = nullptr;    // CoupledBuffers (coupled.hip)
};

namespace fl {

// Geometry scalars for the kernels, with the exact reciprocal of a power-of-two cell width.
inline Geometry make_geometry(double ox, double oy, double oz, double h) {
  int e = 0;
  const bool pow2 = std::isnormal(h) && std::frexp(h, &e) == 0.5;
  return Geometry{ox, oy, oz, h, pow2 ? 1.0 / h : 0.0};
}

inline unsigned grid_for(uint64_t n, int threads = kThreads) {
  const uint64_t g = (n + threads - 1) / threads;
  return static_cast<unsigned>(g == 0 ? 1 : g);
}

// Topology (topology.hip)
int bin_particles(fl_ctx* ctx, const ParticleView& p, uint64_t n, bool deposit_checks, bool gather_checks);
// Bin with the input checks of one caller: plain bin query, quantized/FP64 deposit (mass, velocity,
// slopes), gather, compatible deposit (mass and velocity; enabled slope rows are checked per axis), or
// boundary gather (velocity).
enum { kBinPlain = 0, kBinDeposit = 1, kBinGather = 2, kBinCompatible = 3, kBinBoundaryGather = 4 };
int bin_particles_mode(fl_ctx* ctx, const ParticleView& p, uint64_t n, int mode);
int sort_particles(fl_ctx* ctx, uint64_t n, const ParticleView& src, bool permute, uint32_t permute_columns = ~0u);
int build_lookup(fl_ctx* ctx, const int32_t* host_blocks, uint64_t nb);
int activate_blocks(fl_ctx* ctx, uint64_t n, const int32_t* d_cells, uint64_t cell_count);
int write_sort_entries(fl_ctx* ctx, uint64_t n, fl_sort_entry* d_out);
GridView grid_view(fl_ctx* ctx);
BinList bin_list(fl_ctx* ctx);

// Transfers (transfer.hip)
int p2g_quanta(fl_ctx* ctx, uint64_t n, double inv_qm, double inv_qp, double limit, hipEvent_t before = nullptr,
               hipEvent_t after = nullptr, bool f64 = false);
int normalize_forces(fl_ctx* ctx, uint64_t nf, double qm, double qp, double min_mass, const double g[3], double dt);
// k_normalize_forces on explicit device arrays and stream (no context); err holds two words preset to ~0.
int launch_normalize_forces(hipStream_t s, const int64_t* qm, const int64_t* qp, const int64_t* qs, uint64_t nf,
                            double sqm, double sqp, double min_mass, const double g[3], double dt, double* mass,
                            double* u0, uint8_t* validity, double* ustar, uint64_t* err);
int apply_forces(fl_ctx* ctx, const double* d_old, uint64_t nf, const double g[3], double dt, double* d_pred);
int g2p(fl_ctx* ctx, uint64_t n, const double* d_old, const double* d_new, hipEvent_t before = nullptr,
        hipEvent_t after = nullptr);
int unpack_particles(fl_ctx* ctx, uint64_t n);
// Compatible deposit of the sorted particles into ctx->qm / ctx->qp; slot tags per owned slot,
// particle side tags and affine switches in caller order (device arrays).
int p2g_compatible(fl_ctx* ctx, uint64_t n, const int64_t* slot_side, const uint8_t* slot_ok, const int64_t* side,
                   const uint8_t* affine, double inv_qm, double inv_qp, double limit);
// Device arrays of one boundary gather: per history (0 old, 1 new) the face velocities, validity,
// side tags, eligibility and walls; particle side tags and affine switches in caller order; the
// retention and flag columns; the continuation list and its counter.
struct BoundaryDevice {
  const double* u[2];
  const uint8_t* validity[2];
  const int64_t* side[2];
  const uint8_t* eligible[2];
  const fl_wall_sample* walls[2];
  int same;
  const int64_t* pside;
  const uint8_t* paffine;
  double* extra;
  uint64_t extra_cap;
  uint32_t* slow;
  uint32_t* slow_count;
};
// Boundary gather of the sorted particles into ctx->gather (24 columns) and d.extra (4 columns);
// *slow_count receives the number of particles that needed a continuation.
int g2p_boundary(fl_ctx* ctx, uint64_t n, const BoundaryDevice& d, uint32_t* slow_count);
// Compatible stencils of caller particles [lo, lo + count) from ctx->raw and ctx->bin (81 per particle).
int compatible_stencils(fl_ctx* ctx, uint64_t lo, uint64_t count, const int64_t* slot_side, const uint8_t* slot_ok,
                        const int64_t* side, const uint8_t* affine, fl_compatible_sample* d_out);

// Topology preparation (history.hip). Checks write the first failing (phase, ordinal, code) into
// ctx->err: previous blocks, cells and faces (phases 0-2), then particles and pressure cells (3-4).
int check_history(fl_ctx* ctx, const fl_block_state* b, uint64_t nb, const fl_cell_state* c, uint64_t nc,
                  const fl_face_state* f, uint64_t nf);
int check_topology_inputs(fl_ctx* ctx, uint64_t n, const int32_t* cells, uint64_t nc);
// Owner-block box of the pressure cells' faces (d_red: 6 device ints of scratch).
int cell_owner_box(fl_ctx* ctx, const int32_t* cells, uint64_t nc, int32_t* d_red, int32_t lo[3], int32_t hi[3]);
// Canonical unique cells of the (checked) pressure cells into out; words holds cap.lookup_blocks
// 64-bit words and d_count one device word.
int unique_cells(fl_ctx* ctx, const int32_t* cells, uint64_t nc, void* words, int32_t* out, uint64_t* d_count,
                 uint64_t* count);
// Block, cell and face records of ctx->blocks / cells with the previous state's history.
int remap_topology(fl_ctx* ctx, const fl_block_state* pb, uint64_t pnb, const fl_cell_state* pc, uint64_t pnc,
                   const fl_face_state* pf, uint64_t pnf, uint64_t generation, const int32_t* cells, uint64_t nc,
                   int64_t* old, fl_block_state* ob, fl_cell_state* oc, fl_face_state* of);

// Typed pressure links of a frozen view (links.hip): device copies of the view and the fields.
struct LinkDevice {
  const fl_int3* blocks;
  uint64_t nb;
  const fl_center_distance* centers;
  uint64_t ncenters;
  const fl_frozen_cell* cells;
  uint64_t ncells;
  const fl_frozen_face* faces;
  uint64_t nfaces;
  double h, step, density, sigma;
  const fl_cell_fields* fields;
  const double* predicted;
  double ambient;
  const fl_optional_flux* wetted;
  uint64_t nwetted;
  const fl_air_pressure* air;
  uint64_t nair;
  const fl_internal_interface* interfaces;
  uint64_t ninterfaces;
};
// Frozen-geometry device arrays: inputs, per-candidate staging, the face-slot bitmap and the outputs.
struct FreezeDevice {
  const fl_int3* blocks;
  uint64_t nb;
  fl_vec3 origin;
  double h, full_volume;
  const fl_int3* candidates;
  uint64_t ncand;
  const fl_center_distance* centers;
  uint64_t ncenters;
  const fl_center_distance* corners;
  uint64_t ncorners;
  fl_obstacle obstacle;
  uint64_t* keep;        // ncand
  uint64_t* keep_scan;   // ncand
  fl_frozen_cell* staged;  // ncand
  uint64_t* slots;       // 6 * ncand
  unsigned long long* words;  // nwords = 3 * nb
  uint64_t nwords;
  uint64_t* word_counts;  // nwords
  uint64_t* word_scan;    // nwords
  uint64_t* face_slots;   // 6 * ncand
  fl_frozen_cell* cells;  // ncand
  fl_frozen_face* faces;  // 6 * ncand
  uint8_t* needs;         // ncand
  uint64_t* counts;       // 4: kept cells, unique faces, curvature needed, listed cut cells
  double* phi;            // 9 * ncand: corner and centre samples of the cut cells
  uint64_t* work;         // ncand: the cut cells
  double* uncut_volume;   // 2: the uncut cell's volume, plane mode then sampled mode
  // Sampled solid (sampled != 0): corner samples, dense face normal velocities (count 0: none), and the
  // eight solid corner distances of each listed cell.
  int sampled;
  const fl_solid_corner* solids;
  uint64_t nsolids;
  const double* face_velocities;
  uint64_t nface_velocities;
  double* solid_phi;  // 8 * ncand
};
// Phases 0-2 (sample and candidate checks, kCell, compaction, unique face slots).
int freeze_cells_device(fl_ctx* ctx, const FreezeDevice& d);
// Phases 3-4 over nc kept cells and nf faces (kFace; curvature needs when the surface tension is not zero).
int freeze_faces_device(fl_ctx* ctx, const FreezeDevice& d, uint64_t nc, uint64_t nf, bool needs_curvature);
// Phase 5: the exact curvature, or kCurvature, on cells that need it.
int freeze_curvature_device(fl_ctx* ctx, const FreezeDevice& d, uint64_t nc, bool exact, double curvature);
// validate_only: the view's record checks (ValidateFrozenView); otherwise the field checks and the
// links (LinksCore). Failures go to ctx->err; *d_lower (device) counts raised theta values.
int build_links_device(fl_ctx* ctx, const LinkDevice& d, fl_pressure_cell* out_cells, fl_pressure_face* out_faces,
                       uint64_t* d_lower, bool validate_only);

// The particle-union liquid field (field.hip) of the sorted particles: candidate cells, and centre and
// corner samples with their signed distances, all (z,y,x) sorted device arrays owned by the context.
struct FieldSamples {
  const fl_int3* cells;
  uint64_t cell_count;
  fl_center_distance* centres;
  uint64_t centre_count;
  fl_center_distance* corners;
  uint64_t corner_count;
  fl_int3 cell_box_lo, cell_box_hi;  // a box holding every sample
};
int sample_particle_field(fl_ctx* ctx, double radius, uint64_t max_candidates, FieldSamples* out);
void release_field(fl_ctx* ctx);

// Projection of device link records on the context's grid (pressure.hip): the system and hierarchy
// built on the device, solved with the V-cycle PCG, corrected velocities scattered into unew_slots.
int project_links_device(fl_ctx* ctx, const fl_pressure_cell* cells, uint64_t nc, const fl_pressure_face* faces,
                         uint64_t nf, double dt, const fl_pressure_options& opt, double* unew_slots,
                         fl_pressure_diagnostics* diag);
int last_pressures(fl_ctx* ctx, const double** p, uint64_t* n);

struct ViscosityDevice;

// The prepared pressure system of fl_project's host path (pressure.hip), as device arrays in link and
// cell order, for the coupled viscous pressure solve (coupled.hip).
struct CoupledSystem {
  uint64_t nc, nf;
  const int32_t *tail, *head, *kind, *has_flux, *cell_faces;  // cell_faces: [nc][6] incident links
  const uint64_t* slot;                                       // grid slot of each link, ascending
  const double *area, *scale, *conductance, *ustar, *wall, *lift, *uwall;
  const uint8_t* flag;  // 1 unknown, 2 gauge
  const double *volume, *target, *comp, *prev, *p_init;
  bool p_init_zero;
};
int coupled_prepare(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count, double dt,
                    const fl_pressure_cell* cells, size_t cell_count, const fl_pressure_face* faces,
                    size_t face_count, const fl_pressure_options* options, CoupledSystem* out);
// The same for the typed links of a frozen geometry already on the device (the substep's records), with
// the system and hierarchy built on the device as for project_links_device.
int coupled_prepare_device(fl_ctx* ctx, const fl_pressure_cell* cells, uint64_t nc, const fl_pressure_face* faces,
                           uint64_t nf, double dt, const fl_pressure_options& options, CoupledSystem* out);
// Re-derives the hierarchy's coefficients from the given link conductances (the effective ones of the
// viscous Schur complement's diagonal).
int coupled_coefficients(fl_ctx* ctx, const double* conductance);
// One level-0 V-cycle of the prepared system's hierarchy on the residual r (cell order): z on the
// unknown cells, zero on gauge cells.
int coupled_vcycle(fl_ctx* ctx, const double* r, double* z, const fl_pressure_options& options);
// The pressure error words (residual conversion failures of coupled_vcycle).
int coupled_error(fl_ctx* ctx);

struct CoupledResult {
  uint64_t iterations, velocity_solves, velocity_iterations;
  double residual_l2, initial_residual_l2, max_defect, max_momentum_residual;
};
// SolveCoupled on a prepared system and rows already in the velocity block's arrays (v.rows, face
// ordinals = link ordinals after coupled_rows): the GCR iteration; outputs in p (nc), u (nf) and
// defects (nc), device arrays of coupled.hip's buffers (coupled_outputs).
int coupled_rows(fl_ctx* ctx, const CoupledSystem& S, const ViscosityDevice& v, uint64_t nr, double dt);
int coupled_solve(fl_ctx* ctx, const CoupledSystem& S, const ViscosityDevice& v, uint64_t nr, uint64_t nt, double dt,
                  const fl_pressure_options& options, bool jacobi, CoupledResult* out);
int coupled_outputs(fl_ctx* ctx, const double** p, const double** u, const double** defects);
void release_coupled(fl_ctx* ctx);
// The context's viscosity arrays (capi.hip) sized for n cells, nr rows and nf faces, then for nt packed
// terms; *out views them.
int viscosity_reserve(fl_ctx* ctx, uint64_t n, uint64_t nr, uint64_t nf, ViscosityDevice* out);
int viscosity_reserve_terms(fl_ctx* ctx, uint64_t nt, ViscosityDevice* out);

// The device substep (substep.hip).
int substep(fl_ctx* ctx, const fl_substep_options& o, fl_substep_report& rep);
int substep_download(fl_ctx* ctx, double* positions, double* velocities, double* slopes, double* ratios,
                     uint8_t* flags, uint64_t n);
void release_substep(fl_ctx* ctx);
int project_records_device(fl_ctx* ctx, const fl_pressure_cell* cells, uint64_t nc, const fl_pressure_face* faces,
                           uint64_t nf, double dt, const fl_pressure_options& opt, double* out_pressures,
                           double* out_velocities, fl_pressure_diagnostics* diag);
void substep_scene_uploaded(fl_ctx* ctx);

// Particle update without an obstacle (update.hip): in place on device records, first failure into ctx->err.
int update_particles(fl_ctx* ctx, const fl_particle_update_options& o, const fl_boundary_gather* gathered,
                     const double* distances, uint64_t n, fl_liquid_particle* particles);
// Particle columns of the device substep, in sort order; ordinal (optional) maps a row to the
// particle's ordinal in the step's input order for error words.
struct ParticleColumns {
  double* x[3];
  double* v[3];
  double* c[9];
  double* volume_ratio;
  uint8_t* flags;
  const uint32_t* ordinal;
};
// The update of the columns with GatherMac3D results (24 component-major columns of capacity gcap)
// and one liquid-field value phi for every particle; first failure into ctx->err; the largest updated
// speed's bit pattern into *max_speed (atomic max, optional).
// The open tank runs a clear-particle pass and a contact pass over the particles near a wall, listed
// in list (n entries) with their number in *list_count (device).
int update_columns(fl_ctx* ctx, const fl_particle_update_options& o, uint64_t n, const ParticleColumns& P,
                   const double* gather, uint64_t gcap, double phi, unsigned long long* max_speed, uint32_t* list,
                   uint32_t* list_count);

// Small helpers (topology.hip)
// Variational MAC viscosity (viscosity.hip): device arrays of one strain assembly or velocity-block solve.
struct ViscosityDevice {
  const fl_viscosity_cell* cells;
  fl_strain_row* rows;
  uint64_t row_capacity;
  uint64_t* counts;   // per cell shear rows, then per face incidence counts widened for the scan
  uint64_t* offsets;  // exclusive scan of the shear counts
  const double* inertia;
  const double* rhs;
  double *v, *r, *z, *d, *q, *diag, *strain, *partial, *scal;
  uint64_t partial_capacity;
  uint32_t* flag;  // 2 words
  uint32_t *inc_count, *cursor;
  uint64_t* row_nnz;  // nr: nonzero weighted terms per row
  uint64_t* row_ptr;  // nr + 1: first packed term of each row, then the term count
  uint64_t* inc_ptr;  // nf + 1: first incidence of each face, then the incidence count
  // Packed operator, term_capacity entries each: the rows' nonzero weighted terms in row and term order
  // (face slot, coefficient) and each face's incidences in row order (row, coefficient); entries holds
  // the unsorted (row << 3 | term) keys while the incidence is built.
  uint64_t term_capacity;
  uint32_t *term_slot, *inc_row;
  double *term_coef, *inc_coef;
  uint64_t* entries;
};
// AssembleFullCellMacStrain3D of n cells on the grid of the context's lookup; *total rows written.
int assemble_strain(fl_ctx* ctx, const ViscosityDevice& v, uint64_t n, double h, uint64_t* total);
// First half of SolveMacViscousVelocityBlock3D after the host's input validation: checks and weights
// the rows in place (sqrt(dt) sqrt(w) c, zero on eliminated faces) and scans the nonzero terms per row
// and per face; *terms is the packed operator's size, for which the caller sizes the term arrays.
int weight_viscous_rows(fl_ctx* ctx, const ViscosityDevice& v, uint64_t nr, uint64_t nf, double dt, uint64_t* terms);
// Second half: packs the operator, forms the Jacobi diagonal and runs the PCG (max_rhs from the host's
// validation).
int solve_viscous_block(fl_ctx* ctx, const ViscosityDevice& v, uint64_t nr, uint64_t nf, uint64_t nt, double dt,
                        double max_rhs,
                        uint64_t max_iterations, double relative_tolerance, uint64_t* iterations,
                        double* max_residual);
// The two halves of solve_viscous_block: pack the weighted rows and the face incidence and form the
// Jacobi diagonal (v.diag); then the PCG from zero for the right-hand side in v.rhs (several solves may
// follow one pack).
int pack_viscous_block(fl_ctx* ctx, const ViscosityDevice& v, uint64_t nr, uint64_t nf, uint64_t nt, double dt);
int pcg_viscous_block(fl_ctx* ctx, const ViscosityDevice& v, uint64_t nr, uint64_t nf, double max_rhs,
                      uint64_t max_iterations, double relative_tolerance, uint64_t* iterations,
                      double* max_residual);
int read_error(fl_ctx* ctx, uint64_t* word);
int reset_error(fl_ctx* ctx);
int fill_i64(fl_ctx* ctx, int64_t* p, uint64_t n, int64_t v);
int fill_i32(fl_ctx* ctx, int32_t* p, uint64_t n, int32_t v);
int exclusive_scan_u64(fl_ctx* ctx, const uint64_t* in, uint64_t* out, uint64_t n);

}  // namespace fl
