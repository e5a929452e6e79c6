/* C-ABI of the MI355X-native FluidiTy liquid-step hot path (P2G, G2P, MAC pressure MGPCG).
 *
 * One library, fluidity_amd/libfluidity_hip.so, built for gfx950. Every entry point returns an
 * int status whose values equal absl::StatusCode (0 OK, 3 InvalidArgument, 8 ResourceExhausted,
 * 9 FailedPrecondition, 10 Aborted, 11 OutOfRange, 13 Internal), as the reference operators do
 * (native/include/fluidity/host/transfers_3d.h:17-22, mac_pressure_3d.h:95-100,
 * native/cuda/operator_runtime.h:18-40). On any error, caller outputs are left untouched.
 *
 * Records mirror the reference PODs byte for byte: Vec3, Mat3, BlockIndex3D
 * (native/include/fluidity/base/mac_types_3d.h:9-72), MacTransferParticle3D,
 * MacDepositQuantization3D, MacFaceDepositQuanta3D, MacParticleGatherResult3D
 * (native/include/fluidity/base/mac_transfer_types_3d.h:7-37), LiquidParticleSortEntry3D
 * (native/include/fluidity/host/liquid_step_workspace_3d.h:14-17), and MacPressureCell3D /
 * MacPressureFace3D with the face's std::variant link and std::optional wall flux flattened
 * (native/include/fluidity/host/mac_pressure_3d.h:14-51).
 *
 * Threading: a context is single-threaded and non-reentrant, like the reference workspaces
 * (native/include/fluidity/host/mac_pressure_workspace_3d.h:65). Device memory: fl_create allocates
 * the particle, grid and face arrays and the transfer calls allocate nothing; the projection allocates
 * its arrays on first use and its multigrid levels when the system's block structure changes;
 * fl_learn_setup allocates the network.
 */
#ifndef FLUIDITY_HIP_H_
#define FLUIDITY_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FL_OK = 0,
  FL_INVALID_ARGUMENT = 3,
  FL_RESOURCE_EXHAUSTED = 8,
  FL_FAILED_PRECONDITION = 9,
  FL_ABORTED = 10,
  FL_OUT_OF_RANGE = 11,
  FL_UNIMPLEMENTED = 12,
  FL_INTERNAL = 13,
};

typedef struct fl_vec3 { double x, y, z; } fl_vec3;
typedef struct fl_mat3 { fl_vec3 x_row, y_row, z_row; } fl_mat3;
typedef struct fl_int3 { int32_t x, y, z; } fl_int3;
typedef struct fl_geometry { fl_vec3 world_origin; double cell_width; } fl_geometry;
typedef struct fl_particle {
  fl_vec3 sorted_world_position;
  fl_vec3 trial_world_position;
  double particle_mass;
  fl_vec3 particle_velocity;
  fl_mat3 particle_velocity_slopes;
} fl_particle;
typedef struct fl_quantization {
  double mass_per_quantum;
  double pure_momentum_per_quantum;
  uint64_t declared_particle_capacity;
} fl_quantization;
typedef struct fl_face_deposit {
  double deposited_mass;
  double deposited_pure_momentum;
} fl_face_deposit;
typedef struct fl_face_quanta {
  int64_t mass_quanta_sum;
  int64_t pure_momentum_quanta_sum;
  int64_t stress_force_quanta_sum;
} fl_face_quanta;
typedef struct fl_gather_result {
  fl_vec3 new_pic_velocity;
  fl_vec3 grid_velocity_increment;
  fl_mat3 fitted_velocity_slopes;
  fl_mat3 velocity_gradient;
} fl_gather_result;
typedef struct fl_sort_entry { fl_int3 bin_index; uint32_t pad; uint64_t original_ordinal; } fl_sort_entry;
/* LiquidBoundaryParticle3D, LiquidBoundaryFaceTag3D, CompatibleMacFaceSample3D
 * (native/include/fluidity/base/liquid_boundary_transfer_types_3d.h:4-20). */
typedef struct fl_boundary_particle {
  fl_particle transfer;
  int64_t side_tag;
  uint8_t affine_enabled;
  uint8_t pad[7];
} fl_boundary_particle; /* 168 B */
typedef struct fl_face_tag {
  int64_t side_tag;
  uint8_t structurally_eligible;
  uint8_t pad[7];
} fl_face_tag; /* 16 B */
typedef struct fl_compatible_sample {
  uint64_t face_slot;
  uint8_t compatible;
  uint8_t pad[7];
  double normalized_weight;
  fl_vec3 normalized_weight_gradient;
  fl_vec3 centered_face_offset;
} fl_compatible_sample; /* 72 B */
/* LiquidPrescribedWallSample3D and LiquidBoundaryParticleGather3D
 * (native/include/fluidity/base/liquid_boundary_transfer_types_3d.h:21-36) and the flattened spans of
 * LiquidBoundaryVelocityHistory3D (native/include/fluidity/host/liquid_boundary_transfer_3d.h:83-90).
 * Face validity: 0 missing, 1 valid, 2 predicted, 3 low mass. */
typedef struct fl_wall_sample {
  uint8_t available;
  uint8_t pad[7];
  int64_t side_tag;
  fl_vec3 unit_normal;
  fl_vec3 wall_velocity;
} fl_wall_sample; /* 64 B */
typedef struct fl_velocity_history {
  double velocity_sample_time;
  double wall_sample_time;
  const double* face_velocities;
  const uint8_t* validity;
  const fl_face_tag* face_tags;
  const fl_wall_sample* wall_samples;
  size_t count; /* entries of each array: one per owned slot */
} fl_velocity_history;
typedef struct fl_boundary_gather {
  fl_vec3 new_pic_velocity;
  fl_vec3 grid_velocity_increment;
  fl_vec3 particle_minus_old_pic_velocity;
  fl_mat3 fitted_velocity_slopes;
  fl_mat3 velocity_gradient;
  uint8_t affine_enabled, old_pic_fallback, new_pic_fallback, low_mass_fallback;
  uint8_t pad[4];
} fl_boundary_gather; /* 224 B */

/* link_kind: 0 sealed, 1 internal (lift = pressure jump), 2 Dirichlet (lift = exterior pressure). */
typedef struct fl_pressure_cell {
  fl_int3 cell_index;
  int32_t pad;
  double fluid_volume;
  double target_divergence;
  double bulk_modulus; /* +inf selects zero compliance */
  double previous_pressure;
  double integrated_inward_wall_flux;
} fl_pressure_cell;
/* Frozen boundary geometry and pressure-field records (native/include/fluidity/host/liquid_boundary_geometry_3d.h):
 * LiquidCenterDistance3D, FrozenLiquidCell3D, FrozenLiquidFace3D (its std::optional wall flux as value +
 * flag, the libstdc++ layout), LiquidPressureCellFields3D (optional density likewise),
 * LiquidAirCellPressure3D, LiquidInternalInterface3D and an optional wetted-piston flux. */
typedef struct fl_center_distance {
  fl_int3 cell_index;
  uint32_t pad;
  double signed_distance;
} fl_center_distance; /* 24 B */
typedef struct fl_frozen_cell {
  fl_int3 cell_index;
  uint32_t pad;
  double fluid_volume;
  double center_distance;
  double curvature;
} fl_frozen_cell; /* 40 B */
typedef struct fl_frozen_face {
  uint8_t axis;
  uint8_t pad0[3];
  fl_int3 lattice;
  uint64_t face_slot;
  double open_area;
  double blocked_area;
  double wall_normal_velocity;
  double prescribed_wall_flux;
  uint8_t has_prescribed_wall_flux;
  uint8_t pad1[7];
  uint8_t is_low_x_piston_face;
  uint8_t pad2[7];
} fl_frozen_face; /* 72 B */
typedef struct fl_cell_fields {
  double target_divergence;
  double bulk_modulus;
  double previous_pressure;
  double density;
  uint8_t has_density;
  uint8_t pad[7];
} fl_cell_fields; /* 40 B */
typedef struct fl_air_pressure {
  fl_int3 cell_index;
  uint32_t pad;
  double gas_gauge_pressure;
} fl_air_pressure; /* 24 B */
typedef struct fl_internal_interface {
  uint64_t face_ordinal;
  double tail_fraction;
  double surface_tension;
  double curvature_tail_to_head;
  double normal_viscous_jump_tail_to_head;
} fl_internal_interface; /* 40 B */
typedef struct fl_optional_flux {
  double value;
  uint8_t has_value;
  uint8_t pad[7];
} fl_optional_flux; /* 16 B */
/* FrozenLiquidBoundaryGeometryView3D (native/include/fluidity/host/liquid_boundary_workspace_3d.h:34-43):
 * the grid as its sorted block keys and geometry. */
typedef struct fl_frozen_view {
  const fl_int3* blocks;
  uint64_t block_count;
  fl_geometry geometry;
  const fl_center_distance* center_distances;
  uint64_t center_count;
  const fl_frozen_cell* cells;
  uint64_t cell_count;
  const fl_frozen_face* faces;
  uint64_t face_count;
  double step_duration;
  double liquid_density;
  double surface_tension;
  double capillary_step_bound;
} fl_frozen_view;
/* The frozen-geometry obstacles (native/include/fluidity/host/liquid_boundary_geometry_3d.h:30-43): absent,
 * TranslatingHalfspace3D (solid where unit_normal . x <= coordinate) and TranslatingSlab3D (solid where
 * |unit_normal . x - coordinate| <= half_width). */
enum { FL_OBSTACLE_ABSENT = 0, FL_OBSTACLE_HALFSPACE = 1, FL_OBSTACLE_SLAB = 2, FL_OBSTACLE_TANK = 3 };
/* AxisAlignedLiquidTank3D (native/include/fluidity/base/liquid_boundary_math.h:22-29): walls at the
 * piston x (moving at piston_velocity_x), high x, the floor and low / high z; the top is open. */
typedef struct fl_tank {
  double start_low_x;
  double high_x;
  double floor_y;
  double low_z;
  double high_z;
  double piston_velocity_x;
} fl_tank; /* 48 B */
typedef struct fl_obstacle {
  uint32_t kind;
  uint32_t pad;
  fl_vec3 unit_normal;
  double coordinate;  /* half-space: start plane coordinate; slab: start centre coordinate */
  double half_width;  /* slab only */
  fl_vec3 wall_velocity;
  fl_tank tank;       /* open tank only */
} fl_obstacle; /* 120 B */
/* SolidCornerDistance3D (liquid_boundary_geometry_3d.h:77-81): corner key, signed distance in metres
 * (positive outside the solid) and the wall velocity used when no dense face velocities are given. 48 B. */
typedef struct fl_solid_corner {
  fl_int3 corner_index;
  uint32_t pad;
  double signed_distance;
  fl_vec3 wall_velocity;
} fl_solid_corner;
/* LiquidBoundaryGeometryViewInputs3D (native/include/fluidity/host/liquid_boundary_workspace_3d.h:18-30);
 * corner samples (LiquidCornerDistance3D) use the centre-sample record. max_cells and max_faces give the
 * reference's capacity refusals at its points (0: none). */
typedef struct fl_freeze_inputs {
  const fl_int3* blocks;
  uint64_t block_count;
  fl_geometry geometry;
  const fl_int3* cell_indices;
  uint64_t cell_count;
  const fl_center_distance* center_distances;
  uint64_t center_count;
  const fl_center_distance* corner_distances;
  uint64_t corner_count;
  fl_obstacle obstacle;
  double step_duration;
  double liquid_density;
  double surface_tension;
  double exact_curvature;
  uint8_t has_exact_curvature;
  uint8_t pad[7];
  uint64_t max_cells;
  uint64_t max_faces;
  /* SampledSolidBoundary3D (liquid_boundary_geometry_3d.h:83-91), used when has_sampled_solid is 1 or 2: solid
   * corner samples, distinct and (z,y,x) sorted, covering every candidate's and frozen face's corners;
   * and either no face velocities (count 0: the corner wall velocities are integrated over the blocked
   * part of each face) or one wall normal velocity per owned slot (192 per block). The obstacle must then
   * be absent or a half-space (which clips the sampled geometry). Zero-initialised callers keep the
   * plane-obstacle freeze. has_sampled_solid 2 and 3 have no reference counterpart: 2 applies the
   * wet-piece connectivity refusal (CheckSampledWetConnectivity) only to cells with a corner in the solid
   * (sampled distance <= 0 or on the analytic wall's side), so liquid split by the particle field alone
   * is measured as in the plane-obstacle freeze instead of refused; 3 applies it nowhere (a cell's wet
   * pieces are one pressure cell whether or not a solid separates them). */
  const fl_solid_corner* solid_corners;
  uint64_t solid_corner_count;
  const double* face_normal_velocities;
  uint64_t face_velocity_count;
  uint8_t has_sampled_solid;
  uint8_t pad2[7];
} fl_freeze_inputs;
/* LiquidPressureFields3D (liquid_boundary_geometry_3d.h:170-185); an empty span has count 0. */
typedef struct fl_pressure_fields {
  const fl_cell_fields* cells;
  uint64_t cell_count;
  const double* predicted_face_velocities;
  uint64_t predicted_count;
  double ambient_pressure;
  const fl_optional_flux* wetted_piston_face_fluxes;
  uint64_t wetted_count;
  const fl_air_pressure* air_cell_pressures;
  uint64_t air_count;
  const fl_internal_interface* internal_interfaces;
  uint64_t interface_count;
} fl_pressure_fields;
/* LiquidParticleState3D (native/include/fluidity/host/liquid_state_3d.h:100-120) with its eight
 * LiquidConstituentFraction3D slots (liquid_mixture_3d.h:33-37). */
typedef struct fl_constituent_fraction {
  uint64_t substance_id;
  double volume_fraction;
} fl_constituent_fraction;
typedef struct fl_liquid_particle {
  uint64_t id;
  uint32_t tag;
  uint32_t pad0;
  fl_vec3 world_position;
  fl_vec3 world_velocity;
  fl_mat3 velocity_slopes;
  uint8_t affine_enabled;
  uint8_t pad1[7];
  double mass;
  double reference_volume;
  double volume_ratio;
  int64_t side_tag;
  uint8_t pic_initialized;
  uint8_t pad2[7];
  uint64_t phase_id;
  uint8_t constituent_count;
  uint8_t quadratic_adaptive_transfer;
  uint8_t pad3[6];
  fl_constituent_fraction constituents[8];
} fl_liquid_particle; /* 328 B */
/* The scalars of LiquidParticleUpdateInputs3D (native/include/fluidity/host/liquid_substep_backend_3d.h:16-28). */
typedef struct fl_particle_update_options {
  double cell_width;
  double step_duration;
  double particle_energy_blend;
  double friction_coefficient;
  double penetration_tolerance;
  fl_obstacle obstacle; /* absent (all zero), half-space, slab or open tank */
} fl_particle_update_options; /* 160 B */
/* Topology records of a liquid state (native/include/fluidity/host/liquid_state_3d.h:129-159):
 * LiquidBlockState3D, LiquidCellState3D (occupancy 0 air, 1 liquid, 2 solid) and LiquidFaceState3D
 * (axis 0..2 with the lattice index at offset 4; validity 0 missing, 1 valid, 2 predicted, 3 low mass). */
typedef struct fl_block_state {
  fl_int3 block_index;
  uint32_t pad;
  uint64_t generation;
  uint32_t activation_age;
  uint32_t inactive_age;
} fl_block_state; /* 32 B */
typedef struct fl_cell_state {
  fl_int3 cell_index;
  uint8_t occupancy;
  uint8_t pad[3];
  double previous_pressure;
} fl_cell_state; /* 24 B */
typedef struct fl_face_state {
  uint8_t axis;
  uint8_t pad0[3];
  fl_int3 lattice;
  uint64_t face_slot;
  double carrier_velocity;
  uint8_t validity;
  uint8_t pad1[7];
  int64_t side_tag;
  uint8_t structurally_eligible;
  uint8_t pad2[7];
} fl_face_state; /* 56 B */
/* The topology of a previous LiquidState3D: its geometry, clock epoch and block, cell and face records. */
typedef struct fl_topology_history {
  fl_geometry geometry;
  uint64_t epoch;
  const fl_block_state* blocks;
  uint64_t block_count;
  const fl_cell_state* cells;
  uint64_t cell_count;
  const fl_face_state* faces;
  uint64_t face_count;
} fl_topology_history;
/* LiquidStateCapacities3D without sources. */
typedef struct fl_topology_capacities {
  uint64_t particles, blocks, cells, faces;
} fl_topology_capacities;
typedef struct fl_pressure_face {
  int32_t axis;
  fl_int3 lattice;
  double open_area;
  double inverse_density;
  double pressure_gradient_distance;
  double predicted_normal_velocity;
  double wall_normal_velocity;
  int32_t link_kind;
  int32_t has_prescribed_wall_flux;
  double lift;
  double prescribed_wall_flux;
} fl_pressure_face;

typedef struct fl_pressure_options {
  uint64_t max_iterations;             /* PCG budget; the reference range is [1, 256] */
  double cpu_residual_tolerance_scale; /* scales both reference gates, in [1e-12, 1] */
  double relative_divergence_target;   /* 0 disables; else also ||defect||2 <= target ||defect(p=0)||2 */
  int32_t levels;                      /* geometric multigrid levels, >= 1 */
  int32_t smoothing_sweeps;            /* red-black Gauss-Seidel sweeps before and after the coarse correction */
  int32_t coarse_sweeps;               /* symmetric sweeps on the coarsest level */
  int32_t level_storage;               /* V-cycle coefficients: 1 FP32, 2 BF16 (FP32 accumulation); CG vectors
                                          are FP32, dots and residuals FP64 either way */
  double coarse_scale;                 /* coarse operator = scale * Galerkin crossing sums */
  int32_t preconditioner;              /* 0 multigrid V-cycle (PCG); 1 learned network (flexible CG), falling back
                                          to the V-cycle after 3 stagnating iterations (needs fl_learn_setup) */
  int32_t smoother;                    /* 0 red-black Gauss-Seidel half sweeps over the whole level; 1 block
                                          smoother: red-black sweeps inside each 4^3 block with the block's
                                          neighbours held fixed (block Jacobi between blocks); 2 level-wide
                                          sweeps on the finest level, the block smoother on coarser ones */
} fl_pressure_options;

typedef struct fl_pressure_diagnostics {
  uint64_t iteration_count;
  uint64_t multigrid_level_count;
  double true_residual_l2;
  double max_volume_scaled_residual;
  double max_cell_constraint_defect;
  double rhs_l2;
  double initial_defect_l2;
  double defect_l2;
  double defect_linf;
  uint64_t restarts;
  uint64_t gauge_count;
  uint64_t learned_iterations; /* iterations preconditioned by the learned network */
} fl_pressure_diagnostics;

/* Device time of the call's kernels (hipEvents on the context stream), like the reference's
 * CudaOperatorMeasurement (native/include/fluidity/host/cuda_measurement.h:5-10). */
typedef struct fl_measurement {
  double kernel_ms;   /* kernels only */
  double copy_ms;     /* host<->device staging of the span-based call */
  uint32_t launches;
  char architecture[28];
} fl_measurement;

typedef struct fl_capacities {
  uint64_t particles;       /* per call */
  uint64_t blocks;          /* active 4^3 blocks */
  uint64_t lookup_blocks;   /* bounding-box volume, in blocks, of the dense block and bin tables */
  uint64_t pressure_cells;
  uint64_t pressure_faces;
} fl_capacities;

typedef struct fl_ctx fl_ctx;

/* ---- lifecycle (CudaOperatorRuntime, native/cuda/operator_runtime.h:45-213) ---- */
int fl_create(const fl_capacities* capacities, int device, fl_ctx** out);
void fl_destroy(fl_ctx* ctx);
/* Architecture string of the device the context runs on (gcnArchName). */
int fl_device_name(fl_ctx* ctx, char* out, size_t size);
/* Set-up outside physical attempts (liquid_substep_backend_3d.h:37-46: workspaces allocate only at
 * Create; attempt operations allocate and free nothing). fl_create itself reserves the projection's
 * arrays and its multigrid level pool from the pressure capacities. fl_reserve also reserves, from the
 * capacities, the device buffers that topology preparation, the boundary gather, the frozen geometry,
 * the pressure links and the particle update would otherwise allocate on first use: after it, every
 * call whose inputs stay within the capacities (at most 27 centre and 8 corner samples per candidate
 * cell, link fields at most one per cell or pressure face) allocates and frees no device memory. */
int fl_reserve(fl_ctx* ctx);

/* ---- variational MAC viscosity (native/src/viscosity.cc; host/viscosity.h:47-160) ---- */
/* MacViscosityCell3D, MacStrainTerm3D and MacStrainRow3D (axes 0/1/2). */
typedef struct fl_viscosity_cell {
  fl_int3 cell_index;
  uint32_t pad;
  double dynamic_viscosity; /* Pa s, finite and nonnegative */
} fl_viscosity_cell;       /* 24 B */
typedef struct fl_strain_term {
  uint64_t face_slot;
  double coefficient; /* 1/m */
} fl_strain_term;
typedef struct fl_strain_row {
  fl_int3 low_cell;
  uint8_t first_axis, second_axis, pad[2];
  fl_strain_term terms[8];
  uint64_t term_count;
  double quadrature_weight; /* kg m^2/s */
  double prescribed_strain; /* 1/s */
} fl_strain_row;            /* 168 B */
/* AssembleFullCellMacStrain3D (native/include/fluidity/host/viscosity.h:77-78; viscosity.cc:762-946):
 * three normal rows per cell in cell and axis order, then the interior shear rows (cell, +a, +b and the
 * diagonal cell present) in cell and axis-pair order xy, xz, yz, with harmonic edge viscosity. Cells
 * strictly (z,y,x) sorted. *io_count holds the capacity of out on entry (6 * n always suffices) and the
 * row count on return. Statuses and their order as the reference; out is written only on success. */
int fl_assemble_full_cell_strain(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
                                 const fl_viscosity_cell* cells, size_t cell_count, fl_strain_row* out,
                                 size_t* io_count, fl_measurement* measurement);
/* SolveMacViscousVelocityBlock3D (native/include/fluidity/host/viscosity.h:128-156; viscosity.cc:1331-1650):
 * (M + dt E^T W E) u = rhs by Jacobi PCG in FP64 on the faces of positive inertia (eliminated faces stay
 * zero), stopping on the true residual max|rhs - Hu| <= relative_tolerance max|rhs|. InvalidArgument for
 * malformed inputs or options, OutOfRange for unrepresentable arithmetic, Aborted for a non-positive
 * curvature or an exhausted budget. Outputs are written only on success. */
int fl_solve_viscous_block(fl_ctx* ctx, const fl_strain_row* rows, size_t row_count, const double* face_inertias,
                           size_t face_count, double step_duration, const double* rhs, uint64_t max_iterations,
                           double relative_tolerance, double* out_velocity, uint64_t* out_iterations,
                           double* out_max_residual, fl_measurement* measurement);

/* ProjectMacViscousPressure3D (native/include/fluidity/host/mac_pressure_3d.h; mac_viscous_pressure_3d.cc:
 * 151-459, 558-575, 699-705): the coupled viscous pressure solve of fl_project's cells and faces with the
 * given strain rows (face_slot: grid slots of the faces; AssembleFullCellMacStrain3D's rows, for example).
 * Each pressure iterate's face velocities solve (M + dt E^T W E) u = M u* + dt A (lift + p_tail - p_head)
 * by Jacobi PCG to 1e-10 of the largest right-hand side; the pressure iteration is the reference's GCR
 * (restart 32, budget options->max_iterations in [1, 256]) accepted at |defect| <= 2e-5 per cell and the
 * residual L2 gate. options->preconditioner 0 preconditions the GCR with one V-cycle of the pressure
 * hierarchy of the same links (levels, sweeps, smoother from the options; FP32 level storage); 2 uses the
 * reference's Jacobi diagonal of the Schur complement. Outputs: pressures (cell_count), accepted face
 * velocities (face_count; wall normal velocities on closed faces), cell constraint defects (cell_count).
 * Statuses: the reference's input checks; Aborted when the budget is spent. */
typedef struct fl_viscous_diagnostics {
  uint64_t pressure_iterations;
  double constraint_residual_l2;
  double initial_constraint_residual_l2;
  double max_cell_constraint_defect;
  double max_momentum_residual;
  double inward_prescribed_wall_flux;
  uint64_t velocity_solves;     /* velocity block solves */
  uint64_t velocity_iterations; /* their PCG iterations in total */
} fl_viscous_diagnostics;        /* 64 B */
int fl_project_viscous(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
                       double step_duration, const fl_pressure_cell* cells, size_t cell_count,
                       const fl_pressure_face* faces, size_t face_count, const fl_strain_row* rows, size_t row_count,
                       const fl_pressure_options* options, double* out_pressures, double* out_velocities,
                       double* out_defects, fl_viscous_diagnostics* out_diagnostics, fl_measurement* measurement);

/* ---- topology ---- */
/* ComputeParticleSortBin3D for each position plus the stable (z,y,x)-bin / ordinal order of
 * LiquidStepWorkspace3D::PrepareTopology (native/src/sparse_mac_grid_3d.cc:438-449;
 * native/src/liquid_step_workspace_3d.cc:208-268). out has n entries in sorted order. */
int fl_bin_sort(fl_ctx* ctx, const fl_geometry* geometry, const fl_vec3* positions, size_t n, fl_sort_entry* out,
                fl_measurement* measurement);
/* Block activation: owners of all 81 stencil faces of every particle and of the six faces of every
 * pressure cell, sorted (z,y,x) and unique (native/src/liquid_step_workspace_3d.cc:208-244, :272-276).
 * *io_count holds the capacity of out_blocks on entry and the block count on return. */
int fl_activate(fl_ctx* ctx, const fl_geometry* geometry, const fl_vec3* positions, size_t n,
                const fl_int3* pressure_cells, size_t cell_count, fl_int3* out_blocks, size_t* io_count,
                fl_measurement* measurement);

/* LiquidStepWorkspace3D::PrepareTopology without an imported order (native/src/liquid_step_workspace_3d.cc:171-341):
 * the sorted particle order (n entries), the active blocks (a new block takes generation
 * previous->epoch + 1, or 1; a block of the previous state keeps its record with inactive_age 0), the
 * canonical unique pressure cells (a previous cell keeps its record, a new one is liquid at zero
 * pressure) and one face record per owned slot (the previous record of the same face with the new
 * slot, or a missing face with side tag 0 that is structurally eligible). capacities are the
 * workspace's; out_blocks, out_cells and out_faces hold capacities->blocks, ->cells and ->faces
 * records. previous may be null; its records get the checks of ValidateLiquidState3D for blocks,
 * cells and faces (native/src/liquid_state_3d.cc:559-604), the rest of that state being the caller's
 * to validate. Outputs are written only on success. */
int fl_prepare_topology(fl_ctx* ctx, const fl_geometry* geometry, const fl_topology_capacities* capacities,
                        const fl_vec3* positions, size_t n, const fl_int3* pressure_cells, size_t cell_count,
                        const fl_topology_history* previous, fl_sort_entry* out_order, fl_block_state* out_blocks,
                        uint64_t* out_block_count, fl_cell_state* out_cells, uint64_t* out_cell_count,
                        fl_face_state* out_faces, uint64_t* out_face_count, fl_measurement* measurement);

/* ---- pressure links ---- */
/* BuildLiquidPressureLinks3D on a frozen view (native/src/liquid_boundary_geometry_3d.cc: ValidateFrozenView
 * :1289-1383, then LinksCore :1077-1287): one pressure cell per frozen cell (fluid volume and the cell's
 * fields) and one typed link per frozen face, in frozen order — internal links between two liquid cells
 * (zero jump, or an interface crossing's fma jump with the lerp density after its capillary bound),
 * free-surface Dirichlet links at distance clamp(theta, 0.05, 1) h with theta from the liquid and air
 * centre distances and exterior pressure gas + sigma * curvature
 * (base/liquid_boundary_geometry_math_3d.h:676-689), and sealed faces; the mobility from the cells'
 * densities when present. *out_lower_theta_count receives the number of theta values raised to 0.05.
 * The first failure of the reference's sequential checks decides the status; outputs are written only on
 * success. The reference's bounded workspace caps (4096 cells, 24576 faces) are not applied. */
int fl_build_pressure_links(fl_ctx* ctx, const fl_frozen_view* view, const fl_pressure_fields* fields,
                            fl_pressure_cell* out_cells, fl_pressure_face* out_faces, uint64_t* out_lower_theta_count,
                            fl_measurement* measurement);

/* ---- frozen boundary geometry ---- */
/* FreezeCore (native/src/liquid_boundary_geometry_3d.cc:765-1068) for the absent, half-space, slab and
 * open-tank obstacles and a sampled solid: per candidate cell (strictly sorted) the liquid volume from six
 * body-diagonal tetrahedra clipped by the piecewise-linear corner distances and by the obstacle planes,
 * the tank walls or the sampled solid's piecewise-linear surface; liquid cells (centre distance < 0,
 * volume > 0) become frozen cells in candidate order, their faces frozen faces sorted by slot with open /
 * blocked areas and the wall velocity (the tank's piston flux on its low-x face, the sampled solid's
 * integrated or dense wall velocities); cells on an open face to air get the
 * exact curvature or the centred one of the 19-point centre halo; the capillary step bound
 * sqrt(rho h^3 / (4 pi sigma)) when such a cell exists (infinity otherwise), with FailedPrecondition when
 * the step exceeds it. out_cells holds cell_count records, out_faces 6 * cell_count. The first failure of
 * the reference's sequential checks decides the status; outputs are written only on success. The frozen
 * records and centre samples form the fl_frozen_view that fl_build_pressure_links takes. */
int fl_freeze_geometry(fl_ctx* ctx, const fl_freeze_inputs* inputs, fl_frozen_cell* out_cells, uint64_t* out_cell_count,
                       fl_frozen_face* out_faces, uint64_t* out_face_count, double* out_capillary_step_bound,
                       fl_measurement* measurement);

/* ---- particle update ---- */
/* The liquid particle update of the CPU substep backend without regular-cell transport
 * (native/src/liquid_substep_3d_cpu.cc:239-405) for the absent, half-space and slab obstacles: per
 * particle, alpha = 0.99 e when PIC-initialised and no gather fallback, stored velocity = new PIC +
 * alpha (v - old PIC), advecting velocity blended the same way only under surface expansion (div u > 0,
 * -h <= liquid distance <= 0 and a solid distance above h), then ResolveContact
 * (native/include/fluidity/base/liquid_wall_contact_math_3d.h:365-556): fma(advecting, dt, x) without an
 * obstacle, else the certified swept-plane contact (first hit, one-way normal impulse with Coulomb
 * friction on both velocities, the constrained remainder of the step); a displacement above h/2 Aborted,
 * J *= exp(dt div u) (Aborted unless finite and positive), slopes and the affine switch from the gather.
 * gathered and liquid_distances are in particle order. The first failing particle's status is returned
 * (InvalidArgument for non-finite inputs or invalid options or obstacle, FailedPrecondition for deep or
 * ambiguous penetration, OutOfRange for uncertified contact or a non-finite position, Aborted as above);
 * particles are written only on success. J uses the device exp, which may differ from the host's in the
 * last place. */
int fl_update_particles(fl_ctx* ctx, const fl_particle_update_options* options, const fl_boundary_gather* gathered,
                        const double* liquid_distances, size_t n, fl_liquid_particle* particles,
                        fl_measurement* measurement);

/* ---- transfers ---- */
/* DepositMac3D (native/include/fluidity/host/transfers_3d.h:24-28; native/src/transfers_3d.cc:139-145):
 * FP64 mass and momentum sums per owned slot. The sums are formed in parallel, so they match the
 * reference's particle-order sums to rounding (1e-12 relative), not bit for bit. */
int fl_p2g(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
           const fl_particle* particles, size_t n, fl_face_deposit* out, size_t face_count,
           fl_measurement* measurement);
/* DepositMacQuanta3D (native/include/fluidity/host/transfers_3d.h:36-41; native/src/transfers_3d.cc:147-197).
 * blocks: canonical sorted active blocks (SparseMacGrid3D); out has 192 * block_count records in
 * owned-slot order. The reference's 2^20 particle cap is replaced by the context capacity. */
int fl_p2g_quanta(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
                  const fl_particle* particles, size_t n, fl_quantization quantization, fl_face_quanta* out,
                  size_t face_count, fl_measurement* measurement);
/* DepositCompatibleMacQuanta3D (native/include/fluidity/host/liquid_boundary_transfer_3d.h:34-48;
 * native/src/liquid_boundary_transfer_3d.cc:329-460): the quantized deposit restricted per face axis to
 * the samples whose slot is structurally eligible and carries the particle's side tag, with weights
 * normalised over those samples, the quotient-rule gradient and recentred affine offsets. tags has one
 * entry per owned slot. stencils: null, or n * 3 * 27 records (particle, component, (z,y,x) sample
 * order). The reference's 1024-particle / 24576-face caps are replaced by the context capacities. */
int fl_p2g_compatible(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
                      const fl_boundary_particle* particles, size_t n, const fl_face_tag* tags, size_t tag_count,
                      fl_quantization quantization, fl_face_quanta* out, size_t face_count,
                      fl_compatible_sample* stencils, fl_measurement* measurement);
/* GatherLiquidBoundaryParticles3D (native/include/fluidity/host/liquid_boundary_transfer_3d.h:100-140;
 * native/src/liquid_boundary_transfer_3d.cc:519-651): the gather on the face velocities of two histories
 * in which every stencil sample that is not structurally eligible on the particle's side, or whose
 * validity is neither valid nor predicted, is continued from same-side donors within two cell widths
 * (rank-revealing affine fit, nearest donor when rank-deficient, a compatible prescribed wall as the
 * last resort, and the wall-normal replacement for oblique walls); particle retention, affine switch
 * and fallback flags as the reference. The histories share their samples (one continuation serves
 * both) when all four of their array pointers are equal. out_counts: old-PIC, new-PIC and low-mass
 * fallback particle counts (may be null). */
int fl_g2p_boundary(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
                    const fl_boundary_particle* particles, size_t n, double old_time, double new_time,
                    const fl_velocity_history* old_history, const fl_velocity_history* new_history,
                    fl_boundary_gather* out, uint64_t* out_counts, fl_measurement* measurement);
/* NormalizeLiquidMacFaces3D followed by ApplyForces (native/src/liquid_boundary_transfer_3d.cc:463-516;
 * native/src/liquid_substep_3d_cpu.cc:156-178). validity: 0 missing, 1 valid, 3 low mass.
 * Any output pointer may be null. */
int fl_normalize_forces(fl_ctx* ctx, const fl_face_quanta* quanta, size_t face_count, fl_quantization quantization,
                        double minimum_face_mass, fl_vec3 acceleration, double dt, double* out_mass,
                        double* out_velocity, uint8_t* out_validity, double* out_predicted,
                        fl_measurement* measurement);
/* ApplyForces alone (native/src/liquid_substep_3d_cpu.cc:156-178): predicted = old + dt * g[axis], the
 * axis of owned slot f being (f mod 192) / 64. FL_OUT_OF_RANGE for a non-finite prediction. */
/* No counterpart in the reference: fl_normalize_forces for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It takes no
 * context and allocates and frees nothing: its device memory is a file-scope __device__ array of
 * FL_STATIC_SCRATCH_BYTES on each GPU it runs on, held until the process ends. Same arguments, checks,
 * kernel, results and statuses as fl_normalize_forces, with `device` in place of the context and at
 * most FL_STATIC_MAX_FACES faces (more is FL_RESOURCE_EXHAUSTED). Runs on the device's null stream and
 * returns when the outputs are written. Calls on one GPU share its array: calls from different threads
 * of one process are serialized; calls that overlap on one GPU in any other way race. */
#define FL_STATIC_MAX_FACES 196608u /* 1024 active blocks of 192 owned face slots */
#define FL_STATIC_SCRATCH_BYTES 14352640u
int fl_normalize_forces_static(int device, const fl_face_quanta* quanta, size_t face_count,
                               fl_quantization quantization, double minimum_face_mass, fl_vec3 acceleration, double dt,
                               double* out_mass, double* out_velocity, uint8_t* out_validity, double* out_predicted);
int fl_apply_forces(fl_ctx* ctx, const double* old_face_velocities, size_t face_count, fl_vec3 acceleration,
                    double dt, double* out_predicted, fl_measurement* measurement);
/* GatherMac3D (native/include/fluidity/host/transfers_3d.h:55-61; native/src/transfers_3d.cc:199-261). */
int fl_g2p(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
           const fl_particle* particles, size_t n, const double* old_face_velocities,
           const double* new_face_velocities, size_t face_count, fl_gather_result* out, fl_measurement* measurement);

/* ---- pressure ---- */
/* ProjectMacPressure3D (native/include/fluidity/host/mac_pressure_3d.h:88-102;
 * native/src/mac_pressure_workspace_3d.cc:462-659). Cells in (z,y,x) order; faces are the exact
 * union of incident faces in owned-slot order. Any output pointer may be null. */
int fl_project(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count, double dt,
               const fl_pressure_cell* cells, size_t cell_count, const fl_pressure_face* faces, size_t face_count,
               const fl_pressure_options* options, double* out_pressures, double* out_velocities, double* out_rhs,
               double* out_residuals, double* out_defects, uint64_t* out_gauges,
               fl_pressure_diagnostics* out_diagnostics, fl_measurement* measurement);

/* ---- device-resident liquid step (no host staging inside the timed region) ----
 * The scene is uploaded once (SoA particles, sorted world positions = trial positions), topology and
 * pressure structure are prepared once, and fl_step runs bin-sort -> P2G -> normalize/forces ->
 * projection -> G2P on the device. The span-based calls above reuse the context buffers that hold the
 * scene: after any of them, fl_scene_prepare, fl_step and fl_scene_download return FailedPrecondition
 * until fl_scene_upload runs again. */
typedef struct fl_step_report {
  double sort_ms, p2g_ms, normalize_ms, projection_ms, g2p_ms, total_ms;
  double p2g_kernel_ms, g2p_kernel_ms; /* the P2G and G2P kernels alone (hipEvents around the launch) */
  uint64_t particles, blocks, faces, pressure_cells, nonempty_bins;
  fl_pressure_diagnostics pressure;
  int32_t status;
  int32_t pad;
} fl_step_report;

int fl_scene_upload(fl_ctx* ctx, const fl_geometry* geometry, size_t n, const double* positions_xyz,
                    const double* velocities_xyz, const double* slopes_rowmajor, double particle_mass);
/* Activates blocks on the device and uploads the pressure cells (z,y,x order; faces are derived
 * on the device: liquid-liquid internal, liquid-air Dirichlet p=0 at distance h, and domain
 * walls [0, domain_cells) sealed). */
int fl_scene_prepare(fl_ctx* ctx, const fl_pressure_cell* cells, size_t cell_count, int32_t domain_cells,
                     double density);
int fl_step(fl_ctx* ctx, double dt, fl_vec3 gravity, fl_quantization quantization,
            const fl_pressure_options* options, fl_step_report* report);
/* Copies device results of the last fl_step back (any pointer may be null): face quanta (owned
 * slots), u0, u*, corrected face velocities, sorted order, gather results in caller order. */
int fl_scene_download(fl_ctx* ctx, fl_int3* blocks, size_t block_capacity, fl_face_quanta* quanta, double* u0,
                      double* ustar, double* unew, fl_sort_entry* order, fl_gather_result* gather,
                      double* pressures);

/* ProjectMacPressure3D as fl_project (same records, checks, solver and results), with the system's
 * structure, gauges and multigrid hierarchy built on the device from the uploaded records instead of
 * on the host: the projection path of fl_substep. A component without an anchor (no compliance, no
 * Dirichlet link) has its compatibility test and gauge computed on the host. The learned preconditioner
 * is not available on this path (FailedPrecondition). */
int fl_project_device_built(fl_ctx* ctx, const fl_geometry* geometry, const fl_int3* blocks, size_t block_count,
                            double dt, const fl_pressure_cell* cells, size_t cell_count, const fl_pressure_face* faces,
                            size_t face_count, const fl_pressure_options* options, double* out_pressures,
                            double* out_velocities, fl_pressure_diagnostics* out_diagnostics);

/* The particle-union liquid field (LiquidParticleSphereUnion3D) of PrepareField and SampleWorldBatch
 * (native/src/liquid_substep_3d_staging.cc:759-1000; native/src/liquid_substep_3d_cpu.cc:77-120):
 * the candidate cells (every cell of floor((x -/+ r - o) / h) on each axis around every particle),
 * the centre samples (27 cells around each candidate) and corner samples (8 corners of each candidate),
 * each unique and (z,y,x) sorted, with FieldDistance = min over the particles of hypot(sample - x) - r
 * at every sample. Counts are always written; arrays only when all three are given and large enough.
 * More candidates than candidate_capacity is ResourceExhausted. */
int fl_sample_particle_field(fl_ctx* ctx, const fl_geometry* geometry, const fl_vec3* positions, size_t n,
                             double radius, uint64_t candidate_capacity, fl_int3* out_cells, uint64_t* cell_count,
                             fl_center_distance* out_centres, uint64_t centre_capacity, uint64_t* centre_count,
                             fl_center_distance* out_corners, uint64_t corner_capacity, uint64_t* corner_count,
                             fl_measurement* measurement);

/* Transfer arithmetic of the device-resident paths (fl_step, fl_substep): 64 (the default) keeps the
 * reference's FP64 arithmetic and bit-exact quanta; 32 computes the deposit's stencils and
 * contributions in FP32 (the FP32 transfers of the GPU configurations) into the same integer quanta,
 * rounded per contribution, so the sums stay deterministic and agree with FP64 to about 1e-6 relative.
 * No reference counterpart: the reference's deposit is FP64 only. Span-based entry points are FP64. */
int fl_set_transfer_precision(fl_ctx* ctx, int bits);

/* ---- device-resident substep over moving particles ----
 * One liquid substep on the uploaded scene in the stage order of AttemptImpl
 * (native/src/liquid_substep_3d.cc:1175-2820; the backend members it chains are topology, deposit,
 * normalize, forces, freeze, links, project, gather and update_particles of LiquidSubstepBackend3D,
 * native/include/fluidity/host/liquid_substep_backend_3d.h:47-260, and the field sampling of
 * PrepareField, native/src/liquid_substep_3d_staging.cc:759-1000, for LiquidParticleSphereUnion3D):
 * stable bin sort, the particle-union field (candidate cells, centre and corner samples, signed
 * distances), activation, deposit quanta, normalize + gravity, the frozen boundary geometry with the
 * obstacle, typed pressure links (incompressible cells of one density at rest, the predicted face
 * velocities u*), the MGPCG projection on a device-built system and hierarchy, GatherMac3D of the old
 * and corrected face velocities, and the particle update (ASFLIP with the energy blend, the surface
 * gate with the field value -r at each particle, the obstacle's contact, J). The updated particles,
 * in this substep's sort order, are the scene's particles for the next call (fl_substep_download).
 * Statuses are the failing stage's; report.stage names it (1 sort, 2 field and activation, 3 deposit
 * and forces, 4 freeze, 5 links, 6 projection, 7 gather, 8 update). A failed call leaves the scene's
 * particles as they were, so a caller can retry with a shorter step (the update's Aborted on a move
 * beyond half a cell is the reference scheduler's cue to do so). No reference counterpart exists
 * for the options record as a whole: it gathers the scene and options fields those stages read. */
/* Level-set solids of a scene, sampled by fl_substep at the particle field's corner keys at the start
 * of each substep into a SampledSolidBoundary3D (liquid_boundary_geometry_3d.h:83-91: corner signed
 * distances and wall velocities), which the freeze clips the liquid with. The solid is the union of
 * the parts selected in `parts`: a sphere; a box paddle turned by paddle_angle about the vertical (y)
 * axis through its pivot and rotating at paddle_angular_velocity (rad/s about +y); and the side walls
 * of a container (solid where x or z lies outside [walls_low, walls_high]). Distances are positive
 * outside the solid (the least over the parts); a corner's wall velocity is that of the part nearest
 * to it (omega x r on the paddle, zero otherwise). The caller advances paddle_angle between substeps.
 * No reference counterpart as a record: the reference's substep samples its solids from a snow model
 * (liquid_substep_3d.cc:1286-1311); the sampled-solid freeze is the same. */
typedef struct fl_scene_solid {
  uint32_t parts; /* bit 0 sphere, bit 1 paddle, bit 2 container walls; 0: no sampled solid */
  uint32_t split_rule; /* the freeze's wet-piece refusal, as fl_freeze_inputs' has_sampled_solid: 0 or 1 the
                        * reference's, 2 only in cells a solid reaches, 3 none */
  fl_vec3 sphere_centre;
  double sphere_radius;
  fl_vec3 paddle_pivot;
  fl_vec3 paddle_half_extent; /* body frame */
  double paddle_angle;
  double paddle_angular_velocity;
  fl_vec3 walls_low, walls_high; /* x and z bounds of the container */
} fl_scene_solid;                /* 152 B */
typedef struct fl_substep_options {
  double step_duration;
  fl_vec3 gravity;
  fl_quantization quantization;
  double minimum_face_mass;
  double field_radius; /* LiquidParticleSphereUnion3D::radius, metres */
  double liquid_density;
  double surface_tension;
  double ambient_pressure;
  double particle_energy_blend;
  double friction_coefficient;
  double penetration_tolerance;
  fl_obstacle obstacle; /* absent, half-space, slab or open tank (freeze and contact); absent or
                         * half-space with a sampled solid */
  fl_pressure_options pressure; /* preconditioner 0 (V-cycle) */
  fl_scene_solid solid;         /* parts 0: none */
  /* Uniform dynamic viscosity (Pa s). Positive: the projection is the coupled viscous pressure solve
   * (fl_project_viscous) on AssembleFullCellMacStrain3D's rows of the pressure cells, preconditioned by
   * the V-cycle (pressure.preconditioner 0) or the reference's Jacobi diagonal (2); 0: the inviscid
   * MGPCG projection. */
  double dynamic_viscosity;
} fl_substep_options; /* 456 B */
typedef struct fl_substep_report {
  int32_t status;
  int32_t stage;
  uint64_t error_ordinal; /* first failing element of the failing stage (a particle's ordinal in the
                           * substep's input order for the sort, deposit and update), or ~0 */
  uint64_t particles, blocks, candidates, centre_samples, corner_samples, frozen_cells, frozen_faces;
  uint64_t lower_theta_count;
  double capillary_step_bound;
  fl_pressure_diagnostics pressure;
  double sort_ms, field_ms, deposit_ms, freeze_ms, links_ms, projection_ms, gather_ms, update_ms, total_ms;
  double max_speed; /* largest particle speed after the update (m/s), for the next step's CFL duration */
  fl_viscous_diagnostics viscous; /* the coupled viscous projection's (dynamic_viscosity > 0) */
  uint64_t strain_rows;
} fl_substep_report; /* 336 B */
int fl_substep(fl_ctx* ctx, const fl_substep_options* options, fl_substep_report* report);
/* The scene's particles after substeps, in the order of the last substep's sort (any pointer may be
 * null): positions and velocities (3 per particle), slopes (9, row-major), volume ratios and flags
 * (bit 0: PIC initialised). n must be the scene's particle count. */
int fl_substep_download(fl_ctx* ctx, double* positions_xyz, double* velocities_xyz, double* slopes_rowmajor,
                        double* volume_ratios, uint8_t* flags, size_t n);

/* ---- measurement utility (no counterpart in the reference) ----
 * Bandwidth of a plain float4 device-to-device copy of `bytes` repeated `reps` times (read + write
 * bytes over HIP-event time), the copy ceiling the HBM figures are compared with. Allocates and frees
 * its two buffers. */
int fl_probe_copy(int device, size_t bytes, int reps, double* gbs);

/* Measurement of the projection's kernels on the last system the context projected (fl_project or
 * fl_step): each V-cycle and PCG kernel is launched `reps` times back to back on the context stream
 * between HIP events, on the arrays of that system, and its mean launch time is reported with the
 * level and the number of cell slots (64 per 4^3 block; cells for the evaluation) it covers. The
 * launches overwrite only solver scratch (level vectors, PCG vectors) and recompute the accepted
 * pressure's residuals and defects. FL_FAILED_PRECONDITION before any projection; *count receives the
 * number of entries (FL_RESOURCE_EXHAUSTED when it exceeds capacity). */
enum {
  FL_KERNEL_SMOOTH = 0,            /* red-black Gauss-Seidel half sweep, level 0 */
  FL_KERNEL_RESIDUAL_RED = 1,      /* residual of the red cells, level 0 */
  FL_KERNEL_RESTRICT_RED = 2,      /* sum of the red children into the coarse right-hand side */
  FL_KERNEL_PROLONG = 3,           /* coarse correction added to level 0 */
  FL_KERNEL_BLOCK_SMOOTH_PRE = 4,  /* block smoother, pre-smoothing, level 1 */
  FL_KERNEL_BLOCK_SMOOTH_POST = 5, /* block smoother with prolongation, post-smoothing, level 1 */
  FL_KERNEL_RESIDUAL_ALL = 6,      /* residual of every cell, level 1 */
  FL_KERNEL_RESTRICT_ALL = 7,      /* sum of all children, level 1 to 2 */
  FL_KERNEL_TAIL_VCYCLE = 8,       /* single-workgroup cycle of the smallest levels */
  FL_KERNEL_DIRECTION_APPLY = 9,   /* d = z + beta d, q = A d, d . q (FP32, FP64 partials) */
  FL_KERNEL_UPDATE_XR = 10,        /* x += alpha d, r -= alpha q, residual norms */
  FL_KERNEL_DOT = 11,              /* r . z */
  FL_KERNEL_EVALUATE = 12          /* FP64 residual, defect and identity check */
};
typedef struct fl_kernel_time {
  uint32_t kernel;
  uint32_t level;
  uint64_t slots;
  double ms;
} fl_kernel_time;
int fl_time_projection_kernels(fl_ctx* ctx, int reps, fl_kernel_time* out, size_t capacity, size_t* count);

/* ---- learned preconditioner (no counterpart in the reference) ----
 * A 3-level 3D U-Net (channels 8/16/32, 3x3x3 convolutions) on 8^3 tiles of the level-0 pressure grid
 * maps (residual / its RMS, liquid, air, solid) to a correction; with fl_pressure_options.preconditioner
 * = 1 it preconditions a flexible CG that falls back to the V-cycle after 3 stagnating iterations.
 * Parameters: per layer, weights [cout][cin][27] then biases [cout], FP32. Tiles are [n][4][512]
 * inputs and [n][512] outputs/targets, voxel (z*8 + y)*8 + x. fl_learn_setup allocates the network's
 * device memory for tile_capacity tiles and pair_capacity training pairs. */
size_t fl_learn_param_count(void);
int fl_learn_setup(fl_ctx* ctx, uint64_t tile_capacity, uint64_t pair_capacity);
int fl_learn_set_params(fl_ctx* ctx, const float* params, size_t count);
int fl_learn_get_params(fl_ctx* ctx, float* out, size_t count);
/* FP32 forward pass (the training precision). */
int fl_learn_forward(fl_ctx* ctx, const float* inputs, size_t n, float* outputs);
/* Inference precision used by the preconditioner: the 8^3 layers in BF16 on the matrix cores with
 * FP32 accumulation, the rest FP32. */
int fl_learn_infer(fl_ctx* ctx, const float* inputs, size_t n, float* outputs);
/* Masked mean-squared loss (mask = the liquid channel) and its gradients; either output may be null. */
int fl_learn_loss_grad(fl_ctx* ctx, const float* inputs, const float* targets, size_t n, float* loss,
                       float* grad_params, float* grad_inputs);
int fl_learn_add_pairs(fl_ctx* ctx, const float* inputs, const float* targets, size_t n);
int fl_learn_pairs(fl_ctx* ctx, size_t* count);
/* every > 0: each V-cycle-preconditioned iteration number divisible by `every` appends one
 * (residual, correction) pair per tile of the system to the training set; clear empties it. */
int fl_learn_collect(fl_ctx* ctx, int every, int clear);
/* Mean masked loss of the current parameters over the stored pairs (forward only). */
int fl_learn_eval(fl_ctx* ctx, size_t batch, float* mean_loss);
/* Adam over the training set, seeded shuffling; epoch_losses (may be null) receives the mean loss. */
int fl_learn_train(fl_ctx* ctx, int epochs, size_t batch, float learning_rate, uint64_t seed, float* epoch_losses);

#ifdef __cplusplus
}
#endif
#endif /* FLUIDITY_HIP_H_ */
