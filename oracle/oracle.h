// CPU oracle for the FluidiTy liquid-step hot path (P2G, G2P, MAC pressure projection).
//
// TEST INFRASTRUCTURE ONLY. This library is the checker: it is linked or loaded by
// tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing
// that ships. The product path (fluidity_amd/) never calls it.
//
// It is an FP64 C++ restatement of the reference algorithms, with the reference's
// particle/cell caps lifted (callers pass the cap they want). Every function cites the
// reference file:line it follows. Built with -ffp-contract=off -fno-fast-math so the
// rounding sequence equals the reference's (cmake/NativeOptions.cmake:25).
#pragma once

#include <cstddef>
#include <cstdint>

extern "C" {

// Status codes equal absl::StatusCode numbering (absl/status/status.h).
enum flo_status {
  FLO_OK = 0,
  FLO_INVALID_ARGUMENT = 3,
  FLO_RESOURCE_EXHAUSTED = 8,
  FLO_FAILED_PRECONDITION = 9,
  FLO_ABORTED = 10,
  FLO_OUT_OF_RANGE = 11,
  FLO_INTERNAL = 13,
};

// Byte-for-byte mirrors of native/include/fluidity/base/mac_types_3d.h:9-72 and
// mac_transfer_types_3d.h:7-37.
struct flo_vec3 { double x, y, z; };
struct flo_mat3 { flo_vec3 x_row, y_row, z_row; };
struct flo_int3 { int32_t x, y, z; };
struct flo_geometry { flo_vec3 world_origin; double cell_width; };
struct flo_face_index { uint8_t axis; flo_int3 lattice; };  // 16 B, lattice at offset 4
struct flo_particle {
  flo_vec3 sorted_world_position;
  flo_vec3 trial_world_position;
  double particle_mass;
  flo_vec3 particle_velocity;
  flo_mat3 particle_velocity_slopes;
};  // 152 B
struct flo_quantization { double mass_per_quantum; double pure_momentum_per_quantum; uint64_t declared_particle_capacity; };
struct flo_face_deposit { double deposited_mass; double deposited_pure_momentum; };
struct flo_face_quanta { int64_t mass_quanta_sum; int64_t pure_momentum_quanta_sum; int64_t stress_force_quanta_sum; };
struct flo_gather_result { flo_vec3 new_pic_velocity; flo_vec3 grid_velocity_increment; flo_mat3 fitted_velocity_slopes; flo_mat3 velocity_gradient; };
struct flo_stencil_sample { uint64_t face_slot; double weight; flo_vec3 face_offset_from_particle; flo_vec3 weight_gradient; };
// native/include/fluidity/base/liquid_boundary_transfer_types_3d.h:4-20
struct flo_boundary_particle { flo_particle transfer; int64_t side_tag; uint8_t affine_enabled; uint8_t pad[7]; };  // 168 B
struct flo_face_tag { int64_t side_tag; uint8_t structurally_eligible; uint8_t pad[7]; };                          // 16 B
struct flo_compatible_sample {
  uint64_t face_slot;
  uint8_t compatible;
  uint8_t pad[7];
  double normalized_weight;
  flo_vec3 normalized_weight_gradient;
  flo_vec3 centered_face_offset;
};  // 72 B
// LiquidPrescribedWallSample3D, LiquidBoundaryVelocityHistory3D (spans flattened),
// LiquidBoundaryParticleGather3D (liquid_boundary_transfer_types_3d.h:21-36; liquid_boundary_transfer_3d.h:83-90).
// Validity: 0 missing, 1 valid, 2 predicted, 3 low mass.
struct flo_wall_sample { uint8_t available; uint8_t pad[7]; int64_t side_tag; flo_vec3 unit_normal; flo_vec3 wall_velocity; };  // 64 B
struct flo_history {
  double velocity_sample_time;
  double wall_sample_time;
  const double* face_velocities;
  const uint8_t* validity;
  const flo_face_tag* face_tags;
  const flo_wall_sample* wall_samples;
  size_t count;
};
struct flo_boundary_gather {
  flo_vec3 new_pic_velocity;
  flo_vec3 grid_velocity_increment;
  flo_vec3 particle_minus_old_pic_velocity;
  flo_mat3 fitted_velocity_slopes;
  flo_mat3 velocity_gradient;
  uint8_t affine_enabled, old_pic_fallback, new_pic_fallback, low_mass_fallback;
  uint8_t pad[4];
};  // 224 B
struct flo_sort_entry { flo_int3 bin_index; uint32_t pad; uint64_t original_ordinal; };  // 24 B, as LiquidParticleSortEntry3D

// Flattened MacPressureCell3D / MacPressureFace3D (native/include/fluidity/host/mac_pressure_3d.h:14-51).
// link_kind: 0 sealed, 1 internal (lift = jump), 2 Dirichlet (lift = exterior pressure).
struct flo_pressure_cell {
  flo_int3 cell_index;
  int32_t pad;
  double fluid_volume;
  double target_divergence;
  double bulk_modulus;  // +inf selects zero compliance
  double previous_pressure;
  double integrated_inward_wall_flux;
};  // 56 B
struct flo_pressure_face {
  int32_t axis;
  flo_int3 lattice;
  double open_area;
  double inverse_density;
  double pressure_gradient_distance;
  double predicted_normal_velocity;
  double wall_normal_velocity;
  int32_t link_kind;
  int32_t has_prescribed_wall_flux;
  double lift;
  double prescribed_wall_flux;
};  // 80 B

struct flo_pressure_options {
  uint64_t max_iterations;             // PCG iteration budget
  double cpu_residual_tolerance_scale; // reference option, [1e-12, 1]
  double relative_divergence_target;   // 0 disables; else also require ||defect||/||defect(p=0)|| <= target
  int32_t levels;                      // geometric levels (>= 1)
  int32_t smoothing_sweeps;            // RBGS sweeps before and after the coarse correction
  int32_t coarse_sweeps;               // symmetric RBGS sweeps on the coarsest level
  int32_t fp32_vcycle;                 // 1: V-cycle and CG vectors in FP32 (mixed), 0: all FP64
  double coarse_scale;                 // Galerkin coarse operator scale (1 = P^T A P)
  int32_t preconditioner;              // 0: the geometric V-cycle above; 1: the reference's aggregation
                                       // multigrid (levels, sweeps and coarse_scale then unused)
  int32_t pad;
};

// One level of the reference's aggregation hierarchy (mac_pressure_3d_multigrid.cc:218-301).
struct flo_hierarchy_level {
  uint64_t level_count;      // levels of the whole hierarchy
  uint64_t unknowns;         // unknowns of this level
  uint64_t edges;            // edges of this level
  uint64_t terminal_blocks;  // Cholesky blocks (last level only)
  uint64_t max_terminal;     // largest terminal block
};

struct flo_pressure_diagnostics {
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
};

// ---- grid (sparse_mac_grid_3d.cc) ----
int flo_compute_bin(const flo_geometry* geometry, flo_vec3 position, flo_int3* out_bin);
int flo_grid_sort_blocks(const flo_geometry* geometry, const flo_int3* blocks, size_t count, flo_int3* out_sorted);
int flo_face_slot(const flo_int3* sorted_blocks, size_t count, flo_face_index face, uint64_t* out_slot);
int flo_slot_face(const flo_int3* sorted_blocks, size_t count, uint64_t slot, flo_face_index* out_face);
int flo_face_world_anchor(const flo_geometry* geometry, flo_face_index face, flo_vec3* out);
int flo_owned_stencil(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t count, int axis,
                      flo_vec3 sorted_position, flo_vec3 trial_position, flo_stencil_sample* out27);

// ---- transfers (transfers_3d.cc) ----
int flo_deposit(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count,
                const flo_particle* particles, size_t n, flo_face_deposit* out, size_t face_count);
// max_capacity: the reference limit 2^20 (transfers_3d.cc:22) or any larger value for the uncapped form.
int flo_deposit_quanta(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count,
                       const flo_particle* particles, size_t n, flo_quantization quantization,
                       uint64_t max_capacity, flo_face_quanta* out, size_t face_count, int threads);
int flo_gather(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count,
               const flo_particle* particles, size_t n, const double* old_faces, const double* new_faces,
               size_t face_count, flo_gather_result* out, int threads);
// liquid_boundary_transfer_3d.cc:329-460 (DepositCore). stencils: null, or n * 3 * 27 samples in
// particle, component, (z,y,x) order. max_capacity replaces the reference's 1024-particle cap.
int flo_deposit_compatible(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count,
                           const flo_boundary_particle* particles, size_t n, const flo_face_tag* tags,
                           size_t tag_count, flo_quantization quantization, uint64_t max_capacity,
                           flo_face_quanta* out, size_t face_count, flo_compatible_sample* stencils);
// liquid_boundary_transfer_3d.cc:519-651 (GatherCore). out_counts: old-PIC, new-PIC and low-mass fallback
// particle counts. The histories alias (same samples) when all four pointers are equal.
int flo_fit_ghost_affine(const flo_vec3* offsets, const double* velocities, size_t n, int* full_rank,
                         double* at_target, flo_vec3* slopes);
int flo_gather_boundary(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count,
                        const flo_boundary_particle* particles, size_t n, double old_time, double new_time,
                        const flo_history* old_history, const flo_history* new_history,
                        flo_boundary_gather* out, uint64_t* out_counts);
int flo_normalize(const flo_face_quanta* quanta, size_t face_count, flo_quantization quantization,
                  double minimum_face_mass, double* out_mass, double* out_velocity, uint8_t* out_validity);
int flo_apply_forces(const flo_int3* sorted_blocks, size_t block_count, const double* old_velocities,
                     size_t face_count, flo_vec3 acceleration, double dt, double* out_predicted);

// ---- variational MAC viscosity (native/src/viscosity.cc; host/viscosity.h:47-160) ----
// MacViscosityCell3D, MacStrainTerm3D and MacStrainRow3D (axes as 0/1/2; prescribed strain carried).
struct flo_viscosity_cell { flo_int3 cell_index; uint32_t pad; double dynamic_viscosity; };  // 24 B
struct flo_strain_term { uint64_t face_slot; double coefficient; };                            // 16 B
struct flo_strain_row {
  flo_int3 low_cell;
  uint8_t first_axis, second_axis, pad[2];
  flo_strain_term terms[8];
  uint64_t term_count;
  double quadrature_weight;
  double prescribed_strain;
};  // 168 B
// AssembleFullCellMacStrain3D: three normal rows per cell in cell and axis order, then the interior
// shear rows in low-cell and axis-pair order (xy, xz, yz). Cells strictly (z,y,x) sorted.
int flo_assemble_full_cell_strain(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count,
                                  const flo_viscosity_cell* cells, size_t n, flo_strain_row* out, size_t capacity,
                                  size_t* out_count);
// SolveMacViscousVelocityBlock3D: (M + dt E^T W E) u = rhs by Jacobi PCG on positive-inertia faces.
int flo_solve_viscous_block(const flo_strain_row* rows, size_t row_count, const double* inertia, size_t face_count,
                            double dt, const double* rhs, uint64_t max_iterations, double relative_tolerance,
                            double* out_velocity, uint64_t* out_iterations, double* out_max_residual);

// ProjectMacViscousPressure3D (mac_viscous_pressure_3d.cc:151-459, 558-575, 699-705): the coupled
// viscous pressure solve on full or supplied strain rows (face_slot: grid slots), GCR with the Jacobi
// diagonal of the viscous Schur complement, restart length 32, iteration budget in [1, 256].
struct flo_viscous_diagnostics {
  uint64_t pressure_iterations;
  double constraint_residual_l2, initial_constraint_residual_l2, max_cell_constraint_defect;
  double max_momentum_residual, inward_prescribed_wall_flux;
  uint64_t velocity_solves;
};
int flo_project_viscous(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count, double dt,
                        const flo_pressure_cell* cells, size_t cell_count, const flo_pressure_face* faces,
                        size_t face_count, const flo_strain_row* rows, size_t row_count, uint64_t max_iterations,
                        double* out_pressures, double* out_velocities, double* out_defects,
                        flo_viscous_diagnostics* out_diagnostics);

// ---- pressure links (liquid_boundary_geometry_3d.cc:1077-1383) ----
// LiquidCenterDistance3D, FrozenLiquidCell3D, FrozenLiquidFace3D (with its optional wall flux as value +
// flag, the libstdc++ layout of std::optional<double>), LiquidPressureCellFields3D,
// LiquidAirCellPressure3D and LiquidInternalInterface3D (native/include/fluidity/host/liquid_boundary_geometry_3d.h).
struct flo_center_distance { flo_int3 cell_index; uint32_t pad; double signed_distance; };  // 24 B
struct flo_frozen_cell { flo_int3 cell_index; uint32_t pad; double fluid_volume, center_distance, curvature; };  // 40 B
struct flo_frozen_face {
  flo_face_index face_index;
  uint64_t face_slot;
  double open_area, blocked_area, wall_normal_velocity;
  double prescribed_wall_flux;
  uint8_t has_prescribed_wall_flux, pad1[7];
  uint8_t is_low_x_piston_face, pad2[7];
};  // 72 B
struct flo_cell_fields {
  double target_divergence, bulk_modulus, previous_pressure;
  double density;
  uint8_t has_density, pad[7];
};  // 40 B
struct flo_air_pressure { flo_int3 cell_index; uint32_t pad; double gas_gauge_pressure; };  // 24 B
struct flo_internal_interface {
  uint64_t face_ordinal;
  double tail_fraction, surface_tension, curvature_tail_to_head, normal_viscous_jump_tail_to_head;
};  // 40 B
struct flo_optional_flux { double value; uint8_t has_value, pad[7]; };  // 16 B
// FrozenLiquidBoundaryGeometryView3D (native/include/fluidity/host/liquid_boundary_workspace_3d.h:34-43).
struct flo_frozen_view {
  const flo_int3* blocks;
  uint64_t block_count;
  flo_geometry geometry;
  const flo_center_distance* center_distances;
  uint64_t center_count;
  const flo_frozen_cell* cells;
  uint64_t cell_count;
  const flo_frozen_face* faces;
  uint64_t face_count;
  double step_duration, liquid_density, surface_tension, capillary_step_bound;
};
// LiquidPressureFields3D (liquid_boundary_geometry_3d.h:170-185); empty spans have count 0.
struct flo_pressure_fields {
  const flo_cell_fields* cells;
  uint64_t cell_count;
  const double* predicted_face_velocities;
  uint64_t predicted_count;
  double ambient_pressure;
  const flo_optional_flux* wetted_piston_face_fluxes;
  uint64_t wetted_count;
  const flo_air_pressure* air_cell_pressures;
  uint64_t air_count;
  const flo_internal_interface* internal_interfaces;
  uint64_t interface_count;
};
// BuildLiquidPressureLinks3D on a frozen view (ValidateFrozenView, then LinksCore): one pressure cell
// per frozen cell and one typed link per frozen face, plus the count of theta values raised to 0.05.
// Outputs are written only on success. The reference's 4096-cell / 24576-face workspace caps are not
// applied.
int flo_build_pressure_links(const flo_frozen_view* view, const flo_pressure_fields* fields,
                             flo_pressure_cell* out_cells, flo_pressure_face* out_faces,
                             uint64_t* out_lower_theta_count);

// Frozen-geometry inputs (LiquidBoundaryGeometryViewInputs3D, native/include/fluidity/host/
// liquid_boundary_workspace_3d.h:18-30) for the absent, half-space and slab obstacles
// (liquid_boundary_geometry_3d.h:30-43). Corner samples use the centre-sample record.
enum { FLO_OBSTACLE_ABSENT = 0, FLO_OBSTACLE_HALFSPACE = 1, FLO_OBSTACLE_SLAB = 2, FLO_OBSTACLE_TANK = 3 };
// AxisAlignedLiquidTank3D (native/include/fluidity/base/liquid_boundary_math.h:22-29).
struct flo_tank { double start_low_x, high_x, floor_y, low_z, high_z, piston_velocity_x; };  // 48 B
struct flo_obstacle {
  uint32_t kind, pad;
  flo_vec3 unit_normal;
  double coordinate;  // half-space: start plane coordinate; slab: start centre coordinate
  double half_width;  // slab only
  flo_vec3 wall_velocity;
  flo_tank tank;      // open tank only (the particle update; the freeze refuses it)
};  // 120 B
// SolidCornerDistance3D: corner key, signed distance (positive outside the solid), wall velocity. 48 B.
struct flo_solid_corner { flo_int3 corner_index; uint32_t pad; double signed_distance; flo_vec3 wall_velocity; };
struct flo_freeze_inputs {
  const flo_int3* blocks;
  uint64_t block_count;
  flo_geometry geometry;
  const flo_int3* cell_indices;
  uint64_t cell_count;
  const flo_center_distance* center_distances;
  uint64_t center_count;
  const flo_center_distance* corner_distances;
  uint64_t corner_count;
  flo_obstacle obstacle;
  double step_duration, liquid_density, surface_tension;
  double exact_curvature;
  uint8_t has_exact_curvature, pad[7];
  uint64_t max_cells;  // ResourceExhausted above this many candidates (0: no limit)
  uint64_t max_faces;  // ResourceExhausted above this many frozen faces (0: no limit)
  // SampledSolidBoundary3D (host/liquid_boundary_geometry_3d.h:77-91), used when has_sampled_solid:
  // solid corner samples (distinct, (z,y,x) sorted) and, when face_velocity_count is not 0, one wall
  // normal velocity per owned slot.
  const struct flo_solid_corner* solid_corners;
  uint64_t solid_corner_count;
  const double* face_normal_velocities;
  uint64_t face_velocity_count;
  uint8_t has_sampled_solid, pad2[7];  // 2: the wet-piece refusal only in cells with a corner in the solid, 3: none
};
// FreezeCore (liquid_boundary_geometry_3d.cc:765-1068) with the scalar geometry queries of
// base/liquid_boundary_geometry_math_3d.h (kCell, kFace, kCurvature, kCapillary). out_cells holds
// cell_count records and out_faces 6 * cell_count; the counts and the capillary step bound are
// written only on success.
int flo_freeze_geometry(const flo_freeze_inputs* inputs, flo_frozen_cell* out_cells, uint64_t* out_cell_count,
                        flo_frozen_face* out_faces, uint64_t* out_face_count, double* out_capillary_step_bound);
// MeasureSampledSolidLiquidCell (base/sampled_solid_liquid_geometry_3d.h:118-218): wet volume and the six
// open face areas (low/high x, y, z) of a cell with liquid and solid corner distances; analytic (may be
// null) is an open half-space {normal, coordinate} in cell-width coordinates. Returns 0, or
// FLO_INVALID_ARGUMENT / FLO_OUT_OF_RANGE / FLO_FAILED_PRECONDITION for the reference's kInvalid /
// kRange / kPrecondition.
int flo_measure_sampled_cell(double width, const double liquid[8], const double solid[8], const double* analytic,
                             double* out_wet_volume, double out_open_area[6]);
// MeasureSampledSolidFaceWeights (:220-318): corner area weights of the blocked part of a face.
int flo_sampled_face_weights(double width, const double solid[4], const double* analytic, double out_weights[4]);

// ---- particle update (liquid_substep_3d_cpu.cc:239-405) ----
struct flo_constituent { uint64_t substance_id; double volume_fraction; };
// LiquidParticleState3D (native/include/fluidity/host/liquid_state_3d.h:100-120), 328 B.
struct flo_liquid_particle {
  uint64_t id;
  uint32_t tag, pad0;
  flo_vec3 world_position, world_velocity;
  flo_mat3 velocity_slopes;
  uint8_t affine_enabled, pad1[7];
  double mass, reference_volume, volume_ratio;
  int64_t side_tag;
  uint8_t pic_initialized, pad2[7];
  uint64_t phase_id;
  uint8_t constituent_count, quadratic_adaptive_transfer, pad3[6];
  flo_constituent constituents[8];
};
// LiquidParticleUpdateInputs3D scalars (native/include/fluidity/host/liquid_substep_backend_3d.h:16-28).
struct flo_update_options {
  double cell_width, step_duration, particle_energy_blend, friction_coefficient, penetration_tolerance;
  flo_obstacle obstacle;  // absent, half-space or slab
};  // 112 B
// The particle update without an obstacle or regular-cell transport: stored velocity
// new PIC + 0.99 e (v - old PIC), advecting velocity with the surface-expansion blend, advection
// by fma, the half-cell displacement guard and J *= exp(dt div u). Particles are updated in order and
// the first failing particle's status is returned (earlier particles stay updated, as in the reference).
int flo_update_particles(const flo_update_options* options, const flo_boundary_gather* gathered,
                         const double* liquid_distances, size_t n, flo_liquid_particle* particles);

// ---- topology (liquid_step_workspace_3d.cc:171-300) ----
// LiquidBlockState3D, LiquidCellState3D and LiquidFaceState3D (native/include/fluidity/host/liquid_state_3d.h:129-159).
struct flo_block_state {
  flo_int3 block_index;
  uint32_t pad;
  uint64_t generation;
  uint32_t activation_age, inactive_age;
};  // 32 B
struct flo_cell_state {
  flo_int3 cell_index;
  uint8_t occupancy;  // 0 air, 1 liquid, 2 solid
  uint8_t pad[3];
  double previous_pressure;
};  // 24 B
struct flo_face_state {
  flo_face_index face_index;
  uint64_t face_slot;
  double carrier_velocity;
  uint8_t validity;  // 0 missing, 1 valid, 2 predicted, 3 low mass
  uint8_t pad[7];
  int64_t side_tag;
  uint8_t structurally_eligible;
  uint8_t pad2[7];
};  // 56 B
// The topology records of a previous LiquidState3D with its geometry and clock epoch.
struct flo_topology_history {
  flo_geometry geometry;
  uint64_t epoch;
  const flo_block_state* blocks;
  uint64_t block_count;
  const flo_cell_state* cells;
  uint64_t cell_count;
  const flo_face_state* faces;
  uint64_t face_count;
};
// LiquidStateCapacities3D without sources.
struct flo_topology_capacities { uint64_t particles, blocks, cells, faces; };
// LiquidStepWorkspace3D::PrepareTopology without an imported order (liquid_step_workspace_3d.cc:171-341):
// sorted particle order, active blocks (generation from the previous epoch, or the previous block
// with its inactive age cleared), canonical unique cells and one face record per owned slot, each
// copied from the previous state when the same identity existed there. previous may be null. Outputs
// are written only on success; out arrays hold n, capacities->blocks, ->cells and ->faces records.
int flo_prepare_topology(const flo_geometry* geometry, const flo_topology_capacities* capacities,
                         const flo_vec3* positions, size_t n, const flo_int3* pressure_cells, size_t cell_count,
                         const flo_topology_history* previous, flo_sort_entry* out_order, flo_block_state* out_blocks,
                         uint64_t* out_block_count, flo_cell_state* out_cells, uint64_t* out_cell_count,
                         flo_face_state* out_faces, uint64_t* out_face_count);
int flo_sort_particles(const flo_geometry* geometry, const flo_vec3* positions, size_t n, flo_sort_entry* out);
// Writes the sorted unique active blocks; *io_count holds the capacity on entry, the count on exit.
int flo_activate_blocks(const flo_geometry* geometry, const flo_vec3* positions, size_t n,
                        const flo_int3* pressure_cells, size_t cell_count, flo_int3* out_blocks, size_t* io_count);

// ---- pressure (mac_pressure_3d_arithmetic.h, mac_pressure_workspace_3d.cc) ----
// Outputs may be null. gauges receives up to cell_count ordinals.
int flo_project(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count, double dt,
                const flo_pressure_cell* cells, size_t cell_count, const flo_pressure_face* faces, size_t face_count,
                const flo_pressure_options* options, double* out_pressures, double* out_velocities,
                double* out_rhs, double* out_residuals, double* out_defects, uint64_t* out_gauges,
                flo_pressure_diagnostics* out_diagnostics);
// Evaluate a supplied pressure vector with the FP64 original operator (workspace.cc:399-461).
int flo_evaluate(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count, double dt,
                 const flo_pressure_cell* cells, size_t cell_count, const flo_pressure_face* faces, size_t face_count,
                 const double* pressures, double* out_velocities, double* out_residuals, double* out_defects,
                 flo_pressure_diagnostics* out_diagnostics);

// The reference hierarchy of a prepared system, level `level`: coordinates (3 per unknown), anchor
// and diagonal, edges (lower, higher, conductance), aggregates into the next level (SIZE_MAX on the
// last level) and, on level 0, the cell of each unknown. Arrays hold at most `capacity` entries
// (edges: 3 * capacity); any may be null.
int flo_reference_hierarchy(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count, double dt,
                            const flo_pressure_cell* cells, size_t cell_count, const flo_pressure_face* faces,
                            size_t face_count, uint64_t level, uint64_t capacity, flo_hierarchy_level* out_level,
                            int32_t* coords, double* anchors, double* diagonal, uint64_t* edge_lo, uint64_t* edge_hi,
                            double* edge_k, uint64_t* aggregates, uint64_t* unknown_to_cell);
// One application of the reference V-cycle in FP64 (z = M r over the unknowns of level 0).
int flo_reference_vcycle(const flo_geometry* geometry, const flo_int3* sorted_blocks, size_t block_count, double dt,
                         const flo_pressure_cell* cells, size_t cell_count, const flo_pressure_face* faces,
                         size_t face_count, const double* r, double* z);

// Compiler that built the oracle (for the CPU-baseline record).
// The particle-union liquid field (field.cc): PrepareField's candidate cells, centre and corner samples
// (z,y,x sorted, unique) and FieldDistance at each sample. With null output arrays only the three
// counts are written.
int flo_particle_field(const flo_geometry* geometry, const flo_vec3* positions, size_t n, double radius,
                       uint64_t candidate_capacity, flo_int3* out_cells, uint64_t* cell_count,
                       flo_center_distance* out_centres, uint64_t centre_capacity, uint64_t* centre_count,
                       flo_center_distance* out_corners, uint64_t corner_capacity, uint64_t* corner_count);

const char* flo_compiler(void);

// ---- learned preconditioner network (unet.cc): 3-level U-Net on 8^3 tiles, FP64 ----
// Parameters: per convolution layer, weights [cout][cin][27] then biases [cout]. Inputs [n][4][512]
// (normalised residual, liquid, air, solid), outputs [n][512].
size_t flo_unet_param_count(void);
int flo_unet_forward(const double* params, size_t n, const double* in, double* out);
// Masked mean-squared loss (mask = input channel 1) and its gradient; grad_in may be null.
int flo_unet_loss_grad(const double* params, size_t n, const double* in, const double* target, double* loss,
                       double* grad_params, double* grad_in);

}  // extern "C"
