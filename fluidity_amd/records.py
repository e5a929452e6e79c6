"""numpy record layouts of the C-ABI structs in include/fluidity_hip.h (byte for byte).

They mirror the reference PODs: Vec3/Mat3/BlockIndex3D (native/include/fluidity/base/mac_types_3d.h:9-72),
MacTransferParticle3D / MacFaceDepositQuanta3D / MacParticleGatherResult3D
(native/include/fluidity/base/mac_transfer_types_3d.h:7-37), LiquidParticleSortEntry3D
(native/include/fluidity/host/liquid_step_workspace_3d.h:14-17) and the flattened
MacPressureCell3D / MacPressureFace3D (native/include/fluidity/host/mac_pressure_3d.h:14-51).
"""
import numpy as np

VEC3 = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8")])
MAT3 = np.dtype([("x_row", VEC3), ("y_row", VEC3), ("z_row", VEC3)])
INT3 = np.dtype([("x", "i4"), ("y", "i4"), ("z", "i4")])
GEOMETRY = np.dtype([("world_origin", VEC3), ("cell_width", "f8")])
PARTICLE = np.dtype([("sorted_world_position", VEC3), ("trial_world_position", VEC3), ("particle_mass", "f8"),
                     ("particle_velocity", VEC3), ("particle_velocity_slopes", MAT3)])
QUANTIZATION = np.dtype([("mass_per_quantum", "f8"), ("pure_momentum_per_quantum", "f8"),
                         ("declared_particle_capacity", "u8")])
FACE_DEPOSIT = np.dtype([("deposited_mass", "f8"), ("deposited_pure_momentum", "f8")])
FACE_QUANTA = np.dtype([("mass_quanta_sum", "i8"), ("pure_momentum_quanta_sum", "i8"),
                        ("stress_force_quanta_sum", "i8")])
GATHER_RESULT = np.dtype([("new_pic_velocity", VEC3), ("grid_velocity_increment", VEC3),
                          ("fitted_velocity_slopes", MAT3), ("velocity_gradient", MAT3)])
BOUNDARY_PARTICLE = np.dtype({"names": ["transfer", "side_tag", "affine_enabled"],
                              "formats": [PARTICLE, "i8", "u1"], "offsets": [0, 152, 160], "itemsize": 168})
FACE_TAG = np.dtype({"names": ["side_tag", "structurally_eligible"], "formats": ["i8", "u1"], "offsets": [0, 8],
                     "itemsize": 16})
COMPATIBLE_SAMPLE = np.dtype({"names": ["face_slot", "compatible", "normalized_weight", "normalized_weight_gradient",
                                        "centered_face_offset"],
                              "formats": ["u8", "u1", "f8", VEC3, VEC3], "offsets": [0, 8, 16, 24, 48],
                              "itemsize": 72})
WALL_SAMPLE = np.dtype({"names": ["available", "side_tag", "unit_normal", "wall_velocity"],
                        "formats": ["u1", "i8", VEC3, VEC3], "offsets": [0, 8, 16, 40], "itemsize": 64})
BOUNDARY_GATHER = np.dtype({"names": ["new_pic_velocity", "grid_velocity_increment", "particle_minus_old_pic_velocity",
                                      "fitted_velocity_slopes", "velocity_gradient", "affine_enabled",
                                      "old_pic_fallback", "new_pic_fallback", "low_mass_fallback"],
                            "formats": [VEC3, VEC3, VEC3, MAT3, MAT3, "u1", "u1", "u1", "u1"],
                            "offsets": [0, 24, 48, 72, 144, 216, 217, 218, 219], "itemsize": 224})
SORT_ENTRY = np.dtype([("bin_index", INT3), ("pad", "u4"), ("original_ordinal", "u8")])
PRESSURE_CELL = np.dtype([("cell_index", INT3), ("pad", "i4"), ("fluid_volume", "f8"), ("target_divergence", "f8"),
                          ("bulk_modulus", "f8"), ("previous_pressure", "f8"), ("integrated_inward_wall_flux", "f8")])
PRESSURE_FACE = np.dtype([("axis", "i4"), ("lattice", INT3), ("open_area", "f8"), ("inverse_density", "f8"),
                          ("pressure_gradient_distance", "f8"), ("predicted_normal_velocity", "f8"),
                          ("wall_normal_velocity", "f8"), ("link_kind", "i4"), ("has_prescribed_wall_flux", "i4"),
                          ("lift", "f8"), ("prescribed_wall_flux", "f8")])
PRESSURE_OPTIONS = np.dtype([("max_iterations", "u8"), ("cpu_residual_tolerance_scale", "f8"),
                             ("relative_divergence_target", "f8"), ("levels", "i4"), ("smoothing_sweeps", "i4"),
                             ("coarse_sweeps", "i4"), ("level_storage", "i4"), ("coarse_scale", "f8"),
                             ("preconditioner", "i4"), ("smoother", "i4")])
PRESSURE_DIAGNOSTICS = np.dtype([("iteration_count", "u8"), ("multigrid_level_count", "u8"),
                                 ("true_residual_l2", "f8"), ("max_volume_scaled_residual", "f8"),
                                 ("max_cell_constraint_defect", "f8"), ("rhs_l2", "f8"),
                                 ("initial_defect_l2", "f8"), ("defect_l2", "f8"), ("defect_linf", "f8"),
                                 ("restarts", "u8"), ("gauge_count", "u8"), ("learned_iterations", "u8")])
VISCOUS_DIAGNOSTICS = np.dtype([("pressure_iterations", "u8"), ("constraint_residual_l2", "f8"),
                                ("initial_constraint_residual_l2", "f8"), ("max_cell_constraint_defect", "f8"),
                                ("max_momentum_residual", "f8"), ("inward_prescribed_wall_flux", "f8"),
                                ("velocity_solves", "u8"), ("velocity_iterations", "u8")])
assert VISCOUS_DIAGNOSTICS.itemsize == 64

# LiquidBlockState3D / LiquidCellState3D / LiquidFaceState3D (native/include/fluidity/host/liquid_state_3d.h:129-159).
BLOCK_STATE = np.dtype([("block_index", INT3), ("pad", "u4"), ("generation", "u8"), ("activation_age", "u4"),
                        ("inactive_age", "u4")])
CELL_STATE = np.dtype({"names": ["cell_index", "occupancy", "previous_pressure"], "formats": [INT3, "u1", "f8"],
                       "offsets": [0, 12, 16], "itemsize": 24})
FACE_STATE = np.dtype({"names": ["axis", "lattice", "face_slot", "carrier_velocity", "validity", "side_tag",
                                 "structurally_eligible"],
                       "formats": ["u1", INT3, "u8", "f8", "u1", "i8", "u1"],
                       "offsets": [0, 4, 16, 24, 32, 40, 48], "itemsize": 56})
# Frozen boundary geometry and pressure fields (native/include/fluidity/host/liquid_boundary_geometry_3d.h).
FACE_INDEX = np.dtype({"names": ["axis", "lattice"], "formats": ["u1", INT3], "offsets": [0, 4], "itemsize": 16})
CENTER_DISTANCE = np.dtype({"names": ["cell_index", "signed_distance"], "formats": [INT3, "f8"], "offsets": [0, 16],
                            "itemsize": 24})
# SolidCornerDistance3D (liquid_boundary_geometry_3d.h:77-81), fl_solid_corner.
SOLID_CORNER = np.dtype({"names": ["corner_index", "signed_distance", "wall_velocity"], "formats": [INT3, "f8", VEC3],
                         "offsets": [0, 16, 24], "itemsize": 48})
FROZEN_CELL = np.dtype({"names": ["cell_index", "fluid_volume", "center_distance", "curvature"],
                        "formats": [INT3, "f8", "f8", "f8"], "offsets": [0, 16, 24, 32], "itemsize": 40})
FROZEN_FACE = np.dtype({"names": ["face_index", "face_slot", "open_area", "blocked_area", "wall_normal_velocity",
                                  "prescribed_wall_flux", "has_prescribed_wall_flux", "is_low_x_piston_face"],
                        "formats": [FACE_INDEX, "u8", "f8", "f8", "f8", "f8", "u1", "u1"],
                        "offsets": [0, 16, 24, 32, 40, 48, 56, 64], "itemsize": 72})
CELL_FIELDS = np.dtype({"names": ["target_divergence", "bulk_modulus", "previous_pressure", "density", "has_density"],
                        "formats": ["f8", "f8", "f8", "f8", "u1"], "offsets": [0, 8, 16, 24, 32], "itemsize": 40})
AIR_PRESSURE = np.dtype({"names": ["cell_index", "gas_gauge_pressure"], "formats": [INT3, "f8"], "offsets": [0, 16],
                         "itemsize": 24})
INTERNAL_INTERFACE = np.dtype([("face_ordinal", "u8"), ("tail_fraction", "f8"), ("surface_tension", "f8"),
                               ("curvature_tail_to_head", "f8"), ("normal_viscous_jump_tail_to_head", "f8")])
OPTIONAL_FLUX = np.dtype({"names": ["value", "has_value"], "formats": ["f8", "u1"], "offsets": [0, 8], "itemsize": 16})
CONSTITUENT = np.dtype([("substance_id", "u8"), ("volume_fraction", "f8")])
# LiquidParticleState3D (native/include/fluidity/host/liquid_state_3d.h:100-120).
LIQUID_PARTICLE = np.dtype({"names": ["id", "tag", "world_position", "world_velocity", "velocity_slopes",
                                      "affine_enabled", "mass", "reference_volume", "volume_ratio", "side_tag",
                                      "pic_initialized", "phase_id", "constituent_count",
                                      "quadratic_adaptive_transfer", "constituents"],
                            "formats": ["u8", "u4", VEC3, VEC3, MAT3, "u1", "f8", "f8", "f8", "i8", "u1", "u8", "u1",
                                        "u1", (CONSTITUENT, 8)],
                            "offsets": [0, 8, 16, 40, 64, 136, 144, 152, 160, 168, 176, 184, 192, 193, 200],
                            "itemsize": 328})
TANK = np.dtype([("start_low_x", "f8"), ("high_x", "f8"), ("floor_y", "f8"), ("low_z", "f8"), ("high_z", "f8"),
                 ("piston_velocity_x", "f8")])
OBSTACLE = np.dtype([("kind", "u4"), ("pad", "u4"), ("unit_normal", VEC3), ("coordinate", "f8"), ("half_width", "f8"),
                     ("wall_velocity", VEC3), ("tank", TANK)])
UPDATE_OPTIONS = np.dtype([("cell_width", "f8"), ("step_duration", "f8"), ("particle_energy_blend", "f8"),
                           ("friction_coefficient", "f8"), ("penetration_tolerance", "f8"), ("obstacle", OBSTACLE)])
TOPOLOGY_CAPACITIES = np.dtype([("particles", "u8"), ("blocks", "u8"), ("cells", "u8"), ("faces", "u8")])
# fl_scene_solid: the level-set solids sampled by the device substep.
SCENE_SOLID = np.dtype([("parts", "u4"), ("split_rule", "u4"), ("sphere_centre", VEC3), ("sphere_radius", "f8"),
                        ("paddle_pivot", VEC3), ("paddle_half_extent", VEC3), ("paddle_angle", "f8"),
                        ("paddle_angular_velocity", "f8"), ("walls_low", VEC3), ("walls_high", VEC3)])
# fl_substep_options / fl_substep_report (the device substep).
SUBSTEP_OPTIONS = np.dtype([("step_duration", "f8"), ("gravity", VEC3), ("quantization", QUANTIZATION),
                            ("minimum_face_mass", "f8"), ("field_radius", "f8"), ("liquid_density", "f8"),
                            ("surface_tension", "f8"), ("ambient_pressure", "f8"), ("particle_energy_blend", "f8"),
                            ("friction_coefficient", "f8"), ("penetration_tolerance", "f8"), ("obstacle", OBSTACLE),
                            ("pressure", PRESSURE_OPTIONS), ("solid", SCENE_SOLID), ("dynamic_viscosity", "f8")])
SUBSTEP_REPORT = np.dtype([("status", "i4"), ("stage", "i4"), ("error_ordinal", "u8"), ("particles", "u8"), ("blocks", "u8"),
                           ("candidates", "u8"), ("centre_samples", "u8"), ("corner_samples", "u8"),
                           ("frozen_cells", "u8"), ("frozen_faces", "u8"), ("lower_theta_count", "u8"),
                           ("capillary_step_bound", "f8"), ("pressure", PRESSURE_DIAGNOSTICS), ("sort_ms", "f8"),
                           ("field_ms", "f8"), ("deposit_ms", "f8"), ("freeze_ms", "f8"), ("links_ms", "f8"),
                           ("projection_ms", "f8"), ("gather_ms", "f8"), ("update_ms", "f8"), ("total_ms", "f8"),
                           ("max_speed", "f8"), ("viscous", VISCOUS_DIAGNOSTICS), ("strain_rows", "u8")])

SEALED, INTERNAL, DIRICHLET = 0, 1, 2
LINK_KINDS = {"sealed": SEALED, "internal": INTERNAL, "dirichlet": DIRICHLET}

assert PARTICLE.itemsize == 152 and GATHER_RESULT.itemsize == 192 and SORT_ENTRY.itemsize == 24
assert PRESSURE_CELL.itemsize == 56 and PRESSURE_FACE.itemsize == 80 and FACE_QUANTA.itemsize == 24
assert BLOCK_STATE.itemsize == 32 and CELL_STATE.itemsize == 24 and FACE_STATE.itemsize == 56
assert LIQUID_PARTICLE.itemsize == 328 and UPDATE_OPTIONS.itemsize == 160 and OBSTACLE.itemsize == 120
assert FROZEN_FACE.itemsize == 72 and FROZEN_CELL.itemsize == 40 and CELL_FIELDS.itemsize == 40
assert SUBSTEP_OPTIONS.itemsize == 456 and SUBSTEP_REPORT.itemsize == 336 and SCENE_SOLID.itemsize == 152

# MacViscosityCell3D, MacStrainTerm3D, MacStrainRow3D (native/include/fluidity/host/viscosity.h:47-66).
VISCOSITY_CELL = np.dtype({"names": ["cell_index", "dynamic_viscosity"], "formats": [INT3, "f8"], "offsets": [0, 16],
                           "itemsize": 24})
STRAIN_TERM = np.dtype([("face_slot", "u8"), ("coefficient", "f8")])
STRAIN_ROW = np.dtype({"names": ["low_cell", "first_axis", "second_axis", "terms", "term_count", "quadrature_weight",
                                 "prescribed_strain"],
                       "formats": [INT3, "u1", "u1", (STRAIN_TERM, 8), "u8", "f8", "f8"],
                       "offsets": [0, 12, 13, 16, 144, 152, 160], "itemsize": 168})
