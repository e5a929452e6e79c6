"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and by the
cpu_baseline leg of bench.py, as the checker. The product package never imports it.
Record layouts are numpy structured dtypes that match oracle/oracle.h byte for byte.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

OK, INVALID_ARGUMENT, RESOURCE_EXHAUSTED, FAILED_PRECONDITION, ABORTED, OUT_OF_RANGE, INTERNAL = 0, 3, 8, 9, 10, 11, 13

VEC3 = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8")])
MAT3 = np.dtype([("x_row", VEC3), ("y_row", VEC3), ("z_row", VEC3)])
INT3 = np.dtype([("x", "i4"), ("y", "i4"), ("z", "i4")])
GEOMETRY = np.dtype([("world_origin", VEC3), ("cell_width", "f8")])
FACE_INDEX = np.dtype({"names": ["axis", "lattice"], "formats": ["u1", INT3], "offsets": [0, 4], "itemsize": 16})
PARTICLE = np.dtype([("sorted_world_position", VEC3), ("trial_world_position", VEC3), ("particle_mass", "f8"),
                     ("particle_velocity", VEC3), ("particle_velocity_slopes", MAT3)])
QUANTIZATION = np.dtype([("mass_per_quantum", "f8"), ("pure_momentum_per_quantum", "f8"),
                         ("declared_particle_capacity", "u8")])
FACE_DEPOSIT = np.dtype([("deposited_mass", "f8"), ("deposited_pure_momentum", "f8")])
FACE_QUANTA = np.dtype([("mass_quanta_sum", "i8"), ("pure_momentum_quanta_sum", "i8"),
                        ("stress_force_quanta_sum", "i8")])
GATHER_RESULT = np.dtype([("new_pic_velocity", VEC3), ("grid_velocity_increment", VEC3),
                          ("fitted_velocity_slopes", MAT3), ("velocity_gradient", MAT3)])
STENCIL_SAMPLE = np.dtype([("face_slot", "u8"), ("weight", "f8"), ("face_offset_from_particle", VEC3),
                           ("weight_gradient", VEC3)])
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
                             ("coarse_sweeps", "i4"), ("fp32_vcycle", "i4"), ("coarse_scale", "f8"),
                             ("preconditioner", "i4"), ("pad", "i4")])
PRESSURE_DIAGNOSTICS = np.dtype([("iteration_count", "u8"), ("multigrid_level_count", "u8"),
                                 ("true_residual_l2", "f8"), ("max_volume_scaled_residual", "f8"),
                                 ("max_cell_constraint_defect", "f8"), ("rhs_l2", "f8"),
                                 ("initial_defect_l2", "f8"), ("defect_l2", "f8"), ("defect_linf", "f8"),
                                 ("restarts", "u8"), ("gauge_count", "u8")])

# LiquidBlockState3D, LiquidCellState3D, LiquidFaceState3D (native/include/fluidity/host/liquid_state_3d.h:129-159).
BLOCK_STATE = np.dtype([("block_index", INT3), ("pad", "u4"), ("generation", "u8"), ("activation_age", "u4"),
                        ("inactive_age", "u4")])
CELL_STATE = np.dtype({"names": ["cell_index", "occupancy", "previous_pressure"], "formats": [INT3, "u1", "f8"],
                       "offsets": [0, 12, 16], "itemsize": 24})
FACE_STATE = np.dtype({"names": ["face_index", "face_slot", "carrier_velocity", "validity", "side_tag",
                                 "structurally_eligible"],
                       "formats": [FACE_INDEX, "u8", "f8", "u1", "i8", "u1"],
                       "offsets": [0, 16, 24, 32, 40, 48], "itemsize": 56})
# Frozen boundary geometry and pressure-field records (native/include/fluidity/host/liquid_boundary_geometry_3d.h).
CENTER_DISTANCE = np.dtype({"names": ["cell_index", "signed_distance"], "formats": [INT3, "f8"], "offsets": [0, 16],
                            "itemsize": 24})
# SolidCornerDistance3D (host/liquid_boundary_geometry_3d.h:77-81).
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
VISCOSITY_CELL = np.dtype({"names": ["cell_index", "dynamic_viscosity"], "formats": [INT3, "f8"], "offsets": [0, 16],
                           "itemsize": 24})
STRAIN_TERM = np.dtype([("face_slot", "u8"), ("coefficient", "f8")])
STRAIN_ROW = np.dtype({"names": ["low_cell", "first_axis", "second_axis", "terms", "term_count", "quadrature_weight",
                                 "prescribed_strain"],
                       "formats": [INT3, "u1", "u1", (STRAIN_TERM, 8), "u8", "f8", "f8"],
                       "offsets": [0, 12, 13, 16, 144, 152, 160], "itemsize": 168})
TOPOLOGY_CAPACITIES = np.dtype([("particles", "u8"), ("blocks", "u8"), ("cells", "u8"), ("faces", "u8")])
assert BLOCK_STATE.itemsize == 32 and CELL_STATE.itemsize == 24 and FACE_STATE.itemsize == 56
assert PARTICLE.itemsize == 152 and GATHER_RESULT.itemsize == 192 and SORT_ENTRY.itemsize == 24
assert PRESSURE_CELL.itemsize == 56 and PRESSURE_FACE.itemsize == 80 and FACE_QUANTA.itemsize == 24


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def geometry(origin=(0.0, 0.0, 0.0), h=1.0):
    g = np.zeros(1, GEOMETRY)
    g["world_origin"]["x"], g["world_origin"]["y"], g["world_origin"]["z"] = origin
    g["cell_width"] = h
    return g


def int3(rows):
    a = np.zeros(len(rows), INT3)
    for i, (x, y, z) in enumerate(rows):
        a[i] = (x, y, z)
    return a


def vec3(x, y, z):
    v = np.zeros(1, VEC3)
    v[0] = (x, y, z)
    return v


def _vec3_struct():
    class V(C.Structure):
        _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]
    return V


Vec3C = _vec3_struct()


class Int3C(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32)]


class FaceIndexC(C.Structure):
    _fields_ = [("axis", C.c_uint8), ("lattice", Int3C)]


class QuantC(C.Structure):
    _fields_ = [("mass_per_quantum", C.c_double), ("pure_momentum_per_quantum", C.c_double),
                ("declared_particle_capacity", C.c_uint64)]


def compute_bin(geom, pos):
    out = Int3C()
    s = lib().flo_compute_bin(_p(geom), Vec3C(*map(float, pos)), C.byref(out))
    return s, (out.x, out.y, out.z)


def sort_blocks(geom, blocks):
    out = np.zeros(len(blocks), INT3)
    s = lib().flo_grid_sort_blocks(_p(geom), _p(blocks), C.c_size_t(len(blocks)), _p(out))
    return s, out


def face_slot(blocks, axis, lattice):
    out = C.c_uint64()
    f = FaceIndexC(axis, Int3C(*lattice))
    s = lib().flo_face_slot(_p(blocks), C.c_size_t(len(blocks)), f, C.byref(out))
    return s, out.value


def slot_face(blocks, slot):
    out = FaceIndexC()
    s = lib().flo_slot_face(_p(blocks), C.c_size_t(len(blocks)), C.c_uint64(slot), C.byref(out))
    return s, (out.axis, (out.lattice.x, out.lattice.y, out.lattice.z))


def face_world_anchor(geom, axis, lattice):
    out = Vec3C()
    s = lib().flo_face_world_anchor(_p(geom), FaceIndexC(axis, Int3C(*lattice)), C.byref(out))
    return s, (out.x, out.y, out.z)


def owned_stencil(geom, blocks, axis, sorted_pos, trial_pos):
    out = np.zeros(27, STENCIL_SAMPLE)
    s = lib().flo_owned_stencil(_p(geom), _p(blocks), C.c_size_t(len(blocks)), C.c_int(axis),
                                Vec3C(*map(float, sorted_pos)), Vec3C(*map(float, trial_pos)), _p(out))
    return s, out


def deposit(geom, blocks, particles, out=None):
    nf = 192 * len(blocks)
    if out is None:
        out = np.zeros(nf, FACE_DEPOSIT)
    s = lib().flo_deposit(_p(geom), _p(blocks), C.c_size_t(len(blocks)), _p(particles), C.c_size_t(len(particles)),
                          _p(out), C.c_size_t(len(out)))
    return s, out


def deposit_quanta(geom, blocks, particles, mass_q, mom_q, cap, max_capacity=1 << 20, out=None, threads=1):
    if out is None:
        out = np.zeros(192 * len(blocks), FACE_QUANTA)
    s = lib().flo_deposit_quanta(_p(geom), _p(blocks), C.c_size_t(len(blocks)), _p(particles),
                                 C.c_size_t(len(particles)), QuantC(mass_q, mom_q, cap), C.c_uint64(max_capacity),
                                 _p(out), C.c_size_t(len(out)), C.c_int(threads))
    return s, out


def deposit_compatible(geom, blocks, particles, tags, mass_q, mom_q, cap, max_capacity=1 << 20, stencils=True):
    """DepositCore: returns (status, face quanta, compatible stencils [n][3][27] or None)."""
    out = np.zeros(192 * len(blocks), FACE_QUANTA)
    st = np.zeros((len(particles), 3, 27), COMPATIBLE_SAMPLE) if stencils else None
    s = lib().flo_deposit_compatible(_p(geom), _p(blocks), C.c_size_t(len(blocks)), _p(particles),
                                     C.c_size_t(len(particles)), _p(tags), C.c_size_t(len(tags)),
                                     QuantC(mass_q, mom_q, cap), C.c_uint64(max_capacity), _p(out),
                                     C.c_size_t(len(out)), _p(st) if st is not None else None)
    return s, out, st


class HistoryC(C.Structure):
    _fields_ = [("velocity_sample_time", C.c_double), ("wall_sample_time", C.c_double),
                ("face_velocities", C.c_void_p), ("validity", C.c_void_p), ("face_tags", C.c_void_p),
                ("wall_samples", C.c_void_p), ("count", C.c_size_t)]


def history(time, velocities, validity, tags, walls, wall_time=None):
    """A LiquidBoundaryVelocityHistory3D view; the arrays must outlive the call that uses it."""
    h = HistoryC(time, time if wall_time is None else wall_time, _p(velocities), _p(validity), _p(tags), _p(walls),
                 len(velocities))
    h._keep = (velocities, validity, tags, walls)
    return h


def gather_boundary(geom, blocks, particles, old_time, new_time, old_history, new_history):
    """GatherCore: (status, records, [old-PIC, new-PIC, low-mass fallback counts])."""
    out = np.zeros(len(particles), BOUNDARY_GATHER)
    counts = np.zeros(3, np.uint64)
    s = lib().flo_gather_boundary(_p(geom), _p(blocks), C.c_size_t(len(blocks)), _p(particles),
                                  C.c_size_t(len(particles)), C.c_double(old_time), C.c_double(new_time),
                                  C.byref(old_history), C.byref(new_history), _p(out), _p(counts))
    return s, out, counts


def fit_ghost_affine(offsets, velocities):
    """FitLiquidGhostAffine3D: (status, full_rank, value at target, slopes per cell width)."""
    off = np.ascontiguousarray(offsets, dtype=np.float64).reshape(-1, 3)
    vel = np.ascontiguousarray(velocities, dtype=np.float64)
    full, at = C.c_int(0), C.c_double(0.0)
    slopes = np.zeros(3)
    s = lib().flo_fit_ghost_affine(_p(off), _p(vel), C.c_size_t(len(vel)), C.byref(full), C.byref(at), _p(slopes))
    return s, bool(full.value), at.value, slopes


def gather(geom, blocks, particles, old, new, out=None, threads=1):
    if out is None:
        out = np.zeros(len(particles), GATHER_RESULT)
    old = np.ascontiguousarray(old, dtype=np.float64)
    new = np.ascontiguousarray(new, dtype=np.float64)
    s = lib().flo_gather(_p(geom), _p(blocks), C.c_size_t(len(blocks)), _p(particles), C.c_size_t(len(particles)),
                         _p(old), _p(new), C.c_size_t(len(old)), _p(out), C.c_int(threads))
    return s, out


def normalize(quanta, mass_q, mom_q, min_mass=0.0):
    nf = len(quanta)
    m, u, v = np.zeros(nf), np.zeros(nf), np.zeros(nf, np.uint8)
    s = lib().flo_normalize(_p(quanta), C.c_size_t(nf), QuantC(mass_q, mom_q, 0), C.c_double(min_mass), _p(m), _p(u),
                            _p(v))
    return s, m, u, v


def apply_forces(blocks, old, accel, dt):
    out = np.zeros(len(old))
    old = np.ascontiguousarray(old, dtype=np.float64)
    s = lib().flo_apply_forces(_p(blocks), C.c_size_t(len(blocks)), _p(old), C.c_size_t(len(old)),
                               Vec3C(*map(float, accel)), C.c_double(dt), _p(out))
    return s, out


def sort_particles(geom, positions):
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(len(pos), SORT_ENTRY)
    s = lib().flo_sort_particles(_p(geom), _p(pos), C.c_size_t(len(pos)), _p(out))
    return s, out


def activate_blocks(geom, positions, cells=None, capacity=1 << 24):
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    cells = np.zeros(0, INT3) if cells is None else cells
    out = np.zeros(capacity, INT3)
    cnt = C.c_size_t(capacity)
    s = lib().flo_activate_blocks(_p(geom), _p(pos), C.c_size_t(len(pos)), _p(cells), C.c_size_t(len(cells)), _p(out),
                                  C.byref(cnt))
    return s, out[:cnt.value].copy()


def update_options(cell_width, step_duration, blend, friction=0.0, penetration=0.0, obstacle=None):
    """Particle update options; obstacle: None, or a tuple as made by halfspace() or slab()."""
    o = np.zeros(1, UPDATE_OPTIONS)
    o[0]["cell_width"], o[0]["step_duration"], o[0]["particle_energy_blend"] = cell_width, step_duration, blend
    o[0]["friction_coefficient"], o[0]["penetration_tolerance"] = friction, penetration
    if obstacle is not None:
        kind, n, c, w, v = obstacle[:5]
        ob = o["obstacle"]
        ob["kind"], ob["coordinate"], ob["half_width"] = kind, c, w
        for k, a, b in zip("xyz", n, v):
            ob["unit_normal"][k], ob["wall_velocity"][k] = a, b
        if len(obstacle) > 5:
            for name, value in zip(TANK.names, obstacle[5]):
                ob["tank"][name] = value
        o["obstacle"] = ob
    return o


def update_particles(options, gathered, distances, particles):
    """flo_update_particles on a copy: (status, particles after the call)."""
    out = np.ascontiguousarray(particles).copy()
    g = np.ascontiguousarray(gathered)
    d = np.ascontiguousarray(distances, dtype=np.float64)
    s = lib().flo_update_particles(_p(options), _p(g), _p(d), C.c_size_t(len(out)), _p(out))
    return s, out


class FrozenViewC(C.Structure):
    _fields_ = [("blocks", C.c_void_p), ("block_count", C.c_uint64), ("origin", C.c_double * 3),
                ("cell_width", C.c_double), ("center_distances", C.c_void_p), ("center_count", C.c_uint64),
                ("cells", C.c_void_p), ("cell_count", C.c_uint64), ("faces", C.c_void_p), ("face_count", C.c_uint64),
                ("step_duration", C.c_double), ("liquid_density", C.c_double), ("surface_tension", C.c_double),
                ("capillary_step_bound", C.c_double)]


class PressureFieldsC(C.Structure):
    _fields_ = [("cells", C.c_void_p), ("cell_count", C.c_uint64), ("predicted", C.c_void_p),
                ("predicted_count", C.c_uint64), ("ambient_pressure", C.c_double), ("wetted", C.c_void_p),
                ("wetted_count", C.c_uint64), ("air", C.c_void_p), ("air_count", C.c_uint64),
                ("interfaces", C.c_void_p), ("interface_count", C.c_uint64)]


class LinkProblem:
    """A frozen view and its pressure fields as numpy arrays (the C structures point into them)."""

    def __init__(self, geom, blocks, centers, cells, faces, step=0.01, density=1000.0, sigma=0.0, fields=None,
                 predicted=None, ambient=0.0, wetted=None, air=None, interfaces=None):
        self.geom, self.blocks = geom, np.ascontiguousarray(blocks)
        self.centers, self.cells, self.faces = (np.ascontiguousarray(a) for a in (centers, cells, faces))
        self.step, self.density, self.sigma, self.ambient = step, density, sigma, ambient
        if fields is None:
            fields = np.zeros(len(cells), CELL_FIELDS)
            fields["bulk_modulus"] = np.inf
        self.fields = np.ascontiguousarray(fields)
        self.predicted = np.zeros(len(faces)) if predicted is None else np.ascontiguousarray(predicted, dtype=np.float64)
        self.wetted = np.zeros(0, OPTIONAL_FLUX) if wetted is None else np.ascontiguousarray(wetted)
        self.air = np.zeros(0, AIR_PRESSURE) if air is None else np.ascontiguousarray(air)
        self.interfaces = np.zeros(0, INTERNAL_INTERFACE) if interfaces is None else np.ascontiguousarray(interfaces)

    def c_view(self):
        o = self.geom[0]["world_origin"]
        return FrozenViewC(_p(self.blocks), len(self.blocks), (C.c_double * 3)(o["x"], o["y"], o["z"]),
                           self.geom[0]["cell_width"], _p(self.centers), len(self.centers), _p(self.cells),
                           len(self.cells), _p(self.faces), len(self.faces), self.step, self.density, self.sigma, 0.0)

    def c_fields(self):
        return PressureFieldsC(_p(self.fields), len(self.fields), _p(self.predicted), len(self.predicted), self.ambient,
                               _p(self.wetted), len(self.wetted), _p(self.air), len(self.air), _p(self.interfaces),
                               len(self.interfaces))


def build_pressure_links(problem):
    """flo_build_pressure_links: (status, cells, faces, lower-theta count)."""
    cells = np.zeros(len(problem.cells), PRESSURE_CELL)
    faces = np.zeros(len(problem.faces), PRESSURE_FACE)
    lower = C.c_uint64()
    v, f = problem.c_view(), problem.c_fields()
    s = lib().flo_build_pressure_links(C.byref(v), C.byref(f), _p(cells), _p(faces), C.byref(lower))
    return s, cells, faces, lower.value


OBSTACLE_ABSENT, OBSTACLE_HALFSPACE, OBSTACLE_SLAB, OBSTACLE_TANK = 0, 1, 2, 3


class ObstacleC(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("pad", C.c_uint32), ("unit_normal", C.c_double * 3), ("coordinate", C.c_double),
                ("half_width", C.c_double), ("wall_velocity", C.c_double * 3), ("tank", C.c_double * 6)]


class FreezeInputsC(C.Structure):
    _fields_ = [("blocks", C.c_void_p), ("block_count", C.c_uint64), ("origin", C.c_double * 3),
                ("cell_width", C.c_double), ("cell_indices", C.c_void_p), ("cell_count", C.c_uint64),
                ("center_distances", C.c_void_p), ("center_count", C.c_uint64), ("corner_distances", C.c_void_p),
                ("corner_count", C.c_uint64), ("obstacle", ObstacleC), ("step_duration", C.c_double),
                ("liquid_density", C.c_double), ("surface_tension", C.c_double), ("exact_curvature", C.c_double),
                ("has_exact_curvature", C.c_uint8), ("pad", C.c_uint8 * 7), ("max_cells", C.c_uint64),
                ("max_faces", C.c_uint64), ("solid_corners", C.c_void_p), ("solid_corner_count", C.c_uint64),
                ("face_normal_velocities", C.c_void_p), ("face_velocity_count", C.c_uint64),
                ("has_sampled_solid", C.c_uint8), ("pad2", C.c_uint8 * 7)]


assert C.sizeof(ObstacleC) == 120


def halfspace(normal, plane, velocity=(0.0, 0.0, 0.0)):
    return (OBSTACLE_HALFSPACE, tuple(normal), plane, 0.0, tuple(velocity))


def slab(normal, center, half_width, velocity=(0.0, 0.0, 0.0)):
    return (OBSTACLE_SLAB, tuple(normal), center, half_width, tuple(velocity))


def tank(start_low_x, high_x, floor_y, low_z, high_z, piston_velocity_x=0.0):
    """An open tank (particle update only); the tuple's last element holds its six fields."""
    return (OBSTACLE_TANK, (0.0, 0.0, 0.0), 0.0, 0.0, (0.0, 0.0, 0.0),
            (start_low_x, high_x, floor_y, low_z, high_z, piston_velocity_x))


class FreezeProblem:
    """Freeze inputs as numpy arrays (the C structure points into them). obstacle: None, or a tuple
    (kind, unit normal, coordinate, half width, wall velocity) as made by halfspace() or slab()."""

    def __init__(self, geom, blocks, candidates, centers, corners, obstacle=None, step=0.01, density=1000.0,
                 sigma=0.0, exact_curvature=None, max_cells=0, max_faces=0, solid_corners=None,
                 face_normal_velocities=None, split_rule=1):
        """solid_corners: SOLID_CORNER records selecting the sampled-solid mode (None: no sampled solid);
        face_normal_velocities: None (empty span) or one wall normal velocity per owned slot;
        split_rule: the wet-piece refusal, 1 the reference's, 2 only in cells with a corner in the solid,
        3 none."""
        self.split_rule = split_rule
        self.geom, self.blocks = geom, np.ascontiguousarray(blocks)
        self.candidates, self.centers, self.corners = (np.ascontiguousarray(a) for a in (candidates, centers, corners))
        self.obstacle = obstacle
        self.step, self.density, self.sigma, self.exact_curvature = step, density, sigma, exact_curvature
        self.max_cells, self.max_faces = max_cells, max_faces
        self.solid_corners = None if solid_corners is None else np.ascontiguousarray(solid_corners, dtype=SOLID_CORNER)
        self.face_normal_velocities = (None if face_normal_velocities is None
                                       else np.ascontiguousarray(face_normal_velocities, dtype=np.float64))

    def c_obstacle(self):
        if self.obstacle is None:
            return ObstacleC()
        kind, n, c, w, v = self.obstacle[:5]
        t = self.obstacle[5] if len(self.obstacle) > 5 else (0.0,) * 6
        return ObstacleC(kind, 0, (C.c_double * 3)(*n), c, w, (C.c_double * 3)(*v), (C.c_double * 6)(*t))

    def c_inputs(self):
        o = self.geom[0]["world_origin"]
        k = self.exact_curvature
        return FreezeInputsC(_p(self.blocks), len(self.blocks), (C.c_double * 3)(o["x"], o["y"], o["z"]),
                             self.geom[0]["cell_width"], _p(self.candidates), len(self.candidates), _p(self.centers),
                             len(self.centers), _p(self.corners), len(self.corners), self.c_obstacle(), self.step,
                             self.density, self.sigma, 0.0 if k is None else k, 0 if k is None else 1,
                             (C.c_uint8 * 7)(), self.max_cells, self.max_faces,
                             _p(self.solid_corners), 0 if self.solid_corners is None else len(self.solid_corners),
                             _p(self.face_normal_velocities),
                             0 if self.face_normal_velocities is None else len(self.face_normal_velocities),
                             0 if self.solid_corners is None else self.split_rule,
                             (C.c_uint8 * 7)())


def measure_sampled_cell(width, liquid, solid, analytic=None):
    """flo_measure_sampled_cell: (status, wet volume, six open areas); analytic = (nx, ny, nz, coordinate)."""
    l = np.ascontiguousarray(liquid, dtype=np.float64)
    so = np.ascontiguousarray(solid, dtype=np.float64)
    a = None if analytic is None else np.ascontiguousarray(analytic, dtype=np.float64)
    wet, open_area = C.c_double(), np.zeros(6)
    s = lib().flo_measure_sampled_cell(C.c_double(width), _p(l), _p(so), _p(a), C.byref(wet), _p(open_area))
    return s, wet.value, open_area


def sampled_face_weights(width, solid, analytic=None):
    """flo_sampled_face_weights: (status, four corner weights); analytic = (n0, n1, 0, coordinate) in face
    coordinates."""
    so = np.ascontiguousarray(solid, dtype=np.float64)
    a = None if analytic is None else np.ascontiguousarray(analytic, dtype=np.float64)
    w = np.zeros(4)
    s = lib().flo_sampled_face_weights(C.c_double(width), _p(so), _p(a), _p(w))
    return s, w


def freeze_geometry(problem):
    """flo_freeze_geometry: (status, frozen cells, frozen faces, capillary step bound)."""
    n = len(problem.candidates)
    cells = np.zeros(n, FROZEN_CELL)
    faces = np.zeros(6 * n, FROZEN_FACE)
    nc, nf, bound = C.c_uint64(), C.c_uint64(), C.c_double()
    inputs = problem.c_inputs()
    s = lib().flo_freeze_geometry(C.byref(inputs), _p(cells), C.byref(nc), _p(faces), C.byref(nf), C.byref(bound))
    if s != OK:
        return s, None, None, None
    return s, cells[:nc.value].copy(), faces[:nf.value].copy(), bound.value


class TopologyHistoryC(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("cell_width", C.c_double), ("epoch", C.c_uint64),
                ("blocks", C.c_void_p), ("block_count", C.c_uint64), ("cells", C.c_void_p), ("cell_count", C.c_uint64),
                ("faces", C.c_void_p), ("face_count", C.c_uint64)]


def topology_history(geom, epoch, blocks, cells, faces):
    """The previous state's topology records; the arrays must outlive the returned structure."""
    g = geom[0]
    o = g["world_origin"]
    return TopologyHistoryC((C.c_double * 3)(o["x"], o["y"], o["z"]), g["cell_width"], epoch, _p(blocks), len(blocks),
                            _p(cells), len(cells), _p(faces), len(faces))


def topology_capacities(particles=1 << 20, blocks=1 << 16, cells=1 << 20, faces=192 << 16):
    c = np.zeros(1, TOPOLOGY_CAPACITIES)
    c[0] = (particles, blocks, cells, faces)
    return c


def prepare_topology(geom, positions, cells=None, previous=None, capacities=None):
    """flo_prepare_topology: (status, order, blocks, cells, faces); previous is a TopologyHistoryC or None."""
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    cells = np.zeros(0, INT3) if cells is None else np.ascontiguousarray(cells)
    caps = topology_capacities() if capacities is None else capacities
    c = caps[0]
    nb_cap = min(int(c["blocks"]), 27 * 3 * len(pos) + 4 * len(cells) + 1)
    order = np.zeros(len(pos), SORT_ENTRY)
    ob = np.zeros(nb_cap, BLOCK_STATE)
    oc = np.zeros(min(int(c["cells"]), len(cells) + 1), CELL_STATE)
    of = np.zeros(192 * nb_cap, FACE_STATE)
    nb, nc, nf = C.c_uint64(), C.c_uint64(), C.c_uint64()
    s = lib().flo_prepare_topology(_p(geom), _p(caps), _p(pos), C.c_size_t(len(pos)), _p(cells), C.c_size_t(len(cells)),
                                   None if previous is None else C.byref(previous), _p(order), _p(ob), C.byref(nb),
                                   _p(oc), C.byref(nc), _p(of), C.byref(nf))
    if s != OK:
        return s, None, None, None, None
    return s, order, ob[:nb.value].copy(), oc[:nc.value].copy(), of[:nf.value].copy()


def pressure_options(max_iterations=256, tolerance_scale=1.0, relative_target=0.0, levels=4, sweeps=2,
                     coarse_sweeps=16, fp32=True, coarse_scale=0.5, reference=False):
    """Solver options; reference=True selects the reference's aggregation multigrid (Jacobi 2/3, FP64
    Cholesky terminal blocks, restarts only at recursive convergence) instead of the geometric V-cycle."""
    o = np.zeros(1, PRESSURE_OPTIONS)
    o[0] = (max_iterations, tolerance_scale, relative_target, levels, sweeps, coarse_sweeps, int(fp32), coarse_scale,
            int(reference), 0)
    return o


def project(geom, blocks, dt, cells, faces, options):
    nc, nf = len(cells), len(faces)
    p, u, b, r, d = (np.zeros(nc), np.zeros(nf), np.zeros(nc), np.zeros(nc), np.zeros(nc))
    gauges = np.zeros(nc, np.uint64)
    diag = np.zeros(1, PRESSURE_DIAGNOSTICS)
    s = lib().flo_project(_p(geom), _p(blocks), C.c_size_t(len(blocks)), C.c_double(dt), _p(cells), C.c_size_t(nc),
                          _p(faces), C.c_size_t(nf), _p(options), _p(p), _p(u), _p(b), _p(r), _p(d), _p(gauges),
                          _p(diag))
    g = gauges[: int(diag["gauge_count"][0])].copy()
    return s, dict(pressures=p, velocities=u, rhs=b, residuals=r, defects=d, gauges=g, diagnostics=diag[0])


def evaluate(geom, blocks, dt, cells, faces, pressures):
    nc, nf = len(cells), len(faces)
    u, r, d = np.zeros(nf), np.zeros(nc), np.zeros(nc)
    diag = np.zeros(1, PRESSURE_DIAGNOSTICS)
    pressures = np.ascontiguousarray(pressures, dtype=np.float64)
    s = lib().flo_evaluate(_p(geom), _p(blocks), C.c_size_t(len(blocks)), C.c_double(dt), _p(cells), C.c_size_t(nc),
                           _p(faces), C.c_size_t(nf), _p(pressures), _p(u), _p(r), _p(d), _p(diag))
    return s, dict(velocities=u, residuals=r, defects=d, diagnostics=diag[0])


def compiler():
    lib().flo_compiler.restype = C.c_char_p
    return lib().flo_compiler().decode()


HIERARCHY_LEVEL = np.dtype([("level_count", "u8"), ("unknowns", "u8"), ("edges", "u8"), ("terminal_blocks", "u8"),
                            ("max_terminal", "u8")])


def reference_hierarchy(geom, blocks, dt, cells, faces, level):
    """Level `level` of the reference aggregation hierarchy: a dict of its counts and arrays."""
    cap = len(cells)
    info = np.zeros(1, HIERARCHY_LEVEL)
    coords = np.zeros((cap, 3), np.int32)
    anchors, diag = np.zeros(cap), np.zeros(cap)
    lo, hi, k = np.zeros(3 * cap, np.uint64), np.zeros(3 * cap, np.uint64), np.zeros(3 * cap)
    agg, u2c = np.zeros(cap, np.uint64), np.zeros(cap, np.uint64)
    s = lib().flo_reference_hierarchy(_p(geom), _p(blocks), C.c_size_t(len(blocks)), C.c_double(dt), _p(cells),
                                      C.c_size_t(len(cells)), _p(faces), C.c_size_t(len(faces)), C.c_uint64(level),
                                      C.c_uint64(cap), _p(info), _p(coords), _p(anchors), _p(diag), _p(lo), _p(hi),
                                      _p(k), _p(agg), _p(u2c))
    if s != OK:
        return s, None
    i = info[0]
    n, ne = int(i["unknowns"]), int(i["edges"])
    return s, dict(level_count=int(i["level_count"]), terminal_blocks=int(i["terminal_blocks"]),
                   max_terminal=int(i["max_terminal"]), coords=coords[:n], anchors=anchors[:n], diagonal=diag[:n],
                   edges=np.stack([lo[:ne], hi[:ne]], 1).astype(np.int64), conductances=k[:ne],
                   aggregates=agg[:n], unknown_to_cell=u2c[:n])


def reference_vcycle(geom, blocks, dt, cells, faces, r):
    """z = M r for the reference V-cycle in FP64 over the level-0 unknowns."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.zeros_like(r)
    s = lib().flo_reference_vcycle(_p(geom), _p(blocks), C.c_size_t(len(blocks)), C.c_double(dt), _p(cells),
                                   C.c_size_t(len(cells)), _p(faces), C.c_size_t(len(faces)), _p(r), _p(z))
    return s, z


# ---- learned preconditioner network (unet.cc) ----
def unet_param_count():
    lib().flo_unet_param_count.restype = C.c_size_t
    return int(lib().flo_unet_param_count())


def unet_forward(params, inputs):
    params = np.ascontiguousarray(params, dtype=np.float64)
    inputs = np.ascontiguousarray(inputs, dtype=np.float64).reshape(-1, 4, 512)
    out = np.zeros((len(inputs), 512))
    s = lib().flo_unet_forward(_p(params), C.c_size_t(len(inputs)), _p(inputs), _p(out))
    assert s == OK
    return out


def unet_loss_grad(params, inputs, targets, want_input_grad=False):
    params = np.ascontiguousarray(params, dtype=np.float64)
    inputs = np.ascontiguousarray(inputs, dtype=np.float64).reshape(-1, 4, 512)
    targets = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 512)
    loss = C.c_double()
    gp = np.zeros(len(params))
    gi = np.zeros(inputs.shape) if want_input_grad else None
    s = lib().flo_unet_loss_grad(_p(params), C.c_size_t(len(inputs)), _p(inputs), _p(targets), C.byref(loss), _p(gp),
                                  _p(gi))
    assert s == OK
    return loss.value, gp, gi


def unet_init(seed=0):
    """He-normal weights, zero biases, in the parameter layout of unet.cc."""
    layers = [(4, 8), (8, 8), (8, 16), (16, 16), (16, 32), (32, 32), (48, 16), (16, 16), (24, 8), (8, 1)]
    rng = np.random.default_rng(seed)
    parts = []
    for cin, cout in layers:
        parts.append(rng.normal(0.0, np.sqrt(2.0 / (27 * cin)), cout * cin * 27))
        parts.append(np.zeros(cout))
    p = np.concatenate(parts)
    assert len(p) == unet_param_count()
    return p


def viscosity_cells(indices, viscosities):
    """Viscosity cell records from (n, 3) indices and n viscosities."""
    c = np.zeros(len(indices), VISCOSITY_CELL)
    idx = np.asarray(indices, dtype=np.int32).reshape(-1, 3)
    c["cell_index"]["x"], c["cell_index"]["y"], c["cell_index"]["z"] = idx[:, 0], idx[:, 1], idx[:, 2]
    c["dynamic_viscosity"] = viscosities
    return c


def assemble_full_cell_strain(geom, blocks, cells):
    """AssembleFullCellMacStrain3D (native/src/viscosity.cc:762-946): (status, strain rows)."""
    out = np.zeros(max(6 * len(cells), 1), STRAIN_ROW)
    count = C.c_size_t(0)
    s = lib().flo_assemble_full_cell_strain(_p(geom), _p(blocks), C.c_size_t(len(blocks)), _p(cells),
                                            C.c_size_t(len(cells)), _p(out), C.c_size_t(len(out)), C.byref(count))
    return s, out[:count.value] if s == OK else None


def solve_viscous_block(rows, inertia, dt, rhs, max_iterations=256, relative_tolerance=1e-10):
    """SolveMacViscousVelocityBlock3D (native/src/viscosity.cc:1331-1650): (status, velocity, iterations,
    max residual)."""
    rows = np.ascontiguousarray(rows, dtype=STRAIN_ROW)
    inertia = np.ascontiguousarray(inertia, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    v = np.zeros(len(inertia))
    it = C.c_uint64(0)
    mr = C.c_double(0.0)
    s = lib().flo_solve_viscous_block(_p(rows), C.c_size_t(len(rows)), _p(inertia), C.c_size_t(len(inertia)),
                                      C.c_double(dt), _p(rhs), C.c_uint64(max_iterations),
                                      C.c_double(relative_tolerance), _p(v), C.byref(it), C.byref(mr))
    return s, v, it.value, mr.value


VISCOUS_DIAGNOSTICS = np.dtype([("pressure_iterations", "u8"), ("constraint_residual_l2", "f8"),
                                ("initial_constraint_residual_l2", "f8"), ("max_cell_constraint_defect", "f8"),
                                ("max_momentum_residual", "f8"), ("inward_prescribed_wall_flux", "f8"),
                                ("velocity_solves", "u8")])


def project_viscous(geom, blocks, dt, cells, faces, rows, max_iterations=256):
    """ProjectMacViscousPressure3D (native/src/mac_viscous_pressure_3d.cc:151-459): (status, dict of
    pressures, face velocities, cell constraint defects and diagnostics)."""
    nc, nf = len(cells), len(faces)
    rows = np.ascontiguousarray(rows, dtype=STRAIN_ROW)
    p, u, d = np.zeros(nc), np.zeros(nf), np.zeros(nc)
    diag = np.zeros(1, VISCOUS_DIAGNOSTICS)
    s = lib().flo_project_viscous(_p(geom), _p(blocks), C.c_size_t(len(blocks)), C.c_double(dt), _p(cells),
                                  C.c_size_t(nc), _p(faces), C.c_size_t(nf), _p(rows), C.c_size_t(len(rows)),
                                  C.c_uint64(max_iterations), _p(p), _p(u), _p(d), _p(diag))
    return s, dict(pressures=p, velocities=u, defects=d, diagnostics=diag[0])


def strain_row(terms, weight, first_axis=0, second_axis=0):
    """One strain row from [(slot, coefficient), ...] and its quadrature weight."""
    r = np.zeros(1, STRAIN_ROW)
    r["first_axis"], r["second_axis"] = first_axis, second_axis
    for i, (slot, c) in enumerate(terms):
        r["terms"][0, i] = (slot, c)
    r["term_count"] = len(terms)
    r["quadrature_weight"] = weight
    return r[0]


def particle_field(geom, positions, radius, candidate_capacity=1 << 24):
    """flo_particle_field: (status, candidate cells, centre samples, corner samples) of the particle union."""
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    n = len(pos)
    nc, ncen, ncor = C.c_uint64(), C.c_uint64(), C.c_uint64()
    s = lib().flo_particle_field(_p(geom), _p(pos), C.c_size_t(n), C.c_double(radius), C.c_uint64(candidate_capacity),
                                 None, C.byref(nc), None, C.c_uint64(0), C.byref(ncen), None, C.c_uint64(0),
                                 C.byref(ncor))
    if s != OK:
        return s, None, None, None
    cells = np.zeros(max(nc.value, 1), INT3)
    centres = np.zeros(max(ncen.value, 1), CENTER_DISTANCE)
    corners = np.zeros(max(ncor.value, 1), CENTER_DISTANCE)
    s = lib().flo_particle_field(_p(geom), _p(pos), C.c_size_t(n), C.c_double(radius), C.c_uint64(candidate_capacity),
                                 _p(cells), C.byref(nc), _p(centres), C.c_uint64(len(centres)), C.byref(ncen),
                                 _p(corners), C.c_uint64(len(corners)), C.byref(ncor))
    if s != OK:
        return s, None, None, None
    return s, cells[:nc.value].copy(), centres[:ncen.value].copy(), corners[:ncor.value].copy()
