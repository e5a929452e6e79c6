"""ctypes binding of fluidity_amd/libfluidity_hip.so (the C-ABI of include/fluidity_hip.h).

Loading fails loudly when the library is missing: there is no CPU fallback behind this module.
Every function returns the C status code first, mirroring the reference's absl::Status returns.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import records as R

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfluidity_hip.so")

OK, INVALID_ARGUMENT, RESOURCE_EXHAUSTED, FAILED_PRECONDITION, ABORTED, OUT_OF_RANGE, UNIMPLEMENTED, INTERNAL = (
    0, 3, 8, 9, 10, 11, 12, 13)


class Vec3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class Quantization(C.Structure):
    _fields_ = [("mass_per_quantum", C.c_double), ("pure_momentum_per_quantum", C.c_double),
                ("declared_particle_capacity", C.c_uint64)]


class Measurement(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("copy_ms", C.c_double), ("launches", C.c_uint32),
                ("architecture", C.c_char * 28)]


class Capacities(C.Structure):
    _fields_ = [("particles", C.c_uint64), ("blocks", C.c_uint64), ("lookup_blocks", C.c_uint64),
                ("pressure_cells", C.c_uint64), ("pressure_faces", C.c_uint64)]


class StepReport(C.Structure):
    _fields_ = [("sort_ms", C.c_double), ("p2g_ms", C.c_double), ("normalize_ms", C.c_double),
                ("projection_ms", C.c_double), ("g2p_ms", C.c_double), ("total_ms", C.c_double),
                ("p2g_kernel_ms", C.c_double), ("g2p_kernel_ms", C.c_double),
                ("particles", C.c_uint64), ("blocks", C.c_uint64), ("faces", C.c_uint64),
                ("pressure_cells", C.c_uint64), ("nonempty_bins", C.c_uint64),
                ("pressure", C.c_byte * R.PRESSURE_DIAGNOSTICS.itemsize), ("status", C.c_int32), ("pad", C.c_int32)]

    def diagnostics(self):
        return np.frombuffer(bytes(self.pressure), dtype=R.PRESSURE_DIAGNOSTICS)[0]


EXPORTS = ("fl_create", "fl_destroy", "fl_device_name", "fl_reserve", "fl_bin_sort", "fl_activate", "fl_prepare_topology", "fl_update_particles", "fl_build_pressure_links", "fl_freeze_geometry", "fl_p2g", "fl_p2g_quanta", "fl_p2g_compatible", "fl_g2p_boundary",
           "fl_normalize_forces", "fl_apply_forces", "fl_g2p", "fl_project", "fl_scene_upload", "fl_scene_prepare", "fl_step",
           "fl_scene_download", "fl_probe_copy", "fl_learn_param_count", "fl_learn_setup", "fl_learn_set_params", "fl_learn_get_params",
           "fl_learn_forward", "fl_learn_infer", "fl_learn_loss_grad", "fl_learn_add_pairs", "fl_learn_pairs", "fl_learn_collect", "fl_learn_eval",
           "fl_learn_train", "fl_normalize_forces_static", "fl_time_projection_kernels", "fl_assemble_full_cell_strain", "fl_solve_viscous_block", "fl_substep",
           "fl_substep_download", "fl_sample_particle_field", "fl_project_device_built",
           "fl_set_transfer_precision", "fl_project_viscous")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no CPU fallback exists)")
        _lib = C.CDLL(LIB_PATH)
        _lib.fl_destroy.restype = None
        _lib.fl_learn_param_count.restype = C.c_size_t
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _vec3(v):
    return Vec3(*map(float, v))


class TopologyHistory(C.Structure):
    """fl_topology_history; keep the record arrays alive while it is in use (held in _keep)."""
    _fields_ = [("origin", C.c_double * 3), ("cell_width", C.c_double), ("epoch", C.c_uint64),
                ("blocks", C.c_void_p), ("block_count", C.c_uint64), ("cells", C.c_void_p), ("cell_count", C.c_uint64),
                ("faces", C.c_void_p), ("face_count", C.c_uint64)]


def topology_history(geometry, epoch, blocks, cells, faces):
    g = geometry[0]
    o = g["world_origin"]
    h = TopologyHistory((C.c_double * 3)(o["x"], o["y"], o["z"]), g["cell_width"], epoch, _p(blocks), len(blocks),
                        _p(cells), len(cells), _p(faces), len(faces))
    h._keep = (blocks, cells, faces)
    return h


def topology_capacities(particles=1 << 20, blocks=1 << 16, cells=1 << 20, faces=192 << 16):
    c = np.zeros(1, R.TOPOLOGY_CAPACITIES)
    c[0] = (particles, blocks, cells, faces)
    return c


class Obstacle(C.Structure):
    """fl_obstacle: kind 0 absent, 1 half-space, 2 slab, 3 open tank (particle update only)."""
    _fields_ = [("kind", C.c_uint32), ("pad", C.c_uint32), ("unit_normal", C.c_double * 3), ("coordinate", C.c_double),
                ("half_width", C.c_double), ("wall_velocity", C.c_double * 3), ("tank", C.c_double * 6)]


class FreezeInputs(C.Structure):
    _fields_ = [("blocks", C.c_void_p), ("block_count", C.c_uint64), ("origin", C.c_double * 3),
                ("cell_width", C.c_double), ("cell_indices", C.c_void_p), ("cell_count", C.c_uint64),
                ("center_distances", C.c_void_p), ("center_count", C.c_uint64), ("corner_distances", C.c_void_p),
                ("corner_count", C.c_uint64), ("obstacle", Obstacle), ("step_duration", C.c_double),
                ("liquid_density", C.c_double), ("surface_tension", C.c_double), ("exact_curvature", C.c_double),
                ("has_exact_curvature", C.c_uint8), ("pad", C.c_uint8 * 7), ("max_cells", C.c_uint64),
                ("max_faces", C.c_uint64), ("solid_corners", C.c_void_p), ("solid_corner_count", C.c_uint64),
                ("face_normal_velocities", C.c_void_p), ("face_velocity_count", C.c_uint64),
                ("has_sampled_solid", C.c_uint8), ("pad2", C.c_uint8 * 7)]


class FrozenView(C.Structure):
    """fl_frozen_view."""
    _fields_ = [("blocks", C.c_void_p), ("block_count", C.c_uint64), ("origin", C.c_double * 3),
                ("cell_width", C.c_double), ("center_distances", C.c_void_p), ("center_count", C.c_uint64),
                ("cells", C.c_void_p), ("cell_count", C.c_uint64), ("faces", C.c_void_p), ("face_count", C.c_uint64),
                ("step_duration", C.c_double), ("liquid_density", C.c_double), ("surface_tension", C.c_double),
                ("capillary_step_bound", C.c_double)]


class PressureFields(C.Structure):
    """fl_pressure_fields."""
    _fields_ = [("cells", C.c_void_p), ("cell_count", C.c_uint64), ("predicted", C.c_void_p),
                ("predicted_count", C.c_uint64), ("ambient_pressure", C.c_double), ("wetted", C.c_void_p),
                ("wetted_count", C.c_uint64), ("air", C.c_void_p), ("air_count", C.c_uint64),
                ("interfaces", C.c_void_p), ("interface_count", C.c_uint64)]


class VelocityHistory(C.Structure):
    _fields_ = [("velocity_sample_time", C.c_double), ("wall_sample_time", C.c_double),
                ("face_velocities", C.c_void_p), ("validity", C.c_void_p), ("face_tags", C.c_void_p),
                ("wall_samples", C.c_void_p), ("count", C.c_size_t)]


def velocity_history(time, velocities, validity, tags, walls, wall_time=None):
    """fl_velocity_history over caller arrays (kept alive by the returned object)."""
    h = VelocityHistory(time, time if wall_time is None else wall_time, velocities.ctypes.data,
                        validity.ctypes.data, tags.ctypes.data, walls.ctypes.data, len(velocities))
    h._keep = (velocities, validity, tags, walls)
    return h


class KernelTime(C.Structure):
    _fields_ = [("kernel", C.c_uint32), ("level", C.c_uint32), ("slots", C.c_uint64), ("ms", C.c_double)]


KERNEL_NAMES = ("k_smooth", "k_residual_red", "k_restrict_red", "k_prolong", "k_block_smooth_pre",
                "k_block_smooth_post", "k_residual_all", "k_restrict_all", "k_tail_vcycle", "k_direction_apply",
                "k_update_xr", "k_dot", "k_evaluate")


class Context:
    """One device context (fl_create/fl_destroy); all device memory is sized by the capacities."""

    def __init__(self, particles=1 << 16, blocks=1 << 14, lookup_blocks=1 << 20, pressure_cells=1 << 18,
                 pressure_faces=1 << 20, device=0):
        self.handle = C.c_void_p()
        cap = Capacities(particles, blocks, lookup_blocks, pressure_cells, pressure_faces)
        s = lib().fl_create(C.byref(cap), C.c_int(device), C.byref(self.handle))
        if s != OK:
            raise RuntimeError(f"fl_create failed with status {s}")
        self.last = Measurement()

    def close(self):
        if self.handle:
            lib().fl_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_name(self):
        buf = C.create_string_buffer(64)
        lib().fl_device_name(self.handle, buf, C.c_size_t(64))
        return buf.value.decode()

    def bin_sort(self, geometry, positions):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(pos), R.SORT_ENTRY)
        s = lib().fl_bin_sort(self.handle, _p(geometry), _p(pos), C.c_size_t(len(pos)), _p(out), C.byref(self.last))
        return s, out

    def activate(self, geometry, positions, cells=None, capacity=1 << 22):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        cells = np.zeros(0, R.INT3) if cells is None else np.ascontiguousarray(cells)
        out = np.zeros(capacity, R.INT3)
        cnt = C.c_size_t(capacity)
        s = lib().fl_activate(self.handle, _p(geometry), _p(pos), C.c_size_t(len(pos)), _p(cells),
                              C.c_size_t(len(cells)), _p(out), C.byref(cnt), C.byref(self.last))
        return s, out[: cnt.value].copy()

    def prepare_topology(self, geometry, positions, cells=None, previous=None, capacities=None):
        """fl_prepare_topology: (status, order, blocks, cells, faces); previous is a TopologyHistory or None."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        cells = np.zeros(0, R.INT3) if cells is None else np.ascontiguousarray(cells)
        caps = topology_capacities() if capacities is None else capacities
        c = caps[0]
        nb_cap = min(int(c["blocks"]), 27 * 3 * len(pos) + 4 * len(cells) + 1)
        order = np.zeros(len(pos), R.SORT_ENTRY)
        ob = np.zeros(nb_cap, R.BLOCK_STATE)
        oc = np.zeros(min(int(c["cells"]), len(cells) + 1), R.CELL_STATE)
        of = np.zeros(min(int(c["faces"]), 192 * nb_cap), R.FACE_STATE)
        nb, nc, nf = C.c_uint64(), C.c_uint64(), C.c_uint64()
        s = lib().fl_prepare_topology(self.handle, _p(geometry), _p(caps), _p(pos), C.c_size_t(len(pos)), _p(cells),
                                      C.c_size_t(len(cells)), None if previous is None else C.byref(previous),
                                      _p(order), _p(ob), C.byref(nb), _p(oc), C.byref(nc), _p(of), C.byref(nf),
                                      C.byref(self.last))
        if s != OK:
            return s, None, None, None, None
        return s, order, ob[:nb.value].copy(), oc[:nc.value].copy(), of[:nf.value].copy()

    def build_pressure_links(self, p):
        """fl_build_pressure_links on a problem with numpy record arrays (geom, blocks, centers, cells,
        faces, fields, predicted, wetted, air, interfaces) and scalars (step, density, sigma, ambient):
        (status, cells, faces, lower-theta count)."""
        o = p.geom[0]["world_origin"]
        arrs = [np.ascontiguousarray(a) for a in (p.blocks, p.centers, p.cells, p.faces, p.fields, p.predicted, p.wetted,
                                                    p.air, p.interfaces)]
        blocks, centers, cells, faces, fields, predicted, wetted, air, interfaces = arrs
        view = FrozenView(_p(blocks), len(blocks), (C.c_double * 3)(o["x"], o["y"], o["z"]), p.geom[0]["cell_width"],
                          _p(centers), len(centers), _p(cells), len(cells), _p(faces), len(faces), p.step, p.density,
                          p.sigma, 0.0)
        flds = PressureFields(_p(fields), len(fields), _p(predicted), len(predicted), p.ambient, _p(wetted), len(wetted),
                              _p(air), len(air), _p(interfaces), len(interfaces))
        out_c = np.zeros(len(cells), R.PRESSURE_CELL)
        out_f = np.zeros(len(faces), R.PRESSURE_FACE)
        lower = C.c_uint64()
        s = lib().fl_build_pressure_links(self.handle, C.byref(view), C.byref(flds), _p(out_c), _p(out_f),
                                          C.byref(lower), C.byref(self.last))
        return s, out_c, out_f, lower.value

    def freeze_geometry(self, p):
        """fl_freeze_geometry on a problem with numpy record arrays (geom, blocks, candidates, centers,
        corners), an obstacle tuple (kind, unit normal, coordinate, half width, wall velocity) or None,
        scalars (step, density, sigma, exact_curvature or None, max_cells, max_faces) and optionally a
        sampled solid (solid_corners: SOLID_CORNER records or None; face_normal_velocities: None or one
        per owned slot; split_rule: the wet-piece refusal rule, 1 by default): (status, frozen cells, frozen
        faces, capillary step bound)."""
        o = p.geom[0]["world_origin"]
        blocks, cands, centers, corners = (np.ascontiguousarray(a) for a in (p.blocks, p.candidates, p.centers,
                                                                               p.corners))
        obs = Obstacle()
        if p.obstacle is not None:
            kind, n, c, w, v = p.obstacle[:5]
            t = p.obstacle[5] if len(p.obstacle) > 5 else (0.0,) * 6
            obs = Obstacle(kind, 0, (C.c_double * 3)(*n), c, w, (C.c_double * 3)(*v), (C.c_double * 6)(*t))
        k = p.exact_curvature
        inputs = FreezeInputs(_p(blocks), len(blocks), (C.c_double * 3)(o["x"], o["y"], o["z"]),
                              p.geom[0]["cell_width"], _p(cands), len(cands), _p(centers), len(centers), _p(corners),
                              len(corners), obs, p.step, p.density, p.sigma, 0.0 if k is None else k,
                              0 if k is None else 1, (C.c_uint8 * 7)(), p.max_cells, p.max_faces)
        solids = getattr(p, "solid_corners", None)
        velocities = getattr(p, "face_normal_velocities", None)
        if solids is not None:
            solids = np.ascontiguousarray(solids, dtype=R.SOLID_CORNER)
            inputs.solid_corners, inputs.solid_corner_count = _p(solids), len(solids)
            inputs.has_sampled_solid = int(getattr(p, "split_rule", 1))
            if velocities is not None:
                velocities = np.ascontiguousarray(velocities, dtype=np.float64)
                inputs.face_normal_velocities, inputs.face_velocity_count = _p(velocities), len(velocities)
        cells = np.zeros(len(cands), R.FROZEN_CELL)
        faces = np.zeros(6 * len(cands), R.FROZEN_FACE)
        nc, nf, bound = C.c_uint64(), C.c_uint64(), C.c_double()
        s = lib().fl_freeze_geometry(self.handle, C.byref(inputs), _p(cells), C.byref(nc), _p(faces), C.byref(nf),
                                     C.byref(bound), C.byref(self.last))
        if s != OK:
            return s, None, None, None
        return s, cells[:nc.value].copy(), faces[:nf.value].copy(), bound.value

    def update_particles(self, options, gathered, distances, particles):
        """fl_update_particles on a copy of the particles: (status, particles after the call)."""
        out = np.ascontiguousarray(particles).copy()
        g = np.ascontiguousarray(gathered)
        d = np.ascontiguousarray(distances, dtype=np.float64)
        s = lib().fl_update_particles(self.handle, _p(options), _p(g), _p(d), C.c_size_t(len(out)), _p(out),
                                      C.byref(self.last))
        return s, out

    def p2g(self, geometry, blocks, particles, out=None):
        nf = 192 * len(blocks)
        if out is None:
            out = np.zeros(nf, R.FACE_DEPOSIT)
        s = lib().fl_p2g(self.handle, _p(geometry), _p(blocks), C.c_size_t(len(blocks)), _p(particles),
                         C.c_size_t(len(particles)), _p(out), C.c_size_t(len(out)), C.byref(self.last))
        return s, out

    def p2g_quanta(self, geometry, blocks, particles, mass_q, mom_q, cap, out=None):
        nf = 192 * len(blocks)
        if out is None:
            out = np.zeros(nf, R.FACE_QUANTA)
        s = lib().fl_p2g_quanta(self.handle, _p(geometry), _p(blocks), C.c_size_t(len(blocks)), _p(particles),
                                C.c_size_t(len(particles)), Quantization(mass_q, mom_q, cap), _p(out),
                                C.c_size_t(len(out)), C.byref(self.last))
        return s, out

    def p2g_compatible(self, geometry, blocks, particles, tags, mass_q, mom_q, cap, stencils=False):
        """DepositCompatibleMacQuanta3D: (status, face quanta, stencils [n][3][27] or None)."""
        out = np.zeros(192 * len(blocks), R.FACE_QUANTA)
        st = np.zeros((len(particles), 3, 27), R.COMPATIBLE_SAMPLE) if stencils else None
        s = lib().fl_p2g_compatible(self.handle, _p(geometry), _p(blocks), C.c_size_t(len(blocks)), _p(particles),
                                    C.c_size_t(len(particles)), _p(tags), C.c_size_t(len(tags)),
                                    Quantization(mass_q, mom_q, cap), _p(out), C.c_size_t(len(out)), _p(st),
                                    C.byref(self.last))
        return s, out, st

    def g2p_boundary(self, geometry, blocks, particles, old_time, new_time, old_history, new_history):
        """GatherLiquidBoundaryParticles3D: (status, records, [old-PIC, new-PIC, low-mass fallback counts])."""
        out = np.zeros(len(particles), R.BOUNDARY_GATHER)
        counts = np.zeros(3, np.uint64)
        s = lib().fl_g2p_boundary(self.handle, _p(geometry), _p(blocks), C.c_size_t(len(blocks)), _p(particles),
                                  C.c_size_t(len(particles)), C.c_double(old_time), C.c_double(new_time),
                                  C.byref(old_history), C.byref(new_history), _p(out), _p(counts),
                                  C.byref(self.last))
        return s, out, counts

    def normalize_forces(self, quanta, mass_q, mom_q, min_mass, accel, dt):
        nf = len(quanta)
        m, u, v, p = np.zeros(nf), np.zeros(nf), np.zeros(nf, np.uint8), np.zeros(nf)
        s = lib().fl_normalize_forces(self.handle, _p(quanta), C.c_size_t(nf), Quantization(mass_q, mom_q, 0),
                                      C.c_double(min_mass), _vec3(accel), C.c_double(dt), _p(m), _p(u), _p(v), _p(p),
                                      C.byref(self.last))
        return s, m, u, v, p

    def apply_forces(self, old, accel, dt):
        old = np.ascontiguousarray(old, dtype=np.float64)
        pred = np.zeros(len(old))
        s = lib().fl_apply_forces(self.handle, _p(old), C.c_size_t(len(old)), _vec3(accel), C.c_double(dt), _p(pred),
                                  C.byref(self.last))
        return s, pred

    def g2p(self, geometry, blocks, particles, old, new, out=None):
        if out is None:
            out = np.zeros(len(particles), R.GATHER_RESULT)
        old = np.ascontiguousarray(old, dtype=np.float64)
        new = np.ascontiguousarray(new, dtype=np.float64)
        s = lib().fl_g2p(self.handle, _p(geometry), _p(blocks), C.c_size_t(len(blocks)), _p(particles),
                         C.c_size_t(len(particles)), _p(old), _p(new), C.c_size_t(len(old)), _p(out),
                         C.byref(self.last))
        return s, out

    def project(self, geometry, blocks, dt, cells, faces, options):
        nc, nf = len(cells), len(faces)
        p, u, b, r, d = (np.zeros(nc), np.zeros(nf), np.zeros(nc), np.zeros(nc), np.zeros(nc))
        gauges = np.zeros(nc, np.uint64)
        diag = np.zeros(1, R.PRESSURE_DIAGNOSTICS)
        s = lib().fl_project(self.handle, _p(geometry), _p(blocks), C.c_size_t(len(blocks)), C.c_double(dt),
                             _p(cells), C.c_size_t(nc), _p(faces), C.c_size_t(nf), _p(options), _p(p), _p(u), _p(b),
                             _p(r), _p(d), _p(gauges), _p(diag), C.byref(self.last))
        g = gauges[: int(diag["gauge_count"][0])].copy()
        return s, dict(pressures=p, velocities=u, rhs=b, residuals=r, defects=d, gauges=g, diagnostics=diag[0])

    def project_viscous(self, geometry, blocks, dt, cells, faces, rows, options):
        """fl_project_viscous: (status, dict of pressures, face velocities, cell defects, diagnostics)."""
        nc, nf = len(cells), len(faces)
        rows = np.ascontiguousarray(rows, dtype=R.STRAIN_ROW)
        p, u, d = np.zeros(nc), np.zeros(nf), np.zeros(nc)
        diag = np.zeros(1, R.VISCOUS_DIAGNOSTICS)
        s = lib().fl_project_viscous(self.handle, _p(geometry), _p(blocks), C.c_size_t(len(blocks)), C.c_double(dt),
                                     _p(cells), C.c_size_t(nc), _p(faces), C.c_size_t(nf), _p(rows),
                                     C.c_size_t(len(rows)), _p(options), _p(p), _p(u), _p(d), _p(diag),
                                     C.byref(self.last))
        return s, dict(pressures=p, velocities=u, defects=d, diagnostics=diag[0])

    # ---- device-resident scene ----
    def scene_upload(self, geometry, positions, velocities, slopes, mass):
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        vel = np.ascontiguousarray(velocities, dtype=np.float64)
        c = np.ascontiguousarray(slopes, dtype=np.float64)
        return lib().fl_scene_upload(self.handle, _p(geometry), C.c_size_t(len(pos)), _p(pos), _p(vel), _p(c),
                                     C.c_double(mass))

    def scene_prepare(self, cells, domain_cells, density):
        return lib().fl_scene_prepare(self.handle, _p(cells), C.c_size_t(len(cells)), C.c_int32(domain_cells),
                                      C.c_double(density))

    def assemble_full_cell_strain(self, geometry, blocks, cells):
        """fl_assemble_full_cell_strain: (status, strain rows)."""
        cells = np.ascontiguousarray(cells, dtype=R.VISCOSITY_CELL)
        out = np.zeros(max(6 * len(cells), 1), R.STRAIN_ROW)
        count = C.c_size_t(len(out))
        s = lib().fl_assemble_full_cell_strain(self.handle, _p(geometry), _p(blocks), C.c_size_t(len(blocks)),
                                               _p(cells), C.c_size_t(len(cells)), _p(out), C.byref(count),
                                               C.byref(self.last))
        return s, out[:count.value] if s == OK else None

    def solve_viscous_block(self, rows, inertia, dt, rhs, max_iterations=256, relative_tolerance=1e-10):
        """fl_solve_viscous_block: (status, velocity, iterations, max residual)."""
        rows = np.ascontiguousarray(rows, dtype=R.STRAIN_ROW)
        inertia = np.ascontiguousarray(inertia, dtype=np.float64)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        v = np.zeros(len(inertia))
        it, mr = C.c_uint64(0), C.c_double(0.0)
        s = lib().fl_solve_viscous_block(self.handle, _p(rows), C.c_size_t(len(rows)), _p(inertia),
                                         C.c_size_t(len(inertia)), C.c_double(dt), _p(rhs), C.c_uint64(max_iterations),
                                         C.c_double(relative_tolerance), _p(v), C.byref(it), C.byref(mr),
                                         C.byref(self.last))
        return s, v, it.value, mr.value

    def time_projection_kernels(self, reps=20):
        """fl_time_projection_kernels: [(kernel, level, slots, ms)] for the last projected system."""
        cap = 32
        out = (KernelTime * cap)()
        count = C.c_size_t(0)
        s = lib().fl_time_projection_kernels(self.handle, C.c_int(reps), out, C.c_size_t(cap), C.byref(count))
        return s, [(int(r.kernel), int(r.level), int(r.slots), float(r.ms)) for r in out[:count.value]]

    def step(self, dt, gravity, mass_q, mom_q, cap, options):
        rep = StepReport()
        s = lib().fl_step(self.handle, C.c_double(dt), _vec3(gravity), Quantization(mass_q, mom_q, cap), _p(options),
                          C.byref(rep))
        return s, rep

    def scene_download(self, n_blocks, n_particles, n_cells, what=("blocks", "quanta", "u0", "ustar", "unew", "order",
                                                                    "gather", "pressures")):
        nf = 192 * n_blocks
        out = {}
        out["blocks"] = np.zeros(n_blocks, R.INT3) if "blocks" in what else None
        out["quanta"] = np.zeros(nf, R.FACE_QUANTA) if "quanta" in what else None
        for k in ("u0", "ustar", "unew"):
            out[k] = np.zeros(nf) if k in what else None
        out["order"] = np.zeros(n_particles, R.SORT_ENTRY) if "order" in what else None
        out["gather"] = np.zeros(n_particles, R.GATHER_RESULT) if "gather" in what else None
        out["pressures"] = np.zeros(n_cells) if "pressures" in what else None
        s = lib().fl_scene_download(self.handle, _p(out["blocks"]), C.c_size_t(n_blocks), _p(out["quanta"]),
                                    _p(out["u0"]), _p(out["ustar"]), _p(out["unew"]), _p(out["order"]),
                                    _p(out["gather"]), _p(out["pressures"]))
        return s, out


    def project_device_built(self, geometry, blocks, dt, cells, faces, options):
        """fl_project_device_built: (status, pressures, face velocities, diagnostics)."""
        p, u = np.zeros(len(cells)), np.zeros(len(faces))
        diag = np.zeros(1, R.PRESSURE_DIAGNOSTICS)
        s = lib().fl_project_device_built(self.handle, _p(geometry), _p(blocks), C.c_size_t(len(blocks)),
                                          C.c_double(dt), _p(cells), C.c_size_t(len(cells)), _p(faces),
                                          C.c_size_t(len(faces)), _p(options), _p(p), _p(u), _p(diag))
        return s, p, u, diag[0]

    def sample_particle_field(self, geometry, positions, radius, capacity=1 << 24):
        """fl_sample_particle_field: (status, candidate cells, centre samples, corner samples)."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        nc, ncen, ncor = C.c_uint64(), C.c_uint64(), C.c_uint64()
        s = lib().fl_sample_particle_field(self.handle, _p(geometry), _p(pos), C.c_size_t(len(pos)), C.c_double(radius),
                                           C.c_uint64(capacity), None, C.byref(nc), None, C.c_uint64(0),
                                           C.byref(ncen), None, C.c_uint64(0), C.byref(ncor), C.byref(self.last))
        if s != OK:
            return s, None, None, None
        cells = np.zeros(max(nc.value, 1), R.INT3)
        cen = np.zeros(max(ncen.value, 1), R.CENTER_DISTANCE)
        cor = np.zeros(max(ncor.value, 1), R.CENTER_DISTANCE)
        s = lib().fl_sample_particle_field(self.handle, _p(geometry), _p(pos), C.c_size_t(len(pos)), C.c_double(radius),
                                           C.c_uint64(capacity), _p(cells), C.byref(nc), _p(cen), C.c_uint64(len(cen)),
                                           C.byref(ncen), _p(cor), C.c_uint64(len(cor)), C.byref(ncor),
                                           C.byref(self.last))
        if s != OK:
            return s, None, None, None
        return s, cells[:nc.value], cen[:ncen.value], cor[:ncor.value]

    def set_transfer_precision(self, bits):
        """fl_set_transfer_precision: 64 (bit-exact FP64 deposit) or 32 (FP32 arithmetic, integer quanta)."""
        return lib().fl_set_transfer_precision(self.handle, C.c_int(bits))

    def substep(self, options):
        """fl_substep on the uploaded scene: (status, report record)."""
        rep = np.zeros(1, R.SUBSTEP_REPORT)
        o = np.ascontiguousarray(options, dtype=R.SUBSTEP_OPTIONS)
        s = lib().fl_substep(self.handle, _p(o), _p(rep))
        return s, rep[0]

    def substep_download(self, n):
        """fl_substep_download: (status, positions, velocities, slopes, volume ratios, flags)."""
        x, v, c, j, f = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 9)), np.zeros(n), np.zeros(n, np.uint8)
        s = lib().fl_substep_download(self.handle, _p(x), _p(v), _p(c), _p(j), _p(f), C.c_size_t(n))
        return s, x, v, c, j, f


class Network:
    """The learned preconditioner of a context (fl_learn_*): FP32 parameters, training pairs, training."""

    def __init__(self, ctx, tiles, pairs):
        self.ctx = ctx
        s = lib().fl_learn_setup(ctx.handle, C.c_uint64(tiles), C.c_uint64(pairs))
        if s != OK:
            raise RuntimeError(f"fl_learn_setup failed with status {s}")
        self.count = int(lib().fl_learn_param_count())

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        return lib().fl_learn_set_params(self.ctx.handle, _p(p), C.c_size_t(len(p)))

    def params(self):
        out = np.zeros(self.count, np.float32)
        assert lib().fl_learn_get_params(self.ctx.handle, _p(out), C.c_size_t(self.count)) == OK
        return out

    def forward(self, inputs):
        x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 4, 512)
        out = np.zeros((len(x), 512), np.float32)
        s = lib().fl_learn_forward(self.ctx.handle, _p(x), C.c_size_t(len(x)), _p(out))
        return s, out

    def infer(self, inputs):
        x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 4, 512)
        out = np.zeros((len(x), 512), np.float32)
        s = lib().fl_learn_infer(self.ctx.handle, _p(x), C.c_size_t(len(x)), _p(out))
        return s, out

    def loss_grad(self, inputs, targets):
        x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 4, 512)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 512)
        loss = C.c_float()
        gp = np.zeros(self.count, np.float32)
        gi = np.zeros(x.shape, np.float32)
        s = lib().fl_learn_loss_grad(self.ctx.handle, _p(x), _p(t), C.c_size_t(len(x)), C.byref(loss), _p(gp), _p(gi))
        return s, loss.value, gp, gi

    def add_pairs(self, inputs, targets):
        x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 4, 512)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 512)
        return lib().fl_learn_add_pairs(self.ctx.handle, _p(x), _p(t), C.c_size_t(len(x)))

    def pairs(self):
        n = C.c_size_t()
        assert lib().fl_learn_pairs(self.ctx.handle, C.byref(n)) == OK
        return n.value

    def collect(self, every, clear=False):
        return lib().fl_learn_collect(self.ctx.handle, C.c_int(every), C.c_int(1 if clear else 0))

    def eval(self, batch=256):
        loss = C.c_float()
        s = lib().fl_learn_eval(self.ctx.handle, C.c_size_t(batch), C.byref(loss))
        return s, loss.value

    def train(self, epochs, batch=256, lr=1e-3, seed=0):
        losses = np.zeros(epochs, np.float32)
        s = lib().fl_learn_train(self.ctx.handle, C.c_int(epochs), C.c_size_t(batch), C.c_float(lr), C.c_uint64(seed),
                                 _p(losses))
        return s, losses


def pressure_options(max_iterations=256, tolerance_scale=1.0, relative_target=0.0, levels=4, sweeps=2,
                     coarse_sweeps=16, coarse_scale=0.5, bf16_levels=False, learned=False, smoother=0):
    """smoother: 0 level-wide red-black Gauss-Seidel, 1 block smoother on every level, 2 block smoother below
    the finest level."""
    o = np.zeros(1, R.PRESSURE_OPTIONS)
    o[0] = (max_iterations, tolerance_scale, relative_target, levels, sweeps, coarse_sweeps, 2 if bf16_levels else 1,
            coarse_scale, 1 if learned else 0, int(smoother))
    return o


def substep_options(dt, gravity=(0.0, -9.81, 0.0), mass_q=2.0 ** -40, mom_q=2.0 ** -40, capacity=1 << 20,
                    minimum_face_mass=0.0, field_radius=1.0, density=1000.0, surface_tension=0.0, ambient=0.0,
                    blend=1.0, friction=0.0, penetration=0.0, obstacle=None, pressure=None, solid=None,
                    dynamic_viscosity=0.0):
    """fl_substep_options record (obstacle: an OBSTACLE record; pressure: a pressure_options record;
    solid: a SCENE_SOLID record, see scene_solid)."""
    o = np.zeros(1, R.SUBSTEP_OPTIONS)
    o["step_duration"] = dt
    o["gravity"] = tuple(gravity)
    o["quantization"] = (mass_q, mom_q, capacity)
    o["minimum_face_mass"] = minimum_face_mass
    o["field_radius"] = field_radius
    o["liquid_density"] = density
    o["surface_tension"] = surface_tension
    o["ambient_pressure"] = ambient
    o["particle_energy_blend"] = blend
    o["friction_coefficient"] = friction
    o["penetration_tolerance"] = penetration
    if obstacle is not None:
        o["obstacle"] = obstacle
    o["pressure"] = pressure_options() if pressure is None else pressure
    if solid is not None:
        o["solid"] = solid
    o["dynamic_viscosity"] = dynamic_viscosity
    return o


def scene_solid(sphere=None, paddle=None, walls=None, split_rule=1):
    """fl_scene_solid record: sphere = (centre, radius); paddle = (pivot, half extent, angle, angular
    velocity about +y); walls = (low, high) bounds of the container in x and z; split_rule: the freeze's
    wet-piece refusal (1 the reference's, 2 only in cells a solid reaches, 3 none)."""
    r = np.zeros(1, R.SCENE_SOLID)
    r["split_rule"] = split_rule
    parts = 0
    if sphere is not None:
        parts |= 1
        r["sphere_centre"], r["sphere_radius"] = tuple(sphere[0]), sphere[1]
    if paddle is not None:
        parts |= 2
        r["paddle_pivot"], r["paddle_half_extent"] = tuple(paddle[0]), tuple(paddle[1])
        r["paddle_angle"], r["paddle_angular_velocity"] = paddle[2], paddle[3]
    if walls is not None:
        parts |= 4
        r["walls_low"], r["walls_high"] = tuple(walls[0]), tuple(walls[1])
    r["parts"] = parts
    return r


STATIC_MAX_FACES = 196608


def normalize_forces_static(device, quanta, mass_q, mom_q, min_mass, accel, dt):
    """fl_normalize_forces_static: the context-free, allocation-free normalize + forces on `device`."""
    nf = len(quanta)
    m, u, v, p = np.zeros(nf), np.zeros(nf), np.zeros(nf, np.uint8), np.zeros(nf)
    s = lib().fl_normalize_forces_static(C.c_int(device), _p(quanta), C.c_size_t(nf), Quantization(mass_q, mom_q, 0),
                                         C.c_double(min_mass), _vec3(accel), C.c_double(dt), _p(m), _p(u), _p(v),
                                         _p(p))
    return s, m, u, v, p
