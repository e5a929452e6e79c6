#!/usr/bin/env python3
"""Benchmark of the liquid-step hot path on one MI355X.

One step = one pass of the hot path over the scene, device-resident: bin + stable counting sort,
quantized P2G (FP64 contributions, int64 quanta), normalize + gravity, MAC pressure projection by
mixed-precision MGPCG to a relative divergence of 1e-6 (plus the reference residual gates), and
G2P (FP64). The default workload is two colliding spheres (r = 0.2, +-2 m/s) over a 10 % pool on
a 256^3 sparse grid, 8 particles per liquid cell (22,627,520 particles, 2,828,440 pressure cells).

Prints one JSON line: value = particle-steps/s of the whole step (particles x steps / wall time),
with projection ms, per-stage times, the P2G kernel roofline and a CPU baseline (the oracle,
one thread, on a bounded sample of the same scene) alongside.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from fluidity_amd import hip as H  # noqa: E402
from fluidity_amd import records as R  # noqa: E402
from fluidity_amd import scenes  # noqa: E402

METRIC = "projection ms to 1e-6 rel divergence (learned vs MGPCG); P2G/G2P particle-steps/s"
QUANTUM = 2.0 ** -40
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def p2g_algorithmic_bytes(n_particles, n_faces):
    # Per particle: position (3), mass (1), velocity (3), affine rows (9) as FP64 = 128 B read once;
    # per owned face slot: mass and momentum quanta (2 x int64) written once.
    return 128 * n_particles + 16 * n_faces


def g2p_algorithmic_bytes(n_particles, n_faces):
    # Per particle: position (3 x FP64) read, 24 FP64 results written = 216 B;
    # per owned face slot: old and new velocity (2 x FP64) read once.
    return 216 * n_particles + 16 * n_faces


# Algorithmic bytes per cell slot (64 per 4^3 block) of the projection kernels, each array counted once
# per launch; cb is the byte width of a V-cycle coefficient (2 for BF16 storage, 4 for FP32). None: the
# single-workgroup tail cycle over the smallest levels is latency-bound and not priced.
def projection_bytes_per_slot(name, cb):
    return {
        # other colour's x, own b and diagonal, the three coefficient arrays, own x written
        "k_smooth": 8 + 3 * cb,
        # every x, three coefficient arrays, red anchor, b and diagonal, red residual written
        "k_residual_red": 12 + 3 * cb,
        # red residuals read, coarse right-hand side (1/8 of the slots) written
        "k_restrict_red": 2.5,
        # diagonal and x read, coarse x (1/8) read, x written
        "k_prolong": 12.5,
        # b, diagonal and coefficients read, x written (level 1)
        "k_block_smooth_pre": 12 + 3 * cb,
        # x, b, diagonal, coefficients and coarse x (1/8) read, second buffer written (level 1)
        "k_block_smooth_post": 16.5 + 3 * cb,
        # x, coefficients, anchor, b and diagonal read, residual written (level 1)
        "k_residual_all": 20 + 3 * cb,
        # residuals read, coarse right-hand side written (level 1 to 2)
        "k_restrict_all": 4.5,
        "k_tail_vcycle": None,
        # z, d, three FP32 coefficients, anchor and diagonal read, next d and q written
        "k_direction_apply": 36,
        # d, q, diagonal, volume, x and r read, x and r written
        "k_update_xr": 32,
        # r and z read
        "k_dot": 8,
        # per cell: p and seven FP64 cell arrays, six face ordinals, three 64-byte face records (each
        # face is shared by two cells), residual and defect written
        "k_evaluate": 8 + 56 + 24 + 192 + 16,
    }[name]


def projection_roofline(ctx, level_storage, reps=20):
    s, rows = ctx.time_projection_kernels(reps)
    assert s == H.OK, s
    cb = 2 if level_storage == "bf16" else 4
    out = []
    for kernel, level, slots, ms in rows:
        name = H.KERNEL_NAMES[kernel]
        per = projection_bytes_per_slot(name, cb)
        row = {"kernel": name, "level": level, "slots": slots, "ms": ms}
        if per is not None:
            b = per * slots
            row.update(algorithmic_bytes=b, achieved_gbs=b / (ms * 1e-3) / 1e9,
                       frac=b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)
        out.append(row)
    return out


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE passes scaled by a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"traffic_bytes_per_launch {kernel} {workload}:"):
            return int(line.split(":")[1])
    return None


def copy_ceiling(nbytes=1 << 30, reps=5):
    """Plain float4 device-to-device copy of `nbytes` (fl_probe_copy): bandwidth in GB/s.

    The same copy, of known size, calibrates the FETCH_SIZE / WRITE_SIZE counters.
    """
    import ctypes as C

    gbs = C.c_double()
    assert H.lib().fl_probe_copy(0, C.c_size_t(nbytes), reps, C.byref(gbs)) == H.OK
    return gbs.value


def runtime_copy_ceiling(nbytes=1 << 30, reps=5):
    """The HIP runtime's own device-to-device copy (hipMemcpyAsync, DeviceToDevice) of `nbytes`, timed
    with HIP events: a cross-check of the float4 copy kernel's ceiling. GB/s counts read + write."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    e0, e1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    hip.hipMemsetAsync(a, 0, C.c_size_t(nbytes), None)
    hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None)
    hip.hipEventRecord(e0, None)
    for _ in range(reps):
        hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None)
    hip.hipEventRecord(e1, None)
    hip.hipEventSynchronize(e1)
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), e0, e1)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def projection_inputs(scene, ctx, report):
    """The projection's inputs exactly as the device step froze them: the step's block list and u* at the
    liquid cells' incident faces (owned-slot order), plus the GPU's accepted pressures and face velocities."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    nb, nc = int(report.blocks), int(report.pressure_cells)
    s, dl = ctx.scene_download(nb, 0, nc, what=("blocks", "u0", "ustar", "unew", "pressures"))
    assert s == H.OK
    slots = scene.pressure_face_slots(dl["blocks"])[0]
    return dict(geom=scene.geometry_record().view(O.GEOMETRY), blocks=dl["blocks"].view(O.INT3),
                cells=scene.pressure_cells().view(O.PRESSURE_CELL),
                faces=scene.pressure_faces(dl["blocks"], dl["ustar"]).view(O.PRESSURE_FACE), slots=slots, dl=dl)


def cpu_baseline(scene, pin, sample, threads=1):
    """The oracle (C++ restatement of the reference operators) timed on this host.

    Transfers are timed on `sample` particles of the scene and scaled to the full count, on one thread
    (the reference's semantics) and again on `threads` threads (per-thread integer accumulators, the
    same bits). The projection of every pressure cell is timed twice on one thread (the oracle's solvers
    are serial): with the reference's own preconditioner (aggregation multigrid, weighted Jacobi, FP64
    Cholesky blocks) and with the geometric V-cycle the GPU runs; `value` uses the reference's.
    """
    import pyoracle as O

    sel = np.arange(min(sample, scene.count))
    parts = scene.particles_record(sel).view(O.PARTICLE)
    geom, blocks, dl = pin["geom"], pin["blocks"], pin["dl"]

    def transfers(nthreads):
        t0 = time.perf_counter()
        s, _ = O.deposit_quanta(geom, blocks, parts, QUANTUM, QUANTUM, len(parts), max_capacity=1 << 40,
                                threads=nthreads)
        t_p2g = time.perf_counter() - t0
        assert s == O.OK, s
        t0 = time.perf_counter()
        s, _ = O.gather(geom, blocks, parts, dl["u0"], dl["ustar"], threads=nthreads)
        t_g2p = time.perf_counter() - t0
        assert s == O.OK, s
        return t_p2g, t_g2p

    t_p2g, t_g2p = transfers(1)
    args = (geom, blocks, scene.dt, pin["cells"], pin["faces"])
    legs = {}
    for name, opts in (("reference", O.pressure_options(relative_target=1e-6, reference=True)),
                       ("geometric", O.pressure_options(relative_target=1e-6, levels=6, coarse_sweeps=2))):
        t0 = time.perf_counter()
        s, out = O.project(*args, opts)
        t = time.perf_counter() - t0
        assert s == O.OK, (name, s)
        d = out["diagnostics"]
        legs[name] = {"ms": t * 1e3, "iterations": int(d["iteration_count"]), "levels": int(d["multigrid_level_count"]),
                      "relative_divergence": float(d["defect_l2"] / d["initial_defect_l2"])}
        log(f"cpu projection ({name} preconditioner): {t:.1f}s, {legs[name]['iterations']} iterations")
    t_proj = legs["reference"]["ms"] * 1e-3
    step_s = (t_p2g + t_g2p) / len(parts) * scene.count + t_proj
    nc = len(pin["cells"])
    result = {
        "value": scene.count / step_s,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{scene.name}: P2G+G2P on {len(parts):,} of {scene.count:,} particles scaled to all, plus the "
                   f"full projection of {nc:,} cells with the reference's preconditioner "
                   f"({legs['reference']['iterations']} PCG iterations); oracle C++ FP64, 1 thread"),
        "p2g_s_per_particle": t_p2g / len(parts),
        "g2p_s_per_particle": t_g2p / len(parts),
        "projection_ms": legs["reference"]["ms"],
        "projection": legs,
        "cpu": cpu_model(),
        "compiler": O.compiler(),
        "nproc": os.cpu_count(),
    }
    if threads > 1:
        m_p2g, m_g2p = transfers(threads)
        m_step = (m_p2g + m_g2p) / len(parts) * scene.count + t_proj
        result["all_cores"] = {"value": scene.count / m_step, "cores": threads,
                               "p2g_s_per_particle": m_p2g / len(parts), "g2p_s_per_particle": m_g2p / len(parts),
                               "note": "transfers threaded; projection serial (same time as above)"}
    return result


def fp64_oracle_error(scene, pin):
    """The GPU's accepted projection against a tight FP64 oracle solve of the same frozen system
    (MGPCG in FP64 to 1e-12 relative divergence), and the GPU pressures re-evaluated by the oracle's FP64
    Evaluate: relative pressure and velocity errors, divergence L2 / Linf."""
    import pyoracle as O

    args = (pin["geom"], pin["blocks"], scene.dt, pin["cells"], pin["faces"])
    t0 = time.perf_counter()
    s, ref = O.project(*args, O.pressure_options(relative_target=1e-12, fp32=False, levels=6, coarse_sweeps=2,
                                                 tolerance_scale=1e-6))
    assert s == O.OK, s
    t_ref = time.perf_counter() - t0
    p_gpu, p_ref = pin["dl"]["pressures"], ref["pressures"]
    u_gpu, u_ref = pin["dl"]["unew"][pin["slots"]], ref["velocities"]
    s, ev = O.evaluate(*args, p_gpu)
    assert s == O.OK, s
    d, rd = ev["diagnostics"], ref["diagnostics"]
    out = {
        "reference_solve": "oracle MGPCG, FP64 vectors and V-cycle, 1e-12 relative divergence",
        "reference_relative_divergence": float(rd["defect_l2"] / rd["initial_defect_l2"]),
        "pressure_rel_linf": float(np.abs(p_gpu - p_ref).max() / np.abs(p_ref).max()),
        "pressure_rel_l2": float(np.linalg.norm(p_gpu - p_ref) / np.linalg.norm(p_ref)),
        "velocity_rel_linf": float(np.abs(u_gpu - u_ref).max() / np.abs(u_ref).max()),
        "velocity_rel_l2": float(np.linalg.norm(u_gpu - u_ref) / np.linalg.norm(u_ref)),
        "divergence_l2": float(d["defect_l2"]),
        "divergence_linf": float(d["defect_linf"]),
        "relative_divergence": float(d["defect_l2"] / d["initial_defect_l2"]),
        "oracle_solve_s": t_ref,
    }
    log(f"fp64 oracle: pressure error {out['pressure_rel_linf']:.2e} (Linf, relative), velocity "
        f"{out['velocity_rel_l2']:.2e} (L2, relative), divergence {out['relative_divergence']:.2e} relative")
    return out


def learned_comparison(ctx, scene, dt, opts, args):
    """Projection with the learned preconditioner (flexible CG, V-cycle fallback) against plain MGPCG on
    the same step: training pairs come from the MGPCG solve of one step, the network is trained on the
    device, then one step is timed with each preconditioner."""
    # 8^3 tiles of the level-0 pressure blocks (blocks holding a pressure cell).
    cells = scene.pressure_cells()
    c = np.stack([cells["cell_index"]["x"], cells["cell_index"]["y"], cells["cell_index"]["z"]], 1)
    tiles = len(np.unique(np.unique(np.floor_divide(c, 4), axis=0) // 2, axis=0))
    net = H.Network(ctx, tiles=max(tiles, 1024), pairs=args.learn_pairs)
    rng = np.random.default_rng(0)
    init = []
    for cin, cout in [(4, 8), (8, 8), (8, 16), (16, 16), (16, 32), (32, 32), (48, 16), (16, 16), (24, 8), (8, 1)]:
        init += [rng.normal(0.0, np.sqrt(2.0 / (27 * cin)), cout * cin * 27), np.zeros(cout)]
    assert net.set_params(np.concatenate(init).astype(np.float32)) == H.OK
    assert net.collect(1, clear=True) == H.OK
    s, rep = ctx.step(dt, scene.gravity, QUANTUM, QUANTUM, scene.count, opts)
    assert s == H.OK, s
    assert net.collect(0) == H.OK
    pairs = net.pairs()
    t0 = time.perf_counter()
    s, losses = net.train(epochs=args.learn_epochs, batch=256, lr=2e-3, seed=1)
    train_s = time.perf_counter() - t0
    assert s == H.OK, s
    learned_opts = opts.copy()
    learned_opts["preconditioner"] = 1
    s, lrep = ctx.step(dt, scene.gravity, QUANTUM, QUANTUM, scene.count, learned_opts)
    assert s == H.OK, s
    s, mrep = ctx.step(dt, scene.gravity, QUANTUM, QUANTUM, scene.count, opts)
    assert s == H.OK, s
    ld, md = lrep.diagnostics(), mrep.diagnostics()
    log(f"learned: {pairs} pairs, {args.learn_epochs} epochs in {train_s:.1f}s, loss {losses[0]:.3g} -> {losses[-1]:.3g}; "
        f"projection {lrep.projection_ms:.2f} ms vs MGPCG {mrep.projection_ms:.2f} ms")
    return {
        "projection_ms": lrep.projection_ms,
        "mgpcg_projection_ms": mrep.projection_ms,
        "iterations": int(ld["iteration_count"]),
        "learned_iterations": int(ld["learned_iterations"]),
        "mgpcg_iterations": int(md["iteration_count"]),
        "relative_divergence": float(ld["defect_l2"] / ld["initial_defect_l2"]),
        "training_pairs": int(pairs),
        "training_epochs": args.learn_epochs,
        "training_s": train_s,
        "loss_first_last": [float(losses[0]), float(losses[-1])],
        "network": "3-level 3D U-Net 8/16/32 on 8^3 tiles: BF16 MFMA inference with FP32 accumulation, FP32 training",
    }


def topology_measurement(ctx, scene, cells, dt):
    """fl_prepare_topology at the workload's size: particles advected by one step, with the topology of
    the unmoved particles as the previous state (blocks, cells and faces remapped by identity)."""
    geom = scene.geometry_record()
    cl = np.ascontiguousarray(cells["cell_index"])
    caps = H.topology_capacities(particles=scene.count, blocks=1 << 17, cells=len(cells) + 64, faces=192 << 17)
    s, _, b0, c0, f0 = ctx.prepare_topology(geom, scene.positions, cl, None, caps)
    assert s == H.OK, s
    c0["previous_pressure"] = 1.0
    moved = scene.positions + dt * scene.velocities
    s, _, b1, c1, f1 = ctx.prepare_topology(geom, moved, cl, H.topology_history(geom, 1, b0, c0, f0), caps)
    assert s == H.OK, s
    m = ctx.last
    log(f"topology: {len(b1):,} blocks, {len(c1):,} cells, {len(f1):,} faces in {m.kernel_ms:.2f} ms on the device "
        f"(+{m.copy_ms:.1f} ms of host copies)")
    return {"kernel_ms": m.kernel_ms, "copy_ms": m.copy_ms, "particles": scene.count, "blocks": len(b1),
            "cells": len(c1), "faces": len(f1), "previous_faces": len(f0)}


def update_measurement(ctx, n=1 << 21, h=1.0 / 256, dt=2e-4):
    """fl_update_particles on n synthetic particles (gathered velocities N(0,1), gradients N(0,3),
    PIC-initialised, liquid distances in [-2h, h]); the device interval of the second of two calls."""
    rng = np.random.default_rng(5)
    g = np.zeros(n, R.BOUNDARY_GATHER)
    for k in "xyz":
        g["new_pic_velocity"][k] = rng.normal(0, 1.0, n)
        g["particle_minus_old_pic_velocity"][k] = rng.normal(0, 0.2, n)
        for r in ("x_row", "y_row", "z_row"):
            g["velocity_gradient"][r][k] = rng.normal(0, 3.0, n)
            g["fitted_velocity_slopes"][r][k] = rng.normal(0, 1.0, n)
    g["affine_enabled"] = 1
    p = np.zeros(n, R.LIQUID_PARTICLE)
    for k in "xyz":
        p["world_position"][k] = rng.uniform(0.0, 1.0, n)
    p["volume_ratio"] = 1.0
    p["pic_initialized"] = 1
    d = rng.uniform(-2 * h, h, n)
    opts = np.zeros(1, R.UPDATE_OPTIONS)  # no obstacle
    opts["cell_width"], opts["step_duration"], opts["particle_energy_blend"] = h, dt, 0.95
    for _ in range(2):  # the first call also pays the first touch of its device buffers
        s, _ = ctx.update_particles(opts, g, d, p)
        assert s == H.OK, s
    m = ctx.last
    log(f"particle update: {n:,} particles in {m.kernel_ms:.3f} ms on the device (+{m.copy_ms:.1f} ms of host copies)")
    return {"particles": n, "kernel_ms": m.kernel_ms, "copy_ms": m.copy_ms,
            "particles_per_s": n / (m.kernel_ms * 1e-3)}


class _Problem:
    """Attribute bag for the freeze and link calls of the boundary measurement."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def boundary_measurement(ctx, scene, cells, dt, cpu_candidates=1 << 17):
    """fl_freeze_geometry then fl_build_pressure_links on the workload's liquid: centre and corner samples of
    its analytic signed distance (two spheres of radius 0.2 over the pool y < 0.1) on the whole grid, the
    centres outside the domain taken as air, every cell with a negative centre distance a candidate (the
    workload's liquid cells up to rounding at the surface), no obstacle and a water surface tension; the
    oracle freeze of the first cpu_candidates candidates, on one host thread, is its CPU rate."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    n, h = scene.n, scene.h
    c = np.arange(-1, n + 1)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")

    def phi(px, py, pz):
        d1 = np.sqrt((px - 0.3) ** 2 + (py - 0.5) ** 2 + (pz - 0.5) ** 2) - 0.2
        d2 = np.sqrt((px - 0.7) ** 2 + (py - 0.5) ** 2 + (pz - 0.5) ** 2) - 0.2
        return np.minimum(np.minimum(d1, d2), py - 0.1)

    samples = []
    for offset in (0.5, 0.0):
        a = np.zeros(x.size, R.CENTER_DISTANCE)
        for k, v in zip("xyz", (x, y, z)):
            a["cell_index"][k] = v.ravel()
        a["signed_distance"] = phi((x + offset) * h, (y + offset) * h, (z + offset) * h).ravel()
        samples.append(a)
    centers, corners = samples
    outside = ((x < 0) | (x >= n) | (y < 0) | (y >= n) | (z < 0) | (z >= n)).ravel()
    centers["signed_distance"][outside] = h
    cand = np.ascontiguousarray(centers["cell_index"][centers["signed_distance"] < 0])
    del x, y, z, outside
    nb = n // 4
    bz, by, bx = np.meshgrid(np.arange(-1, nb + 1), np.arange(-1, nb + 1), np.arange(-1, nb + 1), indexing="ij")
    blocks = np.zeros(bx.size, R.INT3)
    for k, v in zip("xyz", (bx, by, bz)):
        blocks[k] = v.ravel()
    p = _Problem(geom=scene.geometry_record(), blocks=blocks, candidates=cand, centers=centers, corners=corners,
                 obstacle=None, step=dt, density=scene.rho, sigma=0.072, exact_curvature=None, max_cells=0,
                 max_faces=0)
    for _ in range(2):  # the first call also pays the first touch of its device buffers
        s, fc, ff, bound = ctx.freeze_geometry(p)
        assert s == H.OK, s
    fm = (ctx.last.kernel_ms, ctx.last.copy_ms)
    rng = np.random.default_rng(7)
    fields = np.zeros(len(fc), R.CELL_FIELDS)
    fields["bulk_modulus"] = np.inf
    lp = _Problem(geom=p.geom, blocks=blocks, centers=centers, cells=fc, faces=ff, step=dt, density=scene.rho,
                  sigma=0.072, fields=fields, predicted=rng.normal(0.0, 1.0, len(ff)), ambient=0.0,
                  wetted=np.zeros(0, R.OPTIONAL_FLUX), air=np.zeros(0, R.AIR_PRESSURE),
                  interfaces=np.zeros(0, R.INTERNAL_INTERFACE))
    s, _, _, lower = ctx.build_pressure_links(lp)
    assert s == H.OK, s
    lm = (ctx.last.kernel_ms, ctx.last.copy_ms)
    log(f"boundary: {len(cand):,} candidates -> {len(fc):,} cells, {len(ff):,} faces; freeze {fm[0]:.2f} ms, links "
        f"{lm[0]:.2f} ms on the device (+{fm[1] + lm[1]:.1f} ms of host copies)")
    k = min(cpu_candidates, len(cand))
    sample = O.FreezeProblem(p.geom.view(O.GEOMETRY), blocks.view(O.INT3), cand[:k].view(O.INT3),
                             centers.view(O.CENTER_DISTANCE), corners.view(O.CENTER_DISTANCE), step=dt,
                             density=scene.rho, sigma=0.072)
    t0 = time.perf_counter()
    s, _, _, _ = O.freeze_geometry(sample)
    cpu_s = time.perf_counter() - t0
    assert s == O.OK, s
    return {"candidates": len(cand), "cells": len(fc), "faces": len(ff), "capillary_step_bound": bound,
            "lower_theta_count": lower, "freeze_kernel_ms": fm[0], "freeze_copy_ms": fm[1], "links_kernel_ms": lm[0],
            "links_copy_ms": lm[1], "candidates_per_s": len(cand) / (fm[0] * 1e-3),
            "cpu_freeze_candidates_per_s": k / cpu_s, "cpu_sample": f"first {k} candidates, 1 thread"}


def viscosity_measurement(ctx, scene, dt, report, nu=1e-3, cpu_cells=1 << 16):
    """fl_assemble_full_cell_strain then fl_solve_viscous_block on the workload's liquid cells with the
    kinematic viscosity nu (dynamic viscosity rho * nu per cell): the backward-Euler viscous velocity block
    with the step's face masses as inertia (massless faces eliminated) and mass * u* as right-hand side,
    tolerance 1e-10 of the largest |rhs|. The oracle's assembly and solve of the first cpu_cells cells,
    on one host thread, is its CPU rate."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    nb = int(report.blocks)
    s, dl = ctx.scene_download(nb, 0, 0, what=("blocks", "quanta", "ustar"))
    assert s == H.OK, s
    pc = scene.pressure_cells()
    ci = pc["cell_index"]
    order = np.lexsort((ci["x"], ci["y"], ci["z"]))
    cells = np.zeros(len(pc), R.VISCOSITY_CELL)
    cells["cell_index"] = ci[order]
    cells["dynamic_viscosity"] = scene.rho * nu
    geom = scene.geometry_record()
    for _ in range(2):  # the first call also pays the first touch of its device buffers
        s, rows = ctx.assemble_full_cell_strain(geom, dl["blocks"], cells)
        if s != H.OK:
            log(f"viscosity: assembly status {s}")
            return {"status": int(s)}
    am = (ctx.last.kernel_ms, ctx.last.copy_ms)
    mass = dl["quanta"]["mass_quanta_sum"].astype(np.float64) * QUANTUM
    rhs = mass * dl["ustar"]
    blocks = dl["blocks"]
    del dl
    for _ in range(2):
        s, v, it, mr = ctx.solve_viscous_block(rows, mass, dt, rhs)
        assert s == H.OK, s
    sm = (ctx.last.kernel_ms, ctx.last.copy_ms)
    log(f"viscosity: {len(cells):,} cells -> {len(rows):,} strain rows in {am[0]:.2f} ms; velocity block "
        f"{it} iterations in {sm[0]:.2f} ms on the device (+{am[1] + sm[1]:.1f} ms of host copies)")
    k = min(cpu_cells, len(cells))
    t0 = time.perf_counter()
    so, orows = O.assemble_full_cell_strain(geom.view(O.GEOMETRY), blocks.view(O.INT3),
                                            cells[:k].view(O.VISCOSITY_CELL))
    assert so == O.OK, so
    so, _, oit, _ = O.solve_viscous_block(orows, mass, dt, rhs)
    cpu_s = time.perf_counter() - t0
    assert so == O.OK, so
    return {"cells": len(cells), "strain_rows": len(rows), "faces": len(mass), "kinematic_viscosity": nu,
            "step_duration": dt, "assemble_kernel_ms": am[0], "solve_kernel_ms": sm[0], "iterations": int(it),
            "max_residual_over_max_rhs": mr / max(float(np.abs(rhs).max()), 1e-300),
            "copy_ms": am[1] + sm[1], "cells_per_s": len(cells) / ((am[0] + sm[0]) * 1e-3),
            "cpu_cells_per_s": k / cpu_s, "cpu_iterations": int(oit),
            "cpu_sample": f"first {k} cells (z,y,x order), assembly + solve, 1 thread"}


def fp32_transfer_measurement(ctx, scene, dt, opts, args, n, nf):
    """The same steps with the FP32-arithmetic deposit (fl_set_transfer_precision 32): FP32 stencils and
    contributions into the same integer quanta (deterministic; within 1e-5 of the FP64 quanta)."""
    assert ctx.set_transfer_precision(32) == H.OK
    reps = []
    for k in range(args.warmup + args.steps):
        s, rep = ctx.step(dt, scene.gravity, QUANTUM, QUANTUM, scene.count, opts)
        assert s == H.OK, s
        if k >= args.warmup:
            reps.append(rep)
    assert ctx.set_transfer_precision(64) == H.OK
    p2g_ms = float(np.mean([r.p2g_kernel_ms for r in reps]))
    achieved = p2g_algorithmic_bytes(n, nf) / (p2g_ms * 1e-3) / 1e9
    return {"kernel": "k_p2g_f32", "p2g_kernel_ms": p2g_ms, "achieved_gbs": achieved,
            "frac": achieved / HBM_PEAK_GBS, "p2g_particle_steps_per_s": n / (p2g_ms * 1e-3),
            "step_total_ms": float(np.mean([r.total_ms for r in reps]))}


def substep_measurement(scene, cells, args):
    """Consecutive device substeps over moving particles (fl_substep) of the workload in the unit-cube
    open tank: the particle-union field (r = 0.9 h), freeze, typed links, MGPCG on a device-built
    hierarchy, gather and particle update every step; dt = 0.5 h / max(1, |u|max), halved and retried
    when the update refuses a move beyond h / 2."""
    ctx = H.Context(particles=scene.count + 1024, blocks=1 << 17, lookup_blocks=1 << 22,
                    pressure_cells=2 * len(cells) + 4096, pressure_faces=8 * len(cells) + 4096)
    assert ctx.scene_upload(scene.geometry_record(), scene.positions, scene.velocities, scene.slopes,
                            scene.mass) == H.OK
    umax = float(np.sqrt((scene.velocities ** 2).sum(1)).max())
    dt = 0.5 * scene.h / max(1.0, umax)
    tank = np.zeros(1, R.OBSTACLE)
    tank["kind"] = 3
    tank["tank"] = (0.0, 1.0, 0.0, 0.0, 1.0, 0.0)
    opts = H.pressure_options(relative_target=1e-6, levels=args.levels, sweeps=args.sweeps,
                              coarse_sweeps=args.coarse_sweeps, bf16_levels=args.level_storage == "bf16",
                              smoother=["rbgs", "block", "hybrid"].index(args.smoother))
    o = H.substep_options(dt, field_radius=0.9 * scene.h, capacity=scene.count + 1024, obstacle=tank, pressure=opts)
    reports, status, failed_stage, retries, durations = [], H.OK, 0, 0, []
    t0 = time.perf_counter()
    for k in range(args.substeps):
        # The scheduler's rule: a move beyond half a cell aborts the attempt, which is retried with half
        # the duration (the scene is left as it was); the next duration follows the CFL bound
        # 0.5 h / max(1, |u|max) of the updated particles.
        for attempt in range(8):
            o["step_duration"] = dt
            s, rep = ctx.substep(o)
            if not (s == H.ABORTED and int(rep["stage"]) == 8):
                break
            retries += 1
            dt *= 0.5
        if s != H.OK:
            status, failed_stage = s, int(rep["stage"])
            break
        reports.append(rep)
        durations.append(dt)
        dt = 0.5 * scene.h / max(1.0, float(rep["max_speed"]))
        if k == 0:
            log(f"substep 0: {int(rep['candidates']):,} candidates, {int(rep['frozen_cells']):,} cells, "
                f"{int(rep['frozen_faces']):,} faces, {rep['total_ms']:.2f} ms", )
    wall = time.perf_counter() - t0
    ctx.close()
    out = {"steps": len(reports), "requested_steps": args.substeps, "status": int(status), "retries": retries,
           "durations": [float(d) for d in durations], "obstacle": "unit-cube open tank",
           "field": "particle sphere union, r = 0.9 h", "simulated_s": float(sum(durations))}
    if failed_stage:
        out["failed_stage"] = failed_stage
    if not reports:
        return out
    timed = reports[1:] if len(reports) > 1 else reports  # the first substep also pays first-touch growth
    stages = ("sort_ms", "field_ms", "deposit_ms", "freeze_ms", "links_ms", "projection_ms", "gather_ms", "update_ms",
              "total_ms")
    out["stage_ms"] = {k: float(np.mean([r[k] for r in timed])) for k in stages}
    out["wall_ms_per_step"] = wall / len(reports) * 1e3
    out["particle_steps_per_s"] = scene.count / (out["stage_ms"]["total_ms"] * 1e-3)
    last = reports[-1]
    out["last"] = {k: int(last[k]) for k in ("candidates", "centre_samples", "corner_samples", "frozen_cells",
                                            "frozen_faces", "blocks", "lower_theta_count")}
    out["projection_iterations"] = [int(r["pressure"]["iteration_count"]) for r in reports]
    out["relative_divergence_max"] = float(max(r["pressure"]["defect_l2"] / r["pressure"]["initial_defect_l2"]
                                               for r in reports))
    log(f"substeps: {len(reports)} in {wall:.2f}s, {out['stage_ms']['total_ms']:.2f} ms per substep on the device")
    return out


def stirred_measurement(args):
    """The full step with implicit viscosity and level-set solids at 256^3: a dam break (x < 0.4, y < 0.3,
    8 particles per cell) in the unit box with a static sphere r = 0.15 at the centre, a box paddle turning
    at 2 rad/s and the side walls sampled as a solid at every substep's corner keys (split rule 3: no
    wet-piece refusal), the floor as a half-space at y = -h/4 (off the lattice planes, whose zero corner
    distances would leave sliver cells unanchored), FP32 deposit arithmetic, and the coupled viscous
    projection (dynamic viscosity 1 Pa s, kinematic 1e-3 m^2/s at 1000 kg/m^3) on the full-cell strain rows
    of the pressure cells, V-cycle preconditioned GCR; consecutive substeps with the scheduler's retry."""
    scene, solid = scenes.stirred_dam_break(256, seed=0)
    log(f"stirred scene: {scene.count:,} particles")
    ctx = H.Context(particles=scene.count + 1024, blocks=1 << 17, lookup_blocks=1 << 22,
                    pressure_cells=3 * 1024 * 1024, pressure_faces=12 * 1024 * 1024)
    assert ctx.scene_upload(scene.geometry_record(), scene.positions, scene.velocities, scene.slopes,
                            scene.mass) == H.OK
    assert ctx.set_transfer_precision(32) == H.OK
    floor = np.zeros(1, R.OBSTACLE)
    floor["kind"] = 1
    floor["unit_normal"] = (0.0, 1.0, 0.0)
    floor["coordinate"] = -0.25 * scene.h
    opts = H.pressure_options(levels=args.levels, sweeps=3, coarse_sweeps=args.coarse_sweeps, smoother=2)
    (c, r), (pv, e, angle, w), (lo, hi) = solid["sphere"], solid["paddle"], solid["walls"]
    dt = 0.5 * scene.h
    o = H.substep_options(dt, field_radius=0.9 * scene.h, capacity=scene.count + 1024, obstacle=floor,
                          pressure=opts, dynamic_viscosity=1.0)
    reports, status, failed_stage, retries, durations = [], H.OK, 0, 0, []
    t0 = time.perf_counter()
    for k in range(args.stirred_steps):
        for attempt in range(8):
            o["step_duration"] = dt
            o["solid"] = H.scene_solid(sphere=(c, r), paddle=(pv, e, angle, w), walls=(lo, hi), split_rule=3)
            s, rep = ctx.substep(o)
            if not (s == H.ABORTED and int(rep["stage"]) == 8):
                break
            retries += 1
            dt *= 0.5
        if s != H.OK:
            status, failed_stage = s, int(rep["stage"])
            break
        reports.append(rep)
        durations.append(dt)
        angle += w * dt
        dt = 0.5 * scene.h / max(1.0, float(rep["max_speed"]))
        if k % 10 == 0:
            log(f"stirred substep {k}: {int(rep['frozen_cells']):,} cells, {int(rep['strain_rows']):,} strain rows, "
                f"GCR {int(rep['viscous']['pressure_iterations'])} iterations, "
                f"{int(rep['viscous']['velocity_iterations'])} velocity PCG iterations, {rep['total_ms']:.1f} ms")
    wall = time.perf_counter() - t0
    assert ctx.set_transfer_precision(64) == H.OK
    ctx.close()
    out = {"workload": "stirred_dam_break_256", "particles": scene.count, "steps": len(reports),
           "requested_steps": args.stirred_steps, "status": int(status), "retries": retries,
           "simulated_s": float(sum(durations)), "dynamic_viscosity_pa_s": 1.0,
           "solid": "sphere r 0.15 at the centre, box paddle 0.36 x 0.2 x 0.04 turning at 2 rad/s about y "
                    "through (0.7, 0.12, 0.5), side walls, sampled at the corner keys (split rule 3); floor "
                    "half-space at y = -h/4",
           "strain": "full-cell strain rows of every pressure cell (cut cells not octant-weighted)",
           "transfers": "FP32 deposit arithmetic, FP64 gather", "preconditioner": "V-cycle of the effective "
           "conductances in the reference's GCR (restart 32)"}
    if failed_stage:
        out["failed_stage"] = failed_stage
    if not reports:
        return out
    timed = reports[1:] if len(reports) > 1 else reports
    stages = ("sort_ms", "field_ms", "deposit_ms", "freeze_ms", "links_ms", "projection_ms", "gather_ms", "update_ms",
              "total_ms")
    out["stage_ms"] = {k: float(np.mean([r[k] for r in timed])) for k in stages}
    out["wall_ms_per_step"] = wall / len(reports) * 1e3
    out["particle_steps_per_s"] = scene.count / (out["stage_ms"]["total_ms"] * 1e-3)
    out["gcr_iterations_mean"] = float(np.mean([int(r["viscous"]["pressure_iterations"]) for r in reports]))
    out["velocity_pcg_iterations_mean"] = float(np.mean([int(r["viscous"]["velocity_iterations"]) for r in reports]))
    out["max_cell_defect_max"] = float(max(r["viscous"]["max_cell_constraint_defect"] for r in reports))
    last = reports[-1]
    out["last"] = {k: int(last[k]) for k in ("candidates", "frozen_cells", "frozen_faces", "blocks", "strain_rows")}
    log(f"stirred: {len(reports)} substeps in {wall:.1f}s, {out['stage_ms']['total_ms']:.1f} ms per substep")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="two_spheres_256",
                    choices=["two_spheres_256", "dam_break_128", "dam_break_128_random", "dam_break_64"])
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--coarse-sweeps", type=int, default=2)
    ap.add_argument("--sweeps", type=int, default=3,
                    help="smoothing sweeps before and after each coarse correction (3: fewest milliseconds at 256^3)")
    ap.add_argument("--level-storage", choices=["fp32", "bf16"], default="bf16",
                    help="V-cycle coefficient storage (BF16 keeps FP32 accumulation)")
    ap.add_argument("--smoother", choices=["rbgs", "block", "hybrid"], default="hybrid",
                    help="V-cycle smoother: level-wide red-black Gauss-Seidel, red-black sweeps per 4^3 block, or "
                         "level-wide on the finest level and per block below")
    ap.add_argument("--cpu-sample", type=int, default=1 << 20)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="threads of the all-cores CPU transfer baseline (16 = the host share of one GPU)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-error", action="store_true",
                    help="skip the comparison of the GPU projection with a tight FP64 oracle solve")
    ap.add_argument("--copy-probe", action="store_true", help="only measure the plain-copy ceiling")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner comparison")
    ap.add_argument("--no-topology", action="store_true", help="skip the topology-preparation measurement")
    ap.add_argument("--no-viscosity", action="store_true", help="skip the viscosity measurement")
    ap.add_argument("--substeps", type=int, default=20,
                    help="consecutive device substeps over moving particles (0: skip)")
    ap.add_argument("--stirred-steps", type=int, default=100,
                    help="substeps of the viscous stirred dam break with level-set solids (0 skips)")
    ap.add_argument("--only-stirred", action="store_true", help="only the viscous stirred dam-break leg")
    ap.add_argument("--only-substep", action="store_true",
                    help="run only the substep leg (profiling): one JSON line with the substep object")
    ap.add_argument("--learn-pairs", type=int, default=50000)
    ap.add_argument("--learn-epochs", type=int, default=8)
    args = ap.parse_args()
    if args.copy_probe:
        print(json.dumps({"copy_ceiling_gbs": copy_ceiling(), "runtime_copy_gbs": runtime_copy_ceiling(),
                          "bytes_per_copy": 1 << 30}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; runs on more than one GPU are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.only_stirred:
        print(json.dumps({"stirred": stirred_measurement(args)}), flush=True)
        return
    t0 = time.perf_counter()
    scene = scenes.by_name(args.workload)
    cells = scene.pressure_cells()
    log(f"scene {scene.name}: {scene.count:,} particles, {len(cells):,} liquid cells ({time.perf_counter() - t0:.1f}s)")
    if args.only_substep:
        print(json.dumps({"workload": scene.name, "substep": substep_measurement(scene, cells, args)}), flush=True)
        return
    ctx = H.Context(particles=scene.count + 1024, blocks=1 << 17, lookup_blocks=1 << 22,
                    pressure_cells=len(cells) + 64, pressure_faces=4 * len(cells) + 4096)
    device = ctx.device_name()
    assert ctx.scene_upload(scene.geometry_record(), scene.positions, scene.velocities, scene.slopes,
                            scene.mass) == H.OK
    s = ctx.scene_prepare(cells, scene.n, scene.rho)
    assert s == H.OK, s
    opts = H.pressure_options(relative_target=1e-6, levels=args.levels, sweeps=args.sweeps, coarse_sweeps=args.coarse_sweeps,
                              bf16_levels=args.level_storage == "bf16", smoother=["rbgs", "block", "hybrid"].index(args.smoother))
    dt = scene.dt
    for _ in range(args.warmup):
        s, rep = ctx.step(dt, scene.gravity, QUANTUM, QUANTUM, scene.count, opts)
        assert s == H.OK, s
    reports = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s, rep = ctx.step(dt, scene.gravity, QUANTUM, QUANTUM, scene.count, opts)
        assert s == H.OK, s
        reports.append(rep)
    wall = time.perf_counter() - t0  # fl_step returns after its last kernel completed
    log(f"timed {args.steps} steps in {wall:.3f}s on {device}")

    def mean(field):
        return float(np.mean([getattr(r, field) for r in reports]))

    last = reports[-1]
    d = last.diagnostics()
    n, nf = int(last.particles), int(last.blocks) * 192
    p2g_ms, g2p_ms = mean("p2g_kernel_ms"), mean("g2p_kernel_ms")
    p2g_bytes = p2g_algorithmic_bytes(n, nf)
    achieved = p2g_bytes / (p2g_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": n * args.steps / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"f64 transfers; {args.level_storage} V-cycle coefficients, f32 CG vectors, f64 dots/residuals",
        "data": "synthetic (seeded scene generator)",
        "config": {
            "workload": scene.name,
            "particles": n,
            "pressure_cells": int(last.pressure_cells),
            "pressure_faces": int(last.faces),
            "active_blocks": int(last.blocks),
            "nonempty_bins": int(last.nonempty_bins),
            "grid": f"{scene.n}^3 sparse, 4^3 blocks",
            "multigrid_levels": int(d["multigrid_level_count"]),
            "level_storage": args.level_storage,
            "smoother": args.smoother,
            "smoothing_sweeps": args.sweeps,
            "relative_divergence_target": 1e-6,
            "device": device,
            "step": "bin + stable sort, P2G quanta, normalize + gravity, MGPCG projection, G2P on the scene's fixed "
                    "pressure structure (first-order free-surface links built once); particles are not advected",
        },
        "projection_ms": mean("projection_ms"),
        "projection_iterations": int(d["iteration_count"]),
        "relative_divergence": float(d["defect_l2"] / d["initial_defect_l2"]),
        "max_cell_defect_per_s": float(d["max_cell_constraint_defect"]),
        "p2g_particle_steps_per_s": n / (p2g_ms * 1e-3),
        "g2p_particle_steps_per_s": n / (g2p_ms * 1e-3),
        "stage_ms": {k: mean(k) for k in ("sort_ms", "p2g_ms", "normalize_ms", "projection_ms", "g2p_ms", "total_ms",
                                          "p2g_kernel_ms", "g2p_kernel_ms")},
        "roofline": {
            "kernel": "k_p2g",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic("k_p2g", scene.name),
            "algorithmic_bytes": p2g_bytes,
            "g2p_achieved": g2p_algorithmic_bytes(n, nf) / (g2p_ms * 1e-3) / 1e9,
            "copy_ceiling_gbs": copy_ceiling(),
            "runtime_copy_gbs": runtime_copy_ceiling(),
        },
    }
    out["roofline"]["vcycle"] = projection_roofline(ctx, args.level_storage)
    out["fp32_transfers"] = fp32_transfer_measurement(ctx, scene, dt, opts, args, n, nf)
    if not args.no_learned:
        out["learned"] = learned_comparison(ctx, scene, dt, opts, args)
    if not (args.no_cpu_baseline and args.no_oracle_error):
        pin = projection_inputs(scene, ctx, last)
        if not args.no_oracle_error:
            out["fp64_oracle_error"] = fp64_oracle_error(scene, pin)
        if not args.no_cpu_baseline:
            t0 = time.perf_counter()
            out["cpu_baseline"] = cpu_baseline(scene, pin, args.cpu_sample, args.cpu_threads)
            log(f"cpu baseline took {time.perf_counter() - t0:.1f}s")
        del pin
    if not args.no_viscosity:
        out["viscosity"] = viscosity_measurement(ctx, scene, dt, last)
    if not args.no_topology:  # last: span-based calls reuse the resident scene's buffers
        out["topology"] = topology_measurement(ctx, scene, cells, dt)
        out["particle_update"] = update_measurement(ctx)
        out["boundary"] = boundary_measurement(ctx, scene, cells, dt)
    ctx.close()
    if args.substeps > 0:
        out["substep"] = substep_measurement(scene, cells, args)
    if args.stirred_steps > 0:
        out["stirred"] = stirred_measurement(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
