"""GPU parity of the device substep (fl_substep) and its new stages against the oracle.

* The particle-union field (fl_sample_particle_field): candidate cells, centre and corner samples and
  their distances bit-identical to the oracle's restatement of PrepareField / FieldDistance, on the
  reference's two-particle case (liquid_tank_staging_3d_cases.h:15-58: 42 candidates, 170 centres, 92
  corners), on a dam break and on a non-power-of-two, offset grid; the capacity refusal.
* The projection with the system and hierarchy built on the device (fl_project_device_built) equals
  fl_project's bit for bit (same hierarchy, same arithmetic) on a dam break's typed links.
* Three substeps of a 16^3 dam break in the unit-cube tank (8 particles per cell, r = 0.9 h) against the
  oracle's chain of the same stages: every integer count equal at every step, the first step's
  candidates, blocks, frozen cells and faces and links identical; particle positions within 1e-9 h,
  velocities within 1e-7 of the largest speed and volume ratios within 1e-10 of the oracle's after
  each step (the device projection is the FP32 V-cycle PCG with FP64 restarts, solved to a relative
  divergence of 1e-10, against the oracle's FP64 solve to 1e-12).
* Three substeps of a 16^3 dam break stirred by a rotating paddle: the scene solid (sphere, paddle, side
  walls) sampled at the corner keys for the sampled-solid freeze, the floor as a half-space obstacle, and
  the coupled viscous projection on full-cell strain rows, against the oracle chain of the same stages.
"""
import numpy as np
import pytest

import pyoracle as O
import substep_chain as SC
from fluidity_amd import hip as H
from fluidity_amd import records as R
from fluidity_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(particles=1 << 16, blocks=1 << 12, lookup_blocks=1 << 16, pressure_cells=1 << 15,
                  pressure_faces=1 << 17)
    print("device", c.device_name(), flush=True)
    yield c
    c.close()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def check_field(ctx, geom, pos, radius, capacity=1 << 24):
    s_want, wc, wcen, wcor = O.particle_field(geom, pos, radius, capacity)
    s_got, gc, gcen, gcor = ctx.sample_particle_field(geom, pos, radius, capacity)
    assert s_got == s_want, (s_got, s_want)
    if s_want != O.OK:
        return None
    assert len(gc) == len(wc) and len(gcen) == len(wcen) and len(gcor) == len(wcor)
    assert np.array_equal(bits(gc), bits(wc))
    for got, want in ((gcen, wcen), (gcor, wcor)):
        assert np.array_equal(bits(got["cell_index"]), bits(want["cell_index"]))
        g, w = got["signed_distance"], want["signed_distance"]
        bad = np.nonzero(g.view(np.uint64) != w.view(np.uint64))[0]
        assert len(bad) == 0, (len(bad), len(g), bad[:5], g[bad[:5]], w[bad[:5]], np.abs(g - w).max())
    return wc, wcen, wcor


def test_field_reference_case(ctx):
    geom = O.geometry((0.0, 0.0, 0.0), 1.0)
    pos = np.array([[-0.5, 0.5, 3.5], [0.5, 0.5, 4.5], [-0.5, 0.5, 3.5]])
    c, cen, cor = check_field(ctx, geom, pos, 0.75)
    assert (len(c), len(cen), len(cor)) == (42, 170, 92)


def test_field_dam_break(ctx):
    sc = scenes.dam_break(16, seed=5)
    geom = O.geometry((0.0, 0.0, 0.0), sc.h)
    check_field(ctx, geom, sc.positions, 0.9 * sc.h)


def test_field_offset_grid(ctx):
    rng = np.random.default_rng(11)
    h = 0.07
    geom = O.geometry((-0.3, 0.1, 0.2), h)
    pos = np.array([-0.3, 0.1, 0.2]) + rng.random((3000, 3)) * h * np.array([9.0, 7.0, 11.0])
    check_field(ctx, geom, pos, 1.3 * h)
    check_field(ctx, geom, pos[:17], 0.45 * h)


def test_field_capacity(ctx):
    geom = O.geometry((0.0, 0.0, 0.0), 1.0)
    pos = np.array([[-0.5, 0.5, 3.5], [0.5, 0.5, 4.5]])
    s, *_ = ctx.sample_particle_field(geom, pos, 0.75, capacity=41)
    assert s == H.RESOURCE_EXHAUSTED


def first_step_links(sc, dt):
    geom = O.geometry((0.0, 0.0, 0.0), sc.h)
    st = SC.State(sc.positions, sc.velocities, sc.slopes, sc.mass)
    _, info = SC.oracle_substep(geom, st, dt, 0.9 * sc.h, SC.tank_obstacle(),
                                O.pressure_options(relative_target=1e-12, levels=3, fp32=False))
    return geom, info


def test_device_built_projection_equals_host_built(ctx):
    sc = scenes.dam_break(16, seed=7)
    dt = 0.25 * sc.h
    geom, info = first_step_links(sc, dt)
    blocks, cells, faces = info["blocks"], info["pressure_cells"], info["pressure_faces"]
    for levels, storage in ((3, False), (2, True)):
        opts = H.pressure_options(relative_target=1e-8, levels=levels, sweeps=2, bf16_levels=storage)
        s, host = ctx.project(geom, blocks, dt, cells, faces, opts)
        assert s == H.OK
        s, p, u, diag = ctx.project_device_built(geom, blocks, dt, cells, faces, opts)
        assert s == H.OK
        assert diag["iteration_count"] == host["diagnostics"]["iteration_count"]
        assert diag["multigrid_level_count"] == host["diagnostics"]["multigrid_level_count"]
        assert np.array_equal(p.view(np.uint64), host["pressures"].view(np.uint64))
        assert np.array_equal(u.view(np.uint64), host["velocities"].view(np.uint64))
    want = info["projection"]
    assert np.max(np.abs(p - want["pressures"])) <= 1e-5 * np.max(np.abs(want["pressures"]))


def test_substep_chain_matches_oracle(ctx):
    sc = scenes.dam_break(16, seed=3)
    geom = O.geometry((0.0, 0.0, 0.0), sc.h)
    h = sc.h
    radius = 0.9 * h
    obstacle = SC.tank_obstacle()
    assert ctx.scene_upload(geom, sc.positions, sc.velocities, sc.slopes.reshape(-1, 9), sc.mass) == H.OK
    st = SC.State(sc.positions, sc.velocities, sc.slopes, sc.mass)
    ob = SC.obstacle_record(obstacle, R.OBSTACLE)
    for step in range(3):
        vmax = float(np.sqrt((st.v ** 2).sum(1)).max())
        dt = 0.25 * h / max(1.0, vmax)
        print("step", step, "dt", dt, flush=True)
        o = H.substep_options(dt, field_radius=radius, capacity=1 << 16, obstacle=ob,
                              pressure=H.pressure_options(relative_target=1e-10, levels=3, sweeps=2))
        s, rep = ctx.substep(o)
        assert s == H.OK, (s, int(rep["stage"]))
        st, info = SC.oracle_substep(geom, st, dt, radius, obstacle,
                                     O.pressure_options(relative_target=1e-12, levels=3, fp32=False),
                                     capacity=1 << 16)
        assert rep["candidates"] == len(info["candidates"])
        assert rep["centre_samples"] == len(info["centres"]) and rep["corner_samples"] == len(info["corners"])
        assert rep["blocks"] == len(info["blocks"])
        assert rep["frozen_cells"] == len(info["frozen_cells"]) and rep["frozen_faces"] == len(info["frozen_faces"])
        assert rep["lower_theta_count"] == info["lower"]
        assert rep["pressure"]["defect_l2"] <= 1e-10 * rep["pressure"]["initial_defect_l2"]
        s, x, v, c, j, f = ctx.substep_download(len(st.x))
        assert s == H.OK
        speed = max(float(np.abs(st.v).max()), 1e-3)
        dx = np.abs(x - st.x).max()
        dv = np.abs(v - st.v).max()
        dj = np.abs(j - st.ratio).max()
        print(f"step {step}: |dx| {dx / h:.3e} h, |dv| {dv / speed:.3e} of max speed, |dJ| {dj:.3e}, "
              f"iterations {int(rep['pressure']['iteration_count'])}", flush=True)
        assert dx <= 1e-9 * h
        assert dv <= 1e-7 * speed
        assert dj <= 1e-10
        assert np.all(f == 1)


def test_failed_substep_leaves_the_scene(ctx):
    """A step too long for the half-cell guard fails in the update (Aborted, stage 8) and leaves the
    particles untouched; the same scene then steps with a shorter duration."""
    sc = scenes.dam_break(16, seed=4)
    sc.velocities[:, 0] = 3.0
    geom = O.geometry((0.0, 0.0, 0.0), sc.h)
    assert ctx.scene_upload(geom, sc.positions, sc.velocities, sc.slopes.reshape(-1, 9), sc.mass) == H.OK
    ob = SC.obstacle_record(SC.tank_obstacle(), R.OBSTACLE)
    o = H.substep_options(0.5 * sc.h, field_radius=0.9 * sc.h, capacity=1 << 16, obstacle=ob,
                          pressure=H.pressure_options(relative_target=1e-6, levels=3))
    s, rep = ctx.substep(o)
    assert s == H.ABORTED and int(rep["stage"]) == 8
    s, x, v, c, j, f = ctx.substep_download(sc.count)
    assert s == H.OK
    assert np.array_equal(x, sc.positions) and np.array_equal(v, sc.velocities)
    assert np.all(j == 1.0) and np.all(f == 0)
    o["step_duration"] = 0.1 * sc.h
    s, rep = ctx.substep(o)
    assert s == H.OK
    assert 0 < rep["max_speed"] < 10.0


def floor_record():
    return SC.obstacle_record(O.halfspace((0.0, 1.0, 0.0), 0.0), R.OBSTACLE)


def solid_record(solid, angle):
    (c, r), (pv, e, _, w), (lo, hi) = solid["sphere"], solid["paddle"], solid["walls"]
    return H.scene_solid(sphere=(c, r), paddle=(pv, e, angle, w), walls=(lo, hi))


@pytest.mark.parametrize("jacobi", [True, False])
def test_substep_chain_with_solids_and_viscosity(ctx, jacobi):
    """Three substeps of a 16^3 dam break stirred by a rotating paddle (sampled-solid freeze of the scene
    solid: sphere, paddle, side walls; floor half-space) with the coupled viscous projection (1 Pa s)
    against the oracle chain of the same stages. With the reference's Jacobi preconditioner the device
    follows the oracle's GCR over the three steps: every count equal, positions within 1e-9 h, velocities
    1e-7 of the largest speed. With the V-cycle both accept at the reference's gates, so one step agrees
    to the gates' accuracy (positions within 1e-6 h, velocities 1e-4); later steps are not compared, as a
    difference of that size can move one cell across the field's classification."""
    sc, solid = scenes.stirred_dam_break(16, seed=3, paddle_half_extent=(0.15, 0.1, 0.06),
                                         paddle_pivot=(0.35, 0.12, 0.5))
    geom = O.geometry((0.0, 0.0, 0.0), sc.h)
    h = sc.h
    radius = 0.9 * h
    floor = O.halfspace((0.0, 1.0, 0.0), 0.0)
    assert ctx.scene_upload(geom, sc.positions, sc.velocities, sc.slopes.reshape(-1, 9), sc.mass) == H.OK
    st = SC.State(sc.positions, sc.velocities, sc.slopes, sc.mass)
    angle = 0.0
    popt = H.pressure_options(levels=3, sweeps=2)
    if jacobi:
        popt["preconditioner"] = 2
    for step in range(3 if jacobi else 1):
        vmax = float(np.sqrt((st.v ** 2).sum(1)).max())
        dt = 0.25 * h / max(1.0, vmax)
        print("step", step, "dt", dt, "angle", angle, flush=True)
        o = H.substep_options(dt, field_radius=radius, capacity=1 << 16, obstacle=floor_record(), pressure=popt,
                              solid=solid_record(solid, angle), dynamic_viscosity=1.0)
        s, rep = ctx.substep(o)
        assert s == H.OK, (s, int(rep["stage"]))
        sol = dict(solid)
        sol["paddle"] = (solid["paddle"][0], solid["paddle"][1], angle, solid["paddle"][3])
        st, info = SC.oracle_substep(geom, st, dt, radius, floor, None, capacity=1 << 16, solid=sol, viscosity=1.0)
        angle += solid["paddle"][3] * dt
        assert rep["candidates"] == len(info["candidates"])
        assert rep["corner_samples"] == len(info["corners"])
        assert rep["blocks"] == len(info["blocks"])
        assert rep["frozen_cells"] == len(info["frozen_cells"]) and rep["frozen_faces"] == len(info["frozen_faces"])
        assert rep["strain_rows"] == len(info["rows"])
        v = rep["viscous"]
        want_it = int(info["projection"]["diagnostics"]["pressure_iterations"])
        print(f"step {step}: coupled iterations {int(v['pressure_iterations'])} (oracle {want_it}), velocity solves "
              f"{int(v['velocity_solves'])}, max defect {v['max_cell_constraint_defect']:.3e}", flush=True)
        assert v["max_cell_constraint_defect"] <= 2e-5
        if jacobi:
            assert abs(int(v["pressure_iterations"]) - want_it) <= 1
        s, x, vel, c, j, f = ctx.substep_download(len(st.x))
        assert s == H.OK
        speed = max(float(np.abs(st.v).max()), 1e-3)
        dx, dv, dj = np.abs(x - st.x).max(), np.abs(vel - st.v).max(), np.abs(j - st.ratio).max()
        print(f"step {step}: |dx| {dx / h:.3e} h, |dv| {dv / speed:.3e}, |dJ| {dj:.3e}", flush=True)
        if jacobi:
            assert dx <= 1e-9 * h and dv <= 1e-7 * speed and dj <= 1e-10
        else:
            assert dx <= 1e-6 * h and dv <= 1e-4 * speed and dj <= 1e-7
