"""GPU parity of the MAC pressure projection (fl_project) and of the device-resident step (fl_step).

Bars: the integrated RHS agrees with the oracle to 2e-12 (it is an FP64 compensated sum of the same
terms); projected pressures and velocities agree with the FP64 dense or tight-tolerance oracle to the
stated tolerance of each test; every accepted projection meets the reference gates
(mac_pressure_3d_multigrid.cc:341-367) and the requested relative divergence.
"""
import math

import numpy as np
import pytest

import cases as K
import pyoracle as O
from fluidity_amd import hip as H
from fluidity_amd import records as R
from fluidity_amd import scenes

pytestmark = pytest.mark.gpu
FIXTURE = K.load_pressure_fixture()


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(particles=1 << 23, blocks=1 << 16, lookup_blocks=1 << 21, pressure_cells=1 << 21,
                  pressure_faces=1 << 23)
    print("device", c.device_name(), flush=True)
    yield c
    c.close()


@pytest.mark.parametrize("name", sorted(FIXTURE))
def test_manufactured_cases(ctx, name):
    case = FIXTURE[name]
    cells, faces, blocks = K.pressure_case_records(case)
    geom = O.geometry((0.0, 0.0, 0.0), 1.0)
    sb = K.sorted_blocks(geom, blocks)
    s, out = ctx.project(geom, sb, case["dt"], cells, faces, H.pressure_options(relative_target=1e-13, levels=3))
    assert s == H.OK, s
    e = case["expected"]
    np.testing.assert_allclose(out["rhs"], e["rhs"], rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(out["pressures"], e["pressures"], rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(out["velocities"], e["velocities"], rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(out["defects"], e["defects"], atol=1e-10)
    assert list(out["gauges"]) == e["gauges"]
    np.testing.assert_allclose(out["residuals"], -cells["fluid_volume"] * out["defects"], atol=1e-12)


@pytest.mark.parametrize("anchored", [False, True])
def test_affine_field_with_all_six_boundaries(ctx, anchored):
    geom, sb, cells, faces, pressure = K.affine_problem(-3, (4, 4, 4), anchored)
    s, out = ctx.project(geom, sb, 1.0, cells, faces, H.pressure_options(levels=3))
    assert s == H.OK
    d = out["diagnostics"]
    assert d["multigrid_level_count"] >= 2 and d["iteration_count"] > 0
    assert d["max_volume_scaled_residual"] <= 2e-5 and d["max_cell_constraint_defect"] <= 2e-5
    assert d["true_residual_l2"] <= 2e-6 + 2e-6 * np.linalg.norm(out["rhs"])
    want = np.array([pressure(tuple(c)) for c in cells["cell_index"].tolist()]) + (0.0 if anchored else 21.0)
    np.testing.assert_allclose(out["pressures"], want, atol=5e-5)
    np.testing.assert_allclose(out["velocities"], 0.0, atol=5e-5)
    np.testing.assert_allclose(out["residuals"], -out["defects"], atol=1e-12)


def test_rejects_invalid_inputs_like_the_oracle(ctx):
    geom, sb, cells, faces, _ = K.affine_problem(-3, (4, 4, 4), True)
    assert ctx.project(geom, sb, 1.0, cells, faces, H.pressure_options(max_iterations=0))[0] == H.INVALID_ARGUMENT
    assert ctx.project(geom, sb, 1.0, cells, faces, H.pressure_options(max_iterations=257))[0] == H.INVALID_ARGUMENT
    swapped = faces.copy()
    swapped[[0, 1]] = swapped[[1, 0]]
    assert ctx.project(geom, sb, 1.0, cells, swapped, H.pressure_options())[0] == H.INVALID_ARGUMENT
    assert O.project(geom, sb, 1.0, cells, swapped, O.pressure_options())[0] == O.INVALID_ARGUMENT
    missing = sb[1:].copy()
    assert ctx.project(geom, missing, 1.0, cells, faces, H.pressure_options())[0] == H.FAILED_PRECONDITION
    bad = cells.copy()
    bad[3]["fluid_volume"] = -1.0
    assert ctx.project(geom, sb, 1.0, bad, faces, H.pressure_options())[0] == H.INVALID_ARGUMENT
    # one iteration cannot meet a 1e-12 relative target: non-convergence is Aborted
    s, _ = ctx.project(geom, sb, 1.0, cells, faces, H.pressure_options(max_iterations=1, relative_target=1e-12))
    assert s == H.ABORTED


@pytest.mark.parametrize("seed", [0, 3])
def test_scene_projection_matches_fp64_oracle(ctx, seed):
    sc = scenes.dam_break(32, seed=seed, random_velocity=(seed == 3))
    st = K.oracle_step_inputs(sc)
    args = (st["geom"], st["blocks"], st["dt"], st["cells"], st["faces"])
    s, ref = O.project(*args, O.pressure_options(relative_target=1e-12, fp32=False, levels=4))
    assert s == O.OK
    s, out = ctx.project(*args, H.pressure_options(relative_target=1e-6, levels=4))
    assert s == H.OK
    d = out["diagnostics"]
    print("gpu iterations", d["iteration_count"], "restarts", d["restarts"], "rel", d["defect_l2"] / d["initial_defect_l2"],
          flush=True)
    assert d["defect_l2"] <= 1e-6 * d["initial_defect_l2"]
    np.testing.assert_allclose(out["rhs"], ref["rhs"], rtol=2e-12, atol=2e-12 * np.abs(ref["rhs"]).max())
    # Pressure and velocity error against the FP64 oracle, relative to their ranges.
    p_err = np.abs(out["pressures"] - ref["pressures"]).max() / np.abs(ref["pressures"]).max()
    u_err = np.abs(out["velocities"] - ref["velocities"]).max() / np.abs(ref["velocities"]).max()
    print("relative pressure error", p_err, "velocity error", u_err, flush=True)
    assert p_err <= 1e-5 and u_err <= 1e-5
    s, fp32 = O.project(*args, O.pressure_options(relative_target=1e-6, fp32=True, levels=4))
    assert s == O.OK
    assert abs(int(d["iteration_count"]) - int(fp32["diagnostics"]["iteration_count"])) <= 3


def test_bf16_level_storage_meets_the_same_gates(ctx):
    # BF16 V-cycle coefficients change the preconditioner only: the accepted projection satisfies the
    # same FP64 gates and agrees with the FP64 oracle to the same tolerance.
    sc = scenes.dam_break(32, seed=0)
    st = K.oracle_step_inputs(sc)
    args = (st["geom"], st["blocks"], st["dt"], st["cells"], st["faces"])
    s, ref = O.project(*args, O.pressure_options(relative_target=1e-12, fp32=False, levels=4))
    assert s == O.OK
    s, f32 = ctx.project(*args, H.pressure_options(relative_target=1e-6, levels=4))
    assert s == H.OK
    s, b16 = ctx.project(*args, H.pressure_options(relative_target=1e-6, levels=4, bf16_levels=True))
    assert s == H.OK
    d = b16["diagnostics"]
    print("iterations fp32", f32["diagnostics"]["iteration_count"], "bf16", d["iteration_count"], flush=True)
    assert d["defect_l2"] <= 1e-6 * d["initial_defect_l2"]
    p_err = np.abs(b16["pressures"] - ref["pressures"]).max() / np.abs(ref["pressures"]).max()
    assert p_err <= 1e-5
    assert int(d["iteration_count"]) <= int(f32["diagnostics"]["iteration_count"]) + 3


@pytest.mark.parametrize("bf16,smoother", [(False, 1), (True, 1), (True, 2)], ids=["fp32_block", "bf16_block",
                                                                               "bf16_hybrid"])
def test_block_smoother_meets_the_gates_deterministically(ctx, bf16, smoother):
    # Red-black sweeps inside each 4^3 block with the block's neighbours held fixed: a different
    # (symmetric) preconditioner, so the accepted projection meets the same FP64 gates, agrees with the
    # FP64 oracle, and repeats bit for bit (no block reads a halo that another block is writing).
    sc = scenes.dam_break(64, seed=1)
    st = K.oracle_step_inputs(sc)
    args = (st["geom"], st["blocks"], st["dt"], st["cells"], st["faces"])
    s, ref = O.project(*args, O.pressure_options(relative_target=1e-12, fp32=False, levels=5))
    assert s == O.OK
    s, rb = ctx.project(*args, H.pressure_options(relative_target=1e-6, levels=5, coarse_sweeps=2, bf16_levels=bf16))
    assert s == H.OK
    opts = H.pressure_options(relative_target=1e-6, levels=5, coarse_sweeps=2, bf16_levels=bf16, smoother=smoother)
    s, blk = ctx.project(*args, opts)
    assert s == H.OK
    s, again = ctx.project(*args, opts)
    assert s == H.OK
    d = blk["diagnostics"]
    print("iterations rbgs", rb["diagnostics"]["iteration_count"], "block", d["iteration_count"], flush=True)
    assert d["defect_l2"] <= 1e-6 * d["initial_defect_l2"]
    p_err = np.abs(blk["pressures"] - ref["pressures"]).max() / np.abs(ref["pressures"]).max()
    assert p_err <= 1e-5
    assert np.array_equal(blk["pressures"], again["pressures"])
    assert int(d["iteration_count"]) <= 2 * int(rb["diagnostics"]["iteration_count"]) + 4


def test_device_step_matches_oracle_pipeline(ctx):
    sc = scenes.dam_break(24, seed=2, random_velocity=True)
    st = K.oracle_step_inputs(sc)
    geom = sc.geometry_record()
    assert ctx.scene_upload(geom, sc.positions, sc.velocities, sc.slopes, sc.mass) == H.OK
    assert ctx.scene_prepare(sc.pressure_cells(), sc.n, sc.rho) == H.OK
    opts = H.pressure_options(relative_target=1e-6, levels=4)
    s, rep = ctx.step(st["dt"], sc.gravity, K.QUANTUM, K.QUANTUM, sc.count, opts)
    assert s == H.OK, s
    nb = int(rep.blocks)
    s, got = ctx.scene_download(nb, sc.count, len(st["cells"]))
    assert s == H.OK
    assert np.array_equal(got["blocks"].view(np.int32), st["blocks"].view(np.int32))
    assert np.array_equal(got["quanta"].view(np.int64), st["quanta"].view(np.int64))
    assert np.array_equal(got["ustar"], st["ustar"])
    s, want_order = O.sort_particles(st["geom"], sc.positions)
    assert np.array_equal(got["order"].view(np.uint8), want_order.view(np.uint8))
    # projection on the device-derived faces equals the oracle's projection of the same inputs
    s, ref = O.project(st["geom"], st["blocks"], st["dt"], st["cells"], st["faces"],
                       O.pressure_options(relative_target=1e-12, fp32=False, levels=4))
    assert s == O.OK
    slots = np.array([O.face_slot(st["blocks"], int(f["axis"]), tuple(int(v) for v in f["lattice"]))[1]
                      for f in st["faces"]])
    u_err = np.abs(got["unew"][slots] - ref["velocities"]).max() / np.abs(ref["velocities"]).max()
    assert u_err <= 1e-5
    d = rep.diagnostics()
    assert d["defect_l2"] <= 1e-6 * d["initial_defect_l2"]
    # G2P of the step against the oracle gather of the same face fields
    s, want = O.gather(st["geom"], st["blocks"], st["particles"], got["u0"], got["unew"], threads=8)
    assert s == O.OK
    a, b = got["gather"].view(np.float64).reshape(-1, 24), want.view(np.float64).reshape(-1, 24)
    scale = np.maximum(np.abs(b).max(0), 1.0)
    assert np.all(np.abs(a - b) <= 2e-12 * np.abs(b) + 2e-12 * scale)


@pytest.mark.parametrize("bf16", [False, True], ids=["fp32_levels", "bf16_levels"])
@pytest.mark.parametrize("name", ["dam_break_128", "two_spheres_256"])
def test_full_size_step_meets_divergence_target(name, bf16):
    # The benchmark configurations at their real sizes: 5 levels, the hybrid smoother, FP32 or BF16 level
    # coefficients (FP32 accumulation). At 128^3 the accepted projection is also compared with a tight FP64
    # oracle solve of the same frozen system.
    sc = scenes.by_name(name)
    print(name, "particles", sc.count, "bf16", bf16, flush=True)
    cells = sc.pressure_cells()
    with H.Context(particles=sc.count + 1024, blocks=1 << 17, lookup_blocks=1 << 22, pressure_cells=len(cells) + 64,
                   pressure_faces=4 * len(cells) + 4096) as c:
        assert c.scene_upload(sc.geometry_record(), sc.positions, sc.velocities, sc.slopes, sc.mass) == H.OK
        assert c.scene_prepare(cells, sc.n, sc.rho) == H.OK
        opts = H.pressure_options(relative_target=1e-6, levels=5, coarse_sweeps=2, bf16_levels=bf16, smoother=2)
        s, rep = c.step(sc.dt, sc.gravity, K.QUANTUM, K.QUANTUM, sc.count, opts)
        assert s == H.OK, s
        d = rep.diagnostics()
        print("step ms", rep.total_ms, "projection ms", rep.projection_ms, "iterations", d["iteration_count"],
              "rel", d["defect_l2"] / d["initial_defect_l2"], flush=True)
        assert d["defect_l2"] <= 1e-6 * d["initial_defect_l2"]
        assert d["max_cell_constraint_defect"] <= 2e-5 and d["max_volume_scaled_residual"] <= 2e-5
        # determinism: a second step reproduces the quanta and the projection bit for bit
        s, first = c.scene_download(int(rep.blocks), 0, len(cells), what=("quanta", "unew"))
        s, rep2 = c.step(sc.dt, sc.gravity, K.QUANTUM, K.QUANTUM, sc.count, opts)
        assert s == H.OK
        s, second = c.scene_download(int(rep.blocks), 0, len(cells), what=("quanta", "unew"))
        assert np.array_equal(first["quanta"], second["quanta"]) and np.array_equal(first["unew"], second["unew"])
        if name == "dam_break_128":
            args, p, u = K.device_step_system(c, sc, rep)
            K.assert_matches_fp64_oracle(args, p, u)
