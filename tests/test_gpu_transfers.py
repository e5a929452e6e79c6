"""GPU parity of the bin sort, activation, quantized P2G, normalize/forces and G2P against the oracle.

Bar: bit-exact for bins, sort order, active blocks and int64 face quanta; bit-exact for the decoded
face velocities; G2P (FP64) within 2e-12 relative plus 2e-12 absolute, the reference test tolerance
(native/tests/test_transfers_3d.cc:57-59), scaled by the field magnitude for random fields.
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

UNIT = O.geometry((0.0, 0.0, 0.0), 1.0)
TRANSLATED = O.geometry((1.0, -3.0, 5.0), 2.0)


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(particles=1 << 23, blocks=1 << 16, lookup_blocks=1 << 21)
    print("device", c.device_name(), flush=True)
    yield c
    c.close()


def corner_blocks(geom=UNIT):
    return K.sorted_blocks(geom, K.CORNER_BLOCKS)


def test_corner_quanta_bit_exact(ctx):
    blocks = corner_blocks()
    p = K.corner_particle()
    for q in (1.0, 2.0 ** -20, 2.0 ** -40):
        s, want = O.deposit_quanta(UNIT, blocks, p, q, q, 1)
        assert s == O.OK
        s, got = ctx.p2g_quanta(UNIT, blocks, p, q, q, 1)
        assert s == H.OK
        assert np.array_equal(got.view(np.int64), want.view(np.int64))


def test_ties_guard_and_error_codes_match_oracle(ctx):
    blocks = corner_blocks()
    for v in (-2.5, -1.5, -0.5, 0.5, 1.5, 2.5):
        p = K.tie_particle(8.0, v)
        assert np.array_equal(ctx.p2g_quanta(UNIT, blocks, p, 1.0, 1.0, 1)[1], O.deposit_quanta(UNIT, blocks, p, 1.0, 1.0, 1)[1])
    guard = K.tie_particle(8.0, 0.0)
    guard[0]["particle_velocity_slopes"]["x_row"]["x"] = -(2.0 ** 61)
    four = np.concatenate([guard] * 4)
    three = four[:3].copy()
    s, got = ctx.p2g_quanta(UNIT, blocks, three, 1.0, 1.0, 3)
    assert s == H.OK and np.array_equal(got, O.deposit_quanta(UNIT, blocks, three, 1.0, 1.0, 3)[1])
    three[2]["particle_velocity_slopes"]["x_row"]["x"] = np.nextafter(-(2.0 ** 61), -math.inf)
    sentinel = np.zeros(192 * len(blocks), R.FACE_QUANTA)
    sentinel[:] = (17, -18, 19)
    out = sentinel.copy()
    assert ctx.p2g_quanta(UNIT, blocks, three, 1.0, 1.0, 3, out=out)[0] == H.OUT_OF_RANGE
    assert np.array_equal(out, sentinel)
    assert ctx.p2g_quanta(UNIT, blocks, four, 1.0, 1.0, 3)[0] == H.RESOURCE_EXHAUSTED
    for bad in (0.0, -1.0, 3.0, 5e-324, math.inf):
        assert ctx.p2g_quanta(UNIT, blocks, four[:2], bad, 1.0, 3)[0] == H.INVALID_ARGUMENT
    for bad_mass in (-1.0, math.inf, math.nan):
        pair = four[:2].copy()
        pair[1]["particle_mass"] = bad_mass
        assert ctx.p2g_quanta(UNIT, blocks, pair, 1.0, 1.0, 3)[0] == H.INVALID_ARGUMENT
    # incomplete ownership and excess motion are FailedPrecondition, as in the oracle
    part = K.sorted_blocks(UNIT, K.CORNER_BLOCKS[:1] + K.CORNER_BLOCKS[2:])
    assert ctx.p2g_quanta(UNIT, part, K.corner_particle(), 1.0, 1.0, 1)[0] == H.FAILED_PRECONDITION
    p = K.corner_particle()
    p[0]["trial_world_position"]["x"] += 0.5000001
    assert ctx.p2g_quanta(UNIT, blocks, p, 1.0, 1.0, 1)[0] == H.FAILED_PRECONDITION


def test_first_failing_particle_decides_the_status(ctx):
    # The reference stops at the first failing particle in caller order (transfers_3d.cc:75-101).
    blocks = corner_blocks()
    ps = np.concatenate([K.corner_particle()] * 6)
    ps[5]["particle_mass"] = math.nan                       # InvalidArgument at ordinal 5
    ps[3]["trial_world_position"]["x"] += 0.75              # FailedPrecondition (motion) at ordinal 3
    for order in (np.arange(6), np.arange(6)[::-1]):
        sub = ps[order].copy()
        want = O.deposit_quanta(UNIT, blocks, sub, 1.0, 1.0, 6)[0]
        assert ctx.p2g_quanta(UNIT, blocks, sub, 1.0, 1.0, 6)[0] == want


def random_scene_particles(seed, n=16, count=None):
    sc = scenes.dam_break(n, seed=seed, random_velocity=True)
    return sc


@pytest.mark.parametrize("seed", [0, 1])
def test_scene_quanta_sort_and_activation_bit_exact(ctx, seed):
    sc = scenes.dam_break(24, seed=seed, random_velocity=True)
    st = K.oracle_step_inputs(sc)
    g = st["geom"]
    pos = sc.positions
    # sort
    s, got = ctx.bin_sort(g, pos)
    assert s == H.OK
    s, want = O.sort_particles(g, pos)
    assert s == O.OK
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    # activation (particles plus pressure cells)
    cells = np.zeros(len(sc.liquid_cells()), R.INT3)
    cells["x"], cells["y"], cells["z"] = sc.liquid_cells().T
    s, blocks = ctx.activate(g, pos, cells)
    assert s == H.OK
    assert np.array_equal(blocks.view(np.int32), st["blocks"].view(np.int32))
    # quanta, shuffled caller order must not change a bit
    parts = st["particles"]
    s, q = ctx.p2g_quanta(g, st["blocks"], parts, K.QUANTUM, K.QUANTUM, len(parts))
    assert s == H.OK
    assert np.array_equal(q.view(np.int64), st["quanta"].view(np.int64))
    perm = np.random.default_rng(seed).permutation(len(parts))
    s, q2 = ctx.p2g_quanta(g, st["blocks"], parts[perm].copy(), K.QUANTUM, K.QUANTUM, len(parts))
    assert s == H.OK and np.array_equal(q2, q)
    # normalize + forces
    s, m, u0, v, ustar = ctx.normalize_forces(q, K.QUANTUM, K.QUANTUM, 0.0, sc.gravity, st["dt"])
    assert s == H.OK
    assert np.array_equal(m, st["mass"]) and np.array_equal(u0, st["u0"]) and np.array_equal(v, st["validity"])
    assert np.array_equal(ustar, st["ustar"])


def test_sort_known_answers(ctx):
    pos = [(2.0, 2.0, 2.0), (-2.0, 2.0, 2.0), (2.1, 2.0, 2.0), (2.0, -2.0, 2.0), (2.0, 2.0, -2.0), (-2.1, 2.0, 2.0)]
    s, e = ctx.bin_sort(UNIT, np.array(pos))
    assert s == H.OK and list(e["original_ordinal"]) == [4, 3, 1, 5, 0, 2]
    s, e = ctx.bin_sort(O.geometry((-1.5, -1.5, -1.5), 1.0), np.array([(0.0, 0.0, 0.0), (-2.0 ** -54, 0.0, 0.0)]))
    assert s == H.OK and list(e["original_ordinal"]) == [1, 0]
    assert e["bin_index"]["x"][0] == -1 and e["bin_index"]["x"][1] == 0
    assert ctx.bin_sort(UNIT, np.array([(math.nan, 0.0, 0.0)]))[0] == H.INVALID_ARGUMENT
    s, blocks = ctx.activate(UNIT, np.array([(3.0, 3.0, 3.0)]))
    assert s == H.OK
    assert [tuple(b) for b in blocks] == [(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)]


def test_large_bin_segments_sort_stably(ctx):
    # 10k particles in a single bin exercise the merge path of the in-segment sort.
    rng = np.random.default_rng(3)
    pos = rng.uniform(1.6, 5.4, size=(10000, 3))
    s, got = ctx.bin_sort(UNIT, pos)
    assert s == H.OK
    s, want = O.sort_particles(UNIT, pos)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("geom", [UNIT, TRANSLATED], ids=["unit", "translated"])
def test_affine_and_quadratic_gather_known_answers(ctx, geom):
    blocks = corner_blocks(geom)
    org = np.array([geom["world_origin"][c][0] for c in "xyz"])
    h = float(geom["cell_width"][0])
    part = K.corner_particle(tuple(org), h)
    xp = K.vec(part[0]["trial_world_position"])
    old = K.sample_field(geom, blocks, lambda a, anc: K.OLD_VELOCITY[a])
    new = K.sample_field(geom, blocks, lambda a, anc: K.PARTICLE_VELOCITY[a] + np.dot(K.AFFINE_ROWS[a], anc - xp))
    s, res = ctx.g2p(geom, blocks, part, old, new)
    assert s == H.OK
    r = res[0]
    np.testing.assert_allclose(K.vec(r["new_pic_velocity"]), K.PARTICLE_VELOCITY, rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(K.vec(r["grid_velocity_increment"]), (-1.0, -5.0, 8.0), rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(K.mat(r["fitted_velocity_slopes"]), K.AFFINE_ROWS, rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(K.mat(r["velocity_gradient"]), K.AFFINE_ROWS, rtol=2e-12, atol=2e-12)
    K.set_vec(part[0], "sorted_world_position", (xp[0] - h / 2, xp[1] + h / 2, xp[2] - h / 2))
    new = K.sample_field(geom, blocks, lambda a, anc: (anc[a] - xp[a]) ** 2)
    s, res = ctx.g2p(geom, blocks, part, np.zeros_like(new), new)
    assert s == H.OK
    pic, slope = (0.25, -0.1875) if h == 1.0 else (1.0, -0.375)
    np.testing.assert_allclose(K.vec(res[0]["new_pic_velocity"]), [pic] * 3, rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(K.mat(res[0]["fitted_velocity_slopes"]), np.diag([slope] * 3), rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(K.mat(res[0]["velocity_gradient"]), 0.0, atol=2e-12)


def test_scene_gather_matches_oracle(ctx):
    sc = scenes.dam_break(24, seed=5, random_velocity=True)
    st = K.oracle_step_inputs(sc)
    rng = np.random.default_rng(9)
    nf = len(st["u0"])
    old = rng.normal(0.0, 1.0, nf)
    new = rng.normal(0.0, 1.0, nf)
    s, want = O.gather(st["geom"], st["blocks"], st["particles"], old, new, threads=8)
    assert s == O.OK
    s, got = ctx.g2p(st["geom"], st["blocks"], st["particles"], old, new)
    assert s == H.OK
    a = got.view(np.float64).reshape(-1, 24)
    b = want.view(np.float64).reshape(-1, 24)
    # Gradient and slope rows carry a factor 1/h; compare against each column's scale.
    scale = np.maximum(np.abs(b).max(0), 1.0)
    assert np.all(np.abs(a - b) <= 2e-12 * np.abs(b) + 2e-12 * scale)
    bad = old.copy()
    bad[7] = math.nan
    assert ctx.g2p(st["geom"], st["blocks"], st["particles"], bad, new)[0] == H.INVALID_ARGUMENT


def test_fp64_deposit_matches_oracle(ctx):
    # One particle: every slot receives one contribution, so the sums are exact and bit-identical.
    blocks = corner_blocks()
    p = K.corner_particle()
    s, want = O.deposit(UNIT, blocks, p)
    assert s == O.OK
    s, got = ctx.p2g(UNIT, blocks, p)
    assert s == H.OK
    assert np.array_equal(got.view(np.float64), want.view(np.float64))
    # A scene: parallel FP64 sums agree with the particle-order sums to rounding.
    sc = scenes.dam_break(24, seed=5, random_velocity=True)
    st = K.oracle_step_inputs(sc)
    s, want = O.deposit(st["geom"], st["blocks"], st["particles"])
    assert s == O.OK
    s, got = ctx.p2g(st["geom"], st["blocks"], st["particles"])
    assert s == H.OK
    a, b = got.view(np.float64).reshape(-1, 2), want.view(np.float64).reshape(-1, 2)
    scale = np.abs(b).max(0)
    assert np.all(np.abs(a - b) <= 1e-12 * np.abs(b) + 1e-14 * scale)
    total = st["particles"]["particle_mass"].sum()
    for axis in range(3):
        m = a.reshape(-1, 3, 64, 2)[:, axis, :, 0].sum()
        assert abs(m - total) <= 1e-12 * total
    # Failures keep the output and report the reference's codes.
    bad = st["particles"].copy()
    bad[3]["particle_mass"] = -1.0
    assert ctx.p2g(st["geom"], st["blocks"], bad)[0] == H.INVALID_ARGUMENT
    bad = st["particles"].copy()
    bad[5]["particle_mass"] = 1e300
    bad[5]["particle_velocity_slopes"]["x_row"]["y"] = 1e300
    out = np.zeros(192 * len(st["blocks"]), R.FACE_DEPOSIT)
    out["deposited_mass"] = 3.0
    assert ctx.p2g(st["geom"], st["blocks"], bad, out)[0] == H.OUT_OF_RANGE
    assert O.deposit(st["geom"], st["blocks"], bad)[0] == O.OUT_OF_RANGE
    assert np.all(out["deposited_mass"] == 3.0)


def test_normalize_error_codes(ctx):
    q = np.zeros(192, R.FACE_QUANTA)
    q[5] = (3, 1, 0)
    q[9]["stress_force_quanta_sum"] = 1
    s, *_ = ctx.normalize_forces(q, 1.0, 1.0, 0.0, (0.0, -9.81, 0.0), 0.1)
    assert s == O.normalize(q.view(O.FACE_QUANTA), 1.0, 1.0)[0] == H.INVALID_ARGUMENT
    q[9]["stress_force_quanta_sum"] = 0
    s, m, u, v, p = ctx.normalize_forces(q, 1.0, 1.0, 0.0, (0.0, -9.81, 0.0), 0.1)
    assert s == H.OK and u[5] == 1.0 / 3.0 and v[5] == 1


@pytest.mark.parametrize("name", ["dam_break_128", "dam_break_128_random"])
def test_full_size_quanta_match_oracle_and_conserve_mass(ctx, name):
    # 128^3 sparse dam break (~4.0M particles): bit-exact quanta against the multithreaded oracle,
    # plus exact conservation of the quantized mass per face axis (F2).
    sc = scenes.by_name(name, seed=0)
    print(name, sc.count, flush=True)
    st = K.oracle_step_inputs(sc, threads=16)
    s, q = ctx.p2g_quanta(st["geom"], st["blocks"], st["particles"], K.QUANTUM, K.QUANTUM, len(st["particles"]))
    assert s == H.OK
    print("gpu kernel ms", ctx.last.kernel_ms, flush=True)
    assert np.array_equal(q.view(np.int64), st["quanta"].view(np.int64))
    axis_of = (np.arange(len(q)) % 192) // 64
    per_particle = round(sc.mass / K.QUANTUM)  # the particle mass is an exact multiple of the quantum here
    for a in range(3):
        total = int(q["mass_quanta_sum"][axis_of == a].sum())
        assert abs(total - per_particle * sc.count) <= 27 * sc.count // 2
