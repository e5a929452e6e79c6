"""Pins the CPU oracle's grid, stencil and transfer restatement to the reference's known answers.

Each test names the reference test whose expectations it reproduces; values that the reference
generates at build time are recomputed independently (tests/golden/make_transfer_fixture.py).
"""
import math

import numpy as np
import pytest

import cases as K
import pyoracle as O

UNIT = O.geometry((0.0, 0.0, 0.0), 1.0)
TRANSLATED = O.geometry((1.0, -3.0, 5.0), 2.0)


def corner_blocks(geom=UNIT):
    return K.sorted_blocks(geom, K.CORNER_BLOCKS)


def test_corner_deposit_matches_exact_fixture():
    # MacTransfer3D.AffineCornerMatchesOracle (native/tests/test_transfers_3d.cc:145-203)
    fx = K.load_transfer_fixture()
    blocks = corner_blocks()
    s, dep = O.deposit(UNIT, blocks, K.corner_particle())
    assert s == O.OK
    seen = np.zeros(len(dep), bool)
    for (axis, x, y, z), m, q in zip(fx["faces"], fx["mass_contribution"], fx["momentum_contribution"]):
        s, slot = O.face_slot(blocks, axis, (x, y, z))
        assert s == O.OK and not seen[slot]
        seen[slot] = True
        assert dep[slot]["deposited_mass"] == pytest.approx(m, rel=2e-12, abs=2e-12)
        assert dep[slot]["deposited_pure_momentum"] == pytest.approx(q, rel=2e-12, abs=2e-12)
    assert np.all(dep["deposited_mass"][~seen] == 0.0)
    assert np.all(dep["deposited_pure_momentum"][~seen] == 0.0)


def test_corner_stencil_matches_exact_weights_offsets_and_gradients():
    fx = K.load_transfer_fixture()
    blocks = corner_blocks()
    p = K.corner_particle()[0]
    want = {tuple(f): (w, o, g) for f, w, o, g in zip(fx["faces"], fx["weight"], fx["offset"], fx["gradient"])}
    for axis in range(3):
        s, st = O.owned_stencil(UNIT, blocks, axis, K.vec(p["sorted_world_position"]), K.vec(p["trial_world_position"]))
        assert s == O.OK
        for smp in st:
            s, (a, lat) = O.slot_face(blocks, int(smp["face_slot"]))
            assert s == O.OK and a == axis
            w, o, g = want.get((axis, *lat), (0.0, None, None))
            assert smp["weight"] == pytest.approx(w, abs=2e-12)
            if o is not None:
                np.testing.assert_allclose(K.vec(smp["face_offset_from_particle"]), o, atol=2e-12)
                np.testing.assert_allclose(K.vec(smp["weight_gradient"]), g, atol=2e-12)


@pytest.mark.parametrize("geom", [UNIT, TRANSLATED], ids=["unit", "translated"])
def test_affine_gather_reproduces_velocity_and_slopes(geom):
    # ExpectAffineGather (native/tests/test_transfers_3d.cc:110-143)
    blocks = corner_blocks(geom)
    org = np.array([geom["world_origin"][c][0] for c in "xyz"])
    h = float(geom["cell_width"][0])
    part = K.corner_particle(tuple(org), h)
    xp = K.vec(part[0]["trial_world_position"])
    old = K.sample_field(geom, blocks, lambda a, anc: K.OLD_VELOCITY[a])
    new = K.sample_field(geom, blocks, lambda a, anc: K.PARTICLE_VELOCITY[a] + np.dot(K.AFFINE_ROWS[a], anc - xp))
    part[0]["particle_mass"] = math.nan  # gather ignores deposit-only data
    s, res = O.gather(geom, blocks, part, old, new)
    assert s == O.OK
    r = res[0]
    np.testing.assert_allclose(K.vec(r["new_pic_velocity"]), K.PARTICLE_VELOCITY, rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(K.vec(r["grid_velocity_increment"]), (-1.0, -5.0, 8.0), rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(K.mat(r["fitted_velocity_slopes"]), K.AFFINE_ROWS, rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(K.mat(r["velocity_gradient"]), K.AFFINE_ROWS, rtol=2e-12, atol=2e-12)


@pytest.mark.parametrize("geom", [UNIT, TRANSLATED], ids=["unit", "translated"])
def test_quadratic_row_distinguishes_fit_and_gradient(geom):
    # MacTransfer3D.QuadraticRowDistinguishesFitAndGradient (test_transfers_3d.cc:205-244)
    blocks = corner_blocks(geom)
    org = np.array([geom["world_origin"][c][0] for c in "xyz"])
    h = float(geom["cell_width"][0])
    part = K.corner_particle(tuple(org), h)
    xp = K.vec(part[0]["trial_world_position"])
    K.set_vec(part[0], "sorted_world_position", (xp[0] - h / 2, xp[1] + h / 2, xp[2] - h / 2))
    old = np.zeros(192 * len(blocks))
    new = K.sample_field(geom, blocks, lambda a, anc: (anc[a] - xp[a]) ** 2)
    s, res = O.gather(geom, blocks, part, old, new)
    assert s == O.OK
    pic = 0.25 if h == 1.0 else 1.0
    slope = -0.1875 if h == 1.0 else -0.375
    r = res[0]
    np.testing.assert_allclose(K.vec(r["new_pic_velocity"]), [pic] * 3, rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(K.vec(r["grid_velocity_increment"]), [pic] * 3, rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(K.mat(r["fitted_velocity_slopes"]), np.diag([slope] * 3), rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(K.mat(r["velocity_gradient"]), np.zeros((3, 3)), atol=2e-12)


def test_rounds_each_contribution_half_to_even_before_summation():
    # MacTransfer3D.RoundsContributionsBeforeSummation (test_transfers_3d.cc:800-862)
    blocks = corner_blocks()
    s, slot = O.face_slot(blocks, *K.TIE_FACE)
    assert s == O.OK
    two = np.concatenate([K.tie_particle(4.0, 1.0), K.tie_particle(4.0, 1.0)])
    s, q = O.deposit_quanta(UNIT, blocks, two, 1.0, 1.0, 2)
    assert s == O.OK
    assert tuple(q[slot]) == (0, 0, 0)  # two half quanta each round to zero
    for v, want in zip((-2.5, -1.5, -0.5, 0.5, 1.5, 2.5), (-2, -2, 0, 0, 2, 2)):
        s, q = O.deposit_quanta(UNIT, blocks, K.tie_particle(8.0, v), 1.0, 1.0, 1)
        assert s == O.OK
        assert q[slot]["mass_quanta_sum"] == 1 and q[slot]["pure_momentum_quanta_sum"] == want


def test_quanta_are_order_invariant_and_within_half_quantum_per_contributor():
    blocks = corner_blocks()
    ps = np.concatenate([K.tie_particle(8.0, v) for v in (-2.5, -1.5, -0.5, 0.5, 1.5, 2.5)] + [K.tie_particle(6.0, -3.0)])
    s, q = O.deposit_quanta(UNIT, blocks, ps, 0.5, 2.0, len(ps))
    assert s == O.OK
    s, qr = O.deposit_quanta(UNIT, blocks, ps[::-1].copy(), 0.5, 2.0, len(ps))
    assert s == O.OK
    assert np.array_equal(q, qr)
    s, dep = O.deposit(UNIT, blocks, ps)
    assert s == O.OK
    n = np.where(dep["deposited_mass"] > 0, len(ps), 0)
    assert np.all(np.abs(q["mass_quanta_sum"] * 0.5 - dep["deposited_mass"]) <= n * 0.5 / 2)
    assert np.all(np.abs(q["pure_momentum_quanta_sum"] * 2.0 - dep["deposited_pure_momentum"]) <= n * 2.0 / 2)


def test_rejects_capacity_scale_range_and_invalid_mass():
    # MacTransfer3D.RejectsCapacityRangeAndAliasingWithoutPublication (test_transfers_3d.cc:864-1000)
    blocks = corner_blocks()
    s, slot = O.face_slot(blocks, *K.TIE_FACE)
    guard = K.tie_particle(8.0, 0.0)
    guard[0]["particle_velocity_slopes"]["x_row"]["x"] = -(2.0 ** 61)
    four = np.concatenate([guard] * 4)
    sentinel = np.zeros(192 * len(blocks), O.FACE_QUANTA)
    sentinel[:] = (17, -18, 19)
    assert O.deposit_quanta(UNIT, blocks, four[:2], 1.0, 1.0, 0)[0] == O.RESOURCE_EXHAUSTED
    assert O.deposit_quanta(UNIT, blocks, four, 1.0, 1.0, 3)[0] == O.RESOURCE_EXHAUSTED
    assert O.deposit_quanta(UNIT, blocks, four[:0], 1.0, 1.0, (1 << 20) + 1)[0] == O.RESOURCE_EXHAUSTED
    for bad in (0.0, -1.0, 3.0, 5e-324, math.inf):
        assert O.deposit_quanta(UNIT, blocks, four[:2], bad, 1.0, 3)[0] == O.INVALID_ARGUMENT
        assert O.deposit_quanta(UNIT, blocks, four[:2], 1.0, bad, 3)[0] == O.INVALID_ARGUMENT
    three = four[:3].copy()
    three[2]["particle_velocity_slopes"]["x_row"]["x"] = np.nextafter(-(2.0 ** 61), -math.inf)
    out = sentinel.copy()
    assert O.deposit_quanta(UNIT, blocks, three, 1.0, 1.0, 3, out=out)[0] == O.OUT_OF_RANGE
    assert np.array_equal(out, sentinel)  # preserved on failure
    for bad_mass in (-1.0, math.inf, math.nan):
        pair = four[:2].copy()
        pair[1]["particle_mass"] = bad_mass
        assert O.deposit(UNIT, blocks, pair)[0] == O.INVALID_ARGUMENT
        assert O.deposit_quanta(UNIT, blocks, pair, 1.0, 1.0, 3)[0] == O.INVALID_ARGUMENT
    s, q = O.deposit_quanta(UNIT, blocks, four[:3], 1.0, 1.0, 3)
    assert s == O.OK
    assert q[slot]["mass_quanta_sum"] == 3 and q[slot]["pure_momentum_quanta_sum"] == 3 << 60
    s, q = O.deposit_quanta(UNIT, blocks, four[:0], 1.0, 1.0, 0, out=sentinel.copy())
    assert s == O.OK and not q.view(np.int64).any()


def test_translated_support_edge_and_underflow_known_answers():
    # SparseMacGrid3D.PreservesTranslatedSupportEdge ... PreservesRepresentableGradient
    # (native/tests/test_sparse_mac_grid_3d.cc:448-556)
    g = O.geometry((-3.0, 0.0, 0.0), 1.0)
    b = K.sorted_blocks(g, [(0, 0, 0)])
    pt = (-2.0 ** -53, 2.25, 2.25)
    s, st = O.owned_stencil(g, b, 1, pt, pt)
    assert s == O.OK and st[0]["face_slot"] == 85
    assert st[12]["weight"] > 0.0 and st[12]["weight_gradient"]["x"] == float.fromhex("-0x1.e4p-55")

    g = O.geometry((-3.3, 0.0, 0.0), 1.1)
    b = K.sorted_blocks(g, [(0, 0, 0)])
    pt = (np.nextafter(2.0 ** -51, 0.0), 2.5, 2.5)
    s, st = O.owned_stencil(g, b, 1, pt, pt)
    assert s == O.OK and st[0]["face_slot"] == 85 and st[12]["weight_gradient"]["x"] < 0.0

    g = O.geometry((-2.5, 0.0, 0.0), 1.0)
    b = K.sorted_blocks(g, [(0, 0, 0)])
    pt = (-2.0 ** -53, 2.25, 2.25)
    s, st = O.owned_stencil(g, b, 1, pt, pt)
    assert s == O.OK and st[13]["face_offset_from_particle"]["x"] == 2.0 ** -53

    assert O.compute_bin(O.geometry((-5.5, 0.0, 0.0), 1.0), (-2.0 ** -53, 2.25, 2.25)) == (O.OK, (0, 0, 0))

    big = np.finfo(np.float64).max
    assert O.compute_bin(O.geometry((-big, 0.0, 0.0), big), (big, 0.0, 0.0))[1][0] == 0
    s, anchor = O.face_world_anchor(O.geometry((-big, 0.0, 0.0), big), 0, (2, 0, 0))
    assert s == O.OK and anchor == (big, big / 2, big / 2)
    s, anchor = O.face_world_anchor(O.geometry((-0.5, 0.0, 0.0), 0.1), 0, (5, 2, 2))
    assert s == O.OK and anchor[0] == 2.0 ** -55

    h = 2.0 ** -300
    g = O.geometry((0.0, 0.0, 0.0), h)
    blocks = [(x, y, z) for z in (-1, 0) for y in (-1, 0) for x in (-1, 0)]
    b = K.sorted_blocks(g, blocks)
    pt = (2.0 ** -900, h / 4, h / 4)
    s, st = O.owned_stencil(g, b, 2, pt, pt)
    assert s == O.OK and st[11]["weight_gradient"]["y"] == float.fromhex("-0x1.08p-902")


def test_incomplete_support_and_motion_are_failed_preconditions():
    blocks = K.sorted_blocks(UNIT, K.CORNER_BLOCKS[:1] + K.CORNER_BLOCKS[2:])
    p = K.corner_particle()
    assert O.deposit_quanta(UNIT, blocks, p, 1.0, 1.0, 1)[0] == O.FAILED_PRECONDITION
    blocks = corner_blocks()
    p = K.corner_particle()
    p[0]["trial_world_position"]["x"] += 0.5000001
    assert O.deposit_quanta(UNIT, blocks, p, 1.0, 1.0, 1)[0] == O.FAILED_PRECONDITION
    p = K.corner_particle()
    p[0]["trial_world_position"]["y"] = math.inf
    assert O.gather(UNIT, blocks, p, np.zeros(1536), np.zeros(1536))[0] == O.INVALID_ARGUMENT


def test_slot_map_inverse_and_signed_ownership():
    # LiquidStepWorkspace3D.BorrowedGridChecksCanonicalKeysAndSignedSlotInverse
    # (native/tests/test_liquid_step_workspace_3d.cc:350-380)
    blocks = K.sorted_blocks(UNIT, [(-1, -1, -1), (0, -1, -1)])
    assert O.slot_face(blocks, 384)[0] == O.OUT_OF_RANGE
    assert O.face_slot(blocks, 3, (0, 0, 0))[0] == O.INVALID_ARGUMENT
    assert O.face_slot(blocks, 0, (0, 0, 0))[0] == O.FAILED_PRECONDITION
    assert O.slot_face(blocks, 191) == (O.OK, (2, (-1, -1, -1)))
    assert O.slot_face(blocks, 192) == (O.OK, (0, (0, -4, -4)))
    assert O.slot_face(blocks, 383) == (O.OK, (2, (3, -1, -1)))
    for slot in range(384):
        s, (a, lat) = O.slot_face(blocks, slot)
        assert O.face_slot(blocks, a, lat) == (O.OK, slot)
    assert O.sort_blocks(UNIT, O.int3([(0, 0, 0), (0, 0, 0)]))[0] == O.INVALID_ARGUMENT
    assert O.sort_blocks(UNIT, O.int3([(536870912, 0, 0)]))[0] == O.OUT_OF_RANGE


def test_sort_orders_signed_bins_and_keeps_equal_bins_stable():
    # LiquidStepWorkspace3D.NegativeBinsAndEqualBinsPreserveIncomingOrder (test_liquid_step_workspace_3d.cc:52-73)
    pos = [(2.0, 2.0, 2.0), (-2.0, 2.0, 2.0), (2.1, 2.0, 2.0), (2.0, -2.0, 2.0), (2.0, 2.0, -2.0), (-2.1, 2.0, 2.0)]
    s, e = O.sort_particles(UNIT, pos)
    assert s == O.OK
    assert list(e["original_ordinal"]) == [4, 3, 1, 5, 0, 2]
    assert tuple(e["bin_index"][0]) == (0, 0, -1) and tuple(e["bin_index"][2]) == (-1, 0, 0)
    # CompensatedSortBoundaryRetainsNegativeSideOfTranslatedTie (:339-351)
    s, e = O.sort_particles(O.geometry((-1.5, -1.5, -1.5), 1.0), [(0.0, 0.0, 0.0), (-2.0 ** -54, 0.0, 0.0)])
    assert s == O.OK
    assert list(e["original_ordinal"]) == [1, 0]
    assert e["bin_index"]["x"][0] == -1 and e["bin_index"]["x"][1] == 0


def test_activation_includes_zero_weight_support_owners():
    # LiquidStepWorkspace3D.ActivatesZeroWeightOriginalNodesRegardlessOfTag (test_liquid_step_workspace_3d.cc:75-105)
    s, blocks = O.activate_blocks(UNIT, [(3.0, 3.0, 3.0)])
    assert s == O.OK
    assert [tuple(b) for b in blocks] == [(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)]


def random_particles(rng, n, lo, hi, h=1.0, sigma_v=1.0, sigma_c=0.1):
    p = np.zeros(n, O.PARTICLE)
    pos = rng.uniform(lo, hi, size=(n, 3))
    for i, c in enumerate("xyz"):
        p["sorted_world_position"][c] = pos[:, i]
        p["trial_world_position"][c] = pos[:, i]
        p["particle_velocity"][c] = rng.normal(0, sigma_v, n)
        for r in ("x_row", "y_row", "z_row"):
            p["particle_velocity_slopes"][r][c] = rng.normal(0, sigma_c, n)
    p["particle_mass"] = rng.uniform(0.5, 2.0, n)
    return p


def test_partition_of_unity_and_moment_identities():
    # F1 basis identities (docs/references/formal-contracts.md:26-34): sum w = 1, sum w r = 0,
    # sum w r r^T = h^2/4 I, sum grad w = 0, sum r grad w^T = I.
    rng = np.random.default_rng(7)
    for h in (1.0, 0.125):
        g = O.geometry((0.25, -1.0, 3.0), h)
        ps = random_particles(rng, 20, 4.0, 12.0)
        org = np.array([0.25, -1.0, 3.0])
        for c, name in enumerate("xyz"):
            ps["trial_world_position"][name] = org[c] + h * ps["trial_world_position"][name]
            ps["sorted_world_position"][name] = ps["trial_world_position"][name]
        s, blocks = O.activate_blocks(g, np.stack([ps["trial_world_position"][c] for c in "xyz"], 1))
        assert s == O.OK
        for p in ps:
            x = K.vec(p["trial_world_position"])
            for axis in range(3):
                s, st = O.owned_stencil(g, blocks, axis, x, x)
                assert s == O.OK
                w = st["weight"]
                r = np.stack([st["face_offset_from_particle"][c] for c in "xyz"], 1)
                gw = np.stack([st["weight_gradient"][c] for c in "xyz"], 1)
                assert w.sum() == pytest.approx(1.0, abs=1e-14)
                np.testing.assert_allclose(w @ r, 0.0, atol=1e-13 * h)
                np.testing.assert_allclose((w[:, None] * r).T @ r, np.eye(3) * h * h / 4, atol=1e-13 * h * h)
                np.testing.assert_allclose(gw.sum(0), 0.0, atol=1e-12 / h)
                np.testing.assert_allclose(r.T @ gw, np.eye(3), atol=1e-12)


def test_deposit_conserves_mass_and_momentum_per_component():
    # F2 (formal-contracts.md:36-48): per component, sum m_f = sum m_p and sum q_f = sum m_p v_a.
    rng = np.random.default_rng(11)
    ps = random_particles(rng, 200, 2.0, 14.0)
    pos = np.stack([ps["trial_world_position"][c] for c in "xyz"], 1)
    s, blocks = O.activate_blocks(UNIT, pos)
    assert s == O.OK
    s, dep = O.deposit(UNIT, blocks, ps)
    assert s == O.OK
    axis_of = (np.arange(len(dep)) % 192) // 64
    for a, c in enumerate("xyz"):
        sel = axis_of == a
        assert dep["deposited_mass"][sel].sum() == pytest.approx(ps["particle_mass"].sum(), rel=1e-12)
        assert dep["deposited_pure_momentum"][sel].sum() == pytest.approx(
            (ps["particle_mass"] * ps["particle_velocity"][c]).sum(), rel=1e-11, abs=1e-11)


def test_normalize_and_forces_closed_forms():
    q = np.zeros(4, O.FACE_QUANTA)
    q[0] = (0, 0, 0)
    q[1] = (3, 5, 0)
    q[2] = (1 << 20, -(3 << 19), 0)
    q[3] = (1 << 10, 7, 0)
    s, m, u, v = O.normalize(q, 2.0 ** -20, 2.0 ** -20, min_mass=4 * 2.0 ** -20)
    assert s == O.OK
    assert list(v) == [0, 3, 1, 1]
    assert m[2] == 1.0 and u[2] == -1.5 and u[1] == 0.0 and u[3] == 7.0 / 1024
    q[3]["stress_force_quanta_sum"] = 1
    assert O.normalize(q, 2.0 ** -20, 2.0 ** -20)[0] == O.INVALID_ARGUMENT
    blocks = corner_blocks()
    old = np.arange(1536, dtype=np.float64)
    s, pred = O.apply_forces(blocks, old, (0.0, -9.81, 1.0), 0.5)
    assert s == O.OK
    axis_of = (np.arange(1536) % 192) // 64
    np.testing.assert_array_equal(pred, old + 0.5 * np.array([0.0, -9.81, 1.0])[axis_of])
