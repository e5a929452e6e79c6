"""Pins the oracle's compatible (masked) deposit to the reference's known answers and invariants.

Known answers: native/tests/liquid_boundary_transfer_3d_cases.h:66-193
(MaskedDepositRecentersMomentumAndQuotientGradient, RejectsEmptyZeroSupportAndMissingExcludedOwner,
RoundsActualMaskedContributionsAndHonorsAffineOff). Invariants: the normalised compatible weights sum
to one, the recentred offsets have zero weighted mean, the quotient-rule gradients sum to zero, and
mass and momentum are conserved to the quanta's rounding; an all-compatible mask reproduces the plain
quantized deposit to rounding.
"""
import numpy as np
import pytest

import cases as K
import pyoracle as O

UNIT = O.geometry((0.0, 0.0, 0.0), 1.0)
Q = K.BOUNDARY_QUANTUM


def one_block():
    return K.sorted_blocks(UNIT, [(0, 0, 0)])


def slot(blocks, axis, lattice):
    s, v = O.face_slot(blocks, axis, lattice)
    assert s == O.OK
    return v


def deposit(blocks, particles, tags, q=Q, cap=1, stencils=True):
    return O.deposit_compatible(UNIT, blocks, particles, tags, q, q, cap, stencils=stencils)


def test_masked_deposit_recenters_momentum_and_quotient_gradient():
    blocks = one_block()
    p = K.boundary_particle()
    K.set_slopes(p, 0, ((100.0, 2.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)))
    tags = K.boundary_tags(len(blocks))
    for z in range(4):
        for y in range(4):
            tags[slot(blocks, 0, (1, y, z))]["side_tag"] = -7
    s, quanta, st = deposit(blocks, p, tags)
    assert s == O.OK
    center = slot(blocks, 0, (2, 1, 1))
    assert quanta[center]["mass_quanta_sum"] == 9 << 30
    assert quanta[center]["pure_momentum_quanta_sum"] == 27 << 30
    mass = momentum = 0.0
    grad_sum = np.zeros(3)
    centered_sum = np.zeros(3)
    probe = slot(blocks, 0, (2, 0, 1))
    for smp in st[0, 0]:
        fq = quanta[int(smp["face_slot"])]
        mass += float(fq["mass_quanta_sum"]) * Q
        momentum += float(fq["pure_momentum_quanta_sum"]) * Q
        assert fq["stress_force_quanta_sum"] == 0
        grad_sum += K.vec(smp["normalized_weight_gradient"])
        centered_sum += smp["normalized_weight"] * K.vec(smp["centered_face_offset"])
        if int(smp["face_slot"]) == probe:
            assert smp["normalized_weight"] == 0.09375
            np.testing.assert_allclose(K.vec(smp["normalized_weight_gradient"]), (0.0, -0.375, 0.0), atol=2e-12)
            np.testing.assert_allclose(K.vec(smp["centered_face_offset"]), (0.0, -1.0, 0.0), atol=2e-12)
    assert mass == pytest.approx(16.0, rel=4 * 2.0 ** -52)
    assert momentum == pytest.approx(48.0, rel=4 * 2.0 ** -52)
    np.testing.assert_allclose(grad_sum, 0.0, atol=2e-12)
    np.testing.assert_allclose(centered_sum, 0.0, atol=2e-12)
    # A fresh zero-mass field stays eligible: the centre slot normalises to the particle's velocity.
    s, _, u, v = O.normalize(quanta, Q, Q, 0.0)
    assert s == O.OK and v[center] == 1 and u[center] == 3.0


def test_rejects_empty_zero_support_and_missing_excluded_owner():
    blocks = one_block()
    p = K.boundary_particle()
    tags = K.boundary_tags(len(blocks), eligible=False)
    assert deposit(blocks, p, tags)[0] == O.FAILED_PRECONDITION
    # The only eligible sample has weight zero: no positive support.
    tags[slot(blocks, 0, (3, 1, 1))]["structurally_eligible"] = 1
    assert deposit(blocks, p, tags)[0] == O.FAILED_PRECONDITION
    tags = K.boundary_tags(len(blocks))
    K.set_position(p, 0, (3.75, 3.75, 3.75))
    assert deposit(blocks, p, tags)[0] == O.FAILED_PRECONDITION
    # Only zero-weight x = 4 samples lack owners; ownership is still required.
    K.set_position(p, 0, (2.5, 1.5, 1.5))
    assert deposit(blocks, p, tags)[0] == O.FAILED_PRECONDITION
    s, quanta, st = deposit(blocks, p[:0], tags)
    assert s == O.OK and np.all(quanta["mass_quanta_sum"] == 0) and st.shape == (0, 3, 27)


def test_rounds_actual_masked_contributions_and_honors_affine_off():
    blocks = one_block()
    tags = K.boundary_tags(len(blocks), eligible=False)
    for axis in range(3):
        tags[slot(blocks, axis, (1, 1, 1))]["structurally_eligible"] = 1
    p = K.boundary_particle(mass=0.5, velocity=(-3.0, 3.0, 5.0))
    s, quanta, _ = deposit(blocks, p, tags, q=1.0)
    assert s == O.OK
    assert quanta[slot(blocks, 0, (1, 1, 1))]["mass_quanta_sum"] == 0
    assert quanta[slot(blocks, 0, (1, 1, 1))]["pure_momentum_quanta_sum"] == -2
    assert quanta[slot(blocks, 2, (1, 1, 1))]["pure_momentum_quanta_sum"] == 2
    tags[slot(blocks, 0, (2, 1, 1))]["structurally_eligible"] = 1
    p[0]["transfer"]["particle_mass"] = 1.0
    K.set_slopes(p, 0, ((2.0 ** 66, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)))
    assert deposit(blocks, p, tags, q=1.0)[0] == O.OUT_OF_RANGE
    p[0]["affine_enabled"] = 0
    s, quanta, _ = deposit(blocks, p, tags, q=1.0)
    assert s == O.OK
    assert quanta[slot(blocks, 0, (1, 1, 1))]["pure_momentum_quanta_sum"] == -2


def test_disabled_affine_rows_are_ignored_even_when_not_finite():
    # liquid_boundary_transfer_3d.h:40 ("Disabled affine rows are ignored, even if stale"); an enabled
    # non-finite row is InvalidArgument (liquid_boundary_transfer_3d.cc:366-372).
    blocks = one_block()
    tags = K.boundary_tags(len(blocks))
    p = K.boundary_particle()
    K.set_slopes(p, 0, ((0.0, 0.0, 0.0), (np.nan, 0.0, 0.0), (0.0, 0.0, 0.0)))
    assert deposit(blocks, p, tags)[0] == O.INVALID_ARGUMENT
    p[0]["affine_enabled"] = 0
    s, quanta, _ = deposit(blocks, p, tags)
    assert s == O.OK
    p[0]["transfer"]["particle_mass"] = -1.0
    assert deposit(blocks, p, tags)[0] == O.INVALID_ARGUMENT


@pytest.mark.parametrize("seed", [3, 4])
def test_normalised_basis_recentring_and_conservation(seed):
    geom, blocks, p, tags = K.masked_scene(seed, 400)
    q = 2.0 ** -40
    s, quanta, st = O.deposit_compatible(geom, blocks, p, tags, q, q, len(p))
    assert s == O.OK, s
    h = float(geom["cell_width"][0])
    mass_total = np.zeros(3)
    mom_total = np.zeros(3)
    for i in range(len(p)):
        t = p[i]["transfer"]
        m = float(t["particle_mass"])
        v = K.vec(t["particle_velocity"])
        for axis in range(3):
            ss = st[i, axis]
            ok = ss["compatible"] != 0
            assert ok.any()
            nw = ss["normalized_weight"][ok]
            assert abs(nw.sum() - 1.0) <= 64 * 2.0 ** -52
            assert np.all(ss["normalized_weight"][~ok] == 0.0)
            c = np.stack([ss["centered_face_offset"][k] for k in "xyz"], 1)[ok]
            np.testing.assert_allclose((nw[:, None] * c).sum(0), 0.0, atol=1e-14 * h)
            g = np.stack([ss["normalized_weight_gradient"][k] for k in "xyz"], 1)[ok]
            np.testing.assert_allclose(g.sum(0), 0.0, atol=1e-12 / h)
            # N = w / S against the plain stencil's weights; the centred offsets are r minus one
            # common mean.
            s2, plain = O.owned_stencil(geom, blocks, axis, K.vec(t["sorted_world_position"]),
                                        K.vec(t["trial_world_position"]))
            assert s2 == O.OK
            assert np.array_equal(plain["face_slot"], ss["face_slot"])
            w = plain["weight"][ok]
            np.testing.assert_allclose(nw, w / w.sum(), rtol=1e-14, atol=1e-300)
            r = np.stack([plain["face_offset_from_particle"][k] for k in "xyz"], 1)[ok]
            shift = r - c
            np.testing.assert_allclose(shift, np.broadcast_to(shift[0], shift.shape), atol=1e-14 * h)
            mass_total[axis] += m
            mom_total[axis] += m * v[axis]
    # Conservation: every contribution is rounded once to a quantum; one component per face axis.
    axis_of_slot = (np.arange(len(quanta)) % 192) // 64
    contributions = int((st["compatible"] != 0).sum())
    for axis in range(3):
        sel = axis_of_slot == axis
        got_m = float(quanta["mass_quanta_sum"][sel].sum()) * q
        got_p = float(quanta["pure_momentum_quanta_sum"][sel].sum()) * q
        assert abs(got_m - mass_total[axis]) <= 0.5 * q * contributions + 1e-12 * mass_total[axis]
        assert abs(got_p - mom_total[axis]) <= 0.5 * q * contributions + 1e-11 * np.abs(mom_total).max()


def test_all_compatible_mask_reproduces_the_plain_deposit_to_rounding():
    geom, blocks, p, tags = K.masked_scene(5, 300, eligible=1.0, sides=1, affine=1.0)
    q = 2.0 ** -40
    s, compat, _ = O.deposit_compatible(geom, blocks, p, tags, q, q, len(p), stencils=False)
    assert s == O.OK
    plain_particles = np.ascontiguousarray(p["transfer"])
    s, plain = O.deposit_quanta(geom, blocks, plain_particles, q, q, len(p))
    assert s == O.OK
    # S = sum w is one and sum N r is zero in exact arithmetic; in FP64 each contribution may round
    # to a neighbouring quantum, and the recentring moves a momentum by |C| times the rounding of
    # the mean offset.
    dm = np.abs(compat["mass_quanta_sum"] - plain["mass_quanta_sum"])
    dp = np.abs(compat["pure_momentum_quanta_sum"] - plain["pure_momentum_quanta_sum"])
    assert dm.max() <= 64
    assert dp.max() <= 1e-9 * np.abs(plain["pure_momentum_quanta_sum"]).max()
    assert dm.max() > 0 or dp.max() > 0  # the two deposits are distinct computations


def test_quantizes_each_particle_before_summing_and_includes_guard_endpoint():
    # liquid_boundary_transfer_3d_cases.h:423-451
    blocks = one_block()
    tags = K.boundary_tags(len(blocks), eligible=False)
    for axis in range(3):
        tags[slot(blocks, axis, (1, 1, 1))]["structurally_eligible"] = 1
    p = np.zeros(2, O.BOUNDARY_PARTICLE)  # (np.concatenate would drop the record padding)
    p[:] = K.boundary_particle(mass=0.5, velocity=(-3.0, 3.0, 5.0))[0]
    s, quanta, _ = deposit(blocks, p, tags, q=1.0, cap=2)
    x = slot(blocks, 0, (1, 1, 1))
    assert s == O.OK and quanta[x]["mass_quanta_sum"] == 0 and quanta[x]["pure_momentum_quanta_sum"] == -4
    for i in range(2):
        t = p[i]["transfer"]
        t["particle_mass"] = 1.0
        t["particle_velocity"]["x"], t["particle_velocity"]["y"], t["particle_velocity"]["z"] = 2.0 ** 61, 0.0, 0.0
        p[i]["transfer"] = t
    s, quanta, _ = deposit(blocks, p, tags, q=1.0, cap=2)
    assert s == O.OK and quanta[x]["pure_momentum_quanta_sum"] == 1 << 62
    t = p[1]["transfer"]
    t["particle_velocity"]["x"] = np.nextafter(2.0 ** 61, np.inf)
    p[1]["transfer"] = t
    assert deposit(blocks, p, tags, q=1.0, cap=2)[0] == O.OUT_OF_RANGE


def test_complete_boundary_deposit_preserves_unmasked_api_at_signed_translated_scale():
    # liquid_boundary_transfer_3d_cases.h:452-473: one block at (-1,-1,-1), origin (11,-17,31), h = 2.
    geom = O.geometry((11.0, -17.0, 31.0), 2.0)
    blocks = K.sorted_blocks(geom, [(-1, -1, -1)])
    p = K.boundary_particle(position=(6.5, -21.5, 26.5), mass=2.0, velocity=(1.0, 2.0, 3.0), side=-3)
    K.set_slopes(p, 0, ((2.0, -1.0, 3.0), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))
    tags = K.boundary_tags(len(blocks), side=-3)
    q = 2.0 ** -30
    s, plain = O.deposit_quanta(geom, blocks, np.ascontiguousarray(p["transfer"]), q, q, 1)
    assert s == O.OK
    s, quanta, _ = O.deposit_compatible(geom, blocks, p, tags, q, q, 1)
    assert s == O.OK
    assert np.array_equal(quanta["mass_quanta_sum"], plain["mass_quanta_sum"])
    assert np.array_equal(quanta["pure_momentum_quanta_sum"], plain["pure_momentum_quanta_sum"])
