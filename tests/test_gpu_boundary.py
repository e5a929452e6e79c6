"""GPU parity of the compatible (masked) deposit, fl_p2g_compatible, against the oracle.

Bar: bit-exact int64 face quanta, bit-exact compatible-stencil records (slot, flag, normalised weight,
quotient-rule gradient, recentred offset) and the same status code as the oracle, on the reference's
known-answer cases (native/tests/liquid_boundary_transfer_3d_cases.h:66-193), on random masks with
two sides, partial eligibility, disabled affine rows and trial motion, at a non-power-of-two cell
width, and on the error paths.
"""
import numpy as np
import pytest

import cases as K
import pyoracle as O
from fluidity_amd import hip as H
from fluidity_amd import records as R

pytestmark = pytest.mark.gpu

UNIT = O.geometry((0.0, 0.0, 0.0), 1.0)
Q = K.BOUNDARY_QUANTUM
FIELDS = ("face_slot", "compatible", "normalized_weight", "normalized_weight_gradient", "centered_face_offset")


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(particles=1 << 20, blocks=1 << 12, lookup_blocks=1 << 16)
    print("device", c.device_name(), flush=True)
    yield c
    c.close()


def same_records(got, want):
    assert got.shape == want.shape
    for f in FIELDS:
        a, b = got[f], want[f]
        if a.dtype.names:
            for c in a.dtype.names:
                assert np.array_equal(a[c].view(np.uint64), b[c].view(np.uint64)), (f, c)
        elif a.dtype.kind == "f":
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), f
        else:
            assert np.array_equal(a, b), f


def check(ctx, geom, blocks, p, tags, q=Q, cap=None, stencils=True):
    cap = len(p) if cap is None else cap
    s_want, want, st_want = O.deposit_compatible(geom, blocks, p, tags, q, q, max(cap, 1), stencils=stencils)
    s_got, got, st_got = ctx.p2g_compatible(geom, blocks, p.view(R.BOUNDARY_PARTICLE), tags.view(R.FACE_TAG), q, q,
                                            max(cap, 1), stencils=stencils)
    assert s_got == s_want, (s_got, s_want)
    if s_want == O.OK:
        assert np.array_equal(got.view(np.int64), want.view(np.int64))
        if stencils:
            same_records(st_got.view(O.COMPATIBLE_SAMPLE), st_want)
    return s_want, want


def slot(blocks, axis, lattice):
    s, v = O.face_slot(blocks, axis, lattice)
    assert s == O.OK
    return v


def test_reference_known_answers(ctx):
    blocks = K.sorted_blocks(UNIT, [(0, 0, 0)])
    p = K.boundary_particle()
    K.set_slopes(p, 0, ((100.0, 2.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)))
    tags = K.boundary_tags(len(blocks))
    for z in range(4):
        for y in range(4):
            tags[slot(blocks, 0, (1, y, z))]["side_tag"] = -7
    s, quanta = check(ctx, UNIT, blocks, p, tags, cap=1)
    assert s == O.OK
    center = slot(blocks, 0, (2, 1, 1))
    assert quanta[center]["mass_quanta_sum"] == 9 << 30 and quanta[center]["pure_momentum_quanta_sum"] == 27 << 30
    # Rounding of actual masked contributions, affine on and off.
    tags = K.boundary_tags(len(blocks), eligible=False)
    for axis in range(3):
        tags[slot(blocks, axis, (1, 1, 1))]["structurally_eligible"] = 1
    p = K.boundary_particle(mass=0.5, velocity=(-3.0, 3.0, 5.0))
    s, quanta = check(ctx, UNIT, blocks, p, tags, q=1.0, cap=1)
    assert s == O.OK and quanta[slot(blocks, 0, (1, 1, 1))]["pure_momentum_quanta_sum"] == -2
    tags[slot(blocks, 0, (2, 1, 1))]["structurally_eligible"] = 1
    p[0]["transfer"]["particle_mass"] = 1.0
    K.set_slopes(p, 0, ((2.0 ** 66, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)))
    assert check(ctx, UNIT, blocks, p, tags, q=1.0, cap=1)[0] == O.OUT_OF_RANGE
    p[0]["affine_enabled"] = 0
    assert check(ctx, UNIT, blocks, p, tags, q=1.0, cap=1)[0] == O.OK


def test_zero_support_missing_owner_and_empty(ctx):
    blocks = K.sorted_blocks(UNIT, [(0, 0, 0)])
    p = K.boundary_particle()
    tags = K.boundary_tags(len(blocks), eligible=False)
    assert check(ctx, UNIT, blocks, p, tags)[0] == O.FAILED_PRECONDITION
    tags[slot(blocks, 0, (3, 1, 1))]["structurally_eligible"] = 1
    assert check(ctx, UNIT, blocks, p, tags)[0] == O.FAILED_PRECONDITION
    tags = K.boundary_tags(len(blocks))
    for pos in ((3.75, 3.75, 3.75), (2.5, 1.5, 1.5)):
        K.set_position(p, 0, pos)
        assert check(ctx, UNIT, blocks, p, tags)[0] == O.FAILED_PRECONDITION
    s, quanta = check(ctx, UNIT, blocks, p[:0], tags, cap=1)
    assert s == O.OK and not quanta["mass_quanta_sum"].any()


@pytest.mark.parametrize("seed,h", [(11, 1.0), (12, 0.25), (13, 0.3)], ids=["unit", "pow2", "non_pow2"])
def test_random_masks_bit_exact(ctx, seed, h):
    geom, blocks, p, tags = K.masked_scene(seed, 3000, h=h)
    s, _ = check(ctx, geom, blocks, p, tags, q=2.0 ** -40)
    assert s == O.OK


def test_larger_scene_quanta_bit_exact(ctx):
    geom, blocks, p, tags = K.masked_scene(21, 60000, h=0.5, blocks_per_axis=6, sides=2, eligible=0.95)
    s, _ = check(ctx, geom, blocks, p, tags, q=2.0 ** -40, stencils=False)
    assert s == O.OK


def test_first_failing_particle_decides_the_status(ctx):
    geom, blocks, p, tags = K.masked_scene(31, 200)
    # A non-finite enabled affine row is InvalidArgument; the same row disabled is ignored.
    t = p[150]["transfer"]
    t["particle_velocity_slopes"]["y_row"]["z"] = np.inf
    p[150]["transfer"] = t
    p[150]["affine_enabled"] = 1
    assert check(ctx, geom, blocks, p, tags, q=2.0 ** -40)[0] == O.INVALID_ARGUMENT
    p[150]["affine_enabled"] = 0
    assert check(ctx, geom, blocks, p, tags, q=2.0 ** -40)[0] == O.OK
    # No compatible support at one particle (its side matches no slot): FailedPrecondition.
    p[40]["side_tag"] = 99
    assert check(ctx, geom, blocks, p, tags, q=2.0 ** -40)[0] == O.FAILED_PRECONDITION
    # An out-of-range contribution at an earlier particle wins over the later missing support.
    t = p[10]["transfer"]
    t["particle_mass"] = 2.0 ** 40
    p[10]["transfer"] = t
    assert check(ctx, geom, blocks, p, tags, q=2.0 ** -40)[0] == O.OUT_OF_RANGE


def test_argument_checks(ctx):
    geom, blocks, p, tags = K.masked_scene(41, 10)
    q = 2.0 ** -40
    s, _, _ = ctx.p2g_compatible(geom, blocks, p.view(R.BOUNDARY_PARTICLE), tags[:-1].view(R.FACE_TAG), q, q, 10)
    assert s == H.INVALID_ARGUMENT
    s, _, _ = ctx.p2g_compatible(geom, blocks, p.view(R.BOUNDARY_PARTICLE), tags.view(R.FACE_TAG), 3.0, q, 10)
    assert s == H.INVALID_ARGUMENT
    s, _, _ = ctx.p2g_compatible(geom, blocks, p.view(R.BOUNDARY_PARTICLE), tags.view(R.FACE_TAG), q, q, 9)
    assert s == H.RESOURCE_EXHAUSTED
