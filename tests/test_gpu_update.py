"""GPU parity of the particle update (fl_update_particles) against the oracle, without an obstacle and
with half-space and slab obstacles.

Bar: the same status as the oracle; on success every field bit-identical except the volume ratio,
which goes through exp and may differ from the host's exp in the last place (|dJ| <= 2 ulp). Cases:
the closed-form cases of the oracle tests, random particles with mixed PIC initialisation,
fallbacks, surface expansion and compression, the half-cell guard on both sides, every error kind
and the first failing particle, a large batch; the reference's wall-contact cases through one particle
(hits, friction, slab penetration, rejections, co-moving endpoints, oblique tangency, fast walls), and
20,000 particles near moving half-spaces and a slab with thousands of hits.
"""
import math

import numpy as np
import pytest

import pyoracle as O
from fluidity_amd import hip as H
from fluidity_amd import records as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(particles=1 << 16, blocks=1 << 10, lookup_blocks=1 << 14)
    print("device", c.device_name(), flush=True)
    yield c
    c.close()


def random_case(rng, n, h=0.25, dt=0.01):
    g = np.zeros(n, O.BOUNDARY_GATHER)
    for k in "xyz":
        g["new_pic_velocity"][k] = rng.normal(0, 2.0, n)
        g["particle_minus_old_pic_velocity"][k] = rng.normal(0, 0.5, n)
        for r in ("x_row", "y_row", "z_row"):
            g["velocity_gradient"][r][k] = rng.normal(0, 3.0, n)
            g["fitted_velocity_slopes"][r][k] = rng.normal(0, 1.0, n)
    g["affine_enabled"] = rng.integers(0, 2, n)
    g["old_pic_fallback"] = rng.random(n) < 0.1
    g["new_pic_fallback"] = rng.random(n) < 0.1
    g["low_mass_fallback"] = rng.random(n) < 0.1
    p = np.zeros(n, O.LIQUID_PARTICLE)
    p["id"] = np.arange(1, n + 1)
    p["tag"] = rng.integers(0, 1 << 30, n)
    for k in "xyz":
        p["world_position"][k] = rng.uniform(-3.0, 3.0, n)
        p["world_velocity"][k] = rng.normal(0, 1, n)
    p["volume_ratio"] = rng.uniform(0.5, 1.5, n)
    p["mass"] = 0.125
    p["pic_initialized"] = rng.random(n) < 0.8
    p["constituent_count"] = 1
    p["constituents"][:, 0]["substance_id"] = 3
    p["constituents"][:, 0]["volume_fraction"] = 1.0
    d = rng.uniform(-2 * h, h, n)
    return O.update_options(h, dt, rng.uniform(0, 1)), g, d, p


def check(ctx, options, g, d, p):
    s_want, want = O.update_particles(options, g, d, p)
    s_got, got = ctx.update_particles(options.view(R.UPDATE_OPTIONS), g.view(R.BOUNDARY_GATHER), d,
                                      p.view(R.LIQUID_PARTICLE))
    assert s_got == s_want, (s_got, s_want)
    got = got.view(O.LIQUID_PARTICLE)
    if s_want != O.OK:
        for name in O.LIQUID_PARTICLE.names:  # untouched on error
            assert np.array_equal(got[name], p[name]), name
        return s_want
    for name in O.LIQUID_PARTICLE.names:
        if name == "volume_ratio":
            a, b = got[name], want[name]
            assert np.all(np.abs(a - b) <= 2 * np.spacing(np.abs(b))), np.max(np.abs(a - b) / np.spacing(np.abs(b)))
        elif got[name].dtype.names:
            for sub in got[name].dtype.names:
                x, y = got[name][sub], want[name][sub]
                if x.dtype.names:
                    for c in x.dtype.names:
                        assert np.array_equal(x[c].view(np.uint64), y[c].view(np.uint64)), (name, sub, c)
                else:
                    assert np.array_equal(x, y), (name, sub)
        else:
            assert np.array_equal(got[name], want[name]), name
    return s_want


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_particles(ctx, seed):
    rng = np.random.default_rng(seed)
    assert check(ctx, *random_case(rng, 5000)) == O.OK


def test_large_batch(ctx):
    rng = np.random.default_rng(9)
    assert check(ctx, *random_case(rng, 60000)) == O.OK
    print("update kernel ms", ctx.last.kernel_ms, "copies ms", ctx.last.copy_ms, flush=True)


def one(h=0.25, dt=0.01, **kw):
    rng = np.random.default_rng(0)
    o, g, d, p = random_case(rng, 4, h, dt)
    g["velocity_gradient"]["x_row"]["x"] = 0.0
    g["velocity_gradient"]["y_row"]["y"] = 0.0
    g["velocity_gradient"]["z_row"]["z"] = 0.0
    for k in "xyz":
        g["new_pic_velocity"][k] = 0.5
    return o, g, d, p


def test_guard_and_error_kinds(ctx):
    h, dt = 0.25, 0.01
    o, g, d, p = one(h, dt)
    v = 0.5 * h / dt
    g["new_pic_velocity"]["x"][2] = v
    g["particle_minus_old_pic_velocity"]["x"][2] = 0.0
    p["world_position"]["x"][2] = 0.0
    g["old_pic_fallback"][2] = 1
    assert check(ctx, o, g, d, p) == O.OK  # exactly h/2
    g["new_pic_velocity"]["x"][2] = np.nextafter(v, np.inf)
    assert check(ctx, o, g, d, p) == O.ABORTED
    o, g, d, p = one(h, dt)
    g["velocity_gradient"]["y_row"]["y"][1] = 1e5
    assert check(ctx, o, g, d, p) == O.ABORTED  # J overflows
    o, g, d, p = one(h, dt)
    g["particle_minus_old_pic_velocity"]["z"][3] = math.nan
    p["pic_initialized"][3] = 1
    g["low_mass_fallback"][3] = 0
    g["old_pic_fallback"][3] = g["new_pic_fallback"][3] = 0
    assert check(ctx, o, g, d, p) == O.INVALID_ARGUMENT
    o, g, d, p = one(h, dt)
    g["new_pic_velocity"]["x"][0] = 1e308
    p["world_position"]["x"][0] = 1.797e308
    assert check(ctx, o, g, d, p) == O.OUT_OF_RANGE
    o, g, d, p = one(h, dt)
    o["friction_coefficient"] = -1.0
    assert check(ctx, o, g, d, p) == O.INVALID_ARGUMENT
    o, g, d, p = one(h, dt)
    o["step_duration"] = 0.0
    assert check(ctx, o, g, d, p) == O.INVALID_ARGUMENT
    # the first failing particle decides: particle 1 Aborted before particle 3 InvalidArgument
    o, g, d, p = one(h, dt)
    g["new_pic_velocity"]["y"][1] = 1e9
    p["world_position"]["z"][3] = math.inf
    assert check(ctx, o, g, d, p) == O.ABORTED
    # an empty batch checks nothing
    assert ctx.update_particles(o.view(R.UPDATE_OPTIONS), g[:0].view(R.BOUNDARY_GATHER), d[:0],
                                p[:0].view(R.LIQUID_PARTICLE))[0] == H.OK


def test_surface_expansion_boundaries(ctx):
    o, g, d, p = one()
    for k in "xyz":
        g["particle_minus_old_pic_velocity"][k] = 0.75
    g["velocity_gradient"]["x_row"]["x"] = 2.0
    g["old_pic_fallback"] = g["new_pic_fallback"] = g["low_mass_fallback"] = 0
    p["pic_initialized"] = 1
    h = float(o["cell_width"][0])
    for phi in (0.0, -h, -h / 3, np.nextafter(0.0, 1.0), np.nextafter(-h, -np.inf)):
        d[:] = phi
        assert check(ctx, o, g, d, p) == O.OK


def test_reference_contact_cases(ctx):
    """The reference's wall-contact cases through one particle, as in the oracle tests."""
    import test_oracle_contact as K
    wall = O.halfspace((-1, 0, 0), -2)
    cases = [((0, 1, 0), (-3, 2, 0), (4, 2, 0), wall, {}), ((2, 1, 0), (-3, 2, 0), (-4, 2, 0), wall, {}),
             ((2, 1, 0), (1, 5, 0), (1, 3, 0), O.halfspace((1, 0, 0), 2, (1, 1, 0)), dict(dt=2, friction=10)),
             ((0, 0, 0), (-2, 3, 0), (0, 4, 0), O.halfspace((1, 0, 0), 0), dict(friction=1)),
             ((1, 2, 3), (-1, 1, 3), (-3, 5, 3), O.halfspace((1, 0, 0), 0, (1, 2, 3)), dict(friction=0.5)),
             ((-0.875, 2, 0), (1, 0, 0), (-1, 3, 0), O.slab((1, 0, 0), 0, 1), dict(penetration=0.125)),
             ((0.75, 0, 0), (0, 0, 0), (0, 0, 0), O.slab((1, 0, 0), 0, 1), dict(penetration=0.125)),
             ((0, 0, 0), (0, 0, 0), (0, 0, 0), O.slab((1, 0, 0), 0, 1), dict(penetration=2)),
             ((0, 0, 0), (0, 0, 0), (0, 0, 0), O.halfspace((1, 0, 0), 4), dict(penetration=2)),
             ((1, 2, 3), (3, 2, 1), (-2, 0, 2), O.halfspace((1, 0, 0), 0), dict(dt=0.5)),
             ((1, 0, 0), (0, 0, 0), (0, 0, 0), O.halfspace((2, 0, 0), 0), {}),
             ((1, 0, 0), (0, 0, 0), (0, 0, 0), O.slab((1, 0, 0), math.inf, 1), {}),
             ((1, 0, 0), (0, 0, 0), (0, 0, 0), O.slab((1, 0, 0), 0, 1, (0, math.inf, 0)), {}),
             ((K.KMAX, 0, 0), (0, 0, 0), (0, 0, 0), O.halfspace((-1, 0, 0), -K.KMAX, (K.KMAX, 0, 0)), dict(dt=2)),
             ((K.KMAX, 0, 0), (0, 0, 0), (0, 0, 0), O.slab((1, 0, 0), 1e20, 1), dict(dt=2)),
             ((0.125, 0, 0), (1e16, 0, 0), (1e16, 0, 0), O.halfspace((1, 0, 0), 0, (1e16, 0, 0)), dict(h=1e17)),
             ((1, 0, 0), (1e16, 0, 0), (1e16, 0, 0), O.halfspace((1, 0, 0), 0.5, (1e16, 0, 0)), dict(h=1e17)),
             ((0, 0, 0), (0.8, -0.6, 3), (0.8, -0.6, 2), O.halfspace((0.6, 0.8, 0), 0), dict(friction=2)),
             ((1, 2, 0), (-3, 5, 0), (-2, 1, 0), O.halfspace((1, 0, 0), 0, (0, 1e20, 0)), {}),
             ((3, 4, 0), (0, 0, -2), (0, 0, 2), O.halfspace((0.6, 0.8, 0), 0, (3, 4, 1)), dict(dt=1.5))]
    for start, stored, advecting, obstacle, kw in cases:
        check(ctx, *K.one(start, stored, advecting, obstacle, **kw))


def walled_case(rng, n, obstacle, normal, penetration=0.0, h=0.25, dt=0.01, friction=0.3):
    """Particles within 0.1 of the plane through the origin on its open side (a few inside it, within
    the penetration tolerance), velocities N(0, 2)."""
    o, g, d, p = random_case(rng, n, h, dt)
    nrm = np.asarray(normal, float)
    tangent = rng.uniform(-3.0, 3.0, (n, 3))
    tangent -= np.outer(tangent @ nrm, nrm)
    depth = rng.uniform(-penetration, 0.1, n)
    pos = tangent + np.outer(depth, nrm)
    for a, k in enumerate("xyz"):
        p["world_position"][k] = pos[:, a]
        g["new_pic_velocity"][k] = rng.normal(0, 2.0, n)
    o = O.update_options(h, dt, float(o["particle_energy_blend"][0]), friction, penetration, obstacle)
    return o, g, d, p


@pytest.mark.parametrize("seed", [1, 2])
def test_random_walls(ctx, seed):
    rng = np.random.default_rng(seed)
    # An oblique moving wall far from the origin: some hit's start distance cannot be resolved against
    # the rounding of the oblique projection, so the reference refuses the batch (OutOfRange); the status
    # and the first failing particle must agree.
    q = (0.6, 0.8, 0.0)
    check(ctx, *walled_case(rng, 20000, O.halfspace(q, 0.0, (0.5, -0.2, 0.1)), q))
    for obstacle, normal, pen in ((O.halfspace((0, 0, 1), 0.0, (0, 0, 0.4)), (0, 0, 1), 0.01),
                                  (O.slab((1, 0, 0), -0.05, 0.05, (0.2, 0, 0)), (1, 0, 0), 0.0)):
        args = walled_case(rng, 20000, obstacle, normal, pen)
        assert check(ctx, *args) == O.OK
        s, base = O.update_particles(O.update_options(0.25, 0.01, float(args[0]["particle_energy_blend"][0]), 0.3),
                                     *args[1:])
        _, walled = O.update_particles(*args)
        hits = int((walled["world_velocity"]["x"] != base["world_velocity"]["x"]).sum())
        print("contact hits", hits, flush=True)
        assert hits > 100


def test_reference_tank_cases(ctx):
    import test_oracle_contact as K
    for args, kw, _ in K.tank_cases():
        check(ctx, *K.one(*args, **kw))


@pytest.mark.parametrize("seed", [1, 2])
def test_random_tank(ctx, seed):
    """20,000 particles within 0.1 of the walls of an open tank with a moving piston, some 0.01 inside
    (within the penetration tolerance): thousands of frictionless hits."""
    rng = np.random.default_rng(seed)
    n = 20000
    o, g, d, p = random_case(rng, n, 0.25, 0.01)
    lo, hi = np.array([0.0, 0.0, 0.0]), np.array([2.0, 1.0, 1.5])
    pos = rng.uniform(lo, hi, (n, 3))
    axis = rng.integers(0, 3, n)
    side = rng.integers(0, 2, n)
    depth = rng.uniform(-0.01, 0.1, n)
    for k in range(n):
        a = axis[k]
        if a == 1:
            pos[k, 1] = lo[1] + depth[k]  # floor only: the top is open
        else:
            pos[k, a] = lo[a] + depth[k] if side[k] == 0 else hi[a] - depth[k]
    for a, c in enumerate("xyz"):
        p["world_position"][c] = pos[:, a]
        g["new_pic_velocity"][c] = rng.normal(0, 2.0, n)
    o = O.update_options(0.25, 0.01, float(o["particle_energy_blend"][0]), 0.0, 0.01,
                         O.tank(lo[0], hi[0], lo[1], lo[2], hi[2], 0.4))
    assert check(ctx, o, g, d, p) == O.OK
    s, base = O.update_particles(O.update_options(0.25, 0.01, float(o["particle_energy_blend"][0])), g, d, p)
    _, walled = O.update_particles(o, g, d, p)
    hits = int(np.sum(walled["world_velocity"]["x"] != base["world_velocity"]["x"]) +
               np.sum(walled["world_velocity"]["y"] != base["world_velocity"]["y"]))
    print("tank hits", hits, flush=True)
    assert hits > 1000
