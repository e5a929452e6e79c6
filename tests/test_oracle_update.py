"""Oracle: the obstacle-free particle update (flo_update_particles), pinned by closed forms.

The reference's ASFLIP relation v = new PIC + 0.99 e (v_start - old PIC) with J unchanged at zero
divergence (native/tests/test_liquid_substep_3d.cc:627-641), the surface-expansion advection rule and
its four conditions (native/src/liquid_substep_3d_cpu.cc:293-300), advection by one fused multiply-add
per component, the half-cell displacement guard, J *= exp(dt div u), the error codes and their
particle order, and untouched fields.
"""
import math

import numpy as np
import pytest

import pyoracle as O

H = 0.25
DT = 0.01


def gather_record(new_pic=(1.0, -2.0, 0.5), delta=(0.25, 0.5, -1.0), grad=None, fallback=(0, 0, 0), affine=1):
    g = np.zeros(1, O.BOUNDARY_GATHER)
    for k, (a, b) in zip("xyz", zip(new_pic, delta)):
        g["new_pic_velocity"][k] = a
        g["particle_minus_old_pic_velocity"][k] = b
    grad = np.zeros((3, 3)) if grad is None else np.asarray(grad, dtype=np.float64)
    for r, row in enumerate(("x_row", "y_row", "z_row")):
        for c, k in enumerate("xyz"):
            g["velocity_gradient"][row][k] = grad[r, c]
            g["fitted_velocity_slopes"][row][k] = 10.0 * r + c + 0.5
    g["affine_enabled"] = affine
    g["old_pic_fallback"], g["new_pic_fallback"], g["low_mass_fallback"] = fallback
    return g


def particle_record(position=(1.0, 2.0, 3.0), pic=1, volume_ratio=1.25):
    p = np.zeros(1, O.LIQUID_PARTICLE)
    p["id"], p["tag"], p["mass"], p["reference_volume"], p["side_tag"], p["phase_id"] = 77, 5, 0.125, 2.0, -3, 9
    p["constituent_count"] = 2
    p["constituents"][0][0] = (4, 0.5)
    p["constituents"][0][1] = (6, 0.5)
    for k, v in zip("xyz", position):
        p["world_position"][k] = v
        p["world_velocity"][k] = 100.0
    p["volume_ratio"] = volume_ratio
    p["pic_initialized"] = pic
    return p


def vec(r):
    return np.array([r["x"], r["y"], r["z"]], dtype=np.float64).reshape(3)


def run(g, p, distance=1.0, blend=0.8, **kw):
    return O.update_particles(O.update_options(H, DT, blend, **kw), g, np.array([distance]), p)


def test_asflip_stored_velocity_and_advection():
    g, p = gather_record(), particle_record()
    s, out = run(g, p, blend=0.8)
    assert s == O.OK
    new_pic, delta = vec(g["new_pic_velocity"][0]), vec(g["particle_minus_old_pic_velocity"][0])
    alpha = 0.99 * 0.8
    assert np.array_equal(vec(out["world_velocity"][0]), new_pic + alpha * delta)
    # zero divergence: no surface blend, advection with the PIC velocity, J unchanged
    x0 = vec(p["world_position"][0])
    want = np.array([math.fma(v, DT, x) for v, x in zip(new_pic, x0)]) if hasattr(math, "fma") else None
    got = vec(out["world_position"][0])
    assert np.allclose(got, x0 + DT * new_pic, rtol=0, atol=1e-15)
    if want is not None:
        assert np.array_equal(got, want)
    assert out["volume_ratio"][0] == p["volume_ratio"][0]
    assert out["pic_initialized"][0] == 1 and out["affine_enabled"][0] == 1
    for r in ("x_row", "y_row", "z_row"):
        assert np.array_equal(vec(out["velocity_slopes"][r][0]), vec(g["fitted_velocity_slopes"][r][0]))
    # fields the update does not own are untouched
    for k in ("id", "tag", "mass", "reference_volume", "side_tag", "phase_id", "constituent_count", "constituents"):
        assert np.array_equal(out[k], p[k]), k


@pytest.mark.parametrize("pic,fallback", [(0, (0, 0, 0)), (1, (1, 0, 0)), (1, (0, 1, 0)), (1, (0, 0, 1))])
def test_uninitialized_or_fallback_particles_take_the_pic_velocity(pic, fallback):
    g, p = gather_record(fallback=fallback, affine=0), particle_record(pic=pic)
    s, out = run(g, p)
    assert s == O.OK
    assert np.array_equal(vec(out["world_velocity"][0]), vec(g["new_pic_velocity"][0]))
    assert out["affine_enabled"][0] == 0 and out["pic_initialized"][0] == 1


def test_surface_expansion_blends_the_advecting_velocity():
    grad = np.diag([0.5, 0.25, 0.25])  # divergence 1
    new_pic, delta = np.array([1.0, -2.0, 0.5]), np.array([0.25, 0.5, -1.0])
    alpha, blend = 0.99 * 0.6, 0.6
    x0 = np.array([1.0, 2.0, 3.0])
    adv_surface = new_pic + (blend * alpha) * delta
    for distance, expanded in [(0.0, True), (-H, True), (-H * 0.5, True), (1e-300, False), (-H - 1e-12, False)]:
        s, out = run(gather_record(grad=grad), particle_record(), distance=distance, blend=blend)
        assert s == O.OK
        adv = adv_surface if expanded else new_pic
        assert np.allclose(vec(out["world_position"][0]), x0 + DT * adv, rtol=0, atol=1e-15), distance
        assert out["volume_ratio"][0] == 1.25 * math.exp(DT * 1.0)
    # compression never blends the advecting velocity
    s, out = run(gather_record(grad=-grad), particle_record(), distance=-0.1, blend=blend)
    assert s == O.OK and np.allclose(vec(out["world_position"][0]), x0 + DT * new_pic, rtol=0, atol=1e-15)
    assert out["volume_ratio"][0] == 1.25 * math.exp(-DT)


def test_half_cell_displacement_guard():
    # a displacement of exactly h/2 is accepted; the next double above is Aborted
    v = 0.5 * H / DT
    g = gather_record(new_pic=(v, 0.0, 0.0), delta=(0.0, 0.0, 0.0))
    p = particle_record(position=(0.0, 0.0, 0.0))
    s, out = run(g, p)
    assert s == O.OK and out["world_position"]["x"][0] == 0.5 * H
    g = gather_record(new_pic=(np.nextafter(v, np.inf), 0.0, 0.0), delta=(0.0, 0.0, 0.0))
    assert run(g, p)[0] == O.ABORTED


def test_volume_ratio_overflow_and_errors():
    g = gather_record(new_pic=(0.0, 0.0, 0.0), grad=np.diag([1e5, 0.0, 0.0]))
    assert run(g, particle_record())[0] == O.ABORTED  # exp overflow
    g = gather_record(new_pic=(0.0, 0.0, 0.0), grad=np.diag([-1e5, 0.0, 0.0]))
    assert run(g, particle_record())[0] == O.ABORTED  # J underflows to 0
    assert run(gather_record(new_pic=(math.nan, 0, 0)), particle_record())[0] == O.INVALID_ARGUMENT
    assert run(gather_record(delta=(0, math.inf, 0)), particle_record())[0] == O.INVALID_ARGUMENT
    assert run(gather_record(), particle_record(position=(math.inf, 0, 0)))[0] == O.INVALID_ARGUMENT
    assert run(gather_record(new_pic=(1e308, 0, 0), delta=(0, 0, 0)),
               particle_record(position=(1.797e308, 0, 0)))[0] == O.OUT_OF_RANGE  # the advanced position overflows
    assert run(gather_record(), particle_record(), friction=-1.0)[0] == O.INVALID_ARGUMENT
    assert run(gather_record(), particle_record(), penetration=math.nan)[0] == O.INVALID_ARGUMENT
    s, _ = O.update_particles(O.update_options(H, 0.0, 0.5), gather_record(), np.array([1.0]), particle_record())
    assert s == O.INVALID_ARGUMENT
    # no particles: options are not checked (the reference validates them per particle)
    s, _ = O.update_particles(O.update_options(H, -1.0, 0.5), gather_record()[:0], np.zeros(0), particle_record()[:0])
    assert s == O.OK


def stack(records, dtype):
    out = np.zeros(len(records), dtype)  # np.concatenate would drop the records' padding
    for i, r in enumerate(records):
        out[i] = r[0]
    return out


def test_first_failing_particle_decides():
    g = stack([gather_record(), gather_record(new_pic=(1e9, 0, 0)), gather_record(new_pic=(math.nan, 0, 0))],
              O.BOUNDARY_GATHER)
    p = stack([particle_record(), particle_record(), particle_record()], O.LIQUID_PARTICLE)
    s, out = O.update_particles(O.update_options(H, DT, 0.5), g, np.ones(3), p)
    assert s == O.ABORTED
    assert out["pic_initialized"][0] == 1
    for name in O.LIQUID_PARTICLE.names:  # later particles untouched (field values; padding is not a value)
        assert np.array_equal(out[name][1:], p[name][1:]), name
