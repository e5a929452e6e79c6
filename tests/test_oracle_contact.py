"""Pins the oracle's particle update with a half-space or slab obstacle (the certified swept-plane
contact of ResolveContact).

Known answers are the reference's wall-contact cases
(native/tests/liquid_wall_contact_3d_cases.h:228-486), driven through one particle: the gather's new
PIC velocity is the advecting velocity A, its particle-minus-old-PIC velocity is S - A, and the energy
blend 1/0.99 makes alpha = 0.99 e exactly 1, so the stored velocity is exactly S; a zero velocity
gradient keeps the surface-expansion blend off and J unchanged, and a wide cell keeps the half-cell
guard out of the way. Checked: the end position and the stored velocity (within 2e-14, the reference's
tolerance) or the status. Closed forms add a frictionless head-on hit on an oblique moving plane, the
reflection-free stop of a normal impact, and the free flight that never reaches the plane.
"""
import math

import numpy as np
import pytest

import pyoracle as O

KMAX = np.finfo(np.float64).max
BLEND = 1 / 0.99  # 0.99 * BLEND == 1.0 exactly


def one(start, stored, advecting, obstacle, dt=1.0, friction=0.0, penetration=0.0, h=1e3):
    o = O.update_options(h, dt, BLEND, friction, penetration, obstacle)
    g = np.zeros(1, O.BOUNDARY_GATHER)
    p = np.zeros(1, O.LIQUID_PARTICLE)
    for k, s, a, x in zip("xyz", stored, advecting, start):
        g["new_pic_velocity"][k] = a
        g["particle_minus_old_pic_velocity"][k] = s - a
        p["world_position"][k] = x
    p["volume_ratio"] = 1.0
    p["pic_initialized"] = 1
    p["mass"] = 1.0
    p["constituent_count"] = 1
    p["constituents"][:, 0]["volume_fraction"] = 1.0
    return o, g, np.zeros(1), p


def run(*args, **kw):
    o, g, d, p = one(*args, **kw)
    s, out = O.update_particles(o, g, d, p)
    if s != O.OK:
        return s, None, None
    x, v = out[0]["world_position"], out[0]["world_velocity"]
    return s, np.array([x["x"], x["y"], x["z"]]), np.array([v["x"], v["y"], v["z"]])


def near(a, b, tol=2e-14):
    return np.all(np.abs(np.asarray(a, float) - np.asarray(b, float)) <= tol)


def test_one_sided_plane_blocks_only_motion_into_solid():
    wall = O.halfspace((-1, 0, 0), -2)
    s, x, v = run((0, 1, 0), (-3, 2, 0), (4, 2, 0), wall)
    assert s == O.OK and near(x, (2, 3, 0)) and near(v, (-3, 2, 0))
    s, x, v = run((2, 1, 0), (-3, 2, 0), (-4, 2, 0), wall)
    assert s == O.OK and near(x, (-2, 3, 0))


def test_tangential_motion_and_touching_wall():
    s, x, v = run((2, 1, 0), (1, 5, 0), (1, 3, 0), O.halfspace((1, 0, 0), 2, (1, 1, 0)), dt=2, friction=10)
    assert s == O.OK and near(x, (4, 7, 0)) and near(v, (1, 5, 0))
    s, x, v = run((0, 0, 0), (-2, 3, 0), (0, 4, 0), O.halfspace((1, 0, 0), 0), friction=1)
    assert s == O.OK and near(v, (0, 1, 0)) and near(x, (0, 4, 0))


def test_friction_uses_each_incoming_speed():
    s, x, v = run((1, 2, 3), (-1, 1, 3), (-3, 5, 3), O.halfspace((1, 0, 0), 0, (1, 2, 3)), friction=0.5)
    assert s == O.OK and near(v, (1, 2, 3)) and near(x, (1, 5.5, 6))


def test_slab_penetration():
    slab = O.slab((1, 0, 0), 0, 1)
    s, x, v = run((-0.875, 2, 0), (1, 0, 0), (-1, 3, 0), slab, penetration=0.125)
    assert s == O.OK and near(x, (-2, 5, 0)) and near(v, (0, 0, 0))
    assert run((0.75, 0, 0), (0, 0, 0), (0, 0, 0), slab, penetration=0.125)[0] == O.FAILED_PRECONDITION
    assert run((0, 0, 0), (0, 0, 0), (0, 0, 0), slab, penetration=2)[0] == O.FAILED_PRECONDITION
    assert run((0, 0, 0), (0, 0, 0), (0, 0, 0), O.halfspace((1, 0, 0), 4), penetration=2)[0] == O.FAILED_PRECONDITION


def test_endpoint_contact_and_free_flight():
    s, x, _ = run((1, 2, 3), (3, 2, 1), (-2, 0, 2), O.halfspace((1, 0, 0), 0), dt=0.5)
    assert s == O.OK and near(x, (0, 2, 4))
    s, x, v = run((1, 2, 3), (3, 2, 1), (-2, 0, 2), None, dt=0.5)
    assert s == O.OK and near(x, (0, 2, 4)) and near(v, (3, 2, 1))


def test_invalid_inputs():
    wall = O.halfspace((1, 0, 0), 0)
    assert run((math.inf, 0, 0), (0, 0, 0), (0, 0, 0), wall)[0] == O.INVALID_ARGUMENT
    assert run((1, 0, 0), (0, math.nan, 0), (0, 0, 0), wall)[0] == O.INVALID_ARGUMENT
    assert run((1, 0, 0), (0, 0, 0), (0, 0, 0), wall, dt=0.0)[0] == O.INVALID_ARGUMENT
    assert run((1, 0, 0), (0, 0, 0), (0, 0, 0), wall, friction=-1)[0] == O.INVALID_ARGUMENT
    assert run((1, 0, 0), (0, 0, 0), (0, 0, 0), wall, penetration=math.inf)[0] == O.INVALID_ARGUMENT
    for ob in (O.halfspace((2, 0, 0), 0), O.slab((1, 0, 0), math.inf, 1), O.slab((1, 0, 0), 0, 0),
               O.slab((1, 0, 0), 0, 1, (0, math.inf, 0))):
        assert run((1, 0, 0), (0, 0, 0), (0, 0, 0), ob)[0] == O.INVALID_ARGUMENT


def test_unrepresentable_trajectory_and_wall_pose():
    assert run((KMAX, 0, 0), (KMAX, 0, 0), (KMAX, 0, 0), None, dt=2)[0] == O.OUT_OF_RANGE
    assert run((KMAX, 0, 0), (0, 0, 0), (0, 0, 0), O.halfspace((-1, 0, 0), -KMAX, (KMAX, 0, 0)), dt=2)[0] == \
        O.OUT_OF_RANGE
    assert run((KMAX, 0, 0), (0, 0, 0), (0, 0, 0), O.slab((1, 0, 0), 1e20, 1), dt=2)[0] == O.OUT_OF_RANGE
    assert run((KMAX, 0, 0), (KMAX, 0, 0), (KMAX, 0, 0), O.halfspace((1, 0, 0), 0, (-KMAX, 0, 0)), dt=0.5)[0] == \
        O.OUT_OF_RANGE


def test_co_moving_endpoints_and_oblique_tangency():
    s, x, _ = run((0.125, 0, 0), (1e16, 0, 0), (1e16, 0, 0), O.halfspace((1, 0, 0), 0, (1e16, 0, 0)), h=1e17)
    assert s == O.OK and x[0] == 1e16
    assert run((1, 0, 0), (1e16, 0, 0), (1e16, 0, 0), O.halfspace((1, 0, 0), 0.5, (1e16, 0, 0)), h=1e17)[0] == \
        O.OUT_OF_RANGE
    s, x, _ = run((0, 0, 0), (0.8, -0.6, 3), (0.8, -0.6, 2), O.halfspace((0.6, 0.8, 0), 0), friction=2)
    assert s == O.OK and near(x, (0.8, -0.6, 2))


def test_fast_wall_and_moving_oblique_plane():
    s, x, v = run((1, 2, 0), (-3, 5, 0), (-2, 1, 0), O.halfspace((1, 0, 0), 0, (0, 1e20, 0)))
    assert s == O.OK and near(v, (0, 5, 0)) and near(x, (0, 3, 0))
    s, x, v = run((3, 4, 0), (0, 0, -2), (0, 0, 2), O.halfspace((0.6, 0.8, 0), 0, (3, 4, 1)), dt=1.5)
    assert s == O.OK and near(v, (3, 4, -2)) and near(x, (4.5, 6, 3), 1e-13)


@pytest.mark.parametrize("speed", [0.5, 2.0, 7.0])
def test_head_on_stop_closed_form(speed):
    """A particle moving at -speed along x toward the half-space x <= 0 from x = 1 stops on the plane
    when it would cross it within the step (position (0, y + dt*vy), stored normal velocity 0), and
    flies freely to 1 - speed*dt otherwise."""
    dt = 0.25
    s, x, v = run((1, 0.5, 0), (-speed, 1, 0), (-speed, 1, 0), O.halfspace((1, 0, 0), 0), dt=dt)
    assert s == O.OK
    if speed * dt >= 1:
        assert near(x, (0, 0.5 + dt, 0)) and near(v, (0, 1, 0))
    else:
        assert near(x, (1 - speed * dt, 0.5 + dt, 0)) and near(v, (-speed, 1, 0))


def test_solid_distance_gates_surface_expansion():
    """Near a wall (solid distance <= h) the advecting velocity drops the ASFLIP term even under surface
    expansion: the particle moves with the new PIC velocity."""
    h, dt = 1.0, 0.1
    o = O.update_options(h, dt, 0.5, 0.0, 0.0, O.halfspace((0, 1, 0), 0))
    g = np.zeros(2, O.BOUNDARY_GATHER)
    p = np.zeros(2, O.LIQUID_PARTICLE)
    g["new_pic_velocity"]["x"] = 1.0
    g["particle_minus_old_pic_velocity"]["x"] = 1.0
    g["velocity_gradient"]["x_row"]["x"] = 0.5  # divergence > 0
    p["world_position"]["y"] = (0.5, 5.0)  # within h of the wall, then far from it
    p["volume_ratio"], p["pic_initialized"] = 1.0, 1
    s, out = O.update_particles(o, g, np.full(2, -0.5 * h), p)
    assert s == O.OK
    assert out[0]["world_position"]["x"] == np.float64(dt) * 1.0
    assert abs(out[1]["world_position"]["x"] - (1.0 + 0.5 * 0.495) * dt) <= 1e-15


TANK = (0.25, 2, 0.125, 0.125, 0.75, 0.25)  # low x (piston, +0.25 m/s), high x, floor, low z, high z


def tank_cases():
    """The reference's open-tank cases (liquid_wall_contact_3d_cases.h:38-145): (args, kwargs, expected
    position and stored velocity, or a status)."""
    t = O.tank(*TANK)
    return [(((0.5, 0.375, 0.375), (-1, -2, 0.125), (-0.25, -0.5, 0.25), t), {}, ((0.5, 0.125, 0.625), (0.25, 0, 0.125))),
            (((1, 0.375, 0.5), (2, 0.75, 1), (1, 0.5, 0.125), t), {}, ((2, 0.875, 0.625), (0, 0.75, 1))),
            (((1, 0.375, 0.5), (2, 0.75, 1), (1, 0.5, 1), t), {}, ((2, 0.875, 0.75), (0, 0.75, 0))),
            (((0.5, 0.375, 0.375), (-1, -2, 1), (0, 2, 0), t), {}, ((0.5, 2.375, 0.375), (0.25, -2, 1))),
            (((0.125, 0.5, 0.0625), (-1, 2, -0.5), (2, 0.25, -0.25), t), dict(penetration=0.125),
             ((2, 0.75, 0.125), (0, 2, 0))),
            (((0.125, 0.5, 0.0625), (-1, 2, -0.5), (2, 0.25, -0.25), t), dict(penetration=0.0625), O.FAILED_PRECONDITION),
            (((0.5, 0.375, 0.375), (0, 0, 0), (0, 0, 0), t), dict(friction=0.1), O.INVALID_ARGUMENT),
            (((0.5, 0.375, 0.375), (0, 0, 0), (0, 0, 0), O.tank(0.25, 2, 0.125, 0.125, 0.125, 0.25)), {},
             O.INVALID_ARGUMENT),
            (((0.5, 0.375, 0.375), (0, 0, 0), (0, 0, 0), O.tank(0.25, 2, 0.125, 0.125, 0.75, 1.75)), {},
             O.FAILED_PRECONDITION),
            (((0.5, 0.375, 0.375), (0, 0, 0), (0, 0, 0), O.tank(0.25, 2, 0.125, 0.125, 0.75, 2)), {},
             O.FAILED_PRECONDITION),
            (((0.5, 0.375, 0.375), (0, 0, 0), (0, 0, 0), O.tank(0.25, 2, 0.125, 0.125, 0.75, -KMAX)), dict(dt=2),
             O.OUT_OF_RANGE),
            (((0.5, 0.375, 0.375), (0, 0, 0), (0, 0, 0), O.tank(-KMAX, KMAX, 0, -1, 1, 0)), {}, O.OUT_OF_RANGE)]


def test_open_tank_cases():
    for args, kw, want in tank_cases():
        s, x, v = run(*args, **kw)
        if isinstance(want, tuple):
            assert s == O.OK and near(x, want[0]) and near(v, want[1]), (args, s, x, v)
        else:
            assert s == want, (args, s)


def test_tank_solid_distance_gates_surface_expansion():
    """The tank's least inward clearance is the solid distance: a particle 0.5 h from the floor drops the
    ASFLIP term of its advecting velocity, one 5 h away keeps it."""
    h, dt = 0.1, 0.01
    o = O.update_options(h, dt, 0.5, 0.0, 0.0, O.tank(0.0, 10.0, 0.0, 0.0, 10.0))
    g = np.zeros(2, O.BOUNDARY_GATHER)
    p = np.zeros(2, O.LIQUID_PARTICLE)
    g["new_pic_velocity"]["x"] = 1.0
    g["particle_minus_old_pic_velocity"]["x"] = 1.0
    g["velocity_gradient"]["x_row"]["x"] = 0.5
    p["world_position"]["x"], p["world_position"]["z"] = 5.0, 5.0
    p["world_position"]["y"] = (0.05, 0.5)
    p["volume_ratio"], p["pic_initialized"] = 1.0, 1
    s, out = O.update_particles(o, g, np.full(2, -0.5 * h), p)
    assert s == O.OK
    assert out[0]["world_position"]["x"] == 5.0 + dt
    assert abs(out[1]["world_position"]["x"] - (5.0 + (1.0 + 0.5 * 0.495) * dt)) <= 1e-15
