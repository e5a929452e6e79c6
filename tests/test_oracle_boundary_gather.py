"""Pins the oracle's boundary gather (ghost continuation, rank-revealing QR, wall projection, fallbacks)
to the reference's known answers: native/tests/liquid_boundary_transfer_3d_cases.h:217-460
(each test names the case it reproduces)."""
import math

import numpy as np
import pytest

import cases as K
import pyoracle as O


def vec(a):
    return K.vec(a)


def test_same_side_affine_ghosts_exclude_poisoned_donors_and_separate_histories():
    f = K.BoundaryFixture()
    for axis in range(3):
        for z in range(4):
            for y in range(4):
                for x in range(4):
                    a = f.anchor(axis, x, y, z)
                    s = f.slot(axis, x, y, z)
                    f.old_values[s] = 1.0 + a[0] + 2.0 * a[1] - a[2]
                    f.new_values[s] = 4.0 + 2.0 * a[0] - a[1] + 3.0 * a[2]
        poison = f.slot(axis, 1, 1, 1)
        f.tags[poison]["side_tag"] = -7
        f.old_values[poison] = -1e8
        f.new_values[poison] = 1e9
    s, out, counts = f.gather()
    assert s == O.OK
    r = out[0]
    np.testing.assert_allclose(vec(r["new_pic_velocity"]), (10.0, 10.0, 10.0), atol=2e-12)
    np.testing.assert_allclose(vec(r["grid_velocity_increment"]), (6.0, 6.0, 6.0), atol=2e-12)
    np.testing.assert_allclose(vec(r["particle_minus_old_pic_velocity"]), (-1.0, 0.0, 1.0), atol=2e-12)
    np.testing.assert_allclose(vec(r["fitted_velocity_slopes"]["x_row"]), (2.0, -1.0, 3.0), atol=2e-12)
    np.testing.assert_allclose(vec(r["fitted_velocity_slopes"]["y_row"]), (2.0, -1.0, 3.0), atol=2e-12)
    np.testing.assert_allclose(vec(r["velocity_gradient"]["z_row"]), (2.0, -1.0, 3.0), atol=2e-12)
    assert r["affine_enabled"] == 1 and counts[0] == 0 and counts[1] == 0


def test_identical_histories_keep_particle_retention_distinct_from_grid_increment():
    f = K.BoundaryFixture()
    f.new_values = f.old_values.copy()
    s, out, _ = f.gather()
    assert s == O.OK
    np.testing.assert_allclose(vec(out[0]["grid_velocity_increment"]), 0.0, atol=2e-12)
    np.testing.assert_allclose(vec(out[0]["particle_minus_old_pic_velocity"]), (1.0, 2.0, 3.0), atol=2e-12)


def test_wall_histories_use_separate_times_and_constrain_oblique_normal():
    f = K.BoundaryFixture()
    f.validity[:] = 0
    for slot in range(len(f.validity)):
        K.set_wall(f.old_walls, slot, 7, (1.0, 0.0, 0.0), (2.0, 4.0, 6.0))
        K.set_wall(f.new_walls, slot, 7, (1.0, 0.0, 0.0), (5.0, 8.0, 9.0))
    s, out, counts = f.gather()
    assert s == O.OK
    np.testing.assert_allclose(vec(out[0]["new_pic_velocity"]), (5.0, 8.0, 9.0), atol=2e-12)
    np.testing.assert_allclose(vec(out[0]["grid_velocity_increment"]), (3.0, 4.0, 3.0), atol=2e-12)
    np.testing.assert_allclose(vec(out[0]["particle_minus_old_pic_velocity"]), 0.0, atol=2e-12)
    assert out[0]["affine_enabled"] == 0 and counts[0] == 1 and counts[1] == 1
    assert f.gather(old_wall_time=1.0)[0] == O.INVALID_ARGUMENT
    # Oblique wall at the missing x-face (1, 1.5, 1.5): every component is evaluated there, then the
    # normal part is replaced.
    f.validity[:] = 1
    fields = {0: ((1.0, 1.0, 2.0, -1.0), (4.0, 2.0, -1.0, 3.0)), 1: ((-2.0, 3.0, -1.0, 2.0), (7.0, -1.0, 4.0, -1.0)),
              2: ((4.0, -1.0, 1.0, 3.0), (-3.0, 3.0, 2.0, -2.0))}
    for axis in range(3):
        (o0, ox, oy, oz), (n0, nx, ny, nz) = fields[axis]
        for z in range(4):
            for y in range(4):
                for x in range(4):
                    a = f.anchor(axis, x, y, z)
                    s = f.slot(axis, x, y, z)
                    f.old_values[s] = o0 + ox * a[0] + oy * a[1] + oz * a[2]
                    f.new_values[s] = n0 + nx * a[0] + ny * a[1] + nz * a[2]
    nc = math.sqrt(0.5)
    target = f.slot(0, 1, 1, 1)
    K.set_wall(f.old_walls, target, 7, (nc, nc, 0.0), (2.0, -1.0, 4.0))
    K.set_wall(f.new_walls, target, 7, (nc, nc, 0.0), (5.0, 1.0, -2.0))
    f.validity[target] = 0
    old_proj = 3.5 - (3.5 + 2.5 - 2.0 + 1.0) / 2.0
    new_proj = 9.0 - (9.0 + 10.5 - 5.0 - 1.0) / 2.0
    w = 0.5 * 0.75 * 0.75
    old_pic_x = 4.0 + w * (old_proj - 3.5)
    new_pic_x = 10.0 + w * (new_proj - 9.0)
    s, out, _ = f.gather()
    assert s == O.OK
    np.testing.assert_allclose(vec(out[0]["new_pic_velocity"]), (new_pic_x, 10.0, 1.5), atol=2e-12)
    np.testing.assert_allclose(vec(out[0]["grid_velocity_increment"]), (new_pic_x - old_pic_x, 6.0, -7.0), atol=2e-12)
    np.testing.assert_allclose(vec(out[0]["particle_minus_old_pic_velocity"]), (3.0 - old_pic_x, 0.0, -3.5),
                               atol=2e-12)
    assert out[0]["affine_enabled"] == 1


def test_singular_donors_use_nearest_continuation_and_suppress_retention():
    f = K.BoundaryFixture()
    f.validity[:] = 0
    for axis in range(3):
        first = f.slot(axis, 1, 1, 1)
        second = f.slot(axis, 2 if axis == 0 else 1, 2 if axis == 1 else 1, 2 if axis == 2 else 1)
        f.validity[first] = 1
        f.validity[second] = 2
        f.old_values[first] = f.old_values[second] = 2.0
        f.new_values[first] = f.new_values[second] = 4.0
    s, out, counts = f.gather()
    assert s == O.OK
    np.testing.assert_allclose(vec(out[0]["new_pic_velocity"]), (4.0, 4.0, 4.0), atol=2e-12)
    np.testing.assert_allclose(vec(out[0]["grid_velocity_increment"]), (2.0, 2.0, 2.0), atol=2e-12)
    np.testing.assert_allclose(vec(out[0]["particle_minus_old_pic_velocity"]), 0.0, atol=2e-12)
    assert out[0]["affine_enabled"] == 0 and counts[0] == 1 and counts[1] == 1
    f.validity[:] = 0
    assert f.gather()[0] == O.FAILED_PRECONDITION


def test_low_mass_disables_affine_and_retention_even_with_full_rank_ghost():
    f = K.BoundaryFixture()
    old_validity = f.validity.copy()
    new_validity = f.validity.copy()
    new_validity[f.slot(0, 1, 1, 1)] = 3
    s, out, counts = f.gather(old_validity=old_validity, new_validity=new_validity)
    assert s == O.OK
    np.testing.assert_allclose(vec(out[0]["new_pic_velocity"]), (3.0, 3.0, 3.0), atol=2e-12)
    np.testing.assert_allclose(vec(out[0]["grid_velocity_increment"]), (1.0, 1.0, 1.0), atol=2e-12)
    np.testing.assert_allclose(vec(out[0]["particle_minus_old_pic_velocity"]), 0.0, atol=2e-12)
    assert out[0]["affine_enabled"] == 0 and counts[2] == 1 and counts[0] == 0 and counts[1] == 0


def test_nonaffine_complete_field_reports_fit_separately_from_true_gradient():
    f = K.BoundaryFixture()
    K.set_position(f.particle, 0, (1.75, 1.75, 1.75))
    for axis in range(3):
        for z in range(4):
            for y in range(4):
                for x in range(4):
                    a = f.anchor(axis, x, y, z)
                    f.new_values[f.slot(axis, x, y, z)] = (a[axis] - 1.75) ** 2
    s, out, _ = f.gather()
    assert s == O.OK
    np.testing.assert_allclose(vec(out[0]["new_pic_velocity"]), (0.25, 0.25, 0.25), atol=2e-12)
    np.testing.assert_allclose(vec(out[0]["fitted_velocity_slopes"]["x_row"]), (-0.1875, 0.0, 0.0), atol=2e-12)
    np.testing.assert_allclose(vec(out[0]["velocity_gradient"]["x_row"]), 0.0, atol=2e-12)


@pytest.mark.parametrize("height", [0.0, 1e-12, 1e-8])
def test_rank_revealing_qr_rejects_near_dependent_normalized_direction(height):
    off = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (1.0, 1.0, height)]
    vel = [7.0, 9.0, 10.0, 12.0 + 4.0 * height]
    s, full, at, slopes = O.fit_ghost_affine(off, vel)
    assert s == O.OK and full == (height == 1e-8)
    if full:
        assert abs(at - 7.0) <= 1e-12
        np.testing.assert_allclose(slopes, (2.0, 3.0, 4.0), atol=1e-6)


def test_nearest_fallback_breaks_distance_ties_by_owned_slot_only_in_new_stage():
    f = K.BoundaryFixture()
    K.set_position(f.particle, 0, (2.0, 1.5, 1.5))
    new_validity = f.validity.copy()
    for z in range(4):
        for y in range(4):
            for x in range(4):
                new_validity[f.slot(0, x, y, z)] = 0
    low, high = f.slot(0, 1, 1, 1), f.slot(0, 3, 1, 1)
    new_validity[low] = 1
    new_validity[high] = 2
    f.new_values[low] = 2.0
    f.new_values[high] = 8.0
    s, out, counts = f.gather(new_validity=new_validity)
    assert s == O.OK
    assert abs(out[0]["new_pic_velocity"]["x"] - 2.75) <= 1e-12
    assert abs(out[0]["grid_velocity_increment"]["x"] - 0.75) <= 1e-12
    assert counts[0] == 0 and counts[1] == 1
    np.testing.assert_allclose(vec(out[0]["particle_minus_old_pic_velocity"]), 0.0, atol=2e-12)
    assert out[0]["affine_enabled"] == 0
    assert abs(out[0]["velocity_gradient"]["x_row"]["x"] - 3.0) <= 1e-12


def test_rejects_malformed_history_and_arithmetic():
    f = K.BoundaryFixture()
    K.set_wall(f.old_walls, 0, 7, (2.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    assert f.gather()[0] == O.INVALID_ARGUMENT
    f.old_walls[0] = np.zeros(1, O.WALL_SAMPLE)[0]
    f.old_values[0] = np.nan
    assert f.gather()[0] == O.INVALID_ARGUMENT
    # Eligibility never hides a non-finite payload marked valid or predicted.
    f.tags[0]["side_tag"] = -7
    f.validity[0] = 2
    assert f.gather()[0] == O.INVALID_ARGUMENT
    f.validity[0] = 0
    assert f.gather()[0] == O.OK
    f.tags[0]["side_tag"] = 7
    f.validity[0] = 1
    t = f.particle[0]["transfer"]
    t["particle_velocity"]["x"] = np.finfo(np.float64).max
    f.particle[0]["transfer"] = t
    f.old_values[:] = -np.finfo(np.float64).max / 2.0
    f.new_values[:] = 0.0
    assert f.gather()[0] == O.OUT_OF_RANGE
