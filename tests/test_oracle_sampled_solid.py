"""Pins the oracle's sampled-solid geometry (the sampled-solid mode of flo_freeze_geometry and its two
measures, flo_measure_sampled_cell and flo_sampled_face_weights).

Known answers are the reference's (native/tests/test_liquid_boundary_geometry_3d.cc:22-77, 273-296,
351-362, and the case SampledSolidFreezeSharesOnePressureFace of
native/tests/liquid_boundary_geometry_3d_cases.h:419-518): an oblique solid sharing one aperture across
two cells (wet 1/8 and 23/32, apertures 3/4 and 9/32), a solid corner keeping its piecewise-linear
measure, the blocked-face weights of an affine wall (11/48, 1/48, 7/48, 5/48) and their closed and
open limits, an analytic platen clipping the wet volume and the weights, the closed zero plateau, the
two-cell freeze with its shared face (open 3/4, blocked 1/4, wall velocity 0.2; affine corner
velocities giving the flux 5/8 and mean velocity 5/2), and the missing-sample and split-lobe
refusals. Mode 2 (no reference counterpart): a cell the particle field alone splits, with no corner in
the solid, is measured as the plane-obstacle freeze measures it instead of refused. Independent closed form: a solid sampled from an exact plane (affine, so its piecewise-
linear interpolation is exact) gives the analytic half-space freeze's volumes and areas to rounding.
"""
import math

import numpy as np
import pytest

import pyoracle as O

from test_oracle_freeze import grid, problem, sample


def cell_measure(width, liquid, solid, analytic=None):
    s, wet, open_area = O.measure_sampled_cell(width, liquid, solid, analytic)
    assert s == O.OK, s
    return wet, open_area


def test_oblique_solid_shares_aperture_across_cells():
    first_l, second_l, first_s, second_s = [], [], [], []
    for k in range(8):
        x, y = k & 1, (k >> 1) & 1
        first_l.append(y - 0.75)
        second_l.append(y - 0.75)
        first_s.append((x + y - 1.25) / math.sqrt(2))
        second_s.append((x + y - 0.25) / math.sqrt(2))
    lw, lo = cell_measure(1.0, first_l, first_s)
    rw, ro = cell_measure(1.0, second_l, second_s)
    assert abs(lw - 0.125) <= 2e-14 and abs(rw - 0.71875) <= 2e-14
    assert abs(lo[1] - 0.75) <= 2e-14 and lo[1] == ro[0]
    assert abs(lo[4] - 0.28125) <= 2e-14 and lo[0] == 0.0


def test_solid_corner_keeps_piecewise_linear_measure():
    solid = [-1.0] * 7 + [1.0]
    wet, open_area = cell_measure(2.0, [-1.0] * 8, solid)
    assert abs(wet - 1.0) <= 2e-14
    for a in range(3):
        assert open_area[2 * a] == 0.0 and abs(open_area[2 * a + 1] - 1.0) <= 2e-14


def test_face_weights_integrate_affine_wall_flux():
    s, w = O.sampled_face_weights(1.0, [-0.5, 0.5, -0.5, 0.5])
    assert s == O.OK
    expected = [11 / 48, 1 / 48, 7 / 48, 5 / 48]
    assert np.all(np.abs(w - expected) <= 2e-14)
    assert abs(w.sum() - 0.5) <= 2e-14 and abs(w @ [2.0, 5.0, -2.0, 1.0] - 0.375) <= 2e-14
    s, closed = O.sampled_face_weights(2.0, [0.0] * 4)
    assert s == O.OK and abs(closed.sum() - 4.0) <= 2e-14
    s, opened = O.sampled_face_weights(1.0, [1.0] * 4)
    assert s == O.OK and opened.sum() == 0.0


def test_analytic_platen_clips_wet_volume_and_weights():
    solid = [(k & 1) - 0.5 for k in range(8)]
    wet, open_area = cell_measure(1.0, [-1.0] * 8, solid, (0.0, 1.0, 0.0, 0.75))
    assert abs(wet - 0.375) <= 2e-14 and abs(open_area[4] - 0.375) <= 2e-14 and open_area[3] == 0.0
    s, w = O.sampled_face_weights(1.0, [-0.5, 0.5, -0.5, 0.5], (0.0, 1.0, 0.0, 0.75))
    assert s == O.OK and abs(w.sum() - 0.375) <= 2e-14 and abs(w @ [2.0, 6.0, 2.0, 6.0] - 1.125) <= 2e-14


def test_zero_plateau_is_closed_and_nan_is_invalid():
    wet, open_area = cell_measure(1.0, [-1.0] * 8, [0.0] * 8)
    assert wet == 0.0 and not open_area.any()
    solid = [float("nan")] + [0.0] * 7
    assert O.measure_sampled_cell(1.0, [-1.0] * 8, solid)[0] == O.INVALID_ARGUMENT
    assert O.measure_sampled_cell(0.0, [-1.0] * 8, [0.0] * 8)[0] == O.INVALID_ARGUMENT


def solid_samples(corners, distance, velocity=lambda c: (0.0, 0.0, 0.0)):
    out = np.zeros(len(corners), O.SOLID_CORNER)
    for i, c in enumerate(corners["cell_index"]):
        idx = (int(c["x"]), int(c["y"]), int(c["z"]))
        out[i]["corner_index"] = idx
        out[i]["signed_distance"] = distance(idx)
        v = velocity(idx)
        out[i]["wall_velocity"]["x"], out[i]["wall_velocity"]["y"], out[i]["wall_velocity"]["z"] = v
    return out


def shared_face_problem(**kw):
    centers, corners = sample(lambda x, y, z: y - 0.75)
    for i, c in enumerate(centers["cell_index"]):
        if not (int(c["y"]) == -2 and int(c["z"]) == -2 and int(c["x"]) in (-2, -1)):
            centers["signed_distance"][i] = 1.0
    solid = solid_samples(corners, lambda c: (c[0] + c[1] + 2.75) / math.sqrt(2), **kw)
    return centers, corners, solid


def test_freeze_shares_one_pressure_face():
    centers, corners, solid = shared_face_problem()
    geom, blocks = grid()
    s, shared = O.face_slot(blocks, 0, (-1, -2, -2))
    assert s == O.OK
    velocities = np.zeros(192 * len(blocks))
    velocities[shared] = 0.2
    p = problem(centers=centers, corners=corners, candidates=((-2, -2, -2), (-1, -2, -2)), solid_corners=solid,
                face_normal_velocities=velocities)
    s, cells, faces, _ = O.freeze_geometry(p)
    assert s == O.OK and len(cells) == 2 and len(faces) == 11
    assert abs(cells[0]["fluid_volume"] - 0.125) <= 2e-14 and abs(cells[1]["fluid_volume"] - 0.71875) <= 2e-14
    f = faces[faces["face_slot"] == shared]
    assert len(f) == 1
    f = f[0]
    assert abs(f["open_area"] - 0.75) <= 2e-14 and abs(f["blocked_area"] - 0.25) <= 2e-14
    assert f["wall_normal_velocity"] == 0.2 and f["has_prescribed_wall_flux"] == 0
    # Links: the shared face is internal, its wall flux (1 - open) u_wall = 0.05.
    lp = O.LinkProblem(geom, blocks, centers, cells, faces)
    s, _, pf, _ = O.build_pressure_links(lp)
    assert s == O.OK
    k = int(np.nonzero(faces["face_slot"] == shared)[0][0])
    assert pf[k]["link_kind"] == 1 and abs(pf[k]["open_area"] - 0.75) <= 2e-14
    assert abs((1.0 - pf[k]["open_area"]) * pf[k]["wall_normal_velocity"] - 0.05) <= 2e-14

    # Affine corner velocities u_x = 2 + 4 (y + 2) with no dense face velocities: flux 5/8, mean 5/2.
    centers, corners, solid = shared_face_problem(velocity=lambda c: (2.0 + 4.0 * (c[1] + 2.0), 0.0, 0.0))
    p = problem(centers=centers, corners=corners, candidates=((-2, -2, -2), (-1, -2, -2)), solid_corners=solid)
    s, cells, faces, _ = O.freeze_geometry(p)
    assert s == O.OK
    f = faces[faces["face_slot"] == shared][0]
    assert f["has_prescribed_wall_flux"] == 1 and abs(f["prescribed_wall_flux"] - 0.625) <= 2e-14
    assert abs(f["open_area"] - 0.75) <= 2e-14 and abs(f["wall_normal_velocity"] - 2.5) <= 2e-14
    s, _, pf, _ = O.build_pressure_links(O.LinkProblem(geom, blocks, centers, cells, faces))
    assert s == O.OK and abs(pf[k]["prescribed_wall_flux"] - 0.625) <= 2e-14 and pf[k]["has_prescribed_wall_flux"]

    # A solid sample missing for a candidate corner.
    p = problem(centers=centers, corners=corners, candidates=((-2, -2, -2), (-1, -2, -2)), solid_corners=solid[:1])
    assert O.freeze_geometry(p)[0] == O.FAILED_PRECONDITION


def test_split_wet_lobes_refused():
    centers, corners, solid = shared_face_problem()
    for i, c in enumerate(solid["corner_index"]):
        idx = (int(c["x"]), int(c["y"]), int(c["z"]))
        solid["signed_distance"][i] = 1.0 if idx in ((-1, -2, -2), (-2, -1, -1)) else -1.0
    corners["signed_distance"] = -1.0
    p = problem(centers=centers, corners=corners, candidates=((-2, -2, -2),), solid_corners=solid)
    assert O.freeze_geometry(p)[0] == O.FAILED_PRECONDITION


def field_split_cell():
    """Cell (-2,-2,-2) whose liquid is two lobes at its corners (-1,-2,-2) and (-2,-1,-1) (corner ordinals
    1 and 6), every other corner dry: the particle field alone splits it; the solid is far (+1 everywhere)."""
    centers, corners = sample(lambda x, y, z: 1.0)
    for i, c in enumerate(corners["cell_index"]):
        if tuple(int(v) for v in c) in ((-1, -2, -2), (-2, -1, -1)):
            corners["signed_distance"][i] = -0.5
    centers["signed_distance"][:] = 1.0
    for i, c in enumerate(centers["cell_index"]):
        if tuple(int(v) for v in c) == (-2, -2, -2):
            centers["signed_distance"][i] = -0.1
    solid = solid_samples(corners, lambda c: 1.0)
    return centers, corners, solid


def test_split_only_where_a_solid_reaches():
    """The reference refuses the field-split cell in sampled mode; mode 2 measures it as the plane-obstacle
    freeze does (same kept cell, volume to rounding), and still refuses the solid-split lobes above."""
    centers, corners, solid = field_split_cell()
    p = problem(centers=centers, corners=corners, candidates=((-2, -2, -2),), solid_corners=solid)
    assert O.freeze_geometry(p)[0] == O.FAILED_PRECONDITION
    p = problem(centers=centers, corners=corners, candidates=((-2, -2, -2),), solid_corners=solid,
                split_rule=2)
    s, cells, faces, _ = O.freeze_geometry(p)
    assert s == O.OK
    s, plain, _, _ = O.freeze_geometry(problem(centers=centers, corners=corners, candidates=((-2, -2, -2),)))
    assert s == O.OK
    assert len(cells) == len(plain) == 1
    assert 0.0 < cells[0]["fluid_volume"] < 0.5
    assert abs(cells[0]["fluid_volume"] - plain[0]["fluid_volume"]) <= 1e-15
    # a corner in the solid keeps the reference's refusal
    centers, corners, solid = shared_face_problem()
    for i, c in enumerate(solid["corner_index"]):
        idx = (int(c["x"]), int(c["y"]), int(c["z"]))
        solid["signed_distance"][i] = 1.0 if idx in ((-1, -2, -2), (-2, -1, -1)) else -1.0
    corners["signed_distance"] = -1.0
    p = problem(centers=centers, corners=corners, candidates=((-2, -2, -2),), solid_corners=solid,
                split_rule=2)
    assert O.freeze_geometry(p)[0] == O.FAILED_PRECONDITION
    # rule 3 keeps the solid-split cell as one cell whose wet volume is the sum of its lobes
    p = problem(centers=centers, corners=corners, candidates=((-2, -2, -2),), solid_corners=solid,
                split_rule=3)
    s, cells, _, _ = O.freeze_geometry(p)
    assert s == O.OK and len(cells) == 1 and 0.0 < cells[0]["fluid_volume"] < 1.0
    p = problem(centers=centers, corners=corners, candidates=((-2, -2, -2),), solid_corners=solid,
                split_rule=4)
    assert O.freeze_geometry(p)[0] == O.INVALID_ARGUMENT


def test_inputs_rejected():
    centers, corners, solid = shared_face_problem()
    cands = ((-2, -2, -2), (-1, -2, -2))
    assert O.freeze_geometry(problem(centers=centers, corners=corners, candidates=cands, solid_corners=solid,
                                     face_normal_velocities=np.zeros(5)))[0] == O.INVALID_ARGUMENT
    bad = np.zeros(192 * 8)
    bad[3] = np.inf
    assert O.freeze_geometry(problem(centers=centers, corners=corners, candidates=cands, solid_corners=solid,
                                     face_normal_velocities=bad))[0] == O.INVALID_ARGUMENT
    assert O.freeze_geometry(problem(centers=centers, corners=corners, candidates=cands, solid_corners=solid,
                                     obstacle=O.slab((0, 1, 0), 0.0, 0.25)))[0] == O.INVALID_ARGUMENT
    unsorted = solid[::-1].copy()
    assert O.freeze_geometry(problem(centers=centers, corners=corners, candidates=cands,
                                     solid_corners=unsorted))[0] == O.INVALID_ARGUMENT


@pytest.mark.parametrize("normal,plane", [((0.0, 1.0, 0.0), 0.3), ((0.6, 0.8, 0.0), 0.1),
                                          ((1 / math.sqrt(3),) * 3, 0.2)])
def test_sampled_plane_matches_analytic_halfspace(normal, plane):
    """A solid sampled from the exact signed distance to a plane n.x = c (solid below, as the half-space
    obstacle) freezes to the half-space's volumes and open areas to rounding: the piecewise-linear
    interpolation of an affine distance is exact, so the two clip paths measure the same regions."""
    liquid = lambda x, y, z: 0.37 * x + 0.21 * y - 0.13 * z - 0.9
    cands = [(x, y, z) for z in (-3, -2, -1) for y in (-3, -2, -1) for x in (-3, -2, -1)]
    centers, corners = sample(liquid)
    n = np.asarray(normal)
    solid = solid_samples(corners, lambda c: float(n @ (np.asarray(c, dtype=float) + 2.0) - plane))
    s1, c1, f1, _ = O.freeze_geometry(problem(centers=centers, corners=corners, candidates=cands,
                                              obstacle=O.halfspace(normal, plane)))
    s2, c2, f2, _ = O.freeze_geometry(problem(centers=centers, corners=corners, candidates=cands,
                                              solid_corners=solid, face_normal_velocities=np.zeros(192 * 8)))
    assert s1 == s2 == O.OK and len(c1) == len(c2) > 0 and len(f1) == len(f2)
    assert np.all(c1["cell_index"] == c2["cell_index"])
    assert np.abs(c1["fluid_volume"] - c2["fluid_volume"]).max() <= 1e-12
    assert np.all(f1["face_slot"] == f2["face_slot"])
    assert np.abs(f1["open_area"] - f2["open_area"]).max() <= 1e-12
