"""Pins the oracle's frozen boundary geometry (flo_freeze_geometry).

Known answers are the reference's geometry cases (native/tests/liquid_boundary_geometry_3d_cases.h)
on its 2x2x2-block grid with the candidate cell (-2, -2, -2) whose low corner is the world origin:
the plane fixture (reference/src/fluidity_reference/liquid_boundary_3d_fixtures.py:26-29, open width
1 - d for a plane at d), oblique and simultaneous liquid/solid clips, the thin slab and its split-cell
refusal, the shared face, plane and sphere curvature, the capillary bound, zero plateaus, tiny and
underflowing measures, and every rejection those cases hold. Closed forms add: the volume of a liquid
half-space cut at height t along a unit axis (t h^3), the scaled corner copies of the six
tetrahedra around one negative corner, and the area / volume of all-liquid cells without an
obstacle (h^2, h^3).
"""
import math

import numpy as np
import pytest

import pyoracle as O

ALL = [(x, y, z) for z in (-1, 0) for y in (-1, 0) for x in (-1, 0)]


def grid(h=1.0):
    return O.geometry((2 * h, 2 * h, 2 * h), h), O.int3(ALL)


def sample(field, lo=-4, hi=1):
    """Centre samples at field(i + 2.5), corner samples at field(i + 2), in (z, y, x) order."""
    idx = [(x, y, z) for z in range(lo, hi + 1) for y in range(lo, hi + 1) for x in range(lo, hi + 1)]
    centers = np.zeros(len(idx), O.CENTER_DISTANCE)
    corners = np.zeros(len(idx), O.CENTER_DISTANCE)
    for k, (x, y, z) in enumerate(idx):
        centers[k] = ((x, y, z), field(x + 2.5, y + 2.5, z + 2.5))
        corners[k] = ((x, y, z), field(x + 2.0, y + 2.0, z + 2.0))
    return centers, corners


def problem(field=None, centers=None, corners=None, candidates=((-2, -2, -2),), h=1.0, **kw):
    if field is not None:
        centers, corners = sample(field)
    geom, blocks = grid(h)
    return O.FreezeProblem(geom, blocks, O.int3(list(candidates)), centers, corners, **kw)


def isolated(center=-0.01):
    centers, corners = sample(lambda x, y, z: 0.99)
    corners["signed_distance"] = -1.0
    k = [i for i, c in enumerate(centers["cell_index"]) if tuple(c) == (-2, -2, -2)][0]
    centers["signed_distance"][k] = center
    return centers, corners


def freeze(p):
    return O.freeze_geometry(p)


def axis_coordinate(face):
    return int(face["face_index"]["lattice"][["x", "y", "z"][int(face["face_index"]["axis"])]])


@pytest.mark.parametrize("offset", [0.0, 0.25, 1.0 - 2.0 ** -20])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_plane_fixture(axis, offset):
    normal = [0.0, 0.0, 0.0]
    normal[axis] = 1.0
    s, cells, faces, bound = freeze(problem(lambda x, y, z: -1.0, obstacle=O.halfspace(normal, offset, (2, 3, 4))))
    assert s == O.OK and len(cells) == 1 and len(faces) == 6 and bound == math.inf
    assert abs(cells[0]["fluid_volume"] - (1 - offset)) <= 2e-14
    for f in faces:
        assert f["blocked_area"] == 1.0 - f["open_area"]
        assert f["wall_normal_velocity"] == 2.0 + int(f["face_index"]["axis"])
        if int(f["face_index"]["axis"]) != axis:
            assert abs(f["open_area"] - (1 - offset)) <= 2e-14


def test_oblique_and_simultaneous_clips():
    centers, corners = sample(lambda x, y, z: x + y + z - 1.0)
    centers["signed_distance"] = -1.0
    s, cells, _, _ = freeze(problem(centers=centers, corners=corners, obstacle=O.halfspace((1, 0, 0), 0.25)))
    assert s == O.OK and abs(cells[0]["fluid_volume"] - 9 / 128) <= 2e-14
    q = 1 / math.sqrt(3)
    s, cells, faces, _ = freeze(problem(lambda x, y, z: -1.0, obstacle=O.halfspace((q, q, q), q)))
    assert s == O.OK and abs(cells[0]["fluid_volume"] - 5 / 6) <= 2e-14
    for f in faces:
        assert abs(f["open_area"] - (0.5 if axis_coordinate(f) == -2 else 1.0)) <= 2e-14


def test_thin_slab_and_split_cell():
    slab = O.slab((1, 0, 0), 0.5, 1 / 32)
    s, cells, faces, _ = freeze(problem(lambda x, y, z: x - 0.52, obstacle=slab))
    assert s == O.OK and abs(cells[0]["fluid_volume"] - 15 / 32) <= 2e-14
    for f in faces:
        if int(f["face_index"]["axis"]) != 0:
            assert abs(f["blocked_area"] - 1 / 16) <= 2e-14
    assert freeze(problem(lambda x, y, z: -1.0, obstacle=slab))[0] == O.FAILED_PRECONDITION


def test_shared_face_one_owner():
    q = 1 / math.sqrt(2)
    s, cells, faces, _ = freeze(problem(lambda x, y, z: -1.0, candidates=((-2, -2, -2), (-1, -2, -2)),
                                        obstacle=O.halfspace((0, q, q), q)))
    assert s == O.OK and len(cells) == 2 and len(faces) == 11
    assert all(abs(c["fluid_volume"] - 0.5) <= 2e-14 for c in cells)
    assert np.all(np.diff(faces["face_slot"].astype(np.int64)) > 0)
    for f in faces:
        if int(f["face_index"]["axis"]) == 0:
            assert abs(f["open_area"] - 0.5) <= 2e-14


def test_plane_and_sphere_curvature():
    s, cells, _, _ = freeze(problem(lambda x, y, z: x - 0.75, sigma=0.072))
    assert s == O.OK and abs(cells[0]["curvature"]) <= 1e-14
    worst = 0.0
    for radius in (8.0, 16.0, 32.0):
        for d in ((1, 0, 0), (0.6, 0.8, 0), (1 / math.sqrt(3),) * 3):
            f = lambda x, y, z: math.hypot(x - 0.5 + radius * d[0], y - 0.5 + radius * d[1],
                                           z - 0.5 + radius * d[2]) - radius - 0.01
            s, cells, _, _ = freeze(problem(f, sigma=0.072))
            assert s == O.OK
            worst = max(worst, abs(cells[0]["curvature"] / (2 / radius) - 1))
    assert worst < 0.05


def test_rejections():
    centers, corners = isolated()
    assert freeze(problem(centers=centers, corners=corners[:0]))[0] == O.FAILED_PRECONDITION
    dup = np.concatenate([centers, centers[:1]])
    assert freeze(problem(centers=dup, corners=corners))[0] == O.INVALID_ARGUMENT
    assert freeze(problem(centers=centers, corners=corners, obstacle=O.halfspace((2, 0, 0), 0)))[0] == \
        O.INVALID_ARGUMENT
    geom, _ = grid()
    empty = O.FreezeProblem(geom, np.zeros(0, O.INT3), O.int3([(-2, -2, -2)]), centers, corners)
    assert freeze(empty)[0] == O.FAILED_PRECONDITION
    assert freeze(problem(centers=centers, corners=corners, sigma=-1))[0] == O.INVALID_ARGUMENT
    assert freeze(problem(centers=centers, corners=corners, step=0))[0] == O.INVALID_ARGUMENT
    assert freeze(problem(centers=centers, corners=corners, density=0))[0] == O.INVALID_ARGUMENT
    assert freeze(problem(centers=centers, corners=corners, h=1e-120))[0] == O.OUT_OF_RANGE
    assert freeze(problem(centers=centers, corners=corners, obstacle=O.halfspace((1, 0, 0), 1e-20)))[0] == \
        O.OUT_OF_RANGE
    assert freeze(problem(centers=centers, corners=corners, obstacle=O.slab((1, 0, 0), 1e30, 1e-20)))[0] == \
        O.OUT_OF_RANGE
    bad = centers.copy()
    bad["signed_distance"][0] = math.nan
    assert freeze(problem(centers=bad, corners=corners))[0] == O.INVALID_ARGUMENT
    assert freeze(problem(centers=centers, corners=corners, candidates=[(-2, -2, -2)] * 4097, max_cells=4096))[0] == \
        O.RESOURCE_EXHAUSTED
    # unsorted candidates, slab half width, exact curvature
    assert freeze(problem(centers=centers, corners=corners, candidates=((-1, -2, -2), (-2, -2, -2))))[0] == \
        O.INVALID_ARGUMENT
    assert freeze(problem(centers=centers, corners=corners, obstacle=O.slab((1, 0, 0), 0.5, 0.0)))[0] == \
        O.INVALID_ARGUMENT
    assert freeze(problem(centers=centers, corners=corners, exact_curvature=math.inf))[0] == O.INVALID_ARGUMENT


def test_capillary_bound():
    centers, corners = isolated()
    s, _, _, bound = freeze(problem(centers=centers, corners=corners))
    assert s == O.OK and bound == math.inf
    # a lone liquid centre has a zero centred gradient
    assert freeze(problem(centers=centers, corners=corners, sigma=0.072))[0] == O.FAILED_PRECONDITION
    centers, corners = sample(lambda x, y, z: x - 0.75)
    s, _, _, bound = freeze(problem(centers=centers, corners=corners, sigma=0.072))
    want = math.sqrt(1000.0 / (4.0 * math.acos(-1.0) * 0.072))
    assert s == O.OK and abs(bound - want) <= 1e-13
    assert freeze(problem(centers=centers, corners=corners, sigma=0.072, step=want * 1.001))[0] == \
        O.FAILED_PRECONDITION
    short = np.delete(centers, 87)
    assert freeze(problem(centers=short, corners=corners, sigma=0.072))[0] != O.OK
    # the exact curvature replaces the centred one; its lift must be normal
    s, cells, _, _ = freeze(problem(centers=centers, corners=corners, sigma=0.072, exact_curvature=-3.0))
    assert s == O.OK and cells[0]["curvature"] == -3.0
    assert freeze(problem(centers=centers, corners=corners, sigma=0.072, exact_curvature=1e-320))[0] == O.OUT_OF_RANGE


def test_interior_positive_sigma_needs_no_curvature():
    centers, corners = sample(lambda x, y, z: -1.0)
    keep = [abs(c["x"] + 2) + abs(c["y"] + 2) + abs(c["z"] + 2) <= 1 for c in centers["cell_index"]]
    s, cells, _, bound = freeze(problem(centers=centers[keep], corners=corners, sigma=0.072, step=1e6))
    assert s == O.OK and len(cells) == 1 and bound == math.inf


def test_plateaus_and_tiny_measures():
    centers, corners = isolated()
    zero = corners.copy()
    zero["signed_distance"] = 0.0
    s, cells, faces, _ = freeze(problem(centers=centers, corners=zero))
    assert s == O.OK and len(cells) == 0 and len(faces) == 0
    mixed = corners.copy()
    mixed["signed_distance"] = [-1.0 if tuple(c) == (-1, -2, -2) else 0.0 for c in corners["cell_index"]]
    s, cells, _, _ = freeze(problem(centers=centers, corners=mixed))
    assert s == O.OK and abs(cells[0]["fluid_volume"] - 1 / 3) <= 2e-14
    c2, k2 = sample(lambda x, y, z: x + y + z - 1e-100)
    c2["signed_distance"] = -1.0
    s, cells, _, _ = freeze(problem(centers=c2, corners=k2))
    assert s == O.OK and abs(cells[0]["fluid_volume"] / 1e-300 - 1 / 6) <= 1e-14
    c3, k3 = sample(lambda x, y, z: x + y + z - 1e-110)
    c3["signed_distance"] = -1.0
    s, cells, _, _ = freeze(problem(centers=c3, corners=k3, h=1e100))
    assert s == O.OK and abs(cells[0]["fluid_volume"] / 1e-30 - 1 / 6) <= 2e-14
    d = 1 / math.sqrt(2)
    c4, k4 = sample(lambda x, y, z: -1.0)
    s, cells, faces, _ = freeze(problem(centers=c4, corners=k4, h=1e100, obstacle=O.halfspace((-d, -d, 0), -d * 1e-70)))
    assert s == O.OK and abs(cells[0]["fluid_volume"] / 1e-40 - 0.5) <= 2e-14
    z = [f for f in faces if int(f["face_index"]["axis"]) == 2]
    assert len(z) == 2
    for f in z:
        assert abs(f["open_area"] / 1e-140 - 0.5) <= 2e-14 and f["blocked_area"] == 1e200 - f["open_area"]


@pytest.mark.parametrize("t", [0.125, 0.3, 0.5, 0.9])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_axis_plane_volume_closed_form(axis, t):
    """A liquid half-space below height t along one axis fills t h^3 of the cell (h = 0.5)."""
    h = 0.5
    centers, corners = sample(lambda *p: p[axis] - t)
    centers["signed_distance"] = -1.0  # centre samples are independent of the corners
    s, cells, faces, _ = freeze(problem(centers=centers, corners=corners, h=h))
    assert s == O.OK and len(cells) == 1
    assert abs(cells[0]["fluid_volume"] - t * h ** 3) <= 8 * np.spacing(h ** 3)
    assert np.all(np.abs(faces["open_area"] - h * h) <= 4 * np.spacing(h * h))


def test_full_cells_without_obstacle():
    """Every all-liquid candidate is one frozen cell of volume h^3 with six open faces of area h^2."""
    cands = [(x, y, z) for z in (-3, -2) for y in (-3, -2) for x in (-3, -2)]
    s, cells, faces, _ = freeze(problem(lambda x, y, z: -1.0, candidates=cands, h=0.25))
    assert s == O.OK and len(cells) == 8 and len(faces) == 36
    # twenty-four clipped tetrahedra summed in rounding arithmetic: within a few ulp of h^3
    assert np.all(np.abs(cells["fluid_volume"] - 0.25 ** 3) <= 16 * np.spacing(0.25 ** 3))
    assert np.all(np.abs(faces["open_area"] - 0.0625) <= 4 * np.spacing(0.0625))
    assert np.all(faces["blocked_area"] == 0.0625 - faces["open_area"])
    assert np.all(faces["wall_normal_velocity"] == 0.0)


def test_single_tetrahedron_closed_form():
    """Only corner 0 negative (phi = -a at 0, 1 elsewhere): every one of the six body-diagonal
    tetrahedra holds corner 0, and its linear interpolant is negative on the copy of itself scaled by
    a/(1+a) about that corner, so the liquid volume is (a/(1+a))^3 h^3."""
    a = 0.5
    centers, corners = isolated(-0.5)
    corners["signed_distance"] = [-a if tuple(c) == (-2, -2, -2) else 1.0 for c in corners["cell_index"]]
    s, cells, _, _ = freeze(problem(centers=centers, corners=corners))
    leg = a / (1 + a)
    assert s == O.OK and abs(cells[0]["fluid_volume"] - leg ** 3) <= 1e-15


def test_air_centre_or_empty_volume_is_skipped():
    centers, corners = sample(lambda x, y, z: -1.0)
    k = [i for i, c in enumerate(centers["cell_index"]) if tuple(c) == (-2, -2, -2)][0]
    centers["signed_distance"][k] = 0.0
    s, cells, faces, _ = freeze(problem(centers=centers, corners=corners))
    assert s == O.OK and len(cells) == 0 and len(faces) == 0
