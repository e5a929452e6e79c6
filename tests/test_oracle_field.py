"""The oracle's particle-union liquid field (flo_particle_field), pinned without the reference.

* The reference's two-particle case (native/tests/liquid_tank_staging_3d_cases.h:15-58): particles
  (-0.5, 0.5, 3.5), (0.5, 0.5, 4.5) and a repeat of the first, radius 0.75 on the unit grid give 42
  candidates, 170 centre samples and 92 corner samples, each strictly (z,y,x) increasing, and every
  distance is FieldDistance (min over the particles of the distance minus the radius).
* Closed forms: a particle's candidate box is floor((x -/+ r - o) / h) on each axis, checked in exact
  rational arithmetic; the centre set is the 27-neighbourhood and the corner set the 8 corners of the
  candidates, checked by brute force; distances against Python's hypot.
* Statuses: a non-finite particle or radius is InvalidArgument, an unrepresentable box OutOfRange, more
  candidates than the capacity ResourceExhausted.
"""
import math
from fractions import Fraction

import numpy as np

import pyoracle as O


def keys(records):
    if records.dtype.names and "cell_index" in records.dtype.names:
        records = records["cell_index"]
    return [(int(r["z"]), int(r["y"]), int(r["x"])) for r in records]


def field_distance(positions, radius, world):
    return min(math.hypot(world[0] - p[0], world[1] - p[1], world[2] - p[2]) - radius for p in positions)


def check_distances(geom, positions, radius, samples, offset):
    h = float(geom[0]["cell_width"])
    o = geom[0]["world_origin"]
    origin = (float(o["x"]), float(o["y"]), float(o["z"]))
    for rec in samples:
        c = rec["cell_index"]
        world = tuple(origin[a] + (int(c[k]) + offset) * h for a, k in enumerate("xyz"))
        want = field_distance(positions, radius, world)
        assert abs(rec["signed_distance"] - want) <= 4e-16 * max(1.0, abs(want)), (rec, want)


def test_reference_two_particle_case():
    geom = O.geometry((0.0, 0.0, 0.0), 1.0)
    pos = np.array([[-0.5, 0.5, 3.5], [0.5, 0.5, 4.5], [-0.5, 0.5, 3.5]])
    s, cells, centres, corners = O.particle_field(geom, pos, 0.75)
    assert s == O.OK
    assert (len(cells), len(centres), len(corners)) == (42, 170, 92)
    for recs in (cells, centres, corners):
        k = keys(recs)
        assert all(a < b for a, b in zip(k, k[1:]))
    check_distances(geom, pos, 0.75, centres, 0.5)
    check_distances(geom, pos, 0.75, corners, 0.0)


def test_boxes_and_sample_sets_by_brute_force():
    rng = np.random.default_rng(5)
    h, origin, radius = 0.5, (-1.0, 0.0, 0.25), 0.4
    geom = O.geometry(origin, h)
    pos = np.array(origin) + rng.random((40, 3)) * 3.0
    s, cells, centres, corners = O.particle_field(geom, pos, radius)
    assert s == O.OK
    want = set()
    for p in pos:
        lo, hi = [], []
        for a in range(3):
            c, r, o = Fraction(p[a]), Fraction(radius), Fraction(origin[a])
            # floor of the rounded quotient: the reference divides the rounded difference by h
            lo.append(math.floor(((p[a] - radius) - origin[a]) / h))
            hi.append(math.floor(((p[a] + radius) - origin[a]) / h))
            # the rounded and the exact box agree away from cell edges
            assert abs(lo[-1] - math.floor((c - r - o) / Fraction(h))) <= 1
        for z in range(lo[2], hi[2] + 1):
            for y in range(lo[1], hi[1] + 1):
                for x in range(lo[0], hi[0] + 1):
                    want.add((z, y, x))
    assert keys(cells) == sorted(want)
    cen = sorted({(z + dz, y + dy, x + dx) for z, y, x in want for dz in (-1, 0, 1) for dy in (-1, 0, 1)
                  for dx in (-1, 0, 1)})
    cor = sorted({(z + dz, y + dy, x + dx) for z, y, x in want for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)})
    assert keys(centres) == cen
    assert keys(corners) == cor
    check_distances(geom, pos, radius, centres[::7], 0.5)
    check_distances(geom, pos, radius, corners[::7], 0.0)


def test_statuses():
    geom = O.geometry((0.0, 0.0, 0.0), 1.0)
    pos = np.array([[-0.5, 0.5, 3.5], [0.5, 0.5, 4.5]])
    assert O.particle_field(geom, pos, 0.75, candidate_capacity=41)[0] == O.RESOURCE_EXHAUSTED
    assert O.particle_field(geom, pos, 0.75, candidate_capacity=42)[0] == O.OK
    assert O.particle_field(geom, pos, 0.0)[0] == O.INVALID_ARGUMENT
    assert O.particle_field(geom, pos, float("nan"))[0] == O.INVALID_ARGUMENT
    bad = pos.copy()
    bad[1, 2] = float("inf")
    assert O.particle_field(geom, bad, 0.75)[0] == O.INVALID_ARGUMENT
    far = pos.copy()
    far[0, 0] = 3e9
    assert O.particle_field(geom, far, 0.75)[0] == O.OUT_OF_RANGE
