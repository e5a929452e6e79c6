"""Write tests/golden/corner_transfer_3d.json: the single-particle corner deposit in exact rationals.

The case is the reference's known-answer test (native/tests/test_transfers_3d.cc:145-203, data
from reference/src/fluidity_reference/transfer_3d_fixtures.py:23-43): unit cells at origin 0,
eight blocks {0,1}x{-1,0}x{0,1}, particle at (15/4, -1/4, 15/4), mass 2, velocity (1,-2,4) and
affine rows (2,4,-1), (-5,1,5), (7,-2,3). Values are recomputed here from the quadratic
B-spline definition with Python Fractions: w(d) = 3/4 - d^2 for |d| <= 1/2,
(3/2 - |d|)^2 / 2 for 1/2 < |d| < 3/2, else 0, with d the particle-minus-face offset in cells.
Run: python tests/golden/make_transfer_fixture.py
"""
import json
import os
from fractions import Fraction as F
from itertools import product

HALF = F(1, 2)


def spline(d):
    """Weight and derivative w.r.t. the particle coordinate, in cell units."""
    a = abs(d)
    if a <= HALF:
        return F(3, 4) - a * a, -2 * d
    if a < F(3, 2):
        e = F(3, 2) - a
        return e * e / 2, (e if d < 0 else -e)
    return F(0), F(0)


def main():
    pos = (F(15, 4), F(-1, 4), F(15, 4))
    mass = F(2)
    vel = (F(1), F(-2), F(4))
    rows = ((2, 4, -1), (-5, 1, 5), (7, -2, 3))
    out = {"position": [float(v) for v in pos], "mass": 2.0, "velocity": [1.0, -2.0, 4.0],
           "affine_rows": [[float(v) for v in r] for r in rows],
           "blocks": [[x, y, z] for z in (0, 1) for y in (-1, 0) for x in (0, 1)],
           "faces": [], "weight": [], "offset": [], "gradient": [], "mass_contribution": [],
           "momentum_contribution": []}
    for axis in range(3):
        for lz, ly, lx in product(range(8), range(-4, 4), range(8)):
            lat = (lx, ly, lz)
            anchor = [F(lat[c]) + (0 if c == axis else HALF) for c in range(3)]
            r = [anchor[c] - pos[c] for c in range(3)]
            wd = [spline(-r[c]) for c in range(3)]
            w = wd[0][0] * wd[1][0] * wd[2][0]
            if w == 0:
                continue
            grad = [wd[c][1] * wd[(c + 1) % 3][0] * wd[(c + 2) % 3][0] for c in range(3)]
            m = w * mass
            v = vel[axis] + sum(F(rows[axis][c]) * r[c] for c in range(3))
            out["faces"].append([axis, *lat])
            out["weight"].append(float(w))
            out["offset"].append([float(x) for x in r])
            out["gradient"].append([float(x) for x in grad])
            out["mass_contribution"].append(float(m))
            out["momentum_contribution"].append(float(m * v))
    assert len(out["faces"]) == 81
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "corner_transfer_3d.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
