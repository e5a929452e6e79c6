"""GPU parity of the open-tank freeze (fl_freeze_geometry with the tank obstacle) against the oracle.

Bar: the same status and, on success, bit-identical frozen cells and faces, on the reference's tank
cases (the clipped volume and the piston-only flux, a retreating piston, a closed high-x wall, the
liquid-clipped, nearly dry and 2^-45-wide cells, the unresolvable local cut and the tank step refusals)
and on a liquid pool filling a tank in a 48^3 grid whose walls and piston cut cells at non-lattice
positions; the frozen view then gives the oracle's links of the oracle's freeze.
"""
import numpy as np
import pytest

import pyoracle as O
import test_oracle_freeze as F
import test_oracle_tank_freeze as T
from test_gpu_freeze import check, same

from fluidity_amd import hip as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(particles=1 << 10, blocks=1 << 12, lookup_blocks=1 << 16)
    print("device", c.device_name(), flush=True)
    yield c
    c.close()


def test_reference_tank_cases(ctx):
    centers, corners = T.only_centre_liquid()
    for tank in (T.tank(), T.tank(piston=-0.125), T.tank(high_x=1.0)):
        assert check(ctx, F.problem(centers=centers, corners=corners, obstacle=tank, step=0.5, density=1.0)) == O.OK
    s, cells, faces, _ = ctx.freeze_geometry(F.problem(centers=centers, corners=corners, obstacle=T.tank(), step=0.5,
                                                       density=1.0))
    assert abs(cells[0]["fluid_volume"] - 0.41015625) <= 2e-14 and faces["prescribed_wall_flux"].max() == 0.068359375
    centers, corners = F.sample(lambda x, y, z: y - 0.75)
    assert check(ctx, F.problem(centers=centers, corners=corners, obstacle=T.tank())) == O.OK
    thin_c, thin_k = F.sample(lambda x, y, z: -1.0 if y == 0.5 else y - (0.125 + 2.0 ** -30))
    assert check(ctx, F.problem(centers=thin_c, corners=thin_k, obstacle=T.tank())) == O.OK
    narrow = T.tank(start_low_x=0.0, high_x=2.0 ** -45, piston=0.0)
    assert check(ctx, F.problem(centers=centers, corners=corners, obstacle=narrow)) == O.OK
    collapsed = T.tank(-2.0 ** -60, 0.0, 0.125, 0.125, 0.75, 0.0)
    assert check(ctx, F.problem(centers=centers, corners=corners, candidates=((-3, -2, -2),),
                                obstacle=collapsed)) == O.OUT_OF_RANGE
    for tank, step, code in ((T.tank(0.0, 0.0, piston=0.0), 0.01, O.INVALID_ARGUMENT),
                             (T.tank(0.0, 1.0, high_z=float("nan"), piston=0.0), 0.01, O.INVALID_ARGUMENT),
                             (T.tank(0.0, 1.0, piston=1.0), 1.0, O.FAILED_PRECONDITION),
                             (T.tank(0.0, 1.0, piston=float("inf")), 0.01, O.INVALID_ARGUMENT),
                             (T.tank(0.0, 1.0, piston=-np.finfo(np.float64).max), 2.0, O.OUT_OF_RANGE)):
        assert check(ctx, F.problem(centers=centers, corners=corners, obstacle=tank, step=step)) == code


def pool_in_tank(n=48):
    """A pool below y = 0.6137 in a 48^3 grid inside a tank whose piston, far wall, floor and z walls cut
    cells at non-lattice positions; the piston moves in +x."""
    from test_gpu_freeze import sphere
    p = sphere(n=n, sigma=0.0, step=1e-3)
    h = 1.0 / n
    for rec, off in ((p.centers, 0.5), (p.corners, 0.0)):
        y = (rec["cell_index"]["y"] + off) * h
        rec["signed_distance"] = y - 0.6137
    ci = p.centers["cell_index"]
    inner = ((ci["x"] >= 0) & (ci["x"] < n) & (ci["y"] >= 0) & (ci["y"] < n) & (ci["z"] >= 0) & (ci["z"] < n) &
             (p.centers["signed_distance"] < 1.5 * h))
    p.candidates = np.ascontiguousarray(ci[inner])
    p.obstacle = T.tank(0.1031, 0.8917, 0.0613, 0.0719, 0.9283, 0.37)
    return p


def test_pool_in_tank(ctx):
    p = pool_in_tank()
    assert check(ctx, p) == O.OK
    s, cells, faces, _ = ctx.freeze_geometry(p)
    full = (1 / 48) ** 3
    partial = int(((cells["fluid_volume"] > 0) & (cells["fluid_volume"] < full)).sum())
    pistons = int(faces["is_low_x_piston_face"].sum())
    print("tank", len(p.candidates), "candidates", len(cells), "cells", partial, "partial", pistons, "piston faces",
          ctx.last.kernel_ms, "ms", flush=True)
    assert partial > 1000 and pistons > 100
    links = O.LinkProblem(p.geom, p.blocks, p.centers, cells.view(O.FROZEN_CELL), faces.view(O.FROZEN_FACE),
                          step=p.step, density=p.density, sigma=p.sigma,
                          predicted=np.random.default_rng(2).normal(0, 1, len(faces)))
    s_want, wc, wf, wl = O.build_pressure_links(links)
    s_got, gc, gf, gl = ctx.build_pressure_links(links)
    assert s_want == O.OK and s_got == s_want and gl == wl
    same(gc.view(O.PRESSURE_CELL), wc)
    same(gf.view(O.PRESSURE_FACE), wf)
