"""GPU parity of the sampled-solid freeze (fl_freeze_geometry with a SampledSolidBoundary3D) against the
oracle.

Bar: the same status and, on success, bit-identical frozen cells and faces, on the reference's sampled-
solid freeze case (the shared aperture with dense and with affine corner wall velocities, the missing
sample, the split wet lobes and the input rejections) and on a liquid sphere in a 48^3 grid around a
static solid sphere (radius 0.15 at the centre) and a paddle (a thin box through the centre, turning
at 2 rad/s about z), whose signed distances are sampled at the cell corners: with the corner wall
velocities integrated, with dense face normal velocities, and clipped by an analytic half-space. A
solid sampled from an exact plane freezes to the analytic half-space's volumes and areas. The frozen
view then gives the oracle's links of the oracle's freeze.
"""
import math

import numpy as np
import pytest

import pyoracle as O
import test_oracle_freeze as F
import test_oracle_sampled_solid as S
from test_gpu_freeze import same, check, sphere

from fluidity_amd import hip as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(particles=1 << 10, blocks=1 << 12, lookup_blocks=1 << 16)
    print("device", c.device_name(), flush=True)
    yield c
    c.close()


def test_reference_sampled_case(ctx):
    centers, corners, solid = S.shared_face_problem()
    geom, blocks = F.grid()
    s, shared = O.face_slot(blocks, 0, (-1, -2, -2))
    velocities = np.zeros(192 * len(blocks))
    velocities[shared] = 0.2
    cands = ((-2, -2, -2), (-1, -2, -2))
    p = F.problem(centers=centers, corners=corners, candidates=cands, solid_corners=solid,
                  face_normal_velocities=velocities)
    assert check(ctx, p) == O.OK
    s, cells, faces, _ = ctx.freeze_geometry(p)
    f = faces[faces["face_slot"] == shared][0]
    assert abs(f["open_area"] - 0.75) <= 2e-14 and f["wall_normal_velocity"] == 0.2
    centers, corners, solid = S.shared_face_problem(velocity=lambda c: (2.0 + 4.0 * (c[1] + 2.0), 0.0, 0.0))
    p = F.problem(centers=centers, corners=corners, candidates=cands, solid_corners=solid)
    assert check(ctx, p) == O.OK
    s, cells, faces, _ = ctx.freeze_geometry(p)
    f = faces[faces["face_slot"] == shared][0]
    assert abs(f["prescribed_wall_flux"] - 0.625) <= 2e-14 and abs(f["wall_normal_velocity"] - 2.5) <= 2e-14
    assert check(ctx, F.problem(centers=centers, corners=corners, candidates=cands,
                                solid_corners=solid[:1])) == O.FAILED_PRECONDITION
    assert check(ctx, F.problem(centers=centers, corners=corners, candidates=cands, solid_corners=solid,
                                face_normal_velocities=np.zeros(5))) == O.INVALID_ARGUMENT
    assert check(ctx, F.problem(centers=centers, corners=corners, candidates=cands, solid_corners=solid[::-1].copy())
                 ) == O.INVALID_ARGUMENT
    assert check(ctx, F.problem(centers=centers, corners=corners, candidates=cands, solid_corners=solid,
                                obstacle=O.slab((0, 1, 0), 0.0, 0.25))) == O.INVALID_ARGUMENT
    lobes = solid.copy()
    for i, c in enumerate(lobes["corner_index"]):
        idx = (int(c["x"]), int(c["y"]), int(c["z"]))
        lobes["signed_distance"][i] = 1.0 if idx in ((-1, -2, -2), (-2, -1, -1)) else -1.0
    split_corners = corners.copy()
    split_corners["signed_distance"] = -1.0
    assert check(ctx, F.problem(centers=centers, corners=split_corners, candidates=((-2, -2, -2),),
                                solid_corners=lobes)) == O.FAILED_PRECONDITION


def paddle_distance(p, theta, centre=(0.5, 0.5, 0.5), half=(0.3, 0.03, 0.2)):
    """Signed distance to a box of half extents `half` turned by theta about z through `centre`."""
    x, y, z = p[..., 0] - centre[0], p[..., 1] - centre[1], p[..., 2] - centre[2]
    c, s = math.cos(theta), math.sin(theta)
    u, v = c * x + s * y, -s * x + c * y
    q = np.stack([np.abs(u) - half[0], np.abs(v) - half[1], np.abs(z) - half[2]], -1)
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(np.max(q, axis=-1), 0.0)
    return outside + inside


def config5_problem(n=48, mode="corners", omega=2.0, theta=0.3):
    """Liquid sphere (radius 0.42) around the solid sphere (radius 0.15) and paddle of the full-step
    configuration; solid distances and the rotating wall velocity sampled at the corners."""
    obstacle = O.halfspace((0.0, 1.0, 0.0), 0.2, (0.0, 0.5, 0.0)) if mode == "halfspace" else None
    p = sphere(n=n, sigma=0.0, step=1e-3, obstacle=obstacle)
    h = 1.0 / n
    ci = p.corners["cell_index"]
    pts = np.stack([ci["x"] * h, ci["y"] * h, ci["z"] * h], -1).astype(np.float64)
    ball = np.linalg.norm(pts - 0.5, axis=-1) - 0.15
    solid = np.minimum(ball, paddle_distance(pts, theta))
    rec = np.zeros(len(pts), O.SOLID_CORNER)
    rec["corner_index"] = ci
    rec["signed_distance"] = solid
    rec["wall_velocity"]["x"] = -omega * (pts[:, 1] - 0.5)
    rec["wall_velocity"]["y"] = omega * (pts[:, 0] - 0.5)
    p.solid_corners = rec
    if mode in ("dense", "halfspace"):
        nf = 192 * len(p.blocks)
        slots = np.arange(nf)
        b = p.blocks[slots // 192]
        local = (slots % 192) % 64
        axis = (slots % 192) // 64
        lx, ly, lz = 4 * b["x"] + local % 4, 4 * b["y"] + (local // 4) % 4, 4 * b["z"] + local // 16
        cx = (lx + np.where(axis == 0, 0.0, 0.5)) * h
        cy = (ly + np.where(axis == 1, 0.0, 0.5)) * h
        vel = np.where(axis == 0, -omega * (cy - 0.5), np.where(axis == 1, omega * (cx - 0.5), 0.0))
        p.face_normal_velocities = vel.astype(np.float64)
    return p


@pytest.mark.parametrize("mode", ["corners", "dense", "halfspace"])
def test_config5_solids(ctx, mode):
    p = config5_problem(mode=mode)
    assert check(ctx, p) == O.OK
    s, cells, faces, _ = ctx.freeze_geometry(p)
    partial = int(((cells["fluid_volume"] > 0) & (cells["fluid_volume"] < (1 / 48) ** 3)).sum())
    blocked = int((faces["blocked_area"] > 0).sum())
    print("sampled", mode, len(p.candidates), "candidates", len(cells), "cells", partial, "partial", blocked,
          "blocked faces", ctx.last.kernel_ms, "ms", flush=True)
    assert partial > 100 and blocked > 100
    if mode != "dense":
        assert faces["has_prescribed_wall_flux"].all()


def test_sampled_plane_matches_halfspace(ctx):
    normal, plane = (0.6, 0.8, 0.0), 0.1
    liquid = lambda x, y, z: 0.37 * x + 0.21 * y - 0.13 * z - 0.9
    cands = [(x, y, z) for z in (-3, -2, -1) for y in (-3, -2, -1) for x in (-3, -2, -1)]
    centers, corners = F.sample(liquid)
    n = np.asarray(normal)
    solid = S.solid_samples(corners, lambda c: float(n @ (np.asarray(c, dtype=float) + 2.0) - plane))
    ph = F.problem(centers=centers, corners=corners, candidates=cands, obstacle=O.halfspace(normal, plane))
    ps = F.problem(centers=centers, corners=corners, candidates=cands, solid_corners=solid,
                   face_normal_velocities=np.zeros(192 * 8))
    assert check(ctx, ph) == O.OK and check(ctx, ps) == O.OK
    _, c1, f1, _ = ctx.freeze_geometry(ph)
    _, c2, f2, _ = ctx.freeze_geometry(ps)
    assert np.abs(c1["fluid_volume"] - c2["fluid_volume"]).max() <= 1e-12
    assert np.abs(f1["open_area"] - f2["open_area"]).max() <= 1e-12


def test_sampled_freeze_then_links(ctx):
    p = config5_problem(mode="corners")
    s, cells, faces, _ = ctx.freeze_geometry(p)
    assert s == O.OK
    rng = np.random.default_rng(5)
    links = O.LinkProblem(p.geom, p.blocks, p.centers, cells.view(O.FROZEN_CELL), faces.view(O.FROZEN_FACE),
                          step=p.step, density=p.density, sigma=p.sigma, predicted=rng.normal(0, 1, len(faces)))
    s_want, wc, wf, wl = O.build_pressure_links(links)
    s_got, gc, gf, gl = ctx.build_pressure_links(links)
    assert s_want == O.OK and s_got == s_want and gl == wl
    same(gc.view(O.PRESSURE_CELL), wc)
    same(gf.view(O.PRESSURE_FACE), wf)


@pytest.mark.parametrize("rule", [1, 2, 3])
def test_split_rules(ctx, rule):
    """The wet-piece refusal rules, bit-exact with the oracle: a cell the particle field alone splits is
    refused by the reference's rule (1) and measured by rules 2 and 3; lobes a solid separates are
    refused by rules 1 and 2 and measured as one cell by rule 3."""
    centers, corners, solid = S.field_split_cell()
    want = O.FAILED_PRECONDITION if rule == 1 else O.OK
    assert check(ctx, F.problem(centers=centers, corners=corners, candidates=((-2, -2, -2),), solid_corners=solid,
                                split_rule=rule)) == want
    centers, corners, solid = S.shared_face_problem()
    for i, c in enumerate(solid["corner_index"]):
        idx = (int(c["x"]), int(c["y"]), int(c["z"]))
        solid["signed_distance"][i] = 1.0 if idx in ((-1, -2, -2), (-2, -1, -1)) else -1.0
    corners["signed_distance"] = -1.0
    want = O.OK if rule == 3 else O.FAILED_PRECONDITION
    assert check(ctx, F.problem(centers=centers, corners=corners, candidates=((-2, -2, -2),), solid_corners=solid,
                                split_rule=rule)) == want
