"""GPU parity of the typed pressure links of a frozen view (fl_build_pressure_links) against the oracle.

Bar: the same status and, on success, bit-identical cell and face records and the same count of
raised theta values, on the oracle tests' cases (the reference's unit-cube fixture and rejections,
free-surface closed forms, capillary lift and gas pressure, internal links with densities and an
interface crossing, sealed faces and piston fluxes, field and sample rejections) and on a
liquid sphere of 34,416 cells with analytic centre distances, cell densities, interface crossings and
air-cell pressures.
"""
import math

import numpy as np
import pytest

import pyoracle as O
import test_oracle_links as L
from fluidity_amd import hip as H
from fluidity_amd import records as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(particles=1 << 10, blocks=1 << 10, lookup_blocks=1 << 14)
    print("device", c.device_name(), flush=True)
    yield c
    c.close()


def same(a, b):
    for name in a.dtype.names:
        x, y = a[name], b[name]
        if x.dtype.names:
            same(x, y)
        elif x.dtype.kind == "f":
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), name
        else:
            assert np.array_equal(x, y), name


def check(ctx, p):
    s_want, wc, wf, wl = O.build_pressure_links(p)
    s_got, gc, gf, gl = ctx.build_pressure_links(p)
    assert s_got == s_want, (s_got, s_want)
    if s_want == O.OK:
        same(gc.view(O.PRESSURE_CELL), wc)
        same(gf.view(O.PRESSURE_FACE), wf)
        assert gl == wl
    return s_want


def test_reference_cases(ctx):
    assert check(ctx, L.fixture(ambient=9.0)) == O.OK
    for density in (-1.0, 5e-324, 1.7976931348623157e308):
        check(ctx, L.fixture(density=density))
    p = L.fixture()
    p.faces[0]["face_index"]["axis"] = 99
    assert check(ctx, p) == O.INVALID_ARGUMENT
    p = L.fixture()
    p.faces, p.predicted = p.faces[:5].copy(), p.predicted[:5].copy()
    assert check(ctx, p) == O.FAILED_PRECONDITION
    for h in (2.0 ** -600, 2.0 ** 400):
        empty = O.LinkProblem(O.geometry((0.0, 0.0, 0.0), h), np.zeros(0, O.INT3), np.zeros(0, O.CENTER_DISTANCE),
                              np.zeros(0, O.FROZEN_CELL), np.zeros(0, O.FROZEN_FACE))
        assert check(ctx, empty) == O.OUT_OF_RANGE


@pytest.mark.parametrize("phi_l,phi_a,h", [(-0.5, 0.5, 1.0), (-0.2, 0.6, 0.25), (-1e-3, 0.9, 0.5), (-0.7, 1e-300, 2.0),
                                           (-1e200, 1e200, 1.0), (-3e-200, 1e-200, 0.125)])
def test_free_surface_links(ctx, phi_l, phi_a, h):
    p = L.make_problem([(1, 1, 1)], centers={(1, 1, 1): phi_l, (2, 1, 1): phi_a}, h=h, ambient=2.5)
    assert check(ctx, p) == O.OK


def test_lift_interfaces_sealed_and_rejections(ctx):
    p = L.make_problem([(1, 1, 1)], curvature=-3.0, sigma=2.0, ambient=10.0)
    air = np.zeros(1, O.AIR_PRESSURE)
    air[0]["cell_index"], air[0]["gas_gauge_pressure"] = (1, 1, 2), 4.0
    p.air = air
    assert check(ctx, p) == O.OK
    for kappa, sigma in ((1e300, 1e10), (1e-300, 1e-100), (1e-300, 1e-10)):
        assert check(ctx, L.make_problem([(1, 1, 1)], curvature=kappa, sigma=sigma)) == O.OUT_OF_RANGE
    fields = np.zeros(2, O.CELL_FIELDS)
    fields["bulk_modulus"] = np.inf
    fields["density"] = (800.0, 1200.0)
    fields["has_density"] = 1
    p = L.make_problem([(1, 1, 1), (2, 1, 1)], fields=fields)
    assert check(ctx, p) == O.OK
    i = [k for k, f in enumerate(p.faces) if tuple(f["face_index"]["lattice"]) == (2, 1, 1) and
         f["face_index"]["axis"] == 0][0]
    itf = np.zeros(1, O.INTERNAL_INTERFACE)
    itf[0] = (i, 0.25, 0.1, 0.3, 0.03)
    p.interfaces = itf
    assert check(ctx, p) == O.OK
    bound = math.sqrt(2000.0 / (4 * math.pi * 0.1))
    p.step = bound * (1 + 1e-12)
    assert check(ctx, p) == O.ABORTED
    p = L.make_problem([(1, 1, 1)], open_area={(0, (1, 1, 1)): 0.0})
    k = [j for j, f in enumerate(p.faces) if f["open_area"] == 0.0][0]
    p.faces[k]["has_prescribed_wall_flux"], p.faces[k]["prescribed_wall_flux"] = 1, -0.75
    p.faces[k]["wall_normal_velocity"] = -0.75
    p.faces[k]["is_low_x_piston_face"] = 1
    wet = np.zeros(len(p.faces), O.OPTIONAL_FLUX)
    wet[k] = (-0.5, 1)
    p.wetted = wet
    assert check(ctx, p) == O.OK
    wet[k] = (0.5, 1)
    assert check(ctx, p) == O.INVALID_ARGUMENT
    p = L.make_problem([(1, 1, 1)])
    p.predicted[3] = math.nan
    assert check(ctx, p) == O.INVALID_ARGUMENT
    p = L.make_problem([(1, 1, 1)], centers={(1, 1, 2): -0.1})
    assert check(ctx, p) == O.FAILED_PRECONDITION
    p = L.make_problem([(1, 1, 1)])
    p.centers = p.centers[p.centers["cell_index"]["z"] != 2].copy()
    assert check(ctx, p) == O.FAILED_PRECONDITION


def sphere_problem(n=48, radius=0.42, seed=4):
    """Liquid cells of a sphere in an n^3 grid (h = 1/n) with analytic centre distances."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    c = np.arange(n + 2) - 1
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    phi = np.sqrt(((x + 0.5) * h - 0.5) ** 2 + ((y + 0.5) * h - 0.5) ** 2 + ((z + 0.5) * h - 0.5) ** 2) - radius
    liquid = [tuple(v) for v in np.stack([x[phi < 0], y[phi < 0], z[phi < 0]], axis=1)]
    centers = {tuple(v): float(d) for v, d in zip(np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1), phi.ravel())}
    p = L.make_problem(liquid, centers=centers, h=h, sigma=0.07, ambient=101.0, curvature=-2.0 / radius)
    p.fields["density"] = rng.uniform(900.0, 1100.0, len(p.fields))
    p.fields["has_density"] = rng.random(len(p.fields)) < 0.5
    p.fields["previous_pressure"] = rng.normal(0, 10, len(p.fields))
    p.fields["target_divergence"] = rng.normal(0, 1, len(p.fields))
    p.predicted[:] = rng.normal(0, 1, len(p.predicted))
    # interface crossings on internal faces whose two cells carry densities
    cell_index = {tuple(int(v) for v in c["cell_index"]): k for k, c in enumerate(p.cells)}
    chosen = []
    for j, f in enumerate(p.faces):
        a, l = int(f["face_index"]["axis"]), tuple(int(v) for v in f["face_index"]["lattice"])
        t = list(l)
        t[a] -= 1
        hk, tk = cell_index.get(l), cell_index.get(tuple(t))
        if hk is not None and tk is not None and p.fields["has_density"][hk] and p.fields["has_density"][tk] and \
                rng.random() < 0.05:
            chosen.append(j)
    itf = np.zeros(len(chosen), O.INTERNAL_INTERFACE)
    itf["face_ordinal"] = chosen
    itf["tail_fraction"] = rng.uniform(0, 1, len(chosen))
    itf["surface_tension"] = 1e-4
    itf["curvature_tail_to_head"] = rng.normal(0, 5, len(chosen))
    itf["normal_viscous_jump_tail_to_head"] = rng.normal(0, 1, len(chosen))
    p.interfaces = itf
    air_cells = sorted([k for k, d in centers.items() if d > 0 and rng.random() < 0.1], key=lambda c: (c[2], c[1], c[0]))
    air = np.zeros(len(air_cells), O.AIR_PRESSURE)
    air["cell_index"] = air_cells
    air["gas_gauge_pressure"] = rng.normal(100.0, 5.0, len(air_cells))
    p.air = air
    p.step = 1e-4
    return p


def test_sphere(ctx):
    p = sphere_problem()
    assert check(ctx, p) == O.OK
    print("links", len(p.cells), "cells", len(p.faces), "faces", ctx.last.kernel_ms, "ms", flush=True)
