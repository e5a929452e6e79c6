"""Oracle: typed pressure links of a frozen view (flo_build_pressure_links).

Pinned by the reference's workspace known answers (native/tests/test_liquid_boundary_workspace_3d.cc:
the unit-cube fixture's six Dirichlet links at distance 0.5 with the ambient pressure, unrepresentable
and invalid densities, a missing incident face, an extra sealed face at the minimum signed index,
unrepresentable cell widths), by the closed forms of the free-surface link (theta = -phi_l / (phi_a -
phi_l), d = clamp(theta, 0.05, 1) h, p_ext = p_gas + sigma kappa), of the internal link (midpoint
density, interface lerp density and fma jump, the capillary bound sqrt((rho_t + rho_h) h^3 / (4 pi
sigma))), and by every rejection.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import cases as K
import pyoracle as O

UNIT = O.geometry((0.0, 0.0, 0.0), 1.0)


def owner(v):
    return tuple(int(math.floor(c / 4)) for c in v)


def make_problem(liquid, centers=None, h=1.0, open_area=None, curvature=0.0, **kw):
    """A frozen view over the liquid cells: their incident faces (open, or open_area[(axis, lattice)]),
    block keys covering those faces, and center samples of every liquid cell and neighbour."""
    geom = O.geometry((0.0, 0.0, 0.0), h)
    liquid = sorted(set(map(tuple, liquid)), key=lambda c: (c[2], c[1], c[0]))
    faces = set()
    for c in liquid:
        for a in range(3):
            for side in (0, 1):
                l = list(c)
                l[a] += side
                faces.add((a, tuple(l)))
    blocks = K.sorted_blocks(geom, sorted({owner(l) for _, l in faces}, key=lambda b: (b[2], b[1], b[0])))
    recs = []
    for a, l in faces:
        s, slot = O.face_slot(blocks, a, l)
        assert s == O.OK
        recs.append((slot, a, l))
    recs.sort()
    fr = np.zeros(len(recs), O.FROZEN_FACE)
    for i, (slot, a, l) in enumerate(recs):
        fr[i]["face_index"]["axis"] = a
        fr[i]["face_index"]["lattice"] = l
        fr[i]["face_slot"] = slot
        area = h * h if open_area is None else open_area.get((a, l), h * h)
        fr[i]["open_area"] = area
        fr[i]["blocked_area"] = h * h - area
    cells = np.zeros(len(liquid), O.FROZEN_CELL)
    for i, c in enumerate(liquid):
        cells[i]["cell_index"] = c
        cells[i]["fluid_volume"] = h ** 3
        cells[i]["center_distance"] = -0.5 * h
        cells[i]["curvature"] = curvature
    pts = {c: -0.5 * h for c in liquid}
    for c in liquid:
        for a in range(3):
            for d in (-1, 1):
                n = list(c)
                n[a] += d
                pts.setdefault(tuple(n), 0.5 * h)
    if centers:
        pts.update(centers)
        for i, c in enumerate(liquid):
            if c in centers:
                cells[i]["center_distance"] = centers[c]
    keys = sorted(pts, key=lambda c: (c[2], c[1], c[0]))
    cd = np.zeros(len(keys), O.CENTER_DISTANCE)
    for i, c in enumerate(keys):
        cd[i]["cell_index"] = c
        cd[i]["signed_distance"] = pts[c]
    return O.LinkProblem(geom, blocks, cd, cells, fr, **kw)


def fixture(**kw):
    fields = np.zeros(1, O.CELL_FIELDS)
    fields["bulk_modulus"] = np.inf
    fields["previous_pressure"] = 7
    return make_problem([(1, 1, 1)], fields=fields, **kw)


def test_reference_unit_cube_links():
    # FrozenGeometryAndTypedLinksMatchOwningAndUnitCube (test_liquid_boundary_workspace_3d.cc:75-101)
    p = fixture(ambient=9.0)
    assert list(p.faces["face_slot"]) == [21, 22, 85, 89, 149, 165]
    s, cells, faces, lower = O.build_pressure_links(p)
    assert s == O.OK and lower == 0
    assert cells[0]["previous_pressure"] == 7 and cells[0]["fluid_volume"] == 1.0
    assert (faces["pressure_gradient_distance"] == 0.5).all()
    assert (faces["link_kind"] == 2).all() and (faces["lift"] == 9.0).all()
    assert (faces["inverse_density"] == 0.001).all() and (faces["open_area"] == 1.0).all()


def test_reference_view_rejections():
    # MalformedFrozenView..., UnrepresentableDensity... (:531-583)
    for density, code in [(-1.0, O.INVALID_ARGUMENT), (5e-324, O.OUT_OF_RANGE), (1.7976931348623157e308, O.OUT_OF_RANGE)]:
        assert O.build_pressure_links(fixture(density=density))[0] == code
    p = fixture()
    p.faces[0]["face_index"]["axis"] = 99
    assert O.build_pressure_links(p)[0] == O.INVALID_ARGUMENT
    # MissingIncidentFacePreservesPublishedPressureLinks (:601-620)
    p = fixture()
    p.faces, p.predicted = p.faces[:5].copy(), p.predicted[:5].copy()
    assert O.build_pressure_links(p)[0] == O.FAILED_PRECONDITION
    # UnrepresentablePhysicalAreaOrVolumeRejectsEmptyFrozenView (:584-600)
    for h in (2.0 ** -600, 2.0 ** -520, 2.0 ** -400, 2.0 ** -350, 2.0 ** 400, 2.0 ** 600):
        empty = O.LinkProblem(O.geometry((0.0, 0.0, 0.0), h), np.zeros(0, O.INT3), np.zeros(0, O.CENTER_DISTANCE),
                              np.zeros(0, O.FROZEN_CELL), np.zeros(0, O.FROZEN_FACE))
        assert O.build_pressure_links(empty)[0] == O.OUT_OF_RANGE
    # ExtraSealedFaceAtMinimumSignedIndexIsNotAnIncidentUnion (:621-640)
    lo = -2 ** 31
    blocks = K.sorted_blocks(UNIT, [(lo // 4, 0, 0)])
    s, slot = O.face_slot(blocks, 0, (lo, 0, 0))
    assert s == O.OK
    f = np.zeros(1, O.FROZEN_FACE)
    f[0]["face_index"]["axis"], f[0]["face_index"]["lattice"], f[0]["face_slot"] = 0, (lo, 0, 0), slot
    f[0]["blocked_area"] = 1.0
    extra = O.LinkProblem(UNIT, blocks, np.zeros(0, O.CENTER_DISTANCE), np.zeros(0, O.FROZEN_CELL), f)
    assert O.build_pressure_links(extra)[0] == O.FAILED_PRECONDITION


@pytest.mark.parametrize("phi_l,phi_a,h", [(-0.5, 0.5, 1.0), (-0.2, 0.6, 0.25), (-1e-3, 0.9, 0.5), (-0.7, 1e-300, 2.0),
                                           (-1e200, 1e200, 1.0), (-3e-200, 1e-200, 0.125)])
def test_free_surface_link_closed_form(phi_l, phi_a, h):
    air = (2, 1, 1)
    p = make_problem([(1, 1, 1)], centers={(1, 1, 1): phi_l, air: phi_a}, h=h, ambient=2.5)
    s, _, faces, lower = O.build_pressure_links(p)
    assert s == O.OK
    m = max(-phi_l, phi_a)
    theta = (-phi_l / m) / ((-phi_l / m) + phi_a / m)
    assert abs(theta - (-phi_l) / (phi_a - phi_l)) <= 4e-16 * theta
    i = [k for k, f in enumerate(p.faces) if tuple(f["face_index"]["lattice"]) == (2, 1, 1) and
         f["face_index"]["axis"] == 0][0]
    assert faces[i]["pressure_gradient_distance"] == min(max(theta, 0.05), 1.0) * h
    assert faces[i]["link_kind"] == 2 and faces[i]["lift"] == 2.5
    # the other five faces see air samples at +h/2
    mo = max(-phi_l, 0.5 * h)
    theta_other = (-phi_l / mo) / ((-phi_l / mo) + 0.5 * h / mo)
    assert lower == int(theta < 0.05) + 5 * int(theta_other < 0.05)


def test_capillary_lift_gas_pressure_and_range():
    p = make_problem([(1, 1, 1)], curvature=-3.0, sigma=2.0, ambient=10.0)
    air = np.zeros(1, O.AIR_PRESSURE)
    air[0]["cell_index"], air[0]["gas_gauge_pressure"] = (1, 1, 2), 4.0
    p.air = air
    s, _, faces, _ = O.build_pressure_links(p)
    assert s == O.OK
    for f in faces:
        # the face at z = 2 borders the air cell with its own gas pressure
        want = (4.0 if tuple(f["lattice"]) == (1, 1, 2) and f["axis"] == 2 else 10.0) + 2.0 * -3.0
        assert f["lift"] == want
    assert O.build_pressure_links(make_problem([(1, 1, 1)], curvature=1e300, sigma=1e10))[0] == O.OUT_OF_RANGE
    assert O.build_pressure_links(make_problem([(1, 1, 1)], curvature=1e-300, sigma=1e-100))[0] == O.OUT_OF_RANGE
    assert O.build_pressure_links(make_problem([(1, 1, 1)], curvature=1e-300, sigma=1e-10))[0] == O.OUT_OF_RANGE


def test_internal_links_and_densities():
    fields = np.zeros(2, O.CELL_FIELDS)
    fields["bulk_modulus"] = np.inf
    fields["density"] = (800.0, 1200.0)
    fields["has_density"] = 1
    p = make_problem([(1, 1, 1), (2, 1, 1)], fields=fields)
    s, cells, faces, _ = O.build_pressure_links(p)
    assert s == O.OK
    for f in faces:
        lat, a = tuple(f["lattice"]), int(f["axis"])
        if (a, lat) == (0, (2, 1, 1)):
            assert f["link_kind"] == 1 and f["lift"] == 0.0 and f["inverse_density"] == 1.0 / 1000.0
        elif f["link_kind"] == 2:
            head_in = lat in [(1, 1, 1), (2, 1, 1)]
            tail = list(lat)
            tail[a] -= 1
            liquid = lat if head_in else tuple(tail)
            assert f["inverse_density"] == 1.0 / (800.0 if liquid == (1, 1, 1) else 1200.0)
    # an interface crossing on the shared face
    i = [k for k, f in enumerate(p.faces) if tuple(f["face_index"]["lattice"]) == (2, 1, 1) and
         f["face_index"]["axis"] == 0][0]
    itf = np.zeros(1, O.INTERNAL_INTERFACE)
    sigma, kappa, viscous = 0.1, 0.3, 0.03  # fma(-sigma, kappa, viscous) differs from the rounded product sum
    itf[0] = (i, 0.25, sigma, kappa, viscous)
    p.interfaces = itf
    s, _, faces, _ = O.build_pressure_links(p)
    assert s == O.OK
    head, tail = 1200.0, 800.0
    assert faces[i]["inverse_density"] == 1.0 / (head + 0.25 * (tail - head))
    exact = float(Fraction(-sigma) * Fraction(kappa) + Fraction(viscous))
    assert exact != -sigma * kappa + viscous
    assert faces[i]["lift"] == exact and faces[i]["link_kind"] == 1
    bound = math.sqrt((tail + head) * 1.0 / (4 * math.pi * sigma))
    p.step = bound * (1 + 1e-12)
    assert O.build_pressure_links(p)[0] == O.ABORTED
    p.step = bound * (1 - 1e-12)
    assert O.build_pressure_links(p)[0] == O.OK
    p.fields["has_density"][1] = 0
    assert O.build_pressure_links(p)[0] == O.INVALID_ARGUMENT


def test_sealed_faces_wall_flux_and_piston():
    closed = {(0, (1, 1, 1)): 0.0}
    p = make_problem([(1, 1, 1)], open_area=closed)
    k = [j for j, f in enumerate(p.faces) if f["open_area"] == 0.0][0]
    p.faces[k]["has_prescribed_wall_flux"], p.faces[k]["prescribed_wall_flux"] = 1, -0.75
    p.faces[k]["wall_normal_velocity"] = -0.75
    s, _, faces, _ = O.build_pressure_links(p)
    assert s == O.OK and faces[k]["link_kind"] == 0 and faces[k]["has_prescribed_wall_flux"] == 1
    assert faces[k]["prescribed_wall_flux"] == -0.75 and faces[k]["wall_normal_velocity"] == -0.75
    # a wetted piston flux on the low-x receiver face replaces the wall flux
    p.faces[k]["is_low_x_piston_face"] = 1
    wet = np.zeros(len(p.faces), O.OPTIONAL_FLUX)
    wet[k] = (-0.5, 1)
    p.wetted = wet
    s, _, faces, _ = O.build_pressure_links(p)
    assert s == O.OK and faces[k]["prescribed_wall_flux"] == -0.5
    wet[k] = (0.5, 1)  # against the wall velocity
    assert O.build_pressure_links(p)[0] == O.INVALID_ARGUMENT
    wet[k] = (0.0, 0)  # a receiver without a flux
    assert O.build_pressure_links(p)[0] == O.INVALID_ARGUMENT


def test_field_and_sample_rejections():
    p = make_problem([(1, 1, 1)])
    p.predicted = p.predicted[:5].copy()
    assert O.build_pressure_links(p)[0] == O.INVALID_ARGUMENT
    p = make_problem([(1, 1, 1)])
    p.predicted[3] = math.nan
    assert O.build_pressure_links(p)[0] == O.INVALID_ARGUMENT
    p = make_problem([(1, 1, 1)])
    p.fields["bulk_modulus"] = 0.0
    assert O.build_pressure_links(p)[0] == O.INVALID_ARGUMENT
    p = make_problem([(1, 1, 1)], centers={(1, 1, 2): -0.1})  # a liquid-side air sample
    assert O.build_pressure_links(p)[0] == O.FAILED_PRECONDITION
    p = make_problem([(1, 1, 1)])
    p.centers = p.centers[p.centers["cell_index"]["z"] != 2].copy()  # missing air sample
    assert O.build_pressure_links(p)[0] == O.FAILED_PRECONDITION
    p = make_problem([(1, 1, 1)])
    air = np.zeros(1, O.AIR_PRESSURE)
    air[0]["cell_index"] = (1, 1, 1)  # a liquid cell, not an air center
    p.air = air
    assert O.build_pressure_links(p)[0] == O.INVALID_ARGUMENT
    p = make_problem([(1, 1, 1)])
    p.cells["center_distance"] = 0.0
    assert O.build_pressure_links(p)[0] == O.INVALID_ARGUMENT
    p = make_problem([(1, 1, 1)])
    p.faces["blocked_area"][2] = 0.5
    assert O.build_pressure_links(p)[0] == O.INVALID_ARGUMENT
