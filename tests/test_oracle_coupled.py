"""The oracle's coupled viscous pressure projection (flo_project_viscous, SolveCoupled), pinned without
the reference.

* The reference's known answers (native/tests/test_mac_pressure_3d.cc): one compliant cell between two
  Dirichlet faces with unit viscosity gives p = 4, face velocities 1 and 3 (:96-122); a 64-cell column
  whose predicted velocities are the differences of an analytic pressure recovers that pressure within
  1e-3 and zero velocities within 2e-5 after more than 32 GCR iterations, across the restart (:505-550).
* Zero viscosity reduces the coupled system to the inviscid projection: pressures and velocities equal
  the oracle's FP64 MGPCG solve of the same links to the stopping gates.
* The accepted state balances momentum: (M + dt E^T W E) u = M u* + dt A (grad p + lift) on every open
  face, recomputed here from the rows with numpy, and the cell constraint defects are within 2e-5.
* Statuses: the iteration budget outside [1, 256] and a strain term on a face outside the links are
  InvalidArgument.
"""
import math

import numpy as np
import pytest

import cases as K
import pyoracle as O


def column_problem(n, face_fn, h=1.0):
    """Cells (0..n-1, 0, 0) of volume h^3; face records from face_fn(axis, lattice) in slot order."""
    idx = [(x, 0, 0) for x in range(n)]
    cells = np.zeros(n, O.PRESSURE_CELL)
    for i, c in enumerate(idx):
        cells[i]["cell_index"] = c
        cells[i]["fluid_volume"] = h ** 3
        cells[i]["bulk_modulus"] = math.inf
    keys = set()
    for (x, y, z) in idx:
        for a in range(3):
            for side in range(2):
                lat = [x, y, z]
                lat[a] += side
                keys.add((a,) + tuple(lat))
    blocks = sorted({(K.owner(x), K.owner(y), K.owner(z)) for (_, x, y, z) in keys}, key=lambda b: b[::-1])
    recs = []
    for (a, x, y, z) in keys:
        f = np.zeros(1, O.PRESSURE_FACE)[0]
        f["axis"], f["lattice"] = a, (x, y, z)
        face_fn(f, a, (x, y, z))
        recs.append(f)
    faces = np.array(recs, dtype=O.PRESSURE_FACE)
    geom = O.geometry((0.0, 0.0, 0.0), h)
    sb = K.sorted_blocks(geom, blocks)
    slots = [O.face_slot(sb, int(f["axis"]), tuple(int(v) for v in f["lattice"]))[1] for f in faces]
    return geom, sb, cells, faces[np.argsort(np.array(slots))]


def full_strain(geom, sb, cells, mu):
    vc = O.viscosity_cells([tuple(int(v) for v in c["cell_index"].tolist()) for c in cells], [mu] * len(cells))
    s, rows = O.assemble_full_cell_strain(geom, sb, vc)
    assert s == O.OK
    return rows


def test_reference_compliant_cell():
    def face(f, a, lat):
        if a == 0:
            f["open_area"], f["inverse_density"], f["pressure_gradient_distance"] = 1.0, 1.0, 1.0
            f["link_kind"] = 2
            f["predicted_normal_velocity"], f["lift"] = (-1.0, 2.0) if lat[0] == 0 else (8.0, 5.0)
    geom, sb, cells, faces = column_problem(1, face)
    cells[0]["target_divergence"], cells[0]["bulk_modulus"], cells[0]["previous_pressure"] = 3.0, 2.0, 2.0
    rows = full_strain(geom, sb, cells, 1.0)
    s, out = O.project_viscous(geom, sb, 1.0, cells, faces, rows)
    assert s == O.OK
    assert abs(out["pressures"][0] - 4.0) <= 1e-9
    xs = [i for i, f in enumerate(faces) if f["axis"] == 0]
    lo, hi = sorted(xs, key=lambda i: int(faces[i]["lattice"]["x"]))
    assert abs(out["velocities"][lo] - 1.0) <= 1e-9 and abs(out["velocities"][hi] - 3.0) <= 1e-9
    assert abs(out["defects"][0]) <= 1e-9
    assert out["diagnostics"]["max_momentum_residual"] < 1e-9


def reference_pressure(i, n=64):
    if i < 0 or i >= n:
        return 0.0
    x = float(i + 1)
    return 0.1 * math.sin(0.32 * x) + 0.05 * math.cos(1.73 * x) + 0.01 * (i % 5)


def test_reference_analytic_column_across_restart():
    n = 64

    def face(f, a, lat):
        if a != 0:
            return
        head = lat[0]
        f["open_area"], f["inverse_density"], f["pressure_gradient_distance"] = 1.0, 1.0, 1.0
        f["predicted_normal_velocity"] = reference_pressure(head) - reference_pressure(head - 1)
        f["link_kind"] = 2 if head in (0, n) else 1
    geom, sb, cells, faces = column_problem(n, face)
    rows = full_strain(geom, sb, cells, 0.2)
    s, out = O.project_viscous(geom, sb, 1.0, cells, faces, rows)
    assert s == O.OK
    assert out["diagnostics"]["pressure_iterations"] > 32
    want = np.array([reference_pressure(int(c["cell_index"]["x"])) for c in cells])
    assert np.max(np.abs(out["pressures"] - want)) <= 1e-3
    assert np.max(np.abs(out["velocities"])) <= 2e-5


def random_affine(seed, mu):
    geom, sb, cells, faces, _ = K.affine_problem(minimum=0, extents=(5, 4, 3), anchored=True)
    rng = np.random.default_rng(seed)
    open_ = faces["open_area"] > 0
    faces["predicted_normal_velocity"][open_] = rng.normal(0.0, 1.0, int(open_.sum()))
    rows = full_strain(geom, sb, cells, mu)
    return geom, sb, cells, faces, rows


def test_zero_viscosity_is_the_inviscid_projection():
    geom, sb, cells, faces, rows = random_affine(3, 0.0)
    s, out = O.project_viscous(geom, sb, 1.0, cells, faces, rows)
    assert s == O.OK
    s, ref = O.project(geom, sb, 1.0, cells, faces, O.pressure_options(relative_target=1e-13, levels=2, fp32=False))
    assert s == O.OK
    scale = np.max(np.abs(ref["pressures"]))
    assert np.max(np.abs(out["pressures"] - ref["pressures"])) <= 1e-5 * scale
    assert np.max(np.abs(out["velocities"] - ref["velocities"])) <= 1e-5 * np.max(np.abs(ref["velocities"]))


def momentum_defect(geom, sb, cells, faces, rows, dt, p, u):
    """max |(M + dt E^T W E) u - M u* - dt A (p_tail - p_head + lift)| over open faces, relative."""
    nf = len(faces)
    slot_of = {}
    for i, f in enumerate(faces):
        slot_of[O.face_slot(sb, int(f["axis"]), tuple(int(v) for v in f["lattice"]))[1]] = i
    cell_of = {tuple(int(v) for v in c["cell_index"].tolist()): i for i, c in enumerate(cells)}
    lhs = np.zeros(nf)
    rhs = np.zeros(nf)
    area = faces["open_area"]
    inertia = np.where(area > 0, dt * area / np.where(area > 0, dt * faces["inverse_density"] /
                                                      faces["pressure_gradient_distance"], 1.0), 0.0)
    for i, f in enumerate(faces):
        if area[i] == 0:
            continue
        lat = [int(f["lattice"][k]) for k in "xyz"]
        a = int(f["axis"])
        tail = list(lat)
        tail[a] -= 1
        pt = p[cell_of[tuple(tail)]] if tuple(tail) in cell_of else None
        ph = p[cell_of[tuple(lat)]] if tuple(lat) in cell_of else None
        lift = f["lift"] if f["link_kind"] == 1 else 0.0
        if f["link_kind"] == 2:
            pt = f["lift"] if pt is None else pt
            ph = f["lift"] if ph is None else ph
        lhs[i] = inertia[i] * u[i]
        rhs[i] = inertia[i] * f["predicted_normal_velocity"] + dt * area[i] * (pt - ph + lift)
    for r in rows:
        terms = [(slot_of[int(t["face_slot"])], float(t["coefficient"])) for t in r["terms"][: int(r["term_count"])]]
        strain = sum(c * u[j] for j, c in terms if area[j] > 0)
        for j, c in terms:
            if area[j] > 0:
                lhs[j] += dt * r["quadrature_weight"] * c * strain
    return np.max(np.abs(lhs - rhs)) / max(np.max(np.abs(rhs)), 1e-300)


@pytest.mark.parametrize("mu", [0.05, 1.0])
def test_accepted_state_balances_momentum_and_constraint(mu):
    geom, sb, cells, faces, rows = random_affine(7, mu)
    s, out = O.project_viscous(geom, sb, 1.0, cells, faces, rows)
    assert s == O.OK
    assert momentum_defect(geom, sb, cells, faces, rows, 1.0, out["pressures"], out["velocities"]) <= 1e-8
    assert np.max(np.abs(out["defects"])) <= 2e-5
    # viscosity changes the answer: the inviscid solve does not balance these rows
    s, ref = O.project(geom, sb, 1.0, cells, faces, O.pressure_options(relative_target=1e-13, levels=2, fp32=False))
    assert momentum_defect(geom, sb, cells, faces, rows, 1.0, ref["pressures"], ref["velocities"]) > 1e-3


def test_statuses():
    geom, sb, cells, faces, rows = random_affine(1, 1.0)
    assert O.project_viscous(geom, sb, 1.0, cells, faces, rows, max_iterations=0)[0] == O.INVALID_ARGUMENT
    assert O.project_viscous(geom, sb, 1.0, cells, faces, rows, max_iterations=257)[0] == O.INVALID_ARGUMENT
    bad = rows.copy()
    bad[0]["terms"][0]["face_slot"] = 10 ** 9
    assert O.project_viscous(geom, sb, 1.0, cells, faces, bad)[0] == O.INVALID_ARGUMENT
