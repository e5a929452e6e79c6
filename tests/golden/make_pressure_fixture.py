"""Write tests/golden/pressure_cases_3d.json: seven manufactured MAC pressure systems.

The case definitions are the reference's known-answer set (native/tests/test_mac_pressure_3d.cc:128;
data from reference/src/fluidity_reference/pressure_3d_fixtures.py:197-249): a unit pressure jump
across an x, y or z face, a low and a high Dirichlet face, two disconnected Neumann cells and one
isolated compressible cell, plus their union translated block by block along x. Expected values
are solved here independently: the integrated operator A = R diag(k) R^T + diag(V/(K dt)) and
right-hand side are assembled densely from the face list, the first cell of each unanchored
component is pinned, and the reduced system is solved with numpy.linalg.solve. Corrected
velocities and physical defects follow u = u* + (dt beta/d)(p_tail - p_head + jump) and
defect = (sum of outward face fluxes - V s + C (p - p_prev)) / V.
Run: python tests/golden/make_pressure_fixture.py
"""
import json
import math
import os

import numpy as np

SEALED, INTERNAL, DIRICHLET = 0, 1, 2


def cell(idx, V=1.0, s=0.0, K=math.inf, prev=0.0):
    return dict(index=tuple(idx), V=V, s=s, K=K, prev=prev)


def face(axis, lat, kind=SEALED, area=0.0, beta=0.0, d=0.0, u=0.0, uw=0.0, lift=0.0):
    return dict(axis=axis, lat=tuple(lat), kind=kind, area=area, beta=beta, d=d, u=u, uw=uw, lift=lift)


def shift(idx, axis, k):
    return tuple(v + k if a == axis else v for a, v in enumerate(idx))


def slot_key(axis, lat):
    x, y, z = lat
    return (z // 4, y // 4, x // 4, axis, z % 4, y % 4, x % 4)


def make_case(cells, special=(), dt=1.0):
    incident = set()
    for c in cells:
        for a in range(3):
            incident.add((a, c["index"]))
            incident.add((a, shift(c["index"], a, 1)))
    over = {(f["axis"], f["lat"]): f for f in special}
    assert set(over) <= incident
    faces = [over.get(k, face(k[0], k[1])) for k in sorted(incident, key=lambda k: slot_key(*k))]
    return dict(cells=sorted(cells, key=lambda c: c["index"][::-1]), faces=faces, dt=dt)


def solve_case(case):
    cells, faces, dt = case["cells"], case["faces"], case["dt"]
    nc, nf = len(cells), len(faces)
    ordinal = {c["index"]: i for i, c in enumerate(cells)}
    R = np.zeros((nc, nf))
    k = np.zeros(nf)
    scale = np.zeros(nf)
    for j, f in enumerate(faces):
        head = f["lat"]
        tail = shift(head, f["axis"], -1)
        if tail in ordinal:
            R[ordinal[tail], j] = 1.0
        if head in ordinal:
            R[ordinal[head], j] = -1.0
        if f["area"] > 0:
            scale[j] = dt * f["beta"] / f["d"]
            k[j] = f["area"] * scale[j]
    comp = np.array([c["V"] / (c["K"] * dt) for c in cells])
    A = (R * k) @ R.T + np.diag(comp)
    b = np.array([c["V"] * c["s"] for c in cells]) + comp * np.array([c["prev"] for c in cells])
    for j, f in enumerate(faces):
        flux = f["area"] * f["u"] + (1.0 - f["area"]) * f["uw"]
        b -= R[:, j] * flux
        if f["kind"] == INTERNAL:
            b -= R[:, j] * k[j] * f["lift"]
        elif f["kind"] == DIRICHLET:
            b += np.abs(R[:, j]) * k[j] * f["lift"]
    # Components over positive-conductance internal faces; pin the lowest unanchored cell.
    parent = list(range(nc))

    def root(i):
        while parent[i] != i:
            i = parent[i]
        return i

    for j, f in enumerate(faces):
        rows = np.flatnonzero(R[:, j])
        if len(rows) == 2 and k[j] > 0:
            a, c = root(rows[0]), root(rows[1])
            parent[max(a, c)] = min(a, c)
    anchored = {root(i) for i in range(nc) if comp[i] > 0}
    for j, f in enumerate(faces):
        if f["kind"] == DIRICHLET:
            anchored |= {root(i) for i in np.flatnonzero(R[:, j])}
    gauges = sorted({root(i) for i in range(nc)} - anchored)
    free = [i for i in range(nc) if i not in gauges]
    p = np.zeros(nc)
    if free:
        p[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
    vel = []
    flux = np.zeros(nf)
    for j, f in enumerate(faces):
        u = f["u"]
        if f["area"] > 0:
            head = f["lat"]
            tail = shift(head, f["axis"], -1)
            lift = f["lift"] if f["kind"] == DIRICHLET else 0.0
            jump = f["lift"] if f["kind"] == INTERNAL else 0.0
            pt = p[ordinal[tail]] if tail in ordinal else lift
            ph = p[ordinal[head]] if head in ordinal else lift
            u += scale[j] * (pt - ph + jump)
        vel.append(u)
        flux[j] = f["area"] * u + (1.0 - f["area"]) * f["uw"]
    defects = []
    for i, c in enumerate(cells):
        out = sum(R[i, j] * flux[j] for j in range(nf))
        defects.append((out - c["V"] * c["s"] + comp[i] * (p[i] - c["prev"])) / c["V"])
    return dict(pressures=p.tolist(), rhs=b.tolist(), residuals=(b - A @ p).tolist(), velocities=vel,
                defects=defects, gauges=gauges)


def cases():
    origin = (0, 0, 0)
    out = {}
    for axis, name in enumerate("xyz"):
        jump = face(axis, origin, INTERNAL, 1.0, 1.0, 1.0, lift=1.0)
        out[f"jump_across_{name}_block"] = make_case([cell(shift(origin, axis, -1)), cell(origin)], [jump])
    bnd = dict(kind=DIRICHLET, area=0.5, beta=1.0, d=0.5, u=1.0, uw=2.0, lift=3.0)
    out["low_dirichlet"] = make_case([cell(origin)], [face(0, origin, **bnd)])
    out["high_dirichlet"] = make_case([cell(origin)], [face(0, (1, 0, 0), **bnd)])
    sep = face(0, origin, INTERNAL, 0.0, 1.0, 1.0, u=0.75, uw=0.25)
    out["disconnected_neumann"] = make_case([cell((-1, 0, 0), s=0.25), cell(origin, s=-0.25)], [sep])
    out["isolated_compliance"] = make_case([cell(origin, s=0.5, K=2.0, prev=4.0)])
    mixed_cells, mixed_faces = [], []
    for bx, c in enumerate(list(out.values())):
        sh = 4 * bx
        mixed_cells += [dict(cc, index=shift(cc["index"], 0, sh)) for cc in c["cells"]]
        mixed_faces += [dict(ff, lat=shift(ff["lat"], 0, sh)) for ff in c["faces"]]
    out["mixed"] = make_case(mixed_cells, mixed_faces)
    return out


def main():
    result = {}
    for name, c in cases().items():
        sol = solve_case(c)
        result[name] = dict(
            dt=c["dt"],
            cells=[dict(index=list(x["index"]), V=x["V"], s=x["s"], K=(None if math.isinf(x["K"]) else x["K"]),
                        prev=x["prev"]) for x in c["cells"]],
            faces=[dict(f, lat=list(f["lat"])) for f in c["faces"]],
            expected=sol)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pressure_cases_3d.json")
    with open(path, "w") as f:
        json.dump(result, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
