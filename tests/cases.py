"""Shared test inputs: the reference's known-answer cases rebuilt as record arrays.

Sources: native/tests/test_transfers_3d.cc:22-85 (corner blocks, corner/tie particles),
native/tests/mac_viscous_traction_cases.h:33-91 (affine pressure problem),
native/tests/matched_error_pressure_3d.cc:29-35 (sealed cube p = 0.125 + 0.75 x).
"""
import json
import math
import os

import numpy as np

import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
CORNER_BLOCKS = [(0, -1, 0), (1, -1, 0), (0, 0, 0), (1, 0, 0), (0, -1, 1), (1, -1, 1), (0, 0, 1), (1, 0, 1)]
CORNER_POSITION = (3.75, -0.25, 3.75)
PARTICLE_VELOCITY = (1.0, -2.0, 4.0)
OLD_VELOCITY = (2.0, 3.0, -4.0)
AFFINE_ROWS = ((2.0, 4.0, -1.0), (-5.0, 1.0, 5.0), (7.0, -2.0, 3.0))
TIE_POSITION = (3.5, 0.0, 4.0)
TIE_FACE = (0, (3, -1, 3))


def sorted_blocks(geom, blocks):
    s, b = O.sort_blocks(geom, O.int3(blocks))
    assert s == O.OK
    return b


def particles(n):
    return np.zeros(n, O.PARTICLE)


def set_vec(rec, field, v):
    rec[field]["x"], rec[field]["y"], rec[field]["z"] = v


def set_mat(rec, field, rows):
    for name, r in zip(("x_row", "y_row", "z_row"), rows):
        rec[field][name]["x"], rec[field][name]["y"], rec[field][name]["z"] = r


def corner_particle(origin=(0.0, 0.0, 0.0), h=1.0):
    p = particles(1)
    pos = tuple(o + h * c for o, c in zip(origin, CORNER_POSITION))
    set_vec(p[0], "sorted_world_position", pos)
    set_vec(p[0], "trial_world_position", pos)
    p[0]["particle_mass"] = 2.0
    set_vec(p[0], "particle_velocity", PARTICLE_VELOCITY)
    set_mat(p[0], "particle_velocity_slopes", AFFINE_ROWS)
    return p


def tie_particle(mass, normal_velocity):
    p = particles(1)
    set_vec(p[0], "sorted_world_position", TIE_POSITION)
    set_vec(p[0], "trial_world_position", TIE_POSITION)
    p[0]["particle_mass"] = mass
    set_vec(p[0], "particle_velocity", (normal_velocity, 0.0, 0.0))
    return p


def vec(a):
    return np.array([a["x"], a["y"], a["z"]], dtype=np.float64)


def mat(a):
    return np.array([vec(a["x_row"]), vec(a["y_row"]), vec(a["z_row"])])


def sample_field(geom, blocks, field):
    """field(axis, anchor) at every owned face anchor, in slot order."""
    values = np.zeros(192 * len(blocks))
    h = float(geom["cell_width"][0])
    org = np.array([geom["world_origin"][c][0] for c in "xyz"])
    for bi, b in enumerate(blocks):
        for axis in range(3):
            for lz in range(4):
                for ly in range(4):
                    for lx in range(4):
                        lat = (4 * int(b["x"]) + lx, 4 * int(b["y"]) + ly, 4 * int(b["z"]) + lz)
                        s, anchor = O.face_world_anchor(geom, axis, lat)
                        assert s == O.OK
                        values[bi * 192 + axis * 64 + lz * 16 + ly * 4 + lx] = field(axis, np.array(anchor))
    return values


def load_transfer_fixture():
    with open(os.path.join(HERE, "golden", "corner_transfer_3d.json")) as f:
        return json.load(f)


def load_pressure_fixture():
    with open(os.path.join(HERE, "golden", "pressure_cases_3d.json")) as f:
        return json.load(f)


def pressure_case_records(case):
    cells = np.zeros(len(case["cells"]), O.PRESSURE_CELL)
    for i, c in enumerate(case["cells"]):
        cells[i]["cell_index"] = tuple(c["index"])
        cells[i]["fluid_volume"] = c["V"]
        cells[i]["target_divergence"] = c["s"]
        cells[i]["bulk_modulus"] = math.inf if c["K"] is None else c["K"]
        cells[i]["previous_pressure"] = c["prev"]
    faces = np.zeros(len(case["faces"]), O.PRESSURE_FACE)
    for i, f in enumerate(case["faces"]):
        faces[i]["axis"] = f["axis"]
        faces[i]["lattice"] = tuple(f["lat"])
        faces[i]["open_area"] = f["area"]
        faces[i]["inverse_density"] = f["beta"]
        faces[i]["pressure_gradient_distance"] = f["d"]
        faces[i]["predicted_normal_velocity"] = f["u"]
        faces[i]["wall_normal_velocity"] = f["uw"]
        faces[i]["link_kind"] = f["kind"]
        faces[i]["lift"] = f["lift"]
    blocks = sorted({(f["lat"][0] // 4, f["lat"][1] // 4, f["lat"][2] // 4) for f in case["faces"]})
    return cells, faces, blocks


def owner(v):
    return v // 4


def affine_problem(minimum=-3, extents=(4, 4, 4), anchored=True, h=1.0, dt=1.0):
    """Box of cells; interior faces carry u* = 2^axis; p = x + 2y + 4z zeroes the corrected field."""
    def pressure(c):
        return c[0] + 2.0 * c[1] + 4.0 * c[2]

    cells_idx = [(x, y, z) for z in range(minimum, minimum + extents[2]) for y in range(minimum, minimum + extents[1])
                 for x in range(minimum, minimum + extents[0])]
    face_keys = set()
    for (x, y, z) in cells_idx:
        for a in range(3):
            for side in range(2):
                face_keys.add((a, x + (side if a == 0 else 0), y + (side if a == 1 else 0), z + (side if a == 2 else 0)))
    blocks = sorted({(owner(x), owner(y), owner(z)) for (_, x, y, z) in face_keys}, key=lambda b: b[::-1])
    cells = np.zeros(len(cells_idx), O.PRESSURE_CELL)
    for i, c in enumerate(cells_idx):
        cells[i]["cell_index"] = c
        cells[i]["fluid_volume"] = h ** 3
        cells[i]["bulk_modulus"] = math.inf
    recs = []
    for (a, x, y, z) in face_keys:
        co = (x, y, z)
        low = co[a] == minimum
        high = co[a] == minimum + extents[a]
        boundary = low or high
        f = np.zeros(1, O.PRESSURE_FACE)[0]
        f["axis"] = a
        f["lattice"] = co
        f["open_area"] = 0.0 if (boundary and not anchored) else h * h
        f["inverse_density"] = 1.0
        f["pressure_gradient_distance"] = h
        f["predicted_normal_velocity"] = 0.0 if (boundary and not anchored) else float(1 << a) * dt / dt
        f["link_kind"] = 1
        if boundary and anchored:
            ext = list(co)
            if low:
                ext[a] -= 1
            f["link_kind"] = 2
            f["lift"] = pressure(ext)
        elif boundary:
            f["link_kind"] = 0
        recs.append(f)
    faces = np.array(recs, dtype=O.PRESSURE_FACE)
    geom = O.geometry((0.0, 0.0, 0.0), h)
    sb = sorted_blocks(geom, blocks)
    slots = []
    for f in faces:
        s, slot = O.face_slot(sb, int(f["axis"]), tuple(int(v) for v in f["lattice"]))
        assert s == O.OK
        slots.append(slot)
    faces = faces[np.argsort(np.array(slots))]
    return geom, sb, cells, faces, pressure


QUANTUM = 2.0 ** -40


def oracle_step_inputs(scene, threads=8):
    """Run the oracle through activation, quantized P2G, normalize and forces for a scene.

    Returns a dict with the geometry, blocks, particles, face quanta, u0, u*, and the pressure
    cell/face records built from u* (the inputs of the projection stage).
    """
    geom = scene.geometry_record().view(O.GEOMETRY)
    cells_idx = scene.liquid_cells()
    cells3 = np.zeros(len(cells_idx), O.INT3)
    cells3["x"], cells3["y"], cells3["z"] = cells_idx.T
    s, blocks = O.activate_blocks(geom, scene.positions, cells3, capacity=1 << 22)
    assert s == O.OK
    parts = scene.particles_record().view(O.PARTICLE)
    s, q = O.deposit_quanta(geom, blocks, parts, QUANTUM, QUANTUM, len(parts), max_capacity=1 << 40, threads=threads)
    assert s == O.OK
    s, mass, u0, validity = O.normalize(q, QUANTUM, QUANTUM, 0.0)
    assert s == O.OK
    dt = scene.dt
    s, ustar = O.apply_forces(blocks, u0, scene.gravity, dt)
    assert s == O.OK
    faces = scene.pressure_faces(blocks.view(np.dtype([("x", "i4"), ("y", "i4"), ("z", "i4")])), ustar).view(
        O.PRESSURE_FACE)
    cells = scene.pressure_cells().view(O.PRESSURE_CELL)
    return dict(geom=geom, blocks=blocks, particles=parts, quanta=q, mass=mass, u0=u0, validity=validity, ustar=ustar,
                dt=dt, cells=cells, faces=faces)


# Compatible (masked) deposit: native/tests/liquid_boundary_transfer_3d_cases.h:19-36 — one block at
# the origin, unit cells, one particle at (1.5, 1.5, 1.5) with mass 16, velocity (3, 4, 5), side 7,
# every slot eligible on side 7, quanta 2^-30 with declared capacity 1.
BOUNDARY_QUANTUM = 2.0 ** -30


def boundary_particle(position=(1.5, 1.5, 1.5), mass=16.0, velocity=(3.0, 4.0, 5.0), side=7, affine=True):
    p = np.zeros(1, O.BOUNDARY_PARTICLE)
    t = p[0]["transfer"]
    set_vec(t, "sorted_world_position", position)
    set_vec(t, "trial_world_position", position)
    t["particle_mass"] = mass
    set_vec(t, "particle_velocity", velocity)
    p[0]["transfer"] = t
    p[0]["side_tag"] = side
    p[0]["affine_enabled"] = 1 if affine else 0
    return p


def boundary_tags(block_count, side=7, eligible=True):
    tags = np.zeros(192 * block_count, O.FACE_TAG)
    tags["side_tag"] = side
    tags["structurally_eligible"] = 1 if eligible else 0
    return tags


def set_slopes(particles, i, rows):
    t = particles[i]["transfer"]
    set_mat(t, "particle_velocity_slopes", rows)
    particles[i]["transfer"] = t


def set_position(particles, i, pos):
    t = particles[i]["transfer"]
    set_vec(t, "sorted_world_position", pos)
    set_vec(t, "trial_world_position", pos)
    particles[i]["transfer"] = t


def masked_scene(seed, n, h=1.0, blocks_per_axis=3, sides=2, eligible=0.85, affine=0.8, trial=True):
    """Random compatible-deposit inputs inside a cube of blocks, every stencil owned.

    Positions keep every normal and transverse sample inside the allocated cube; trial positions move
    up to 0.45 h from the sorted ones. Slot and particle sides are drawn from `sides` tags, slots are
    eligible with probability `eligible`, affine rows are enabled with probability `affine`.
    """
    rng = np.random.default_rng(seed)
    geom = O.geometry((0.25 * h, -0.5 * h, 0.125 * h), h)
    cube = [(x, y, z) for z in range(blocks_per_axis) for y in range(blocks_per_axis) for x in range(blocks_per_axis)]
    blocks = sorted_blocks(geom, cube)
    cells = 4 * blocks_per_axis
    org = np.array([0.25 * h, -0.5 * h, 0.125 * h])
    lo, hi = 2.0, cells - 2.0
    sorted_pos = org + h * rng.uniform(lo, hi, (n, 3))
    trial_pos = sorted_pos + (h * rng.uniform(-0.45, 0.45, (n, 3)) if trial else 0.0)
    p = np.zeros(n, O.BOUNDARY_PARTICLE)
    t = p["transfer"]
    for a, c in enumerate("xyz"):
        t["sorted_world_position"][c] = sorted_pos[:, a]
        t["trial_world_position"][c] = trial_pos[:, a]
        t["particle_velocity"][c] = rng.normal(0.0, 1.0, n)
    t["particle_mass"] = rng.uniform(0.5, 2.0, n) * h ** 3
    for r in ("x_row", "y_row", "z_row"):
        for c in "xyz":
            t["particle_velocity_slopes"][r][c] = rng.normal(0.0, 0.1 / h, n)
    p["transfer"] = t
    p["side_tag"] = rng.integers(0, sides, n)
    p["affine_enabled"] = rng.random(n) < affine
    tags = np.zeros(192 * len(blocks), O.FACE_TAG)
    tags["side_tag"] = rng.integers(0, sides, len(tags))
    tags["structurally_eligible"] = rng.random(len(tags)) < eligible
    return geom, blocks, p, tags


class BoundaryFixture:
    """native/tests/liquid_boundary_transfer_3d_cases.h:23-50: one block at the origin, unit cells, the
    boundary particle, every slot eligible on side 7, old face velocities 2, new 3, all valid, no walls,
    old history at t = 0 and new at t = 1 (separate arrays)."""

    def __init__(self):
        self.geom = O.geometry((0.0, 0.0, 0.0), 1.0)
        self.blocks = sorted_blocks(self.geom, [(0, 0, 0)])
        self.particle = boundary_particle()
        nf = 192 * len(self.blocks)
        self.tags = boundary_tags(len(self.blocks))
        self.old_values = np.full(nf, 2.0)
        self.new_values = np.full(nf, 3.0)
        self.validity = np.ones(nf, np.uint8)
        self.old_walls = np.zeros(nf, O.WALL_SAMPLE)
        self.new_walls = np.zeros(nf, O.WALL_SAMPLE)

    def slot(self, axis, x, y, z):
        s, v = O.face_slot(self.blocks, axis, (x, y, z))
        assert s == O.OK
        return v

    def anchor(self, axis, x, y, z):
        s, a = O.face_world_anchor(self.geom, axis, (x, y, z))
        assert s == O.OK
        return np.array(a)

    def histories(self, old_validity=None, new_validity=None, old_wall_time=None):
        """(time, velocities, validity, tags, walls, wall time) of the old and the new history."""
        old = (0.0, self.old_values, self.validity if old_validity is None else old_validity, self.tags,
               self.old_walls, 0.0 if old_wall_time is None else old_wall_time)
        new = (1.0, self.new_values, self.validity if new_validity is None else new_validity, self.tags,
               self.new_walls, 1.0)
        return old, new

    def gather(self, **kw):
        old, new = self.histories(**kw)
        return O.gather_boundary(self.geom, self.blocks, self.particle, 0.0, 1.0,
                                 O.history(old[0], *old[1:5], wall_time=old[5]),
                                 O.history(new[0], *new[1:5], wall_time=new[5]))


def set_wall(walls, slot, side, normal, velocity):
    w = walls[slot]
    w["available"] = 1
    w["side_tag"] = side
    set_vec(w, "unit_normal", normal)
    set_vec(w, "wall_velocity", velocity)
    walls[slot] = w


def gather_boundary_cases():
    """The reference's boundary-gather cases (liquid_boundary_transfer_3d_cases.h:217-420) as
    (name, fixture, keyword arguments of BoundaryFixture.histories)."""
    out = []
    f = BoundaryFixture()
    for axis in range(3):
        for z in range(4):
            for y in range(4):
                for x in range(4):
                    a = f.anchor(axis, x, y, z)
                    s = f.slot(axis, x, y, z)
                    f.old_values[s] = 1.0 + a[0] + 2.0 * a[1] - a[2]
                    f.new_values[s] = 4.0 + 2.0 * a[0] - a[1] + 3.0 * a[2]
        poison = f.slot(axis, 1, 1, 1)
        f.tags[poison]["side_tag"] = -7
        f.old_values[poison] = -1e8
        f.new_values[poison] = 1e9
    out.append(("same_side_affine_ghosts", f, {}))
    f = BoundaryFixture()
    f.new_values = f.old_values.copy()
    out.append(("identical_histories", f, {}))
    f = BoundaryFixture()
    f.validity[:] = 0
    for slot in range(len(f.validity)):
        set_wall(f.old_walls, slot, 7, (1.0, 0.0, 0.0), (2.0, 4.0, 6.0))
        set_wall(f.new_walls, slot, 7, (1.0, 0.0, 0.0), (5.0, 8.0, 9.0))
    out.append(("walls_everywhere", f, {}))
    f = BoundaryFixture()
    fields = {0: ((1.0, 1.0, 2.0, -1.0), (4.0, 2.0, -1.0, 3.0)), 1: ((-2.0, 3.0, -1.0, 2.0), (7.0, -1.0, 4.0, -1.0)),
              2: ((4.0, -1.0, 1.0, 3.0), (-3.0, 3.0, 2.0, -2.0))}
    for axis in range(3):
        (o0, ox, oy, oz), (n0, nx, ny, nz) = fields[axis]
        for z in range(4):
            for y in range(4):
                for x in range(4):
                    a = f.anchor(axis, x, y, z)
                    s = f.slot(axis, x, y, z)
                    f.old_values[s] = o0 + ox * a[0] + oy * a[1] + oz * a[2]
                    f.new_values[s] = n0 + nx * a[0] + ny * a[1] + nz * a[2]
    nc = math.sqrt(0.5)
    target = f.slot(0, 1, 1, 1)
    set_wall(f.old_walls, target, 7, (nc, nc, 0.0), (2.0, -1.0, 4.0))
    set_wall(f.new_walls, target, 7, (nc, nc, 0.0), (5.0, 1.0, -2.0))
    f.validity[target] = 0
    out.append(("oblique_wall", f, {}))
    f = BoundaryFixture()
    f.validity[:] = 0
    for axis in range(3):
        first = f.slot(axis, 1, 1, 1)
        second = f.slot(axis, 2 if axis == 0 else 1, 2 if axis == 1 else 1, 2 if axis == 2 else 1)
        f.validity[first] = 1
        f.validity[second] = 2
        f.old_values[first] = f.old_values[second] = 2.0
        f.new_values[first] = f.new_values[second] = 4.0
    out.append(("singular_donors", f, {}))
    f = BoundaryFixture()
    f.validity[:] = 0
    out.append(("no_donors", f, {}))
    f = BoundaryFixture()
    nv = f.validity.copy()
    nv[f.slot(0, 1, 1, 1)] = 3
    out.append(("low_mass", f, {"old_validity": f.validity.copy(), "new_validity": nv}))
    f = BoundaryFixture()
    set_position(f.particle, 0, (1.75, 1.75, 1.75))
    for axis in range(3):
        for z in range(4):
            for y in range(4):
                for x in range(4):
                    a = f.anchor(axis, x, y, z)
                    f.new_values[f.slot(axis, x, y, z)] = (a[axis] - 1.75) ** 2
    out.append(("nonaffine_field", f, {}))
    f = BoundaryFixture()
    set_position(f.particle, 0, (2.0, 1.5, 1.5))
    nv = f.validity.copy()
    for z in range(4):
        for y in range(4):
            for x in range(4):
                nv[f.slot(0, x, y, z)] = 0
    low, high = f.slot(0, 1, 1, 1), f.slot(0, 3, 1, 1)
    nv[low], nv[high] = 1, 2
    f.new_values[low], f.new_values[high] = 2.0, 8.0
    out.append(("nearest_tie", f, {"new_validity": nv}))
    return out


def boundary_gather_scene(seed, n, h=0.5, blocks_per_axis=3, missing=0.05, low_mass=0.002, sides=2, eligible=0.9,
                          walls=0.2):
    """Random boundary-gather inputs: the masked-deposit scene plus two histories with random face
    velocities, validity (missing with probability `missing`, low mass with `low_mass`, otherwise valid
    or predicted), and unit-normal prescribed walls on a fraction `walls` of the slots."""
    geom, blocks, p, tags = masked_scene(seed, n, h=h, blocks_per_axis=blocks_per_axis, sides=sides,
                                         eligible=eligible)
    rng = np.random.default_rng(seed + 1000)
    nf = len(tags)
    hist = []
    for k in range(2):
        u = rng.normal(0.0, 1.0, nf)
        v = rng.choice(np.array([1, 2], np.uint8), nf, p=[0.9, 0.1])
        v[rng.random(nf) < low_mass] = 3
        v[rng.random(nf) < missing] = 0
        w = np.zeros(nf, O.WALL_SAMPLE)
        sel = np.nonzero(rng.random(nf) < walls)[0]
        nrm = rng.normal(0.0, 1.0, (len(sel), 3))
        nrm /= np.linalg.norm(nrm, axis=1)[:, None]
        for j, sl in enumerate(sel):
            set_wall(w, sl, int(rng.integers(0, sides)), tuple(nrm[j]), tuple(rng.normal(0.0, 1.0, 3)))
        hist.append((float(k), u, v, tags, w, float(k)))
    return geom, blocks, p, hist[0], hist[1]


def device_step_system(c, sc, report):
    """The projection a device step froze: its block list and u* at the liquid cells' incident faces, with
    the step's accepted pressures and corrected face velocities (at the same faces)."""
    nb, nc = int(report.blocks), int(report.pressure_cells)
    s, dl = c.scene_download(nb, 0, nc, what=("blocks", "ustar", "unew", "pressures"))
    assert s == 0
    slots = sc.pressure_face_slots(dl["blocks"])[0]
    args = (sc.geometry_record().view(O.GEOMETRY), dl["blocks"].view(O.INT3), sc.dt,
            sc.pressure_cells().view(O.PRESSURE_CELL), sc.pressure_faces(dl["blocks"], dl["ustar"]).view(O.PRESSURE_FACE))
    return args, dl["pressures"], dl["unew"][slots]


def assert_matches_fp64_oracle(args, pressures, velocities, tol=1e-5):
    """Tight FP64 oracle solve of the same system; the GPU pressures re-evaluated by the oracle's FP64 Evaluate
    meet the reference gates and the relative-divergence target."""
    s, ref = O.project(*args, O.pressure_options(relative_target=1e-12, fp32=False, levels=6, coarse_sweeps=2))
    assert s == O.OK
    p_err = np.abs(pressures - ref["pressures"]).max() / np.abs(ref["pressures"]).max()
    u_err = np.linalg.norm(velocities - ref["velocities"]) / np.linalg.norm(ref["velocities"])
    s, ev = O.evaluate(*args, pressures)
    assert s == O.OK
    d = ev["diagnostics"]
    print("fp64 oracle: pressure error", p_err, "velocity error", u_err, "divergence", d["defect_l2"],
          "relative", d["defect_l2"] / d["initial_defect_l2"], flush=True)
    assert p_err <= tol and u_err <= tol
    assert d["defect_l2"] <= 1e-6 * d["initial_defect_l2"]
    assert d["max_cell_constraint_defect"] <= 2e-5 and d["max_volume_scaled_residual"] <= 2e-5
    return p_err, u_err
