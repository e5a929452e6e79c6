"""Seeded synthetic liquid scenes for the hot path (dam break, colliding spheres over a pool).

Units are SI on the unit cube with origin 0 and cell width h = 1/n. Liquid cells are those whose
centre lies inside the liquid region; each gets 8 particles, one per octant of the cell, placed
uniformly at random inside its octant (numpy Philox stream, seeded). Particle mass is rho h^3 / 8,
gravity (0, -9.81, 0), rho = 1000, dt = 0.5 h / max(1, |u|max).

Pressure inputs follow the reference's record format (native/include/fluidity/host/mac_pressure_3d.h:14-51):
liquid cells are the unknowns; a face between two liquid cells is an internal link (area h^2,
jump 0); a face between a liquid cell and an air cell inside the domain is a Dirichlet link with
exterior pressure 0 and gradient distance h (first-order free surface); a face on the domain
boundary is a sealed link (zero area, zero wall velocity).
"""
from __future__ import annotations

import dataclasses

import math

import numpy as np

from . import records as R

RHO = 1000.0
GRAVITY = (0.0, -9.81, 0.0)


@dataclasses.dataclass
class Scene:
    name: str
    n: int
    seed: int
    liquid: np.ndarray          # (n, n, n) bool indexed [z, y, x]
    positions: np.ndarray       # (N, 3) float64
    velocities: np.ndarray      # (N, 3) float64
    slopes: np.ndarray          # (N, 3, 3) float64, rows by velocity component
    mass: float
    rho: float = RHO
    gravity: tuple = GRAVITY

    @property
    def h(self) -> float:
        return 1.0 / self.n

    @property
    def count(self) -> int:
        return len(self.positions)

    @property
    def dt(self) -> float:
        umax = float(np.sqrt((self.velocities ** 2).sum(1)).max()) if self.count else 0.0
        return 0.5 * self.h / max(1.0, umax)

    def geometry_record(self) -> np.ndarray:
        g = np.zeros(1, R.GEOMETRY)
        g["cell_width"] = self.h
        return g

    def particles_record(self, index=None) -> np.ndarray:
        sel = slice(None) if index is None else index
        pos, vel, c = self.positions[sel], self.velocities[sel], self.slopes[sel]
        p = np.zeros(len(pos), R.PARTICLE)
        for i, k in enumerate("xyz"):
            p["sorted_world_position"][k] = pos[:, i]
            p["trial_world_position"][k] = pos[:, i]
            p["particle_velocity"][k] = vel[:, i]
            for j, row in enumerate(("x_row", "y_row", "z_row")):
                p["particle_velocity_slopes"][row][k] = c[:, j, i]
        p["particle_mass"] = self.mass
        return p

    def liquid_cells(self) -> np.ndarray:
        """(M, 3) int32 cell indices in (z, y, x) order."""
        z, y, x = np.nonzero(self.liquid)
        return np.stack([x, y, z], 1).astype(np.int32)

    def pressure_cells(self) -> np.ndarray:
        cells = self.liquid_cells()
        rec = np.zeros(len(cells), R.PRESSURE_CELL)
        rec["cell_index"]["x"], rec["cell_index"]["y"], rec["cell_index"]["z"] = cells.T
        rec["fluid_volume"] = self.h ** 3
        rec["bulk_modulus"] = np.inf
        return rec

    def pressure_faces(self, blocks: np.ndarray, predicted: np.ndarray | None = None) -> np.ndarray:
        """Incident faces of the liquid cells in owned-slot order for the sorted block list `blocks`.

        `predicted` optionally holds the face velocity per owned slot (the forces stage output).
        """
        slots, axis, lat = self.pressure_face_slots(blocks)
        n, h = self.n, self.h
        tail = lat.copy()
        tail[np.arange(len(lat)), axis] -= 1
        on_wall = (lat[np.arange(len(lat)), axis] == 0) | (lat[np.arange(len(lat)), axis] == n)

        def is_liquid(c):
            inside = np.all((c >= 0) & (c < n), axis=1)
            out = np.zeros(len(c), bool)
            ci = c[inside]
            out[inside] = self.liquid[ci[:, 2], ci[:, 1], ci[:, 0]]
            return out

        both = is_liquid(tail) & is_liquid(lat)
        rec = np.zeros(len(slots), R.PRESSURE_FACE)
        rec["axis"] = axis
        rec["lattice"]["x"], rec["lattice"]["y"], rec["lattice"]["z"] = lat.T
        kind = np.where(on_wall, R.SEALED, np.where(both, R.INTERNAL, R.DIRICHLET))
        rec["link_kind"] = kind
        rec["open_area"] = np.where(kind == R.SEALED, 0.0, h * h)
        rec["inverse_density"] = 1.0 / self.rho
        rec["pressure_gradient_distance"] = h
        if predicted is not None:
            rec["predicted_normal_velocity"] = predicted[slots]
        return rec

    def pressure_face_slots(self, blocks: np.ndarray):
        """Owned slots, normal axes and head lattice indices of the liquid cells' incident faces."""
        cells = self.liquid_cells().astype(np.int64)
        lut, lo, dims = block_lookup(blocks)
        faces = []
        for a in range(3):
            for side in (0, 1):
                f = cells.copy()
                f[:, a] += side
                faces.append(np.concatenate([np.full((len(f), 1), a), f], 1))
        faces = np.concatenate(faces)
        owner = np.floor_divide(faces[:, 1:], 4)
        b = lookup(lut, lo, dims, owner)
        if np.any(b < 0):
            raise ValueError("a pressure face has no active owner block")
        local = faces[:, 1:] - 4 * owner
        slots = b * 192 + faces[:, 0] * 64 + local[:, 2] * 16 + local[:, 1] * 4 + local[:, 0]
        slots, first = np.unique(slots, return_index=True)
        return slots, faces[first, 0].astype(np.int32), faces[first, 1:].astype(np.int32)


def block_lookup(blocks: np.ndarray):
    """Dense ordinal table over the bounding box of a sorted block list (structured INT3 or (M,3))."""
    b = _as_int3(blocks)
    lo = b.min(0)
    dims = b.max(0) - lo + 1
    lut = np.full(int(np.prod(dims)), -1, np.int64)
    rel = b - lo
    lut[(rel[:, 2] * dims[1] + rel[:, 1]) * dims[0] + rel[:, 0]] = np.arange(len(b))
    return lut, lo, dims


def lookup(lut, lo, dims, keys):
    rel = keys - lo
    inside = np.all((rel >= 0) & (rel < dims), axis=1)
    out = np.full(len(keys), -1, np.int64)
    r = rel[inside]
    out[inside] = lut[(r[:, 2] * dims[1] + r[:, 1]) * dims[0] + r[:, 0]]
    return out


def _as_int3(blocks):
    if blocks.dtype.names:
        return np.stack([blocks["x"], blocks["y"], blocks["z"]], 1).astype(np.int64)
    return np.asarray(blocks, np.int64).reshape(-1, 3)


def _seed_particles(liquid: np.ndarray, n: int, seed: int):
    """8 particles per liquid cell, one per octant, uniform inside the octant."""
    h = 1.0 / n
    cells = np.stack(np.nonzero(liquid)[::-1], 1).astype(np.float64)  # (M, 3) x, y, z
    rng = np.random.Generator(np.random.Philox(seed))
    octants = np.array([[ox, oy, oz] for oz in (0, 1) for oy in (0, 1) for ox in (0, 1)], np.float64)
    jitter = rng.random((len(cells), 8, 3))
    pos = (cells[:, None, :] + (octants[None, :, :] + jitter) * 0.5) * h
    return pos.reshape(-1, 3), rng


def _centres(n):
    c = (np.arange(n) + 0.5) / n
    return c[None, None, :], c[None, :, None], c[:, None, None]  # x, y, z broadcast over [z, y, x]


def dam_break(n: int = 64, seed: int = 0, random_velocity: bool = False, dense: bool = False) -> Scene:
    """Dam break: liquid where x < 0.4 and y < 0.6; v = 0, C = 0 (or N(0,1) and N(0,0.1) if random)."""
    x, y, z = _centres(n)
    liquid = np.broadcast_to((x < 0.4) & (y < 0.6), (n, n, n)).copy()
    pos, rng = _seed_particles(liquid, n, seed)
    count = len(pos)
    if random_velocity:
        vel = rng.normal(0.0, 1.0, (count, 3))
        slopes = rng.normal(0.0, 0.1, (count, 3, 3))
    else:
        vel = np.zeros((count, 3))
        slopes = np.zeros((count, 3, 3))
    name = f"dam_break_{n}" + ("_random" if random_velocity else "") + ("_dense" if dense else "")
    return Scene(name, n, seed, liquid, pos, vel, slopes, RHO * (1.0 / n) ** 3 / 8)


def two_spheres(n: int = 256, seed: int = 0) -> Scene:
    """Spheres r=0.2 at (0.3,0.5,0.5) moving +2 in x and (0.7,0.5,0.5) moving -2, over a pool y < 0.1."""
    x, y, z = _centres(n)
    s1 = (x - 0.3) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2 < 0.04
    s2 = (x - 0.7) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2 < 0.04
    pool = np.broadcast_to(y < 0.1, (n, n, n))
    liquid = s1 | s2 | pool
    pos, _ = _seed_particles(liquid, n, seed)
    cells = np.floor(pos * n).astype(np.int64)
    in1 = s1[cells[:, 2], cells[:, 1], cells[:, 0]]
    in2 = s2[cells[:, 2], cells[:, 1], cells[:, 0]]
    vel = np.zeros((len(pos), 3))
    vel[in1, 0] = 2.0
    vel[in2, 0] = -2.0
    return Scene(f"two_spheres_{n}", n, seed, liquid, pos, vel, np.zeros((len(pos), 3, 3)), RHO * (1.0 / n) ** 3 / 8)


def solid_distance(points, sphere=None, paddle=None, walls=None):
    """Signed distance (positive outside) of the union of a sphere (centre, radius), a box paddle (pivot,
    half extent, angle about +y, angular velocity) and the side walls of a container ((low, high) in x and
    z), and the wall velocity of the nearest part; the same operations, in the same order, as the device
    substep's sampling of the scene solid. points: (m, 3)."""
    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    best = np.full(len(points), np.inf)
    vx = np.zeros(len(points))
    vz = np.zeros(len(points))
    if sphere is not None:
        (cx, cy, cz), r = sphere
        dx, dy, dz = x - cx, y - cy, z - cz
        best = np.sqrt(dx * dx + dy * dy + dz * dz) - r
    if paddle is not None:
        (px, py, pz), (ex, ey, ez), angle, omega = paddle
        ca, sa = math.cos(angle), math.sin(angle)
        rx, ry, rz = x - px, y - py, z - pz
        bx, bz = ca * rx - sa * rz, sa * rx + ca * rz
        qx, qy, qz = np.abs(bx) - ex, np.abs(ry) - ey, np.abs(bz) - ez
        mx, my, mz = np.maximum(qx, 0.0), np.maximum(qy, 0.0), np.maximum(qz, 0.0)
        d = np.sqrt(mx * mx + my * my + mz * mz) + np.minimum(np.maximum(qx, np.maximum(qy, qz)), 0.0)
        take = d < best
        best = np.where(take, d, best)
        vx = np.where(take, omega * rz, vx)
        vz = np.where(take, -omega * rx, vz)
    if walls is not None:
        (lx, _, lz), (hx, _, hz) = walls
        d = np.minimum(np.minimum(x - lx, hx - x), np.minimum(z - lz, hz - z))
        take = d < best
        best = np.where(take, d, best)
        vx = np.where(take, 0.0, vx)
        vz = np.where(take, 0.0, vz)
    return best, vx, vz


def stirred_dam_break(n: int = 256, seed: int = 0, paddle_half_extent=(0.18, 0.1, 0.02), paddle_pivot=(0.7, 0.12, 0.5)):
    """Dam break (liquid where x < 0.4 and y < 0.3) in the unit box with a static sphere r = 0.15 at the
    centre and a box paddle turning at 2 rad/s about the vertical through its pivot; the side walls
    are part of the solid, the floor y = 0 is a half-space obstacle. Particles seeded inside the solid are
    dropped. Returns (scene, solid description: dict of sphere, paddle, walls)."""
    solid = dict(sphere=((0.5, 0.5, 0.5), 0.15), paddle=(tuple(paddle_pivot), tuple(paddle_half_extent), 0.0, 2.0),
                 walls=((0.0, 0.0, 0.0), (1.0, 0.0, 1.0)))
    x, y, z = _centres(n)
    liquid = np.broadcast_to((x < 0.4) & (y < 0.3), (n, n, n)).copy()
    pos, _ = _seed_particles(liquid, n, seed)
    keep = solid_distance(pos, **solid)[0] > 0.0
    pos = pos[keep]
    sc = Scene(f"stirred_dam_break_{n}", n, seed, liquid, pos, np.zeros((len(pos), 3)), np.zeros((len(pos), 3, 3)),
               RHO * (1.0 / n) ** 3 / 8)
    return sc, solid


def by_name(name: str, seed: int = 0) -> Scene:
    if name == "dam_break_64":
        return dam_break(64, seed)
    if name == "dam_break_128":
        return dam_break(128, seed)
    if name == "dam_break_128_random":
        return dam_break(128, seed, random_velocity=True)
    if name == "two_spheres_256":
        return two_spheres(256, seed)
    raise KeyError(name)
