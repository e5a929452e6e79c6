"""Variational MAC viscosity on the GPU against the oracle (native/src/viscosity.cc).

Strain rows are compared byte for byte (same rows, slots, coefficients and weights in the reference's
order), on the reference's cases and on a dam-break liquid with random viscosities; the velocity block is
compared with the oracle's sequential Jacobi PCG (FP64, different summation order in the dots: velocities
within 1e-12 of the solution's scale, iteration counts within one) and on the reference's known answers.
"""
import numpy as np
import pytest

import cases as K
import pyoracle as O
from fluidity_amd import hip as H
from fluidity_amd import records as R
from fluidity_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(particles=1 << 16, blocks=1 << 14, lookup_blocks=1 << 18, pressure_cells=1 << 18,
                  pressure_faces=1 << 20)
    print("device", c.device_name(), flush=True)
    yield c
    c.close()


def _rows(rows):
    """Device-layout strain rows from oracle records (np.array of records drops the padding)."""
    return np.array(rows, dtype=O.STRAIN_ROW).view(R.STRAIN_ROW)


def _same(a, b):
    return a.tobytes() == b.tobytes()


def test_reference_strain_cases_bit_exact(ctx):
    blocks = np.zeros(8, O.INT3)
    for i, k in enumerate([(-1, -1, -1), (0, -1, -1), (-1, 0, -1), (0, 0, -1), (-1, -1, 0), (0, -1, 0), (-1, 0, 0),
                           (0, 0, 0)]):
        blocks[i] = k
    g = O.geometry((0.0, 0.0, 0.0), 0.5)
    _, blocks = O.sort_blocks(g, blocks)
    cells = O.viscosity_cells([(x, y, z) for z in (-1, 0) for y in (-1, 0) for x in (-1, 0)], 2.0)
    s, got = ctx.assemble_full_cell_strain(g.view(R.GEOMETRY), blocks.view(R.INT3), cells.view(R.VISCOSITY_CELL))
    so, want = O.assemble_full_cell_strain(g, blocks, cells)
    assert s == so == H.OK and len(got) == 30 and _same(got, want)
    one = np.zeros(1, O.INT3)
    g1 = O.geometry((0.0, 0.0, 0.0), 1.0)
    phases = O.viscosity_cells([(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)], [1.0, 1.0, 1000.0, 1000.0])
    s, got = ctx.assemble_full_cell_strain(g1.view(R.GEOMETRY), one.view(R.INT3), phases.view(R.VISCOSITY_CELL))
    so, want = O.assemble_full_cell_strain(g1, one, phases)
    assert s == so == H.OK and len(got) == 13 and _same(got, want)
    for bad in (O.viscosity_cells([(3, 0, 0), (3, 0, 0)], 1.0), O.viscosity_cells([(3, 0, 0)], 1.0),
                O.viscosity_cells([(0, 0, 0)], -1.0), O.viscosity_cells([(0, 0, 2**31 - 1)], 1.0),
                O.viscosity_cells([(1, 0, 0), (0, 0, 0)], 1.0)):
        s, _ = ctx.assemble_full_cell_strain(g1.view(R.GEOMETRY), one.view(R.INT3), bad.view(R.VISCOSITY_CELL))
        assert s == O.assemble_full_cell_strain(g1, one, bad)[0] != H.OK


def _dam_cells(n, seed):
    sc = scenes.dam_break(n, seed=seed)
    st = K.oracle_step_inputs(sc)
    idx = sc.liquid_cells()
    rng = np.random.default_rng(seed)
    mu = np.exp(rng.normal(0.0, 1.0, len(idx)))
    mu[rng.random(len(idx)) < 0.05] = 0.0  # some inviscid cells give zero edge viscosity
    return sc, st, O.viscosity_cells(idx, mu)


@pytest.mark.parametrize("seed", [1, 2])
def test_strain_rows_of_a_dam_break_bit_exact(ctx, seed):
    sc, st, cells = _dam_cells(24, seed)
    s, got = ctx.assemble_full_cell_strain(st["geom"].view(R.GEOMETRY), st["blocks"].view(R.INT3),
                                           cells.view(R.VISCOSITY_CELL))
    so, want = O.assemble_full_cell_strain(st["geom"], st["blocks"], cells)
    print("cells", len(cells), "rows", len(want), flush=True)
    assert s == so == H.OK and len(got) == len(want) and _same(got, want)


def test_velocity_block_known_answers(ctx):
    rows = _rows([O.strain_row([(0, 1.0), (1, -1.0)], 2.0)])
    s, v, it, mr = ctx.solve_viscous_block(rows, [1.0, 3.0, 0.0], 0.5, [3.0, 2.0, 0.0])
    assert s == H.OK and np.allclose(v, [2.0, 1.0, 0.0], atol=1e-10, rtol=0) and v[2] == 0.0
    s, v, _, _ = ctx.solve_viscous_block(rows, [1.0, 0.0, 0.0], 0.5, [3.0, 0.0, 0.0])
    assert s == H.OK and abs(v[0] - 1.5) <= 1e-10 and v[1] == 0.0
    assert ctx.solve_viscous_block(rows, [1.0, 0.0, 0.0], 0.5, [3.0, 2.0, 0.0])[0] == H.INVALID_ARGUMENT
    assert ctx.solve_viscous_block(rows, [1.0, 3.0, 0.0], 0.5, [3.0, 2.0, 0.0], max_iterations=1)[0] == H.ABORTED
    rows2 = _rows([O.strain_row([(0, 1.0), (1, -1.0)], 2.0),
                      O.strain_row([(0, 0.5), (1, -0.5), (2, 0.5), (3, -0.5)], 8.0, 0, 1)])
    s, v, _, _ = ctx.solve_viscous_block(rows2, [1.0, 2.0, 3.0, 5.0], 0.5, [13.0, -13.0, 17.0, -18.0])
    assert s == H.OK and np.allclose(v, [2.0, -1.0, 3.0, -2.0], atol=1e-10, rtol=0)
    rows3 = _rows([O.strain_row([(0, -2.0), (1, 1.0)], 4.0), O.strain_row([(1, 2.0), (2, -1.0)], 0.125),
                      O.strain_row([(2, 1.0), (3, -2.0)], 8.0)])
    rhs = [3.0, -2.0, 5.0, 0.0]
    s, v, it, mr = ctx.solve_viscous_block(rows3, [1.0, 2.0, 3.0, 0.0], 0.5, rhs)
    so, vo, ito, _ = O.solve_viscous_block(rows3.view(O.STRAIN_ROW), [1.0, 2.0, 3.0, 0.0], 0.5, rhs)
    assert s == so == H.OK and it == ito and np.allclose(v, vo, atol=1e-13, rtol=0) and v[3] == 0.0 and mr <= 5e-10
    s, v, it, mr = ctx.solve_viscous_block(rows3, [1.0, 2.0, 3.0, 0.0], 0.5, [0.0] * 4)
    assert s == H.OK and it == 0 and mr == 0.0 and not v.any()
    dup = _rows([O.strain_row([(0, 1.0), (0, -1.0)], 4.0)])
    assert ctx.solve_viscous_block(dup, [1.0, 1.0, 0.0], 1.0, [1.0, 0.0, 0.0])[0] == H.INVALID_ARGUMENT


@pytest.mark.parametrize("seed", [1, 2])
def test_velocity_block_of_a_dam_break_matches_oracle(ctx, seed):
    """Backward-Euler viscous step of a dam break's faces: inertia = face mass (zero on massless faces,
    which are eliminated), rhs = mass * u*, random viscosity per liquid cell. At dt = 1e-3 (a few
    iterations) the GPU iterates follow the oracle's to 1e-12; at dt = 0.5 (about 70 iterations) both
    solutions are compared with a tight oracle solve (relative tolerance 1e-14) to the accuracy the
    1e-10 stopping rule allows."""
    sc, st, cells = _dam_cells(24, seed)
    so, rows = O.assemble_full_cell_strain(st["geom"], st["blocks"], cells)
    assert so == O.OK
    mass = np.asarray(st["mass"], dtype=np.float64)
    rhs = mass * np.asarray(st["ustar"], dtype=np.float64)
    s, v, it, mr = ctx.solve_viscous_block(rows.view(R.STRAIN_ROW), mass, 1e-3, rhs)
    so, vo, ito, mro = O.solve_viscous_block(rows, mass, 1e-3, rhs)
    print("faces", len(mass), "rows", len(rows), "iterations", it, ito, flush=True)
    assert s == so == H.OK and abs(int(it) - int(ito)) <= 1
    scale = max(np.abs(vo).max(), 1e-300)
    assert np.all(np.abs(v - vo) <= 1e-12 * scale) and np.all(v[mass == 0.0] == 0.0)
    s, v, it, mr = ctx.solve_viscous_block(rows.view(R.STRAIN_ROW), mass, 0.5, rhs)
    so, vo, ito, mro = O.solve_viscous_block(rows, mass, 0.5, rhs)
    st_, ve, _, _ = O.solve_viscous_block(rows, mass, 0.5, rhs, max_iterations=256, relative_tolerance=1e-14)
    print("dt 0.5 iterations", it, ito, "max residual", mr, mro, flush=True)
    assert s == so == st_ == H.OK and abs(int(it) - int(ito)) <= 3 and mr <= 1e-10 * np.abs(rhs).max()
    scale = np.abs(ve).max()
    assert np.abs(v - ve).max() <= 1e-8 * scale and np.abs(vo - ve).max() <= 1e-8 * scale
