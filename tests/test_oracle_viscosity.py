"""Oracle of the variational MAC viscosity against the reference's known answers.

Strain assembly: native/tests/test_viscosity.cc (FullCell3DStrain..., 2x2x2 cells of h = 0.5 with 24
normal and 6 shear rows whose strain of affine velocities is exact; harmonic shear across phases; the
duplicate-cell and missing-owner refusals). Velocity block: the same file's sparse block with an
eliminated face, the normal-plus-shear block with a known solution, and native/tests/test_cuda_viscosity.cc's
three-row block with its true-residual identity and refusals.
"""
import numpy as np

import pyoracle as O


def _grid_2x2x2():
    blocks = np.zeros(8, O.INT3)
    for i, (x, y, z) in enumerate([(-1, -1, -1), (0, -1, -1), (-1, 0, -1), (0, 0, -1), (-1, -1, 0), (0, -1, 0),
                                   (-1, 0, 0), (0, 0, 0)]):
        blocks[i] = (x, y, z)
    g = O.geometry((0.0, 0.0, 0.0), 0.5)
    s, sb = O.sort_blocks(g, blocks)
    assert s == O.OK
    return g, sb


def test_full_cell_strain_rows_weights_and_affine_exactness():
    g, blocks = _grid_2x2x2()
    idx = [(x, y, z) for z in (-1, 0) for y in (-1, 0) for x in (-1, 0)]
    s, rows = O.assemble_full_cell_strain(g, blocks, O.viscosity_cells(idx, 2.0))
    assert s == O.OK and len(rows) == 30
    normal_d = (2.0, 7.0, 13.0)
    shear_d = {(0, 1): 3.0, (0, 2): 5.0, (1, 2): 11.0}
    for k, row in enumerate(rows):
        normal = k < 24
        assert row["term_count"] == (2 if normal else 4)
        assert row["quadrature_weight"] == (0.5 if normal else 1.0)
        rot = sym = 0.0
        for t in range(int(row["term_count"])):
            slot, c = int(row["terms"][t]["face_slot"]), float(row["terms"][t]["coefficient"])
            s, (axis, lat) = O.slot_face(blocks, slot)
            s2, p = O.face_world_anchor(g, axis, lat)
            assert s == O.OK and s2 == O.OK
            x, y, z = p
            rot += c * (2 * x + 3 * y + 5 * z, -3 * x + 7 * y + 11 * z, -5 * x - 11 * y + 13 * z)[axis]
            sym += c * (2 * x + 3 * y + 5 * z, 3 * x + 7 * y + 11 * z, 5 * x + 11 * y + 13 * z)[axis]
        a, b = int(row["first_axis"]), int(row["second_axis"])
        assert rot == (normal_d[k % 3] if normal else 0.0)
        assert sym == (normal_d[k % 3] if normal else shear_d[(a, b)])


def test_harmonic_shear_across_phases():
    blocks = np.zeros(1, O.INT3)
    g = O.geometry((0.0, 0.0, 0.0), 1.0)
    cells = O.viscosity_cells([(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)], [1.0, 1.0, 1000.0, 1000.0])
    s, rows = O.assemble_full_cell_strain(g, blocks, cells)
    assert s == O.OK and len(rows) == 13
    shear = rows[-1]
    assert shear["term_count"] == 4 and abs(shear["quadrature_weight"] - 8.0 / 1.001) <= 1e-12
    traction = 1.0 / 1.001
    strain = 0.0
    for t in range(4):
        slot, c = int(shear["terms"][t]["face_slot"]), float(shear["terms"][t]["coefficient"])
        _, (axis, lat) = O.slot_face(blocks, slot)
        _, p = O.face_world_anchor(g, axis, lat)
        if axis == 0:
            strain += c * traction * (min(p[1], 1.0) + max(p[1] - 1.0, 0.0) / 1000.0)
    assert abs(shear["quadrature_weight"] * strain / 2.0 - traction) <= 1e-12


def test_strain_refusals():
    blocks = np.zeros(1, O.INT3)
    g = O.geometry((0.0, 0.0, 0.0), 1.0)
    cell = O.viscosity_cells([(3, 0, 0)], [1.0])
    assert O.assemble_full_cell_strain(g, blocks, np.concatenate([cell, cell]))[0] == O.INVALID_ARGUMENT
    assert O.assemble_full_cell_strain(g, blocks, cell)[0] == O.FAILED_PRECONDITION
    assert O.assemble_full_cell_strain(g, blocks, O.viscosity_cells([(0, 0, 0)], [-1.0]))[0] == O.INVALID_ARGUMENT
    assert O.assemble_full_cell_strain(g, blocks, O.viscosity_cells([(0, 0, 2**31 - 1)], [1.0]))[0] == O.OUT_OF_RANGE
    s, rows = O.assemble_full_cell_strain(g, blocks, cell[:0])
    assert s == O.OK and len(rows) == 0


def test_velocity_block_known_answers():
    rows = np.array([O.strain_row([(0, 1.0), (1, -1.0)], 2.0)])
    s, v, it, mr = O.solve_viscous_block(rows, [1.0, 3.0, 0.0], 0.5, [3.0, 2.0, 0.0])
    assert s == O.OK and np.allclose(v, [2.0, 1.0, 0.0], atol=1e-10, rtol=0) and v[2] == 0.0 and mr <= 1e-10
    s, v, _, _ = O.solve_viscous_block(rows, [1.0, 0.0, 0.0], 0.5, [3.0, 0.0, 0.0])
    assert s == O.OK and abs(v[0] - 1.5) <= 1e-10 and v[1] == 0.0
    assert O.solve_viscous_block(rows, [1.0, 0.0, 0.0], 0.5, [3.0, 2.0, 0.0])[0] == O.INVALID_ARGUMENT
    assert O.solve_viscous_block(rows, [1.0, 3.0, 0.0], 0.5, [3.0, 2.0, 0.0], max_iterations=1)[0] == O.ABORTED
    rows2 = np.array([O.strain_row([(0, 1.0), (1, -1.0)], 2.0),
                      O.strain_row([(0, 0.5), (1, -0.5), (2, 0.5), (3, -0.5)], 8.0, 0, 1)])
    s, v, _, mr = O.solve_viscous_block(rows2, [1.0, 2.0, 3.0, 5.0], 0.5, [13.0, -13.0, 17.0, -18.0])
    assert s == O.OK and np.allclose(v, [2.0, -1.0, 3.0, -2.0], atol=1e-10, rtol=0)


def test_velocity_block_true_residual_and_refusals():
    rows = np.array([O.strain_row([(0, -2.0), (1, 1.0)], 4.0), O.strain_row([(1, 2.0), (2, -1.0)], 0.125),
                     O.strain_row([(2, 1.0), (3, -2.0)], 8.0)])
    rhs = [3.0, -2.0, 5.0, 0.0]
    s, v, it, mr = O.solve_viscous_block(rows, [1.0, 2.0, 3.0, 0.0], 0.5, rhs)
    assert s == O.OK and v[3] == 0.0 and mr <= 5e-10
    defects = (rhs[0] - (9 * v[0] - 4 * v[1]), rhs[1] - (-4 * v[0] + 4.25 * v[1] - 0.125 * v[2]),
               rhs[2] - (-0.125 * v[1] + 7.0625 * v[2]))
    assert max(abs(d) for d in defects) <= 5e-10
    one = np.array([O.strain_row([(0, 1.0), (1, -1.0)], 4.0)])
    assert O.solve_viscous_block(one, [1.0, 1.0, 0.0], 1.0, [1.0, 0.0, 0.0], max_iterations=1)[0] == O.ABORTED
    assert O.solve_viscous_block(one, [1.0, 1.0, 0.0], 1.0, [1.0, 0.0, 1.0])[0] == O.INVALID_ARGUMENT
    s, v, it, mr = O.solve_viscous_block(one, [1.0, 1.0, 0.0], 1.0, [0.0, 0.0, 0.0])
    assert s == O.OK and it == 0 and mr == 0.0 and not v.any()
    dup = np.array([O.strain_row([(0, 1.0), (0, -1.0)], 4.0)])
    assert O.solve_viscous_block(dup, [1.0, 1.0, 0.0], 1.0, [1.0, 0.0, 0.0])[0] == O.INVALID_ARGUMENT
