"""GPU parity of topology preparation with history remap (fl_prepare_topology) against the oracle.

Bar: the same status as the oracle and, on success, bit-identical sort entries and block, cell and
face records, on the reference's known-answer cases (native/tests/test_liquid_step_workspace_3d.cc),
on the record checks of a previous state, on random histories with shifted blocks, sparse previous
faces and duplicate cells, and on a 32^3 dam-break scene stepped twice.
"""
import math

import numpy as np
import pytest

import pyoracle as O
from fluidity_amd import hip as H
from fluidity_amd import records as R
from fluidity_amd import scenes

pytestmark = pytest.mark.gpu

UNIT = O.geometry((0.0, 0.0, 0.0), 1.0)
SMALL = dict(particles=16, blocks=32, cells=16, faces=32 * 192)


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(particles=1 << 20, blocks=1 << 14, lookup_blocks=1 << 16, pressure_cells=1 << 18)
    print("device", c.device_name(), flush=True)
    yield c
    c.close()


def same(a, b):
    assert a.shape == b.shape
    for name in a.dtype.names:
        x, y = a[name], b[name]
        if x.dtype.names:
            same(x, y)
        elif x.dtype.kind == "f":
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), name
        else:
            assert np.array_equal(x, y), name


def run(ctx, positions=(), cells=(), previous=None, epoch=0, caps=None, geom=UNIT, prev_geom=None):
    """Both implementations on the same inputs; previous is (blocks, cells, faces) in oracle layout."""
    caps = caps or dict(particles=1 << 20, blocks=1 << 14, cells=1 << 18, faces=192 << 14)
    pos = np.array(positions, dtype=np.float64).reshape(-1, 3)
    cl = O.int3([tuple(c) for c in cells]) if len(cells) else np.zeros(0, O.INT3)
    pg = geom if prev_geom is None else prev_geom
    oh = gh = None
    if previous is not None:
        pb, pc, pf = previous
        oh = O.topology_history(pg, epoch, pb, pc, pf)
        gh = H.topology_history(pg, epoch, pb.view(R.BLOCK_STATE), pc.view(R.CELL_STATE), pf.view(R.FACE_STATE))
    s_want, *want = O.prepare_topology(geom, pos, cl, oh, O.topology_capacities(**caps))
    s_got, *got = ctx.prepare_topology(geom, pos, cl.view(R.INT3), gh, H.topology_capacities(**caps))
    assert s_got == s_want, (s_got, s_want)
    if s_want == O.OK:
        same(got[0], want[0])
        same(got[1].view(O.BLOCK_STATE), want[1])
        same(got[2].view(O.CELL_STATE), want[2])
        same(got[3].view(O.FACE_STATE), want[3])
    return s_want, want


def test_reference_known_answers(ctx):
    pos = [(2.0, 2.0, 2.0), (-2.0, 2.0, 2.0), (2.1, 2.0, 2.0), (2.0, -2.0, 2.0), (2.0, 2.0, -2.0), (-2.1, 2.0, 2.0)]
    s, (order, _, _, _) = run(ctx, pos, caps=SMALL)
    assert s == O.OK and list(order["original_ordinal"]) == [4, 3, 1, 5, 0, 2]
    s, (_, blocks, _, faces) = run(ctx, [(3.0, 3.0, 3.0)], caps=SMALL)
    assert s == O.OK and len(blocks) == 8 and len(faces) == 8 * 192
    s, (_, blocks, cells, _) = run(ctx, cells=[(3, 3, 3), (-1, -1, -1), (3, 3, 3)], caps=SMALL)
    assert s == O.OK and len(blocks) == 8 and len(cells) == 2
    # RemapsHistoryByIdentityAcrossShiftedBlockOrdinals (test_liquid_step_workspace_3d.cc:142-190)
    s, (_, blocks, cells, faces) = run(ctx, cells=[(1, 1, 1), (9, 1, 1)], caps=SMALL)
    blocks[0]["generation"], blocks[0]["activation_age"], blocks[0]["inactive_age"] = 4, 7, 2
    cells[0]["previous_pressure"], cells[0]["occupancy"] = 123.5, 2
    faces["carrier_velocity"] = 10.0 + faces["face_slot"].astype(np.float64)
    faces["validity"], faces["side_tag"], faces["structurally_eligible"] = 3, -17, 0
    s, (_, nb, nc, nf) = run(ctx, cells=[(-3, 1, 1), (1, 1, 1)], previous=(blocks, cells, faces), epoch=11, caps=SMALL)
    assert s == O.OK and nb[0]["generation"] == 12 and nb[1]["generation"] == 4 and nc[1]["previous_pressure"] == 123.5
    # EmptyPreparationClearsExtentsAndReactivationInitializesHistory (:296-320)
    empty = (np.zeros(0, O.BLOCK_STATE), np.zeros(0, O.CELL_STATE), np.zeros(0, O.FACE_STATE))
    assert run(ctx, previous=(blocks, cells, faces), caps=SMALL)[0] == O.OK
    s, (_, nb, _, _) = run(ctx, cells=[(1, 1, 1)], previous=empty, epoch=9, caps=SMALL)
    assert s == O.OK and nb[0]["generation"] == 10


def test_capacity_and_input_errors(ctx):
    tiny = dict(particles=1, blocks=1, cells=1, faces=192)
    assert run(ctx, cells=[(1, 1, 1)], caps=tiny)[0] == O.OK
    assert run(ctx, [(3, 3, 3)], caps=tiny)[0] == O.RESOURCE_EXHAUSTED
    assert run(ctx, [(2, 2, 2), (2, 2, 2)], caps=tiny)[0] == O.RESOURCE_EXHAUSTED
    assert run(ctx, cells=[(1, 1, 1), (2, 1, 1)], caps=tiny)[0] == O.RESOURCE_EXHAUSTED
    assert run(ctx, cells=[(1, 1, 1)], caps=dict(SMALL, faces=191))[0] == O.RESOURCE_EXHAUSTED
    assert run(ctx, cells=[(2147483647, 0, 0)])[0] == O.OUT_OF_RANGE
    assert run(ctx, [(2147483647.0, 0.0, 0.0)])[0] == O.OUT_OF_RANGE
    assert run(ctx, [(math.nan, 0.0, 0.0)])[0] == O.INVALID_ARGUMENT
    assert run(ctx, [(1.0, 1.0, 1.0), (math.nan, 0.0, 0.0)], cells=[(2147483647, 0, 0)])[0] == O.INVALID_ARGUMENT
    assert run(ctx, [(1.0, 1.0, 1.0), (-3e9, 0.0, 0.0), (math.nan, 0.0, 0.0)])[0] == O.OUT_OF_RANGE
    empty = (np.zeros(0, O.BLOCK_STATE), np.zeros(0, O.CELL_STATE), np.zeros(0, O.FACE_STATE))
    assert run(ctx, previous=empty, epoch=2 ** 64 - 1)[0] == O.OUT_OF_RANGE
    assert run(ctx, previous=empty, prev_geom=O.geometry((0.0, 0.0, 0.0), 2.0))[0] == O.INVALID_ARGUMENT


def valid_history(ctx):
    s, (_, blocks, cells, faces) = run(ctx, [(1.5, 1.5, 1.5), (6.5, 1.5, 1.5)], cells=[(1, 1, 1), (6, 1, 1)])
    assert s == O.OK
    return blocks, cells, faces


MUTATIONS = {
    "dup_block": (lambda b, c, f: b.__setitem__(0, b[1]), O.INVALID_ARGUMENT),
    "block_range": (lambda b, c, f: b[0]["block_index"].__setitem__("x", 536870912), O.OUT_OF_RANGE),
    "cell_order": (lambda b, c, f: c.__setitem__(0, c[1]), O.INVALID_ARGUMENT),
    "cell_owner": (lambda b, c, f: c[0]["cell_index"].__setitem__("z", 400), O.INVALID_ARGUMENT),
    "occupancy": (lambda b, c, f: c.__setitem__("occupancy", 3), O.INVALID_ARGUMENT),
    "pressure": (lambda b, c, f: c.__setitem__("previous_pressure", math.inf), O.INVALID_ARGUMENT),
    "validity": (lambda b, c, f: f[5].__setitem__("validity", 4), O.INVALID_ARGUMENT),
    "carrier": (lambda b, c, f: f[5].__setitem__("carrier_velocity", math.nan), O.INVALID_ARGUMENT),
    "axis": (lambda b, c, f: f[5]["face_index"].__setitem__("axis", 3), O.INVALID_ARGUMENT),
    "slot": (lambda b, c, f: f[5].__setitem__("face_slot", 6), O.INVALID_ARGUMENT),
    "slot_order": (lambda b, c, f: f.__setitem__(slice(5, 7), f[5:7][::-1].copy()), O.INVALID_ARGUMENT),
}


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_history_record_checks(ctx, name):
    mutate, code = MUTATIONS[name]
    blocks, cells, faces = valid_history(ctx)
    mutate(blocks, cells, faces)
    assert run(ctx, cells=[(1, 1, 1)], previous=(blocks, cells, faces), epoch=1)[0] == code


def test_history_checks_run_in_record_order(ctx):
    blocks, cells, faces = valid_history(ctx)
    cells[0]["occupancy"] = 7
    faces[3]["validity"] = 9
    blocks[1]["block_index"]["y"] = -536870913
    assert run(ctx, previous=(blocks, cells, faces), epoch=1)[0] == O.OUT_OF_RANGE


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_random_histories(ctx, seed):
    rng = np.random.default_rng(seed)
    geom = O.geometry((-0.37, 0.11, 0.5), 0.25 if seed % 2 else 0.3)
    pos0 = rng.uniform(-2.0, 2.0, (3000, 3))
    cells0 = rng.integers(-12, 12, (900, 3))
    s, (_, b0, c0, f0) = run(ctx, pos0, cells0, geom=geom)
    assert s == O.OK
    b0["activation_age"] = rng.integers(0, 9, len(b0))
    b0["inactive_age"] = rng.integers(0, 9, len(b0))
    b0["generation"] = rng.integers(1, 5, len(b0))
    c0["occupancy"] = rng.integers(0, 3, len(c0))
    c0["previous_pressure"] = rng.normal(0, 100, len(c0))
    f0["carrier_velocity"] = rng.normal(0, 1, len(f0))
    f0["validity"] = rng.integers(0, 4, len(f0))
    f0["side_tag"] = rng.integers(-5, 5, len(f0))
    f0["structurally_eligible"] = rng.integers(0, 2, len(f0))
    if seed % 2 == 0:
        f0 = f0[rng.random(len(f0)) < 0.6].copy()  # a sparse previous face list
    pos1 = pos0 + rng.normal(0, 0.4, pos0.shape)
    cells1 = np.concatenate([rng.integers(-14, 10, (1000, 3)), cells0[:50]])
    s, (order, b1, c1, f1) = run(ctx, pos1, cells1, previous=(b0, c0, f0), epoch=5, geom=geom)
    assert s == O.OK
    assert set(order["original_ordinal"]) == set(range(len(pos1)))


def test_dam_break_32_stepped_twice(ctx):
    sc = scenes.dam_break(32, seed=3)
    geom = O.geometry((0.0, 0.0, 0.0), sc.h)
    pos = np.stack([sc.positions[:, 0], sc.positions[:, 1], sc.positions[:, 2]], axis=1)
    cells = np.unique(np.floor(pos / sc.h).astype(np.int32), axis=0)
    s, (_, b0, c0, f0) = run(ctx, pos, cells, geom=geom)
    assert s == O.OK
    c0["previous_pressure"] = np.arange(len(c0), dtype=np.float64)
    f0["carrier_velocity"] = np.linspace(-1.0, 1.0, len(f0))
    moved = pos + np.array([0.0, -0.5 * sc.h, 0.25 * sc.h])
    cells1 = np.unique(np.floor(moved / sc.h).astype(np.int32), axis=0)
    s, (_, b1, c1, f1) = run(ctx, moved, cells1, previous=(b0, c0, f0), epoch=1, geom=geom)
    assert s == O.OK and len(c1) == len(cells1)
    print("topology", len(pos), "particles", len(b1), "blocks", len(c1), "cells", ctx.last.kernel_ms, "ms", flush=True)


def test_span_calls_invalidate_the_resident_scene(ctx):
    sc = scenes.dam_break(16, seed=1)
    geom = sc.geometry_record()
    assert ctx.scene_upload(geom, sc.positions, sc.velocities, sc.slopes, sc.mass) == H.OK
    assert ctx.scene_prepare(sc.pressure_cells(), sc.n, sc.rho) == H.OK
    opts = H.pressure_options(relative_target=1e-6, levels=3)
    assert ctx.step(sc.dt, sc.gravity, 2.0 ** -40, 2.0 ** -40, sc.count, opts)[0] == H.OK
    assert run(ctx, sc.positions[:100])[0] == O.OK  # reuses the scene's buffers
    assert ctx.step(sc.dt, sc.gravity, 2.0 ** -40, 2.0 ** -40, sc.count, opts)[0] == H.FAILED_PRECONDITION
    assert ctx.scene_prepare(sc.pressure_cells(), sc.n, sc.rho) == H.FAILED_PRECONDITION
    assert ctx.scene_upload(geom, sc.positions, sc.velocities, sc.slopes, sc.mass) == H.OK
    assert ctx.scene_prepare(sc.pressure_cells(), sc.n, sc.rho) == H.OK
    assert ctx.step(sc.dt, sc.gravity, 2.0 ** -40, 2.0 ** -40, sc.count, opts)[0] == H.OK
