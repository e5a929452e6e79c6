"""Oracle: topology preparation with history remap (flo_prepare_topology), pinned by the known answers
of native/tests/test_liquid_step_workspace_3d.cc, the record checks of ValidateLiquidState3D
(native/src/liquid_state_3d.cc:559-604) and a dictionary remap of random histories.
"""
import math

import numpy as np
import pytest

import cases as K
import pyoracle as O

UNIT = O.geometry((0.0, 0.0, 0.0), 1.0)
SMALL = O.topology_capacities(particles=16, blocks=32, cells=16, faces=32 * 192)


def prepare(positions=(), cells=(), previous=None, caps=SMALL, geom=UNIT):
    return O.prepare_topology(geom, np.array(positions, dtype=np.float64).reshape(-1, 3), O.int3(list(cells)),
                              previous, caps)


def same(a, b):
    """Field-by-field equality of record arrays (padding bytes are not part of a record's value)."""
    assert a.shape == b.shape
    for name in a.dtype.names:
        x, y = a[name], b[name]
        if x.dtype.names:
            same(x, y)
        elif x.dtype.kind == "f":
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), name
        else:
            assert np.array_equal(x, y), name


def keys(records, field):
    return [tuple(int(v) for v in r) for r in records[field][["x", "y", "z"]]]


def history(geom, epoch, blocks, cells, faces):
    h = O.topology_history(geom, epoch, blocks, cells, faces)
    h._keep = (blocks, cells, faces)  # the arrays must outlive the structure
    return h


def test_sorted_order_and_negative_bins():
    # NegativeBinsAndEqualBinsPreserveIncomingOrder (test_liquid_step_workspace_3d.cc:52-73)
    pos = [(2.0, 2.0, 2.0), (-2.0, 2.0, 2.0), (2.1, 2.0, 2.0), (2.0, -2.0, 2.0), (2.0, 2.0, -2.0), (-2.1, 2.0, 2.0)]
    s, order, blocks, cells, faces = prepare(pos)
    assert s == O.OK
    assert list(order["original_ordinal"]) == [4, 3, 1, 5, 0, 2]
    assert tuple(order["bin_index"][0]) == (0, 0, -1) and tuple(order["bin_index"][2]) == (-1, 0, 0)
    assert len(faces) == 192 * len(blocks) and len(cells) == 0
    assert (blocks["generation"] == 1).all()


def test_zero_weight_support_owners_and_face_records():
    # ActivatesZeroWeightOriginalNodesRegardlessOfTag (test_liquid_step_workspace_3d.cc:75-105)
    s, _, blocks, _, faces = prepare([(3.0, 3.0, 3.0)])
    assert s == O.OK
    assert keys(blocks, "block_index") == [(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)]
    assert len(faces) == 8 * 192
    assert (faces["face_slot"] == np.arange(8 * 192)).all()
    assert (faces["validity"] == 0).all() and (faces["carrier_velocity"] == 0).all()
    assert (faces["structurally_eligible"] == 1).all() and (faces["side_tag"] == 0).all()
    # Every face of the three 27-sample stencils [2,4]^3 has an owned slot.
    for slot, f in enumerate(faces):
        assert O.slot_face(K.sorted_blocks(UNIT, keys(blocks, "block_index")), slot) == \
            (O.OK, (int(f["face_index"]["axis"]), tuple(int(v) for v in f["face_index"]["lattice"])))


def test_pressure_face_owners_and_canonical_unique_cells():
    # ActivatesSixPressureFaceOwnersAndCanonicalUniqueCells (test_liquid_step_workspace_3d.cc:107-140)
    s, _, blocks, cells, faces = prepare(cells=[(3, 3, 3), (-1, -1, -1), (3, 3, 3)])
    assert s == O.OK
    assert keys(blocks, "block_index") == [(-1, -1, -1), (0, -1, -1), (-1, 0, -1), (-1, -1, 0), (0, 0, 0), (1, 0, 0),
                                           (0, 1, 0), (0, 0, 1)]
    assert keys(cells, "cell_index") == [(-1, -1, -1), (3, 3, 3)]
    assert (cells["occupancy"] == 1).all() and (cells["previous_pressure"] == 0).all()
    faces_present = {(int(f["face_index"]["axis"]), tuple(int(v) for v in f["face_index"]["lattice"])) for f in faces}
    for face in [(0, (-1, -1, -1)), (0, (0, -1, -1)), (1, (-1, -1, -1)), (1, (-1, 0, -1)), (2, (-1, -1, -1)),
                 (2, (-1, -1, 0))]:
        assert face in faces_present


def test_history_remaps_by_identity_across_shifted_block_ordinals():
    # RemapsHistoryByIdentityAcrossShiftedBlockOrdinals (test_liquid_step_workspace_3d.cc:142-190)
    s, _, blocks, cells, faces = prepare(cells=[(1, 1, 1), (9, 1, 1)])
    assert s == O.OK and keys(blocks, "block_index") == [(0, 0, 0), (2, 0, 0)]
    blocks[0]["generation"], blocks[0]["activation_age"], blocks[0]["inactive_age"] = 4, 7, 2
    cells[0]["previous_pressure"], cells[0]["occupancy"] = 123.5, 2
    faces["carrier_velocity"] = 10.0 + faces["face_slot"].astype(np.float64)
    faces["validity"], faces["side_tag"], faces["structurally_eligible"] = 3, -17, 0
    snapshot = (blocks.copy(), cells.copy(), faces.copy())
    prev = history(UNIT, 11, blocks, cells, faces)
    s, _, nb, nc, nf = prepare(cells=[(-3, 1, 1), (1, 1, 1)], previous=prev)
    assert s == O.OK
    assert keys(nb, "block_index") == [(-1, 0, 0), (0, 0, 0)]
    assert nb[0]["generation"] == 12 and nb[0]["activation_age"] == 0
    assert nb[1]["generation"] == 4 and nb[1]["activation_age"] == 7 and nb[1]["inactive_age"] == 0
    assert nc[0]["previous_pressure"] == 0 and nc[0]["occupancy"] == 1
    assert nc[1]["previous_pressure"] == 123.5 and nc[1]["occupancy"] == 2
    assert (nf["carrier_velocity"][:192] == 0).all() and (nf["validity"][:192] == 0).all()
    expected = faces[:192].copy()
    expected["face_slot"] += 192
    same(nf[192:], expected)
    for a, b in zip(snapshot, (blocks, cells, faces)):  # the previous state is not modified
        same(a, b)


def test_capacity_failures():
    # CapacityFailuresPreservePublishedValuesAndAddresses (test_liquid_step_workspace_3d.cc:192-225)
    caps = O.topology_capacities(particles=1, blocks=1, cells=1, faces=192)
    assert prepare(cells=[(1, 1, 1)], caps=caps)[0] == O.OK
    assert prepare([(3, 3, 3)], caps=caps)[0] == O.RESOURCE_EXHAUSTED
    assert prepare([(2, 2, 2), (2, 2, 2)], caps=caps)[0] == O.RESOURCE_EXHAUSTED
    assert prepare(cells=[(1, 1, 1), (2, 1, 1)], caps=caps)[0] == O.RESOURCE_EXHAUSTED
    # FaceCapacityIsCheckedBeforePublish (:227-235)
    caps = O.topology_capacities(particles=16, blocks=32, cells=16, faces=191)
    assert prepare(cells=[(1, 1, 1)], caps=caps)[0] == O.RESOURCE_EXHAUSTED


def test_lattice_generation_and_input_errors():
    # RejectsLatticeAndGenerationOverflowAndInvalidInputTransactionally (test_liquid_step_workspace_3d.cc:237-258)
    assert prepare(cells=[(2147483647, 0, 0)])[0] == O.OUT_OF_RANGE
    assert prepare([(2147483647.0, 0.0, 0.0)])[0] == O.OUT_OF_RANGE
    assert prepare([(math.nan, 0.0, 0.0)])[0] == O.INVALID_ARGUMENT
    empty = (np.zeros(0, O.BLOCK_STATE), np.zeros(0, O.CELL_STATE), np.zeros(0, O.FACE_STATE))
    assert prepare(previous=history(UNIT, 2 ** 64 - 1, *empty))[0] == O.OUT_OF_RANGE
    assert prepare(previous=history(O.geometry((0.0, 0.0, 0.0), 2.0), 0, *empty))[0] == O.INVALID_ARGUMENT
    # A particle error comes before a cell error; both come after the history checks.
    assert prepare([(math.nan, 0.0, 0.0)], cells=[(2147483647, 0, 0)])[0] == O.INVALID_ARGUMENT
    assert prepare([(math.nan, 0.0, 0.0)], previous=history(UNIT, 2 ** 64 - 1, *empty))[0] == O.OUT_OF_RANGE


def test_empty_preparation_and_reactivation_start_fresh_history():
    # EmptyPreparationClearsExtentsAndReactivationInitializesHistory (test_liquid_step_workspace_3d.cc:296-320)
    s, _, blocks, cells, faces = prepare(cells=[(1, 1, 1)])
    cells[0]["previous_pressure"] = 99
    faces[0]["carrier_velocity"] = 23
    s, _, nb, nc, nf = prepare(previous=history(UNIT, 0, blocks, cells, faces))
    assert s == O.OK and len(nb) == 0 and len(nf) == 0 and len(nc) == 0
    empty = (np.zeros(0, O.BLOCK_STATE), np.zeros(0, O.CELL_STATE), np.zeros(0, O.FACE_STATE))
    s, _, nb, nc, nf = prepare(cells=[(1, 1, 1)], previous=history(UNIT, 9, *empty))
    assert s == O.OK and nb[0]["generation"] == 10
    assert nc[0]["previous_pressure"] == 0 and nf[0]["carrier_velocity"] == 0 and nf[0]["validity"] == 0


def valid_history():
    s, _, blocks, cells, faces = prepare([(1.5, 1.5, 1.5), (6.5, 1.5, 1.5)], cells=[(1, 1, 1), (6, 1, 1)])
    assert s == O.OK
    return blocks, cells, faces


@pytest.mark.parametrize("mutate,code", [
    (lambda b, c, f: b.__setitem__(0, b[1]), O.INVALID_ARGUMENT),                       # duplicate block
    (lambda b, c, f: b[0]["block_index"].__setitem__("x", 536870912), O.OUT_OF_RANGE),  # block beyond the lattice
    (lambda b, c, f: c.__setitem__(0, c[1]), O.INVALID_ARGUMENT),                       # cells out of order
    (lambda b, c, f: c[0]["cell_index"].__setitem__("z", 400), O.INVALID_ARGUMENT),     # cell without an owner
    (lambda b, c, f: c.__setitem__("occupancy", 3), O.INVALID_ARGUMENT),
    (lambda b, c, f: c.__setitem__("previous_pressure", math.inf), O.INVALID_ARGUMENT),
    (lambda b, c, f: f[5].__setitem__("validity", 4), O.INVALID_ARGUMENT),
    (lambda b, c, f: f[5].__setitem__("carrier_velocity", math.nan), O.INVALID_ARGUMENT),
    (lambda b, c, f: f[5]["face_index"].__setitem__("axis", 3), O.INVALID_ARGUMENT),
    (lambda b, c, f: f[5].__setitem__("face_slot", 6), O.INVALID_ARGUMENT),             # slot of another face
    (lambda b, c, f: f.__setitem__(slice(5, 7), f[5:7][::-1].copy()), O.INVALID_ARGUMENT),  # decreasing slots
], ids=["dup_block", "block_range", "cell_order", "cell_owner", "occupancy", "pressure", "validity", "carrier",
        "axis", "slot", "slot_order"])
def test_history_record_checks(mutate, code):
    blocks, cells, faces = valid_history()
    assert prepare(cells=[(1, 1, 1)], previous=history(UNIT, 1, blocks, cells, faces))[0] == O.OK
    mutate(blocks, cells, faces)
    assert prepare(cells=[(1, 1, 1)], previous=history(UNIT, 1, blocks, cells, faces))[0] == code


def test_history_checks_run_in_record_order():
    blocks, cells, faces = valid_history()
    cells[0]["occupancy"] = 7                       # InvalidArgument at a cell ...
    blocks[1]["block_index"]["y"] = -536870913      # ... but a block range error is found first
    assert prepare(previous=history(UNIT, 1, blocks, cells, faces))[0] == O.OUT_OF_RANGE


def test_sparse_previous_faces_are_found_by_identity():
    # The previous face list may be a subset of its slots (ascending, each at its own slot).
    blocks, cells, faces = valid_history()
    faces["carrier_velocity"] = np.arange(len(faces), dtype=np.float64)
    keep = faces[::7].copy()
    s, _, nb, nc, nf = prepare(cells=[(1, 1, 1), (6, 1, 1)], previous=history(UNIT, 3, blocks, cells, keep))
    assert s == O.OK
    kept = {int(f["face_slot"]) for f in keep}
    for f in nf:
        slot = int(f["face_slot"])
        assert f["carrier_velocity"] == (float(slot) if slot in kept else 0.0)


def dict_remap(geom, positions, cells, prev, epoch):
    """Remap by hash lookups on identities, an independent restatement of the remap rules."""
    s, _, blocks, ncells, faces = O.prepare_topology(geom, positions, cells, None, O.topology_capacities())
    assert s == O.OK
    pb = {tuple(int(v) for v in b["block_index"]): b for b in prev[0]}
    pc = {tuple(int(v) for v in c["cell_index"]): c for c in prev[1]}
    pf = {(int(f["face_index"]["axis"]), tuple(int(v) for v in f["face_index"]["lattice"])): f for f in prev[2]}
    for b in blocks:
        old = pb.get(tuple(int(v) for v in b["block_index"]))
        if old is None:
            b["generation"] = epoch + 1
        else:
            b["generation"], b["activation_age"], b["inactive_age"] = old["generation"], old["activation_age"], 0
    for c in ncells:
        old = pc.get(tuple(int(v) for v in c["cell_index"]))
        if old is not None:
            c["occupancy"], c["previous_pressure"] = old["occupancy"], old["previous_pressure"]
    for f in faces:
        old = pf.get((int(f["face_index"]["axis"]), tuple(int(v) for v in f["face_index"]["lattice"])))
        if old is not None:
            for k in ("carrier_velocity", "validity", "side_tag", "structurally_eligible"):
                f[k] = old[k]
    return blocks, ncells, faces


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_histories_match_a_dictionary_remap(seed):
    rng = np.random.default_rng(seed)
    geom = O.geometry((-0.37, 0.11, 0.5), 0.25)
    pos0 = rng.uniform(-2.0, 2.0, (400, 3))
    cells0 = O.int3([tuple(v) for v in rng.integers(-12, 12, (150, 3))])
    s, _, b0, c0, f0 = O.prepare_topology(geom, pos0, cells0, None, O.topology_capacities())
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
    pos1 = pos0 + rng.normal(0, 0.4, pos0.shape)
    cells1 = O.int3([tuple(v) for v in rng.integers(-14, 10, (170, 3))])
    s, order, b1, c1, f1 = O.prepare_topology(geom, pos1, cells1, history(geom, 5, b0, c0, f0), O.topology_capacities())
    assert s == O.OK
    wb, wc, wf = dict_remap(geom, pos1, cells1, (b0, c0, f0), 5)
    same(b1, wb)
    same(c1, wc)
    same(f1, wf)
    assert set(order["original_ordinal"]) == set(range(len(pos1)))
