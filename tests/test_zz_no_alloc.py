"""fl_normalize_forces_static (context-free, allocation-free) against fl_normalize_forces and the oracle.

Two calls in a row through the C-ABI, the second with fewer faces than the first, each compared bit for
bit with the context entry point on the same quanta; the oversize and bad-input statuses; and a good call
after a failed one (the shared error words start over on every call).
"""
import numpy as np
import pytest

import cases as K
import pyoracle as O
from fluidity_amd import hip as H
from fluidity_amd import records as R
from fluidity_amd import scenes

pytestmark = pytest.mark.gpu


def _quanta(n):
    """Quanta of a random-velocity dam break with about n owned face slots (a multiple of 192)."""
    sc = scenes.dam_break(n, seed=3, random_velocity=True)
    st = K.oracle_step_inputs(sc)
    return sc, st


def _same(a, b):
    return all(np.array_equal(x.view(np.uint8), y.view(np.uint8)) for x, y in zip(a, b))


def test_static_matches_context_entry_point_on_two_calls():
    sc, st = _quanta(32)
    q = st["quanta"].view(R.FACE_QUANTA)
    assert len(q) % 192 == 0 and 192 < len(q) <= H.STATIC_MAX_FACES, len(q)
    print("faces", len(q), flush=True)
    with H.Context(particles=1 << 16, blocks=1 << 12, lookup_blocks=1 << 16) as ctx:
        for nf in (len(q), len(q) // 192 // 3 * 192):  # the second call is smaller than the first
            part = np.ascontiguousarray(q[:nf])
            s, *got = H.normalize_forces_static(0, part, K.QUANTUM, K.QUANTUM, 0.0, sc.gravity, st["dt"])
            assert s == H.OK, s
            s2, *want = ctx.normalize_forces(part, K.QUANTUM, K.QUANTUM, 0.0, sc.gravity, st["dt"])
            assert s2 == H.OK
            assert _same(got, want), nf
            # and the oracle (the first call covers every face of the scene)
            so, mass, u0, validity = O.normalize(part.view(O.FACE_QUANTA), K.QUANTUM, K.QUANTUM, 0.0)
            assert so == O.OK
            assert np.array_equal(got[0], mass) and np.array_equal(got[1], u0) and np.array_equal(got[2], validity)
            print("static call with", nf, "faces matches", flush=True)


def test_static_statuses_and_untouched_outputs():
    q = np.zeros(192, R.FACE_QUANTA)
    q[5] = (3, 1, 0)
    q[9]["stress_force_quanta_sum"] = 1
    s, *_ = H.normalize_forces_static(0, q, 1.0, 1.0, 0.0, (0.0, -9.81, 0.0), 0.1)
    assert s == H.INVALID_ARGUMENT
    q[9]["stress_force_quanta_sum"] = 0
    s, m, u, v, p = H.normalize_forces_static(0, q, 1.0, 1.0, 0.0, (0.0, -9.81, 0.0), 0.1)
    assert s == H.OK and u[5] == 1.0 / 3.0 and v[5] == 1 and p[5] == 1.0 / 3.0 and m[5] == 3.0
    # more faces than the static array holds: refused before any device work, outputs untouched
    big = np.zeros(H.STATIC_MAX_FACES + 192, R.FACE_QUANTA)
    s, m, *_ = H.normalize_forces_static(0, big, 1.0, 1.0, 0.0, (0.0, -9.81, 0.0), 0.1)
    assert s == H.RESOURCE_EXHAUSTED and not m.any()
    # a quantum that is not a power of two, and a face count that is not a multiple of 192
    assert H.normalize_forces_static(0, q, 3.0, 1.0, 0.0, (0.0, 0.0, 0.0), 0.1)[0] == H.INVALID_ARGUMENT
    assert H.normalize_forces_static(0, q[:100].copy(), 1.0, 1.0, 0.0, (0.0, 0.0, 0.0), 0.1)[0] == H.INVALID_ARGUMENT
    # an overflowing prediction is OutOfRange, as through the context entry point
    s, *_ = H.normalize_forces_static(0, q, 1.0, 1.0, 0.0, (0.0, 1e308, 0.0), 1e10)
    assert s == H.OUT_OF_RANGE
    # a device that does not exist
    assert H.normalize_forces_static(1 << 20, q, 1.0, 1.0, 0.0, (0.0, 0.0, 0.0), 0.1)[0] == H.FAILED_PRECONDITION
    # and a good call after the failures
    s, m2, u2, v2, p2 = H.normalize_forces_static(0, q, 1.0, 1.0, 0.0, (0.0, -9.81, 0.0), 0.1)
    assert s == H.OK and np.array_equal(u2, u) and np.array_equal(p2, p)
