"""The FP32-arithmetic deposit of the device-resident step (fl_set_transfer_precision(ctx, 32)).

Bar (the FP32 transfer tolerance): every face's mass and momentum quanta within 1e-5 of the largest
quanta of its channel against the oracle's FP64 deposit (DepositMacQuanta3D) of the same particles on
the same blocks, on a dam break at rest and with random velocities and slopes; the per-axis total mass
within the rounding of one contribution per particle sample; switching back to 64 restores the
bit-exact quanta.
"""
import numpy as np
import pytest

import cases as K
from fluidity_amd import hip as H
from fluidity_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(particles=1 << 20, blocks=1 << 14, lookup_blocks=1 << 18, pressure_cells=1 << 17,
                  pressure_faces=1 << 19)
    print("device", c.device_name(), flush=True)
    yield c
    c.close()


def step_quanta(ctx, sc, bits):
    assert ctx.scene_upload(sc.geometry_record(), sc.positions, sc.velocities, sc.slopes, sc.mass) == H.OK
    assert ctx.scene_prepare(sc.pressure_cells(), sc.n, sc.rho) == H.OK
    assert ctx.set_transfer_precision(bits) == H.OK
    opts = H.pressure_options(relative_target=1e-6, levels=3)
    s, rep = ctx.step(sc.dt, sc.gravity, K.QUANTUM, K.QUANTUM, sc.count, opts)
    assert s == H.OK, s
    s, out = ctx.scene_download(int(rep.blocks), sc.count, int(rep.pressure_cells), what=("blocks", "quanta"))
    assert s == H.OK
    assert ctx.set_transfer_precision(64) == H.OK
    return out


@pytest.mark.parametrize("random_velocity", [False, True])
def test_fp32_deposit_agrees_with_fp64(ctx, random_velocity):
    sc = scenes.dam_break(32, seed=2, random_velocity=random_velocity)
    st = K.oracle_step_inputs(sc)
    out = step_quanta(ctx, sc, 32)
    assert np.array_equal(out["blocks"].view(np.int32), st["blocks"].view(np.int32))
    for ch in ("mass_quanta_sum", "pure_momentum_quanta_sum"):
        got = out["quanta"][ch].astype(np.float64)
        want = st["quanta"][ch].astype(np.float64)
        scale = np.abs(want).max()
        err = np.abs(got - want).max()
        print(ch, "max error", err, "max quanta", scale, flush=True)
        if scale == 0:
            assert err == 0
        else:
            assert err <= 1e-5 * scale
    axis_of = (np.arange(len(out["quanta"])) % 192) // 64
    for a in range(3):
        got = int(out["quanta"]["mass_quanta_sum"][axis_of == a].sum())
        want = int(st["quanta"]["mass_quanta_sum"][axis_of == a].sum())
        assert abs(got - want) <= 27 * sc.count * (1 + 1e-6 * sc.mass / K.QUANTUM)
    back = step_quanta(ctx, sc, 64)
    assert np.array_equal(back["quanta"].view(np.int64), st["quanta"].view(np.int64))


def test_precision_values(ctx):
    assert ctx.set_transfer_precision(16) == H.INVALID_ARGUMENT
    assert ctx.set_transfer_precision(32) == H.OK
    assert ctx.set_transfer_precision(64) == H.OK
