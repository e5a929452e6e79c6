"""GPU learned preconditioner: network parity with the FP64 oracle, training, and the flexible CG.

Bars: the FP32 forward pass within 1e-4 of the FP64 oracle's (relative to the output range) for the
same FP32-representable parameters; the loss within 1e-4 relative and every parameter gradient within
2e-3 of the largest gradient of its layer (FP32 accumulation over up to 27 x 48 x 512 terms); training
lowers the loss; a projection preconditioned by the trained network is accepted by the same FP64 gates
and agrees with the FP64 oracle to 1e-5; and training generalises: on V-cycle corrections of held-out
scenes the trained network's loss is below the untrained network's (the A/B test of the learning).
"""
import numpy as np
import pytest

import cases as K
import pyoracle as O
from fluidity_amd import hip as H
from fluidity_amd import scenes

pytestmark = pytest.mark.gpu

LAYERS = [(4, 8), (8, 8), (8, 16), (16, 16), (16, 32), (32, 32), (48, 16), (16, 16), (24, 8), (8, 1)]


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(particles=1 << 20, blocks=1 << 14, lookup_blocks=1 << 20, pressure_cells=1 << 18,
                  pressure_faces=1 << 20)
    print("device", c.device_name(), flush=True)
    yield c
    c.close()


def random_tiles(n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 4, 512))
    liquid = (rng.random((n, 512)) < 0.7).astype(float)
    x[:, 0] = rng.normal(0.0, 1.0, (n, 512)) * liquid
    x[:, 1] = liquid
    solid = (1 - liquid) * (rng.random((n, 512)) < 0.3)
    x[:, 3] = solid
    x[:, 2] = 1 - liquid - solid
    return x, rng.normal(0.0, 1.0, (n, 512))


def layer_slices():
    out, q = [], 0
    for ci, co in LAYERS:
        out.append(slice(q, q + co * ci * 27 + co))
        q += co * ci * 27 + co
    return out


def test_network_forward_and_gradients_match_the_oracle(ctx):
    net = H.Network(ctx, tiles=64, pairs=64)
    assert net.count == O.unet_param_count()
    p = O.unet_init(11).astype(np.float32)
    p[np.arange(len(p)) % 97 == 0] += 0.05  # nonzero biases and offsets
    assert net.set_params(p) == H.OK
    assert np.array_equal(net.params(), p)
    x, t = random_tiles(5, 12)
    x32, t32 = x.astype(np.float32), t.astype(np.float32)
    s, out = net.forward(x32)
    assert s == H.OK
    want = O.unet_forward(p.astype(np.float64), x32.astype(np.float64))
    assert np.abs(out - want).max() <= 1e-4 * np.abs(want).max()
    # Inference precision: BF16 operands on the matrix cores for the 8^3 layers, FP32 accumulation.
    s, fast = net.infer(x32)
    assert s == H.OK
    assert np.abs(fast - want).max() <= 3e-2 * np.abs(want).max()
    s, loss, gp, gi = net.loss_grad(x32, t32)
    assert s == H.OK
    wl, wg, wi = O.unet_loss_grad(p.astype(np.float64), x32.astype(np.float64), t32.astype(np.float64), True)
    assert abs(loss - wl) <= 1e-4 * abs(wl)
    for sl in layer_slices():
        scale = np.abs(wg[sl]).max()
        assert np.abs(gp[sl] - wg[sl]).max() <= 2e-3 * scale
    assert np.abs(gi - wi).max() <= 2e-3 * np.abs(wi).max()


def test_training_lowers_the_loss(ctx):
    # Targets: a fixed linear smoothing of the residual channel, learnable by the network.
    net = H.Network(ctx, tiles=256, pairs=1024)
    assert net.set_params(O.unet_init(1).astype(np.float32)) == H.OK
    x, _ = random_tiles(512, 2)
    r = x[:, 0].reshape(-1, 8, 8, 8)
    t = r.copy()
    for a in (1, 2, 3):
        t = t + 0.5 * (np.roll(r, 1, a) + np.roll(r, -1, a))
    t = (t.reshape(-1, 512) * x[:, 1]).astype(np.float32)
    assert net.add_pairs(x.astype(np.float32), t) == H.OK
    assert net.pairs() == 512
    s, losses = net.train(epochs=8, batch=128, lr=3e-3, seed=5)
    assert s == H.OK
    print("epoch losses", losses, flush=True)
    assert np.all(np.isfinite(losses)) and losses[-1] < 0.5 * losses[0]


def scene_problem(seed):
    sc = scenes.dam_break(32, seed=seed, random_velocity=True)
    st = K.oracle_step_inputs(sc)
    return st, (st["geom"], st["blocks"], st["dt"], st["cells"], st["faces"])


def test_learned_flexible_cg_meets_the_gates_and_training_generalises(ctx):
    net = H.Network(ctx, tiles=1024, pairs=40000)
    init = O.unet_init(3).astype(np.float32)
    assert net.set_params(init) == H.OK
    st, args = scene_problem(21)
    s, untrained = ctx.project(*args, H.pressure_options(relative_target=1e-6, levels=4, learned=True))
    assert s == H.OK
    # Held-out pairs (seeds 7, 8) and training pairs (seeds 1..6) from plain MGPCG solves.
    def collect(seeds):
        assert net.collect(1, clear=True) == H.OK
        for seed in seeds:
            _, a = scene_problem(seed)
            assert ctx.project(*a, H.pressure_options(relative_target=1e-6, levels=4))[0] == H.OK
        assert net.collect(0) == H.OK
        return net.pairs()

    assert collect((7, 8)) > 0
    s, held_before = net.eval()
    assert s == H.OK
    assert collect((1, 2, 3, 4, 5, 6)) > 0
    s, losses = net.train(epochs=16, batch=256, lr=2e-3, seed=9)
    assert s == H.OK
    print("pairs", net.pairs(), "losses", losses, flush=True)
    trained_params = net.params()
    assert collect((7, 8)) > 0
    s, held_after = net.eval()
    assert s == H.OK
    print("held-out loss before", held_before, "after", held_after, flush=True)
    # A/B of training itself: the trained network predicts unseen V-cycle corrections better.
    assert held_after < held_before
    assert np.array_equal(net.params(), trained_params)
    s, trained = ctx.project(*args, H.pressure_options(relative_target=1e-6, levels=4, learned=True))
    assert s == H.OK
    s, plain = ctx.project(*args, H.pressure_options(relative_target=1e-6, levels=4))
    assert s == H.OK
    s, ref = O.project(*args, O.pressure_options(relative_target=1e-12, fp32=False, levels=4))
    assert s == O.OK
    for name, out in (("untrained", untrained), ("trained", trained), ("mgpcg", plain)):
        d = out["diagnostics"]
        p_err = np.abs(out["pressures"] - ref["pressures"]).max() / np.abs(ref["pressures"]).max()
        print(name, "iterations", d["iteration_count"], "learned", d["learned_iterations"], "p_err", p_err, flush=True)
        # Every accepted projection passes the same FP64 gates whatever preconditioned it.
        assert d["defect_l2"] <= 1e-6 * d["initial_defect_l2"] and p_err <= 1e-5
    assert trained["diagnostics"]["learned_iterations"] > 0


def test_learned_flexible_cg_at_128_meets_the_gates_and_the_fp64_oracle():
    # The learned preconditioner at a benchmark size (dam break, 128^3 sparse, 502,656 liquid cells): the
    # network is trained on pairs logged from the plain MGPCG solve of the same step, then the flexible CG
    # with the network (V-cycle fallback after stagnation) projects the step. The accepted projection meets
    # the FP64 gates and agrees with a tight FP64 oracle solve of the same frozen system.
    sc = scenes.by_name("dam_break_128")
    cells = sc.pressure_cells()
    with H.Context(particles=sc.count + 1024, blocks=1 << 16, lookup_blocks=1 << 21, pressure_cells=len(cells) + 64,
                   pressure_faces=4 * len(cells) + 4096) as c:
        assert c.scene_upload(sc.geometry_record(), sc.positions, sc.velocities, sc.slopes, sc.mass) == H.OK
        assert c.scene_prepare(cells, sc.n, sc.rho) == H.OK
        net = H.Network(c, tiles=4096, pairs=24000)
        assert net.set_params(O.unet_init(4).astype(np.float32)) == H.OK
        opts = H.pressure_options(relative_target=1e-6, levels=5, coarse_sweeps=2, bf16_levels=True, smoother=2)
        assert net.collect(1, clear=True) == H.OK
        s, plain = c.step(sc.dt, sc.gravity, K.QUANTUM, K.QUANTUM, sc.count, opts)
        assert s == H.OK, s
        assert net.collect(0) == H.OK
        print("pairs", net.pairs(), flush=True)
        assert net.pairs() > 0
        s, losses = net.train(epochs=4, batch=256, lr=2e-3, seed=3)
        assert s == H.OK and np.all(np.isfinite(losses))
        learned_opts = opts.copy()
        learned_opts["preconditioner"] = 1
        s, rep = c.step(sc.dt, sc.gravity, K.QUANTUM, K.QUANTUM, sc.count, learned_opts)
        assert s == H.OK, s
        d = rep.diagnostics()
        print("learned", rep.projection_ms, "ms", d["iteration_count"], "iterations (", d["learned_iterations"],
              "learned ); mgpcg", plain.projection_ms, "ms", plain.diagnostics()["iteration_count"], flush=True)
        assert d["learned_iterations"] > 0
        assert d["defect_l2"] <= 1e-6 * d["initial_defect_l2"]
        assert d["max_cell_constraint_defect"] <= 2e-5 and d["max_volume_scaled_residual"] <= 2e-5
        args, p, u = K.device_step_system(c, sc, rep)
        K.assert_matches_fp64_oracle(args, p, u)
