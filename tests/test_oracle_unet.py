"""The oracle's learned-preconditioner network (FP64 U-Net on 8^3 tiles) pinned by finite differences.

The reference has no learned component; the forward pass is pinned by closed forms of its building
blocks (a zero network is zero; translation of a single impulse through a 1-layer receptive field) and
the backward pass by central finite differences of the loss in the parameters and the inputs.
"""
import numpy as np

import pyoracle as O


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


def test_parameter_count_matches_the_layer_table():
    layers = [(4, 8), (8, 8), (8, 16), (16, 16), (16, 32), (32, 32), (48, 16), (16, 16), (24, 8), (8, 1)]
    assert O.unet_param_count() == sum(co * ci * 27 + co for ci, co in layers)


def test_zero_weights_give_the_output_bias():
    p = np.zeros(O.unet_param_count())
    p[-1] = 0.25  # output bias
    x, _ = random_tiles(2, 1)
    assert np.array_equal(O.unet_forward(p, x), np.full((2, 512), 0.25))


def test_parameter_gradient_matches_central_differences():
    p = O.unet_init(3)
    p += np.random.default_rng(4).normal(0.0, 0.01, len(p))  # nonzero biases too
    x, t = random_tiles(2, 5)
    loss, g, _ = O.unet_loss_grad(p, x, t)
    rng = np.random.default_rng(6)
    idx = rng.choice(len(p), 60, replace=False)
    idx = np.concatenate([idx, [len(p) - 1, len(p) - 2]])
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (O.unet_loss_grad(pp, x, t)[0] - O.unet_loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(fd)), (i, fd, g[i])


def test_input_gradient_matches_central_differences():
    p = O.unet_init(7)
    x, t = random_tiles(1, 8)
    _, _, gi = O.unet_loss_grad(p, x, t, want_input_grad=True)
    rng = np.random.default_rng(9)
    for v in rng.choice(512, 20, replace=False):
        h = 1e-6
        xp, xm = x.copy(), x.copy()
        xp[0, 0, v] += h
        xm[0, 0, v] -= h
        fd = (O.unet_loss_grad(p, xp, t)[0] - O.unet_loss_grad(p, xm, t)[0]) / (2 * h)
        assert abs(fd - gi[0, 0, v]) <= 1e-6 * max(1e-3, abs(fd))


def test_single_layer_receptive_field_and_zero_padding():
    # Only the first layer's centre tap is nonzero and everything after it is the identity path of
    # channel 0 through the skip into the output layer: an impulse stays where it is.
    n = O.unet_param_count()
    p = np.zeros(n)
    layers = [(4, 8), (8, 8), (8, 16), (16, 16), (16, 32), (32, 32), (48, 16), (16, 16), (24, 8), (8, 1)]
    off = []
    q = 0
    for ci, co in layers:
        off.append(q)
        q += co * ci * 27 + co
    centre = 13
    p[off[0] + (0 * 4 + 0) * 27 + centre] = 1.0      # layer 0: out0 <- in0
    p[off[1] + (0 * 8 + 0) * 27 + centre] = 1.0      # layer 1: out0 <- in0 (skip s1 channel 0)
    p[off[8] + (0 * 24 + 16) * 27 + centre] = 1.0    # layer 8: out0 <- concat channel 16 = s1 channel 0
    p[off[9] + (0 * 8 + 0) * 27 + centre] = 1.0      # layer 9: out <- channel 0
    x = np.zeros((1, 4, 512))
    x[0, 0, 5 * 64 + 2 * 8 + 7] = 2.0
    out = O.unet_forward(p, x)
    want = np.zeros((1, 512))
    want[0, 5 * 64 + 2 * 8 + 7] = 2.0
    assert np.array_equal(out, want)
