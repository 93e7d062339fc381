"""Pins of the CPU oracle of the Poisson-guess network (oracle/cnn.py) that
do not need a GPU: finite differences, BF16 rounding against torch, direct
convolution sums, shift equivariance, Adam's closed-form first step and the
x-space DoF <-> image placement."""

import numpy as np
import pytest

import oracle as O
from oracle import cnn as C


def _params(seed=7):
    rng = np.random.default_rng(seed)
    p = np.zeros(C.n_params(), dtype=np.float64)
    o = 0
    for ci, co in C.layer_shapes():
        p[o:o + co * ci * 9] = rng.uniform(-1, 1, co * ci * 9) / np.sqrt(ci * 9)
        o += co * ci * 9
        p[o:o + co] = rng.uniform(-0.1, 0.1, co)
        o += co
    return p


def test_parameter_count_and_layout():
    assert C.n_params() == (9 * 32 + 32) + 2 * (32 * 32 * 9 + 32) + (32 * 9 + 1)
    p = np.arange(C.n_params(), dtype=np.float64)
    layers = C.unpack(p)
    assert layers[0][0].shape == (32, 1, 3, 3) and layers[3][0].shape == \
        (1, 32, 3, 3)
    # canonical order: w of layer 0, b of layer 0, w of layer 1, ...
    assert layers[0][1][0] == 288 and layers[1][0][0, 0, 0, 0] == 320


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) *
        np.float32(10.0) ** rng.integers(-30, 30, 20000).astype(np.float32),
        # exact ties between two BF16 values, both parities
        (np.arange(1, 2001, dtype=np.uint32) * 0x10000 + 0x8000 +
         0x3f800000).view(np.float32),
        np.array([0.0, -0.0, np.inf, -np.inf, 3.0e38, -3.0e38],
                 dtype=np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = C.bf16(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.isnan(C.bf16(np.array([np.nan], dtype=np.float32)))[0]


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 3, 4, 5))
    w = rng.standard_normal((2, 3, 3, 3))
    b = rng.standard_normal(2)
    out = C.conv(x, w, b)
    B, Ci, H, W = x.shape
    ref = np.zeros((B, 2, H, W))
    for bb in range(B):
        for o in range(2):
            for y in range(H):
                for xx in range(W):
                    s = b[o]
                    for i in range(Ci):
                        for ky in range(3):
                            for kx in range(3):
                                s += w[o, i, ky, kx] * \
                                    x[bb, i, (y + ky - 1) % H, (xx + kx - 1) % W]
                    ref[bb, o, y, xx] = s
    assert np.allclose(out, ref, rtol=1e-13, atol=1e-13)


def test_network_commutes_with_periodic_shifts():
    p = _params()
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 16, 16))
    out = C.forward(p, x, dtype=np.float64, round_bf16=False)
    xs = np.roll(x, shift=(3, -5), axis=(1, 2))
    outs = C.forward(p, xs, dtype=np.float64, round_bf16=False)
    assert np.allclose(outs, np.roll(out, shift=(3, -5), axis=(1, 2)),
                       rtol=1e-12, atol=1e-12)


def test_backward_matches_finite_differences():
    p = _params()
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 8, 8))
    t = rng.standard_normal((2, 8, 8))
    loss, g = C.loss_and_grad(p, x, t, dtype=np.float64, round_bf16=False)
    # a few entries of every weight and bias block, plus random others
    idx = []
    o = 0
    for ci, co in C.layer_shapes():
        idx += [o, o + co * ci * 9 - 1, o + co * ci * 9, o + co * ci * 9 + co - 1]
        o += co * ci * 9 + co
    idx += list(rng.integers(0, C.n_params(), 24))
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp = C.loss_and_grad(pp, x, t, dtype=np.float64, round_bf16=False)[0]
        lm = C.loss_and_grad(pm, x, t, dtype=np.float64, round_bf16=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_adam_first_steps():
    rng = np.random.default_rng(5)
    p = rng.standard_normal(100).astype(np.float32)
    g = rng.standard_normal(100).astype(np.float32)
    opt = C.Adam(100)
    p1 = opt.step(p, g, 1e-2)
    # first step: m/c1 = g, v/c2 = g^2 -> p - lr g / (|g| + eps)
    assert np.allclose(p1, p - 1e-2 * np.sign(g), rtol=0, atol=1e-6)
    p2 = opt.step(p1, g, 1e-2)
    assert np.allclose(p2, p1 - 1e-2 * np.sign(g), rtol=0, atol=1e-6)
    p3 = opt.step(p2, -g, 1e-2)
    # m = 0.9 (0.1 g + 0.09 g) ... closed form of the moments for +g, +g, -g
    m = 0.1 * (-g) + 0.9 * (0.1 * g + 0.9 * 0.1 * g)
    v = 0.001 * g * g * (1 + 0.999 + 0.999 ** 2)
    ref = p2 - 1e-2 * (m / (1 - 0.9 ** 3)) / (np.sqrt(v / (1 - 0.999 ** 3)) + 1e-8)
    assert np.allclose(p3, ref, rtol=0, atol=1e-6)


def test_initial_parameters_follow_splitmix():
    p = C.init_params(11, O.splitmix_uniform)
    layers = C.unpack(p)
    lim0 = np.sqrt(6.0 / 9)
    assert np.isclose(layers[0][0].ravel()[0], lim0 * O.splitmix_uniform(11, 0),
                      rtol=1e-7)
    assert np.all(np.abs(layers[3][0]) <= 0.1 * np.sqrt(6.0 / 288) + 1e-7)
    assert all(np.all(b == 0) for _, b in layers)


def test_dofs_image_placement():
    n0, n1, N = 3, 2, 4
    # node (cx, cy, ix, iy) -> pixel (cy N + iy, cx N + ix)
    u = np.zeros(n0 * n1 * N * N)
    for cy in range(n1):
        for cx in range(n0):
            for iy in range(N):
                for ix in range(N):
                    dof = (cx + n0 * cy) * N * N + ix + N * iy
                    u[dof] = 1000 * (cy * N + iy) + (cx * N + ix)
    img, rms = C.dofs_to_image(u, n0, n1, N)
    Y, X = np.mgrid[0:n1 * N, 0:n0 * N]
    ref = 1000.0 * Y + X
    ref = (ref - ref.mean()) / np.sqrt(np.mean((ref - ref.mean()) ** 2))
    assert np.allclose(img, ref, rtol=1e-6, atol=1e-6)
    back = C.image_to_dofs(img, n0, n1, N, rms)
    assert np.allclose(back, u - u.mean(), rtol=1e-6, atol=1e-3)
