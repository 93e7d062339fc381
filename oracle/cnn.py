"""CPU oracle of the Poisson-guess network (numpy).

TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, to check results.  The network has no
counterpart in hyper.deal (whose Poisson CG starts from the previous
potential, examples/vlasov_poisson/include/application.h:575-576), so this
module restates the network's own definition:

  four 3x3 convolutions with periodic padding, channels 1 -> 32 -> 32 -> 32
  -> 1, leaky ReLU with slope 1/8 after the first three;
  conv: out[b,o,y,x] = bias[o] + sum_{i,ky,kx} w[o,i,ky,kx]
                                   * in[b,i,(y+ky-1) mod H,(x+kx-1) mod W];
  weights read as BF16 (round to nearest even), biases FP32;
  loss = mean (net(in) - target)^2; gradients of the loss with respect to
  the master weights are taken through the BF16 copies (straight through);
  Adam (beta 0.9, 0.999, eps 1e-8, bias-corrected) on the master values;
  initial weights uniform in +-sqrt(6/fan_in) (last layer scaled by 1/10)
  from SplitMix64 counters 0, 1, ... in canonical order, biases zero.

Parameters are one flat vector in the canonical layout: per layer
w[out][in][3][3] then b[out].

What pins it (tests/test_oracle_cnn.py): the backward pass against central
finite differences of the FP64 forward (no BF16 rounding); the BF16 rounding
against torch's float32 -> bfloat16 conversion; the convolution against a
direct loop sum on a tiny image; the periodic padding against a translated
input (the network commutes with periodic shifts); Adam's first step
against its closed form p - lr * sign(g) (for |g| >> eps).
"""

import numpy as np

CHANNELS = 32
LAYERS = 4
SLOPE = 0.125


def layer_shapes():
    """(in, out) channels per layer."""
    return [(1, CHANNELS), (CHANNELS, CHANNELS), (CHANNELS, CHANNELS),
            (CHANNELS, 1)]


def n_params():
    return sum(ci * co * 9 + co for ci, co in layer_shapes())


def unpack(p):
    """flat parameters -> [(w[Co,Ci,3,3], b[Co])] (views)."""
    out, o = [], 0
    for ci, co in layer_shapes():
        w = p[o:o + co * ci * 9].reshape(co, ci, 3, 3)
        o += co * ci * 9
        b = p[o:o + co]
        o += co
        out.append((w, b))
    return out


def init_params(seed, uniform):
    """uniform(seed, i): the SplitMix64 draw in [-1, 1) for counter i."""
    p = np.zeros(n_params(), dtype=np.float32)
    ctr, o = 0, 0
    for l, (ci, co) in enumerate(layer_shapes()):
        lim = np.sqrt(6.0 / (ci * 9)) * (0.1 if l == LAYERS - 1 else 1.0)
        for k in range(co * ci * 9):
            p[o + k] = np.float32(lim * uniform(seed, ctr))
            ctr += 1
        o += co * ci * 9 + co
    return p


def bf16(x):
    """float32 values rounded to the nearest BF16 (ties to even), as
    float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7fffffff) > 0x7f800000
    r = ((u + 0x7fff + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, u | 0x400000, r).astype(np.uint32)
    return r.view(np.float32).reshape(x.shape)


def _shifted(x, ky, kx):
    """s[..., y, x] = x[..., (y+ky-1) mod H, (x+kx-1) mod W]"""
    return np.roll(x, shift=(1 - ky, 1 - kx), axis=(-2, -1))


def conv(x, w, b=None):
    """x [B,Ci,H,W], w [Co,Ci,3,3] -> [B,Co,H,W]."""
    out = np.zeros((x.shape[0], w.shape[0]) + x.shape[2:], dtype=x.dtype)
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("oi,bihw->bohw", w[:, :, ky, kx],
                             _shifted(x, ky, kx))
    if b is not None:
        out += b[None, :, None, None]
    return out


def conv_input_grad(g, w):
    """d/dx of sum(g * conv(x, w)): g [B,Co,H,W] -> [B,Ci,H,W]."""
    out = np.zeros((g.shape[0], w.shape[1]) + g.shape[2:], dtype=g.dtype)
    for ky in range(3):
        for kx in range(3):
            out += np.roll(np.einsum("oi,bohw->bihw", w[:, :, ky, kx], g),
                           shift=(ky - 1, kx - 1), axis=(-2, -1))
    return out


def conv_weight_grad(x, g):
    """d/dw of sum(g * conv(x, w)) -> [Co,Ci,3,3], and d/db -> [Co]."""
    co, ci = g.shape[1], x.shape[1]
    gw = np.zeros((co, ci, 3, 3), dtype=g.dtype)
    for ky in range(3):
        for kx in range(3):
            gw[:, :, ky, kx] = np.einsum("bohw,bihw->oi", g,
                                         _shifted(x, ky, kx))
    return gw, g.sum(axis=(0, 2, 3))


def forward(p, x, dtype=np.float32, round_bf16=True, cache=False):
    """x [B,H,W] -> out [B,H,W] (and the activations when cache)."""
    a = x[:, None].astype(dtype)
    acts = [a]
    layers = unpack(p)
    for l, (w, b) in enumerate(layers):
        wq = (bf16(w) if round_bf16 else w).astype(dtype)
        z = conv(a, wq, b.astype(dtype))
        a = np.where(z > 0, z, dtype(SLOPE) * z) if l < LAYERS - 1 else z
        acts.append(a)
    out = a[:, 0]
    return (out, acts) if cache else out


def loss_and_grad(p, x, target, dtype=np.float32, round_bf16=True):
    """loss = mean (net(x) - target)^2 (FP64 sum) and d loss / d p."""
    out, acts = forward(p, x, dtype, round_bf16, cache=True)
    d = out - target.astype(dtype)
    n = d.size
    loss = float(np.sum(d.astype(np.float64) ** 2) / n)
    g = (dtype(2.0 / n) * d)[:, None]
    layers = unpack(p)
    grads = [None] * LAYERS
    for l in range(LAYERS - 1, -1, -1):
        w, b = layers[l]
        wq = (bf16(w) if round_bf16 else w).astype(dtype)
        if l < LAYERS - 1:
            g = g * np.where(acts[l + 1] > 0, dtype(1), dtype(SLOPE))
        gw, gb = conv_weight_grad(acts[l], g)
        grads[l] = (gw, gb)
        if l > 0:
            g = conv_input_grad(g, wq)
    flat = np.concatenate([np.concatenate([gw.ravel(), gb])
                           for gw, gb in grads]).astype(dtype)
    return loss, flat


class Adam:
    """Adam on a flat FP32 parameter vector (beta 0.9, 0.999, eps 1e-8)."""

    def __init__(self, n):
        self.m = np.zeros(n, dtype=np.float32)
        self.v = np.zeros(n, dtype=np.float32)
        self.t = 0

    def step(self, p, g, lr):
        self.t += 1
        b1, b2 = np.float32(0.9), np.float32(0.999)
        c1 = np.float32(1.0 - 0.9 ** self.t)
        c2 = np.float32(1.0 - 0.999 ** self.t)
        g = g.astype(np.float32)
        self.m = b1 * self.m + (np.float32(1) - b1) * g
        self.v = b2 * self.v + (np.float32(1) - b2) * g * g
        return (p - np.float32(lr) * (self.m / c1) /
                (np.sqrt(self.v / c2) + np.float32(1e-8))).astype(np.float32)


def dofs_to_image(u, n0, n1, N):
    """2D x-space nodal values (cell-major, x0 cells fastest, node i0 + N i1)
    -> the (n1 N) x (n0 N) image (u - mean) / rms and rms."""
    u = np.asarray(u, dtype=np.float64)
    c = u.reshape(n1, n0, N, N)               # [cy, cx, iy, ix]
    img = c.transpose(0, 2, 1, 3).reshape(n1 * N, n0 * N)
    d = img - u.mean()
    rms = float(np.sqrt(np.mean(d * d)))
    return (d / rms if rms > 0 else 0 * d).astype(np.float32), rms


def image_to_dofs(img, n0, n1, N, scale, continuous=False):
    """image -> 2D x-space nodal values (scale * image); continuous: the
    end nodes of neighbouring cells (one point, two pixels, four at a
    corner) get the mean of their pixels."""
    img = np.asarray(img, dtype=np.float64)
    if continuous and N > 1:
        img = img.copy()
        for axis in (1, 0):
            a = np.moveaxis(img, axis, 0)
            last = np.arange(N - 1, a.shape[0], N)
            first = (last + 1) % a.shape[0]
            m = 0.5 * (a[last] + a[first])
            a[last] = m
            a[first] = m
            img = np.moveaxis(a, 0, axis)
    img = img.reshape(n1, N, n0, N)
    return (scale * img.transpose(0, 2, 1, 3)).reshape(-1)
