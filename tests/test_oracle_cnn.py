"""Pins the oracle's CNN restatement (hpccg_oracle.c: oracle_cnn_forward / _backward) before
the GPU kernels are checked against it: the forward pass against an independent numpy
convolution (shifted-slice sums), the backward pass against central finite differences of
the forward pass, and the BF16 rounding against its definition (round to nearest even)."""
import numpy as np
import pytest

import oracle


def np_conv(inp, W, cin, cout, dims):
    """inp [v][cin] -> out [v][cout] with zero padding, via shifted slices (x fastest)."""
    nx, ny, nz = dims
    a = inp.reshape(nz, ny, nx, cin).astype(np.float64)
    pad = np.zeros((nz + 2, ny + 2, nx + 2, cin))
    pad[1:-1, 1:-1, 1:-1] = a
    out = np.zeros((nz, ny, nx, cout))
    Wr = W.reshape(cout, cin, 27).astype(np.float64)
    for t in range(27):
        dx, dy, dz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
        sh = pad[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
        out += np.einsum("zyxc,oc->zyxo", sh, Wr[:, :, t])
    return out.reshape(-1, cout)


def weights(seed, scale=0.2):
    return (np.random.default_rng(seed).standard_normal(oracle.NW1 + oracle.NW2 + oracle.NW3) * scale).astype(np.float32)


@pytest.mark.parametrize("dims", [(5, 4, 3), (8, 8, 4), (3, 7, 2)])
def test_forward_matches_independent_numpy(dims):
    w = weights(1)
    n = int(np.prod(dims))
    u = np.random.default_rng(2).standard_normal(n).astype(np.float32)
    z, h1, h2 = oracle.cnn_forward(w, u, dims, quant=False)
    W1, W2, W3 = oracle.split_weights(w)
    r1 = np.maximum(np_conv(u[:, None], W1, 1, 16, dims), 0)
    assert np.allclose(h1.reshape(-1, 16), r1, rtol=1e-6, atol=1e-6)
    r2 = np.maximum(np_conv(h1.reshape(-1, 16), W2, 16, 16, dims), 0)
    assert np.allclose(h2.reshape(-1, 16), r2, rtol=1e-6, atol=1e-6)
    r3 = np_conv(h2.reshape(-1, 16), W3, 16, 1, dims)[:, 0]
    assert np.allclose(z, r3, rtol=1e-6, atol=1e-6)


def test_backward_matches_finite_differences():
    """Positive weights and inputs keep every ReLU active, so the loss is linear in each layer's
    weights near w and central differences are exact up to float rounding."""
    dims = (4, 3, 3)
    n = int(np.prod(dims))
    rng = np.random.default_rng(3)
    w = np.abs(weights(3, 0.3))
    u = np.abs(rng.standard_normal(n)).astype(np.float32) + 0.1
    dz = rng.standard_normal(n).astype(np.float32)
    g = oracle.cnn_backward(w, u, dz, dims, quant=False)
    L = lambda ww: float(np.dot(oracle.cnn_forward(ww, u, dims, quant=False)[0].astype(np.float64), dz))
    for lo, hi in [(0, oracle.NW1), (oracle.NW1, oracle.NW1 + oracle.NW2), (oracle.NW1 + oracle.NW2, len(w))]:
        D = np.zeros(len(w), np.float32)
        D[lo:hi] = np.abs(rng.standard_normal(hi - lo)).astype(np.float32) * 0.1
        gd = float(np.dot(g.astype(np.float64), D))
        eps = 1e-2
        fd = (L(w + eps * D) - L(w - eps * D)) / (2 * eps)
        assert abs(fd - gd) <= 1e-4 * abs(gd), (lo, fd, gd)


def test_backward_matches_independent_numpy_backprop():
    """Mixed signs (active and inactive ReLUs): manual backprop with shifted-slice convolutions."""
    dims = (5, 4, 3)
    n = int(np.prod(dims))
    w = weights(11)
    u = np.random.default_rng(12).standard_normal(n).astype(np.float32)
    dz = np.random.default_rng(13).standard_normal(n).astype(np.float32)
    g = oracle.cnn_backward(w, u, dz, dims, quant=False)
    z, h1, h2 = oracle.cnn_forward(w, u, dims, quant=False)
    W1, W2, W3 = (x.astype(np.float64) for x in oracle.split_weights(w))
    h1 = h1.reshape(-1, 16).astype(np.float64)
    h2 = h2.reshape(-1, 16).astype(np.float64)

    def flipT(W, cout, cin):  # W[co][ci][t] -> W'[ci][co][26 - t]
        return W.reshape(cout, cin, 27)[:, :, ::-1].transpose(1, 0, 2).copy()

    def wgrad(dout, inp, cout, cin):
        nx, ny, nz = dims
        a = inp.reshape(nz, ny, nx, cin)
        pad = np.zeros((nz + 2, ny + 2, nx + 2, cin))
        pad[1:-1, 1:-1, 1:-1] = a
        d = dout.reshape(nz, ny, nx, cout)
        out = np.zeros((cout, cin, 27))
        for t in range(27):
            dx, dy, dzz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
            sh = pad[1 + dzz:1 + dzz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
            out[:, :, t] = np.einsum("zyxo,zyxc->oc", d, sh)
        return out.reshape(-1)

    d = dz.astype(np.float64)[:, None]
    g3 = wgrad(d, h2, 1, 16)
    dh2 = np_conv(d, flipT(W3, 1, 16), 1, 16, dims) * (h2 > 0)
    g2 = wgrad(dh2, h1, 16, 16)
    dh1 = np_conv(dh2, flipT(W2, 16, 16), 16, 16, dims) * (h1 > 0)
    g1 = wgrad(dh1, u.astype(np.float64)[:, None], 16, 1)
    ref = np.concatenate([g1, g2, g3])
    assert np.allclose(g, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())


def test_backward_batch_is_sum_of_samples():
    dims = (4, 4, 2)
    n = int(np.prod(dims))
    w = weights(7)
    u = np.random.default_rng(8).standard_normal(2 * n).astype(np.float32)
    dz = np.random.default_rng(9).standard_normal(2 * n).astype(np.float32)
    g = oracle.cnn_backward(w, u, dz, dims, batch=2)
    g0 = oracle.cnn_backward(w, u[:n], dz[:n], dims)
    g1 = oracle.cnn_backward(w, u[n:], dz[n:], dims)
    assert np.allclose(g, g0 + g1, rtol=1e-5, atol=1e-5)


def test_bf16_round_to_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.01171875, -3.5, 1e-20, 65504.0, 3.0e38], np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins; 1 + 3*2^-8 -> 1 + 2^-6
    assert oracle.bf16(vals)[1] == 1.0 and oracle.bf16(vals)[2] == 1.015625
    for v in vals:
        assert oracle.lib().oracle_bf16(float(v)) == oracle.bf16(np.array([v], np.float32))[0]
    q = oracle.bf16(np.random.default_rng(1).standard_normal(1000).astype(np.float32))
    assert np.all(q.view(np.uint32) & 0xFFFF == 0)
