"""Pins of the closure oracle (oracle/closure.py) on the CPU."""
import numpy as np
import pytest

from oracle import closure as C


def test_bf16_rounding_bits():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.0e-20, 65504.0, 1 + 2 ** -7], dtype=np.float32)
    r = C.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # halfway, ties to even (mantissa 0)
    assert r[2] == 1 + 2 ** -6              # halfway, ties to even (mantissa 2)
    assert r[3] == -2.5
    assert r[6] == 1 + 2 ** -7              # representable
    torch = pytest.importorskip("torch")
    y = np.random.default_rng(0).standard_normal(10000).astype(np.float32) * 10.0 ** np.arange(-4, 6).repeat(1000)
    ref = torch.from_numpy(y).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(C.bf16(y), ref)


def test_box_filter_and_stress_loops():
    rng = np.random.default_rng(1)
    u = rng.standard_normal((3, 8, 4, 12))
    r = 2
    ub = C.box_filter(u, r)
    ref = np.zeros((3, 4, 2, 6))
    for c in range(3):
        for i in range(4):
            for j in range(2):
                for k in range(6):
                    ref[c, i, j, k] = u[c, 2 * i:2 * i + 2, 2 * j:2 * j + 2, 2 * k:2 * k + 2].mean()
    assert np.allclose(ub, ref, rtol=0, atol=1e-15)
    tau = C.sgs_stress(u, r)
    blk = u[:, 2:4, 0:2, 4:6].reshape(3, -1)
    cov = (blk[:, None, :] * blk[None, :, :]).mean(-1) - blk.mean(-1)[:, None] * blk.mean(-1)[None, :]
    for m, (i, j) in enumerate(C.PAIRS):
        assert abs(tau[m, 1, 0, 2] - cov[i, j]) < 1e-14
    assert np.all(tau[:3] >= -1e-15)
    # a field constant within each block has no subgrid stress
    v = np.repeat(np.repeat(np.repeat(rng.standard_normal((3, 4, 2, 6)), 2, 1), 2, 2), 2, 3)
    assert np.abs(C.sgs_stress(v, 2)).max() < 1e-14


def test_conv_delta_kernel_is_shift_and_constant_field():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 6, 8, 5))
    W = np.zeros((27, 5, 5))
    d = (1, -1, 0)
    W[C.TAPS.index(d)] = np.eye(5)
    out = C.conv(x, W, np.zeros(5))
    assert np.array_equal(out, np.roll(x, shift=(-1, 1, 0), axis=(0, 1, 2)))
    assert np.array_equal(out[1, 2, 3], x[2, 1, 3])
    Wr = rng.standard_normal((27, 4, 5))
    b = rng.standard_normal(4)
    c = rng.standard_normal(5)
    xc = np.broadcast_to(c, (4, 6, 8, 5)).copy()
    assert np.allclose(C.conv(xc, Wr, b), Wr.sum(0) @ c + b, atol=1e-13)


def test_conv_adjoints():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 6, 8, 3))
    W = rng.standard_normal((27, 5, 3))
    g = rng.standard_normal((4, 6, 8, 5))
    lhs = np.sum(g * C.conv(x, W, np.zeros(5)))
    assert abs(lhs - np.sum(C.conv_input_grad(g, W) * x)) < 1e-10 * abs(lhs)
    assert abs(lhs - np.sum(C.conv_weight_grad(g, x) * W)) < 1e-10 * abs(lhs)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((4, 4, 4, 3))
    t = rng.standard_normal((4, 4, 4, 6)) * 0.1
    W, b = C.init_params(seed=5)
    b = [rng.standard_normal(v.shape) * 0.1 for v in b]
    y, saved = C.forward(x, W, b)
    dW, db = C.backward(y, t, saved)
    h = 1e-6

    def L():
        return C.loss(C.forward(x, W, b)[0], t)

    for layer in range(3):
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in W[layer].shape)
            w0 = W[layer][idx]
            W[layer][idx] = w0 + h
            lp = L()
            W[layer][idx] = w0 - h
            lm = L()
            W[layer][idx] = w0
            fd = (lp - lm) / (2 * h)
            assert abs(fd - dW[layer][idx]) <= 1e-6 * max(abs(fd), 1e-3), (layer, idx, fd, dW[layer][idx])
        j = int(rng.integers(0, b[layer].shape[0]))
        b0 = b[layer][j]
        b[layer][j] = b0 + h
        lp = L()
        b[layer][j] = b0 - h
        lm = L()
        b[layer][j] = b0
        fd = (lp - lm) / (2 * h)
        assert abs(fd - db[layer][j]) <= 1e-6 * max(abs(fd), 1e-3), (layer, j)


def test_bf16_storage_close_to_exact():
    rng = np.random.default_rng(6)
    x = rng.standard_normal((4, 4, 8, 3))
    W, b = C.init_params(seed=7)
    y64, _ = C.forward(x, W, b)
    y16, _ = C.forward(x, W, b, bf16_storage=True)
    rel = np.linalg.norm(y16 - y64) / np.linalg.norm(y64)
    assert 1e-4 < rel < 3e-2   # bf16 storage is visible but small
