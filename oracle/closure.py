"""NumPy restatement of the learned subgrid closure (no reference counterpart).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the HIP closure
kernels, never by the product package.

What it defines (the GPU kernels implement the same arithmetic):
  * ``box_filter(u, r)``: mean over non-overlapping r^3 blocks of a periodic
    physical field (c, N0, N1, N2) -> (c, N0/r, N1/r, N2/r).
  * ``sgs_stress(u, r)``: tau_ij = bar(u_i u_j) - bar(u_i) bar(u_j) in the
    component order (xx, yy, zz, xy, xz, yz).
  * the network: three periodic 3x3x3 cross-correlation layers
    3 -> 32 -> 32 -> 6, bias, SiLU after layers 1 and 2,
        out[p, co] = b[co] + sum_{tap, ci} W[tap, co, ci] in[p + d(tap), ci],
    tap = 9 (dz + 1) + 3 (dy + 1) + (dx + 1) for d = (dz, dy, dx) in {-1,0,1}^3
    on the axes (0, 1, 2) of a channels-last field (N0, N1, N2, C).
  * ``loss``: mean over points and the 6 components of (y - t)^2.
  * ``backward``: analytic gradients of the loss w.r.t. every weight and bias.
  * ``bf16`` rounding (round to nearest even of the float32 value) applied,
    when ``bf16=True``, where the GPU stores bfloat16: the network input, the
    weights and the outputs of layers 1 and 2 (pre- and post-activation) and
    the gradients that enter the backward products.

How it is pinned (tests/test_oracle_closure.py): central finite differences of
``loss`` in float64 against ``backward`` for every parameter group; a delta
kernel reproduces a periodic shift; a constant field gives sum(W) * c + b;
box_filter/sgs_stress against explicit loops and the identity
tau_ii >= 0 (trace of a covariance); bf16 rounding against exact bit
patterns.
"""
import numpy as np

__all__ = ["bf16", "box_filter", "sgs_stress", "init_params", "forward", "loss", "backward", "TAPS",
           "silu", "dsilu"]

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
LAYERS = ((3, 32), (32, 32), (32, 6))  # (cin, cout)


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    b = a.view(np.uint32).astype(np.uint64)
    lsb = (b >> 16) & 1
    r = ((b + 0x7FFF + lsb) >> 16) << 16
    nan = np.isnan(a)
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def box_filter(u, r):
    """Block mean of (c, N0, N1, N2) over r^3 blocks."""
    c, n0, n1, n2 = u.shape
    v = u.reshape(c, n0 // r, r, n1 // r, r, n2 // r, r)
    return v.mean(axis=(2, 4, 6))


PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def sgs_stress(u, r):
    """tau (6, N0/r, N1/r, N2/r) = bar(u_i u_j) - bar(u_i) bar(u_j)."""
    ub = box_filter(u, r)
    tau = np.empty((6,) + ub.shape[1:])
    for m, (i, j) in enumerate(PAIRS):
        tau[m] = box_filter((u[i] * u[j])[None], r)[0] - ub[i] * ub[j]
    return tau


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init_params(seed=0, scale=1.0):
    """Weights W[l] (27, cout, cin) and biases b[l] (cout,), Kaiming-uniform-like."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W, b = [], []
    for cin, cout in LAYERS:
        lim = scale * np.sqrt(3.0 / (27 * cin))
        W.append(rng.uniform(-lim, lim, (27, cout, cin)))
        b.append(np.zeros(cout))
    return W, b


def _shift(x, d):
    """x[p + d] for every p (periodic), x channels-last (N0, N1, N2, C)."""
    return np.roll(x, shift=(-d[0], -d[1], -d[2]), axis=(0, 1, 2))


def conv(x, W, bias):
    out = np.zeros(x.shape[:3] + (W.shape[1],))
    for t, d in enumerate(TAPS):
        out += _shift(x, d) @ W[t].T
    return out + bias


def conv_input_grad(g, W):
    """d/dx of sum(g * conv(x, W)): g_in[q] = sum_tap W[tap]^T g[q - d(tap)]."""
    out = np.zeros(g.shape[:3] + (W.shape[2],))
    for t, d in enumerate(TAPS):
        out += _shift(g, (-d[0], -d[1], -d[2])) @ W[t]
    return out


def conv_weight_grad(g, x):
    """dW[tap, co, ci] = sum_p g[p, co] x[p + d(tap), ci]."""
    n = g.shape[0] * g.shape[1] * g.shape[2]
    g2 = g.reshape(n, -1)
    return np.stack([g2.T @ _shift(x, d).reshape(n, -1) for d in TAPS])


def forward(x, W, b, bf16_storage=False):
    """x: (N0, N1, N2, 3) network input.  Returns the output (N0, N1, N2, 6) and
    the saved tensors (inputs of each layer, pre-activations)."""
    q = bf16 if bf16_storage else (lambda a: np.asarray(a, dtype=np.float64))
    Wq = [q(w) for w in W]
    a0 = q(x)
    z1 = conv(a0, Wq[0], b[0])
    z1s = q(z1)
    a1 = q(silu(z1s))
    z2 = conv(a1, Wq[1], b[1])
    z2s = q(z2)
    a2 = q(silu(z2s))
    y = conv(a2, Wq[2], b[2])
    return y, (a0, z1s, a1, z2s, a2, Wq)


def loss(y, t):
    return float(np.mean((y - t) ** 2))


def backward(y, t, saved, bf16_storage=False):
    """Gradients (dW list, db list) of loss(y, t) with the saved tensors."""
    q = bf16 if bf16_storage else (lambda a: np.asarray(a, dtype=np.float64))
    a0, z1, a1, z2, a2, Wq = saved
    gy = 2.0 * (y - t) / y.size
    gyq = q(gy)
    dW3 = conv_weight_grad(gyq, a2)
    db3 = gy.reshape(-1, gy.shape[-1]).sum(axis=0)
    g_a2 = conv_input_grad(gyq, Wq[2])
    g_z2 = q(g_a2 * dsilu(z2))
    dW2 = conv_weight_grad(g_z2, a1)
    db2 = g_z2.reshape(-1, g_z2.shape[-1]).sum(axis=0)
    g_a1 = conv_input_grad(g_z2, Wq[1])
    g_z1 = q(g_a1 * dsilu(z1))
    dW1 = conv_weight_grad(g_z1, a0)
    db1 = g_z1.reshape(-1, g_z1.shape[-1]).sum(axis=0)
    return [dW1, dW2, dW3], [db1, db2, db3]
