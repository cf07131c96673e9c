"""FP64 reference of the learned sub-grid closure (TEST INFRASTRUCTURE ONLY).

The closure has no counterpart in the reference; this module is the checker for the device
kernels (inference fused into the collision, training forward/backward/Adam).  Its backward
pass is pinned by central finite differences in tests/test_mlp.py.

Network: 9 normalised features -> 32 SiLU -> 32 SiLU -> softplus, giving nu_t / nu_scale.
Flat parameter order: W1[32][9], b1[32], W2[32][32], b2[32], w3[32], b3.
"""
import numpy as np

NF, NH = 9, 32
NPARAM = NH * NF + NH + NH * NH + NH + NH + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NH * NF].reshape(NH, NF); o += NH * NF
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    w3 = p[o:o + NH]; o += NH
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def softplus(z):
    return np.logaddexp(0.0, z)


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    Z1 = X @ W1.T + b1
    H1 = silu(Z1)
    Z2 = H1 @ W2.T + b2
    H2 = silu(Z2)
    z3 = H2 @ w3 + b3
    return softplus(z3), (X, Z1, H1, Z2, H2, z3)


def loss_and_grad(p, X, T):
    """Mean squared error of softplus output against targets, and its gradient."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (X, Z1, H1, Z2, H2, z3) = forward(p, X)
    n = X.shape[0]
    loss = np.mean((y - T) ** 2)
    dy = 2.0 * (y - T) / n
    dz3 = dy * sigmoid(z3)
    dw3 = H2.T @ dz3
    db3 = dz3.sum()
    dH2 = np.outer(dz3, w3)
    s2 = sigmoid(Z2)
    dZ2 = dH2 * s2 * (1.0 + Z2 * (1.0 - s2))
    dW2 = dZ2.T @ H1
    db2 = dZ2.sum(0)
    dH1 = dZ2 @ W2
    s1 = sigmoid(Z1)
    dZ1 = dH1 * s1 * (1.0 + Z1 * (1.0 - s1))
    dW1 = dZ1.T @ X
    db1 = dZ1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, [db3]])
    return loss, g


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def features_from_populations(f, tau0, dims=3):
    """The 9 closure inputs of each node from its populations (FP64), as the device computes
    them: P = Pi_neq / rho; a = sqrt(2 P:P); invariants I1 = tr P, I2 = (I1^2 - P:P)/2,
    I3 = det P; |u|^2, rho - 1, tau0, a tau0, a^2."""
    from oracle.zlbm_oracle import lattice
    cx, cy, cz, w, _ = lattice(dims)
    c = np.stack([cx, cy, cz], 1)[:, :dims]
    rho = f.sum(1)
    u = (f @ c) / rho[:, None]
    cu = u @ c.T
    feq = w * rho[:, None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u ** 2).sum(1)[:, None])
    pi = np.einsum("nq,qa,qb->nab", f - feq, c, c)
    P = pi / rho[:, None, None]
    PP = np.einsum("nab,nab->n", P, P)
    a = np.sqrt(2.0 * PP)
    I1 = np.trace(P, axis1=1, axis2=2)
    I2 = 0.5 * (I1 ** 2 - PP)
    I3 = np.linalg.det(P) if dims == 3 else np.zeros_like(a)
    u2 = (u ** 2).sum(1)
    t0 = np.full_like(a, tau0)
    return np.stack([a, I1, I2, I3, u2, rho - 1.0, t0, a * tau0, a * a], axis=1)


def smagorinsky_nu_t(f, tau0, cs, dims=3):
    """nu_t of the Smagorinsky closure in lattice units, tau_t / 3 with
    tau_t = (sqrt(tau0^2 + 18 cs^2 sqrt(2 Pi:Pi)) - tau0) / 2 (les_models/lbm_les_smagorinsky.cpp:22-46)."""
    from oracle.zlbm_oracle import lattice
    cx, cy, cz, w, _ = lattice(dims)
    c = np.stack([cx, cy, cz], 1)[:, :dims]
    rho = f.sum(1)
    u = (f @ c) / rho[:, None]
    cu = u @ c.T
    feq = w * rho[:, None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u ** 2).sum(1)[:, None])
    pi = np.einsum("nq,qa,qb->nab", f - feq, c, c)
    q = np.sqrt(2.0 * np.einsum("nab,nab->n", pi, pi))
    tau_t = 0.5 * (np.sqrt(tau0 ** 2 + 18 * cs ** 2 * q) - tau0)
    return tau_t / 3.0


def bf16_round(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f16_round(x):
    """Round to the nearest FP16 value (ties to even), returned as float64."""
    return np.asarray(x, dtype=np.float64).astype(np.float16).astype(np.float64)


SILU_SCALE = -np.log2(np.e)


def forward_lowp(p, X, rnd):
    """The network with low-precision operands and wide accumulation, as the matrix-core form
    evaluates it. W1, b1 and b2 are first scaled by c = -log2(e) (in FP64, stored as FP32), so
    both layers' pre-activations are t = c z and an activation is t / (1 + 2^t) = c silu(z).
    Inputs, the scaled W1 / b1, W2 and the layer-1 activations are rounded by `rnd` (b1 enters the
    products as the weight of a constant input 1); the scaled b2 (the layer-2 accumulators'
    starting value, FP32), w3 / c, b3 and the accumulations are exact. Up to those roundings this
    is forward()."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    c = SILU_SCALE

    def sc(a):
        return (c * np.asarray(a, np.float64)).astype(np.float32).astype(np.float64)

    with np.errstate(over="ignore"):          # 2^t -> inf gives the activation 0, as on the device
        T1 = rnd(X) @ rnd(sc(W1)).T + rnd(sc(b1))
        H1 = rnd(T1 / (1.0 + np.exp2(T1)))
        T2 = H1 @ rnd(W2).T + sc(b2)
        return softplus((T2 / (1.0 + np.exp2(T2))) @ (w3 / c) + b3)


def forward_bf16(p, X):
    return forward_lowp(p, X, bf16_round)


def forward_f16(p, X):
    return forward_lowp(p, X, f16_round)


def loss_and_grad_lowp(p, X, T, rnd):
    """loss_and_grad with the roundings of the matrix-core training kernel: X, W1, b1, W2, b2
    and the layer-1 activations are rounded by `rnd` as operands of the forward products; the
    layer-2 delta is rounded as operand of both dW2 and the layer-1 delta, and the layer-1 delta
    as operand of dW1 and db1 (db1 is a column of the same product, against inputs of 1). The
    SiLU arguments, the output layer, db2, dw3, db3 and the loss use unrounded values. With rnd
    the identity this is loss_and_grad (a test checks it)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    n = X.shape[0]
    Xr = rnd(X)
    Z1 = Xr @ rnd(W1).T + rnd(b1)
    H1 = rnd(silu(Z1))
    Z2 = H1 @ rnd(W2).T + rnd(b2)
    H2 = silu(Z2)
    z3 = H2 @ w3 + b3
    y = softplus(z3)
    loss = np.mean((y - T) ** 2)
    dz3 = 2.0 * (y - T) / n * sigmoid(z3)
    dw3 = H2.T @ dz3
    db3 = dz3.sum()
    s2 = sigmoid(Z2)
    dZ2 = np.outer(dz3, w3) * s2 * (1.0 + Z2 * (1.0 - s2))
    db2 = dZ2.sum(0)
    dZ2r = rnd(dZ2)
    dW2 = dZ2r.T @ H1
    s1 = sigmoid(Z1)
    dZ1 = rnd((dZ2r @ rnd(W2)) * s1 * (1.0 + Z1 * (1.0 - s1)))
    dW1 = dZ1.T @ Xr
    db1 = dZ1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, [db3]])
    return loss, g

