"""ORACLE -- test infrastructure only.

Float64 numpy restatement of the learned preconditioner's network (no
reference counterpart: the reference has no CG and no learned parts). Only
tests/ may import it, as the checker of the HIP forward/backward kernels.

Network (per cell c, H hidden units, element-face adjacency N(c)):
    x_c   = [r_c / s_r, zb_c / s_z, g_c]              (Np + Np + 4 features)
    m_c   = sum_{n in N(c)} h_n                        (neighbour sum)
    h^l_c = relu(A_l h^{l-1}_c + B_l m^{l-1}_c + b_l),  l = 1..3, h^0 = x
    y_c   = W_o h^3_c + b_o                            (Np outputs)
with zb = P_BJ r (block-Jacobi) and the preconditioned vector
    z = zb + s_z * c_t * y.
Loss (training): mean over all cells and nodes of (y - t)^2.

bf16=True rounds to bfloat16 wherever the device kernels store bf16 (the
inputs x, every weight, every h^l and every neighbour sum m); everything
else is float64. Pinned by finite differences (tests/test_mpnn_oracle.py).
"""
from __future__ import annotations

import numpy as np


def bf16(a):
    """Round float64/float32 values to the nearest bfloat16 (ties to even)."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def neighbour_sum(h, nbr):
    """h: (K, D); nbr: (K, 3) neighbour cell or -1 at a boundary face."""
    m = np.zeros_like(h)
    for e in range(3):
        ok = nbr[:, e] >= 0
        m[ok] += h[nbr[ok, e]]
    return m


def init_params(din, H, np_out, seed=0):
    """He-style initialisation (deterministic)."""
    rng = np.random.default_rng(seed)
    p = {}
    dims = [din, H, H, H]
    for l in range(1, 4):
        fan = 2 * dims[l - 1]
        p[f"A{l}"] = rng.standard_normal((H, dims[l - 1])) * np.sqrt(2.0 / fan)
        p[f"B{l}"] = rng.standard_normal((H, dims[l - 1])) * np.sqrt(2.0 / fan)
        p[f"b{l}"] = np.zeros(H)
    p["Wo"] = rng.standard_normal((np_out, H)) * np.sqrt(1.0 / H)
    p["bo"] = np.zeros(np_out)
    return p


def forward(p, x, nbr, bf=False):
    rnd = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    q = {k: rnd(v) if k[0] in "ABW" else np.asarray(v, np.float64) for k, v in p.items()}
    cache = {"h0": rnd(x)}
    h = cache["h0"]
    for l in range(1, 4):
        m = rnd(neighbour_sum(h, nbr))
        pre = h @ q[f"A{l}"].T + m @ q[f"B{l}"].T + q[f"b{l}"]
        cache[f"m{l - 1}"] = m
        cache[f"pre{l}"] = pre
        h = rnd(np.maximum(pre, 0.0))
        cache[f"h{l}"] = h
    y = h @ q["Wo"].T + q["bo"]
    cache["q"] = q
    return y, cache


def loss_and_grads(p, x, t, nbr, bf=False):
    """Mean-square loss against targets t and its parameter gradients."""
    y, c = forward(p, x, nbr, bf)
    n = y.size
    loss = float(np.mean((y - t) ** 2))
    g = backward(p, c, 2.0 * (y - t) / n, nbr)
    return loss, g, y


def energy_loss_and_grads(p, x, nbr, zb, r, e, scale, apply_S, bf=False):
    """Energy-norm loss of the preconditioned vector z = zb + scale * y:
        L = (z - e)^T S (z - e) / (e^T S e),  S e = r,
    with dL/dy = scale * 2 (S z - r) / (e^T r). zb, r, e: (K, Np); apply_S maps
    a (K, Np) array to S times it."""
    y, c = forward(p, x, nbr, bf)
    z = zb + scale * y
    Sz = apply_S(z)
    er = float(np.sum(e * r))
    loss = float((np.sum(z * Sz) - 2.0 * np.sum(z * r) + er) / er)
    g = backward(p, c, scale * 2.0 * (Sz - r) / er, nbr)
    return loss, g, y


def backward(p, c, dy, nbr):
    """Parameter gradients for the output gradient dy (K, Np)."""
    q = c["q"]
    g = {"Wo": dy.T @ c["h3"], "bo": dy.sum(0)}
    dh = dy @ q["Wo"]
    for l in range(3, 0, -1):
        dz = dh * (c[f"pre{l}"] > 0)
        g[f"A{l}"] = dz.T @ c[f"h{l - 1}"]
        g[f"B{l}"] = dz.T @ c[f"m{l - 1}"]
        g[f"b{l}"] = dz.sum(0)
        if l > 1:
            dh = dz @ q[f"A{l}"] + neighbour_sum(dz @ q[f"B{l}"], nbr)
    return g


def features(r, zb, metric, Np):
    """x = [r/s_r, zb/s_z, g] with global RMS scales; r, zb: (K, Np), metric: (K, 4)."""
    s_r = np.sqrt(np.mean(r * r))
    s_z = np.sqrt(np.mean(zb * zb))
    g = metric / np.sqrt(np.mean(metric * metric, axis=0))
    return np.concatenate([r / s_r, zb / s_z, g], axis=1), s_r, s_z
