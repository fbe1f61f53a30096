"""Pins the learned-preconditioner oracle (oracle/mpnn_oracle.py) by central
finite differences of its float64 loss, parameter by parameter, and by
invariants of the message passing (adjoint of the neighbour sum)."""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mpnn_oracle as mo  # noqa: E402


def _tiny(seed=0, K=9, Np=3, H=6):
    rng = np.random.default_rng(seed)
    nbr = -np.ones((K, 3), dtype=int)
    # a ring of cells plus one chord, symmetric adjacency
    edges = [(i, (i + 1) % K) for i in range(K)] + [(0, 4)]
    slots = [0] * K
    for a, b in edges:
        nbr[a, slots[a]] = b
        slots[a] += 1
        nbr[b, slots[b]] = a
        slots[b] += 1
    din = 2 * Np + 4
    x = rng.standard_normal((K, din))
    t = rng.standard_normal((K, Np))
    p = mo.init_params(din, H, Np, seed)
    for k in p:
        if k.startswith("b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    return p, x, t, nbr


def test_neighbour_sum_adjoint():
    p, x, t, nbr = _tiny()
    rng = np.random.default_rng(5)
    a = rng.standard_normal((x.shape[0], 4))
    b = rng.standard_normal((x.shape[0], 4))
    assert abs(np.sum(mo.neighbour_sum(a, nbr) * b) - np.sum(a * mo.neighbour_sum(b, nbr))) < 1e-12


@pytest.mark.parametrize("name", ["A1", "B1", "b1", "A2", "B2", "b2", "A3", "B3", "b3", "Wo", "bo"])
def test_gradients_match_finite_differences(name):
    p, x, t, nbr = _tiny()
    _, g, _ = mo.loss_and_grads(p, x, t, nbr)
    rng = np.random.default_rng(1)
    flat = p[name].reshape(-1)
    for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
        h = 1e-6
        old = flat[idx]
        flat[idx] = old + h
        lp, _, _ = mo.loss_and_grads(p, x, t, nbr)
        flat[idx] = old - h
        lm, _, _ = mo.loss_and_grads(p, x, t, nbr)
        flat[idx] = old
        fd = (lp - lm) / (2 * h)
        an = g[name].reshape(-1)[idx]
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, idx, fd, an)


def test_bf16_rounding_helper():
    v = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-9, -3.14159265, 1e-30])
    r = mo.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # halfway rounds to even
    assert r[2] == 1.0 + 2**-7
    assert abs(r[3] + 3.140625) < 1e-12


@pytest.mark.parametrize("name", ["A1", "B2", "b3", "Wo", "bo"])
def test_energy_loss_gradients_match_finite_differences(name):
    p, x, t, nbr = _tiny(seed=2)
    K, Np = t.shape
    rng = np.random.default_rng(3)
    G = rng.standard_normal((K * Np, K * Np))
    S = G @ G.T + K * Np * np.eye(K * Np)  # SPD stand-in for the operator

    def apply_S(v):
        return (S @ v.reshape(-1)).reshape(K, Np)

    e = rng.standard_normal((K, Np))
    r = apply_S(e)
    zb = 0.1 * rng.standard_normal((K, Np))
    _, g, _ = mo.energy_loss_and_grads(p, x, nbr, zb, r, e, 0.7, apply_S)
    flat = p[name].reshape(-1)
    for idx in rng.choice(flat.size, size=min(4, flat.size), replace=False):
        h = 1e-6
        old = flat[idx]
        flat[idx] = old + h
        lp, _, _ = mo.energy_loss_and_grads(p, x, nbr, zb, r, e, 0.7, apply_S)
        flat[idx] = old - h
        lm, _, _ = mo.energy_loss_and_grads(p, x, nbr, zb, r, e, 0.7, apply_S)
        flat[idx] = old
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[name].reshape(-1)[idx]) <= 1e-6 * max(1.0, abs(fd))


def test_energy_loss_is_zero_at_the_exact_error():
    p, x, t, nbr = _tiny(seed=4)
    K, Np = t.shape
    for k in p:
        p[k] = np.zeros_like(p[k])
    e = np.random.default_rng(0).standard_normal((K, Np))
    loss, _, _ = mo.energy_loss_and_grads(p, x, nbr, e, 2.0 * e, e, 1.0, lambda v: 2.0 * v)
    assert abs(loss) < 1e-14
