"""Pins of the oracle's per-pair and per-cell arithmetic against closed forms.

W, dW      kernel.cu:11-38: normalisation, value at 0, derivative, parity, support.
IDIC solve kernel.cu:2294-2437: it must be the backward-Euler solution of the implicit
           two-fluid system, checked against an independent linear solve, the one-species
           dust-box closed form, momentum balance and the eps_2 -> 0 limit.
"""
import math

import numpy as np
import pytest
from scipy import integrate

PI_REF = 3.14159265  # the reference's pi (kernel.cu:13)


@pytest.mark.parametrize("h", [0.1, 0.0024, 1.7])
def test_W_normalisation(oracle_mod, h):
    val, _ = integrate.quad(lambda r: 4 * math.pi * r * r * oracle_mod.W(r, h), 0, 2 * h,
                            points=[h], epsabs=0, epsrel=1e-13, limit=200)
    # with pi = 3.14159265 the integral is pi / 3.14159265 = 1 + 1.14e-9
    assert abs(val - math.pi / PI_REF) < 1e-11
    assert abs(val - 1.0 - 1.14e-9) < 1e-11


@pytest.mark.parametrize("h", [0.1, 0.0048])
def test_W_values_and_support(oracle_mod, h):
    assert oracle_mod.W(0.0, h) == pytest.approx(1.0 / (PI_REF * h ** 3), rel=1e-15)
    assert oracle_mod.W(2 * h, h) == 0.0
    assert oracle_mod.W(2.5 * h, h) == 0.0
    assert oracle_mod.W(h, h) == pytest.approx(0.25 / (PI_REF * h ** 3), rel=1e-15)
    assert oracle_mod.W(-0.7 * h, h) == oracle_mod.W(0.7 * h, h)


@pytest.mark.parametrize("h", [0.1, 0.0048])
def test_dW_is_derivative_and_odd(oracle_mod, h):
    for q in (0.1, 0.37, 0.8, 1.2, 1.55, 1.9):
        r = q * h
        d = 1e-6 * h
        fd = (oracle_mod.W(r + d, h) - oracle_mod.W(r - d, h)) / (2 * d)
        assert oracle_mod.dW(r, h) == pytest.approx(fd, rel=1e-7)
        assert oracle_mod.dW(-r, h) == -oracle_mod.dW(r, h)
    # continuity at q = 1 and q = 2, zero at 0
    assert oracle_mod.dW(h, h) == pytest.approx(-0.75 / (PI_REF * h ** 4), rel=1e-15)
    assert oracle_mod.dW((1 - 1e-12) * h, h) == pytest.approx(oracle_mod.dW((1 + 1e-12) * h, h), rel=1e-9)
    assert oracle_mod.dW(2 * h, h) == 0.0
    assert oracle_mod.dW(0.0, h) == 0.0
    assert oracle_mod.dW(3 * h, h) == 0.0


def implicit_solve(vg, vd, psi, tbar, eps, tau):
    """Backward Euler of  dv_g/dt = psi + sum_k eps_k (v_k - v_g)/t_k,  dv_k/dt = (v_g - v_k)/t_k,
    solved as a dense linear system per component (independent of the reference's algebra)."""
    nd = len(tbar)
    n = 1 + nd
    out = np.zeros((n, 3))
    for a in range(3):
        M = np.zeros((n, n))
        b = np.zeros(n)
        M[0, 0] = 1.0 / tau + sum(eps[k] / tbar[k] for k in range(nd))
        for k in range(nd):
            M[0, 1 + k] = -eps[k] / tbar[k]
            M[1 + k, 1 + k] = 1.0 / tau + 1.0 / tbar[k]
            M[1 + k, 0] = -1.0 / tbar[k]
            b[1 + k] = vd[k][a] / tau
        b[0] = vg[a] / tau + psi[a]
        out[:, a] = np.linalg.solve(M, b)
    return out[0], out[1:]


@pytest.mark.parametrize("nd", [1, 2])
def test_idic_cell_matches_linear_solve(oracle_mod, nd):
    rng = np.random.default_rng(7)
    for _ in range(200):
        vg = rng.normal(size=3)
        vd = rng.normal(size=(nd, 3))
        psi = rng.normal(size=3) * 10
        tbar = np.exp(rng.uniform(np.log(1e-5), np.log(10.0), size=nd))
        eps = rng.uniform(0.01, 10.0, size=nd)
        tau = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2))))
        g, d = oracle_mod.idic_cell(vg, vd, psi, tbar, eps, tau)
        g2, d2 = implicit_solve(vg, vd, psi, tbar, eps, tau)
        scale = max(np.max(np.abs(g2)), np.max(np.abs(d2)), 1e-30)
        assert np.max(np.abs(g - g2)) / scale < 1e-9
        assert np.max(np.abs(d - d2)) / scale < 1e-9
        # momentum: v_g + sum eps_k v_k grows by exactly tau psi
        lhs = g + (eps[:, None] * d).sum(0)
        rhs = vg + (eps[:, None] * vd).sum(0) + tau * psi
        assert np.max(np.abs(lhs - rhs)) <= 1e-12 * max(1.0, np.max(np.abs(rhs)))


def test_idic_dustybox_closed_form(oracle_mod):
    """One species, uniform cell: x = v_g - v_d obeys x' = t (x + tau psi) / (t + tau (1 + eps))."""
    for t, eps, tau in [(1e-3, 1.0, 7e-4), (10.0, 0.1, 1e-3), (1e-6, 5.0, 1e-3)]:
        vg = np.array([0.3, -0.2, 0.1])
        vd = np.array([[-0.1, 0.4, 0.0]])
        psi = np.array([2.0, 0.0, -1.0])
        g, d = oracle_mod.idic_cell(vg, vd, psi, [t], [eps], tau)
        x_new = g - d[0]
        expect = t * (vg - vd[0] + tau * psi) / (t + tau * (1 + eps))
        assert np.allclose(x_new, expect, rtol=1e-12, atol=1e-15)


def test_idic_two_species_limit(oracle_mod):
    """eps_2 -> 0 reduces the two-species solve to the one-species one."""
    vg = np.array([0.5, 0.1, -0.3])
    vd = np.array([[0.0, 0.2, 0.1], [1.0, -1.0, 2.0]])
    psi = np.array([1.0, 2.0, 3.0])
    g1, d1 = oracle_mod.idic_cell(vg, vd[:1], psi, [1e-3], [0.8], 5e-4)
    g2, d2 = oracle_mod.idic_cell(vg, vd, psi, [1e-3, 0.5], [0.8, 1e-12], 5e-4)
    assert np.allclose(g2, g1, rtol=1e-9, atol=1e-12)
    assert np.allclose(d2[0], d1[0], rtol=1e-9, atol=1e-12)
