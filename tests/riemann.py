"""Exact solution of the isothermal Riemann problem (sound speed c, gas initially at rest on
both sides, rho_l > rho_r): a left rarefaction and a right shock around a uniform star state.

Star density from the two wave curves (Toro, Riemann Solvers, ch. 4, isothermal limit):
  rarefaction  u* = -c ln(rho*/rho_l)
  shock        u* = c (rho* - rho_r) / sqrt(rho* rho_r)
"""
import math

import numpy as np


def star_state(rho_l: float, rho_r: float, c: float = 1.0):
    f = lambda r: -c * math.log(r / rho_l) - c * (r - rho_r) / math.sqrt(r * rho_r)
    lo, hi = rho_r, rho_l
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0.0:
            lo = mid
        else:
            hi = mid
    rs = 0.5 * (lo + hi)
    return rs, -c * math.log(rs / rho_l)


def solution(x, t: float, rho_l: float, rho_r: float, c: float = 1.0, x0: float = 0.0):
    """Density and velocity at positions x and time t > 0 (discontinuity initially at x0)."""
    rs, us = star_state(rho_l, rho_r, c)
    s = us * rs / (rs - rho_r)          # shock speed from mass conservation
    xi = (np.asarray(x, dtype=np.float64) - x0) / t
    rho = np.empty_like(xi)
    u = np.empty_like(xi)
    head, tail = -c, us - c
    left = xi <= head
    fan = (xi > head) & (xi <= tail)
    star = (xi > tail) & (xi <= s)
    right = xi > s
    rho[left], u[left] = rho_l, 0.0
    u[fan] = xi[fan] + c
    rho[fan] = rho_l * np.exp(-u[fan] / c)
    rho[star], u[star] = rs, us
    rho[right], u[right] = rho_r, 0.0
    return rho, u


def shock_tube_errors(d, t: float, rho_l: float, c: float, ratio: float = 8.0):
    """One species of a shock tube (initial discontinuity at x = 0, density rho_l / ratio on
    the right) against the isothermal solution with sound speed c: L1 of 90 binned density and
    velocity values over x in [-0.45, 0.45] (relative to the mean exact density / the star
    velocity), the mean density and velocity between half the shock distance and 0.03 behind
    the shock relative to the star state, and the position of the last particle above the
    mid density between star and right states, relative to the exact shock."""
    rho_r = rho_l / ratio
    rs, us = star_state(rho_l, rho_r, c)
    x, act = d["x"], d["ind"] == 0
    bins = np.linspace(-0.45, 0.45, 91)
    xc = 0.5 * (bins[1:] + bins[:-1])
    k = np.digitize(x, bins) - 1
    rho_b = np.array([d["rho"][(k == i) & act].mean() for i in range(90)])
    u_b = np.array([d["vx"][(k == i) & act].mean() for i in range(90)])
    re, ue = solution(xc, t, rho_l, rho_r, c)
    xs = us * rs / (rs - rho_r) * t
    plateau = act & (x > 0.5 * xs) & (x < xs - 0.03)
    shock = float(np.max(x[act & (d["rho"] > 0.5 * (rs + rho_r)) & (x > 0.0)]))
    return {
        "l1_rho": float(np.mean(np.abs(rho_b - re)) / np.mean(re)),
        "l1_u": float(np.mean(np.abs(u_b - ue)) / us),
        "rho_star": float(np.mean(d["rho"][plateau])) / rs - 1.0,
        "u_star": float(np.mean(d["vx"][plateau])) / us - 1.0,
        "shock": shock - xs,
    }
