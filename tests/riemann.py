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
