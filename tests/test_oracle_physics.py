"""Structural pins of the oracle's force pass and step (IDIC.h:198-260, kernel.cu:2294-2523).

For equal-mass particles with no frozen neighbours the SPH pressure + viscosity force is
pairwise antisymmetric and the energy equation balances it:
    sum_i m_i A_i = 0                     (momentum)
    sum_i m_i v_i . A_i + sum_i m_i de_i/dt = 0   (total energy rate)
A dropped term, a wrong sign or a missing factor 1/2 on the viscous heating breaks one
of these.  The set-up tests pin the lattice counts of the benchmark configurations.
"""
import numpy as np
import pytest

from dustyshock3d_amd import configs
from tests.parity_util import STATE_FIELDS


def _blob_state(oracle_mod, seed=3):
    """A jittered lattice block with random velocities, away from the x ends, no frozen walls."""
    spec = configs.tube(0.04, 0.01, 1.0, 1.0, max_steps=5)
    st = {s: oracle_mod.tube_ic(spec, s) for s in (0, 1)}
    rng = np.random.default_rng(seed)
    for s in (0, 1):
        a = st[s]
        keep = np.abs(a["x"]) < 0.2
        for k in list(a):
            a[k] = np.ascontiguousarray(a[k][keep])
        a["ind"][:] = 0
        n = len(a["x"])
        for k in ("vx", "vy", "vz"):
            a[k] = rng.normal(scale=0.3, size=n)
        # equal masses inside a species
        a["m"][:] = a["m"][0]
    return spec, st


def test_force_conserves_momentum_and_energy_rate(oracle_mod):
    spec, st = _blob_state(oracle_mod)
    params = oracle_mod.tube_params(spec)
    params.x_min, params.x_max = -0.3, 0.3
    o = oracle_mod.OracleSim(params)
    for s in (0, 1):
        o.upload(s, {k: st[s][k] for k in STATE_FIELDS})
    o.step(1)
    g = o.download(0)
    m = st[0]["m"]
    A = np.stack([g["ax"], g["ay"], g["az"]], 1)
    v = np.stack([st[0]["vx"], st[0]["vy"], st[0]["vz"]], 1)
    scale = np.sum(m[:, None] * np.abs(A))
    assert scale > 0
    assert np.max(np.abs(np.sum(m[:, None] * A, 0))) < 1e-12 * scale
    work = np.sum(m * np.sum(v * A, 1))
    heat = np.sum(m * g["edot"])
    mag = np.sum(np.abs(m * np.sum(v * A, 1))) + np.sum(np.abs(m * g["edot"]))
    assert abs(work + heat) < 1e-11 * mag
    # viscosity is active in this state (some approaching pairs), so the test covers it
    assert np.any(g["edot"] != 0)


def test_conservation_sums_reported(oracle_mod):
    spec = configs.tube(0.03, 0.01, 1.0, 1.0, max_steps=3)
    o = oracle_mod.OracleSim.from_spec(spec)
    o.step(3)
    cons = o.conservation()
    P = np.zeros(3)
    E = 0.0
    for s in (0, 1):
        a = o.download(s)
        v = np.stack([a["vx"], a["vy"], a["vz"]], 1)
        P += np.sum(a["m"][:, None] * v, 0)
        E += np.sum(0.5 * a["m"] * np.sum(v * v, 1)) + (np.sum(a["m"] * a["e"]) if s == 0 else 0.0)
    assert np.allclose(cons[:3], P, rtol=1e-12, atol=1e-18)
    assert cons[3] == pytest.approx(E, rel=1e-13)


def test_drag_relaxes_relative_velocity(oracle_mod):
    """Stiff drag (K large, t_stop << dt): after one step dust moves with the gas in every cell."""
    spec = configs.tube(0.03, 0.01, 1e6, 1.0, max_steps=2)
    o = oracle_mod.OracleSim.from_spec(spec)
    st = {s: o.download(s) for s in (0, 1)}
    rng = np.random.default_rng(1)
    st[1]["vx"][:] = np.where(st[1]["ind"] == 0, rng.normal(size=len(st[1]["vx"])), 0.0)
    o2 = oracle_mod.OracleSim(o.params)
    for s in (0, 1):
        o2.upload(s, {k: st[s][k] for k in STATE_FIELDS})
    o2.step(1)
    d = o2.download(1)
    cells = o2.avg_cells()
    g = o2.grid()
    n0, n1, f = g["n"][0], g["n"][1], g["f"]
    na0, na1 = n0 // f, n1 // f
    key = d["cell"]
    kx, ky, kz = key % n0, (key // n0) % n1, key // (n0 * n1)
    ac = kx // f + na0 * (ky // f + na1 * (kz // f))
    act = d["ind"] == 0
    vgh = cells["vg_hat"].reshape(-1, 3)[:, 0]
    tb = cells["tbar"][ac]
    tau = o2.status()["dt"]
    # backward Euler against the cell's solved gas velocity (IDIC.h:276):
    # v_new - vg_hat = (v_old - vg_hat) * tbar / (tbar + tau), with tbar/tau ~ 1e-3 here
    expect = (st[1]["vx"] - vgh[ac]) * tb / (tb + tau)
    got = d["vx"] - vgh[ac]
    assert np.allclose(got[act], expect[act], rtol=1e-9, atol=1e-15)
    assert np.max(np.abs(got[act])) < 2e-3 * np.max(np.abs(st[1]["vx"]))


@pytest.mark.parametrize("cfg,count", [(configs.cfg0, 11250), (configs.cfg1, 703125), (configs.cfg2, 2812500)])
def test_lattice_counts(oracle_mod, cfg, count):
    assert oracle_mod.tube_count(cfg()) == count


def test_lattice_setup(oracle_mod):
    spec = configs.cfg0()
    g = oracle_mod.tube_ic(spec, 0)
    d = oracle_mod.tube_ic(spec, 1)
    assert np.all(g["m"] == 1.25e-7 * (1 + 0 * g["m"])) or np.allclose(g["m"], 1.25e-7, rtol=1e-15)
    left = g["x"] <= 0.0 + 1e-3 * 0.005
    assert np.allclose(g["rho"][left], 1.0)
    assert np.allclose(g["rho"][~left], 0.125, rtol=1e-14)
    assert np.all((g["y"] >= 0) & (g["y"] < 0.05) & (g["z"] >= 0) & (g["z"] < 0.05))
    # jitter within 1e-3 of the local spacing, independent streams for gas and dust
    assert not np.array_equal(g["x"], d["x"])
    assert np.max(np.abs(g["x"] - d["x"])) <= 2e-3 * 0.01 + 1e-15
    # frozen x-end layers of 2.5 h
    frozen = g["ind"] == 1
    assert np.all((g["x"][frozen] < -0.5 + 2.5 * 0.006 + 1e-9) | (g["x"][frozen] > 0.5 - 2.5 * 0.012 - 1e-9))
    assert frozen.sum() > 0


# --- Monaghan viscosity (IDIC.h:213-235): properties that fix alpha, beta, c_bar, eps and signs ---
#   mu = hh (v_i - v_j).(r_i - r_j) / (d^2 + eps hh^2)  for approaching pairs, else no viscosity
#   Pi = (-alpha c_bar mu + beta mu^2) / rho_bar,  c_bar = (c_i + c_j)/2,  c = sqrt(gamma p / rho)
# The reference takes gamma into the viscosity sound speed also for the isothermal law
# (c = sqrt(gamma) cs there); the tests below fix that, the linear / quadratic split and the
# placement of eps, so that a dropped factor, a swapped sign or eps on the wrong term fails.

def _gas_rates(oracle_mod, params, st, **over):
    """Gas acceleration (n, 3) and de/dt of the force pass of one step."""
    p = oracle_mod.Params.from_buffer_copy(params)
    for k, v in over.items():
        setattr(p, k, v)
    o = oracle_mod.OracleSim(p)
    for s in st:
        o.upload(s, {k: st[s][k] for k in STATE_FIELDS})
    o.step(1)
    g = o.download(0)
    return np.stack([g["ax"], g["ay"], g["az"]], 1), g["edot"]


def _gas_accel(oracle_mod, params, st, **over):
    return _gas_rates(oracle_mod, params, st, **over)[0]


def _nw(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _blob_params(oracle_mod, spec):
    params = oracle_mod.tube_params(spec)
    params.x_min, params.x_max = -0.3, 0.3
    params.eos_type = 1
    return params


def test_viscosity_is_affine_in_alpha_beta_and_scales_with_velocity(oracle_mod):
    spec, st = _blob_state(oracle_mod, seed=5)
    params = _blob_params(oracle_mod, spec)
    acc = lambda a, b, state=st: _gas_accel(oracle_mod, params, state, alpha=a, beta=b)
    A00, A10, A01 = acc(0.0, 0.0), acc(1.0, 0.0), acc(0.0, 1.0)
    Aa, Ab = A10 - A00, A01 - A00
    assert np.max(np.abs(Aa)) > 0 and np.max(np.abs(Ab)) > 0
    # affine: no cross term, each coefficient enters once
    A = acc(0.7, 1.9)
    assert _nw(A - A00, 0.7 * Aa + 1.9 * Ab) < 1e-11
    # velocities x lam (lam > 0 keeps the approaching pairs): pressure part unchanged bit for
    # bit, alpha part x lam, beta part x lam^2
    lam = 2.5
    sl = {s: dict(st[s]) for s in st}
    for k in ("vx", "vy", "vz"):
        sl[0][k] = st[0][k] * lam
    B00 = acc(0.0, 0.0, sl)
    assert np.array_equal(B00, A00)
    assert _nw(acc(1.0, 0.0, sl) - B00, lam * Aa) < 1e-11
    assert _nw(acc(0.0, 1.0, sl) - B00, lam * lam * Ab) < 1e-11


def test_viscosity_sound_speed_takes_gamma_for_the_isothermal_law(oracle_mod):
    spec, st = _blob_state(oracle_mod, seed=6)
    params = _blob_params(oracle_mod, spec)
    alpha_part = lambda gam: (_gas_accel(oracle_mod, params, st, alpha=1.0, beta=0.0, gamma=gam)
                              - _gas_accel(oracle_mod, params, st, alpha=0.0, beta=0.0, gamma=gam))
    assert _nw(alpha_part(1.4), np.sqrt(1.4) * alpha_part(1.0)) < 1e-11
    # the beta part has no sound speed
    beta_part = lambda gam: (_gas_accel(oracle_mod, params, st, alpha=0.0, beta=1.0, gamma=gam)
                             - _gas_accel(oracle_mod, params, st, alpha=0.0, beta=0.0, gamma=gam))
    assert _nw(beta_part(1.4), beta_part(1.0)) < 1e-12


@pytest.mark.parametrize("H", [1.5, -1.5])
def test_viscosity_acts_on_approaching_pairs_only(oracle_mod, H):
    """vx = H x, vy = vz = 0: every pair has (v_i - v_j).(r_i - r_j) = H dx^2, so an expanding
    flow (H > 0) has no viscosity at all and a contracting one (H < 0) has it on every pair
    with dx != 0.  A flipped sign in mu swaps the two."""
    spec, st = _blob_state(oracle_mod, seed=7)
    params = _blob_params(oracle_mod, spec)
    st[0]["vx"] = H * st[0]["x"]
    st[0]["vy"][:] = 0.0
    st[0]["vz"][:] = 0.0
    A0 = _gas_accel(oracle_mod, params, st, alpha=0.0, beta=0.0)
    A1 = _gas_accel(oracle_mod, params, st, alpha=1.0, beta=2.0)
    if H > 0:
        assert np.array_equal(A1, A0)
    else:
        assert _nw(A1, A0) > 1e-3
        # viscosity dissipates: its work on the flow is negative, and it heats every
        # particle (each pair approaches, so each pair's heating term is >= 0)
        v = np.stack([st[0]["vx"], st[0]["vy"], st[0]["vz"]], 1)
        assert np.sum(st[0]["m"] * np.sum(v * (A1 - A0), 1)) < 0
        e0 = _gas_rates(oracle_mod, params, st, alpha=0.0, beta=0.0)[1]
        e1 = _gas_rates(oracle_mod, params, st, alpha=1.0, beta=2.0)[1]
        heat = e1 - e0
        assert np.all(heat >= -1e-12 * np.max(np.abs(heat)))
        assert np.mean(heat > 0) > 0.9


def _pair_state(d, h, u, e=(1.0, 1.0)):
    """Two gas particles on the x axis, d apart, closing at 2u."""
    a = {k: np.zeros(2) for k in STATE_FIELDS}
    a["ind"] = np.zeros(2, np.int32)
    a["x"][:] = (-0.5 * d, 0.5 * d)
    a["vx"][:] = (u, -u)
    a["m"][:] = 1e-6
    a["h"][:] = h
    a["e"][:] = e
    a["rho"][:] = 1.0
    a["p"][:] = 1.0
    return {0: a}


def _pair_params(oracle_mod, eos_type, gamma):
    params = oracle_mod.tube_params(configs.tube(0.06, 0.01, 1.0, 1.0, max_steps=5))
    params.n_dust = 0
    params.x_min, params.x_max = -0.05, 0.05
    params.eos_type = eos_type
    params.gamma = gamma
    params.h_mode = 0
    return params


@pytest.mark.parametrize("d", [0.004, 0.013, 0.019])
def test_viscosity_of_one_pair_closed_form(oracle_mod, d):
    """One approaching pair (no periodic image within reach, width 0.06 > 2 (h_i + h_j)):
    mu = hh (v_i - v_j).(r_i - r_j) / (d^2 + eps hh^2) = -2 u d h / (d^2 + eps h^2), so
      alpha part / beta part = -alpha c_bar / (beta mu)        (independent of W and rho_bar)
      alpha part (eps1) / alpha part (eps2) = (d^2 + eps2 h^2) / (d^2 + eps1 h^2)."""
    h, u, gam = 0.01, 0.3, 1.4
    params = _pair_params(oracle_mod, 1, gam)
    st = _pair_state(d, h, u)
    part = lambda **kw: (_gas_accel(oracle_mod, params, st, **kw)
                         - _gas_accel(oracle_mod, params, st, alpha=0.0, beta=0.0, eps=kw["eps"]))
    for eps in (0.01, 1.0):
        Aa = part(alpha=1.0, beta=0.0, eps=eps)[:, 0]
        Ab = part(alpha=0.0, beta=1.0, eps=eps)[:, 0]
        mu = -2.0 * u * d * h / (d * d + eps * h * h)
        cbar = np.sqrt(gam) * params.cs
        assert np.all(Aa != 0)
        assert np.allclose(Aa / Ab, -cbar / mu, rtol=1e-12)
        # equal and opposite on the two particles
        assert Aa[0] == pytest.approx(-Aa[1], rel=1e-14)
        # the pair decelerates: particle 0 moves right and is pushed left
        assert Aa[0] < 0
    r = part(alpha=1.0, beta=0.0, eps=0.01)[:, 0] / part(alpha=1.0, beta=0.0, eps=1.0)[:, 0]
    assert np.allclose(r, (d * d + 1.0 * h * h) / (d * d + 0.01 * h * h), rtol=1e-12)


def test_viscosity_sound_speed_is_the_pair_mean(oracle_mod):
    """Adiabatic law, c = sqrt(gamma (gamma - 1) e): the alpha part follows (c_i + c_j) / 2."""
    d, h, u, gam = 0.012, 0.01, 0.3, 5.0 / 3.0
    params = _pair_params(oracle_mod, 0, gam)

    def alpha_part(e):
        st = _pair_state(d, h, u, e)
        return (_gas_accel(oracle_mod, params, st, alpha=1.0, beta=0.0)
                - _gas_accel(oracle_mod, params, st, alpha=0.0, beta=0.0))[:, 0]

    c = lambda e: np.sqrt(gam * (gam - 1.0) * e)
    same, mixed = alpha_part((1.0, 1.0)), alpha_part((1.0, 4.0))
    assert np.allclose(mixed / same, 0.5 * (c(1.0) + c(4.0)) / c(1.0), rtol=1e-12)
