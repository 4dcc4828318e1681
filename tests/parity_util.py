"""Shared helpers of the parity tests: state injection and the comparison norms."""
from __future__ import annotations

import numpy as np

STATE_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m", "rho", "p", "e", "h", "t_stop", "ind")


def take_state(sim, species_list=(0, 1)) -> dict:
    """Full per-species state of a simulation (oracle or GPU), original particle order."""
    return {s: sim.download(s) for s in species_list}


def inject(sim_cls, params, state: dict):
    sim = sim_cls(params)
    for s, arrs in state.items():
        sim.upload(s, {k: arrs[k] for k in STATE_FIELDS})
    return sim


def rel_pointwise(a, b):
    """max |a-b| / |b| over entries with b != 0."""
    nz = b != 0
    if not np.any(nz):
        return float(np.max(np.abs(a - b))) if a.size else 0.0
    return float(np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz])))


def rel_normwise(a, b):
    """||a-b||_inf / ||b||_inf (used for velocities, which are 0 ahead of the shock)."""
    d = float(np.max(np.abs(a - b))) if a.size else 0.0
    s = float(np.max(np.abs(b))) if b.size else 0.0
    return d / s if s > 0 else d


def vec_normwise(a: dict, b: dict, keys=("vx", "vy", "vz"), mask=None):
    """max |a - b| / max |b| over particles and components of a vector field."""
    da = np.stack([a[k] for k in keys], 1)
    db = np.stack([b[k] for k in keys], 1)
    if mask is not None:
        da, db = da[mask], db[mask]
    s = float(np.max(np.abs(db))) if db.size else 0.0
    d = float(np.max(np.abs(da - db))) if db.size else 0.0
    return d / s if s > 0 else d


def momentum_error(cons_a, cons_b, states_b: dict):
    """|dP| / max(|P|, sum m|v|) with P the reference momentum."""
    P = np.asarray(cons_b[:3])
    scale = 0.0
    for arrs in states_b.values():
        scale += float(np.sum(arrs["m"] * np.sqrt(arrs["vx"] ** 2 + arrs["vy"] ** 2 + arrs["vz"] ** 2)))
    scale = max(float(np.linalg.norm(P)), scale, 1e-300)
    return float(np.linalg.norm(np.asarray(cons_a[:3]) - P)) / scale


def energy_error(cons_a, cons_b):
    """|dE| / |E| of the total energy sum (kinetic + internal), the fourth conservation entry."""
    return abs(float(cons_a[3]) - float(cons_b[3])) / max(abs(float(cons_b[3])), 1e-300)


def quantile_normwise(a: dict, b: dict, q: float, keys=("vx", "vy", "vz"), mask=None):
    """q-quantile over particles of max_k |a_k - b_k| / max |b| (a robust companion of vec_normwise)."""
    da = np.stack([a[k] for k in keys], 1)
    db = np.stack([b[k] for k in keys], 1)
    if mask is not None:
        da, db = da[mask], db[mask]
    s = float(np.max(np.abs(db))) if db.size else 0.0
    e = np.max(np.abs(da - db), axis=1) / (s if s > 0 else 1.0)
    return float(np.quantile(e, q)) if e.size else 0.0
