"""Pins of the oracle's neighbour machinery (kernel.cu:1937-1995, IDIC.h:17-45) by brute force.

- sorted order: (cell key, particle id), the order of the reference's linked lists;
- cell keys: every particle lies inside its cell (up to the clamp at the x ends);
- neighbour sets: an O(N^2) search over all particles and periodic images, with the
  predicate d2 < (h_i + h_j)^2, evaluated exactly (fma emulated with rationals) for
  pairs within rounding of the cutoff;
- density: the brute-force sum of m_j W(r_ij, (h_i + h_j)/2) over that set.
"""
from fractions import Fraction

import numpy as np
import pytest

from dustyshock3d_amd import configs
from tests.parity_util import STATE_FIELDS

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
PI_REF = 3.14159265


def mix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def pair_tag(ids, sy, sz):
    ids = np.asarray(ids, dtype=np.uint64)
    sy = np.asarray(np.asarray(sy) + 128, dtype=np.uint64)
    sz = np.asarray(np.asarray(sz) + 128, dtype=np.uint64)
    return mix64((ids << np.uint64(16)) | (sy << np.uint64(8)) | sz)


def exact_pred(dx, dy, dz, hs):
    """d2 = fma(dz,dz,fma(dy,dy,dx*dx)) < fl(hs*hs), with each rounding done exactly."""
    a = float(Fraction(dx) * Fraction(dx))
    b = float(Fraction(dy) * Fraction(dy) + Fraction(a))
    d2 = float(Fraction(dz) * Fraction(dz) + Fraction(b))
    return d2 < hs * hs


def brute_force(state, L, imgs=3):
    x, y, z, h, m, ind = (state[k] for k in ("x", "y", "z", "h", "m", "ind"))
    n = len(x)
    cnt = np.zeros(n, np.int64)
    hsh = np.zeros(n, np.uint64)
    rho = np.zeros(n)
    for i in np.nonzero(ind == 0)[0]:
        for sy in range(-imgs, imgs + 1):
            for sz in range(-imgs, imgs + 1):
                dx = x - x[i]
                dy = (y - y[i]) + sy * L[1]
                dz = (z - z[i]) + sz * L[2]
                hs = h[i] + h
                d2 = dx * dx + dy * dy + dz * dz
                H2 = hs * hs
                near = np.abs(d2 - H2) <= 1e-9 * H2
                inside = (d2 < H2) & ~near
                for j in np.nonzero(near)[0]:
                    inside[j] = exact_pred(dx[j], dy[j], dz[j], hs[j])
                js = np.nonzero(inside)[0]
                if len(js) == 0:
                    continue
                cnt[i] += len(js)
                with np.errstate(over="ignore"):
                    hsh[i] += np.sum(pair_tag(js, sy, sz), dtype=np.uint64)
                hh = 0.5 * (h[i] + h[js])
                q = np.sqrt(d2[js]) / hh
                w = np.where(q < 1, 1 - 1.5 * q * q + 0.75 * q ** 3, np.where(q <= 2, 0.25 * (2 - q) ** 3, 0.0))
                rho[i] += np.sum(m[js] * w / (PI_REF * hh ** 3))
    return cnt, hsh, rho


@pytest.fixture(scope="module")
def small_step(oracle_mod):
    spec = configs.tube(0.03, 0.01, 1.0, 1.0, max_steps=4)
    o = oracle_mod.OracleSim.from_spec(spec)
    o.step(2)   # move the particles off the lattice
    state = {s: o.download(s) for s in (0, 1)}
    o2 = oracle_mod.OracleSim(o.params)
    for s in (0, 1):
        o2.upload(s, {k: state[s][k] for k in STATE_FIELDS})
    o2.step(1)
    return o2, state


def test_sorted_order_is_key_then_id(small_step):
    o, state = small_step
    g = o.grid()
    for s in (0, 1):
        after = o.download(s)
        key = after["cell"]
        ids = o.sorted_ids(s)
        expect = np.lexsort((np.arange(len(key)), key))
        assert np.array_equal(ids, expect)
        # nn is the linked list: next particle in sorted order within the same cell
        nn = after["nn"]
        for k in range(len(ids) - 1):
            a, b = ids[k], ids[k + 1]
            assert nn[a] == (b if key[a] == key[b] else -1)
        assert nn[ids[-1]] == -1
        # every particle lies inside the cell its key names
        n0, n1 = g["n"][0], g["n"][1]
        kx, ky, kz = key % n0, (key // n0) % n1, key // (n0 * n1)
        for a, kk, pos in ((0, kx, state[s]["x"]), (1, ky, state[s]["y"]), (2, kz, state[s]["z"])):
            lo = g["origin"][a] + kk * g["cell"][a]
            hi = lo + g["cell"][a]
            inside = (pos >= lo - 1e-12) & (pos < hi + 1e-12)
            clamped = (a == 0) & (((kk == 0) & (pos < lo)) | ((kk == g["n"][a] - 1) & (pos >= hi)))
            assert np.all(inside | clamped)


def test_neighbour_sets_brute_force(small_step):
    o, state = small_step
    L = o.grid()["length"]
    for s in (0, 1):
        cnt_o, hsh_o = o.neighbour_stats(s)
        cnt_b, hsh_b, rho_b = brute_force(state[s], L)
        act = state[s]["ind"] == 0
        assert cnt_b[act].min() > 20
        assert np.array_equal(cnt_o[act], cnt_b[act])
        assert np.array_equal(hsh_o[act], hsh_b[act])
        rho_o = o.download(s)["rho"]
        assert np.max(np.abs(rho_o[act] - rho_b[act]) / rho_b[act]) < 1e-13


def test_thin_tube_uses_periodic_images(small_step):
    """2h exceeds the tube width on the coarse side, so images of one particle both count."""
    o, state = small_step
    L = o.grid()["length"]
    assert 2 * state[0]["h"].max() > L[1]
    cnt_img, _, _ = brute_force(state[0], L, imgs=3)
    cnt_min, _, _ = brute_force(state[0], L, imgs=0)
    assert np.any(cnt_img > cnt_min)


def test_pair_tag_is_injective_over_ids_and_images():
    """The neighbour-set hash sums one tag per (id, periodic image): distinct neighbours must
    give distinct tags for every image offset the search can reach (at most 64 cells)."""
    ids = np.array([0, 1, 2, 255, 256, 4095, 4096, 65535, 65536, 2812499, 2 ** 31 - 1])
    off = np.arange(-64, 65)
    I, SY, SZ = np.meshgrid(ids, off, off, indexing="ij")
    tags = pair_tag(I.ravel(), SY.ravel(), SZ.ravel())
    assert len(np.unique(tags)) == tags.size
