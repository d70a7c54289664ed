"""Pins of the seeded initial-condition generators (CPU only): Plummer
moments (virial ratio, mass profile, velocity bound), the two-sphere
geometry, the planetary ensemble's orbital elements and the Kepler ->
Cartesian conversion against closed forms."""
import numpy as np
import pytest

from grav_sim_amd import ic
from oracle import oracle as O


def _potential_energy(x, m, G=1.0):
    e = 0.0
    for i in range(len(m) - 1):
        d = np.linalg.norm(x[i + 1:] - x[i], axis=1)
        e -= G * m[i] * np.sum(m[i + 1:] / d)
    return e


@pytest.mark.parametrize("n,seed", [(4096, 1), (8192, 2)])
def test_plummer_virial_equilibrium(n, seed):
    x, v, m = ic.plummer(n, seed)
    K = 0.5 * np.sum(m * np.sum(v * v, axis=1))
    W = _potential_energy(x, m)
    ratio = 2.0 * K / abs(W)
    assert abs(ratio - 1.0) < 3.0 / np.sqrt(n), ratio
    # total energy of a Plummer sphere (b = 1, M = 1, G = 1): -3 pi / 64
    assert abs((K + W) / (-3.0 * np.pi / 64.0) - 1.0) < 5.0 / np.sqrt(n)
    assert O.energy(x, v, m, 1.0) == pytest.approx(K + W, rel=1e-10)


def test_plummer_mass_profile_and_bounds():
    n = 20000
    x, v, m = ic.plummer(n, 3)
    assert np.allclose(m, 1.0 / n) and abs(m.sum() - 1.0) < 1e-12
    # centre of mass and momentum zeroed
    assert np.abs((m[:, None] * x).sum(axis=0)).max() < 1e-12
    assert np.abs((m[:, None] * v).sum(axis=0)).max() < 1e-12
    r = np.sort(np.linalg.norm(x, axis=1))
    # enclosed mass M(<r) = r^3 / (1 + r^2)^{3/2}: Kolmogorov-Smirnov distance
    cdf = r ** 3 / (1.0 + r * r) ** 1.5
    emp = np.arange(1, n + 1) / n
    # (sampling truncated at X <= 0.999: the model CDF is renormalised by 0.999;
    # the centre-of-mass shift moves radii by ~1/sqrt(n))
    assert np.max(np.abs(cdf / 0.999 - emp)) < 1.63 / np.sqrt(n) + 2e-3
    # speeds below the local escape speed
    rr = np.linalg.norm(x, axis=1)
    vesc = np.sqrt(2.0) * (rr * rr + 1.0) ** (-0.25)
    assert np.all(np.linalg.norm(v, axis=1) < vesc * 1.05)  # COM shift is tiny


def test_two_spheres_geometry():
    n = 8192
    x, v, m = ic.two_spheres(n, 4)
    h = n // 2
    c0, c1 = x[:h].mean(axis=0), x[h:].mean(axis=0)
    assert abs((c1 - c0)[0] - 10.0) < 0.2 and np.abs((c1 - c0)[1:]).max() < 0.2
    dv = v[h:].mean(axis=0) - v[:h].mean(axis=0)
    assert abs(dv[0] + 0.5) < 1e-12 and np.abs(dv[1:]).max() < 1e-12
    assert abs(m.sum() - 1.0) < 1e-12
    assert np.abs((m[:, None] * x).sum(axis=0)).max() < 1e-12


def test_keplerian_to_cartesian_closed_forms():
    rng = np.random.default_rng(5)
    k = 200
    a = rng.uniform(0.5, 30.0, k)
    e = rng.uniform(0.0, 0.9, k)
    inc, w, node, f = rng.uniform(0.0, np.pi, k), *rng.uniform(0.0, 2 * np.pi, (3, k))
    M = rng.uniform(0.5, 2.0, k)
    x, v = ic.keplerian_to_cartesian(a, e, inc, w, node, f, M, 1.0)
    r = np.linalg.norm(x, axis=1)
    s = np.linalg.norm(v, axis=1)
    # vis-viva and the orbit equation
    assert np.allclose(s * s, M * (2.0 / r - 1.0 / a), rtol=1e-12)
    assert np.allclose(r, a * (1 - e * e) / (1 + e * np.cos(f)), rtol=1e-12)
    # specific angular momentum and inclination
    hv = np.cross(x, v)
    assert np.allclose(np.linalg.norm(hv, axis=1), np.sqrt(M * a * (1 - e * e)), rtol=1e-12)
    assert np.allclose(np.arccos(np.clip(hv[:, 2] / np.linalg.norm(hv, axis=1), -1, 1)), inc, atol=1e-9)


def test_ensemble_orbital_elements():
    off, X, V, M = ic.ensemble(64, 3)
    counts = np.diff(off)
    assert np.all(counts >= 9) and np.all(counts <= 9 + 2000)
    for s in range(64):
        sl = slice(off[s], off[s + 1])
        x, v, m = X[sl], V[sl], M[sl]
        assert m[0] == 1.0 and np.all(m[9:] == 0.0)
        assert np.all((m[1:9] >= 1e-7) & (m[1:9] <= 1e-3))
        assert np.abs((m[:, None] * x).sum(axis=0)).max() < 1e-12
        # heliocentric elements of the test particles: a in [2, 3.5], e in [0, 0.2]
        rx, rv = x[9:] - x[0], v[9:] - v[0]
        r = np.linalg.norm(rx, axis=1)
        a = 1.0 / (2.0 / r - np.sum(rv * rv, axis=1) / 1.0)
        h = np.linalg.norm(np.cross(rx, rv), axis=1)
        ecc = np.sqrt(np.maximum(0.0, 1.0 - h * h / a))
        assert np.all((a > 2.0 - 1e-6) & (a < 3.5 + 1e-6))
        assert np.all(ecc < 0.2 + 1e-6)
