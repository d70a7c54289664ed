"""Seeded synthetic initial conditions for the force-evaluation workloads.

These generators have no counterpart in the reference (its examples draw
uniform cubes from a time-seeded PCG32, examples/benchmark/c/benchmark.c:27-34);
they produce the deterministic inputs the workloads are defined on:

* ``plummer``      Plummer sphere, G=1, equal masses, Aarseth-Henon-Wielen
                   sampling, truncated at X<=0.999, centre of mass and
                   momentum zeroed.
* ``two_spheres``  two Plummer spheres approaching each other.
* ``ensemble``     independent planetary systems with massless test particles,
                   built with the reference's Kepler->Cartesian conversion
                   (src/utils.c:82-141).
* ``uniform_cube`` the reference benchmark's distribution (uniform in [-1,1)^3).

All use numpy's PCG64 generator, which is reproducible across platforms.
"""
import numpy as np


def _unit_vectors(rng, n):
    cz = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    sz = np.sqrt(1.0 - cz * cz)
    return np.stack([sz * np.cos(phi), sz * np.sin(phi), cz], axis=1)


def _plummer_raw(rng, n, mass, b, G):
    x_m = rng.uniform(0.0, 0.999, n)
    x_m = np.where(x_m <= 0.0, 1e-300, x_m)
    r = b / np.sqrt(x_m ** (-2.0 / 3.0) - 1.0)
    pos = r[:, None] * _unit_vectors(rng, n)
    # speed ratio q = v / v_esc, density g(q) ~ q^2 (1 - q^2)^{7/2} (rejection)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        qa = rng.uniform(0.0, 1.0, todo.size)
        ya = rng.uniform(0.0, 0.1, todo.size)
        ok = ya < qa * qa * (1.0 - qa * qa) ** 3.5
        q[todo[ok]] = qa[ok]
        todo = todo[~ok]
    v_esc = np.sqrt(2.0 * G * mass) * (r * r + b * b) ** (-0.25)
    vel = (q * v_esc)[:, None] * _unit_vectors(rng, n)
    return pos, vel


def _zero_com(x, v, m):
    mt = m.sum()
    x -= (m[:, None] * x).sum(axis=0) / mt
    v -= (m[:, None] * v).sum(axis=0) / mt


def plummer(n, seed, b=1.0, G=1.0, total_mass=1.0):
    """Plummer sphere of n equal-mass particles (mass total_mass/n)."""
    rng = np.random.default_rng(seed)
    m = np.full(n, total_mass / n)
    x, v = _plummer_raw(rng, n, total_mass, b, G)
    _zero_com(x, v, m)
    return np.ascontiguousarray(x), np.ascontiguousarray(v), m


def two_spheres(n, seed, separation=10.0, approach_speed=0.5, b=1.0, G=1.0):
    """Two Plummer spheres of n/2 particles each (every particle mass 1/n),
    centres at (-separation/2, 0, 0) and (+separation/2, 0, 0), moving towards
    each other at relative speed approach_speed."""
    assert n % 2 == 0
    rng = np.random.default_rng(seed)
    h = n // 2
    m = np.full(n, 1.0 / n)
    xs, vs = [], []
    for sgn in (-1.0, 1.0):
        x, v = _plummer_raw(rng, h, 0.5, b, G)
        mm = np.full(h, 1.0 / n)
        _zero_com(x, v, mm)
        x[:, 0] += sgn * separation / 2.0
        v[:, 0] -= sgn * approach_speed / 2.0
        xs.append(x)
        vs.append(v)
    x = np.concatenate(xs)
    v = np.concatenate(vs)
    _zero_com(x, v, m)
    return np.ascontiguousarray(x), np.ascontiguousarray(v), m


def uniform_cube(n, seed, lo=-1.0, hi=1.0):
    """Uniform positions in [lo, hi)^3, unit masses, zero velocity."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(lo, hi, (n, 3))
    return x, np.zeros((n, 3)), np.ones(n)


def keplerian_to_cartesian(a, e, inc, argp, node, f, total_mass, G):
    """Vectorised form of src/utils.c:82-141 (same formulae)."""
    ci, si = np.cos(inc), np.sin(inc)
    cw, sw = np.cos(argp), np.sin(argp)
    cn, sn = np.cos(node), np.sin(node)
    cf, sf = np.cos(f), np.sin(f)
    ecc = np.stack([cn * cw - sn * sw * ci, sn * cw + cn * sw * ci, sw * si], axis=-1)
    qv = np.stack([-cn * sw - sn * cw * ci, -sn * sw + cn * cw * ci, cw * si], axis=-1)
    rr = a * ((1.0 - e * e) / (1.0 + e * cf))
    x = rr[..., None] * (cf[..., None] * ecc + sf[..., None] * qv)
    vv = np.sqrt(G * total_mass / (a * (1.0 - e * e)))
    v = vv[..., None] * (-sf[..., None] * ecc + (e + cf)[..., None] * qv)
    return x, v


def ensemble(num_systems, seed, num_planets=8, max_test=2000, G=1.0):
    """Independent planetary systems (massive particles first, then tests).

    Each system: central mass 1 at the origin; num_planets planets with
    m ~ logU[1e-7, 1e-3], a ~ logU[0.4, 30], e ~ U(0, 0.1), i ~ U(0, 5 deg);
    T ~ U{0..max_test} massless test particles with a ~ U(2, 3.5),
    e ~ U(0, 0.2), i ~ U(0, 5 deg); all angles U(0, 2 pi).  Centre of mass and
    momentum are zeroed per system.

    Returns (offsets, x, v, m): system s holds particles
    offsets[s]..offsets[s+1]-1 of the flat (P,3)/(P,) arrays.
    """
    rng = np.random.default_rng(seed)
    deg = np.pi / 180.0
    counts = 1 + num_planets + rng.integers(0, max_test + 1, num_systems)
    offsets = np.zeros(num_systems + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(counts)
    P = int(offsets[-1])
    x = np.zeros((P, 3))
    v = np.zeros((P, 3))
    m = np.zeros(P)
    for s in range(num_systems):
        o = int(offsets[s])
        nt = int(counts[s]) - 1 - num_planets
        m[o] = 1.0
        mp = np.exp(rng.uniform(np.log(1e-7), np.log(1e-3), num_planets))
        ap = np.exp(rng.uniform(np.log(0.4), np.log(30.0), num_planets))
        ep = rng.uniform(0.0, 0.1, num_planets)
        ip = rng.uniform(0.0, 5.0 * deg, num_planets)
        ang = rng.uniform(0.0, 2.0 * np.pi, (3, num_planets))
        xp, vp = keplerian_to_cartesian(ap, ep, ip, ang[0], ang[1], ang[2], 1.0 + mp, G)
        m[o + 1 : o + 1 + num_planets] = mp
        x[o + 1 : o + 1 + num_planets] = xp
        v[o + 1 : o + 1 + num_planets] = vp
        if nt:
            at = rng.uniform(2.0, 3.5, nt)
            et = rng.uniform(0.0, 0.2, nt)
            it = rng.uniform(0.0, 5.0 * deg, nt)
            angt = rng.uniform(0.0, 2.0 * np.pi, (3, nt))
            xt, vt = keplerian_to_cartesian(at, et, it, angt[0], angt[1], angt[2], np.ones(nt), G)
            x[o + 1 + num_planets : o + counts[s]] = xt
            v[o + 1 + num_planets : o + counts[s]] = vt
        sl = slice(o, o + int(counts[s]))
        xs, vs, ms = x[sl], v[sl], m[sl]
        _zero_com(xs, vs, ms)
    return offsets, x, v, m
