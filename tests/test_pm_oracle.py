"""Pins of the particle-mesh restatement (oracle/pm.py), CPU only.

The reference's PM needs FFTW3, absent from this image, so its outputs
cannot be produced here; each stage is pinned by what the mathematics
fixes instead, chosen so that a dropped factor, a wrong sign, a transposed
axis or an off-by-one in the stencil or the CIC corners fails a test.
"""
import math

import numpy as np
import pytest

from oracle import pm as P


def test_wrap_maps_into_the_box_by_whole_periods():
    rng = np.random.default_rng(1)
    c, w = np.array([0.3, -1.0, 2.0]), 0.75
    x = c + rng.uniform(-7.0, 7.0, size=(5000, 3))
    y = P.wrap(x, c, w)
    assert np.all(np.abs(y - c) <= w * (1 + 1e-12))
    k = (y - x) / (2 * w)
    assert np.allclose(k, np.round(k), atol=1e-9)
    inside = np.abs(x - c) <= w
    assert np.array_equal(y[inside], x[inside])


def test_cic_deposit_corners_centres_and_mass():
    n, L = 8, 2.0
    cell = L / n
    c = np.zeros(3)
    rho = 1.0 / L ** 3
    # a particle on a grid node puts all its mass in that cell
    x = np.array([[-1.0 + 3 * cell, -1.0 + 5 * cell, -1.0 + 1 * cell]])
    d = P.cic_deposit(x, np.array([1.0]), n, c, L, rho)
    dens = (d + 1.0) * cell ** 3 * rho
    assert dens[3 * n * n + 5 * n + 1] == pytest.approx(1.0, abs=1e-15) and np.count_nonzero(dens > 1e-15) == 1
    # at a cell centre it is shared equally by the 8 surrounding nodes
    x = np.array([[-1.0 + 2.5 * cell, -1.0 + 6.5 * cell, -1.0 + 7.5 * cell]])
    dens = (P.cic_deposit(x, np.array([2.0]), n, c, L, rho) + 1.0) * cell ** 3 * rho
    nz = np.flatnonzero(dens > 1e-15)
    assert len(nz) == 8 and np.allclose(dens[nz], 0.25)
    # (the z index wraps at the upper face)
    assert {(i // (n * n), (i // n) % n, i % n) for i in nz} == {(a, b, cc) for a in (2, 3) for b in (6, 7)
                                                                 for cc in (7, 0)}
    # random particles: total mass conserved
    rng = np.random.default_rng(2)
    x = rng.uniform(-1.0, 1.0, size=(3000, 3))
    m = rng.uniform(0.5, 1.5, 3000)
    dens = (P.cic_deposit(x, m, n, c, L, rho) + 1.0) * cell ** 3 * rho
    assert dens.sum() == pytest.approx(m.sum(), rel=1e-13)


@pytest.mark.parametrize("mode", [(1, 0, 0), (0, 2, 0), (1, 1, 3), (4, 0, 0)])
def test_poisson_single_mode_closed_form(mode):
    """delta = cos(k.x) -> phi = -(3/2 Om H0^2 / a) cos(k.x) / |k|^2."""
    n, L, H0, om, a = 16, 3.0, 0.7, 0.3, 0.5
    g = np.arange(n) * (L / n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    k = 2 * np.pi / L * np.array(mode)
    delta = np.cos(k[0] * X + k[1] * Y + k[2] * Z)
    phi = P.potential(delta.ravel(), n, L, H0, om, a).reshape(n, n, n)
    want = -(1.5 * om * H0 * H0 / a) * delta / np.dot(k, k)
    assert np.allclose(phi, want, rtol=0, atol=1e-13 * np.abs(want).max())


@pytest.mark.parametrize("axis,kk", [(0, 1), (1, 3), (2, 5)])
def test_gradient_stencil_transfer_function(axis, kk):
    """The 4th-order stencil on sin(k x) gives (8 sin kh - sin 2kh)/(6h) cos(k x)
    along its own axis and 0 along the others."""
    n, L = 32, 2.0
    h = L / n
    g = np.arange(n) * h
    grids = np.meshgrid(g, g, g, indexing="ij")
    k = 2 * np.pi * kk / L
    phi = np.sin(k * grids[axis])
    acc = P.gradient(phi.ravel(), n, L).reshape(n, n, n, 3)
    want = (8 * np.sin(k * h) - np.sin(2 * k * h)) / (6 * h) * np.cos(k * grids[axis])
    assert np.allclose(acc[..., axis], want, atol=1e-12)
    for other in {0, 1, 2} - {axis}:
        assert np.abs(acc[..., other]).max() < 1e-12


def test_cic_interp_reproduces_linear_fields():
    n, L = 16, 4.0
    h = L / n
    c = np.array([0.0, 0.0, 0.0])
    nodes = -L / 2 + np.arange(n) * h
    X, Y, Z = np.meshgrid(nodes, nodes, nodes, indexing="ij")
    grid = np.stack([1.0 + 2.0 * X - Y, 0.5 * Z, -3.0 + X + Y + Z], axis=-1).reshape(n ** 3, 3)
    rng = np.random.default_rng(3)
    x = rng.uniform(-L / 2, L / 2 - 1.01 * h, size=(2000, 3))  # away from the periodic seam
    a = P.cic_interp(x, grid, n, c, L)
    want = np.stack([1.0 + 2.0 * x[:, 0] - x[:, 1], 0.5 * x[:, 2], -3.0 + x.sum(1)], axis=1)
    assert np.allclose(a, want, atol=1e-12)


def _point_mass_case(n=64, L=1.0, r_cells=(4, 6, 8, 12)):
    c = np.full(3, L / 2)
    M, mt = 1.0, 1e-12
    src = c + 0.5 * L / n  # at a cell centre (off the nodes)
    xs = [src]
    for r in r_cells:
        xs.append(src + np.array([r * L / n, 0.0, 0.0]))
    x = np.array(xs)
    m = np.array([M] + [mt] * len(r_cells))
    return x, m, c, L


def test_point_mass_field_is_newtonian_and_momentum_conserved():
    """A point mass in a periodic box: at a few cells the PM field is the
    Newtonian G_eff M / r^2 with 4 pi G_eff rho = 3/2 Om H0^2 / a, pointing
    away from the mass (acceleration_PM returns +grad phi; the integrator
    subtracts it); the total force on a random set sums to zero."""
    n, H0, om, a = 64, 1.0, 1.0, 1.0
    x, m, c, L = _point_mass_case(n)
    rho = m.sum() / L ** 3
    acc = P.acceleration_pm(x, m, c, L / 2, H0, om, rho, n, a)
    G = 1.5 * om * H0 * H0 / a / (4 * np.pi * rho)
    r = x[1:, 0] - x[0, 0]
    newton = G * m[0] / r ** 2
    rel = acc[1:, 0] / newton
    # softened by CIC at a few cells, periodic images (~1e-3 at r = L/5) beyond
    assert np.all(rel > 0.85) and np.all(rel < 1.05), rel
    assert abs(rel[2] - 1.0) < 0.03, rel
    rng = np.random.default_rng(4)
    x = rng.uniform(0.0, L, size=(4000, 3))
    m = rng.uniform(0.5, 1.5, 4000)
    acc = P.acceleration_pm(x, m, c, L / 2, H0, om, m.sum() / L ** 3, 32, a)
    tot = (m[:, None] * acc).sum(0)
    assert np.abs(tot).max() < 1e-11 * (m[:, None] * np.abs(acc)).sum()


def test_compute_da_closed_forms():
    H0 = 0.7
    for a in (0.1, 0.5, 1.0, 2.0):
        # Einstein-de Sitter: da/dt = H0 a^(-1/2)
        assert P.compute_da(a, H0, 1.0, 0.0) == pytest.approx(H0 / math.sqrt(a), rel=1e-15)
        # de Sitter: da/dt = H0 a
        assert P.compute_da(a, H0, 0.0, 1.0) == pytest.approx(H0 * a, rel=1e-15)


def zeldovich_case(n=16, amp=0.15, a0=0.05, a1=0.25, steps=40):
    """Plane wave in Einstein-de Sitter: x = q + D(a) psi(q), D = a,
    psi = -amp sin(k q) / k along x; p = a^2 v = H0 a^(3/2) psi."""
    L = 1.0
    h = 0.01  # H0 = 1
    H0 = h * 100.0
    q1 = (np.arange(n) + 0.5) * (L / n)
    Q = np.stack(np.meshgrid(q1, q1, q1, indexing="ij"), axis=-1).reshape(-1, 3)
    k = 2 * np.pi / L
    psi = np.zeros_like(Q)
    psi[:, 0] = -amp * np.sin(k * Q[:, 0]) / k
    x0 = Q + a0 * psi
    v0 = H0 * a0 ** -0.5 * psi  # dx/dt = (da/dt) psi
    m = np.full(len(Q), 1.0 / len(Q))
    c = np.full(3, L / 2)
    return dict(Q=Q, psi=psi, x=x0, v=v0, m=m, c=c, bw=L / 2, h=h, a0=a0, a1=a1, dt=(a1 - a0) / steps, n=n)


def test_zeldovich_plane_wave_through_the_integrator():
    """Before shell crossing the Zel'dovich solution is exact for a plane
    wave; the PM integrator must follow it to the discretisation error."""
    z = zeldovich_case()
    x, steps, t = P.leapfrog_cosmology(z["x"], z["v"], z["m"], z["c"], z["bw"], z["h"], 1.0, 0.0, z["dt"],
                                       z["a0"], z["a1"], z["n"])
    assert steps == 40 and t == pytest.approx(z["a1"])
    want = z["Q"] + z["a1"] * z["psi"]
    disp = (x - z["Q"])[:, 0]
    err = np.abs(disp - (want - z["Q"])[:, 0]).max() / np.abs(z["a1"] * z["psi"][:, 0]).max()
    assert err < 0.02, err
    # a plane wave along x: nothing moves across it
    assert np.abs(x[:, 1:] - z["Q"][:, 1:]).max() < 1e-12
