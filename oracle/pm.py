"""CPU restatement of the particle-mesh force and the cosmological Leapfrog.

Test infrastructure only (tests/, bench.py's CPU baseline): nothing in the
library imports this.

Follows, step for step and in the same floating-point operation order:

* ``wrap``               set_periodic_boundary_conditions, src/system.c:1337-1380
* ``cic_deposit``        cloud_in_cell, src/acceleration_PM.c:15-105 (the
                         serial order: particle by particle, corners in the
                         reference's order; the reference's OpenMP loop adds
                         without synchronisation, so it has no fixed order)
* ``potential``          src/acceleration_PM.c:301-350: r2c FFT, multiply by
                         -1/k^2, c2r FFT, scale by 3/2 Omega_m H0^2 / (a n^3)
* ``gradient``           compute_acceleration_with_gradient, :212-272
                         (4th-order central differences, +grad phi)
* ``cic_interp``         get_cloud_in_cell_acceleration, :107-210
* ``acceleration_pm``    acceleration_PM, :274-391
* ``compute_da``         src/cosmology.c:5-26
* ``leapfrog_cosmology`` src/integrator_cosmology_leapfrog.c:27-255 (output
                         disabled)

The reference's FFTs are FFTW3's, which this image does not have: the
reference's PM cannot be built here, and the transforms below are numpy's
(pocketfft), equal to FFTW's to rounding.  The restatement is pinned by
closed forms in tests/test_pm_oracle.py: a single Fourier mode through the
Poisson solve, the stencil's transfer function on a sine, CIC exactness on
corners, centres and linear fields, momentum conservation, the point-mass
field against the Newtonian force, and the Zel'dovich solution of a plane
wave in an Einstein-de Sitter universe through the whole integrator.
"""
import math

import numpy as np

# corners in the reference's order (src/acceleration_PM.c:58-97)
CORNERS = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1))


def wrap(x, center, width):
    """Periodic boundary conditions with the reference's fmod / parity formula."""
    x = np.array(x, dtype=np.float64, copy=True)
    c = np.asarray(center, dtype=np.float64)
    for d in range(3):
        p = x[:, d] - c[d]
        lo, hi = p < -width, p > width
        q = np.fmod(np.trunc(p / width), 2.0) != 0.0
        f = np.fmod(p, width)
        x[lo, d] = np.where(q[lo], f[lo] + width, f[lo]) + c[d]
        x[hi, d] = np.where(q[hi], f[hi] - width, f[hi]) + c[d]
    return x


def _cells(x, n, center, box_length):
    bw = box_length / 2.0
    inv = 1.0 / (box_length / n)
    c = np.asarray(center, dtype=np.float64)
    xn = ((x - c) + bw) * inv
    i = xn.astype(np.int64)  # truncation, as the reference's (int) cast
    w = xn - i
    return i, w, 1.0 - w


def _corner_index(i, n, dx, dy, dz):
    return ((i[:, 0] + dx) % n) * (n * n) + ((i[:, 1] + dy) % n) * n + ((i[:, 2] + dz) % n)


def _corner_weight(w, wm, dx, dy, dz):
    return ((w[:, 0] if dx else wm[:, 0]), (w[:, 1] if dy else wm[:, 1]), (w[:, 2] if dz else wm[:, 2]))


def cic_deposit(x, m, n, center, box_length, mean_bkg_density):
    """Density contrast delta on the n^3 grid (flat, x-major)."""
    i, w, wm = _cells(x, n, center, box_length)
    idx = np.empty((len(m), 8), dtype=np.int64)
    val = np.empty((len(m), 8))
    for c, (dx, dy, dz) in enumerate(CORNERS):
        wx, wy, wz = _corner_weight(w, wm, dx, dy, dz)
        idx[:, c] = _corner_index(i, n, dx, dy, dz)
        val[:, c] = m * wx * wy * wz
    delta = np.zeros(n ** 3)
    np.add.at(delta, idx.ravel(), val.ravel())  # particle by particle, corners in order
    cell = box_length / n
    return (delta / (cell * cell * cell * mean_bkg_density)) - 1.0


def potential(delta, n, box_length, H0, omega_m, scale_factor):
    spec = np.fft.rfftn(np.asarray(delta).reshape(n, n, n))
    half = n // 2 + 1
    i = np.arange(n)
    k1 = np.where(i < half, i, i - n)
    kz = np.arange(half)
    two_pi_l2 = 4.0 * math.pi * math.pi / (box_length * box_length)
    ksq = (k1[:, None, None] * k1[:, None, None] + k1[None, :, None] * k1[None, :, None]
           + kz[None, None, :] * kz[None, None, :]) * two_pi_l2
    with np.errstate(divide="ignore"):
        kernel = np.where(ksq == 0.0, 0.0, -1.0 / ksq)
    spec = spec * kernel
    phi = np.fft.irfftn(spec, s=(n, n, n), axes=(0, 1, 2)) * float(n ** 3)  # unnormalised c2r, as FFTW's
    factor = ((3.0 / 2.0) * omega_m * H0 * H0) / (scale_factor * float(n ** 3))
    return (phi * factor).ravel()


def gradient(phi, n, box_length):
    """(n^3, 3): +grad phi by 4th-order central differences."""
    cell = box_length / n
    p = np.asarray(phi).reshape(n, n, n)
    out = np.empty((n, n, n, 3))
    for ax in range(3):
        out[..., ax] = (np.roll(p, 2, ax) - 8.0 * np.roll(p, 1, ax) + 8.0 * np.roll(p, -1, ax)
                        - np.roll(p, -2, ax)) / (12.0 * cell)
    return out.reshape(n ** 3, 3)


def cic_interp(x, acc_grid, n, center, box_length):
    i, w, wm = _cells(x, n, center, box_length)
    a = np.zeros((len(x), 3))
    for c, (dx, dy, dz) in enumerate(CORNERS):
        wx, wy, wz = _corner_weight(w, wm, dx, dy, dz)
        g = acc_grid[_corner_index(i, n, dx, dy, dz)]
        t = g * wx[:, None] * wy[:, None] * wz[:, None]
        a = t if c == 0 else a + t
    return a


def acceleration_pm(x, m, center, box_width, H0, omega_m, mean_bkg_density, n, scale_factor):
    L = box_width * 2.0
    delta = cic_deposit(x, m, n, center, L, mean_bkg_density)
    phi = potential(delta, n, L, H0, omega_m, scale_factor)
    return cic_interp(x, gradient(phi, n, L), n, center, L)


def compute_da(a, H0, omega_m, omega_lambda):
    omega_k = 1.0 - omega_m - omega_lambda
    return H0 * a * math.sqrt(omega_lambda + omega_m / (a * a * a) + omega_k / (a * a))


def leapfrog_cosmology(x, v, m, center, box_width, h, omega_m, omega_lambda, dt, a_begin, a_final, n):
    """Returns (x, num_steps, t).  Velocities are left as given: the
    reference converts momenta back only for snapshots."""
    H0 = h * 100.0
    x = np.array(x, dtype=np.float64, copy=True)
    L = box_width * 2.0
    total_m = 0.0
    for mi in m:  # vec_sum, src/math_functions.c:13-21
        total_m += mi
    rho = total_m / (L * L * L)
    t = a_begin
    x = wrap(x, center, box_width)
    p = (t * t) * np.asarray(v, dtype=np.float64)
    acc = acceleration_pm(x, m, center, box_width, H0, omega_m, rho, n, t)
    total = int(math.ceil((a_final - a_begin) / dt))
    steps = 0
    while steps < total:
        if t + dt > a_final:
            dt = a_final - t
        da = compute_da(t, H0, omega_m, omega_lambda)
        p = p - (0.5 * dt) * acc / da
        t += 0.5 * dt
        da = compute_da(t, H0, omega_m, omega_lambda)
        x = x + dt * p / ((t * t) * da)
        x = wrap(x, center, box_width)
        acc = acceleration_pm(x, m, center, box_width, H0, omega_m, rho, n, t)
        p = p - (0.5 * dt) * acc / da
        steps += 1
        t = a_begin + steps * dt
    return x, steps, t
