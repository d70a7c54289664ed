"""Particle-mesh gravity and the cosmological Leapfrog on the GPU, through
the C-ABI, against the CPU restatement (oracle/pm.py).

The deposit sums in another order than the oracle's and the transforms are
hipFFT's, so forces agree with the oracle to rounding, not bitwise: the
tolerance is 1e-10 of the largest acceleration.  The step arithmetic of the
integrator is the reference's, so trajectories agree to the same order.
"""
import ctypes as C

import numpy as np
import pytest

from grav_sim_amd import _abi
from grav_sim_amd import lib as L
from oracle import pm as P

pytestmark = pytest.mark.gpu


def _clustered(n, seed, L_box=2.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-L_box / 2, L_box / 2, size=(n, 3))
    k = n // 3
    x[:k] = 0.3 + 0.05 * rng.standard_normal((k, 3))  # a dense clump
    x[k:k + 5] = L_box / 2 - 1e-12  # at the upper faces: wrap to node 0
    return x, rng.uniform(0.5, 1.5, n) / n


@pytest.mark.parametrize("grid,center", [(32, (0.0, 0.0, 0.0)), (64, (0.0, 0.0, 0.0)), (48, (1.0, -2.0, 0.5))])
def test_acceleration_pm_matches_oracle(grid, center):
    print("pm case", grid, center, flush=True)
    x, m = _clustered(20000, grid)
    x = x + np.asarray(center)
    bw, H0, om, a = 1.0, 0.7, 0.3, 0.4
    rho = m.sum() / (2 * bw) ** 3
    ag = L.acceleration_pm(x, m, center, bw, H0, om, rho, grid, a)
    ao = P.acceleration_pm(x, m, np.asarray(center), bw, H0, om, rho, grid, a)
    err = np.abs(ag - ao).max() / np.abs(ao).max()
    assert err < 1e-10, err
    # momentum: the total force vanishes
    tot = (m[:, None] * ag).sum(0)
    assert np.abs(tot).max() < 1e-11 * (m[:, None] * np.abs(ag)).sum()


def test_acceleration_pm_is_bitwise_reproducible():
    """The density is gathered per cell in a fixed order (no atomics), so
    repeated evaluations give the same doubles."""
    x, m = _clustered(50000, 5)
    args = (x, m, (0.0, 0.0, 0.0), 1.0, 0.7, 0.3, m.sum() / 8.0, 64, 0.5)
    a1 = L.acceleration_pm(*args)
    a2 = L.acceleration_pm(*args)
    assert np.array_equal(a1, a2)


def test_point_mass_field_is_newtonian():
    n, H0, om, a = 64, 1.0, 1.0, 1.0
    Lb = 1.0
    c = np.full(3, Lb / 2)
    src = c + 0.5 * Lb / n
    r_cells = (4, 6, 8, 12)
    x = np.array([src] + [src + np.array([r * Lb / n, 0.0, 0.0]) for r in r_cells])
    m = np.array([1.0] + [1e-12] * len(r_cells))
    rho = m.sum() / Lb ** 3
    acc = L.acceleration_pm(x, m, c, Lb / 2, H0, om, rho, n, a)
    G = 1.5 * om * H0 * H0 / a / (4 * np.pi * rho)
    rel = acc[1:, 0] / (G * m[0] / (x[1:, 0] - x[0, 0]) ** 2)
    assert np.all(rel > 0.85) and np.all(rel < 1.05) and abs(rel[2] - 1.0) < 0.03, rel


def _zeldovich(n=16, amp=0.15, a0=0.05, a1=0.25, steps=40):
    Lb, h = 1.0, 0.01
    H0 = h * 100.0
    q1 = (np.arange(n) + 0.5) * (Lb / n)
    Q = np.stack(np.meshgrid(q1, q1, q1, indexing="ij"), axis=-1).reshape(-1, 3)
    k = 2 * np.pi / Lb
    psi = np.zeros_like(Q)
    psi[:, 0] = -amp * np.sin(k * Q[:, 0]) / k
    return dict(Q=Q, psi=psi, x=Q + a0 * psi, v=H0 * a0 ** -0.5 * psi, m=np.full(len(Q), 1.0 / len(Q)),
                c=np.full(3, Lb / 2), bw=Lb / 2, h=h, a0=a0, a1=a1, dt=(a1 - a0) / steps, n=n)


def test_zeldovich_plane_wave_matches_oracle_and_closed_form():
    z = _zeldovich()
    xg, steps, t = L.leapfrog_cosmology(z["x"], z["v"], z["m"], z["c"], z["bw"], z["h"], 1.0, 0.0, z["dt"],
                                        z["a0"], z["a1"], z["n"])
    xo, so, to = P.leapfrog_cosmology(z["x"], z["v"], z["m"], z["c"], z["bw"], z["h"], 1.0, 0.0, z["dt"],
                                      z["a0"], z["a1"], z["n"])
    assert steps == so == 40 and t == to
    assert np.abs(xg - xo).max() < 1e-11
    disp = (xg - z["Q"])[:, 0]
    err = np.abs(disp - z["a1"] * z["psi"][:, 0]).max() / np.abs(z["a1"] * z["psi"][:, 0]).max()
    assert err < 0.02, err


def test_lcdm_run_through_launch_matches_oracle():
    """A clustered box in LambdaCDM (h=0.7, Om=0.3, OL=0.7) from a=0.1 to 0.3,
    particles leaving the box and wrapping, through
    launch_cosmological_simulation's checks."""
    rng = np.random.default_rng(11)
    n, grid = 6000, 32
    c, bw = np.array([50.0, 50.0, 50.0]), 50.0
    x = c + rng.uniform(-bw, bw, size=(n, 3))
    x[: n // 2] = c + 5.0 * rng.standard_normal((n // 2, 3))
    v = 30.0 * rng.standard_normal((n, 3))
    m = np.full(n, 1.0)
    args = (x, v, m, c, bw, 0.7, 0.3, 0.7, 0.01, 0.1, 0.3, grid)
    xg, steps, t = L.leapfrog_cosmology(*args, launch=True)
    xo, so, to = P.leapfrog_cosmology(*args)
    assert steps == so and t == to
    assert np.all(np.abs(xg - c) <= bw)
    d = np.abs(xg - xo)
    d = np.minimum(d, 2 * bw - d)  # a particle on the seam may wrap either way
    assert d.max() < 1e-8 * bw, d.max()


def test_launch_cosmological_simulation_checks():
    lib = L.load()
    x, v, m = np.zeros((4, 3)), np.zeros((4, 3)), np.ones(4)
    se = lib.get_new_settings()
    se.verbose = 0
    st_ = _abi.SimulationStatus()

    def run(op, a0=0.1, a1=0.2, grid=16, **kw):
        s, ids = L.cosmological_system(x, v, m, [0.0, 0.0, 0.0], 1.0, kw.get("h", 0.7), 0.3, kw.get("ol", 0.7))
        r = lib.launch_cosmological_simulation(C.byref(s), C.byref(op), C.byref(st_), C.byref(se), 0.01, a0, a1,
                                               grid)
        code = r.return_code
        lib.free_traceback(C.byref(r))
        return code

    off = lib.get_new_output_param()
    csv = lib.get_new_output_param()
    csv.method = 2
    hdf = lib.get_new_output_param()
    hdf.method, hdf.output_dir, hdf.output_interval = 3, b"/tmp/pm_snaps/", 0.01
    assert run(csv) == _abi.GRAV_VALUE_ERROR
    assert run(hdf) == _abi.GRAV_MEMORY_ERROR
    assert run(off, ol=0.0) == _abi.GRAV_VALUE_ERROR
    assert run(off, a0=0.3, a1=0.2) == _abi.GRAV_VALUE_ERROR
    assert run(off, grid=0) == _abi.GRAV_VALUE_ERROR
    assert run(off, grid=2048) == _abi.GRAV_NOT_IMPLEMENTED_ERROR
    assert run(off) == 0


def test_bench_configuration_matches_oracle():
    """The benchmark's workload at full size: a perturbed 128^3 lattice (2^21
    particles, in random order) on a 256^3 grid, against the numpy
    restatement, to 1e-10 of the largest acceleration."""
    nside, grid, box = 128, 256, 100.0
    rng = np.random.default_rng(13)
    q = (np.arange(nside) + 0.5) * (box / nside)
    x = np.stack(np.meshgrid(q, q, q, indexing="ij"), axis=-1).reshape(-1, 3)
    x = np.mod(x + 0.3 * (box / nside) * rng.standard_normal(x.shape), box)
    x = x[rng.permutation(len(x))]
    m = np.full(len(x), 1.0 / len(x))
    c, bw = np.full(3, box / 2), box / 2
    args = (x, m, c, bw, 70.0, 0.3, m.sum() / box ** 3, grid, 0.5)
    ag = L.acceleration_pm(*args)
    print("gpu done", flush=True)
    ao = P.acceleration_pm(*args)
    err = np.abs(ag - ao).max() / np.abs(ao).max()
    assert err < 1e-10, err
