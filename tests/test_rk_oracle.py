"""Oracle pins of the fixed-step (Euler, Euler-Cromer, RK4) and embedded
Runge-Kutta (RKF45, DOPRI, DVERK, RKF78) integrators (CPU only).

(a) bitwise against the reference's final states in the golden fixture
f9_rk.npz; (b) bitwise against the reference library built from its sources
on other inputs, when that build is present; (c) what the mathematics fixes:
the convergence order of each fixed-step method on a Kepler orbit and the
embedded methods' accuracy against the closed-form orbit."""
import numpy as np
import pytest

from oracle import oracle as O

FIXED = {1: dict(cromer=False), 2: dict(cromer=True)}


def _fixed(integ, x, v, m, G, dt, tf, eps=0.0, method=1):
    if integ in FIXED:
        return O.euler(x, v, m, G, dt, tf, eps=eps, method=method, **FIXED[integ])
    return O.rk4(x, v, m, G, dt, tf, eps=eps, method=method)


@pytest.mark.parametrize("integ", [1, 2, 3])
def test_fixed_step_golden_bitwise(golden, integ):
    f = golden("f9_rk.npz")
    xo, vo, so = _fixed(integ, f["x"], f["v"], f["m"], 1.0, 1e-3, 0.2005, eps=1e-2)
    assert so == int(f[f"i{integ}_steps"])
    assert np.array_equal(xo, f[f"i{integ}_x"]) and np.array_equal(vo, f[f"i{integ}_v"])


def test_rk4_massless_golden_bitwise(golden):
    f = golden("f9_rk.npz")
    xo, vo, so = O.rk4(f["x"], f["v"], f["rk4_massless_m"], 1.0, 1e-3, 0.2005, eps=1e-2, method=2)
    assert so == int(f["rk4_massless_steps"])
    assert np.array_equal(xo, f["rk4_massless_x"]) and np.array_equal(vo, f["rk4_massless_v"])


@pytest.mark.parametrize("integ", [5, 6, 7, 8])
def test_embedded_golden_bitwise(golden, integ):
    f = golden("f9_rk.npz")
    xo, vo, so = O.rk_embedded(f["x"], f["v"], f["m"], 1.0, integ, 1e-9, 1.0, eps=1e-2)
    assert so == int(f[f"i{integ}_steps"])
    assert np.array_equal(xo, f[f"i{integ}_x"]) and np.array_equal(vo, f[f"i{integ}_v"])


def test_dopri_pythagorean_golden_bitwise(golden):
    f = golden("f9_rk.npz")
    xo, vo, so = O.rk_embedded(f["p3_x"], f["p3_v"], f["p3_m"], float(f["p3_G"]), 6, 1e-12, 20.0)
    assert so == int(f["p3_steps"])
    assert np.array_equal(xo, f["p3_x_final"]) and np.array_equal(vo, f["p3_v_final"])


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("integ", [1, 2, 3, 5, 6, 7, 8])
def test_against_reference_build(integ):
    rng = np.random.default_rng(integ)
    x = rng.standard_normal((23, 3))
    v = 0.3 * rng.standard_normal((23, 3))
    m = rng.uniform(0.5, 1.5, 23) / 23
    if integ <= 3:
        xo, vo, so = _fixed(integ, x, v, m, 1.0, 7e-4, 0.05, eps=1e-3)
        ids, xr, vr, mr, sr = O.ref_simulation(x, v, m, 1.0, integ, 7e-4, 0.05, eps=1e-3)
    else:
        xo, vo, so = O.rk_embedded(x, v, m, 1.0, integ, 1e-8, 0.1, eps=1e-3, initial_dt=2e-3)
        ids, xr, vr, mr, sr = O.ref_simulation(x, v, m, 1.0, integ, -1.0, 0.1, eps=1e-3, tolerance=1e-8,
                                               initial_dt=2e-3)
    assert so == sr and np.array_equal(xo, xr) and np.array_equal(vo, vr)


def _kepler(e=0.3):
    """A massless test particle on an ellipse of semi-major axis 1 around a
    unit mass at rest (massless form, so the central mass does not move);
    returns the initial state and the closed-form position at time t."""
    x = np.array([[0.0, 0.0, 0.0], [1.0 - e, 0.0, 0.0]])
    v = np.array([[0.0, 0.0, 0.0], [0.0, np.sqrt((1 + e) / (1 - e)), 0.0]])
    m = np.array([1.0, 0.0])

    def at(t):
        M = t  # mean motion 1 (a = 1, G M = 1), pericentre at t = 0
        E = M
        for _ in range(60):
            E = E - (E - e * np.sin(E) - M) / (1 - e * np.cos(E))
        return np.array([np.cos(E) - e, np.sqrt(1 - e * e) * np.sin(E), 0.0])

    return x, v, m, at


@pytest.mark.parametrize("integ,order", [(1, 1), (2, 1), (3, 4)])
def test_fixed_step_convergence_order(integ, order):
    x, v, m, at = _kepler()
    tf = 1.0
    errs = []
    for dt in (4e-3, 2e-3):
        xo, vo, so = _fixed(integ, x, v, m, 1.0, dt, tf, method=2)
        errs.append(np.linalg.norm(xo[1] - at(tf)))
    rate = np.log2(errs[0] / errs[1])
    assert abs(rate - order) < 0.25, (errs, rate)


@pytest.mark.parametrize("integ", [5, 6, 7, 8])
def test_embedded_meets_tolerance_on_kepler_orbit(integ):
    x, v, m, at = _kepler()
    xo, vo, so = O.rk_embedded(x, v, m, 1.0, integ, 1e-10, 2.0 * np.pi, method=2)
    assert np.linalg.norm(xo[1] - at(2.0 * np.pi)) < 1e-6
    assert so > 5
