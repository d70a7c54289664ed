"""GPU Euler, Euler-Cromer, RK4 and embedded Runge-Kutta (RKF45, DOPRI,
DVERK, RKF78) through the C-ABI (integrator_launch_simulation).

With the bitwise FP64 force methods every update keeps the reference's
operation order and the embedded error norm is summed in index order, so
final states and step counts are bitwise equal to the reference's golden
runs (f9) and to the oracle."""
import numpy as np
import pytest

from grav_sim_amd import ic
from grav_sim_amd import lib as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("integ", [1, 2, 3])
def test_fixed_step_golden_bitwise(golden, integ):
    f = golden("f9_rk.npz")
    ids, x, v, m, steps = L.launch_simulation(f["x"], f["v"], f["m"], 1.0, integ, 1e-3, 0.2005, method=1,
                                              eps=1e-2)
    assert steps == int(f[f"i{integ}_steps"])
    assert np.array_equal(x, f[f"i{integ}_x"]) and np.array_equal(v, f[f"i{integ}_v"])


def test_rk4_massless_golden_bitwise(golden):
    f = golden("f9_rk.npz")
    ids, x, v, m, steps = L.launch_simulation(f["x"], f["v"], f["rk4_massless_m"], 1.0, 3, 1e-3, 0.2005,
                                              method=2, eps=1e-2)
    assert steps == int(f["rk4_massless_steps"])
    assert np.array_equal(x, f["rk4_massless_x"]) and np.array_equal(v, f["rk4_massless_v"])


@pytest.mark.parametrize("integ", [5, 6, 7, 8])
def test_embedded_golden_bitwise(golden, integ):
    f = golden("f9_rk.npz")
    ids, x, v, m, steps = L.launch_simulation(f["x"], f["v"], f["m"], 1.0, integ, -1.0, 1.0, method=1,
                                              eps=1e-2, tolerance=1e-9)
    assert steps == int(f[f"i{integ}_steps"])
    assert np.array_equal(x, f[f"i{integ}_x"]) and np.array_equal(v, f[f"i{integ}_v"])


def test_dopri_pythagorean_three_body_golden_bitwise(golden):
    f = golden("f9_rk.npz")
    ids, x, v, m, steps = L.launch_simulation(f["p3_x"], f["p3_v"], f["p3_m"], float(f["p3_G"]), 6, -1.0,
                                              20.0, method=1, tolerance=1e-12)
    assert steps == int(f["p3_steps"])
    assert np.array_equal(x, f["p3_x_final"]) and np.array_equal(v, f["p3_v_final"])


@pytest.mark.parametrize("integ,method", [(1, 2), (2, 1), (3, 1), (5, 2), (7, 1)])
def test_matches_oracle_bitwise(integ, method):
    x, v, m = ic.plummer(300, 17)
    if method == 2:
        m = m.copy()
        m[150:] = 0.0
    if integ <= 3:
        ids, xg, vg, mg, steps = L.launch_simulation(x, v, m, 1.0, integ, 2e-3, 0.05, method=method, eps=1e-3)
        if integ == 3:
            xo, vo, so = O.rk4(x, v, m, 1.0, 2e-3, 0.05, eps=1e-3, method=method)
        else:
            xo, vo, so = O.euler(x, v, m, 1.0, 2e-3, 0.05, eps=1e-3, method=method, cromer=integ == 2)
    else:
        ids, xg, vg, mg, steps = L.launch_simulation(x, v, m, 1.0, integ, -1.0, 0.1, method=method, eps=1e-3,
                                                     tolerance=1e-8)
        xo, vo, so = O.rk_embedded(x, v, m, 1.0, integ, 1e-8, 0.1, eps=1e-3, method=method)
    assert steps == so
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo)


def test_embedded_refuses_fp32_forces():
    x, v, m = ic.plummer(64, 3)
    with pytest.raises(Exception):
        L.launch_simulation(x, v, m, 1.0, 6, -1.0, 0.1, method=16, eps=1e-3, tolerance=1e-8)


def test_rk4_fp32_forces_track_fp64():
    """RK4 with the FP32 split direct sum at N=4096 (Plummer, eps=1e-2) stays
    within FP32-force rounding of the same run with the FP64 forces: 100 steps
    of dt=1e-3, positions within 1e-7 and energies within 1e-8 relative."""
    x, v, m = ic.plummer(4096, 19)
    ids, x32, v32, m32, s32 = L.launch_simulation(x, v, m, 1.0, 3, 1e-3, 0.1, method=16, eps=1e-2)
    ids, x64, v64, m64, s64 = L.launch_simulation(x, v, m, 1.0, 3, 1e-3, 0.1, method=1, eps=1e-2)
    assert s32 == s64 == 100
    dx = np.abs(x32 - x64).max()
    e32, e64 = O.energy(x32, v32, m, 1.0), O.energy(x64, v64, m, 1.0)
    print(f"RK4 N=4096 FP32 vs FP64 forces: max |dx| {dx:.2e}, energy rel. diff {abs(e32 - e64) / abs(e64):.2e}")
    assert dx < 1e-7 and abs(e32 - e64) / abs(e64) < 1e-8
