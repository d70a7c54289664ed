"""GPU IAS15 (device-resident predictor-corrector) through the C-ABI.

With the bitwise FP64 direct sums every per-component update keeps the
reference's operation order, so final states and step counts are bitwise
equal to the reference's golden runs (f8) and to the oracle."""
import numpy as np
import pytest

from grav_sim_amd import ic
from grav_sim_amd import lib as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu

IAS15 = 9


def test_ias15_pythagorean_three_body_bitwise(golden):
    f = golden("f8_ias15.npz")
    ids, x, v, m, steps = L.launch_simulation(f["p3_x"], f["p3_v"], f["p3_m"], float(f["p3_G"]), IAS15, -1.0,
                                              70.0, method=1, tolerance=1e-9)
    assert steps == int(f["p3_steps"])
    assert np.array_equal(x, f["p3_x_final"]) and np.array_equal(v, f["p3_v_final"])


def test_ias15_plummer_softened_bitwise(golden):
    f = golden("f8_ias15.npz")
    ids, x, v, m, steps = L.launch_simulation(f["pl_x"], f["pl_v"], f["pl_m"], 1.0, IAS15, -1.0, 2.0, method=1,
                                              eps=1e-2, tolerance=1e-9)
    assert steps == int(f["pl_steps"])
    assert np.array_equal(x, f["pl_x_final"]) and np.array_equal(v, f["pl_v_final"])


@pytest.mark.parametrize("n,method,initial_dt", [(300, 1, -1.0), (300, 2, -1.0), (200, 1, 1e-3)])
def test_ias15_matches_oracle_bitwise(n, method, initial_dt):
    x, v, m = ic.plummer(n, 13)
    if method == 2:
        m = m.copy()
        m[n // 2:] = 0.0  # massive first, then test particles
    ids, xg, vg, mg, steps = L.launch_simulation(x, v, m, 1.0, IAS15, -1.0, 0.25, method=method, eps=1e-3,
                                                 tolerance=1e-9, initial_dt=initial_dt)
    xo, vo, so = O.ias15(x, v, m, 1.0, 1e-9, 0.25, eps=1e-3, method=method, initial_dt=initial_dt)
    assert steps == so
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo)


def test_ias15_n1024_bitwise_and_energy():
    """N=1024 Plummer (eps=1e-3) to t=0.05: bitwise equal to the oracle; the
    (unsoftened) energy changes by less than 1e-5."""
    x, v, m = ic.plummer(1024, 14)
    ids, xg, vg, mg, steps = L.launch_simulation(x, v, m, 1.0, IAS15, -1.0, 0.05, method=1, eps=1e-3,
                                                 tolerance=1e-9)
    xo, vo, so = O.ias15(x, v, m, 1.0, 1e-9, 0.05, eps=1e-3)
    assert steps == so and np.array_equal(xg, xo) and np.array_equal(vg, vo)
    e0, e1 = O.energy(x, v, m, 1.0), O.energy(xg, vg, mg, 1.0)
    print(f"N=1024 IAS15 to t=0.05: {steps} steps, |dE/E| {abs(e1 - e0) / abs(e0):.2e}")
    assert abs(e1 - e0) / abs(e0) < 1e-5
