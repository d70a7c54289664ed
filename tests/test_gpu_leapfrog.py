"""GPU Leapfrog through launch_simulation (src/integrator.c:894-1121 semantics).

The device-resident integrator reproduces the reference trajectory bitwise:
N=1024 Plummer, eps=0, dt=1e-3, 1000 steps, against the golden final state
written by the reference build.
"""
import numpy as np
import pytest

from grav_sim_amd import _abi, ic
from grav_sim_amd import lib as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_leapfrog_config1_bitwise(golden):
    f = golden("f5_leapfrog.npz")
    ids, x, v, m, steps = L.launch_simulation(f["x"], f["v"], f["m"], 1.0, _abi.INTEGRATOR_LEAPFROG, 1e-3, 1.0,
                                              method=1, eps=0.0)
    assert steps == 1000
    assert np.array_equal(x, f["x_final"])
    assert np.array_equal(v, f["v_final"])
    e0, ef = float(f["e0"]), float(f["ef"])
    print(f"config1 |dE/E| = {abs((ef - e0) / e0):.3e}")


def test_leapfrog_clamped_last_step_and_massless_matches_oracle():
    off, X, V, M = ic.ensemble(1, 3, max_test=100)
    x, v, m = X[off[0]:off[1]], V[off[0]:off[1]], M[off[0]:off[1]]
    ids, xg, vg, mg, steps = L.launch_simulation(x, v, m, 1.0, _abi.INTEGRATOR_LEAPFROG, 0.03, 1.0,
                                                 method=2, eps=0.0)
    xo, vo, so = O.leapfrog(x, v, m, 1.0, 0.03, 1.0, method=2)
    assert steps == so == 34
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo)


def test_sim_handle_steps_match_launch_simulation():
    x, v, m = ic.plummer(300, 8)
    p = L.param(1, eps=1e-2)
    dt = 2.0 ** -10  # tf = 50 dt exactly: no clamped last step
    sim = L.Sim(x, v, m, 1.0)
    sim.leapfrog_begin(p, dt)
    sim.leapfrog_steps(p, dt, 50)
    sim.leapfrog_end(dt)
    xs, vs = sim.download()
    sim.close()
    xo, vo, so = O.leapfrog(x, v, m, 1.0, dt, 50 * dt, eps=1e-2)
    assert so == 50
    assert np.array_equal(xs, xo) and np.array_equal(vs, vo)


def test_fp32_leapfrog_energy_tracks_fp64():
    """FP32-split forces change the Leapfrog energy error only marginally
    relative to the FP64 path on the same trajectory (N=4096, 100 steps)."""
    x, v, m = ic.plummer(4096, 2)
    drifts = {}
    for method in (1, 16):
        p = L.param(method, eps=1e-3)
        sim = L.Sim(x, v, m, 1.0)
        e0 = sim.energy()
        sim.leapfrog_begin(p, 1e-3)
        sim.leapfrog_steps(p, 1e-3, 100)
        sim.leapfrog_end(1e-3)
        drifts[method] = (sim.energy() - e0) / abs(e0)
        sim.close()
    print(f"|dE/E| after 100 steps: fp64 {abs(drifts[1]):.3e}  fp32-split {abs(drifts[16]):.3e}")
    assert abs(drifts[16]) < 1e-5
    assert abs(drifts[16] - drifts[1]) < 1e-8


@pytest.mark.parametrize("method", [1, 16])
def test_graph_replayed_steps_equal_eager_steps(method):
    """Leapfrog steps replayed from a captured HIP graph (timing off) against
    the same steps launched eagerly (timing on): bitwise equal after every
    chunk, with eager work (energy, accelerations) between the replays."""
    x, v, m = ic.plummer(4096, 31)
    p = L.param(method, eps=1e-3)
    dt = 1e-3
    sg, se = L.Sim(x, v, m, 1.0), L.Sim(x, v, m, 1.0)
    sg.leapfrog_begin(p, dt)
    se.leapfrog_begin(p, dt)
    for chunk in (5, 17, 3, 40):
        L.timing_enable(False)
        sg.leapfrog_steps(p, dt, chunk)
        eg = sg.energy()
        L.timing_enable(True)
        se.leapfrog_steps(p, dt, chunk)
        ee = se.energy()
        L.timing_enable(False)
        xg, vg = sg.download()
        xe, ve = se.download()
        assert np.array_equal(xg, xe) and np.array_equal(vg, ve), chunk
        assert eg == ee
        sg.acceleration(p, copy=False)  # eager work on the stream between replays
    sg.close()
    se.close()
