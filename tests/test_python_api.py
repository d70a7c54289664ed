"""The Python-facing surface that needs no GPU: built-in systems, the
synchronous snapshot writer, output-parameter validation, and the
GravitySimulatorAPI mirror's parameter classes and CSV reader.

Pins: the reference's own outputs in tests/golden/f10_python_api.npz (built-in
systems from its load_built_in_system_python, snapshot files written by its
launch_simulation), compared bitwise / byte for byte.
"""
import ctypes as C
import os

import numpy as np
import pytest

from grav_sim_amd import _abi
from grav_sim_amd import api
from grav_sim_amd import lib as L

BUILT_IN = ["circular_binary_orbit", "eccentric_binary_orbit", "3d_helix", "sun_earth_moon", "figure-8",
            "pyth-3-body", "solar_system", "solar_system_plus"]


@pytest.fixture(scope="module")
def f10(golden):
    return golden("f10_python_api.npz")


@pytest.mark.parametrize("name", BUILT_IN)
def test_built_in_system_bitwise_equal_to_reference(f10, name):
    s = api.System.get_built_in_system(api.load_c_lib(), name)
    assert s.num_particles == len(f10[f"{name}_m"])
    assert np.array_equal(s.particle_ids, f10[f"{name}_ids"])
    for k in ("x", "v", "m"):
        assert np.array_equal(getattr(s, k), f10[f"{name}_{k}"]), k
    assert s.G == float(f10[f"{name}_G"])


def test_built_in_system_barycentric_and_unknown_name():
    lib = api.load_c_lib()
    s = api.System.get_built_in_system(lib, "solar_system")
    # centre of mass and total momentum vanish to rounding
    assert np.abs((s.m[:, None] * s.x).sum(0)).max() < 1e-15
    assert np.abs((s.m[:, None] * s.v).sum(0)).max() < 1e-17
    with pytest.raises(ValueError):
        api.System.get_built_in_system(lib, "no_such_system")


def test_initialize_built_in_system_into_existing_memory():
    """is_memory_initialized: the caller's arrays and G are used; too small -> GRAV_VALUE_ERROR."""
    lib = L.load()
    x, v, m = np.zeros((4, 3)), np.zeros((4, 3)), np.zeros(4)
    s = _abi.make_system(x, v, m, 2.0)
    st = lib.initialize_built_in_system(C.byref(s), b"pyth-3-body", True)
    assert st.return_code == 0
    assert np.array_equal(m[:3], np.array([3.0, 4.0, 5.0]) / 2.0) and m[3] == 0.0
    assert np.array_equal(x[0], [1.0, 3.0, 0.0])
    x2, v2, m2 = np.zeros((2, 3)), np.zeros((2, 3)), np.zeros(2)
    s2 = _abi.make_system(x2, v2, m2, 1.0)
    st = lib.initialize_built_in_system(C.byref(s2), b"solar_system", True)
    assert st.return_code == _abi.GRAV_VALUE_ERROR
    lib.free_traceback(C.byref(st))


def _output_param(lib, d, floats=False, method=2, interval=0.01):
    op = lib.get_new_output_param()
    op.method = method
    op.output_dir = d
    op.output_initial = True
    op.output_interval = interval
    if floats:
        op.coordinate_output_dtype = op.velocity_output_dtype = op.mass_output_dtype = 1
    return op


@pytest.mark.parametrize("case,floats", [("euler", False), ("rk4_float", True)])
def test_output_snapshot_matches_reference_file(f10, tmp_path, case, floats):
    """output_snapshot on the initial state writes the reference's initial
    snapshot byte for byte (metadata, header, %.17g / %.8g columns)."""
    lib = L.load()
    x, v, m = (np.ascontiguousarray(f10[k]) for k in ("csv_x", "csv_v", "csv_m"))
    s = _abi.make_system(x, v, m, 1.0)
    d = (str(tmp_path) + "/").encode()
    op = _output_param(lib, d, floats)
    ip = lib.get_new_integrator_param()
    ap = lib.get_new_acceleration_param()
    status = _abi.SimulationStatus()
    se = lib.get_new_settings()
    st = lib.output_snapshot(C.byref(op), C.byref(s), C.byref(ip), C.byref(ap), C.byref(status), C.byref(se))
    assert st.return_code == 0
    assert op.output_count_ == 1
    assert sorted(os.listdir(tmp_path)) == ["snapshot_00000.csv"]
    assert (tmp_path / "snapshot_00000.csv").read_text() == str(f10[f"csv_{case}_texts"][0])


def test_finalize_output_param_errors_and_directory(tmp_path):
    lib = L.load()
    se = lib.get_new_settings()
    se.verbose = 0
    d = str(tmp_path / "snaps") + "/"
    op = _output_param(lib, d.encode())
    assert lib.finalize_output_param(C.byref(op), C.byref(se)).return_code == 0
    assert os.path.isdir(d)
    # existing directory: accepted (the reference warns when verbose)
    assert lib.finalize_output_param(C.byref(op), C.byref(se)).return_code == 0
    for bad, code in ((_output_param(lib, d.encode(), interval=0.0), _abi.GRAV_VALUE_ERROR),
                      (_output_param(lib, d.rstrip("/").encode()), _abi.GRAV_VALUE_ERROR),
                      (_output_param(lib, d.encode(), method=3), _abi.GRAV_MEMORY_ERROR),
                      (_output_param(lib, d.encode(), method=7), _abi.GRAV_VALUE_ERROR)):
        st = lib.finalize_output_param(C.byref(bad), C.byref(se))
        assert st.return_code == code
        lib.free_traceback(C.byref(st))
    off = lib.get_new_output_param()
    assert lib.finalize_output_param(C.byref(off), C.byref(se)).return_code == 0


def test_parameter_classes_match_reference_encodings():
    a, i, o, s = api.GravitySimulatorAPI.get_new_parameters()
    assert (a._method, a.opening_angle, a.softening_length, a.max_num_particles_per_leaf) == (1, 1.0, 0.0, 1)
    assert i._integrator == 1 and i.dt == -1.0 and i.whfast_remove_invalid_particles is True
    assert o._method == 1 and o.output_interval == 1.0 and o.output_dir.endswith("/")
    assert s._verbose == 2 and s.enable_progress_bar is True
    for name, code in (("euler", 1), ("euler_cromer", 2), ("rk4", 3), ("leapfrog", 4), ("rkf45", 5), ("dopri", 6),
                       ("dverk", 7), ("rkf78", 8), ("ias15", 9), ("whfast", 10)):
        i.integrator = name.upper()
        assert i._integrator == code and i.integrator == name
    for name, code in (("pairwise", 1), ("massless", 2), ("barnes_hut", 3), ("pairwise_fp32", 16),
                       ("barnes_hut_quadrupole", 19), ("barnes_hut_group", 20),
                       ("barnes_hut_group_quadrupole", 21), ("barnes_hut_group_expansion", 22)):
        a.method = f"  {name} "
        assert a._method == code
    o.coordinate_output_dtype = "float"
    assert o._coordinate_output_dtype == 1
    with pytest.raises(ValueError, match="Invalid integrator: nope"):
        i.integrator = "nope"
    with pytest.raises(ValueError, match="Invalid opening angle"):
        a.opening_angle = 0.0
    with pytest.raises(ValueError, match="Invalid softening length"):
        a.softening_length = -1.0
    with pytest.raises(ValueError, match="Invalid output method"):
        o.method = "json"
    with pytest.raises(ValueError, match="Invalid verbosity level"):
        s.verbose = "loud"


def test_system_add_remove_keplerian_and_csv_round_trip(tmp_path, capsys):
    gs = api.GravitySimulatorAPI()
    s = gs.get_new_system()
    s.add([0.0, 0.0, 0.0], [0.0, 0.0, 0.0], 1.0)
    s.add_keplerian(1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1e-3, primary_particle_id=0)
    s.add([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]], [[0.1, 0.2, 0.3], [0.4, 0.5, 0.6]], [1e-9, 2e-9])
    assert s.num_particles == 4 and s.particle_ids.tolist() == [0, 1, 2, 3]
    # circular orbit of radius 1 about the primary with speed sqrt(G (m0 + m))
    assert np.allclose(s.x[1], [1.0, 0.0, 0.0]) and np.isclose(np.linalg.norm(s.v[1]), np.sqrt(s.G * 1.001))
    s.remove([2])
    assert s.particle_ids.tolist() == [0, 1, 3]
    s.center_of_mass_correction()
    assert np.allclose((s.m[:, None] * s.x).sum(0), 0.0, atol=1e-15)
    p = tmp_path / "sys.csv"
    s.save(p)
    t = gs.load_system(p)
    assert np.array_equal(t.particle_ids, s.particle_ids)
    assert np.array_equal(t.x, s.x) and np.array_equal(t.v, s.v) and np.array_equal(t.m, s.m)


def test_read_csv_data_and_momenta(f10, tmp_path):
    """The reference's snapshot files read back into sol_state; linear and
    angular momentum per snapshot through the C helpers agree with numpy
    (the energy helper computes on the GPU: tests/test_gpu_output.py)."""
    names, texts = f10["csv_ias15_names"], f10["csv_ias15_texts"]
    for n, t in zip(names, texts):
        (tmp_path / str(n)).write_text(str(t))
    gs = api.GravitySimulatorAPI()
    G, t, dt, ids, sol = gs.read_csv_data(tmp_path)
    assert G == 1.0 and t[0] == 0.0 and t[-1] == 0.5 and ids.tolist() == list(range(8))
    assert np.array_equal(sol[0, :, 1:4], f10["csv_x"]) and np.array_equal(sol[0, :, 0], f10["csv_m"])
    m, x, v = sol[..., 0], sol[..., 1:4], sol[..., 4:7]
    lm = gs.compute_linear_momentum(sol)
    assert np.allclose(lm, (m * v.sum(-1)).sum(-1), atol=1e-15)
    am = gs.compute_angular_momentum(sol)
    assert np.allclose(am, (m * np.cross(x, v).sum(-1)).sum(-1), atol=1e-15)
    gs.delete_snapshots(tmp_path)
    assert not list(tmp_path.glob("snapshot_*.csv"))


@pytest.mark.parametrize("floats", [False, True])
def test_snapshot_rows_print_like_printf_g(tmp_path, floats):
    """Every column of a snapshot row is the %.17g (%.8g) text of its value:
    random bit patterns (subnormals, signed zeros, extreme exponents) checked
    against Python's own correctly rounded %-formatting."""
    rng = np.random.default_rng(5)
    n = 4000
    bits = rng.integers(0, 2 ** 63 - 1, size=7 * n, dtype=np.int64) * rng.choice([-1, 1], 7 * n)
    vals = bits.view(np.float64)
    vals = np.where(np.isfinite(vals), vals, 0.5)
    vals[:6] = [0.0, -0.0, 5e-324, -2.2250738585072014e-308, 1.7976931348623157e308, 0.1]
    m, x, v = np.ascontiguousarray(vals[:n]), vals[n:4 * n].reshape(n, 3).copy(), vals[4 * n:].reshape(n, 3).copy()
    ids = rng.integers(-2 ** 31, 2 ** 31 - 1, size=n, dtype=np.int32)
    lib = L.load()
    s = _abi.make_system(x, v, m, 1.0, ids)
    op = _output_param(lib, (str(tmp_path) + "/").encode(), floats)
    status = _abi.SimulationStatus(0.25, 1e-3, 0)
    st = lib.output_snapshot(C.byref(op), C.byref(s), C.byref(lib.get_new_integrator_param()),
                             C.byref(lib.get_new_acceleration_param()), C.byref(status),
                             C.byref(lib.get_new_settings()))
    assert st.return_code == 0
    rows = (tmp_path / "snapshot_00000.csv").read_text().splitlines()
    assert rows[:5] == [f"# num_particles: {n}", "# G: 1", "# time: 0.25", "# dt: 0.001",
                        "particle_id,m,x,y,z,vx,vy,vz"]
    fmt = "%.8g" if floats else "%.17g"
    for i, r in enumerate(rows[5:]):
        want = [str(ids[i])] + [fmt % val for val in (m[i], *x[i], *v[i])]
        assert r.split(",") == want, i
