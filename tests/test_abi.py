"""C-ABI checks that need no GPU: the library loads, exports every declared
symbol, keeps the reference struct layouts, and its host-side validation
returns the reference's error codes."""
import ctypes as C
import os
import re
import subprocess

import pytest

from grav_sim_amd import _abi
from grav_sim_amd import lib as L

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _have_gpu():
    try:
        return os.path.exists("/dev/kfd")
    except Exception:
        return False


def test_struct_layouts_match_reference():
    # sizes/offsets of src/system.h:12-20, src/acceleration.h:20-26,
    # src/error.h:31-36, src/common.h:29-36, src/linear_octree.h:20-57 on LP64
    assert C.sizeof(_abi.System) == 48
    assert _abi.System.x.offset == 16 and _abi.System.m.offset == 32 and _abi.System.G.offset == 40
    assert C.sizeof(_abi.AccelerationParam) == 32
    assert _abi.AccelerationParam.max_num_particles_per_leaf.offset == 24
    assert C.sizeof(_abi.ErrorStatus) == 24
    assert C.sizeof(_abi.IntegratorParam) == 40
    assert C.sizeof(_abi.LinearOctree) == 96


def test_header_layout_compiles_with_static_asserts(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include "grav_sim_hip.h"\n#include "grav_sim_hip_ext.h"\n#include <stddef.h>\n'
        "_Static_assert(sizeof(System) == 48, \"System\");\n"
        "_Static_assert(offsetof(System, G) == 40, \"System.G\");\n"
        "_Static_assert(sizeof(AccelerationParam) == 32, \"AccelerationParam\");\n"
        "_Static_assert(sizeof(ErrorStatus) == 24, \"ErrorStatus\");\n"
        "_Static_assert(sizeof(IntegratorParam) == 40, \"IntegratorParam\");\n"
        "_Static_assert(sizeof(LinearOctree) == 96, \"LinearOctree\");\n"
        "int main(void) { return 0; }\n")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(REPO, "include"), str(src),
                        "-o", str(tmp_path / "layout")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_library_exports_every_declared_symbol():
    lib = L.load()
    declared = set()
    for h in ("grav_sim_hip.h", "grav_sim_hip_ext.h"):
        text = open(os.path.join(REPO, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for mm in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = mm.group(1)
            if name not in ("sizeof", "struct", "if", "while", "return", "GS_RESTRICT", "typedef"):
                declared.add(name)
    declared = {d for d in declared if not d.startswith("WRAP_")}
    missing = [d for d in sorted(declared) if not hasattr(lib, d)]
    assert not missing, missing
    assert set(L.EXPORTS) <= declared


def test_defaults_match_reference():
    lib = L.load()
    p = lib.get_new_acceleration_param()
    assert (p.method, p.opening_angle, p.softening_length, p.max_num_particles_per_leaf) == (1, 1.0, 0.0, -1)
    ip = lib.get_new_integrator_param()
    assert (ip.integrator, ip.dt, ip.tolerance, ip.initial_dt, ip.whfast_remove_invalid_particles) == \
        (-1, -1.0, -1.0, -1.0, True)
    op = lib.get_new_output_param()
    assert op.method == _abi.OUTPUT_METHOD_DISABLED and op.output_interval == -1.0
    s = lib.get_new_system()
    assert s.num_particles == 0 and abs(s.G - 0.000295912208284119496676630) < 1e-20


@pytest.mark.parametrize("method,theta,eps,leaf,code,new_leaf", [
    (1, 1.0, 0.0, -1, 0, -1),
    (3, 0.5, 0.0, -1, 0, 1),       # BH: leaf -1 -> 1 (src/acceleration.c:113-116)
    (3, 0.5, 0.0, 4, 0, 4),
    (3, -0.1, 0.0, 1, 2, 1),       # negative opening angle
    (1, 1.0, -1e-3, -1, 2, -1),    # negative softening
    (3, 0.5, 0.0, 0, 2, 0),        # leaf < 1
    (7, 0.5, 0.0, 1, 2, 1),        # unknown method
    (16, 1.0, 1e-3, -1, 0, -1),    # FP32 split (additive method id)
])
def test_finalize_acceleration_param(method, theta, eps, leaf, code, new_leaf):
    lib = L.load()
    p = L.param(method, theta, eps, leaf)
    st = lib.finalize_acceleration_param(C.byref(p))
    assert st.return_code == code
    if code:
        text = L.tb_text(st)
        assert "ValueError" in text
        lib.free_traceback(C.byref(st))
    assert p.max_num_particles_per_leaf == new_leaf


def test_finalize_integration_param():
    lib = L.load()
    ip = lib.get_new_integrator_param()
    ip.integrator = 4
    st = lib.finalize_integration_param(C.byref(ip))
    assert st.return_code == _abi.GRAV_VALUE_ERROR  # dt must be positive
    lib.free_traceback(C.byref(st))
    ip.dt = 0.1
    assert lib.finalize_integration_param(C.byref(ip)).return_code == 0
    ip.integrator = 42
    st = lib.finalize_integration_param(C.byref(ip))
    assert st.return_code == _abi.GRAV_VALUE_ERROR
    lib.free_traceback(C.byref(st))


def test_check_if_included():
    lib = L.load()
    a = (5 << 60) | 123
    b = (5 << 60) | 999
    assert lib.linear_octree_check_if_included(a, b, 1)
    assert not lib.linear_octree_check_if_included(a, b, 21)


def test_unknown_method_rejected_before_device_use():
    import numpy as np
    lib = L.load()
    x = np.zeros((4, 3))
    m = np.ones(4)
    s = _abi.make_system(x, x.copy(), m, 1.0)
    p = L.param(99)
    a = np.zeros((4, 3))
    st = lib.acceleration(_abi.dptr(a), C.byref(s), C.byref(p))
    assert st.return_code == _abi.GRAV_VALUE_ERROR
    assert "Unknown acceleration method" in L.tb_text(st)
    lib.free_traceback(C.byref(st))


@pytest.mark.skipif(_have_gpu(), reason="checks the no-device path")
def test_no_device_fails_loudly():
    """Without a HIP device the library reports GRAV_FAILURE: there is no CPU fallback."""
    import numpy as np
    r = subprocess.run(
        ["python", "-c",
         "import sys; sys.path.insert(0, %r)\n"
         "import numpy as np\nfrom grav_sim_amd import lib\n"
         "try:\n    lib.acceleration(np.zeros((3, 3)) + np.arange(9).reshape(3, 3), np.ones(3))\n"
         "except RuntimeError as e:\n    print('RAISED', 'No HIP device' in str(e))\n"
         % os.path.join(REPO, "gravity-simulator_amd")],
        capture_output=True, text=True, timeout=120)
    assert "RAISED True" in r.stdout, r.stdout + r.stderr
    del np


def test_python_api_symbols_and_host_helpers():
    """Every restype grav_sim/utils.py:38-58 sets resolves, and the host helpers
    agree with the reference formulae."""
    import numpy as np
    lib = L.load()
    for name in ("launch_simulation_python", "compute_energy_python", "compute_linear_momentum_python",
                 "compute_angular_momentum_python", "load_built_in_system_python", "free_memory_int32",
                 "free_memory_double", "keplerian_to_cartesian_python"):
        assert hasattr(lib, name), name
    out = [C.c_double() for _ in range(6)]
    lib.keplerian_to_cartesian_python.argtypes = [C.POINTER(C.c_double)] * 6 + [C.c_double] * 8
    lib.keplerian_to_cartesian_python(*[C.byref(o) for o in out], 1.5, 0.1, 0.2, 0.3, 0.4, 0.5, 1.0, 1.0)
    from grav_sim_amd import ic
    x, v = ic.keplerian_to_cartesian(np.array(1.5), np.array(0.1), np.array(0.2), np.array(0.3), np.array(0.4),
                                     np.array(0.5), 1.0, 1.0)
    assert np.allclose([o.value for o in out], np.concatenate([x, v]), rtol=1e-14, atol=1e-15)
    sol = np.random.default_rng(0).normal(size=(2, 5, 7))
    lm = np.zeros(2)
    lib.compute_linear_momentum_python.argtypes = [_abi.PD, _abi.PD, C.c_int, C.c_int]
    lib.compute_linear_momentum_python(_abi.dptr(lm), _abi.dptr(np.ascontiguousarray(sol)), 2, 5)
    assert np.allclose(lm, (sol[..., 0] * sol[..., 4:7].sum(-1)).sum(-1), rtol=1e-14)


def test_ias15_refuses_fp32_forces():
    """IAS15's corrector and step control need FP64 forces: the FP32 methods
    are refused before any device use (no spin with dt -> 1e-12 tf)."""
    import numpy as np

    lib = L.load()
    x = np.zeros((2, 3))
    x[1, 0] = 1.0
    v = np.zeros((2, 3))
    m = np.ones(2)
    s = _abi.make_system(x, v, m, 1.0)
    ip = lib.get_new_integrator_param()
    ip.integrator = 9
    ip.tolerance = 1e-9
    ap = L.param(16)
    op = lib.get_new_output_param()
    st_ = _abi.SimulationStatus()
    se = lib.get_new_settings()
    se.verbose = 0
    st = lib.ias15(C.byref(s), C.byref(ip), C.byref(ap), C.byref(op), C.byref(st_), C.byref(se), 1.0)
    assert st.return_code == _abi.GRAV_VALUE_ERROR
    assert "FP64" in L.tb_text(st)
    lib.free_traceback(C.byref(st))


def test_cosmological_system_layout_and_host_helpers():
    """CosmologicalSystem is the reference's 128-byte struct; the host-side
    helpers (defaults, validation, periodic wrap, expansion rate) follow
    src/system.c:121-250,1337-1380 and src/cosmology.c:5-26 and need no GPU."""
    import numpy as np
    from oracle import pm as P
    assert C.sizeof(_abi.CosmologicalSystem) == 128
    assert _abi.CosmologicalSystem.box_center.offset == 72 and _abi.CosmologicalSystem.unit_time.offset == 120
    lib = L.load()
    s0 = lib.get_new_cosmological_system()
    assert (s0.h, s0.omega_m, s0.omega_lambda, s0.box_width, s0.unit_mass) == (-1.0, -1.0, -1.0, -1.0, 1.0)
    rng = np.random.default_rng(7)
    c, w = [0.5, -0.25, 3.0], 0.75
    x = np.asarray(c) + rng.uniform(-5.0, 5.0, size=(3000, 3))
    v, m = np.zeros_like(x), np.ones(3000)
    x0 = x.copy()
    s, ids = L.cosmological_system(x, v, m, c, w, 0.7, 0.3, 0.7)
    lib.set_periodic_boundary_conditions(C.byref(s))  # wraps x in place
    assert not np.array_equal(x, x0) and np.array_equal(x, P.wrap(x0, c, w))
    for a in (0.02, 0.5, 1.0):
        assert lib.compute_da(a, 67.0, 0.3, 0.7) == P.compute_da(a, 67.0, 0.3, 0.7)
    assert lib.finalize_cosmological_system(C.byref(s)).return_code == 0
    for field, bad in (("h", 0.0), ("omega_m", -1.0), ("omega_lambda", 0.0), ("box_width", 0.0)):
        t, _ = L.cosmological_system(x, v, m, c, w, 0.7, 0.3, 0.7)
        setattr(t, field, bad)
        st = lib.finalize_cosmological_system(C.byref(t))
        assert st.return_code == _abi.GRAV_VALUE_ERROR, field
        lib.free_traceback(C.byref(st))


def test_group_expansion_ratio_setter_checks_its_range():
    """gs_set_group_expansion_ratio accepts 0 < phi < 1 only (host-side, no
    device use); the default is 0.1."""
    assert L.get_group_expansion_ratio() == 0.1
    for bad in (0.0, 1.0, -0.5, float("nan")):
        with pytest.raises(RuntimeError, match="Group expansion ratio"):
            L.set_group_expansion_ratio(bad)
    L.set_group_expansion_ratio(0.25)
    try:
        assert L.get_group_expansion_ratio() == 0.25
    finally:
        L.set_group_expansion_ratio(0.1)


def test_group_expansion_angle_factor_setter_checks_its_range():
    assert L.get_group_expansion_angle_factor() == 1.25
    for bad in (0.99, 4.5, float("nan")):
        with pytest.raises(RuntimeError, match="Group expansion angle factor"):
            L.set_group_expansion_angle_factor(bad)
    L.set_group_expansion_angle_factor(1.0)
    try:
        assert L.get_group_expansion_angle_factor() == 1.0
    finally:
        L.set_group_expansion_angle_factor(1.25)
