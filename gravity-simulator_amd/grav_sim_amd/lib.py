"""ctypes binding of libgrav_sim.so (the MI355X force-evaluation library).

Mirrors how the reference's Python layer binds its C library
(grav_sim/utils.py:12-58: load the shared library, set restypes, call with
numpy buffers) and raises ``RuntimeError`` on a non-zero return code, as
grav_sim/simulator.py:117-123 does.  The library computes on the GPU only:
without a HIP device every compute call fails with GRAV_FAILURE.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import PD, PI, PI64, AccelerationParam, ErrorStatus, LinearOctree, System

PF = C.POINTER(C.c_float)


def _fptr(a):
    return a.ctypes.data_as(PF)

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libgrav_sim.so")

# every symbol the C-ABI headers (include/grav_sim_hip.h, include/grav_sim_hip_ext.h) declare
EXPORTS = [
    "make_success_error_status", "raise_warning", "raise_warning_fmt", "raise_error", "raise_error_fmt",
    "traceback", "free_traceback", "print_and_free_traceback", "get_new_settings", "get_new_system",
    "get_initialized_system", "finalize_system", "free_system", "get_new_output_param",
    "finalize_output_param", "get_new_acceleration_param", "finalize_acceleration_param", "acceleration",
    "acceleration_barnes_hut", "benchmark_acceleration", "get_new_linear_octree", "construct_octree",
    "free_linear_octree", "linear_octree_check_if_included", "get_new_integrator_param",
    "finalize_integration_param", "integrator_launch_simulation", "launch_simulation", "compute_energy",
    "grav_get_current_time", "launch_simulation_python", "compute_energy_python",
    "gs_device_info", "gs_benchmark_acceleration", "gs_synchronize", "gs_timing_enable", "gs_timing_reset", "gs_timing_query",
    "gs_sim_create", "gs_sim_destroy", "gs_sim_upload", "gs_sim_download", "gs_sim_acceleration",
    "gs_sim_leapfrog_begin", "gs_sim_leapfrog_steps", "gs_sim_leapfrog_end", "gs_sim_energy",
    "gs_sim_accel_error", "gs_octree_sort", "gs_last_walk_interactions", "gs_autotune_fp32",
    "gs_last_walk_expansion_pairs", "gs_set_group_expansion_ratio", "gs_get_group_expansion_ratio",
    "gs_set_group_expansion_angle_factor", "gs_get_group_expansion_angle_factor",
    "gs_set_fp32_config", "gs_get_fp32_config", "gs_acceleration_noalloc", "gs_copy_calibrate", "gs_issue_rates",
    "gs_whfast_ensemble",
    "gs_learned_num_params", "gs_learned_train", "gs_learned_get_params", "gs_learned_set_params",
    "gs_learned_reset", "gs_learned_forward", "gs_learned_gradients", "gs_learned_features",
    "gs_learned_set_policy", "gs_learned_status", "gs_learned_training_report", "gs_learned_training_times",
    "gs_learned_calibrate",
    "ias15", "whfast",
    "rk_embedded",
    "free_memory_int32", "free_memory_double", "load_built_in_system_python", "keplerian_to_cartesian",
    "keplerian_to_cartesian_python", "compute_linear_momentum_python", "compute_angular_momentum_python",
    "initialize_built_in_system", "output_snapshot",
    "compute_omega_k", "compute_da", "get_new_cosmological_system", "get_initialized_cosmological_system",
    "finalize_cosmological_system", "free_cosmological_system", "set_periodic_boundary_conditions",
    "acceleration_PM", "leapfrog_cosmology", "output_snapshot_cosmology", "launch_cosmological_simulation",
]


class GsBenchmarkRecord(C.Structure):
    _fields_ = [("method", C.c_int), ("num_times", C.c_int), ("mean_s", C.c_double), ("std_s", C.c_double),
                ("mae", C.c_double), ("interactions", C.c_double), ("interactions_per_s", C.c_double),
                ("particles_per_s", C.c_double), ("flops_per_s", C.c_double), ("peak_flops_per_s", C.c_double),
                ("roofline_frac", C.c_double)]


class GsDeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("cu_count", C.c_int),
                ("clock_khz", C.c_int), ("fp32_vector_tflops", C.c_double),
                ("fp64_vector_tflops", C.c_double), ("hbm_gbs", C.c_double)]


_lib = None


def load(path=None):
    """Load and configure libgrav_sim.so (build it with __graft_entry__.build())."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(p)
    ES = ErrorStatus
    pS, pP = C.POINTER(System), C.POINTER(AccelerationParam)
    sigs = {
        "acceleration": (ES, [PD, pS, pP]),
        "acceleration_barnes_hut": (ES, [PD, pS, pP]),
        "finalize_acceleration_param": (ES, [pP]),
        "get_new_acceleration_param": (AccelerationParam, []),
        "get_new_integrator_param": (_abi.IntegratorParam, []),
        "finalize_integration_param": (ES, [C.POINTER(_abi.IntegratorParam)]),
        "get_new_output_param": (_abi.OutputParam, []),
        "get_new_settings": (_abi.Settings, []),
        "get_new_system": (System, []),
        "finalize_system": (ES, [pS]),
        "construct_octree": (ES, [C.POINTER(LinearOctree), pS, pP, PD, C.c_double]),
        "free_linear_octree": (None, [C.POINTER(LinearOctree)]),
        "get_new_linear_octree": (LinearOctree, []),
        "linear_octree_check_if_included": (C.c_bool, [C.c_int64, C.c_int64, C.c_int]),
        "launch_simulation": (ES, [pS, C.POINTER(_abi.IntegratorParam), pP, C.POINTER(_abi.OutputParam),
                                   C.POINTER(_abi.SimulationStatus), C.POINTER(_abi.Settings), C.c_double]),
        "compute_energy": (C.c_double, [pS]),
        "free_traceback": (None, [C.POINTER(ES)]),
        "gs_device_info": (ES, [C.POINTER(GsDeviceInfo)]),
        "gs_benchmark_acceleration": (ES, [C.POINTER(System), C.POINTER(AccelerationParam), C.c_int, PI,
                                           C.POINTER(GsBenchmarkRecord)]),
        "gs_synchronize": (ES, []),
        "gs_timing_enable": (ES, [C.c_int]),
        "gs_timing_reset": (ES, []),
        "gs_timing_query": (ES, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
        "gs_sim_create": (ES, [C.POINTER(C.c_void_p), C.c_int, C.c_double]),
        "gs_sim_destroy": (ES, [C.c_void_p]),
        "gs_sim_upload": (ES, [C.c_void_p, PD, PD, PD]),
        "gs_sim_download": (ES, [C.c_void_p, PD, PD]),
        "gs_sim_acceleration": (ES, [C.c_void_p, pP, PD]),
        "gs_sim_leapfrog_begin": (ES, [C.c_void_p, pP, C.c_double]),
        "gs_sim_leapfrog_steps": (ES, [C.c_void_p, pP, C.c_double, C.c_int64]),
        "gs_sim_leapfrog_end": (ES, [C.c_void_p, C.c_double]),
        "gs_sim_energy": (ES, [C.c_void_p, PD]),
        "gs_sim_accel_error": (ES, [C.c_void_p, pP, C.c_int, PD, PD, PD]),
        "gs_octree_sort": (ES, [pS, PI64, PI, PD, PD]),
        "gs_last_walk_interactions": (ES, [C.POINTER(C.c_longlong)]),
        "gs_last_walk_expansion_pairs": (ES, [C.POINTER(C.c_longlong)]),
        "gs_set_group_expansion_ratio": (ES, [C.c_double]),
        "gs_get_group_expansion_ratio": (ES, [C.POINTER(C.c_double)]),
        "gs_set_group_expansion_angle_factor": (ES, [C.c_double]),
        "gs_get_group_expansion_angle_factor": (ES, [C.POINTER(C.c_double)]),
        "gs_autotune_fp32": (ES, [C.c_int, PI, PI, PI]),
        "gs_set_fp32_config": (ES, [C.c_int, C.c_int, C.c_int]),
        "gs_get_fp32_config": (ES, [PI, PI, PI]),
        "gs_acceleration_noalloc": (ES, [C.c_void_p, C.POINTER(System), C.POINTER(AccelerationParam)]),
        "gs_copy_calibrate": (ES, [C.c_size_t, C.c_int, PD]),
        "gs_issue_rates": (ES, [PD, PD, PD]),
        "gs_whfast_ensemble": (ES, [C.c_int, PI64, PI, PD, PD, PD, C.c_double, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_bool, PI]),
        "gs_learned_num_params": (C.c_int, []),
        "gs_learned_train": (ES, [pS, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                  C.c_uint, PD]),
        "gs_learned_get_params": (ES, [PF, PF, PF, PF]),
        "gs_learned_set_params": (ES, [PF, PF, PF, C.c_float]),
        "gs_learned_reset": (ES, []),
        "gs_learned_forward": (ES, [PF, C.c_int, PF]),
        "gs_learned_gradients": (ES, [PF, PF, C.c_int, PF, PD]),
        "gs_learned_features": (ES, [pS, C.c_double, C.c_double, C.c_int, PD, PF]),
        "gs_learned_set_policy": (ES, [C.c_double, C.c_int, C.c_double]),
        "gs_learned_status": (ES, [PD, PI]),
        "gs_learned_training_report": (ES, [PD, PD, PD, PD, PD]),
        "gs_learned_training_times": (ES, [PD, PD, PD, PD]),
        "gs_learned_calibrate": (ES, [pS, C.c_double, C.c_double, C.c_int, C.c_int, C.c_uint]),
        "ias15": (ES, [pS, C.POINTER(_abi.IntegratorParam), pP, C.POINTER(_abi.OutputParam),
                       C.POINTER(_abi.SimulationStatus), C.POINTER(_abi.Settings), C.c_double]),
        "whfast": (ES, [pS, C.POINTER(_abi.IntegratorParam), pP, C.POINTER(_abi.OutputParam),
                        C.POINTER(_abi.SimulationStatus), C.POINTER(_abi.Settings), C.c_double]),
        "rk_embedded": (ES, [pS, C.POINTER(_abi.IntegratorParam), pP, C.POINTER(_abi.OutputParam),
                             C.POINTER(_abi.SimulationStatus), C.POINTER(_abi.Settings), C.c_double]),
        "initialize_built_in_system": (ES, [pS, C.c_char_p, C.c_bool]),
        "output_snapshot": (ES, [C.POINTER(_abi.OutputParam), pS, C.POINTER(_abi.IntegratorParam), pP,
                                 C.POINTER(_abi.SimulationStatus), C.POINTER(_abi.Settings)]),
        "finalize_output_param": (ES, [C.POINTER(_abi.OutputParam), C.POINTER(_abi.Settings)]),
        "compute_omega_k": (C.c_double, [C.c_double, C.c_double]),
        "compute_da": (C.c_double, [C.c_double] * 4),
        "get_new_cosmological_system": (_abi.CosmologicalSystem, []),
        "finalize_cosmological_system": (ES, [C.POINTER(_abi.CosmologicalSystem)]),
        "set_periodic_boundary_conditions": (None, [C.POINTER(_abi.CosmologicalSystem)]),
        "acceleration_PM": (ES, [PD, C.c_int, PD, PD, PD, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.c_double]),
        "leapfrog_cosmology": (ES, [C.POINTER(_abi.CosmologicalSystem), C.POINTER(_abi.OutputParam),
                                    C.POINTER(_abi.SimulationStatus), C.POINTER(_abi.Settings), C.c_double,
                                    C.c_double, C.c_double, C.c_int]),
        "launch_cosmological_simulation": (ES, [C.POINTER(_abi.CosmologicalSystem), C.POINTER(_abi.OutputParam),
                                                C.POINTER(_abi.SimulationStatus), C.POINTER(_abi.Settings),
                                                C.c_double, C.c_double, C.c_double, C.c_int]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = L
    return L


def tb_text(st):
    if not st.traceback:
        return ""
    return C.string_at(st.traceback).decode(errors="replace")


def check(st, what="call"):
    """Raise RuntimeError (with the C traceback) for a failed ErrorStatus."""
    if st.return_code != _abi.GRAV_SUCCESS:
        text = tb_text(st)
        load().free_traceback(C.byref(st))
        raise RuntimeError(f"{what} failed (code {st.return_code}):\n{text}")


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def param(method, theta=1.0, eps=0.0, leaf=-1):
    return AccelerationParam(int(method), float(theta), float(eps), int(leaf))


def acceleration(x, m, method=1, G=1.0, eps=0.0, theta=1.0, leaf=-1):
    """The drop-in acceleration() on host arrays: returns a (N,3) float64."""
    L = load()
    x, m = _c(x), _c(m)
    v = np.zeros_like(x)
    s = _abi.make_system(x, v, m, G)
    p = param(method, theta, eps, leaf)
    a = np.zeros_like(x)
    check(L.acceleration(_abi.dptr(a), C.byref(s), C.byref(p)), "acceleration")
    return a


def construct_octree(x, m, leaf=1, box_center=None, box_width=-1.0):
    """GPU construct_octree; returns a dict of numpy copies of the LinearOctree."""
    L = load()
    x, m = _c(x), _c(m)
    v = np.zeros_like(x)
    s = _abi.make_system(x, v, m, 1.0)
    p = param(3, 0.5, 0.0, leaf)
    t = L.get_new_linear_octree()
    bc = _abi.dptr(_c(box_center)) if box_center is not None else None
    check(L.construct_octree(C.byref(t), C.byref(s), C.byref(p), bc, float(box_width)), "construct_octree")
    n, nn = len(m), t.num_internal_nodes

    def arr(ptr, k, dt):
        return np.ctypeslib.as_array(ptr, shape=(k,)).astype(dt).copy()

    out = dict(width=t.box_width, num_nodes=nn,
               sorted_keys=arr(t.particle_morton_indices_deepest_level, n, np.int64),
               sorted_idx=arr(t.sorted_indices, n, np.int32),
               count=arr(t.tree_num_particles, nn, np.int32),
               nchild=arr(t.tree_num_internal_children, nn, np.int32),
               first=arr(t.tree_first_particle_sorted_idx, nn, np.int32),
               child0=arr(t.tree_first_internal_children_idx, nn, np.int32),
               mass=arr(t.tree_mass, nn, np.float64),
               comx=arr(t.tree_center_of_mass_x, nn, np.float64),
               comy=arr(t.tree_center_of_mass_y, nn, np.float64),
               comz=arr(t.tree_center_of_mass_z, nn, np.float64))
    L.free_linear_octree(C.byref(t))
    return out


def octree_sort(x):
    L = load()
    x = _c(x)
    n = len(x)
    s = _abi.make_system(x, np.zeros_like(x), np.ones(n), 1.0)
    keys = np.zeros(n, dtype=np.int64)
    idx = np.zeros(n, dtype=np.int32)
    c = np.zeros(3)
    w = C.c_double()
    check(L.gs_octree_sort(C.byref(s), keys.ctypes.data_as(PI64), _abi.iptr(idx), _abi.dptr(c), C.byref(w)),
          "gs_octree_sort")
    return keys, idx, c, w.value


def last_walk_interactions():
    n = C.c_longlong()
    check(load().gs_last_walk_interactions(C.byref(n)))
    return n.value


def last_walk_expansion_pairs():
    """(group, node) pairs of the last walk with local expansions (method 22)."""
    n = C.c_longlong()
    check(load().gs_last_walk_expansion_pairs(C.byref(n)))
    return n.value


def set_group_expansion_ratio(phi):
    check(load().gs_set_group_expansion_ratio(float(phi)))


def get_group_expansion_ratio():
    v = C.c_double()
    check(load().gs_get_group_expansion_ratio(C.byref(v)))
    return v.value


def set_group_expansion_angle_factor(k):
    check(load().gs_set_group_expansion_angle_factor(float(k)))


def get_group_expansion_angle_factor():
    v = C.c_double()
    check(load().gs_get_group_expansion_angle_factor(C.byref(v)))
    return v.value


def device_info():
    info = GsDeviceInfo()
    check(load().gs_device_info(C.byref(info)), "gs_device_info")
    return info


def synchronize():
    check(load().gs_synchronize(), "gs_synchronize")


def timing_enable(on=True):
    check(load().gs_timing_enable(int(on)))


def timing_reset():
    check(load().gs_timing_reset())


def timing_query(kernel):
    ms, calls = C.c_double(), C.c_longlong()
    check(load().gs_timing_query(kernel.encode(), C.byref(ms), C.byref(calls)))
    return ms.value, calls.value


def autotune_fp32(n):
    b, i, s = C.c_int(), C.c_int(), C.c_int()
    check(load().gs_autotune_fp32(int(n), C.byref(b), C.byref(i), C.byref(s)), "gs_autotune_fp32")
    return b.value, i.value, s.value


def copy_calibrate(nbytes=1 << 28, reps=5):
    g = C.c_double()
    check(load().gs_copy_calibrate(int(nbytes), int(reps), C.byref(g)), "gs_copy_calibrate")
    return g.value


def issue_rates():
    """dict(fma, mul, rsq): cycles per wave64 instruction of FP32 chains."""
    v = [C.c_double(0.0) for _ in range(3)]
    check(load().gs_issue_rates(*[C.byref(t) for t in v]), "gs_issue_rates")
    return dict(zip(("fma", "mul", "rsq"), [t.value for t in v]))


def set_fp32_config(block, ipt, split):
    check(load().gs_set_fp32_config(int(block), int(ipt), int(split)), "gs_set_fp32_config")


def get_fp32_config():
    b, i, s = C.c_int(), C.c_int(), C.c_int()
    check(load().gs_get_fp32_config(C.byref(b), C.byref(i), C.byref(s)), "gs_get_fp32_config")
    return b.value, i.value, s.value


class Sim:
    """Device-resident simulation state (GsSim)."""

    def __init__(self, x, v, m, G=1.0):
        L = load()
        self.n = len(m)
        self.h = C.c_void_p()
        check(L.gs_sim_create(C.byref(self.h), self.n, float(G)), "gs_sim_create")
        self.upload(x, v, m)

    def upload(self, x=None, v=None, m=None):
        f = lambda a: _abi.dptr(_c(a)) if a is not None else None  # noqa: E731
        keep = [_c(a) for a in (x, v, m) if a is not None]  # noqa: F841
        check(load().gs_sim_upload(self.h, f(x), f(v), f(m)), "gs_sim_upload")

    def download(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        check(load().gs_sim_download(self.h, _abi.dptr(x), _abi.dptr(v)), "gs_sim_download")
        return x, v

    def acceleration(self, p, copy=True):
        a = np.zeros((self.n, 3)) if copy else None
        check(load().gs_sim_acceleration(self.h, C.byref(p), _abi.dptr(a) if copy else None),
              "gs_sim_acceleration")
        return a

    def leapfrog_begin(self, p, dt):
        check(load().gs_sim_leapfrog_begin(self.h, C.byref(p), float(dt)), "leapfrog_begin")

    def leapfrog_steps(self, p, dt, k):
        check(load().gs_sim_leapfrog_steps(self.h, C.byref(p), float(dt), int(k)), "leapfrog_steps")

    def leapfrog_end(self, dt):
        check(load().gs_sim_leapfrog_end(self.h, float(dt)), "leapfrog_end")

    def energy(self):
        e = C.c_double()
        check(load().gs_sim_energy(self.h, C.byref(e)), "gs_sim_energy")
        return e.value

    def accel_error(self, p, samples=0):
        md, p99, mx = C.c_double(), C.c_double(), C.c_double()
        check(load().gs_sim_accel_error(self.h, C.byref(p), int(samples), C.byref(md), C.byref(p99),
                                        C.byref(mx)), "gs_sim_accel_error")
        return md.value, p99.value, mx.value

    def close(self):
        if self.h:
            load().gs_sim_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def launch_simulation(x, v, m, G, integrator, dt, tf, method=1, eps=0.0, theta=1.0, leaf=-1,
                      remove_invalid=True, tolerance=-1.0, initial_dt=-1.0, output_dir=None,
                      output_interval=1.0, output_initial=True):
    """launch_simulation; returns (ids, x, v, m, steps).  With `output_dir`,
    CSV snapshots every `output_interval` are written there."""
    L = load()
    x, v, m = _c(x).copy(), _c(v).copy(), _c(m).copy()
    ids = np.arange(len(m), dtype=np.int32)
    s = _abi.make_system(x, v, m, G, ids)
    ip = L.get_new_integrator_param()
    ip.integrator = integrator
    ip.dt = dt
    ip.whfast_remove_invalid_particles = remove_invalid
    ip.tolerance = tolerance
    ip.initial_dt = initial_dt
    ap = param(method, theta, eps, leaf)
    op = L.get_new_output_param()
    if output_dir is not None:
        d = (str(output_dir).rstrip("/") + "/").encode()
        op.method = 2
        op.output_dir = d
        op.output_interval = output_interval
        op.output_initial = output_initial
    st = _abi.SimulationStatus()
    se = L.get_new_settings()
    se.verbose = 0
    se.enable_progress_bar = False
    check(L.launch_simulation(C.byref(s), C.byref(ip), C.byref(ap), C.byref(op), C.byref(st), C.byref(se),
                              float(tf)), "launch_simulation")
    k = s.num_particles
    return ids[:k].copy(), x[:k].copy(), v[:k].copy(), m[:k].copy(), st.num_steps


def whfast_ensemble(offsets, x, v, m, G, dt, tf, method=2, eps=0.0, remove_invalid=True, ids=None):
    """Batched GPU WHFast.  Returns a list of (ids, x, v, m) per system."""
    L = load()
    offsets = _c(offsets, np.int64)
    x, v, m = _c(x).copy(), _c(v).copy(), _c(m).copy()
    ns = len(offsets) - 1
    if ids is None:
        ids = np.concatenate([np.arange(offsets[s + 1] - offsets[s], dtype=np.int32) for s in range(ns)])
    ids = _c(ids, np.int32).copy()
    cnt = np.zeros(ns, dtype=np.int32)
    check(L.gs_whfast_ensemble(ns, offsets.ctypes.data_as(PI64), _abi.iptr(ids), _abi.dptr(x), _abi.dptr(v),
                               _abi.dptr(m), float(G), int(method), float(eps), float(dt), float(tf),
                               bool(remove_invalid), _abi.iptr(cnt)), "gs_whfast_ensemble")
    out = []
    for s in range(ns):
        o, k = int(offsets[s]), int(cnt[s])
        out.append((ids[o:o + k].copy(), x[o:o + k].copy(), v[o:o + k].copy(), m[o:o + k].copy()))
    return out


# ---------------------------------------------------------------- learned correction
def benchmark_acceleration(x, m, sets, G=1.0):
    """gs_benchmark_acceleration over sets of (method, theta, eps, leaf,
    num_times); returns the records as dicts (the text goes to stdout)."""
    L = load()
    x, m = _c(x), _c(m)
    v = np.zeros_like(x)
    s = _abi.make_system(x, v, m, G)
    k = len(sets)
    params = (AccelerationParam * k)(*[AccelerationParam(me, th, ep, lf) for me, th, ep, lf, _ in sets])
    times = np.array([t for *_, t in sets], dtype=np.int32)
    recs = (GsBenchmarkRecord * k)()
    check(L.gs_benchmark_acceleration(C.byref(s), params, k, times.ctypes.data_as(PI), recs),
          "gs_benchmark_acceleration")
    return [{f: getattr(r, f) for f, _ in GsBenchmarkRecord._fields_} for r in recs]


def learned_num_params():
    return int(load().gs_learned_num_params())


def learned_train(x, m, G=1.0, eps=0.0, theta=0.8, leaf=1, samples=1 << 20, epochs=30, batch=4096,
                  lr=1e-3, seed=1):
    """Train the residual network on (x, m); returns the last batch's loss."""
    L = load()
    x, m = _c(x), _c(m)
    v0 = np.zeros_like(x)
    s = _abi.make_system(x, v0, m, G)
    loss = C.c_double(0.0)
    check(L.gs_learned_train(C.byref(s), float(eps), float(theta), int(leaf), int(samples), int(epochs),
                             int(batch), float(lr), int(seed), C.byref(loss)), "gs_learned_train")
    return loss.value


def learned_calibrate(x, m, G=1.0, eps=0.0, theta=0.8, leaf=1, samples=1 << 16, seed=1):
    """Choose the network's output scale on another particle set (a
    validation realisation); the report then describes these samples."""
    L = load()
    x, m = _c(x), _c(m)
    s = _abi.make_system(x, np.zeros_like(x), m, G)
    check(L.gs_learned_calibrate(C.byref(s), float(eps), float(theta), int(leaf), int(samples), int(seed)),
          "gs_learned_calibrate")


def learned_get_params():
    L = load()
    w = np.zeros(learned_num_params(), dtype=np.float32)
    mu = np.zeros(16, dtype=np.float32)
    sd = np.zeros(16, dtype=np.float32)
    ts = np.zeros(1, dtype=np.float32)
    check(L.gs_learned_get_params(_fptr(w), _fptr(mu), _fptr(sd), _fptr(ts)), "gs_learned_get_params")
    return w, mu, sd, float(ts[0])


def learned_set_params(w, mean=None, std=None, tscale=1.0):
    L = load()
    w = _c(w, np.float32)
    mean = np.zeros(16, np.float32) if mean is None else _c(mean, np.float32)
    std = np.ones(16, np.float32) if std is None else _c(std, np.float32)
    check(L.gs_learned_set_params(_fptr(w), _fptr(mean), _fptr(std), float(tscale)), "gs_learned_set_params")


def learned_reset():
    check(load().gs_learned_reset(), "gs_learned_reset")


def learned_forward(X):
    L = load()
    X = _c(X, np.float32)
    Y = np.zeros((X.shape[0], 3), dtype=np.float32)
    check(L.gs_learned_forward(_fptr(X), X.shape[0], _fptr(Y)), "gs_learned_forward")
    return Y


def learned_gradients(X, T):
    L = load()
    X, T = _c(X, np.float32), _c(T, np.float32)
    g = np.zeros(learned_num_params(), dtype=np.float32)
    loss = C.c_double(0.0)
    check(L.gs_learned_gradients(_fptr(X), _fptr(T), X.shape[0], _fptr(g), C.byref(loss)),
          "gs_learned_gradients")
    return g, loss.value


def learned_features(x, m, G=1.0, eps=0.0, theta=0.8, leaf=1):
    """(coarse accelerations (N,3) float64, raw features (N,16) float32)."""
    L = load()
    x, m = _c(x), _c(m)
    a = np.zeros_like(x)
    s = _abi.make_system(x, a, m, G)  # v is not read
    F = np.zeros((len(m), 16), dtype=np.float32)
    check(L.gs_learned_features(C.byref(s), float(eps), float(theta), int(leaf), _abi.dptr(a), _fptr(F)),
          "gs_learned_features")
    return a, F


def learned_set_policy(tol=1e-3, samples=1024, fallback_theta=0.5):
    check(load().gs_learned_set_policy(float(tol), int(samples), float(fallback_theta)), "gs_learned_set_policy")


def learned_status():
    """(p99 of the result returned, fallback status): 0 no fallback, 1 the
    fallback met the tolerance, 2 it did not."""
    p = C.c_double(0.0)
    f = C.c_int(0)
    check(load().gs_learned_status(C.byref(p), C.byref(f)), "gs_learned_status")
    return p.value, f.value


def learned_training_times():
    """dict(samples_ms, prep_ms, loop_ms, select_ms) of the last training."""
    v = [C.c_double(0.0) for _ in range(4)]
    check(load().gs_learned_training_times(*[C.byref(t) for t in v]), "gs_learned_training_times")
    return dict(zip(("samples_ms", "prep_ms", "loop_ms", "select_ms"), [t.value for t in v]))


def learned_training_report():
    """dict(alpha, p99_net, p99_without, median_net, median_without) of the last training."""
    v = [C.c_double(0.0) for _ in range(5)]
    check(load().gs_learned_training_report(*[C.byref(t) for t in v]), "gs_learned_training_report")
    return dict(zip(("alpha", "p99_net", "p99_without", "median_net", "median_without"), [t.value for t in v]))


# ---------------------------------------------------------------- particle mesh / cosmology
def acceleration_pm(x, m, box_center, box_width, H0, omega_m, mean_bkg_density, grid, scale_factor):
    """acceleration_PM on host arrays (+grad phi, as the reference returns it)."""
    L = load()
    x, m = _c(x), _c(m)
    a = np.zeros_like(x)
    c = _c(box_center)
    check(L.acceleration_PM(_abi.dptr(a), len(m), _abi.dptr(x), _abi.dptr(m), _abi.dptr(c), float(box_width),
                            float(H0), float(omega_m), float(mean_bkg_density), int(grid), float(scale_factor)),
          "acceleration_PM")
    return a


def cosmological_system(x, v, m, box_center, box_width, h, omega_m, omega_lambda, ids=None):
    """A CosmologicalSystem over numpy arrays (kept alive by the caller)."""
    s = _abi.CosmologicalSystem()
    if ids is None:
        ids = np.arange(len(m), dtype=np.int32)
    s.num_particles = len(m)
    s.particle_ids = _abi.iptr(ids)
    s.x, s.v, s.m = _abi.dptr(x), _abi.dptr(v), _abi.dptr(m)
    s.h, s.omega_m, s.omega_lambda = h, omega_m, omega_lambda
    s.omega_k = 1.0 - omega_m - omega_lambda
    for d in range(3):
        s.box_center[d] = box_center[d]
    s.box_width = box_width
    s.unit_mass = s.unit_length = s.unit_time = 1.0
    return s, ids


def leapfrog_cosmology(x, v, m, box_center, box_width, h, omega_m, omega_lambda, dt, a_begin, a_final, grid,
                       launch=False):
    """The cosmological Leapfrog (output disabled); returns (x, num_steps, t).
    launch=True goes through launch_cosmological_simulation's checks."""
    L = load()
    x, v, m = _c(x).copy(), _c(v).copy(), _c(m).copy()
    s, ids = cosmological_system(x, v, m, box_center, box_width, h, omega_m, omega_lambda)
    op = L.get_new_output_param()
    st = _abi.SimulationStatus()
    se = L.get_new_settings()
    se.verbose = 0
    se.enable_progress_bar = False
    f = L.launch_cosmological_simulation if launch else L.leapfrog_cosmology
    check(f(C.byref(s), C.byref(op), C.byref(st), C.byref(se), float(dt), float(a_begin), float(a_final),
            int(grid)), "leapfrog_cosmology")
    return x, st.num_steps, st.t
