"""ctypes mirror of the grav_sim C-ABI structs (include/grav_sim_hip.h).

Layouts are byte-identical to the reference headers:
``System`` src/system.h:12-20, ``AccelerationParam`` src/acceleration.h:20-26,
``ErrorStatus`` src/error.h:31-36, ``IntegratorParam`` / ``OutputParam`` /
``SimulationStatus`` src/common.h:29-56, ``Settings`` src/settings.h:11-16,
``LinearOctree`` src/linear_octree.h:20-57.
"""
import ctypes as C

GRAV_UNKNOWN_ERROR = -1
GRAV_SUCCESS = 0
GRAV_FAILURE = 1
GRAV_VALUE_ERROR = 2
GRAV_POINTER_ERROR = 3
GRAV_MEMORY_ERROR = 4
GRAV_OS_ERROR = 5
GRAV_NOT_IMPLEMENTED_ERROR = 6

ACCELERATION_METHOD_PAIRWISE = 1
ACCELERATION_METHOD_MASSLESS = 2
ACCELERATION_METHOD_BARNES_HUT = 3
# additive GPU methods (include/grav_sim_hip_ext.h)
ACCELERATION_METHOD_PAIRWISE_FP32 = 16
ACCELERATION_METHOD_MASSLESS_FP32 = 17
ACCELERATION_METHOD_BARNES_HUT_LEARNED = 18
ACCELERATION_METHOD_BARNES_HUT_QUADRUPOLE = 19
ACCELERATION_METHOD_BARNES_HUT_GROUP = 20
ACCELERATION_METHOD_BARNES_HUT_GROUP_QUADRUPOLE = 21
ACCELERATION_METHOD_BARNES_HUT_GROUP_EXPANSION = 22

INTEGRATOR_EULER = 1
INTEGRATOR_EULER_CROMER = 2
INTEGRATOR_RK4 = 3
INTEGRATOR_LEAPFROG = 4
INTEGRATOR_RKF45 = 5
INTEGRATOR_DOPRI = 6
INTEGRATOR_DVERK = 7
INTEGRATOR_RKF78 = 8
INTEGRATOR_IAS15 = 9
INTEGRATOR_WHFAST = 10

OUTPUT_METHOD_DISABLED = 1
OUTPUT_METHOD_CSV = 2
OUTPUT_METHOD_HDF5 = 3

MORTON_MAX_LEVEL = 21

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)
PI64 = C.POINTER(C.c_int64)


class ErrorStatus(C.Structure):
    _fields_ = [("return_code", C.c_int), ("traceback", C.c_void_p), ("traceback_code_", C.c_int)]


class System(C.Structure):
    _fields_ = [
        ("num_particles", C.c_int),
        ("particle_ids", PI),
        ("x", PD),
        ("v", PD),
        ("m", PD),
        ("G", C.c_double),
    ]


class AccelerationParam(C.Structure):
    _fields_ = [
        ("method", C.c_int),
        ("opening_angle", C.c_double),
        ("softening_length", C.c_double),
        ("max_num_particles_per_leaf", C.c_int),
    ]


class IntegratorParam(C.Structure):
    _fields_ = [
        ("integrator", C.c_int),
        ("dt", C.c_double),
        ("tolerance", C.c_double),
        ("initial_dt", C.c_double),
        ("whfast_remove_invalid_particles", C.c_bool),
    ]


class OutputParam(C.Structure):
    _fields_ = [
        ("method", C.c_int),
        ("output_dir", C.c_char_p),
        ("output_initial", C.c_bool),
        ("output_interval", C.c_double),
        ("coordinate_output_dtype", C.c_int),
        ("velocity_output_dtype", C.c_int),
        ("mass_output_dtype", C.c_int),
        ("output_count_", C.c_int),
    ]


class SimulationStatus(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("num_steps", C.c_int64)]


class Settings(C.Structure):
    _fields_ = [("verbose", C.c_int), ("enable_progress_bar", C.c_bool), ("is_exit_ptr", C.POINTER(C.c_bool))]


class CosmologicalSystem(C.Structure):
    # src/system.h:22-38, 128 bytes
    _fields_ = [
        ("num_particles", C.c_int),
        ("particle_ids", PI),
        ("x", PD),
        ("v", PD),
        ("m", PD),
        ("h", C.c_double),
        ("omega_m", C.c_double),
        ("omega_lambda", C.c_double),
        ("omega_k", C.c_double),
        ("box_center", C.c_double * 3),
        ("box_width", C.c_double),
        ("unit_mass", C.c_double),
        ("unit_length", C.c_double),
        ("unit_time", C.c_double),
    ]


class LinearOctree(C.Structure):
    _fields_ = [
        ("box_width", C.c_double),
        ("num_internal_nodes", C.c_int),
        ("particle_morton_indices_deepest_level", PI64),
        ("sorted_indices", PI),
        ("tree_num_particles", PI),
        ("tree_num_internal_children", PI),
        ("tree_first_particle_sorted_idx", PI),
        ("tree_first_internal_children_idx", PI),
        ("tree_mass", PD),
        ("tree_center_of_mass_x", PD),
        ("tree_center_of_mass_y", PD),
        ("tree_center_of_mass_z", PD),
    ]


def dptr(arr):
    """Pointer to a C-contiguous float64 numpy array."""
    assert arr.dtype.name == "float64" and arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(PD)


def iptr(arr):
    """Pointer to a C-contiguous int32 numpy array."""
    assert arr.dtype.name == "int32" and arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(PI)


def make_system(x, v, m, G, ids=None):
    """Wrap numpy arrays (x, v: (N,3) float64; m: (N,) float64) as a System.

    The arrays are referenced, not copied (as src/python_interface.c:94-193
    does); keep them alive while the System is in use.
    """
    import numpy as np

    n = int(m.shape[0])
    if ids is None:
        ids = np.arange(n, dtype=np.int32)
    s = System()
    s.num_particles = n
    s.particle_ids = iptr(ids)
    s.x = dptr(x)
    s.v = dptr(v)
    s.m = dptr(m)
    s.G = float(G)
    s._keep = (x, v, m, ids)
    return s
