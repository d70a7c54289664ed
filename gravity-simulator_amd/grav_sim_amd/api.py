"""GravitySimulatorAPI on libgrav_sim.so (the MI355X library).

The user interface of the reference's Python package -- ``GravitySimulatorAPI``
(grav_sim/api.py:23-447), its parameter classes (grav_sim/parameters.py:1-425),
``System`` (grav_sim/system.py:11-275) and ``Simulator``
(grav_sim/simulator.py:16-143) -- with the same names, attribute names,
string encodings, defaults, validation messages and ctypes calls, so scripts
written against the reference run unchanged after
``from grav_sim_amd.api import GravitySimulatorAPI``.

Differences, all additive:

* ``AccelerationParam.method`` also takes the library's GPU methods
  "pairwise_fp32", "massless_fp32" (FP32 tile sums with Kahan compensation),
  "barnes_hut_learned", "barnes_hut_quadrupole" (IDs 16-19), "barnes_hut_group",
  "barnes_hut_group_quadrupole" and "barnes_hut_group_expansion" (IDs 20-22:
  one tree walk per group of at most 64 particles; 22 adds the far nodes
  through a second-order expansion about the group's centre).
* ``GravitySimulatorAPI.launch_ensemble`` runs many WHFast systems in one
  batched GPU call (``gs_whfast_ensemble``).
* Plotting helpers (grav_sim/plotting.py) are not provided.

"particle_mesh" is listed as in the reference, whose library refuses it
unless built with FFTW; this library refuses it as well
(GRAV_VALUE_ERROR from finalize_acceleration_param).
"""
import csv
import ctypes as C
import queue
import threading
import time
from pathlib import Path

import numpy as np

from . import _abi
from . import lib as _lib


def load_c_lib(c_lib_path=None):
    """The library with the restypes grav_sim/utils.py:38-58 sets."""
    L = _lib.load(str(c_lib_path) if c_lib_path else None)
    L.launch_simulation_python.restype = C.c_int
    L.compute_energy_python.restype = None
    L.compute_linear_momentum_python.restype = None
    L.compute_angular_momentum_python.restype = None
    L.load_built_in_system_python.restype = C.c_int
    L.free_memory_int32.restype = None
    L.free_memory_double.restype = None
    L.keplerian_to_cartesian_python.restype = None
    return L


def keplerian_to_cartesian(c_lib, semi_major_axis, eccentricity, inclination, argument_of_periapsis,
                           longitude_of_ascending_node, true_anomaly, total_mass, G):
    """grav_sim/utils.py:61-124: orbital elements -> (x, v) through the C helper."""
    out = [C.c_double() for _ in range(6)]
    c_lib.keplerian_to_cartesian_python(*[C.byref(o) for o in out], C.c_double(semi_major_axis),
                                        C.c_double(eccentricity), C.c_double(inclination),
                                        C.c_double(argument_of_periapsis),
                                        C.c_double(longitude_of_ascending_node), C.c_double(true_anomaly),
                                        C.c_double(total_mass), C.c_double(G))
    return np.array([o.value for o in out[:3]]), np.array([o.value for o in out[3:]])


def _choice(value, allowed, what):
    value = value.lower().strip()
    if value not in allowed:
        raise ValueError(f"Invalid {what}: {value}. Available {_PLURAL[what]}: {allowed}")
    return value


_PLURAL = {"acceleration method": "methods", "integrator": "integrators", "output method": "methods",
           "coordinate output data type": "types", "velocity output data type": "types",
           "mass output data type": "types", "verbosity level": "levels"}


# ---------------------------------------------------------------- parameters
class AccelerationParam:
    """grav_sim/parameters.py:4-121."""

    AVAILABLE_ACCELERATION_METHODS = ["pairwise", "massless", "barnes_hut", "particle_mesh", "pairwise_fp32",
                                      "massless_fp32", "barnes_hut_learned", "barnes_hut_quadrupole",
                                      "barnes_hut_group", "barnes_hut_group_quadrupole",
                                      "barnes_hut_group_expansion"]
    ACCELERATION_METHODS_TO_ENCODING = {"pairwise": 1, "massless": 2, "barnes_hut": 3, "particle_mesh": 4,
                                        "pairwise_fp32": 16, "massless_fp32": 17, "barnes_hut_learned": 18,
                                        "barnes_hut_quadrupole": 19, "barnes_hut_group": 20,
                                        "barnes_hut_group_quadrupole": 21, "barnes_hut_group_expansion": 22}
    ACCELERATION_ENCODING_TO_METHODS = {v: k for k, v in ACCELERATION_METHODS_TO_ENCODING.items()}

    def __init__(self):
        self.method = "pairwise"
        self.opening_angle = 1.0
        self.softening_length = 0.0
        self.pm_grid_size = 128
        self.max_num_particles_per_leaf = 1

    @property
    def method(self):
        return self.ACCELERATION_ENCODING_TO_METHODS[self._method]

    @method.setter
    def method(self, value):
        value = _choice(value, self.AVAILABLE_ACCELERATION_METHODS, "acceleration method")
        self._method = self.ACCELERATION_METHODS_TO_ENCODING[value]

    @property
    def opening_angle(self):
        return self._opening_angle

    @opening_angle.setter
    def opening_angle(self, value):
        value = float(value)
        if value <= 0:
            raise ValueError(f"Invalid opening angle: {value}. Must be positive.")
        self._opening_angle = value

    @property
    def softening_length(self):
        return self._softening_length

    @softening_length.setter
    def softening_length(self, value):
        value = float(value)
        if value < 0:
            raise ValueError(f"Invalid softening length: {value}. Must be non-negative.")
        self._softening_length = value

    @property
    def pm_grid_size(self):
        return self._pm_grid_size

    @pm_grid_size.setter
    def pm_grid_size(self, value):
        value = int(value)
        if value <= 0:
            raise ValueError(f"Invalid grid size: {value}. Must be positive.")
        self._pm_grid_size = value

    @property
    def max_num_particles_per_leaf(self):
        return self._max_num_particles_per_leaf

    @max_num_particles_per_leaf.setter
    def max_num_particles_per_leaf(self, value):
        value = int(value)
        if value <= 0:
            raise ValueError(f"Invalid maximum number of particles per leaf: {value}. Must be positive.")
        self._max_num_particles_per_leaf = value


class IntegratorParam:
    """grav_sim/parameters.py:124-245."""

    AVAILABLE_INTEGRATORS = ["euler", "euler_cromer", "rk4", "leapfrog", "rkf45", "dopri", "dverk", "rkf78",
                             "ias15", "whfast"]
    FIXED_STEP_SIZE_INTEGRATORS = ["euler", "euler_cromer", "rk4", "leapfrog", "whfast"]
    ADAPTIVE_STEP_SIZE_INTEGRATORS = ["rkf45", "dopri", "dverk", "rkf78", "ias15"]
    INTEGRATOR_TO_ENCODING = {name: i + 1 for i, name in enumerate(AVAILABLE_INTEGRATORS)}
    ENCODING_TO_INTEGRATORS = {v: k for k, v in INTEGRATOR_TO_ENCODING.items()}
    RECOMMENDED_IAS15_SETTINGS_BUILT_IN_SYSTEMS = {
        "circular_binary_orbit": [50.0, "days", 1e-9, 1],
        "eccentric_binary_orbit": [2.6, "years", 1e-9, 1],
        "3d_helix": [20.0, "days", 1e-9, 1],
        "sun_earth_moon": [1.0, "years", 1e-9, 1],
        "figure-8": [20.0, "days", 1e-9, 1],
        "pyth-3-body": [70.0, "days", 1e-9, 1],
        "solar_system": [200.0, "years", 1e-9, 1],
        "solar_system_plus": [250.0, "years", 1e-9, 1],
    }

    def __init__(self):
        self.integrator = "euler"
        self.dt = -1.0
        self.tolerance = -1.0
        self.initial_dt = -1.0
        self.whfast_remove_invalid_particles = True

    @property
    def integrator(self):
        return self.ENCODING_TO_INTEGRATORS[self._integrator]

    @integrator.setter
    def integrator(self, value):
        value = _choice(value, self.AVAILABLE_INTEGRATORS, "integrator")
        self._integrator = self.INTEGRATOR_TO_ENCODING[value]

    dt = property(lambda self: self._dt, lambda self, v: setattr(self, "_dt", float(v)))
    tolerance = property(lambda self: self._tolerance, lambda self, v: setattr(self, "_tolerance", float(v)))
    initial_dt = property(lambda self: self._initial_dt, lambda self, v: setattr(self, "_initial_dt", float(v)))
    whfast_remove_invalid_particles = property(
        lambda self: self._whfast_remove_invalid_particles,
        lambda self, v: setattr(self, "_whfast_remove_invalid_particles", bool(v)))


class OutputParam:
    """grav_sim/parameters.py:248-372."""

    AVAILABLE_OUTPUT_METHODS = ["disabled", "csv", "hdf5"]
    AVAILABLE_OUTPUT_DTYPE = ["float", "double"]
    OUTPUT_METHODS_TO_ENCODING = {"disabled": 1, "csv": 2, "hdf5": 3}
    OUTPUT_ENCODING_TO_METHODS = {v: k for k, v in OUTPUT_METHODS_TO_ENCODING.items()}
    OUTPUT_DTYPE_TO_ENCODING = {"float": 1, "double": 2}
    OUTPUT_ENCODING_TO_DTYPE = {v: k for k, v in OUTPUT_DTYPE_TO_ENCODING.items()}

    def __init__(self):
        self.method = "disabled"
        self.output_dir = "tmp/"
        self.output_initial = True
        self.output_interval = 1.0
        self.coordinate_output_dtype = "double"
        self.velocity_output_dtype = "double"
        self.mass_output_dtype = "double"

    @property
    def method(self):
        return self.OUTPUT_ENCODING_TO_METHODS[self._method]

    @method.setter
    def method(self, value):
        value = _choice(value, self.AVAILABLE_OUTPUT_METHODS, "output method")
        self._method = self.OUTPUT_METHODS_TO_ENCODING[value]

    @property
    def output_dir(self):
        return self._output_dir

    @output_dir.setter
    def output_dir(self, value):
        value = str(Path(value).resolve())
        if not value.endswith("/"):
            value += "/"
        self._output_dir = value

    output_initial = property(lambda self: self._output_initial,
                              lambda self, v: setattr(self, "_output_initial", bool(v)))
    output_interval = property(lambda self: self._output_interval,
                               lambda self, v: setattr(self, "_output_interval", float(v)))

    def _dtype(kind):
        attr = f"_{kind}_output_dtype"

        def get(self):
            return OutputParam.OUTPUT_ENCODING_TO_DTYPE[getattr(self, attr)]

        def set_(self, value):
            value = _choice(value, OutputParam.AVAILABLE_OUTPUT_DTYPE, f"{kind} output data type")
            setattr(self, attr, OutputParam.OUTPUT_DTYPE_TO_ENCODING[value])

        return property(get, set_)

    coordinate_output_dtype = _dtype("coordinate")
    velocity_output_dtype = _dtype("velocity")
    mass_output_dtype = _dtype("mass")
    del _dtype


class Settings:
    """grav_sim/parameters.py:375-425."""

    AVAILABLE_VERBOSITY_LEVELS = ["ignore_all", "ignore_info", "normal", "verbose"]
    VERBOSITY_TO_ENCODING = {name: i for i, name in enumerate(AVAILABLE_VERBOSITY_LEVELS)}
    ENCODING_TO_VERBOSITY = {v: k for k, v in VERBOSITY_TO_ENCODING.items()}

    def __init__(self):
        self.verbose = "normal"
        self.enable_progress_bar = True

    @property
    def verbose(self):
        return self.ENCODING_TO_VERBOSITY[self._verbose]

    @verbose.setter
    def verbose(self, value):
        value = _choice(value, self.AVAILABLE_VERBOSITY_LEVELS, "verbosity level")
        self._verbose = self.VERBOSITY_TO_ENCODING[value]

    enable_progress_bar = property(lambda self: self._enable_progress_bar,
                                   lambda self, v: setattr(self, "_enable_progress_bar", bool(v)))


# ---------------------------------------------------------------- system
class System:
    """grav_sim/system.py:11-275 (plotting omitted)."""

    BUILT_IN_SYSTEMS = ["circular_binary_orbit", "eccentric_binary_orbit", "3d_helix", "sun_earth_moon",
                        "figure-8", "pyth-3-body", "solar_system", "solar_system_plus"]
    G_DEFAULT = 0.000295912208284119496676630

    def __init__(self, c_lib):
        self.c_lib = c_lib
        self.num_particles = 0
        self.particle_ids = np.zeros((0,), dtype=np.int32)
        self.x = np.zeros((0, 3), dtype=np.float64)
        self.v = np.zeros((0, 3), dtype=np.float64)
        self.m = np.zeros((0,), dtype=np.float64)
        self.G = self.G_DEFAULT

    def add(self, x, v, m):
        """Append particle(s); ids continue from the largest id present."""
        self.x = np.vstack((self.x, np.array(x, dtype=np.float64)))
        self.v = np.vstack((self.v, np.array(v, dtype=np.float64)))
        self.m = np.hstack((self.m, m))
        k = len(m) if isinstance(m, (list, np.ndarray)) else 1
        first = int(self.particle_ids.max()) + 1 if self.num_particles > 0 else 0
        self.particle_ids = np.hstack((self.particle_ids, np.arange(first, first + k, dtype=np.int32)))
        self.num_particles += k

    def add_keplerian(self, semi_major_axis, eccentricity, inclination, argument_of_periapsis,
                      longitude_of_ascending_node, true_anomaly, m, primary_particle_id):
        """A body on the given orbit about the particle with id primary_particle_id (uses self.G)."""
        idx = np.where(self.particle_ids == primary_particle_id)[0]
        if len(idx) == 0:
            raise ValueError(f"Particle ID {primary_particle_id} not found in the system")
        p = idx[0]
        x, v = keplerian_to_cartesian(self.c_lib, semi_major_axis, eccentricity, inclination,
                                      argument_of_periapsis, longitude_of_ascending_node, true_anomaly,
                                      float(self.m[p] + m), self.G)
        self.add(x + self.x[p], v + self.v[p], m)

    def remove(self, particle_ids):
        if isinstance(particle_ids, int):
            ids = np.array([particle_ids])
        elif isinstance(particle_ids, list):
            ids = np.array(particle_ids)
        elif isinstance(particle_ids, np.ndarray):
            ids = particle_ids
        else:
            raise TypeError("indices must be an int, list or np.ndarray")
        keep = np.isin(self.particle_ids, ids, invert=True)
        self.particle_ids, self.x, self.v, self.m = (self.particle_ids[keep], self.x[keep], self.v[keep],
                                                     self.m[keep])
        self.num_particles = len(self.particle_ids)

    def save(self, file_path=None):
        if file_path is None:
            file_path = Path(__file__).parent / "customized_systems.csv"
        with open(file_path, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(["particle_id", "m", "x", "y", "z", "vx", "vy", "vz"])
            for i in range(self.num_particles):
                w.writerow([self.particle_ids[i], self.m[i], *self.x[i], *self.v[i]])
        print(f'System successfully saved to "{file_path}"')

    @staticmethod
    def load_system(c_lib, file_path):
        file_path = Path(file_path)
        if not file_path.is_file():
            raise FileNotFoundError(f"File not found: {file_path}")
        system = System(c_lib)
        ids = []
        with open(file_path, "r") as f:
            r = csv.reader(f)
            next(r)
            for row in r:
                system.add(x=[float(row[2]), float(row[3]), float(row[4])],
                           v=[float(row[5]), float(row[6]), float(row[7])], m=float(row[1]))
                ids.append(int(row[0]))
        system.particle_ids = np.array(ids, dtype=np.int32)
        return system

    def center_of_mass_correction(self):
        M = np.sum(self.m)
        self.x -= np.sum(self.m[:, np.newaxis] * self.x, axis=0) / M
        self.v -= np.sum(self.m[:, np.newaxis] * self.v, axis=0) / M

    @staticmethod
    def get_built_in_system(c_lib, system_name):
        """grav_sim/system.py:226-275: the C library's built-in system, copied and freed."""
        n = C.c_int()
        ids = C.POINTER(C.c_int32)()
        px, pv, pm = C.POINTER(C.c_double)(), C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
        G = C.c_double()
        rc = c_lib.load_built_in_system_python(system_name.encode("utf-8"), C.byref(n), C.byref(ids),
                                               C.byref(px), C.byref(pv), C.byref(pm), C.byref(G))
        if rc != 0:
            raise ValueError("Failed to load built-in system. Please check the C library traceback.")
        k = n.value
        system = System(c_lib)
        system.num_particles = k
        system.particle_ids = np.ctypeslib.as_array(ids, shape=(k,)).copy()
        system.x = np.ctypeslib.as_array(px, shape=(k, 3)).copy()
        system.v = np.ctypeslib.as_array(pv, shape=(k, 3)).copy()
        system.m = np.ctypeslib.as_array(pm, shape=(k,)).copy()
        system.G = G.value
        c_lib.free_memory_int32(ids)
        for p in (px, pv, pm):
            c_lib.free_memory_double(p)
        return system


# ---------------------------------------------------------------- simulator
class Simulator:
    """grav_sim/simulator.py:16-143."""

    DAYS_PER_YEAR = 365.242189

    def __init__(self, c_lib):
        self.c_lib = c_lib

    def launch_simulation(self, system, acceleration_param, integrator_param, output_param, settings, tf):
        """Run launch_simulation_python on a worker thread with the reference's
        positional arguments (grav_sim/simulator.py:53-102); Ctrl-C sets the exit
        flag the step loops poll and re-raises KeyboardInterrupt.  The system is
        updated in place with the final state (and the particles WHFast kept)."""
        x = np.ascontiguousarray(system.x, dtype=np.float64)
        v = np.ascontiguousarray(system.v, dtype=np.float64)
        m = np.ascontiguousarray(system.m, dtype=np.float64)
        ids = np.ascontiguousarray(system.particle_ids, dtype=np.int32)
        num_particles = C.c_int32(system.num_particles)
        new_ids = C.POINTER(C.c_int32)()
        new_x, new_v, new_m = C.POINTER(C.c_double)(), C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
        is_exit = C.c_bool(False)
        args = (
            C.byref(num_particles), ids.ctypes.data_as(C.POINTER(C.c_int32)), _abi.dptr(x), _abi.dptr(v),
            _abi.dptr(m), C.byref(new_ids), C.byref(new_x), C.byref(new_v), C.byref(new_m),
            C.c_double(system.G), C.c_int32(integrator_param._integrator), C.c_double(integrator_param.dt),
            C.c_double(integrator_param.tolerance), C.c_double(integrator_param.initial_dt),
            C.c_bool(integrator_param.whfast_remove_invalid_particles), C.c_int32(acceleration_param._method),
            C.c_double(acceleration_param.opening_angle), C.c_double(acceleration_param.softening_length),
            C.c_int32(acceleration_param.max_num_particles_per_leaf), C.c_int32(output_param._method),
            output_param.output_dir.encode("utf-8"), C.c_bool(output_param.output_initial),
            C.c_double(output_param.output_interval), C.c_int32(output_param._coordinate_output_dtype),
            C.c_int32(output_param._velocity_output_dtype), C.c_int32(output_param._mass_output_dtype),
            C.c_int32(settings._verbose), C.c_bool(settings.enable_progress_bar), C.byref(is_exit),
            C.c_double(tf),
        )
        result = queue.Queue()
        worker = threading.Thread(target=lambda: result.put(self.c_lib.launch_simulation_python(*args)))
        try:
            worker.start()
            while worker.is_alive():
                time.sleep(0.05)
        except KeyboardInterrupt:
            print("\nKeyboardInterrupt detected. Stopping simulation...\n")
            is_exit.value = True
            worker.join()
            raise
        worker.join()
        if result.get() != 0:
            raise RuntimeError("Simulation failed. Please check the C library traceback.")
        k = num_particles.value
        system.particle_ids = np.ctypeslib.as_array(new_ids, shape=(k,)).copy()
        system.x = np.ctypeslib.as_array(new_x, shape=(k, 3)).copy()
        system.v = np.ctypeslib.as_array(new_v, shape=(k, 3)).copy()
        system.m = np.ctypeslib.as_array(new_m, shape=(k,)).copy()
        system.num_particles = k


# ---------------------------------------------------------------- API
class GravitySimulatorAPI:
    """grav_sim/api.py:23-447 (plotting omitted) plus launch_ensemble."""

    def __init__(self, c_lib_path=None):
        self.c_lib = load_c_lib(c_lib_path)
        self.BUILT_IN_SYSTEMS = System.BUILT_IN_SYSTEMS
        self.simulator = Simulator(c_lib=self.c_lib)
        self.DAYS_PER_YEAR = self.simulator.DAYS_PER_YEAR
        self.launch_simulation = self.simulator.launch_simulation
        self.AVAILABLE_ACCELERATION_METHODS = AccelerationParam.AVAILABLE_ACCELERATION_METHODS
        self.AVAILABLE_INTEGRATORS = IntegratorParam.AVAILABLE_INTEGRATORS
        self.FIXED_STEP_SIZE_INTEGRATORS = IntegratorParam.FIXED_STEP_SIZE_INTEGRATORS
        self.ADAPTIVE_STEP_SIZE_INTEGRATORS = IntegratorParam.ADAPTIVE_STEP_SIZE_INTEGRATORS
        self.AVAILABLE_OUTPUT_METHODS = OutputParam.AVAILABLE_OUTPUT_METHODS
        self.AVAILABLE_OUTPUT_DTYPE = OutputParam.AVAILABLE_OUTPUT_DTYPE

    def get_new_system(self):
        return System(c_lib=self.c_lib)

    def load_system(self, file_path):
        return System.load_system(self.c_lib, file_path)

    def get_built_in_system(self, system_name):
        return System.get_built_in_system(self.c_lib, system_name)

    @staticmethod
    def get_new_parameters():
        return AccelerationParam(), IntegratorParam(), OutputParam(), Settings()

    def days_to_years(self, days):
        return days / self.simulator.DAYS_PER_YEAR

    def years_to_days(self, years):
        return years * self.simulator.DAYS_PER_YEAR

    @staticmethod
    def read_csv_data(output_dir):
        """grav_sim/api.py:124-228: (G, time, dt, particle_ids, sol_state) from the
        snapshot_*.csv files, rows ordered by particle id."""
        output_dir = Path(output_dir)
        if not output_dir.is_dir():
            raise FileNotFoundError(f"Output directory not found: {output_dir}")
        files = sorted(output_dir.glob("snapshot_*.csv"))
        if len(files) == 0:
            raise FileNotFoundError(f"No snapshot files found in: {output_dir}")
        G = -1.0
        t = np.zeros(len(files))
        dt = np.zeros(len(files))
        num_particles = None
        for i, fp in enumerate(files):
            meta = {}
            with open(fp, "r") as f:
                for line in f:
                    line = line.strip()
                    if not line.startswith("#"):
                        break
                    key, _, val = line[1:].partition(":")
                    meta[key.strip()] = val.strip()
            n = int(meta["num_particles"])
            if num_particles is None:
                num_particles = n
            elif n != num_particles:
                raise ValueError(f"Number of particles changed from {num_particles} to {n}")
            G, t[i], dt[i] = float(meta["G"]), float(meta["time"]), float(meta["dt"])
        sol_state = np.zeros((len(files), num_particles, 7))
        particle_ids = None
        for i, fp in enumerate(files):
            data = np.genfromtxt(fp, delimiter=",", skip_header=5).reshape(-1, 8)
            snap_ids = data[:, 0].astype(np.int32)
            if i == 0:
                particle_ids = np.sort(snap_ids)
                if len(np.unique(particle_ids)) != len(particle_ids):
                    raise ValueError(f"Particle IDs are not unique. Particle IDs: {particle_ids}")
            order = np.argsort(snap_ids)
            data = data[order]
            if not np.array_equal(particle_ids, snap_ids[order]):
                raise ValueError(f"Particle IDs do not match in snapshot {i + 1}: {fp}")
            sol_state[i] = data[:, 1:]
        return G, t, dt, particle_ids, sol_state

    @staticmethod
    def delete_snapshots(output_dir):
        output_dir = Path(output_dir)
        if not output_dir.is_dir():
            raise FileNotFoundError(f"Output directory not found: {output_dir}")
        for fp in sorted(output_dir.glob("snapshot_*.csv")):
            fp.unlink()

    def _per_snapshot(self, fn, sol_state, *lead):
        sol = np.ascontiguousarray(sol_state, dtype=np.float64)
        out = np.zeros(sol.shape[0])
        fn(_abi.dptr(out), *lead, _abi.dptr(sol), C.c_int32(sol.shape[0]), C.c_int32(sol.shape[1]))
        return out

    def compute_energy(self, sol_state, G):
        """Total energy per snapshot (src/python_interface.c:195-243)."""
        return self._per_snapshot(self.c_lib.compute_energy_python, sol_state, C.c_double(G))

    def compute_linear_momentum(self, sol_state):
        return self._per_snapshot(self.c_lib.compute_linear_momentum_python, sol_state)

    def compute_angular_momentum(self, sol_state):
        return self._per_snapshot(self.c_lib.compute_angular_momentum_python, sol_state)

    @staticmethod
    def compute_eccentricity(G, sol_state):
        """grav_sim/api.py:363-407: eccentricity of every body about body 0."""
        m0, m = sol_state[0, 0, 0], sol_state[0, 1:, 0]
        x = sol_state[:, 1:, 1:4] - sol_state[:, 0, 1:4].reshape(-1, 1, 3)
        v = sol_state[:, 1:, 4:7] - sol_state[:, 0, 4:7].reshape(-1, 1, 3)
        e = (np.cross(v, np.cross(x, v)) / (G * (m0 + m))[np.newaxis, :, np.newaxis]
             - x / np.linalg.norm(x, axis=2)[:, :, np.newaxis])
        return np.linalg.norm(e, axis=2)

    @staticmethod
    def compute_inclination(sol_state):
        """grav_sim/api.py:409-446: inclination of every orbit about body 0."""
        x = sol_state[:, 1:, 1:4] - sol_state[:, 0, 1:4].reshape(-1, 1, 3)
        v = sol_state[:, 1:, 4:7] - sol_state[:, 0, 4:7].reshape(-1, 1, 3)
        h = np.cross(x, v)
        h = h / np.linalg.norm(h, axis=2)[:, :, np.newaxis]
        return np.arccos(h[..., 2])

    def launch_ensemble(self, systems, acceleration_param, integrator_param, tf):
        """WHFast on many independent systems in one batched GPU call
        (gs_whfast_ensemble); each System is updated in place as
        ``launch_simulation`` would leave it.  All systems share G, dt and the
        acceleration method ("massless" or "pairwise" and their softening)."""
        if integrator_param.integrator != "whfast":
            raise ValueError("launch_ensemble runs the WHFast integrator")
        systems = list(systems)
        if not systems:
            return
        G = systems[0].G
        if any(s.G != G for s in systems):
            raise ValueError("every system of an ensemble must share G")
        counts = [s.num_particles for s in systems]
        offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        cat = lambda k: np.concatenate([np.asarray(getattr(s, k)) for s in systems])  # noqa: E731
        res = _lib.whfast_ensemble(offsets, cat("x").reshape(-1, 3), cat("v").reshape(-1, 3), cat("m"), G,
                                   integrator_param.dt, tf, method=acceleration_param._method,
                                   eps=acceleration_param.softening_length,
                                   remove_invalid=integrator_param.whfast_remove_invalid_particles,
                                   ids=cat("particle_ids").astype(np.int32))
        for s, (ids, x, v, m) in zip(systems, res):
            s.particle_ids, s.x, s.v, s.m, s.num_particles = ids, x, v, m, len(ids)
