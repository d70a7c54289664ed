/*
 * grav_sim_hip.h -- C-ABI of the MI355X force-evaluation library (libgrav_sim.so).
 *
 * Every type and function in this header keeps the name, layout, argument
 * meaning and error behaviour of the grav_sim C library it replaces, so that
 * the existing C integrators and the Python ctypes layer bind to it unchanged.
 * Each declaration cites the reference declaration it replaces
 * (paths relative to the grav_sim source tree).
 *
 * Additive entry points (new method IDs, device-resident calls, the batched
 * WHFast ensemble, the learned Barnes-Hut correction, timing) live in
 * grav_sim_hip_ext.h.
 *
 * Usable from C99 and C++.  Plain pointers and sizes only.
 */
#ifndef GRAV_SIM_HIP_H
#define GRAV_SIM_HIP_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
#define GS_RESTRICT __restrict__
extern "C" {
#else
#define GS_RESTRICT restrict
#endif

/* ------------------------------------------------------------------ */
/* Basic types                      (src/common.h:24-27)               */
/* ------------------------------------------------------------------ */
typedef int32_t int32;
typedef int64_t int64;
typedef uint32_t uint32;
typedef uint64_t uint64;

/* ------------------------------------------------------------------ */
/* Error handling                   (src/error.h:14-36)                */
/* ------------------------------------------------------------------ */
#define GRAV_UNKNOWN_ERROR -1
#define GRAV_SUCCESS 0
#define GRAV_FAILURE 1
#define GRAV_VALUE_ERROR 2
#define GRAV_POINTER_ERROR 3
#define GRAV_MEMORY_ERROR 4
#define GRAV_OS_ERROR 5
#define GRAV_NOT_IMPLEMENTED_ERROR 6

#define GRAV_TRACEBACK_NOT_INITIALIZED -1
#define GRAV_TRACEBACK_SUCCESS 0
#define GRAV_TRACEBACK_MALLOC_FAILED 1
#define GRAV_TRACEBACK_TRUNCATED 2
#define GRAV_TRACEBACK_SNPRINTF_FAILED 3

/* 24 bytes: return_code@0, traceback@8, traceback_code_@16 (src/error.h:31-36) */
typedef struct ErrorStatus
{
    int return_code;
    char *traceback;
    int traceback_code_;
} ErrorStatus;

/* replaces src/error.h:155 */
ErrorStatus make_success_error_status(void);
/* replaces src/error.h:165-170 */
void raise_warning(const char *GS_RESTRICT warning_file, const int warning_line,
                   const char *GS_RESTRICT warning_func, const char *GS_RESTRICT warning_msg);
/* replaces src/error.h:187-193 */
ErrorStatus raise_warning_fmt(const char *GS_RESTRICT warning_file, const int warning_line,
                              const char *GS_RESTRICT warning_func,
                              const char *GS_RESTRICT format, ...);
/* replaces src/error.h:206-212 */
ErrorStatus raise_error(const char *GS_RESTRICT error_file, const int error_line,
                        const char *GS_RESTRICT error_func, const int error_code,
                        const char *GS_RESTRICT error_msg);
/* replaces src/error.h:226-233 */
ErrorStatus raise_error_fmt(const char *GS_RESTRICT error_file, const int error_line,
                            const char *GS_RESTRICT error_func, const int error_code,
                            const char *GS_RESTRICT format, ...);
/* replaces src/error.h:246-252 */
ErrorStatus traceback(ErrorStatus error_status, const char *GS_RESTRICT function_call_source_code,
                      const char *GS_RESTRICT error_file, const int error_line,
                      const char *GS_RESTRICT error_func);
/* replaces src/error.h:257 */
void free_traceback(ErrorStatus *GS_RESTRICT error_status);
/* replaces src/error.h:264 */
void print_and_free_traceback(ErrorStatus *GS_RESTRICT error_status);

#define WRAP_RAISE_WARNING(error_msg) raise_warning(__FILE__, __LINE__, __func__, error_msg)
#define WRAP_RAISE_WARNING_FMT(format, ...) \
    raise_warning_fmt(__FILE__, __LINE__, __func__, format, ##__VA_ARGS__)
#define WRAP_RAISE_ERROR(error_code, error_msg) \
    raise_error(__FILE__, __LINE__, __func__, error_code, error_msg)
#define WRAP_RAISE_ERROR_FMT(error_code, format, ...) \
    raise_error_fmt(__FILE__, __LINE__, __func__, error_code, format, ##__VA_ARGS__)
#define WRAP_TRACEBACK(function_call) \
    traceback(function_call, #function_call, __FILE__, __LINE__, __func__)

/* ------------------------------------------------------------------ */
/* Settings                         (src/settings.h:6-16)              */
/* ------------------------------------------------------------------ */
#define GRAV_VERBOSITY_IGNORE_ALL 0
#define GRAV_VERBOSITY_IGNORE_INFO 1
#define GRAV_VERBOSITY_NORMAL 2
#define GRAV_VERBOSITY_VERBOSE 3

typedef struct Settings
{
    int verbose;
    bool enable_progress_bar;
    bool *is_exit_ptr;
} Settings;

/* replaces src/settings.h:23 (shared static exit flag, src/settings.c:5-14) */
Settings get_new_settings(void);

/* ------------------------------------------------------------------ */
/* System                           (src/system.h:12-20)               */
/* ------------------------------------------------------------------ */
/* 48 bytes: num_particles@0, particle_ids@8, x@16, v@24, m@32, G@40.
 * x and v are AoS double[N][3]; m is double[N]. */
typedef struct System
{
    int num_particles;
    int *particle_ids;
    double *x;
    double *v;
    double *m;
    double G;
} System;

/* replaces src/system.h:45 */
System get_new_system(void);
/* replaces src/system.h:55-58 */
ErrorStatus get_initialized_system(System *GS_RESTRICT system, const int num_particles);
/* replaces src/system.h (finalize_system; definition src/system.c:70-111) */
ErrorStatus finalize_system(System *GS_RESTRICT system);
/* replaces src/system.h (free_system; definition src/system.c:113-119) */
void free_system(System *GS_RESTRICT system);
/* 128 bytes: num_particles@0, particle_ids@8, x@16, v@24, m@32, h@40,
 * omega_m@48, omega_lambda@56, omega_k@64, box_center@72, box_width@96,
 * unit_mass@104, unit_length@112, unit_time@120 (src/system.h:22-38) */
typedef struct CosmologicalSystem
{
    int num_particles;
    int *particle_ids;
    double *x;
    double *v;
    double *m;
    double h;
    double omega_m;
    double omega_lambda;
    double omega_k;
    double box_center[3];
    double box_width;
    double unit_mass;
    double unit_length;
    double unit_time;
} CosmologicalSystem;

/* replaces src/system.h:78 (src/system.c:121-141) */
CosmologicalSystem get_new_cosmological_system(void);
/* replaces src/system.h:88-91 (src/system.c:143-172) */
ErrorStatus get_initialized_cosmological_system(CosmologicalSystem *GS_RESTRICT system,
                                                const int num_particles);
/* replaces src/system.h (finalize_cosmological_system; src/system.c:174-242) */
ErrorStatus finalize_cosmological_system(CosmologicalSystem *GS_RESTRICT system);
/* replaces src/system.h (free_cosmological_system; src/system.c:244-250) */
void free_cosmological_system(CosmologicalSystem *GS_RESTRICT system);
/* replaces src/system.h (set_periodic_boundary_conditions; src/system.c:1337-1380) */
void set_periodic_boundary_conditions(CosmologicalSystem *GS_RESTRICT system);

/* replaces src/system.h:237-241 (src/system.c:579-1095): the named initial conditions circular_binary_orbit,
 * eccentric_binary_orbit, 3d_helix, sun_earth_moon, figure-8, pyth-3-body,
 * solar_system, solar_system_plus, bitwise equal to the reference's. */
ErrorStatus initialize_built_in_system(System *GS_RESTRICT system,
                                       const char *GS_RESTRICT system_name,
                                       const bool is_memory_initialized);

/* ------------------------------------------------------------------ */
/* Integrator / output parameters   (src/common.h:29-56)               */
/* ------------------------------------------------------------------ */
/* 40 bytes */
typedef struct IntegratorParam
{
    int integrator;
    double dt;
    double tolerance;
    double initial_dt;
    bool whfast_remove_invalid_particles;
} IntegratorParam;

typedef struct OutputParam
{
    int method;
    char *output_dir;
    bool output_initial;
    double output_interval;
    int coordinate_output_dtype;
    int velocity_output_dtype;
    int mass_output_dtype;

    int output_count_;
} OutputParam;

typedef struct SimulationStatus
{
    double t;
    double dt;
    int64 num_steps;
} SimulationStatus;

/* Output method and dtype IDs (src/output.h:13-19) */
#define OUTPUT_METHOD_DISABLED 1
#define OUTPUT_METHOD_CSV 2
#define OUTPUT_METHOD_HDF5 3
#define OUTPUT_DTYPE_FLOAT 1
#define OUTPUT_DTYPE_DOUBLE 2

/* replaces src/output.h:34 (defaults src/output.c:85-98) */
OutputParam get_new_output_param(void);
/* replaces src/output.h:44-47 (src/output.c:126-240): method check,
 * positive interval, directory path ending in '/' (snapshots_YYYYMMDD_HHMMSS/
 * when NULL), created here.  HDF5 is not built in: GRAV_MEMORY_ERROR, as in a
 * reference build without HDF5. */
ErrorStatus finalize_output_param(OutputParam *GS_RESTRICT output_param,
                                  const Settings *GS_RESTRICT settings);

/* ------------------------------------------------------------------ */
/* Acceleration                     (src/acceleration.h:16-26)          */
/* ------------------------------------------------------------------ */
#define ACCELERATION_METHOD_PAIRWISE 1
#define ACCELERATION_METHOD_MASSLESS 2
#define ACCELERATION_METHOD_BARNES_HUT 3

/* 32 bytes: method@0, opening_angle@8, softening_length@16,
 * max_num_particles_per_leaf@24 */
typedef struct AccelerationParam
{
    int method;
    double opening_angle;
    double softening_length;
    int max_num_particles_per_leaf;
} AccelerationParam;

/* replaces src/acceleration.h:33 (defaults src/acceleration.c:62-71) */
AccelerationParam get_new_acceleration_param(void);

/* replaces src/acceleration.h:41-43 (validation src/acceleration.c:73-129) */
ErrorStatus finalize_acceleration_param(AccelerationParam *GS_RESTRICT acceleration_param);

/* THE drop-in boundary.  replaces src/acceleration.h:53-57
 * (dispatcher src/acceleration.c:131-154).  `a` (N x 3 doubles, host memory)
 * is fully overwritten; `system` and `acceleration_param` are read only.
 * Methods 1-3 are the reference methods, computed on the GPU; method IDs >= 16
 * (grav_sim_hip_ext.h) select the additional GPU variants. */
ErrorStatus acceleration(double *GS_RESTRICT a, const System *GS_RESTRICT system,
                         const AccelerationParam *GS_RESTRICT acceleration_param);

/* replaces src/acceleration.h:66-70 (src/acceleration_barnes_hut.c:33-76) */
ErrorStatus acceleration_barnes_hut(double *GS_RESTRICT a, const System *GS_RESTRICT system,
                                    const AccelerationParam *GS_RESTRICT acceleration_param);

/* replaces src/acceleration.h:98-103 (src/acceleration.c:369-507).  Same
 * printout; the reference acceleration for the MAE is always computed. */
ErrorStatus benchmark_acceleration(const System *GS_RESTRICT system,
                                   const AccelerationParam *acceleration_params,
                                   const int num_acceleration_params,
                                   const int *GS_RESTRICT num_times_acceleration_param);

/* ------------------------------------------------------------------ */
/* Linear octree                    (src/linear_octree.h:14-101)        */
/* ------------------------------------------------------------------ */
#define MORTON_MAX_LEVEL 21

/* 96 bytes, SoA (src/linear_octree.h:20-57).  Host-side copy of the tree the
 * GPU builds.  Node numbering is level order (children of a node are
 * contiguous, as in the reference); the node SET equals the reference's. */
typedef struct LinearOctree
{
    double box_width;
    int num_internal_nodes;
    int64 *particle_morton_indices_deepest_level;
    int *sorted_indices;
    int *tree_num_particles;
    int *tree_num_internal_children;
    int *tree_first_particle_sorted_idx;
    int *tree_first_internal_children_idx;
    double *tree_mass;
    double *tree_center_of_mass_x;
    double *tree_center_of_mass_y;
    double *tree_center_of_mass_z;
} LinearOctree;

/* replaces src/linear_octree.h:64 (every pointer NULL) */
LinearOctree get_new_linear_octree(void);

/* replaces src/linear_octree.h:75-81 (src/linear_octree.c:825-962).  Builds
 * the tree on the GPU and fills `octree` with host copies (malloc'd; release
 * with free_linear_octree).  Unlike the reference, leaf nodes also carry
 * their mass and centre of mass. */
ErrorStatus construct_octree(LinearOctree *GS_RESTRICT octree, const System *GS_RESTRICT system,
                             const AccelerationParam *GS_RESTRICT acceleration_param,
                             const double *GS_RESTRICT box_center, const double box_width);

/* replaces src/linear_octree.h:88 */
void free_linear_octree(LinearOctree *GS_RESTRICT octree);

/* replaces src/linear_octree.h:97-101 */
bool linear_octree_check_if_included(const int64 morton_index_i, const int64 morton_index_j,
                                     const int level);

/* ------------------------------------------------------------------ */
/* Integrators                      (src/integrator.h:17-139)           */
/* ------------------------------------------------------------------ */
#define INTEGRATOR_EULER 1
#define INTEGRATOR_EULER_CROMER 2
#define INTEGRATOR_RK4 3
#define INTEGRATOR_LEAPFROG 4
#define INTEGRATOR_RKF45 5
#define INTEGRATOR_DOPRI 6
#define INTEGRATOR_DVERK 7
#define INTEGRATOR_RKF78 8
#define INTEGRATOR_IAS15 9
#define INTEGRATOR_WHFAST 10

/* replaces src/integrator.h:33 (defaults src/integrator.c:119-129) */
IntegratorParam get_new_integrator_param(void);
/* replaces src/integrator.h:45 (src/integrator.c:131-189) */
ErrorStatus finalize_integration_param(IntegratorParam *GS_RESTRICT integration_param);

/* replaces src/integrator.h:62-70 (src/integrator.c:191-279).  Every
 * integrator runs device-resident on the GPU with the reference's FP64
 * operation order: EULER, EULER_CROMER, RK4 (src/integrator.c:276-892),
 * LEAPFROG (:894-1121), the embedded RKF45 / DOPRI / DVERK / RKF78 below,
 * IAS15 and WHFAST (src/integrator_whfast.c:200-407 semantics).
 * OUTPUT_METHOD_CSV writes the reference's snapshots (times, files, format,
 * and the state they leave in the system) for every integrator; device ->
 * host copies and CSV formatting run behind the step loop on writer
 * threads. */
ErrorStatus integrator_launch_simulation(System *GS_RESTRICT system,
                                         IntegratorParam *GS_RESTRICT integrator_param,
                                         AccelerationParam *GS_RESTRICT acceleration_param,
                                         OutputParam *GS_RESTRICT output_param,
                                         SimulationStatus *GS_RESTRICT simulation_status,
                                         Settings *GS_RESTRICT settings, const double tf);

/* replaces src/integrator.h:85-93 (src/integrator_rk_embedded.c:522-987):
 * the embedded Runge-Kutta methods (integrator_param->integrator RKF45,
 * DOPRI, DVERK or RKF78) with the reference's tableaus, initial step, error
 * norm (summed in index order) and step-size control, every per-component
 * update an FP64 GPU kernel in the reference's operation order (bitwise equal
 * to the reference with the FP64 direct sums).  The FP32 force methods are
 * refused (GRAV_VALUE_ERROR).  CSV snapshots as integrator_launch_simulation. */
ErrorStatus rk_embedded(System *system, IntegratorParam *integrator_param,
                        AccelerationParam *acceleration_param, OutputParam *output_param,
                        SimulationStatus *simulation_status, Settings *settings, const double tf);

/* replaces src/integrator.h:108-116 (src/integrator_ias15.c:229-615):
 * IAS15 with the reference's predictor-corrector, step-size control and
 * compensated sums, every per-component update an FP64 GPU kernel in the
 * reference's operation order and the forces from acceleration_param on the
 * device (bitwise equal to the reference with the FP64 direct sums).  CSV
 * snapshots as integrator_launch_simulation. */
ErrorStatus ias15(System *system, IntegratorParam *integrator_param,
                  AccelerationParam *acceleration_param, OutputParam *output_param,
                  SimulationStatus *simulation_status, Settings *settings, const double tf);

/* replaces src/integrator.h:131-139 (src/integrator_whfast.c:200-407): the GPU
 * WHFast on one system (gs_whfast_ensemble batches many).  CSV snapshots as
 * integrator_launch_simulation, in the reference's distance order. */
ErrorStatus whfast(System *system, IntegratorParam *integrator_param,
                   AccelerationParam *acceleration_param, OutputParam *output_param,
                   SimulationStatus *simulation_status, Settings *settings, const double tf);

/* replaces src/output.h:61-68 (src/output.c:241-293):
 * one CSV snapshot of the host-side system, snapshot_%05d.csv numbered by
 * output_param->output_count_, which it increments. */
ErrorStatus output_snapshot(OutputParam *output_param, const System *system,
                            const IntegratorParam *integrator_param,
                            const AccelerationParam *acceleration_param,
                            const SimulationStatus *simulation_status, const Settings *settings);

/* ------------------------------------------------------------------ */
/* Particle mesh and cosmology                                         */
/* ------------------------------------------------------------------ */
/* replaces src/cosmology.h (compute_omega_k, compute_da; src/cosmology.c:5-26) */
double compute_omega_k(const double omega_m, const double omega_lambda);
double compute_da(const double a, const double H0, const double omega_m, const double omega_lambda);

/* replaces src/acceleration.h:73-85 (src/acceleration_PM.c:274-391): the PM
 * acceleration (+grad phi of the CIC density contrast's potential, which the
 * cosmological integrator subtracts) on the GPU: CIC deposit with FP64
 * atomics, hipFFT r2c / c2r, 4th-order gradient, CIC interpolation.  Host
 * arrays in and out.  pm_grid_size in 1..1024 (larger: GRAV_NOT_IMPLEMENTED_ERROR). */
ErrorStatus acceleration_PM(double *GS_RESTRICT a, const int num_particles,
                            const double *GS_RESTRICT x, const double *GS_RESTRICT m,
                            const double *GS_RESTRICT box_center, const double box_width,
                            const double H0, const double omega_m, const double mean_bkg_density,
                            const int pm_grid_size, const double scale_factor);

/* replaces src/integrator.h:155-164 (src/integrator_cosmology_leapfrog.c:27-255):
 * device-resident, the step arithmetic of the reference; snapshots are HDF5
 * only in the reference and are not built in, so output must be disabled. */
ErrorStatus leapfrog_cosmology(CosmologicalSystem *GS_RESTRICT system,
                               OutputParam *GS_RESTRICT output_param,
                               SimulationStatus *GS_RESTRICT simulation_status,
                               Settings *GS_RESTRICT settings, double dt, const double a_begin,
                               const double a_final, const int pm_grid_size);

/* replaces src/output.h:80-85 (src/output.c:295-349) */
ErrorStatus output_snapshot_cosmology(OutputParam *output_param, const CosmologicalSystem *system,
                                      const SimulationStatus *simulation_status,
                                      const Settings *settings);

/* replaces src/grav_sim.h:71-80 (src/grav_sim.c:140-260) */
ErrorStatus launch_cosmological_simulation(CosmologicalSystem *GS_RESTRICT system,
                                           OutputParam *GS_RESTRICT output_param,
                                           SimulationStatus *GS_RESTRICT simulation_status,
                                           Settings *GS_RESTRICT settings, double dt,
                                           const double a_begin, const double a_final,
                                           const int pm_grid_size);

/* replaces src/grav_sim.h:47-55 (src/grav_sim.c:61-138) */
ErrorStatus launch_simulation(System *GS_RESTRICT system,
                              IntegratorParam *GS_RESTRICT integrator_param,
                              AccelerationParam *GS_RESTRICT acceleration_param,
                              OutputParam *GS_RESTRICT output_param,
                              SimulationStatus *GS_RESTRICT simulation_status,
                              Settings *GS_RESTRICT settings, const double tf);

/* replaces src/utils.h compute_energy (src/utils.c:27-59); FP64 on the GPU */
double compute_energy(const System *GS_RESTRICT system);

/* replaces src/utils.h grav_get_current_time (src/utils.c:20-25) */
double grav_get_current_time(void);

/* replaces src/python_interface.h:86-117 (src/python_interface.c:94-193).
 * Returns 0 on success, 1 on failure (traceback printed to stderr). */
int launch_simulation_python(int32 *GS_RESTRICT num_particles, int32 *particle_ids, double *x,
                             double *v, double *m, int32 **new_particle_ids_ptr,
                             double **new_x_ptr, double **new_v_ptr, double **new_m_ptr,
                             const double G, const int32 integrator, const double dt,
                             const double tolerance, const double initial_dt,
                             const bool whfast_remove_invalid_particles,
                             const int32 acceleration_method, const double opening_angle,
                             const double softening_length,
                             const int32 max_num_particles_per_leaf, const int32 output_method,
                             char *output_dir, const bool output_initial,
                             const double output_interval, const int32 coordinate_output_dtype,
                             const int32 velocity_output_dtype, const int32 mass_output_dtype,
                             const int32 verbose, const bool enable_progress_bar,
                             bool *is_exit_ptr, const double tf);

/* replaces src/python_interface.h:128-134 (src/python_interface.c:195-243) */
void compute_energy_python(double *GS_RESTRICT energy, const double G,
                           const double *GS_RESTRICT sol_state, const int32 num_snapshots,
                           const int32 num_particles);

/* replaces src/python_interface.h:9-11 */
void free_memory_int32(int32 *GS_RESTRICT ptr);
void free_memory_double(double *GS_RESTRICT ptr);

/* replaces src/python_interface.h:27-35 (src/python_interface.c:18-53): the
 * built-in system in malloc'ed arrays the caller frees with free_memory_*. */
int32 load_built_in_system_python(const char *system_name, int *num_particles_ptr,
                                  int32 **particle_ids_ptr, double **x_ptr, double **v_ptr,
                                  double **m_ptr, double *G_ptr);

/* replaces src/utils.h keplerian_to_cartesian (src/utils.c:82-141) */
void keplerian_to_cartesian(double *GS_RESTRICT x, double *GS_RESTRICT v,
                            const double semi_major_axis, const double eccentricity,
                            const double inclination, const double argument_of_periapsis,
                            const double longitude_of_ascending_node, const double true_anomaly,
                            const double total_mass, const double G);

/* replaces src/python_interface.h:52-67 */
void keplerian_to_cartesian_python(double *GS_RESTRICT x, double *GS_RESTRICT y,
                                   double *GS_RESTRICT z, double *GS_RESTRICT v_x,
                                   double *GS_RESTRICT v_y, double *GS_RESTRICT v_z,
                                   const double semi_major_axis, const double eccentricity,
                                   const double inclination, const double argument_of_periapsis,
                                   const double longitude_of_ascending_node,
                                   const double true_anomaly, const double total_mass,
                                   const double G);

/* replaces src/python_interface.h:144-164 (src/python_interface.c:245-304) */
void compute_linear_momentum_python(double *GS_RESTRICT linear_momentum,
                                    const double *GS_RESTRICT sol_state, const int32 num_snapshots,
                                    const int32 num_particles);
void compute_angular_momentum_python(double *GS_RESTRICT angular_momentum,
                                     const double *GS_RESTRICT sol_state,
                                     const int32 num_snapshots, const int32 num_particles);

#ifdef __cplusplus
}
#endif

#endif /* GRAV_SIM_HIP_H */
