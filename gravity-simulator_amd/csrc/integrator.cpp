// integrator.cpp -- launch_simulation / integrator_launch_simulation on the
// GPU (src/grav_sim.c:61-138, src/integrator.c:191-279).  Leapfrog runs
// device-resident with the reference's operation order
// (src/integrator.c:894-1121); WHFast runs the GPU WHFast
// (src/integrator_whfast.c:200-407); IAS15 the GPU IAS15 (ias15.hip).
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "gs_internal.h"

namespace gs {
ErrorStatus whfast_gpu(System *system, IntegratorParam *ip, AccelerationParam *ap,
                       OutputParam *op, SimulationStatus *status, Settings *settings, double tf);
ErrorStatus ias15_gpu(System *system, IntegratorParam *ip, AccelerationParam *ap,
                      OutputParam *op, SimulationStatus *status, Settings *settings, double tf);
ErrorStatus euler_gpu(System *system, IntegratorParam *ip, AccelerationParam *ap,
                      OutputParam *op, SimulationStatus *status, Settings *settings, double tf,
                      bool cromer);
ErrorStatus rk4_gpu(System *system, IntegratorParam *ip, AccelerationParam *ap,
                    OutputParam *op, SimulationStatus *status, Settings *settings, double tf);
ErrorStatus rk_embedded_gpu(System *system, IntegratorParam *ip, AccelerationParam *ap,
                            OutputParam *op, SimulationStatus *status, Settings *settings,
                            double tf);
ErrorStatus sim_leapfrog_snapshot(GsSim *s, SnapshotWriter *w, double t, double dt);
}

using namespace gs;

static ErrorStatus leapfrog_gpu(System *system, IntegratorParam *ip, AccelerationParam *ap,
                                OutputParam *op, SimulationStatus *status, Settings *settings,
                                const double tf)
{
    const int n = system->num_particles;
    double dt = ip->dt;
    Context *c;
    GS_TRY(context(&c));
    SnapshotWriter snaps;
    GS_TRY(snaps.begin(c, op, system));
    // initial snapshot before the first half kick (src/integrator.c:945-959)
    if (snaps.on() && op->output_initial)
        GS_TRY(snaps.write_host(system->x, system->v, status->t, status->dt));
    GsSim *sim;
    GS_TRY(gs_sim_create(&sim, n, system->G));
    ErrorStatus st = gs_sim_upload(sim, system->x, system->v, system->m);
    if (st.return_code == GRAV_SUCCESS) st = gs_sim_leapfrog_begin(sim, ap, dt);
    const int64 total = static_cast<int64>(ceil(tf / dt));
    status->t = 0.0;
    status->dt = dt;
    status->num_steps = 0;
    // steps 0 .. total-2 end at (i+1) dt <= (total-1) dt < tf, so only the last
    // one can be clamped (src/integrator.c:1001-1006): the others run in
    // chunks (a replayed HIP graph for the direct-sum methods), checking the
    // exit flag between chunks.  A chunk ends early at a step that is due for
    // a snapshot (t = steps dt >= next output time, src/integrator.c:1045).
    const int64 chunk = 1024;
    while (st.return_code == GRAV_SUCCESS && status->num_steps < total - 1) {
        int64 k = std::min(chunk, total - 1 - status->num_steps);
        if (snaps.on())
            for (int64 j = 1; j < k; j++)
                if (snaps.due(static_cast<double>(status->num_steps + j) * dt)) {
                    k = j;
                    break;
                }
        st = gs_sim_leapfrog_steps(sim, ap, dt, k);
        status->num_steps += k;
        status->t = status->num_steps * dt;
        if (st.return_code == GRAV_SUCCESS && snaps.due(status->t))
            st = sim_leapfrog_snapshot(sim, &snaps, status->t, dt);
        if (settings->is_exit_ptr && *settings->is_exit_ptr) break;
    }
    const bool exited = settings->is_exit_ptr && *settings->is_exit_ptr;
    if (st.return_code == GRAV_SUCCESS && !exited && status->num_steps < total) {
        if (status->t + dt > tf) dt = tf - status->t;
        status->dt = dt;
        st = gs_sim_leapfrog_steps(sim, ap, dt, 1);
        status->num_steps++;
        status->t = status->num_steps * dt;
        if (st.return_code == GRAV_SUCCESS && snaps.due(status->t))
            st = sim_leapfrog_snapshot(sim, &snaps, status->t, dt);
    }
    if (st.return_code == GRAV_SUCCESS) st = gs_sim_leapfrog_end(sim, dt);
    if (st.return_code == GRAV_SUCCESS) st = gs_sim_download(sim, system->x, system->v);
    const ErrorStatus fin = snaps.finish();
    gs_sim_destroy(sim);
    if (st.return_code == GRAV_SUCCESS) return fin;
    return st;
}

extern "C" {

ErrorStatus ias15(System *system, IntegratorParam *ip, AccelerationParam *ap, OutputParam *op,
                  SimulationStatus *status, Settings *settings, const double tf)
{
    if (!system || !ip || !ap || !op || !status || !settings)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    return WRAP_TRACEBACK(ias15_gpu(system, ip, ap, op, status, settings, tf));
}

ErrorStatus rk_embedded(System *system, IntegratorParam *ip, AccelerationParam *ap,
                        OutputParam *op, SimulationStatus *status, Settings *settings,
                        const double tf)
{
    if (!system || !ip || !ap || !op || !status || !settings)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    return WRAP_TRACEBACK(rk_embedded_gpu(system, ip, ap, op, status, settings, tf));
}

ErrorStatus whfast(System *system, IntegratorParam *ip, AccelerationParam *ap, OutputParam *op,
                   SimulationStatus *status, Settings *settings, const double tf)
{
    if (!system || !ip || !ap || !op || !status || !settings)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    return WRAP_TRACEBACK(whfast_gpu(system, ip, ap, op, status, settings, tf));
}

ErrorStatus integrator_launch_simulation(System *system, IntegratorParam *ip,
                                         AccelerationParam *ap, OutputParam *op,
                                         SimulationStatus *status, Settings *settings,
                                         const double tf)
{
    if (!system || !ip || !ap || !op || !status || !settings)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    switch (ip->integrator) {
        case INTEGRATOR_LEAPFROG:
            return WRAP_TRACEBACK(leapfrog_gpu(system, ip, ap, op, status, settings, tf));
        case INTEGRATOR_WHFAST:
            return WRAP_TRACEBACK(whfast(system, ip, ap, op, status, settings, tf));
        case INTEGRATOR_IAS15:
            return WRAP_TRACEBACK(ias15_gpu(system, ip, ap, op, status, settings, tf));
        case INTEGRATOR_EULER:
            return WRAP_TRACEBACK(euler_gpu(system, ip, ap, op, status, settings, tf, false));
        case INTEGRATOR_EULER_CROMER:
            return WRAP_TRACEBACK(euler_gpu(system, ip, ap, op, status, settings, tf, true));
        case INTEGRATOR_RK4:
            return WRAP_TRACEBACK(rk4_gpu(system, ip, ap, op, status, settings, tf));
        case INTEGRATOR_RKF45:
        case INTEGRATOR_DOPRI:
        case INTEGRATOR_DVERK:
        case INTEGRATOR_RKF78:
            return WRAP_TRACEBACK(rk_embedded_gpu(system, ip, ap, op, status, settings, tf));
        default:
            return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "Invalid integrator");
    }
}

ErrorStatus launch_simulation(System *system, IntegratorParam *ip, AccelerationParam *ap,
                              OutputParam *op, SimulationStatus *status, Settings *settings,
                              const double tf)
{
    GS_TRY(finalize_system(system));
    GS_TRY(finalize_acceleration_param(ap));
    GS_TRY(finalize_integration_param(ip));
    GS_TRY(finalize_output_param(op, settings));
    if (tf < 0.0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "tf must be non-negative. Got: %g", tf);
    if (settings->verbose >= GRAV_VERBOSITY_IGNORE_INFO) fputs("Launching simulation...\n", stdout);
    return WRAP_TRACEBACK(integrator_launch_simulation(system, ip, ap, op, status, settings, tf));
}

int launch_simulation_python(int32 *num_particles, int32 *particle_ids, double *x, double *v,
                             double *m, int32 **new_particle_ids_ptr, double **new_x_ptr,
                             double **new_v_ptr, double **new_m_ptr, const double G,
                             const int32 integrator, const double dt, const double tolerance,
                             const double initial_dt, const bool whfast_remove_invalid_particles,
                             const int32 acceleration_method, const double opening_angle,
                             const double softening_length,
                             const int32 max_num_particles_per_leaf, const int32 output_method,
                             char *output_dir, const bool output_initial,
                             const double output_interval, const int32 coordinate_output_dtype,
                             const int32 velocity_output_dtype, const int32 mass_output_dtype,
                             const int32 verbose, const bool enable_progress_bar,
                             bool *is_exit_ptr, const double tf)
{
    System system = get_new_system();
    system.num_particles = *num_particles;
    system.particle_ids = particle_ids;
    system.x = x;
    system.v = v;
    system.m = m;
    system.G = G;

    AccelerationParam ap = get_new_acceleration_param();
    ap.method = acceleration_method;
    ap.opening_angle = opening_angle;
    ap.softening_length = softening_length;
    ap.max_num_particles_per_leaf = max_num_particles_per_leaf;

    IntegratorParam ip = get_new_integrator_param();
    ip.integrator = integrator;
    ip.dt = dt;
    ip.tolerance = tolerance;
    ip.initial_dt = initial_dt;
    ip.whfast_remove_invalid_particles = whfast_remove_invalid_particles;

    OutputParam op = get_new_output_param();
    op.method = output_method;
    op.output_dir = output_dir;
    op.output_interval = output_interval;
    op.output_initial = output_initial;
    op.coordinate_output_dtype = coordinate_output_dtype;
    op.velocity_output_dtype = velocity_output_dtype;
    op.mass_output_dtype = mass_output_dtype;

    // (the reference leaves this uninitialised; an initial snapshot prints it)
    SimulationStatus status = {0.0, 0.0, 0};
    Settings settings = get_new_settings();
    settings.verbose = verbose;
    settings.enable_progress_bar = enable_progress_bar;
    settings.is_exit_ptr = is_exit_ptr;

    ErrorStatus st = WRAP_TRACEBACK(launch_simulation(&system, &ip, &ap, &op, &status, &settings, tf));
    if (st.return_code != GRAV_SUCCESS) {
        print_and_free_traceback(&st);
        return 1;
    }
    *num_particles = system.num_particles;
    *new_particle_ids_ptr = system.particle_ids;
    *new_x_ptr = system.x;
    *new_v_ptr = system.v;
    *new_m_ptr = system.m;
    return 0;
}

}  // extern "C"
