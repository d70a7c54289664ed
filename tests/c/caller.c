/* A C99 caller of the force path, written the way a user of the reference
 * library writes one (examples/benchmark/c/benchmark.c:12-67): a System from
 * get_initialized_system, acceleration() with the pairwise and Barnes-Hut
 * methods (src/acceleration.h:53-57), then launch_simulation() with Leapfrog
 * (src/grav_sim.h:47-55).  Built either against this library's header or,
 * with -DUSE_REFERENCE_HEADERS, against the reference's own headers; either
 * way it links libgrav_sim.so.
 *
 *   caller IN OUT
 *     IN:  int32 N, then x (3N doubles), v (3N), m (N)
 *     OUT: a_pairwise (3N), a_barnes_hut (3N), x and v after 10 Leapfrog
 *          steps of dt = 1e-3 with the pairwise sum (3N each)
 * Exit status 0 on success, 1 with the library's traceback on stderr. */
#include <stdio.h>
#include <stdlib.h>

#ifdef USE_REFERENCE_HEADERS
#include "grav_sim.h"
#else
#include "grav_sim_hip.h"
#endif

static int fail(ErrorStatus *st)
{
    print_and_free_traceback(st);
    return 1;
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 2;
    }
    FILE *in = fopen(argv[1], "rb");
    if (!in) return 2;
    int n = 0;
    if (fread(&n, sizeof(int), 1, in) != 1 || n <= 0) return 2;
    System system;
    ErrorStatus st = get_initialized_system(&system, n);
    if (st.return_code != GRAV_SUCCESS) return fail(&st);
    size_t ok = fread(system.x, sizeof(double), 3 * (size_t)n, in);
    ok += fread(system.v, sizeof(double), 3 * (size_t)n, in);
    ok += fread(system.m, sizeof(double), (size_t)n, in);
    fclose(in);
    if (ok != 7 * (size_t)n) return 2;
    system.G = 1.0;

    double *a = malloc(3 * (size_t)n * sizeof(double));
    double *b = malloc(3 * (size_t)n * sizeof(double));
    if (!a || !b) return 2;

    AccelerationParam pairwise = get_new_acceleration_param();
    pairwise.method = ACCELERATION_METHOD_PAIRWISE;
    pairwise.softening_length = 1e-3;
    st = acceleration(a, &system, &pairwise);
    if (st.return_code != GRAV_SUCCESS) return fail(&st);

    AccelerationParam bh = get_new_acceleration_param();
    bh.method = ACCELERATION_METHOD_BARNES_HUT;
    bh.opening_angle = 0.5;
    bh.softening_length = 1e-3;
    bh.max_num_particles_per_leaf = 1;
    st = acceleration(b, &system, &bh);
    if (st.return_code != GRAV_SUCCESS) return fail(&st);

    IntegratorParam ip = get_new_integrator_param();
    ip.integrator = INTEGRATOR_LEAPFROG;
    ip.dt = 1e-3;
    OutputParam op = get_new_output_param();
    SimulationStatus status;
    Settings settings = get_new_settings();
    settings.verbose = 0;
    settings.enable_progress_bar = false;
    st = launch_simulation(&system, &ip, &pairwise, &op, &status, &settings, 10 * 1e-3);
    if (st.return_code != GRAV_SUCCESS) return fail(&st);

    FILE *out = fopen(argv[2], "wb");
    if (!out) return 2;
    fwrite(a, sizeof(double), 3 * (size_t)n, out);
    fwrite(b, sizeof(double), 3 * (size_t)n, out);
    fwrite(system.x, sizeof(double), 3 * (size_t)n, out);
    fwrite(system.v, sizeof(double), 3 * (size_t)n, out);
    fclose(out);
    free(a);
    free(b);
    free_system(&system);
    return 0;
}
