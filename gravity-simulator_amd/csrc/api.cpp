// api.cpp -- the C-ABI entry points: the reference's acceleration / octree /
// integrator / Python-interface functions, implemented on the GPU, plus the
// additive device-resident entry points of grav_sim_hip_ext.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gs_internal.h"
#include "octree.h"

namespace gs {

hipError_t launch_copy(Context *c, const void *src, void *dst, size_t bytes);
ErrorStatus learned_acceleration(Context *c, const double *x, const double *m, int n, double G,
                                 const AccelerationParam *p, double *a, long long *interactions);

static long long g_last_walk_interactions = 0;
static long long g_last_walk_far_pairs = 0;
static double g_group_expansion_ratio = 0.1;
static double g_group_expansion_angle_factor = 1.25;

ErrorStatus device_acceleration(Context *c, const double *x, const double *m, int n, double G,
                                const AccelerationParam *p, double *a)
{
    const double eps = p->softening_length;
    g_last_walk_far_pairs = 0;
    switch (p->method) {
        case ACCELERATION_METHOD_PAIRWISE:
        case ACCELERATION_METHOD_MASSLESS:
            GS_HIP(launch_direct_f64(c, x, m, n, G, eps,
                                     p->method == ACCELERATION_METHOD_MASSLESS, a));
            return make_success_error_status();
        case ACCELERATION_METHOD_PAIRWISE_FP32:
        case ACCELERATION_METHOD_MASSLESS_FP32: {
            // massless sources carry zero mass and contribute exactly zero
            const Fp32Config cfg = fp32_config();
            const size_t pb = direct_f32_partial_bytes(n, cfg, c->cu_count);
            void *part;
            GS_TRY(slot(c, SLOT_PARTIAL, pb, &part));
            GS_HIP(launch_direct_f32(c, x, m, n, G, eps, a, part, pb, cfg));
            return make_success_error_status();
        }
        case ACCELERATION_METHOD_BARNES_HUT:
        case ACCELERATION_METHOD_BARNES_HUT_QUADRUPOLE: {
            DeviceTree *t;
            GS_TRY(tree_build(c, x, m, n, p->max_num_particles_per_leaf, nullptr, -1.0,
                              p->method == ACCELERATION_METHOD_BARNES_HUT_QUADRUPOLE, &t));
            GS_TRY(tree_walk(c, t, x, G, eps, p->opening_angle, a, &g_last_walk_interactions,
                             p->method == ACCELERATION_METHOD_BARNES_HUT_QUADRUPOLE ? 2 : 1));
            return make_success_error_status();
        }
        case ACCELERATION_METHOD_BARNES_HUT_GROUP:
        case ACCELERATION_METHOD_BARNES_HUT_GROUP_QUADRUPOLE: {
            const bool q = p->method == ACCELERATION_METHOD_BARNES_HUT_GROUP_QUADRUPOLE;
            DeviceTree *t;
            GS_TRY(tree_build(c, x, m, n, p->max_num_particles_per_leaf, nullptr, -1.0, q, &t));
            GS_TRY(tree_walk_group(c, t, G, eps, p->opening_angle, a, &g_last_walk_interactions,
                                   q ? 2 : 1));
            return make_success_error_status();
        }
        case ACCELERATION_METHOD_BARNES_HUT_GROUP_EXPANSION: {
            DeviceTree *t;
            GS_TRY(tree_build(c, x, m, n, p->max_num_particles_per_leaf, nullptr, -1.0, true, &t));
            GS_TRY(tree_walk_group(c, t, G, eps, p->opening_angle, a, &g_last_walk_interactions, 1,
                                   g_group_expansion_ratio,
                                   g_group_expansion_angle_factor * p->opening_angle,
                                   &g_last_walk_far_pairs));
            return make_success_error_status();
        }
        case ACCELERATION_METHOD_BARNES_HUT_LEARNED:
            return learned_acceleration(c, x, m, n, G, p, a, &g_last_walk_interactions);
        default:
            return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Unknown acceleration method. Got: %d",
                                        p->method);
    }
}

static bool known_method(int method)
{
    return (method >= ACCELERATION_METHOD_PAIRWISE && method <= ACCELERATION_METHOD_BARNES_HUT) ||
           (method >= ACCELERATION_METHOD_PAIRWISE_FP32 &&
            method <= ACCELERATION_METHOD_BARNES_HUT_GROUP_EXPANSION);
}

// Host-buffer form used by acceleration(): H2D x and m, compute, D2H a.
static ErrorStatus host_acceleration(double *a, const System *system,
                                     const AccelerationParam *p)
{
    if (!a || !system || !p)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    if (!known_method(p->method))
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Unknown acceleration method. Got: %d",
                                    p->method);
    const int n = system->num_particles;
    if (n <= 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Number of particles must be positive. Got: %d", n);
    if (!system->x || !system->m)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System array x or m is NULL");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    void *dx, *dm, *da;
    GS_TRY(slot(c, SLOT_X, sizeof(double) * 3 * n, &dx));
    GS_TRY(slot(c, SLOT_M, sizeof(double) * n, &dm));
    GS_TRY(slot(c, SLOT_A, sizeof(double) * 3 * n, &da));
    GS_HIP(hipMemcpyAsync(dx, system->x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    GS_HIP(hipMemcpyAsync(dm, system->m, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    GS_TRY(device_acceleration(c, static_cast<double *>(dx), static_cast<double *>(dm), n,
                               system->G, p, static_cast<double *>(da)));
    GS_HIP(hipMemcpyAsync(a, da, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    return make_success_error_status();
}

}  // namespace gs

using namespace gs;

// ====================================================================== GsSim
struct GsSim {
    int n = 0;
    double G = 1.0;
    double *x = nullptr, *v = nullptr, *m = nullptr, *a = nullptr;
    double *xe = nullptr, *ve = nullptr;
    // the other half of the ping-pong of the fused FP64 Leapfrog step; x, xe
    // and v hold the state whenever no step call is running
    double *x2 = nullptr, *xe2 = nullptr, *v2 = nullptr;
    // captured Leapfrog step (drift, forces, kick) and what it was captured for
    hipGraphExec_t lf_exec = nullptr;
    struct Key {
        int method = -1;
        double eps = 0.0, dt = 0.0;
        Fp32Config cfg;
        void *partial = nullptr;
        double *x = nullptr;  // the state buffers current when captured
        bool operator==(const Key &o) const
        {
            return method == o.method && eps == o.eps && dt == o.dt && cfg.kind == o.cfg.kind &&
                   cfg.block == o.cfg.block && cfg.ipt == o.cfg.ipt && cfg.jsplit == o.cfg.jsplit &&
                   partial == o.partial && x == o.x;
        }
    } lf_key;
};

namespace gs {

// Methods whose force evaluation is a fixed sequence of kernel launches on
// the stream (no host synchronisation, allocation or memset), so a Leapfrog
// step can be captured into a HIP graph and replayed.
static bool graph_method(int method)
{
    return method == ACCELERATION_METHOD_PAIRWISE || method == ACCELERATION_METHOD_MASSLESS ||
           method == ACCELERATION_METHOD_PAIRWISE_FP32 ||
           method == ACCELERATION_METHOD_MASSLESS_FP32;
}

// The FP64 direct sums run the whole step as one kernel (drift, forces and
// kick fused, into the other half of the ping-pong); the other methods as
// drift, force evaluation and kick.
static bool fused_step_method(int method)
{
    return method == ACCELERATION_METHOD_PAIRWISE || method == ACCELERATION_METHOD_MASSLESS;
}

// parity: which half of the fused step's ping-pong holds the state (0: x,
// xe, v; 1: x2, xe2, v2); the fused step flips it.
static ErrorStatus lf_step(Context *c, GsSim *s, const AccelerationParam *p, double dt,
                           int &parity)
{
    if (fused_step_method(p->method)) {
        double *x = parity ? s->x2 : s->x, *xe = parity ? s->xe2 : s->xe, *v = parity ? s->v2 : s->v;
        double *xo = parity ? s->x : s->x2, *xeo = parity ? s->xe : s->xe2, *vo = parity ? s->v : s->v2;
        GS_HIP(launch_lf_step_f64(c, x, xe, v, s->m, s->n, s->G, p->softening_length,
                                  p->method == ACCELERATION_METHOD_MASSLESS, dt, s->a, xo, xeo, vo,
                                  s->ve));
        parity ^= 1;
        return make_success_error_status();
    }
    GS_HIP(launch_lf_drift(c, s->x, s->xe, s->v, s->n, dt));
    GS_TRY(device_acceleration(c, s->x, s->m, s->n, s->G, p, s->a));
    GS_HIP(launch_lf_kick(c, s->v, s->ve, s->a, s->n, dt));
    return make_success_error_status();
}

// Steps one graph replay performs: two for the fused step (there and back
// through the ping-pong, starting from and ending at parity 1), one
// otherwise.
static int lf_graph_steps(const AccelerationParam *p) { return fused_step_method(p->method) ? 2 : 1; }

// Capture one step into s->lf_exec unless the cached graph was captured for
// the same parameters and buffers.  Called after an eager step, so every
// buffer the step needs already exists and nothing allocates while capturing.
static ErrorStatus lf_graph(Context *c, GsSim *s, const AccelerationParam *p, double dt)
{
    GsSim::Key key;
    key.method = p->method;
    key.eps = p->softening_length;
    key.dt = dt;
    key.cfg = fp32_config();
    key.partial = c->slots[SLOT_PARTIAL].ptr;
    key.x = s->x;
    if (s->lf_exec && key == s->lf_key) return make_success_error_status();
    int parity = 1;  // a captured fused pair starts after one eager step
    if (s->lf_exec) {
        GS_HIP(hipGraphExecDestroy(s->lf_exec));
        s->lf_exec = nullptr;
    }
    GS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    ErrorStatus st = make_success_error_status();
    for (int r = 0; r < lf_graph_steps(p) && st.return_code == GRAV_SUCCESS; r++)
        st = lf_step(c, s, p, dt, parity);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (st.return_code != GRAV_SUCCESS) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    GS_HIP(e);
    const hipError_t ei = hipGraphInstantiate(&s->lf_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    GS_HIP(ei);
    if (key.partial != c->slots[SLOT_PARTIAL].ptr) {
        // a buffer moved while capturing: the graph would point at freed memory
        (void)hipGraphExecDestroy(s->lf_exec);
        s->lf_exec = nullptr;
        return WRAP_RAISE_ERROR(GRAV_FAILURE, "Leapfrog graph capture saw a buffer reallocation");
    }
    s->lf_key = key;
    return make_success_error_status();
}

}  // namespace gs

extern "C" {

// ---------------------------------------------------------------- acceleration
ErrorStatus acceleration(double *a, const System *system, const AccelerationParam *p)
{
    return host_acceleration(a, system, p);
}

ErrorStatus acceleration_barnes_hut(double *a, const System *system, const AccelerationParam *p)
{
    if (!p) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "Acceleration parameter pointer is NULL");
    AccelerationParam q = *p;
    q.method = ACCELERATION_METHOD_BARNES_HUT;
    return host_acceleration(a, system, &q);
}

ErrorStatus construct_octree(LinearOctree *octree, const System *system,
                             const AccelerationParam *p, const double *box_center,
                             const double box_width)
{
    if (!octree) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "Octree pointer is NULL");
    if (!system) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System pointer is NULL");
    if (!p) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "Acceleration parameter pointer is NULL");
    const int n = system->num_particles;
    if (n <= 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Number of particles must be positive. Got: %d", n);
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    void *dx, *dm;
    GS_TRY(slot(c, SLOT_X, sizeof(double) * 3 * n, &dx));
    GS_TRY(slot(c, SLOT_M, sizeof(double) * n, &dm));
    GS_HIP(hipMemcpyAsync(dx, system->x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    GS_HIP(hipMemcpyAsync(dm, system->m, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    DeviceTree *t;
    GS_TRY(tree_build(c, static_cast<double *>(dx), static_cast<double *>(dm), n,
                      p->max_num_particles_per_leaf, box_center, box_width, false, &t));
    return tree_to_host(c, t, octree);
}

static const char *benchmark_method_name(int method)
{
    switch (method) {
        case ACCELERATION_METHOD_PAIRWISE: return "Pairwise";
        case ACCELERATION_METHOD_MASSLESS: return "Massless";
        case ACCELERATION_METHOD_BARNES_HUT: return "Barnes-Hut";
        case ACCELERATION_METHOD_PAIRWISE_FP32: return "Pairwise (FP32 split)";
        case ACCELERATION_METHOD_MASSLESS_FP32: return "Massless (FP32 split)";
        case ACCELERATION_METHOD_BARNES_HUT_LEARNED: return "Barnes-Hut (learned correction)";
        case ACCELERATION_METHOD_BARNES_HUT_QUADRUPOLE: return "Barnes-Hut (quadrupole)";
        case ACCELERATION_METHOD_BARNES_HUT_GROUP: return "Barnes-Hut (group walk)";
        case ACCELERATION_METHOD_BARNES_HUT_GROUP_QUADRUPOLE:
            return "Barnes-Hut (group walk, quadrupole)";
        case ACCELERATION_METHOD_BARNES_HUT_GROUP_EXPANSION:
            return "Barnes-Hut (group walk, local expansions)";
        default: return nullptr;
    }
}

// Directed pair interactions of one evaluation: all ordered pairs for the
// pairwise sums, massive sources on every other particle for the massless
// forms, the walk's own count for the tree methods.
static double benchmark_interactions(const System *system, int method)
{
    const double n = system->num_particles;
    switch (method) {
        case ACCELERATION_METHOD_PAIRWISE:
        case ACCELERATION_METHOD_PAIRWISE_FP32: return n * (n - 1.0);
        case ACCELERATION_METHOD_MASSLESS:
        case ACCELERATION_METHOD_MASSLESS_FP32: {
            double massive = 0.0;
            for (int i = 0; i < system->num_particles; i++) massive += system->m[i] != 0.0;
            return massive * (n - 1.0);
        }
        default: return static_cast<double>(g_last_walk_interactions);
    }
}

ErrorStatus gs_benchmark_acceleration(const System *system, const AccelerationParam *params,
                                      const int num_params, const int *num_times,
                                      GsBenchmarkRecord *records)
{
    if (!system || !params || !num_times)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    const int n = system->num_particles;
    std::vector<double> ref(static_cast<size_t>(3) * n), a(static_cast<size_t>(3) * n);
    // The MAE reference is test 0's first run (src/acceleration.c:419-435).
    // When test 0 is skipped the reference would compare against memory
    // nothing wrote; here that reference is then evaluated once, untimed.
    if (num_params > 0 && num_times[0] <= 0) GS_TRY(acceleration(ref.data(), system, &params[0]));
    GsDeviceInfo info;
    GS_TRY(gs_device_info(&info));
    fputs("Benchmarking acceleration...\n", stdout);
    for (int i = 0; i < num_params; i++) {
        const int times = num_times[i];
        GsBenchmarkRecord rec;
        memset(&rec, 0, sizeof(rec));
        rec.method = params[i].method;
        rec.num_times = times;
        if (times <= 0) {
            printf("Test %d:    Skipped since num_times: %d <= 0\n\n", i, times);
            if (records) records[i] = rec;
            continue;
        }
        const char *name = benchmark_method_name(params[i].method);
        if (!name)
            return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Unknown acceleration method. Got: %d",
                                        params[i].method);
        std::vector<double> rt(times);
        double mae = 0.0;
        for (int j = 0; j < times; j++) {
            double *out = (i == 0 && j == 0) ? ref.data() : a.data();
            const double t0 = grav_get_current_time();
            GS_TRY(acceleration(out, system, &params[i]));
            rt[j] = grav_get_current_time() - t0;
            if (i != 0 && j == 0) {
                for (int k = 0; k < 3 * n; k++) mae += fabs(ref[k] - a[k]);
                mae /= n;
            }
        }
        double mean = 0.0;
        for (double t : rt) mean += t;
        mean /= times;
        double var = 0.0;
        for (double t : rt) var += (t - mean) * (t - mean);
        const double sd = times > 1 ? sqrt(var / (times - 1)) : 0.0;
        printf("Test %d:", i);
        printf("    Method: %s\n", name);
        printf("    Number of times: %d\n", times);
        printf("    Avg time: %.3g (+- %.3g) s\n", mean, sd);
        printf("    MAE: %.3g\n", mae);
        printf("\n");
        rec.mean_s = mean;
        rec.std_s = sd;
        rec.mae = mae;
        rec.interactions = benchmark_interactions(system, params[i].method);
        rec.interactions_per_s = mean > 0.0 ? rec.interactions / mean : 0.0;
        rec.particles_per_s = mean > 0.0 ? n / mean : 0.0;
        rec.flops_per_s = 20.0 * rec.interactions_per_s;
        const bool fp64 = params[i].method == ACCELERATION_METHOD_PAIRWISE ||
                          params[i].method == ACCELERATION_METHOD_MASSLESS;
        rec.peak_flops_per_s = 1e12 * (fp64 ? info.fp64_vector_tflops : info.fp32_vector_tflops);
        rec.roofline_frac = rec.flops_per_s / rec.peak_flops_per_s;
        if (records) records[i] = rec;
    }
    fflush(stdout);
    return make_success_error_status();
}

ErrorStatus benchmark_acceleration(const System *system, const AccelerationParam *params,
                                   const int num_params, const int *num_times)
{
    return gs_benchmark_acceleration(system, params, num_params, num_times, nullptr);
}

// ---------------------------------------------------------------- energy
double compute_energy(const System *system)
{
    if (!system || system->num_particles <= 0) return 0.0;
    Context *c;
    ErrorStatus st = context(&c);
    if (st.return_code != GRAV_SUCCESS) {
        print_and_free_traceback(&st);
        return NAN;
    }
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    const int n = system->num_particles;
    void *dx, *dv, *dm, *dp;
    const int nparts = (n + 255) / 256;
    if (slot(c, SLOT_X, sizeof(double) * 3 * n, &dx).return_code ||
        slot(c, SLOT_V, sizeof(double) * 3 * n, &dv).return_code ||
        slot(c, SLOT_M, sizeof(double) * n, &dm).return_code ||
        slot(c, SLOT_PARTIAL, sizeof(double) * nparts, &dp).return_code)
        return NAN;
    std::vector<double> part(nparts);
    int np = 0;
    if (hipMemcpyAsync(dx, system->x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(dv, system->v, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream) ||
        hipMemcpyAsync(dm, system->m, sizeof(double) * n, hipMemcpyHostToDevice, c->stream) ||
        launch_energy(c, static_cast<double *>(dx), static_cast<double *>(dv),
                      static_cast<double *>(dm), n, system->G, static_cast<double *>(dp), &np) ||
        hipMemcpyAsync(part.data(), dp, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream) ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return NAN;
    double e = 0.0;
    for (double v : part) e += v;
    return e;
}

void compute_energy_python(double *energy, const double G, const double *sol_state,
                           const int32 num_snapshots, const int32 num_particles)
{
    std::vector<double> x(3 * static_cast<size_t>(num_particles)),
        v(3 * static_cast<size_t>(num_particles)), m(num_particles);
    for (int s = 0; s < num_snapshots; s++) {
        const double *snap = sol_state + static_cast<size_t>(s) * num_particles * 7;
        for (int i = 0; i < num_particles; i++) {
            m[i] = snap[i * 7 + 0];
            for (int d = 0; d < 3; d++) {
                x[3 * i + d] = snap[i * 7 + 1 + d];
                v[3 * i + d] = snap[i * 7 + 4 + d];
            }
        }
        System sys = get_new_system();
        sys.num_particles = num_particles;
        sys.x = x.data();
        sys.v = v.data();
        sys.m = m.data();
        sys.G = G;
        energy[s] = compute_energy(&sys);
    }
}

// ---------------------------------------------------------------- GsSim
ErrorStatus gs_sim_create(GsSim **out, int n, double G)
{
    if (!out) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "out is NULL");
    if (n <= 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Number of particles must be positive. Got: %d", n);
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    GsSim *s = new GsSim();
    s->n = n;
    s->G = G;
    const size_t b3 = sizeof(double) * 3 * n, b1 = sizeof(double) * n;
    hipError_t e = hipMalloc(&s->x, b3);
    if (!e) e = hipMalloc(&s->v, b3);
    if (!e) e = hipMalloc(&s->m, b1);
    if (!e) e = hipMalloc(&s->a, b3);
    if (!e) e = hipMalloc(&s->xe, b3);
    if (!e) e = hipMalloc(&s->ve, b3);
    if (!e) e = hipMalloc(&s->x2, b3);
    if (!e) e = hipMalloc(&s->xe2, b3);
    if (!e) e = hipMalloc(&s->v2, b3);
    if (!e) e = hipMemsetAsync(s->xe, 0, b3, c->stream);
    if (!e) e = hipMemsetAsync(s->ve, 0, b3, c->stream);
    if (!e) e = hipMemsetAsync(s->a, 0, b3, c->stream);
    if (e) {
        gs_sim_destroy(s);
        return hip_error(e, "gs_sim_create allocation", __FILE__, __LINE__, __func__);
    }
    *out = s;
    return make_success_error_status();
}

ErrorStatus gs_sim_destroy(GsSim *s)
{
    if (!s) return make_success_error_status();
    Context *c;
    if (context(&c).return_code == GRAV_SUCCESS) {
        std::lock_guard<std::recursive_mutex> lk(c->mu);
        (void)hipStreamSynchronize(c->stream);
    }
    if (s->lf_exec) (void)hipGraphExecDestroy(s->lf_exec);
    (void)hipFree(s->x);
    (void)hipFree(s->v);
    (void)hipFree(s->m);
    (void)hipFree(s->a);
    (void)hipFree(s->xe);
    (void)hipFree(s->ve);
    (void)hipFree(s->x2);
    (void)hipFree(s->xe2);
    (void)hipFree(s->v2);
    delete s;
    return make_success_error_status();
}

ErrorStatus gs_sim_upload(GsSim *s, const double *x, const double *v, const double *m)
{
    if (!s) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "sim is NULL");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    const size_t b3 = sizeof(double) * 3 * s->n, b1 = sizeof(double) * s->n;
    if (x) GS_HIP(hipMemcpyAsync(s->x, x, b3, hipMemcpyHostToDevice, c->stream));
    if (v) GS_HIP(hipMemcpyAsync(s->v, v, b3, hipMemcpyHostToDevice, c->stream));
    if (m) GS_HIP(hipMemcpyAsync(s->m, m, b1, hipMemcpyHostToDevice, c->stream));
    GS_HIP(hipMemsetAsync(s->xe, 0, b3, c->stream));
    GS_HIP(hipMemsetAsync(s->ve, 0, b3, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    return make_success_error_status();
}

ErrorStatus gs_sim_download(GsSim *s, double *x, double *v)
{
    if (!s) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "sim is NULL");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    const size_t b3 = sizeof(double) * 3 * s->n;
    if (x) GS_HIP(hipMemcpyAsync(x, s->x, b3, hipMemcpyDeviceToHost, c->stream));
    if (v) GS_HIP(hipMemcpyAsync(v, s->v, b3, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    return make_success_error_status();
}

}  // extern "C"

namespace gs {

// Snapshot of a Leapfrog run between steps: x, and v_{n+1/2} brought back to
// v_n = v - 0.5 a dt in a scratch copy (src/integrator.c:1045-1072); the run's
// own v is left as it is.
ErrorStatus sim_leapfrog_snapshot(GsSim *s, SnapshotWriter *w, double t, double dt)
{
    if (!s || !w) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    double *vs;
    GS_TRY(w->scratch(&vs));
    GS_HIP(hipMemcpyAsync(vs, s->v, sizeof(double) * 3 * s->n, hipMemcpyDeviceToDevice,
                          c->stream));
    GS_HIP(launch_lf_sync(c, vs, s->a, s->n, dt));
    return w->write_device(s->x, vs, t, dt);
}

}  // namespace gs

extern "C" {

ErrorStatus gs_sim_acceleration(GsSim *s, const AccelerationParam *p, double *a_host)
{
    if (!s || !p) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    GS_TRY(device_acceleration(c, s->x, s->m, s->n, s->G, p, s->a));
    if (a_host) {
        GS_HIP(hipMemcpyAsync(a_host, s->a, sizeof(double) * 3 * s->n, hipMemcpyDeviceToHost,
                              c->stream));
        GS_HIP(hipStreamSynchronize(c->stream));
    }
    return make_success_error_status();
}

ErrorStatus gs_sim_leapfrog_begin(GsSim *s, const AccelerationParam *p, double dt)
{
    if (!s || !p) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    GS_TRY(device_acceleration(c, s->x, s->m, s->n, s->G, p, s->a));
    GS_HIP(launch_lf_half_kick(c, s->v, s->ve, s->a, s->n, dt));
    return make_success_error_status();
}

ErrorStatus gs_sim_leapfrog_steps(GsSim *s, const AccelerationParam *p, double dt, int64 k)
{
    if (!s || !p) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    // with per-kernel timing on, every step runs eagerly so that its kernels
    // carry their own events
    const int per = lf_graph_steps(p);
    const bool graph = k >= 1 + 2 * per && graph_method(p->method) && !c->timing;
    int64 i = 0;
    int parity = 0;
    if (graph) {
        GS_TRY(lf_step(c, s, p, dt, parity));  // allocates what the step needs
        i = 1;
        GS_TRY(lf_graph(c, s, p, dt));
        for (; i + per <= k; i += per) GS_HIP(hipGraphLaunch(s->lf_exec, c->stream));
    }
    for (; i < k; i++) GS_TRY(lf_step(c, s, p, dt, parity));
    if (parity) {
        // an odd number of fused steps: the state is in the second half
        const size_t b3 = sizeof(double) * 3 * s->n;
        GS_HIP(hipMemcpyAsync(s->x, s->x2, b3, hipMemcpyDeviceToDevice, c->stream));
        GS_HIP(hipMemcpyAsync(s->xe, s->xe2, b3, hipMemcpyDeviceToDevice, c->stream));
        GS_HIP(hipMemcpyAsync(s->v, s->v2, b3, hipMemcpyDeviceToDevice, c->stream));
    }
    return make_success_error_status();
}

ErrorStatus gs_sim_leapfrog_end(GsSim *s, double dt)
{
    if (!s) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "sim is NULL");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    GS_HIP(launch_lf_sync(c, s->v, s->a, s->n, dt));
    return make_success_error_status();
}

ErrorStatus gs_sim_energy(GsSim *s, double *energy)
{
    if (!s || !energy) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    const int nparts = (s->n + 255) / 256;
    void *dp;
    GS_TRY(slot(c, SLOT_PARTIAL, sizeof(double) * nparts, &dp));
    int np = 0;
    GS_HIP(launch_energy(c, s->x, s->v, s->m, s->n, s->G, static_cast<double *>(dp), &np));
    std::vector<double> part(np);
    GS_HIP(hipMemcpyAsync(part.data(), dp, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    double e = 0.0;
    for (double v : part) e += v;
    *energy = e;
    return make_success_error_status();
}

ErrorStatus gs_sim_accel_error(GsSim *s, const AccelerationParam *p, int num_samples,
                               double *median, double *p99, double *max)
{
    if (!s || !p || !median || !p99 || !max)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    const int n = s->n;
    const int ns = (num_samples <= 0 || num_samples >= n) ? n : num_samples;
    std::vector<int> targets(ns);
    for (int t = 0; t < ns; t++)
        targets[t] = static_cast<int>((static_cast<long long>(t) * n) / ns);
    GS_TRY(device_acceleration(c, s->x, s->m, s->n, s->G, p, s->a));
    std::vector<double> a(static_cast<size_t>(3) * n), ref(static_cast<size_t>(3) * ns);
    void *dt, *dr;
    GS_TRY(slot(c, SLOT_IDS, sizeof(int) * ns, &dt));
    GS_TRY(slot(c, SLOT_V, sizeof(double) * 3 * ns, &dr));
    GS_HIP(hipMemcpyAsync(dt, targets.data(), sizeof(int) * ns, hipMemcpyHostToDevice, c->stream));
    const bool skip = p->method == ACCELERATION_METHOD_MASSLESS ||
                      p->method == ACCELERATION_METHOD_MASSLESS_FP32;
    const size_t pb = direct_f64_sampled_bytes(n, ns);
    void *part;
    GS_TRY(slot(c, SLOT_PARTIAL, pb, &part));
    GS_HIP(launch_direct_f64_sampled(c, s->x, s->m, n, s->G, p->softening_length, skip,
                                     static_cast<int *>(dt), ns, static_cast<double *>(dr), part,
                                     pb));
    GS_HIP(hipMemcpyAsync(a.data(), s->a, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipMemcpyAsync(ref.data(), dr, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost,
                          c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    std::vector<double> err;
    err.reserve(ns);
    for (int t = 0; t < ns; t++) {
        const int i = targets[t];
        double dn = 0.0, rn = 0.0;
        for (int d = 0; d < 3; d++) {
            const double df = a[3 * i + d] - ref[3 * t + d];
            dn += df * df;
            rn += ref[3 * t + d] * ref[3 * t + d];
        }
        if (rn > 0.0) err.push_back(sqrt(dn / rn));
    }
    if (err.empty()) {
        *median = *p99 = *max = 0.0;
        return make_success_error_status();
    }
    std::sort(err.begin(), err.end());
    const size_t k = err.size();
    *median = k % 2 ? err[k / 2] : 0.5 * (err[k / 2 - 1] + err[k / 2]);
    *p99 = err[std::min(k - 1, static_cast<size_t>(std::ceil(0.99 * k)) - 1)];
    *max = err.back();
    return make_success_error_status();
}

// ---------------------------------------------------------------- octree extras
ErrorStatus gs_octree_sort(const System *system, int64 *sorted_keys, int *sorted_idx,
                           double *center, double *width)
{
    if (!system || !sorted_keys || !sorted_idx)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    const int n = system->num_particles;
    if (n <= 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Number of particles must be positive. Got: %d", n);
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    void *dx;
    GS_TRY(slot(c, SLOT_X, sizeof(double) * 3 * n, &dx));
    GS_HIP(hipMemcpyAsync(dx, system->x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    DeviceTree t;
    GS_TRY(tree_sort(c, static_cast<double *>(dx), n, nullptr, -1.0, &t));
    GS_HIP(hipMemcpyAsync(sorted_keys, t.skeys, sizeof(int64) * n, hipMemcpyDeviceToHost,
                          c->stream));
    GS_HIP(hipMemcpyAsync(sorted_idx, t.sidx, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    if (center)
        for (int d = 0; d < 3; d++) center[d] = t.center[d];
    if (width) *width = t.width;
    return make_success_error_status();
}

ErrorStatus gs_last_walk_interactions(long long *count)
{
    if (!count) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "count is NULL");
    *count = g_last_walk_interactions;
    return make_success_error_status();
}

ErrorStatus gs_last_walk_expansion_pairs(long long *count)
{
    if (!count) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "count is NULL");
    *count = g_last_walk_far_pairs;
    return make_success_error_status();
}

ErrorStatus gs_set_group_expansion_ratio(double phi)
{
    if (!(phi > 0.0 && phi < 1.0))
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Group expansion ratio must be in (0, 1). Got: %g", phi);
    g_group_expansion_ratio = phi;
    return make_success_error_status();
}

ErrorStatus gs_get_group_expansion_ratio(double *phi)
{
    if (!phi) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "phi is NULL");
    *phi = g_group_expansion_ratio;
    return make_success_error_status();
}

ErrorStatus gs_set_group_expansion_angle_factor(double k)
{
    if (!(k >= 1.0 && k <= 4.0))
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Group expansion angle factor must be in [1, 4]. Got: %g", k);
    g_group_expansion_angle_factor = k;
    return make_success_error_status();
}

ErrorStatus gs_get_group_expansion_angle_factor(double *k)
{
    if (!k) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "k is NULL");
    *k = g_group_expansion_angle_factor;
    return make_success_error_status();
}

// ---------------------------------------------------------------- copy calibration
ErrorStatus gs_copy_calibrate(size_t bytes, int reps, double *gbs)
{
    if (!gbs) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "gbs is NULL");
    if (bytes < 16 || reps < 1) return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "bytes >= 16, reps >= 1");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    bytes &= ~static_cast<size_t>(15);
    void *src = nullptr, *dst = nullptr;
    GS_HIP(hipMalloc(&src, bytes));
    hipError_t e = hipMalloc(&dst, bytes);
    if (!e) e = hipMemsetAsync(src, 0, bytes, c->stream);
    if (!e) e = launch_copy(c, src, dst, bytes);  // warm-up
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!e) e = hipEventCreate(&e0);
    if (!e) e = hipEventCreate(&e1);
    if (!e) e = hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps && !e; r++) e = launch_copy(c, src, dst, bytes);
    if (!e) e = hipEventRecord(e1, c->stream);
    if (!e) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (!e) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (e) return hip_error(e, "copy calibration", __FILE__, __LINE__, __func__);
    *gbs = 2.0 * static_cast<double>(bytes) * reps / (ms * 1e-3) / 1e9;
    return make_success_error_status();
}

ErrorStatus gs_issue_rates(double *fma_cycles, double *mul_cycles, double *rsq_cycles)
{
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    const int blocks = 8 * 4 * c->cu_count;  // 32 wavefronts per CU, several rounds
    void *p;
    GS_TRY(slot(c, SLOT_MISC, sizeof(float) * 256 * static_cast<size_t>(blocks), &p));
    float *out = static_cast<float *>(p);
    hipDeviceProp_t prop;
    GS_HIP(hipGetDeviceProperties(&prop, c->device));
    const double clock_hz = static_cast<double>(prop.clockRate) * 1e3;
    const double simds = 4.0 * c->cu_count;
    double *res[3] = {fma_cycles, mul_cycles, rsq_cycles};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (!e) e = hipEventCreate(&e1);
    for (int kind = 0; kind < 3 && !e; kind++) {
        e = launch_issue_probe(c, kind, out, blocks);  // warm-up
        float best = 0.0f;
        for (int r = 0; r < 3 && !e; r++) {
            e = hipEventRecord(e0, c->stream);
            if (!e) e = launch_issue_probe(c, kind, out, blocks);
            if (!e) e = hipEventRecord(e1, c->stream);
            if (!e) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (!e) e = hipEventElapsedTime(&ms, e0, e1);
            if (!e && (r == 0 || ms < best)) best = ms;
        }
        if (!e && res[kind])
            *res[kind] = best * 1e-3 * clock_hz * simds / issue_probe_instructions(blocks);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e) return hip_error(e, "issue-rate probe", __FILE__, __LINE__, __func__);
    return make_success_error_status();
}

// ---------------------------------------------------------------- FP32 autotuner
// (block, targets per thread): block 4/8/16 = the shared-target
// scalar-source kernel with that many wavefronts per 64 targets, 0 = the
// scalar-source kernel, 128/256/512 = the LDS-tile kernel
static const int kShapes[][2] = {{8, 1},   {4, 1},   {16, 1},  {0, 1},   {0, 2},   {0, 4}, {256, 1},
                                  {256, 2}, {256, 4}, {128, 2}, {128, 4}, {512, 1}, {512, 2}};

static bool shared_target_block(int block) { return block == 4 || block == 8 || block == 16; }

static Fp32Config shape_config(int block, int ipt, int split)
{
    Fp32Config cfg;
    cfg.kind = shared_target_block(block) ? 2 : block == 0 ? 1 : 0;
    cfg.block = block == 0 ? 256 : block;
    cfg.ipt = ipt;
    cfg.jsplit = split;
    return cfg;
}

static int shape_block(const Fp32Config &cfg) { return cfg.kind == 1 ? 0 : cfg.block; }

ErrorStatus gs_set_fp32_config(int block, int ipt, int split)
{
    bool ok = false;
    for (auto &sh : kShapes) ok = ok || (sh[0] == block && sh[1] == ipt);
    if (!ok || split < 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Unsupported FP32 shape (%d, %d, %d)", block,
                                    ipt, split);
    fp32_config() = shape_config(block, ipt, split);
    return make_success_error_status();
}

ErrorStatus gs_get_fp32_config(int *block, int *ipt, int *split)
{
    const Fp32Config &cfg = fp32_config();
    if (block) *block = shape_block(cfg);
    if (ipt) *ipt = cfg.ipt;
    if (split) *split = cfg.jsplit;
    return make_success_error_status();
}

ErrorStatus gs_autotune_fp32(int n, int *block, int *ipt, int *split)
{
    if (n <= 0) return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "n must be positive");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    // synthetic input: a deterministic unit cube lattice jittered by an LCG
    std::vector<double> hx(static_cast<size_t>(3) * n), hm(n, 1.0 / n);
    unsigned long long st = 12345;
    for (auto &v : hx) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        v = static_cast<double>(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
    GsSim *s;
    GS_TRY(gs_sim_create(&s, n, 1.0));
    ErrorStatus es = gs_sim_upload(s, hx.data(), nullptr, hm.data());
    if (es.return_code) {
        gs_sim_destroy(s);
        return es;
    }
    const int splits[] = {0, 4, 8, 16, 32, 64};
    struct Timed {
        Fp32Config cfg;
        float ms;
        size_t bytes;
    };
    std::vector<Timed> timed;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (auto &sh : kShapes)
        for (int sp : splits) {
            const Fp32Config cfg = shape_config(sh[0], sh[1], sp);
            const size_t pb = direct_f32_partial_bytes(n, cfg, c->cu_count);
            void *part;
            ErrorStatus e2 = slot(c, SLOT_PARTIAL, pb, &part);
            if (e2.return_code) {
                free_traceback(&e2);
                continue;
            }
            if (launch_direct_f32(c, s->x, s->m, n, 1.0, 1e-3, s->a, part, pb, cfg)) continue;
            (void)hipEventRecord(e0, c->stream);
            for (int r = 0; r < 3; r++)
                (void)launch_direct_f32(c, s->x, s->m, n, 1.0, 1e-3, s->a, part, pb, cfg);
            (void)hipEventRecord(e1, c->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            timed.push_back({cfg, ms, pb});
        }
    // the fastest shape, except that among the shapes within 1 % of it (the
    // run-to-run spread of these timings) the one with the fewest scratch
    // bytes -- the FP64 partial sums it writes and reads back -- wins
    Fp32Config best = fp32_config();
    float best_ms = 1e30f;
    for (const Timed &t : timed) best_ms = std::min(best_ms, t.ms);
    size_t best_bytes = ~size_t(0);
    float best_tie_ms = 1e30f;
    for (const Timed &t : timed)
        if (t.ms <= 1.01f * best_ms &&
            (t.bytes < best_bytes || (t.bytes == best_bytes && t.ms < best_tie_ms))) {
            best = t.cfg;
            best_bytes = t.bytes;
            best_tie_ms = t.ms;
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    gs_sim_destroy(s);
    fp32_config() = best;
    if (block) *block = shape_block(best);
    if (ipt) *ipt = best.ipt;
    if (split) *split = best.jsplit;
    return make_success_error_status();
}

}  // extern "C"
