// learned.cpp -- host side of the learned Barnes-Hut correction: the
// BARNES_HUT_LEARNED evaluation (coarse walk + MLP + spot check + fallback),
// training-set generation from the library's own direct-sum kernel, and the
// gs_learned_* entry points.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "gs_internal.h"
#include "mlp.h"
#include "octree.h"

namespace gs {

static long long g_learned_interactions = 0;

// Evenly strided spot-check targets (the same rule as the error sampler).
static std::vector<int> strided_targets(int n, int ns)
{
    std::vector<int> t(ns);
    for (int k = 0; k < ns; k++) t[k] = static_cast<int>((static_cast<long long>(k) * n) / ns);
    return t;
}

static double percentile99(std::vector<double> &e)
{
    if (e.empty()) return 0.0;
    std::sort(e.begin(), e.end());
    const size_t k = e.size();
    return e[std::min(k - 1, static_cast<size_t>(std::ceil(0.99 * k)) - 1)];
}

// 99th-percentile relative error of a (device, all n) against the FP32
// split-precision direct sum on ns strided targets.
static ErrorStatus spot_check(Context *c, const double *x, const double *m, int n, double G,
                              double eps, const double *a, int ns, double *p99)
{
    const std::vector<int> tg = strided_targets(n, ns);
    const size_t need_bytes = direct_f32_targets_bytes(n, ns);
    void *buf, *sp;
    GS_TRY(slot(c, SLOT_PARTIAL, need_bytes, &buf));
    const size_t off = ((sizeof(int) * ns + 255) / 256) * 256;
    GS_TRY(slot(c, SLOT_SPOT, off + 2 * sizeof(double) * 3 * ns, &sp));
    int *dt = static_cast<int *>(sp);
    double *dr = reinterpret_cast<double *>(static_cast<char *>(sp) + off);
    double *dg = dr + 3 * static_cast<size_t>(ns);
    GS_HIP(hipMemcpyAsync(dt, tg.data(), sizeof(int) * ns, hipMemcpyHostToDevice, c->stream));
    const int rec = timing_begin(c, "spot_check");
    GS_HIP(launch_direct_f32_targets(c, x, m, n, G, eps, dt, ns, dr, buf, need_bytes));
    timing_end(c, rec);
    GS_HIP(launch_gather_rows3(c, a, dt, ns, dg));
    // reference and gathered result: one copy
    std::vector<double> rg(6 * static_cast<size_t>(ns));
    GS_HIP(hipMemcpyAsync(rg.data(), dr, sizeof(double) * 6 * ns, hipMemcpyDeviceToHost,
                          c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    const double *ref = rg.data(), *got = rg.data() + 3 * static_cast<size_t>(ns);
    std::vector<double> err;
    err.reserve(ns);
    for (int k = 0; k < ns; k++) {
        double dn = 0.0, rn = 0.0;
        for (int d = 0; d < 3; d++) {
            const double df = got[3 * k + d] - ref[3 * k + d];
            dn += df * df;
            rn += ref[3 * k + d] * ref[3 * k + d];
        }
        if (rn > 0.0) err.push_back(std::sqrt(dn / rn));
    }
    *p99 = percentile99(err);
    return make_success_error_status();
}

// Coarse walk with features: a (device), raw or normalised features into F.
static ErrorStatus coarse_walk(Context *c, const double *x, const double *m, int n, double G,
                               double eps, double theta, int leaf, double *a, bool normalise,
                               float **F_out, DeviceTree **t_out, float **wf_out = nullptr)
{
    DeviceTree *t;
    GS_TRY(tree_build(c, x, m, n, leaf, nullptr, -1.0, true, &t, true));
    void *wf, *F;
    GS_TRY(slot(c, SLOT_FEAT, sizeof(float) * MLP_WALK_FEAT * static_cast<size_t>(n), &wf));
    GS_TRY(slot(c, SLOT_MLP_X, sizeof(float) * MLP_F * static_cast<size_t>(n), &F));
    GS_TRY(tree_walk(c, t, x, G, eps, theta, a, &g_learned_interactions, 1,
                     static_cast<float *>(wf)));
    const int rec = timing_begin(c, "learned_features");
    GS_TRY(mlp_features(c, a, static_cast<float *>(wf), n, G, normalise, static_cast<float *>(F)));
    timing_end(c, rec);
    *F_out = static_cast<float *>(F);
    if (t_out) *t_out = t;
    if (wf_out) *wf_out = static_cast<float *>(wf);
    return make_success_error_status();
}

ErrorStatus learned_acceleration(Context *c, const double *x, const double *m, int n, double G,
                                 const AccelerationParam *p, double *a, long long *interactions)
{
    MlpModel &M = mlp_model();
    if (!M.trained)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR,
                                "The learned Barnes-Hut correction has no trained network "
                                "(train it with gs_learned_train or install parameters with "
                                "gs_learned_set_params)");
    const double eps = p->softening_length;
    float *F, *wf;
    DeviceTree *t;
    GS_TRY(coarse_walk(c, x, m, n, G, eps, p->opening_angle, p->max_num_particles_per_leaf, a,
                       true, &F, &t, &wf));
    long long cnt = g_learned_interactions;
    const int rec = timing_begin(c, "learned_mlp");
    GS_TRY(mlp_apply(c, F, wf, G, n, a));
    timing_end(c, rec);
    M.last_fell_back = 0;
    M.last_p99 = -1.0;
    M.last_p99_coarse = -1.0;
    if (M.spot_samples > 0) {
        const int ns = std::min(M.spot_samples, n);
        GS_TRY(spot_check(c, x, m, n, G, eps, a, ns, &M.last_p99));
        M.last_p99_coarse = M.last_p99;
        if (!(M.last_p99 <= M.spot_tol)) {
            // the group walk with quadrupole moments at the fallback angle
            // (the tree already holds the second moments), checked again on
            // the same particles; last_p99 is then the fallback's error
            long long c2 = 0;
            GS_TRY(tree_walk_group(c, t, G, eps, M.fallback_theta, a, &c2, 2));
            cnt += c2;
            GS_TRY(spot_check(c, x, m, n, G, eps, a, ns, &M.last_p99));
            M.last_fell_back = M.last_p99 <= M.spot_tol ? 1 : 2;
        }
    }
    if (interactions) *interactions = cnt;
    return make_success_error_status();
}

}  // namespace gs

using namespace gs;

static ErrorStatus check_system(const System *s)
{
    if (!s) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System pointer is NULL");
    if (s->num_particles <= 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Number of particles must be positive. Got: %d",
                                    s->num_particles);
    if (!s->x || !s->m) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System array x or m is NULL");
    return make_success_error_status();
}

static ErrorStatus upload_xm(Context *c, const System *s, double **dx, double **dm, double **da)
{
    const int n = s->num_particles;
    void *px, *pm, *pa;
    GS_TRY(slot(c, SLOT_X, sizeof(double) * 3 * n, &px));
    GS_TRY(slot(c, SLOT_M, sizeof(double) * n, &pm));
    GS_TRY(slot(c, SLOT_A, sizeof(double) * 3 * n, &pa));
    GS_HIP(hipMemcpyAsync(px, s->x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    GS_HIP(hipMemcpyAsync(pm, s->m, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    *dx = static_cast<double *>(px);
    *dm = static_cast<double *>(pm);
    *da = static_cast<double *>(pa);
    return make_success_error_status();
}

static int leaf_or_default(int leaf) { return leaf < 1 ? 1 : leaf; }

// Relative errors of held-out samples nt .. nt+nh-1 with the network's
// correction scaled by alpha in {0, 1/8, ..., 1}; keeps the alpha of the
// smallest 99th percentile (ties: smallest median) and records the held-out
// statistics with and without the network.  ac: the coarse accelerations of
// the samples (sample order).
static ErrorStatus choose_output_scale(const std::vector<double> &ac, const std::vector<double> &ref,
                                       const std::vector<float> &ws, int nt, int nh,
                                       const std::vector<float> &Yh, double G, double tscale,
                                       double *alpha_out)
{
    MlpModel &M = mlp_model();
    double best_p99 = 0.0, best_med = 0.0, best_alpha = 0.0;
    // per sample: the prediction without the network (coarse + G (q + o)), the
    // network's full-scale term, the reference and its squared norm
    std::vector<double> base(3 * static_cast<size_t>(nh)), dir(3 * static_cast<size_t>(nh));
    std::vector<double> rn2(nh);
    for (int h = 0; h < nh; h++) {
        const int k = nt + h;
        const double *ak = &ac[3 * static_cast<size_t>(k)];
        const float *wk = &ws[static_cast<size_t>(MLP_WALK_FEAT) * k];
        const double q[3] = {wk[0], wk[1], wk[2]}, v[3] = {wk[3], wk[4], wk[5]};
        double e1[3], e2[3], e3[3];
        mlp_local_frame(ak, q, v, e1, e2, e3);
        const double s = G * static_cast<double>(wk[7]) * tscale;
        const double o[3] = {wk[10], wk[11], wk[12]};
        const float *y = &Yh[3 * static_cast<size_t>(h)];
        double r2 = 0.0;
        for (int d = 0; d < 3; d++) {
            base[3 * h + d] = ak[d] + G * (q[d] + o[d]);
            dir[3 * h + d] = s * (y[0] * e1[d] + y[1] * e2[d] + y[2] * e3[d]);
            const double rf = ref[3 * static_cast<size_t>(k) + d];
            r2 += rf * rf;
        }
        rn2[h] = r2;
    }
    std::vector<double> err(nh);
    for (int step = 0; step <= 8; step++) {
        const double alpha = step / 8.0;
        for (int h = 0; h < nh; h++) {
            const int k = nt + h;
            double dn = 0.0;
            for (int d = 0; d < 3; d++) {
                const double pd = base[3 * h + d] + alpha * dir[3 * h + d];
                const double rf = ref[3 * static_cast<size_t>(k) + d];
                dn += (pd - rf) * (pd - rf);
            }
            err[h] = rn2[h] > 0.0 ? std::sqrt(dn / rn2[h]) : 0.0;
        }
        const double p99 = percentile99(err);  // sorts err
        const double med = err[err.size() / 2];
        if (step == 0) {
            M.hold_p99_skip = p99;
            M.hold_med_skip = med;
        }
        if (step == 0 || p99 < best_p99 || (p99 == best_p99 && med < best_med)) {
            best_p99 = p99;
            best_med = med;
            best_alpha = alpha;
        }
    }
    M.hold_alpha = best_alpha;
    M.hold_p99_net = best_p99;
    M.hold_med_net = best_med;
    *alpha_out = best_alpha;
    return make_success_error_status();
}

extern "C" {

int gs_learned_num_params(void) { return MLP_NPARAM; }

// Training / calibration samples of one particle set: ns particles drawn
// without replacement (partial Fisher-Yates, 64-bit LCG), their FP32
// split-precision direct-sum labels, their coarse walk accelerations and
// their raw and walk features (host copies, sample order).
struct Samples {
    int n = 0, ns = 0;
    std::vector<int> idx;
    std::vector<double> ref, ac;
    std::vector<float> fs, ws;
};

static ErrorStatus build_samples(Context *c, const System *system, double eps, double theta,
                                 int leaf, int samples, unsigned int seed, Samples &S)
{
    const int n = system->num_particles;
    const int ns = std::min(samples, n);
    const double G = system->G;
    S.n = n;
    S.ns = ns;
    double *x, *m, *a;
    GS_TRY(upload_xm(c, system, &x, &m, &a));
    float *F, *wf;
    GS_TRY(coarse_walk(c, x, m, n, G, eps, theta, leaf_or_default(leaf), a, false, &F, nullptr, &wf));
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    unsigned long long st = 0x853C49E6748FEA9Bull ^ (static_cast<unsigned long long>(seed) << 1);
    for (int k = 0; k < ns; k++) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        const int j = k + static_cast<int>((st >> 33) % static_cast<unsigned long long>(n - k));
        std::swap(perm[k], perm[j]);
    }
    S.idx.assign(perm.begin(), perm.begin() + ns);
    // labels from the FP32 split-precision direct sum
    void *sp, *buf;
    const size_t lb = direct_f32_targets_bytes(n, ns);
    GS_TRY(slot(c, SLOT_PARTIAL, lb, &buf));
    auto up = [](size_t b) { return ((b + 255) / 256) * 256; };
    const size_t off_ref = up(sizeof(int) * ns);
    const size_t off_feat = off_ref + up(sizeof(double) * 3 * ns);
    const size_t off_wf = off_feat + up(sizeof(float) * MLP_F * static_cast<size_t>(ns));
    const size_t off_ac = off_wf + up(sizeof(float) * MLP_WALK_FEAT * static_cast<size_t>(ns));
    GS_TRY(slot(c, SLOT_SPOT, off_ac + sizeof(double) * 3 * static_cast<size_t>(ns), &sp));
    int *didx = static_cast<int *>(sp);
    double *dref = reinterpret_cast<double *>(static_cast<char *>(sp) + off_ref);
    float *dfs = reinterpret_cast<float *>(static_cast<char *>(sp) + off_feat);
    float *dwf = reinterpret_cast<float *>(static_cast<char *>(sp) + off_wf);
    double *dac = reinterpret_cast<double *>(static_cast<char *>(sp) + off_ac);
    GS_HIP(hipMemcpyAsync(didx, S.idx.data(), sizeof(int) * ns, hipMemcpyHostToDevice, c->stream));
    GS_HIP(launch_direct_f32_targets(c, x, m, n, G, eps, didx, ns, dref, buf, lb));
    GS_HIP(mlp_gather_rows(c, F, didx, ns, MLP_F, dfs));
    GS_HIP(mlp_gather_rows(c, wf, didx, ns, MLP_WALK_FEAT, dwf));
    GS_HIP(launch_gather_rows3(c, a, didx, ns, dac));
    S.ref.resize(3 * static_cast<size_t>(ns));
    S.ac.resize(3 * static_cast<size_t>(ns));
    S.fs.resize(static_cast<size_t>(MLP_F) * ns);
    S.ws.resize(static_cast<size_t>(MLP_WALK_FEAT) * ns);
    GS_HIP(hipMemcpyAsync(S.ref.data(), dref, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost,
                          c->stream));
    GS_HIP(hipMemcpyAsync(S.ac.data(), dac, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipMemcpyAsync(S.fs.data(), dfs, sizeof(float) * MLP_F * ns, hipMemcpyDeviceToHost,
                          c->stream));
    GS_HIP(hipMemcpyAsync(S.ws.data(), dwf, sizeof(float) * MLP_WALK_FEAT * ns,
                          hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    return make_success_error_status();
}

// Network outputs (host) of samples [k0, k0 + nh) with the model's feature
// normalisation.
static ErrorStatus forward_samples(Context *c, const Samples &S, int k0, int nh,
                                   std::vector<float> &Yh)
{
    MlpModel &M = mlp_model();
    std::vector<float> Xh(static_cast<size_t>(nh) * MLP_F);
    for (int h = 0; h < nh; h++)
        for (int f = 0; f < MLP_F; f++)
            Xh[static_cast<size_t>(MLP_F) * h + f] =
                (S.fs[static_cast<size_t>(MLP_F) * (k0 + h) + f] - M.fmean[f]) / M.fstd[f];
    const size_t xb = ((sizeof(float) * Xh.size() + 255) / 256) * 256;
    void *p;
    GS_TRY(slot(c, SLOT_MLP_X, xb + sizeof(float) * 3 * static_cast<size_t>(nh), &p));
    float *dX = static_cast<float *>(p);
    float *dY = reinterpret_cast<float *>(static_cast<char *>(p) + xb);
    GS_HIP(hipMemcpyAsync(dX, Xh.data(), sizeof(float) * Xh.size(), hipMemcpyHostToDevice, c->stream));
    GS_TRY(mlp_forward_rows(c, dX, nh, dY));
    Yh.resize(3 * static_cast<size_t>(nh));
    GS_HIP(hipMemcpyAsync(Yh.data(), dY, sizeof(float) * 3 * nh, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    return make_success_error_status();
}

ErrorStatus gs_learned_train(const System *system, double softening_length, double theta_coarse,
                             int max_num_particles_per_leaf, int samples, int epochs, int batch,
                             double lr, unsigned int seed, double *final_loss)
{
    GS_TRY(check_system(system));
    if (samples <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0.0) || theta_coarse < 0.0 ||
        softening_length < 0.0)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR,
                                "samples, epochs, batch and lr must be positive; theta and "
                                "softening non-negative");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    GS_TRY(mlp_ensure(c));
    MlpModel &M = mlp_model();
    const double G = system->G;
    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point t) {
        return std::chrono::duration<double, std::milli>(clk::now() - t).count();
    };
    const clk::time_point t0 = clk::now();
    Samples S;
    GS_TRY(build_samples(c, system, softening_length, theta_coarse, max_num_particles_per_leaf,
                         samples, seed, S));
    M.time_samples_ms = ms_since(t0);
    const clk::time_point t1 = clk::now();
    const int ns = S.ns;
    const std::vector<double> &ref = S.ref, &ac = S.ac;
    const std::vector<float> &fs = S.fs, &ws = S.ws;

    // targets: the residual left after the quadrupole and octupole sums, in
    // units of the next orders' scale G s3 and in the particle's local frame,
    //   r = (a_direct - a_coarse - G (q + o)) / (G s3),  T = (r.e1, r.e2, r.e3),
    // with each sample's norm clipped at the 99.9th percentile, scaled to unit
    // RMS per component
    std::vector<double> r(3 * static_cast<size_t>(ns)), rn(ns);
    for (int k = 0; k < ns; k++) {
        const double *ak = &ac[3 * static_cast<size_t>(k)];
        const float *wk = &ws[static_cast<size_t>(MLP_WALK_FEAT) * k];
        const double q[3] = {wk[0], wk[1], wk[2]}, v[3] = {wk[3], wk[4], wk[5]};
        double e1[3], e2[3], e3[3];
        mlp_local_frame(ak, q, v, e1, e2, e3);
        const double s3 = G * static_cast<double>(wk[7]);
        const double o[3] = {wk[10], wk[11], wk[12]};
        double rg[3];
        for (int d = 0; d < 3; d++)
            rg[d] = s3 > 0.0 ? (ref[3 * k + d] - ak[d] - G * (q[d] + o[d])) / s3 : 0.0;
        r[3 * k + 0] = rg[0] * e1[0] + rg[1] * e1[1] + rg[2] * e1[2];
        r[3 * k + 1] = rg[0] * e2[0] + rg[1] * e2[1] + rg[2] * e2[2];
        r[3 * k + 2] = rg[0] * e3[0] + rg[1] * e3[1] + rg[2] * e3[2];
        rn[k] = std::sqrt(rg[0] * rg[0] + rg[1] * rg[1] + rg[2] * rg[2]);
    }
    std::vector<double> srt(rn);
    const size_t kc = std::min(srt.size() - 1, static_cast<size_t>(0.999 * srt.size()));
    std::nth_element(srt.begin(), srt.begin() + kc, srt.end());
    const double clip = srt[kc];
    double ss = 0.0;
    for (int k = 0; k < ns; k++) {
        if (rn[k] > clip && rn[k] > 0.0)
            for (int d = 0; d < 3; d++) r[3 * k + d] *= clip / rn[k];
        for (int d = 0; d < 3; d++) ss += r[3 * k + d] * r[3 * k + d];
    }
    double tscale = std::sqrt(ss / (3.0 * ns));
    if (!(tscale > 0.0)) tscale = 1.0;
    // feature normalisation
    for (int f = 0; f < MLP_F; f++) {
        double s1 = 0.0, s2 = 0.0;
        for (int k = 0; k < ns; k++) s1 += fs[static_cast<size_t>(MLP_F) * k + f];
        const double mu = s1 / ns;
        for (int k = 0; k < ns; k++) {
            const double d = fs[static_cast<size_t>(MLP_F) * k + f] - mu;
            s2 += d * d;
        }
        const double sd = std::sqrt(s2 / ns);
        M.fmean[f] = static_cast<float>(mu);
        M.fstd[f] = static_cast<float>(sd > 1e-6 ? sd : 1.0);
    }
    M.tscale = static_cast<float>(tscale);
    std::vector<float> XT(static_cast<size_t>(ns) * (MLP_F + MLP_O));
    float *Xh = XT.data(), *Th = XT.data() + static_cast<size_t>(ns) * MLP_F;
    for (int k = 0; k < ns; k++) {
        for (int f = 0; f < MLP_F; f++)
            Xh[static_cast<size_t>(MLP_F) * k + f] =
                (fs[static_cast<size_t>(MLP_F) * k + f] - M.fmean[f]) / M.fstd[f];
        for (int d = 0; d < 3; d++) Th[3 * static_cast<size_t>(k) + d] = static_cast<float>(r[3 * k + d] / tscale);
    }
    // F (SLOT_MLP_X) is no longer needed: it holds the training set now
    void *dX;
    GS_TRY(slot(c, SLOT_MLP_X, sizeof(float) * XT.size(), &dX));
    GS_HIP(hipMemcpyAsync(dX, XT.data(), sizeof(float) * XT.size(), hipMemcpyHostToDevice, c->stream));
    float *X = static_cast<float *>(dX);
    M.trained = false;
    // the last tenth of the (randomly ordered) sample is held out to choose
    // the network's output scale
    const int nh = ns >= 20 ? ns / 10 : 0;
    const int nt = ns - nh;
    float *T = X + static_cast<size_t>(ns) * MLP_F;
    GS_HIP(hipStreamSynchronize(c->stream));
    M.time_prep_ms = ms_since(t1);
    const clk::time_point t2 = clk::now();
    GS_TRY(mlp_train(c, X, T, nt, epochs, std::min(batch, nt), lr, seed, final_loss));
    GS_HIP(hipStreamSynchronize(c->stream));
    M.time_loop_ms = ms_since(t2);
    const clk::time_point t3 = clk::now();
    double alpha = 1.0;
    if (nh > 0) {
        std::vector<float> Yh(3 * static_cast<size_t>(nh));
        void *py;
        GS_TRY(slot(c, SLOT_SPOT, sizeof(float) * 3 * static_cast<size_t>(nh), &py));
        GS_TRY(mlp_forward_rows(c, X + static_cast<size_t>(nt) * MLP_F, nh, static_cast<float *>(py)));
        GS_HIP(hipMemcpyAsync(Yh.data(), py, sizeof(float) * 3 * nh, hipMemcpyDeviceToHost, c->stream));
        GS_HIP(hipStreamSynchronize(c->stream));
        GS_TRY(choose_output_scale(ac, ref, ws, nt, nh, Yh, G, tscale, &alpha));
    }
    M.tscale_base = static_cast<float>(tscale);
    M.tscale = static_cast<float>(tscale * alpha);
    GS_HIP(hipStreamSynchronize(c->stream));
    M.time_select_ms = ms_since(t3);
    M.trained = true;
    return make_success_error_status();
}

ErrorStatus gs_learned_calibrate(const System *system, double softening_length, double theta_coarse,
                                 int max_num_particles_per_leaf, int samples, unsigned int seed)
{
    GS_TRY(check_system(system));
    if (samples < 20 || theta_coarse < 0.0 || softening_length < 0.0)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR,
                                "samples must be at least 20; theta and softening non-negative");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    MlpModel &M = mlp_model();
    if (!M.trained)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR,
                                "The learned Barnes-Hut correction has no trained network to "
                                "calibrate");
    Samples S;
    GS_TRY(build_samples(c, system, softening_length, theta_coarse, max_num_particles_per_leaf,
                         samples, seed, S));
    std::vector<float> Yh;
    GS_TRY(forward_samples(c, S, 0, S.ns, Yh));
    double alpha = 0.0;
    GS_TRY(choose_output_scale(S.ac, S.ref, S.ws, 0, S.ns, Yh, system->G, M.tscale_base, &alpha));
    M.tscale = static_cast<float>(M.tscale_base * alpha);
    return make_success_error_status();
}

ErrorStatus gs_learned_get_params(float *params, float *feature_mean, float *feature_std,
                                  float *target_scale)
{
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    GS_TRY(mlp_ensure(c));
    MlpModel &M = mlp_model();
    if (params) {
        GS_HIP(hipMemcpyAsync(params, M.w, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost,
                              c->stream));
        GS_HIP(hipStreamSynchronize(c->stream));
    }
    if (feature_mean) std::memcpy(feature_mean, M.fmean, sizeof(M.fmean));
    if (feature_std) std::memcpy(feature_std, M.fstd, sizeof(M.fstd));
    if (target_scale) *target_scale = M.tscale;
    return make_success_error_status();
}

ErrorStatus gs_learned_set_params(const float *params, const float *feature_mean,
                                  const float *feature_std, float target_scale)
{
    if (!params || !feature_mean || !feature_std)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    for (int f = 0; f < MLP_F; f++)
        if (!(feature_std[f] > 0.0f))
            return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                        "Feature standard deviations must be positive. Got "
                                        "%g at %d",
                                        static_cast<double>(feature_std[f]), f);
    if (!(target_scale >= 0.0f))
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "Target scale must be non-negative");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    GS_TRY(mlp_ensure(c));
    MlpModel &M = mlp_model();
    GS_HIP(hipMemcpyAsync(M.w, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice,
                          c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(M.fmean, feature_mean, sizeof(M.fmean));
    std::memcpy(M.fstd, feature_std, sizeof(M.fstd));
    M.tscale = target_scale;
    M.tscale_base = target_scale;
    M.trained = true;
    return make_success_error_status();
}

ErrorStatus gs_learned_reset(void)
{
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    mlp_model().trained = false;
    return make_success_error_status();
}

ErrorStatus gs_learned_forward(const float *X, int rows, float *Y)
{
    if (!X || !Y) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    if (rows <= 0) return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "rows must be positive");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    void *p;
    GS_TRY(slot(c, SLOT_MLP_X, sizeof(float) * static_cast<size_t>(rows) * (MLP_F + MLP_O), &p));
    float *dX = static_cast<float *>(p), *dY = dX + static_cast<size_t>(rows) * MLP_F;
    GS_HIP(hipMemcpyAsync(dX, X, sizeof(float) * rows * MLP_F, hipMemcpyHostToDevice, c->stream));
    GS_TRY(mlp_forward_rows(c, dX, rows, dY));
    GS_HIP(hipMemcpyAsync(Y, dY, sizeof(float) * rows * MLP_O, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    return make_success_error_status();
}

ErrorStatus gs_learned_gradients(const float *X, const float *T, int rows, float *grad,
                                 double *loss)
{
    if (!X || !T || !grad || !loss) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    if (rows <= 0) return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "rows must be positive");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    GS_TRY(mlp_ensure(c));
    void *p;
    GS_TRY(slot(c, SLOT_MLP_X, sizeof(float) * static_cast<size_t>(rows) * (MLP_F + MLP_O), &p));
    float *dX = static_cast<float *>(p), *dT = dX + static_cast<size_t>(rows) * MLP_F;
    GS_HIP(hipMemcpyAsync(dX, X, sizeof(float) * rows * MLP_F, hipMemcpyHostToDevice, c->stream));
    GS_HIP(hipMemcpyAsync(dT, T, sizeof(float) * rows * MLP_O, hipMemcpyHostToDevice, c->stream));
    MlpModel &M = mlp_model();
    GS_TRY(mlp_gradients(c, dX, dT, rows, M.g, loss));
    GS_HIP(hipMemcpyAsync(grad, M.g, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    return make_success_error_status();
}

ErrorStatus gs_learned_features(const System *system, double softening_length, double theta,
                                int max_num_particles_per_leaf, double *a_coarse,
                                float *features)
{
    GS_TRY(check_system(system));
    if (!a_coarse || !features) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "NULL argument");
    Context *c;
    GS_TRY(context(&c));
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    GS_TRY(mlp_ensure(c));
    const int n = system->num_particles;
    double *x, *m, *a;
    GS_TRY(upload_xm(c, system, &x, &m, &a));
    float *F;
    GS_TRY(coarse_walk(c, x, m, n, system->G, softening_length, theta,
                       leaf_or_default(max_num_particles_per_leaf), a, false, &F, nullptr));
    GS_HIP(hipMemcpyAsync(a_coarse, a, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipMemcpyAsync(features, F, sizeof(float) * MLP_F * static_cast<size_t>(n),
                          hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    return make_success_error_status();
}

ErrorStatus gs_learned_set_policy(double spot_tol, int spot_samples, double fallback_theta)
{
    if (!(spot_tol > 0.0) || spot_samples < 0 || !(fallback_theta >= 0.0))
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR,
                                "spot_tol must be positive, spot_samples and fallback_theta "
                                "non-negative");
    MlpModel &M = mlp_model();
    M.spot_tol = spot_tol;
    M.spot_samples = spot_samples;
    M.fallback_theta = fallback_theta;
    return make_success_error_status();
}

ErrorStatus gs_learned_training_report(double *alpha, double *p99_net, double *p99_without,
                                       double *median_net, double *median_without)
{
    MlpModel &M = mlp_model();
    if (alpha) *alpha = M.hold_alpha;
    if (p99_net) *p99_net = M.hold_p99_net;
    if (p99_without) *p99_without = M.hold_p99_skip;
    if (median_net) *median_net = M.hold_med_net;
    if (median_without) *median_without = M.hold_med_skip;
    return make_success_error_status();
}

ErrorStatus gs_learned_training_times(double *samples_ms, double *prep_ms, double *loop_ms,
                                      double *select_ms)
{
    const MlpModel &M = mlp_model();
    if (M.time_loop_ms < 0.0)
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR, "No training has run");
    if (samples_ms) *samples_ms = M.time_samples_ms;
    if (prep_ms) *prep_ms = M.time_prep_ms;
    if (loop_ms) *loop_ms = M.time_loop_ms;
    if (select_ms) *select_ms = M.time_select_ms;
    return make_success_error_status();
}

ErrorStatus gs_learned_status(double *last_p99, int *fell_back)
{
    MlpModel &M = mlp_model();
    if (last_p99) *last_p99 = M.last_p99;
    if (fell_back) *fell_back = M.last_fell_back;
    return make_success_error_status();
}

}  // extern "C"
