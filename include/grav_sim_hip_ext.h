/*
 * grav_sim_hip_ext.h -- additive entry points of libgrav_sim.so.
 *
 * None of these has a counterpart in the reference library; they extend it
 * without changing any reference struct.  All return ErrorStatus (see
 * grav_sim_hip.h) and follow the reference's ownership rules: the caller owns
 * every buffer it passes; the library owns device memory behind handles.
 */
#ifndef GRAV_SIM_HIP_EXT_H
#define GRAV_SIM_HIP_EXT_H

#include <stddef.h>

#include "grav_sim_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- additional acceleration methods (AccelerationParam.method) ----------
 * IDs >= 16 so the reference IDs 1..3 (src/acceleration.h:16-18) and the
 * struct layout stay untouched.                                              */

/* Direct sum in FP32 on coordinates relative to each target block's centre,
 * tile partial sums combined with Kahan compensation, result in FP64.
 * Same physics as ACCELERATION_METHOD_PAIRWISE (src/acceleration.c:177-234). */
#define ACCELERATION_METHOD_PAIRWISE_FP32 16
/* As above, massive sources only (src/acceleration.c:236-367 physics). */
#define ACCELERATION_METHOD_MASSLESS_FP32 17
/* Barnes-Hut walk at a coarse opening angle plus a learned residual
 * correction; falls back to the group walk with quadrupoles at theta = 0.5
 * when a direct-sum spot check fails (see gs_learned_*). */
#define ACCELERATION_METHOD_BARNES_HUT_LEARNED 18
/* Barnes-Hut with quadrupole moments (same tree and opening criterion as
 * ACCELERATION_METHOD_BARNES_HUT, multipole order 2). */
#define ACCELERATION_METHOD_BARNES_HUT_QUADRUPOLE 19
/* Barnes-Hut group walk: each wavefront walks the tree once for a group of
 * at most 64 particles (children of at most 64 particles of a tree node with
 * more, packed in order; groups wider than 4 w (64/n)^(1/3) in some axis,
 * w the tree's box width, are split into single particles) and accepts a
 * node for all of them
 * when l^2 < theta^2 d^2 with d the distance from the node's centre of mass
 * to the group's bounding box (none of the group inside the node).  Every
 * particle would accept such a node on its own, so the interactions refine
 * those of ACCELERATION_METHOD_BARNES_HUT at the same theta (more
 * interactions, smaller errors); the force sums run as dense FP32 tiles
 * relative to the group centre.  Monopoles (20) or with quadrupoles (21). */
#define ACCELERATION_METHOD_BARNES_HUT_GROUP 20
#define ACCELERATION_METHOD_BARNES_HUT_GROUP_QUADRUPOLE 21
/* Group walk (as 20) with local expansions: a node whose centre of mass lies
 * farther than r_g / phi from the group's box centre c (r_g the box's
 * half-diagonal; phi from gs_set_group_expansion_ratio, 0.1 by default) is
 * far.  Far nodes are accepted at the opening angle k theta (k from
 * gs_set_group_expansion_angle_factor, 1.25 by default; near nodes at theta,
 * as in method 20) and are not summed per particle: the monopole field of
 * each is expanded about c to second order (value, gradient and second
 * derivatives; truncation error of order phi^3 of that node's term) and the
 * value of its quadrupole field at c is added (the tree carries second
 * moments), the expansions of all far nodes are added once per group and
 * evaluated at each particle.  With k = 1 the accepted set, and so the
 * interaction count, is that of method 20 at the same theta. */
#define ACCELERATION_METHOD_BARNES_HUT_GROUP_EXPANSION 22

/* ---- device information --------------------------------------------------- */
typedef struct GsDeviceInfo
{
    char name[256];
    char arch[64];
    int cu_count;
    int clock_khz;
    double fp32_vector_tflops; /* published dense vector peak, packed FP32 */
    double fp64_vector_tflops;
    double hbm_gbs;
} GsDeviceInfo;

ErrorStatus gs_device_info(GsDeviceInfo *info);

/* ---- benchmark_acceleration with a machine-readable record ----------------
 * benchmark_acceleration (src/acceleration.c:369-507) prints the same text
 * and, when records is not NULL, fills records[i] for each parameter set:
 * timings of the host call (device copies included), the MAE against test 0,
 * the directed pair interactions of one evaluation (all ordered pairs for the
 * pairwise sums, massive sources on every other particle for the massless
 * forms, the walk's count for the tree methods), their rate, particles/s,
 * 20 flops per interaction and the fraction of the FP64 (methods 1, 2) or FP32
 * (every other method) vector peak.  Skipped tests keep method and num_times
 * and zeros elsewhere. */
typedef struct GsBenchmarkRecord
{
    int method;
    int num_times;
    double mean_s;
    double std_s;
    double mae;
    double interactions;
    double interactions_per_s;
    double particles_per_s;
    double flops_per_s;
    double peak_flops_per_s;
    double roofline_frac;
} GsBenchmarkRecord;

ErrorStatus gs_benchmark_acceleration(const System *system,
                                      const AccelerationParam *acceleration_params,
                                      int num_acceleration_params,
                                      const int *num_times_acceleration_param,
                                      GsBenchmarkRecord *records);

/* Wait for all work the library has queued. */
ErrorStatus gs_synchronize(void);

/* ---- kernel timing (HIP events on the library's stream) ------------------ */
ErrorStatus gs_timing_enable(int on);
ErrorStatus gs_timing_reset(void);
/* Total device milliseconds and launch count of the named kernel since the
 * last reset (synchronises).  Unknown names give 0 / 0. */
ErrorStatus gs_timing_query(const char *kernel, double *total_ms, long long *calls);

/* ---- device-resident simulation handle ------------------------------------
 * Holds x, v, m (FP64, AoS) and the integrator state in device memory so
 * that force evaluations and Leapfrog steps run without host transfers.     */
typedef struct GsSim GsSim;

ErrorStatus gs_sim_create(GsSim **out, int num_particles, double G);
ErrorStatus gs_sim_destroy(GsSim *sim);
ErrorStatus gs_sim_upload(GsSim *sim, const double *x, const double *v, const double *m);
ErrorStatus gs_sim_download(GsSim *sim, double *x, double *v);
/* a(x) with the given parameters into device storage; copied to a_host
 * (N x 3) if not NULL. */
ErrorStatus gs_sim_acceleration(GsSim *sim, const AccelerationParam *param, double *a_host);
/* Leapfrog pieces with the reference's operation order
 * (src/integrator.c:962-1094): begin = a(x0) + compensated half kick;
 * steps = k x (compensated drift, a(x), compensated kick);
 * end = velocity synchronisation v -= a dt / 2. */
ErrorStatus gs_sim_leapfrog_begin(GsSim *sim, const AccelerationParam *param, double dt);
ErrorStatus gs_sim_leapfrog_steps(GsSim *sim, const AccelerationParam *param, double dt, int64 k);
ErrorStatus gs_sim_leapfrog_end(GsSim *sim, double dt);
/* Total energy of the device state (src/utils.c:27-59 formula). */
ErrorStatus gs_sim_energy(GsSim *sim, double *energy);
/* Relative acceleration error of `param` against the FP64 direct sum on the
 * GPU, over `num_samples` targets (evenly strided; all if <= 0 or >= N).
 * Writes median, 99th percentile and maximum of |a - a64| / |a64|. */
ErrorStatus gs_sim_accel_error(GsSim *sim, const AccelerationParam *param, int num_samples,
                               double *median, double *p99, double *max);

/* ---- batched WHFast ------------------------------------------------------------
 * Integrates num_systems independent systems with WHFast
 * (src/integrator_whfast.c:200-407 semantics, one GPU workgroup per system).
 * System s holds particles offsets[s] .. offsets[s+1]-1 of the flat arrays
 * (x, v: N x 3; m, ids: N), with ids local to the system (the particle with
 * id 0 is the primary the reference sorts distances from).  On return each
 * system's survivors are stored, in the reference's final order, from
 * offsets[s] onwards and out_counts[s] gives their number.  method is
 * ACCELERATION_METHOD_PAIRWISE or _MASSLESS; up to 2048 particles and 64
 * massive particles per system. */
ErrorStatus gs_whfast_ensemble(int num_systems, const int64 *offsets, int32 *ids, double *x,
                               double *v, double *m, double G, int method, double eps, double dt,
                               double tf, bool remove_invalid, int *out_counts);

/* ---- octree pieces ----------------------------------------------------------
 * Morton keys (sorted) and the stable sort permutation of system->x, computed
 * on the GPU (src/linear_octree.c:113-326).  Arrays are N long. */
ErrorStatus gs_octree_sort(const System *system, int64 *sorted_keys, int *sorted_idx,
                           double *center, double *width);

/* Number of interactions (accepted nodes + leaf particle pairs) of the last
 * Barnes-Hut walk. */
ErrorStatus gs_last_walk_interactions(long long *count);

/* Number of (group, node) pairs the last group walk with local expansions
 * (ACCELERATION_METHOD_BARNES_HUT_GROUP_EXPANSION) took through the
 * expansion; 0 after any other method.  No counterpart in the reference. */
ErrorStatus gs_last_walk_expansion_pairs(long long *count);

/* Expansion ratio phi of ACCELERATION_METHOD_BARNES_HUT_GROUP_EXPANSION
 * (0 < phi < 1; default 0.1).  Process-wide.  No counterpart in the
 * reference. */
ErrorStatus gs_set_group_expansion_ratio(double phi);
ErrorStatus gs_get_group_expansion_ratio(double *phi);

/* Opening-angle factor k of the far nodes of
 * ACCELERATION_METHOD_BARNES_HUT_GROUP_EXPANSION (1 <= k <= 4; default 1.25).
 * Process-wide.  No counterpart in the reference. */
ErrorStatus gs_set_group_expansion_angle_factor(double k);
ErrorStatus gs_get_group_expansion_angle_factor(double *k);

/* Streams `bytes` through a plain device copy kernel `reps` times and writes
 * the achieved bandwidth (read + write bytes / time).  Its kernel ("copy")
 * is the reference point for the HBM byte counters. */
ErrorStatus gs_copy_calibrate(size_t bytes, int reps, double *gbs);

/* Cycles per wave64 instruction (per SIMD, at the device's nominal clock)
 * of independent chains of v_fmac_f32, v_mul_f32 and v_rsq_f32 over the
 * whole GPU: the measured issue rates behind the FP32 kernels' instruction
 * mix ceilings.  Any pointer may be NULL.  No counterpart in the reference. */
ErrorStatus gs_issue_rates(double *fma_cycles, double *mul_cycles, double *rsq_cycles);

/* ---- learned Barnes-Hut correction (ACCELERATION_METHOD_BARNES_HUT_LEARNED) --
 * A monopole walk at the caller's opening angle (0.8 is the intended coarse
 * value) accumulates per-particle features of the accepted nodes (quadrupole
 * and octupole force sums G q and G o from the nodes' second and third
 * moments, size-weighted sums such as s3 = sum M l^3 / d^5, node counts).
 * The residual a_direct - a_coarse is predicted as
 * G (q + o) + G s3 * tscale * net(F): the accepted nodes' quadrupole and
 * octupole terms plus, in units of the next multipole orders' scale, what an
 * MLP
 *   16 -> 128 -> 128 -> 128 -> 3, tanh-GELU, BF16 matrix-core products with
 *   FP32 accumulation, FP32 master weights
 * learns from the normalised features F; the prediction is added to the
 * coarse result.  Every evaluation then spot-checks
 * spot_samples evenly strided particles against the FP32 split-precision
 * direct sum; if their 99th-percentile relative error exceeds spot_tol the
 * result is replaced by the group walk with quadrupole terms at
 * fallback_theta, which is checked again on the same particles.
 * Parameters: weights W1 (16x128), b1, W2 (128x128), b2, W3 (128x128), b3,
 * W4 (128x3), b4, each matrix row-major as (inputs, outputs), in one flat
 * FP32 array of gs_learned_num_params() values.                              */
#define GS_LEARNED_NUM_FEATURES 16
#define GS_LEARNED_NUM_OUTPUTS 3
int gs_learned_num_params(void);

/* Trains the network on one particle set: coarse walk at theta_coarse,
 * FP32 direct-sum labels for `samples` particles drawn without replacement
 * with `seed`, feature normalisation and target scale from that sample
 * (targets: the residual left after G (q + o) divided by G s3, each
 * sample's norm clipped at the sample's 99.9th percentile, scaled to unit
 * RMS), then `epochs` epochs of Adam (batches of `batch` rows, learning rate
 * lr with cosine decay) from a deterministic initialisation on nine tenths
 * of the sample.  On the held-out tenth the output scale alpha in
 * {0, 1/8, ..., 1} that gives the smallest 99th-percentile relative error is
 * chosen (alpha 0 leaves the quadrupole and octupole sums alone) and folded
 * into the target scale.  final_loss (may be NULL): mean squared error of the last epoch's last
 * batch. */
ErrorStatus gs_learned_train(const System *system, double softening_length, double theta_coarse,
                             int max_num_particles_per_leaf, int samples, int epochs, int batch,
                             double lr, unsigned int seed, double *final_loss);
/* Held-out statistics of the last gs_learned_train: the chosen output scale
 * and the 99th-percentile and median relative errors with the network at that
 * scale and without it.  Any pointer may be NULL. */
ErrorStatus gs_learned_training_report(double *alpha, double *p99_net, double *p99_without,
                                       double *median_net, double *median_without);
/* Wall-clock milliseconds of the last gs_learned_train: the samples (coarse
 * walk, direct-sum labels, copies to the host), the host's preparation of the
 * training set (targets, normalisation, upload), the training loop on the GPU
 * and the held-out choice of the output scale.  Any pointer may be NULL;
 * GRAV_VALUE_ERROR before the first training. */
ErrorStatus gs_learned_training_times(double *samples_ms, double *prep_ms, double *loop_ms,
                                      double *select_ms);
/* Chooses the output scale again, on `samples` particles of another particle
 * set (a validation realisation): the same candidates and rule as in
 * gs_learned_train, the trained weights unchanged; the report then holds
 * these samples' statistics. */
ErrorStatus gs_learned_calibrate(const System *system, double softening_length, double theta_coarse,
                                 int max_num_particles_per_leaf, int samples, unsigned int seed);
/* Copies out / installs the parameters, the feature normalisation (16 means
 * and 16 standard deviations) and the target scale.  set marks the network
 * trained; reset marks it untrained (method 18 then fails with a value
 * error). */
ErrorStatus gs_learned_get_params(float *params, float *feature_mean, float *feature_std,
                                  float *target_scale);
ErrorStatus gs_learned_set_params(const float *params, const float *feature_mean,
                                  const float *feature_std, float target_scale);
ErrorStatus gs_learned_reset(void);
/* Network output Y (rows x 3) for normalised features X (rows x 16). */
ErrorStatus gs_learned_forward(const float *X, int rows, float *Y);
/* Gradient of mean over rows of |net(X) - T|^2 with respect to the
 * parameters (flat layout above) and that loss. */
ErrorStatus gs_learned_gradients(const float *X, const float *T, int rows, float *grad,
                                 double *loss);
/* Coarse walk accelerations (N x 3) and raw (unnormalised) features (N x 16)
 * of a particle set. */
ErrorStatus gs_learned_features(const System *system, double softening_length, double theta,
                                int max_num_particles_per_leaf, double *a_coarse,
                                float *features);
/* Spot-check policy (defaults 1e-3, 1024, 0.5; spot_samples 0 disables the
 * check) and the outcome of the last evaluation: last_p99 the spot check's
 * 99th-percentile error of the result returned (of the fallback when there
 * was one); fell_back 0 none, 1 fallback within spot_tol, 2 fallback still
 * above spot_tol. */
ErrorStatus gs_learned_set_policy(double spot_tol, int spot_samples, double fallback_theta);
ErrorStatus gs_learned_status(double *last_p99, int *fell_back);

/* ---- FP32 direct-sum autotuner ---------------------------------------------
 * Times the candidate (block, targets-per-thread, source-split) shapes on an
 * N-particle input resident on the device and keeps the fastest for later
 * calls -- among the shapes within 1 % of the fastest, the one with the
 * smallest scratch (partial sums) -- and writes the chosen shape.  block = 0 denotes the scalar-source
 * kernel (independent wavefronts reading pre-converted source tiles through
 * wave-uniform loads); block = 4, 8 or 16 the shared-target scalar-source
 * kernel (that many wavefronts per 64 targets, each over its share of the
 * source chunk, their sums combined in LDS in a fixed order;
 * targets_per_thread 1 only);
 * block = 128/256/512 the LDS-tile kernel.  A source split of 0 picks the
 * split automatically. */
ErrorStatus gs_autotune_fp32(int num_particles, int *block, int *targets_per_thread,
                             int *source_split);
ErrorStatus gs_set_fp32_config(int block, int targets_per_thread, int source_split);
/* The shape later FP32 direct sums use (default: block 16, 1, automatic split). */
ErrorStatus gs_get_fp32_config(int *block, int *targets_per_thread, int *source_split);

/* ---- direct sums without device allocation ---------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  acceleration() for methods 1, 2, 16 and 17 with
 * system->x, system->m and a in device memory of the library's device; the
 * result is complete when the call returns.  It calls no hipMalloc or
 * hipFree: the FP32 sums' scratch is a static device array of 32 MiB that
 * the GPU holds until the process ends, enough for about 800 000 particles
 * (larger N: GRAV_VALUE_ERROR; past the default shape's source-chunk count
 * the chunks are fewer, which changes the FP32 rounding, not the bound).
 * Calls that overlap on one GPU share that array and race. */
ErrorStatus gs_acceleration_noalloc(double *a, const System *system,
                                    const AccelerationParam *acceleration_param);

#ifdef __cplusplus
}
#endif

#endif /* GRAV_SIM_HIP_EXT_H */
