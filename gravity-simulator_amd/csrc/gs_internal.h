// gs_internal.h -- internal declarations shared by the library's host and
// device sources.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "grav_sim_hip.h"
#include "grav_sim_hip_ext.h"

namespace gs {

// ---------------------------------------------------------------- errors
// Turn a failing HIP call into the library's ErrorStatus convention:
// allocation failures -> GRAV_MEMORY_ERROR, everything else -> GRAV_FAILURE.
ErrorStatus hip_error(hipError_t e, const char *what, const char *file, int line,
                      const char *func);

#define GS_HIP(call)                                                                \
    do {                                                                            \
        hipError_t gs_e_ = (call);                                                  \
        if (gs_e_ != hipSuccess)                                                    \
            return ::gs::hip_error(gs_e_, #call, __FILE__, __LINE__, __func__);     \
    } while (0)

#define GS_TRY(call)                                                                \
    do {                                                                            \
        ErrorStatus gs_s_ = WRAP_TRACEBACK(call);                                   \
        if (gs_s_.return_code != GRAV_SUCCESS)                                      \
            return gs_s_;                                                           \
    } while (0)

// ---------------------------------------------------------------- device context
// One context per process (device 0 unless GS_DEVICE is set): a non-blocking
// stream at default priority, grow-only device buffers addressed by slot, and
// the kernel-timing registry.  Calls that use the context hold its mutex.
struct Buffer {
    void *ptr = nullptr;
    size_t bytes = 0;
};

enum Slot : int {
    // host-entry staging
    SLOT_X = 0, SLOT_V, SLOT_M, SLOT_A, SLOT_IDS,
    // direct-sum partials
    SLOT_PARTIAL,
    // octree
    SLOT_BBOX, SLOT_KEYS, SLOT_KEYS_ALT, SLOT_IDX, SLOT_IDX_ALT, SLOT_HIST, SLOT_HIST_SCAN,
    SLOT_NODES, SLOT_NODE_TMP, SLOT_LEVELS, SLOT_PACKED_P, SLOT_SORTED_P, SLOT_MISC,
    SLOT_GROUP_PARENTS, SLOT_GROUP_FLAGS,
    // learned correction
    SLOT_FEAT, SLOT_MLP_W, SLOT_MLP_ACT, SLOT_MLP_GRAD, SLOT_MLP_X, SLOT_SPOT, SLOT_MLP_TS, SLOT_MLP_IMG,
    SLOT_COUNT
};

struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    int clock_khz = 0;
    char arch[64] = {0};
    char name[256] = {0};
    std::recursive_mutex mu;
    Buffer slots[SLOT_COUNT];
    // pinned host staging for small readbacks
    void *pinned = nullptr;
    size_t pinned_bytes = 0;

    // timing
    bool timing = false;
    struct Rec {
        int kernel;
        hipEvent_t start, stop;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::string> kernel_names;
    std::vector<double> kernel_ms;
    std::vector<long long> kernel_calls;
};

// Returns the initialised context or an error (no HIP device, init failure).
ErrorStatus context(Context **out);
// Looks up, once per device with that device current, the device address of
// the no-allocation entry point's static scratch array.
hipError_t noalloc_scratch_init(int device);

// Grow-only device buffer for a slot.
ErrorStatus slot(Context *c, Slot s, size_t bytes, void **out);
ErrorStatus pinned(Context *c, size_t bytes, void **out);

// Kernel timing: wrap a launch as
//   int t = timing_begin(c, "name");  launch...;  timing_end(c, t);
int timing_begin(Context *c, const char *kernel);
void timing_end(Context *c, int rec);
ErrorStatus timing_flush(Context *c);

// ---------------------------------------------------------------- kernels (launchers)
// Every launcher enqueues on c->stream and returns hipSuccess or the launch error.

// FP64 direct sum, reference summation order (bitwise).  skip_massless: sources
// with m == 0 are skipped (massless method).
hipError_t launch_direct_f64(Context *c, const double *x, const double *m, int n, double G,
                             double eps, bool skip_massless, double *a);

// FP32 tile-relative direct sum with tile-level Kahan compensation.
struct Fp32Config {
    int kind = 2;        // 0: LDS source tiles, 1: scalar-loaded source tiles,
                         // 2: scalar-loaded source tiles, shared target groups
    int block = 16;      // threads per workgroup (kind 0), wavefronts per target group (kind 2)
    int ipt = 1;         // targets per thread
    int jsplit = 0;      // source chunks (0 = automatic)
};
hipError_t launch_direct_f32(Context *c, const double *x, const double *m, int n, double G,
                             double eps, double *a, void *partial_buf, size_t partial_bytes,
                             const Fp32Config &cfg);
size_t direct_f32_partial_bytes(int n, const Fp32Config &cfg, int cu_count);
// Same arithmetic for the listed targets only: out[t] = a(targets[t]).
hipError_t launch_direct_f32_targets(Context *c, const double *x, const double *m, int n, double G,
                                     double eps, const int *targets, int nt, double *out,
                                     void *buf, size_t bytes);
size_t direct_f32_targets_bytes(int n, int nt);
// out[t] = src[idx[t]] for rows of 3 doubles.
hipError_t launch_gather_rows3(Context *c, const double *src, const int *idx, int nt, double *out);
// FP32 issue-rate probe: kind 0 v_fmac_f32, 1 v_mul_f32, 2 v_rsq_f32 chains,
// `blocks` workgroups of 256 threads, one float each into out; and the
// number of wave64 instructions of that form one launch issues.
hipError_t launch_issue_probe(Context *c, int kind, float *out, int blocks);
double issue_probe_instructions(int blocks);
// FP64 sums for the listed targets split over source chunks (the sampled
// error reference: same terms, chunked summation order).
hipError_t launch_direct_f64_sampled(Context *c, const double *x, const double *m, int n, double G,
                                     double eps, bool skip_massless, const int *targets, int nt,
                                     double *out, void *buf, size_t bytes);
size_t direct_f64_sampled_bytes(int n, int nt);
// FP64 reference-order sums for the listed targets.
hipError_t launch_direct_f64_targets(Context *c, const double *x, const double *m, int n,
                                     double G, double eps, bool skip_massless, const int *targets,
                                     int num_targets, double *out);
Fp32Config &fp32_config();

// FP64 energy (src/utils.c:27-59 formula; reduction order differs).
hipError_t launch_energy(Context *c, const double *x, const double *v, const double *m, int n,
                         double G, double *partials, int *num_partials);

// One Leapfrog step with the FP64 direct sum in one launch: drift (x, xe, v)
// into (xo, xeo), forces of the drifted positions into a, kick v into vo with
// ve updated in place -- the arithmetic of launch_lf_drift, launch_direct_f64
// and launch_lf_kick in that order.  The output buffers must not alias the
// inputs.
hipError_t launch_lf_step_f64(Context *c, const double *x, const double *xe, const double *v,
                              const double *m, int n, double G, double eps, bool skip_massless,
                              double dt, double *a, double *xo, double *xeo, double *vo, double *ve);

// Leapfrog updates (src/integrator.c:973-1094 op order).
hipError_t launch_lf_half_kick(Context *c, double *v, double *ve, const double *a, int n,
                               double dt);
hipError_t launch_lf_drift(Context *c, double *x, double *xe, const double *v, int n, double dt);
hipError_t launch_lf_kick(Context *c, double *v, double *ve, const double *a, int n, double dt);
hipError_t launch_lf_sync(Context *c, double *v, const double *a, int n, double dt);

// ---------------------------------------------------------------- octree
struct DeviceTree;  // defined in octree.hip

// Build the tree of (x, m) (device arrays, AoS) into the context's buffers.
// box_center/box_width: NULL/<=0 -> bounding box.  with_quad: also the second
// moments, which the quadrupole and feature walks need.
ErrorStatus tree_build(Context *c, const double *x, const double *m, int n, int leaf_cap,
                       const double *box_center, double box_width, bool with_quad,
                       DeviceTree **out, bool with_oct = false);
// Walk; writes a (device, AoS, original order).  order 1: monopoles (the
// reference's), 2: with quadrupole terms.  feat != NULL: monopole walk that
// also writes 16 FP32 features per particle for the learned correction (the
// tree built with second and third moments).
ErrorStatus tree_walk(Context *c, DeviceTree *t, const double *x, double G, double eps,
                      double theta, double *a, long long *interactions, int order = 1,
                      float *feat = nullptr);
// Group walk (one wavefront per 64 sorted particles, group opening test);
// order 1: monopoles, 2: with quadrupole terms (tree built with_quad).
// phi > 0 (monopole walk only): nodes with rg < phi |com - c| (rg the
// half-diagonal of the group's box, c its centre) are far; they are accepted
// at the opening angle theta_far (>= theta) and enter a second-order local
// expansion of their monopole fields about c, plus the value of their
// quadrupole fields at c (the tree must carry second moments), instead of the
// per-particle node terms; *far_pairs gets the number of such (group, node)
// pairs.
ErrorStatus tree_walk_group(Context *c, DeviceTree *t, double G, double eps, double theta,
                            double *a, long long *interactions, int order, double phi = 0.0,
                            double theta_far = 0.0, long long *far_pairs = nullptr);
// Host copy of the tree in the reference struct layout.
ErrorStatus tree_to_host(Context *c, DeviceTree *t, LinearOctree *out);

// Device-resident acceleration of (x, m) into a (all device pointers).
ErrorStatus device_acceleration(Context *c, const double *x, const double *m, int n, double G,
                                const AccelerationParam *p, double *a);

// ---------------------------------------------------------------- snapshots
// CSV snapshots of a device-resident run (file name, metadata, header and row
// formats of src/output.c:351-469; output times of src/integrator.c:400-415).
// A snapshot is opened by the caller's thread (so an unwritable directory
// fails at the step that asks for it), copied device -> pinned host on the
// run's stream behind an event, and formatted by a pool of writer threads
// while the run continues.  Particle ids and masses are taken once at
// begin() for the runs that keep them fixed (write_device), or copied with
// every snapshot (write_device_particles).
class SnapshotWriter {
public:
    SnapshotWriter() = default;
    ~SnapshotWriter();
    SnapshotWriter(const SnapshotWriter &) = delete;
    SnapshotWriter &operator=(const SnapshotWriter &) = delete;
    // op == NULL or OUTPUT_METHOD_DISABLED: off, every call below is a no-op.
    ErrorStatus begin(Context *c, OutputParam *op, const System *sys);
    bool on() const { return impl_ != nullptr; }
    // the reference's test `t >= next_output_time`
    bool due(double t) const;
    // state in host memory (the initial snapshot)
    ErrorStatus write_host(const double *x, const double *v, double t, double dt);
    // state in device memory, copied on c->stream behind the work queued so far
    ErrorStatus write_device(const double *x, const double *v, double t, double dt);
    // the same for a run that reorders and removes particles: masses, ids and
    // the particle count (*count <= the count at begin()) come from the device
    ErrorStatus write_device_particles(const double *x, const double *v, const double *m,
                                       const int32 *ids, const int *count, double t, double dt);
    // device buffer of 3 n doubles for a derived velocity (allocated on first use)
    ErrorStatus scratch(double **out);
    // waits for every queued snapshot; the first writer error, if any
    ErrorStatus finish();

private:
    struct Impl;
    Impl *impl_ = nullptr;
};

// One CSV snapshot written synchronously from host arrays (the body of
// output_snapshot_csv, src/output.c:351-469).
ErrorStatus write_csv_snapshot(const char *path, int n, double G, double t, double dt,
                               const int32 *ids, const double *m, const double *x,
                               const double *v, const OutputParam *op);

}  // namespace gs
