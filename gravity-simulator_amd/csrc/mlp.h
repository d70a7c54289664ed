// mlp.h -- the residual network of the learned Barnes-Hut correction
// (internal; the C-ABI entry points are gs_learned_* in grav_sim_hip_ext.h).
//
// 16 -> 128 -> 128 -> 128 -> 3, tanh-GELU after each hidden layer.  FP32
// master weights in one flat array, each weight matrix row-major as
// (inputs, outputs): W[k * outputs + n].
#pragma once

#include "gs_internal.h"

namespace gs {

constexpr int MLP_F = 16;   // input features
constexpr int MLP_H = 128;  // hidden width
constexpr int MLP_O = 3;    // outputs (residual / (|a| * tscale))
constexpr int MLP_WALK_FEAT = 16;  // walk accumulators per particle

constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_F * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H * MLP_H;
constexpr int MLP_OFF_W4 = MLP_OFF_B3 + MLP_H;
constexpr int MLP_OFF_B4 = MLP_OFF_W4 + MLP_H * MLP_O;
constexpr int MLP_NPARAM = MLP_OFF_B4 + MLP_O;

// ---------------------------------------------------------------- local frame
// The network sees only rotation-invariant quantities and predicts the
// residual's components in a per-particle orthonormal frame, so that it
// cannot tie its output to absolute directions (positions) of one particle
// realisation:
//   e1 = a / |a| (coarse acceleration),
//   e2 = the part of q orthogonal to e1, normalised (else that of v, else
//        that of the coordinate axis least aligned with e1),
//   e3 = e1 x e2,
// with q the walk's quadrupole sum and v = sum (M l^3 / d^6) R.  FP64, the
// same operations on host and device.
__host__ __device__ inline void mlp_local_frame(const double a[3], const double q[3],
                                                const double v[3], double e1[3], double e2[3],
                                                double e3[3])
{
    const double an = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (an > 0.0) {
        e1[0] = a[0] / an;
        e1[1] = a[1] / an;
        e1[2] = a[2] / an;
    } else {
        e1[0] = 1.0;
        e1[1] = 0.0;
        e1[2] = 0.0;
    }
    double u[3];
    double un = 0.0;
    const double *cand[2] = {q, v};
    for (int c = 0; c < 2 && !(un > 0.0); c++) {
        const double *w = cand[c];
        const double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        const double we = w[0] * e1[0] + w[1] * e1[1] + w[2] * e1[2];
        for (int d = 0; d < 3; d++) u[d] = w[d] - we * e1[d];
        un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        if (!(un > 1e-6 * wn)) un = 0.0;
    }
    if (!(un > 0.0)) {
        int k = 0;
        for (int d = 1; d < 3; d++)
            if (fabs(e1[d]) < fabs(e1[k])) k = d;
        for (int d = 0; d < 3; d++) u[d] = (d == k ? 1.0 : 0.0) - e1[k] * e1[d];
        un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    }
    for (int d = 0; d < 3; d++) e2[d] = u[d] / un;
    e3[0] = e1[1] * e2[2] - e1[2] * e2[1];
    e3[1] = e1[2] * e2[0] - e1[0] * e2[2];
    e3[2] = e1[0] * e2[1] - e1[1] * e2[0];
}

struct MlpModel {
    float *w = nullptr;  // device: master weights
    float *g = nullptr;  // device: gradients
    float *m = nullptr;  // device: Adam first moments
    float *v = nullptr;  // device: Adam second moments
    float fmean[MLP_F];  // feature normalisation (host copy)
    float fstd[MLP_F];
    float tscale = 1.0f;  // scale of the targets (residual / (G s3)) times the output scale
    float tscale_base = 1.0f;  // the training's target scale (output scale 1)
    bool trained = false;
    // held-out evaluation of the last training run (relative errors against
    // the direct sum): output scale chosen, p99 / median with the network at
    // that scale and without it (quadrupole sum only)
    double hold_alpha = -1.0, hold_p99_net = -1.0, hold_p99_skip = -1.0;
    double hold_med_net = -1.0, hold_med_skip = -1.0;
    // wall-clock milliseconds of the last training: samples (coarse walk,
    // direct-sum labels, copies), host preparation of the training set, the
    // training loop, the held-out choice of the output scale
    double time_samples_ms = -1.0, time_prep_ms = -1.0, time_loop_ms = -1.0, time_select_ms = -1.0;
    // runtime policy of method BARNES_HUT_LEARNED
    double spot_tol = 1e-3;      // p99 relative error that triggers the fallback
    int spot_samples = 1024;     // particles checked against the FP32 direct sum
    double fallback_theta = 0.5;  // group walk with quadrupoles used on fallback
    // status of the last evaluation: p99 of the returned result and of the
    // coarse corrected walk; 0 no fallback, 1 fell back and the fallback met
    // the tolerance, 2 fell back and the fallback missed it too
    double last_p99 = -1.0;
    double last_p99_coarse = -1.0;
    int last_fell_back = 0;
};

MlpModel &mlp_model();

// Raw (normalise = false) or normalised features of n particles from the
// coarse acceleration a (device, AoS) and the walk accumulators walkfeat
// (device, MLP_WALK_FEAT floats per particle).  F: n x MLP_F, device.
ErrorStatus mlp_features(Context *c, const double *a, const float *walkfeat, int n, double G,
                         bool normalise, float *F);
// Y = net(X) for rows x MLP_F normalised features (device arrays).
ErrorStatus mlp_forward_rows(Context *c, const float *X, int rows, float *Y);
// Gradient of the mean over rows of |net(X) - T|^2 (device arrays).
ErrorStatus mlp_gradients(Context *c, const float *X, const float *T, int rows, float *grad,
                          double *loss);
// Adam training from a deterministic initialisation (device arrays).
ErrorStatus mlp_train(Context *c, const float *X, const float *T, int S, int epochs, int batch,
                      double lr, unsigned seed, double *final_loss);
// a += G q + G s3 * tscale * net(F) for n particles (device arrays), q the
// quadrupole sum and s3 = sum M l^3 / d^5 held in the walk accumulators wf.
ErrorStatus mlp_apply(Context *c, const float *F, const float *wf, double G, int n, double *a);
ErrorStatus mlp_ensure(Context *c);
// out[r] = X[idx[r]] for rows of F floats (device arrays).
hipError_t mlp_gather_rows(Context *c, const float *X, const int *idx, int rows, int F, float *out);

}  // namespace gs
