/*
 * oracle.h -- CPU restatement of grav_sim's force-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the GPU
 * library: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it.  Nothing in the product links or calls it.
 *
 * Every function names the reference code (paths relative to the grav_sim
 * source tree) whose behaviour it restates.  Results are FP64 and, where the
 * reference fixes an operation order, follow that order exactly (built with
 * -ffp-contract=off), so they are bitwise comparable with the reference
 * build and with the GPU's FP64 kernels.
 *
 * Parity status of each function (how it is pinned) is listed in the
 * docstring of oracle/oracle.py, which wraps this library for the tests.
 */
#ifndef GS_ORACLE_H
#define GS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Direct pairwise accelerations (src/acceleration.c:177-234), restated per
 * target in the summation order the reference's i<j loop produces for every
 * particle: sources l = 0..N-1 ascending, l != k, each term
 * fl(fl(G / fl(fl(r*r)*r) * fl(x_l - x_k)) * m_l). */
void orc_pairwise(double *a, int n, const double *x, const double *m, double G, double eps);

/* Massive + massless accelerations (src/acceleration.c:236-367) with the
 * physically correct source mass m[idx_i] (the reference reads m[i], the
 * ordinal among massive particles; both agree when massive come first).
 * Targets: every particle; sources: massive particles in index order. */
void orc_massless(double *a, int n, const double *x, const double *m, double G, double eps);

/* Bounding box (src/linear_octree.c:113-146). */
void orc_bbox(double center[3], double *width, int n, const double *x);

/* 63-bit Morton keys (src/linear_octree.c:159-202): same FP64 quantisation,
 * bits interleaved one at a time instead of by magic numbers. */
void orc_morton(int64_t *keys, int n, const double *x, const double center[3], double width);

/* Stable sort of keys with the identity payload (src/linear_octree.c:216-326,
 * :913-916).  Merge sort: any stable sort yields the reference permutation. */
void orc_sort(int64_t *keys, int32_t *idx, int n);

/* Octree node set (src/linear_octree.c:342-823).  Nodes in level order
 * (depth, prefix ascending): node 0 is the root.  A node at depth d >= 1 exists
 * for every non-empty octant of an internal parent; a node is internal iff it
 * is the root, or holds more than leaf_cap particles and d < 21.
 * Returns the node count, or -1 if max_nodes is too small.
 * Multipoles: leaf = sequential sums over its particles in sorted order;
 * internal = sequential sums over its children; com = (sum m x) / M. */
int orc_octree(int max_nodes, int n, const double *x, const double *m,
               const int64_t *sorted_keys, const int32_t *sorted_idx, int leaf_cap,
               int *depth, int64_t *prefix, int *count, int *first, int *nchild, int *child0,
               double *mass, double *mx, double *my, double *mz,
               double *comx, double *comy, double *comz);

/* Barnes-Hut walk (src/acceleration_barnes_hut.c:78-248) with the reference's
 * quirks fixed: leaf nodes carry their own mass and centre of mass, and the
 * "target inside node" test compares the target's own key with the node's
 * prefix.  Per target, depth-first over children in octant order:
 *   included  = (key_i >> 3(21-d)) == prefix
 *   accept    = !included && (w/2^d)^2 < theta^2 * |x_i - com|^2
 *   accept    -> a -= G M R / (|R|^2 + eps^2)^{3/2}
 *   leaf      -> direct sum over its particles, skipping i
 *   otherwise -> descend.
 * Writes a[idx*3..] for every particle and, if n_int != NULL, the number of
 * accepted nodes + leaf-particle interactions per particle.
 * With quad / oct != NULL the quadrupole / octupole terms of accepted nodes
 * are added.  With feat != NULL (and quad given) the walk stays
 * monopole-only and writes 16 features per particle (original order) for the
 * learned correction: quadrupole-acceleration sum (3), sum (M l^3 / r^6) R
 * (3), sum M l^2 / r^4, sum M l^3 / r^5, accepted nodes, particle
 * interactions, octupole-acceleration sum (3, zero without oct), 0, 0, 0 --
 * without G; l is the node side, R = x_i - com, r^2 = |R|^2 + eps^2; a leaf
 * holding one particle counts as a particle interaction. */
void orc_bh_walk(double *a, int64_t *n_int, int n, const double *x, const double *m, double G,
                 double eps, double theta, double width, const int64_t *sorted_keys,
                 const int32_t *sorted_idx, const int *depth, const int64_t *prefix,
                 const int *count, const int *first, const int *nchild, const int *child0,
                 const double *mass, const double *comx, const double *comy, const double *comz,
                 const double *quad, const double *oct, double *feat);

/* The walk of groups of at most `group` consecutive sorted particles (the
 * children of at most `group` particles of each node with more, packed in
 * order; a run wider than 4 width (64/n)^(1/3) in some axis is split into
 * single particles): a node is accepted for the whole group when it holds none of the
 * group's particles and (l)^2 < theta^2 d^2 with d the distance from its
 * centre of mass to the group's bounding box; group = 1 is orc_bh_walk
 * (without features). */
void orc_bh_walk_group(double *a, int64_t *n_int, int n, const double *x, const double *m,
                       double G, double eps, double theta, double width,
                       const int64_t *sorted_keys, const int32_t *sorted_idx, const int *depth,
                       const int64_t *prefix, const int *count, const int *first,
                       const int *nchild, const int *child0, const double *mass,
                       const double *comx, const double *comy, const double *comz,
                       const double *quad, int group);

/* orc_bh_walk_group with local expansions (phi > 0; near nodes monopoles
 * only, quad must be NULL): a node whose centre of mass lies farther than
 * rg / phi from the group's box centre c (rg the box's half-diagonal) is far;
 * it is accepted when its side is below theta or, being far, below theta_far
 * (used when larger than theta) times the distance to the group's box, and
 * is not summed per particle: its monopole field is expanded about c to
 * second order (orc_far_add) and, with quad_far (num_nodes x 6 second
 * moments, may be NULL), the value of its quadrupole field at c is added to
 * the expansion's constant term; the sum of the expansions is evaluated at
 * each particle (orc_far_eval).  Such a node counts as one interaction of
 * each particle of the group; *far_pairs (may be NULL) gets the number of
 * (group, node) pairs expanded. */
void orc_bh_walk_group_le(double *a, int64_t *n_int, int n, const double *x, const double *m,
                          double G, double eps, double theta, double width,
                          const int64_t *sorted_keys, const int32_t *sorted_idx, const int *depth,
                          const int64_t *prefix, const int *count, const int *first,
                          const int *nchild, const int *child0, const double *mass,
                          const double *comx, const double *comy, const double *comz,
                          const double *quad, int group, double phi, double theta_far,
                          const double *quad_far, int64_t *far_pairs);

/* Moments of one source for the second-order local expansion (23 sums, see
 * oracle.c) and the expansion's value at offset d from the centre. */
void orc_far_add(double *F, double y0, double y1, double y2, double M, double eps2);
void orc_far_eval(const double *F, double d0, double d1, double d2, double out[3]);
/* Second central moments per node, Q_ab = sum m (x - com)_a (x - com)_b, in
 * the order xx yy zz xy xz yz: leaves over their particles, internal nodes
 * from their children by the parallel-axis shift.  An extension (the
 * reference keeps monopoles only); with quad != NULL, orc_bh_walk adds the
 * quadrupole term of every accepted node. */
void orc_octree_quad(int num, const double *x, const double *m, const int32_t *sorted_idx,
                     const int *count, const int *first, const int *nchild, const int *child0,
                     const double *mass, const double *comx, const double *comy,
                     const double *comz, double *quad);

/* Third central moments per node, T_abc = sum m s_a s_b s_c (s = x - com),
 * in the order xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz: leaves over their
 * particles, internal nodes from their children (parallel-axis shift with
 * the children's second moments, quad from orc_octree_quad). */
void orc_octree_oct(int num, const double *x, const double *m, const int32_t *sorted_idx,
                    const int *count, const int *first, const int *nchild, const int *child0,
                    const double *mass, const double *comx, const double *comy,
                    const double *comz, const double *quad, double *oct);

/* Leapfrog KDK with compensated summation (src/integrator.c:894-1121), using
 * orc_pairwise (method 1) or orc_massless (method 2) for the accelerations.
 * Runs ceil(tf/dt) steps with the reference's last-step clamp and final
 * velocity synchronisation.  x, v updated in place. Returns steps taken. */
int64_t orc_leapfrog(double *x, double *v, int n, const double *m, double G, int method,
                     double eps, double dt, double tf);

/* Total energy (src/utils.c:27-59): KE via vec_norm_3d squared, PE over i<j
 * without softening, one running sum in the reference order. */
double orc_energy(int n, const double *x, const double *v, const double *m, double G);

/* ---- WHFast (src/integrator_whfast.c) ---- */

/* Stumpff functions c0..c3 (src/integrator_whfast.c:774-815). */
void orc_stumpff(double z, double c[4]);

/* Full WHFast run of one system (src/integrator_whfast.c:200-407):
 * initial sort by distance, per-step sort-and-permute of the Jacobi arrays
 * (stable order for ties), Kepler drift with removal of invalid particles in
 * ascending index order, Jacobi<->Cartesian, massless-form accelerations with
 * the auxiliary sums zero-initialised.  State arrays are updated in place;
 * *n is updated when particles are removed.  method: 1 pairwise, 2 massless.
 * Returns 0, or 1 on an invalid Stumpff argument when removal is disabled. */
int orc_whfast(int *n, int32_t *ids, double *x, double *v, double *m, double G, int method,
               double eps, double dt, double tf, int remove_invalid);

/* IAS15 (src/integrator_ias15.c:229-1382): Gauss-Radau predictor-corrector
 * with adaptive steps and compensated sums, the reference's operation order,
 * using orc_pairwise (method 1) or orc_massless (2).  initial_dt <= 0 picks
 * the reference's automatic initial step.  x, v updated in place; returns the
 * number of accepted steps. */
int64_t orc_ias15(double *x, double *v, int n, const double *m, double G, int method, double eps,
                  double tol, double initial_dt, double tf);

/* Euler (cromer 0) / Euler-Cromer (cromer 1), RK4, embedded RK (integrator
 * ids 5-8); each returns the number of accepted steps */
int64_t orc_euler(double *x, double *v, int n, const double *m, double G, int method, double eps,
                  double dt, double tf, int cromer);
int64_t orc_rk4(double *x, double *v, int n, const double *m, double G, int method, double eps,
                double dt, double tf);
int64_t orc_rk_embedded(double *x, double *v, int n, const double *m, double G, int method,
                        double eps, int integrator, double tol, double initial_dt, double tf);

#ifdef __cplusplus
}
#endif

#endif
