/*
 * oracle.c -- CPU restatement of grav_sim's force-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Build with -ffp-contract=off:
 * every FP64 expression below is written in the association order of the
 * reference expression it restates, so results are bitwise reproducible.
 */
#include "oracle.h"
#include "ias15_coeffs.h"
#include "rk_tableaus.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAX_LEVEL 21

/* src/math_functions.c:28-31 */
static double norm3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* ------------------------------------------------------------------ */
/* Direct sums                                                        */
/* ------------------------------------------------------------------ */

/* One source term as the reference forms it (src/acceleration.c:206-229):
 * R = x_l - x_k (the sign flip of the j>k orientation is exact in IEEE),
 * r = sqrt(((R0^2 + R1^2) + R2^2) + eps^2), t = G / ((r*r)*r),
 * term = (t * R) * m_l. */
static void add_term(double acc[3], const double *xk, const double *xl, double ml, double G,
                     double eps2)
{
    const double R0 = xl[0] - xk[0];
    const double R1 = xl[1] - xk[1];
    const double R2 = xl[2] - xk[2];
    const double r = sqrt(R0 * R0 + R1 * R1 + R2 * R2 + eps2);
    const double t = G / (r * r * r);
    acc[0] += (t * R0) * ml;
    acc[1] += (t * R1) * ml;
    acc[2] += (t * R2) * ml;
}

void orc_pairwise(double *a, int n, const double *x, const double *m, double G, double eps)
{
    const double eps2 = eps * eps;
    for (int k = 0; k < n; k++)
    {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int l = 0; l < n; l++)
        {
            if (l == k)
                continue;
            add_term(acc, &x[k * 3], &x[l * 3], m[l], G, eps2);
        }
        a[k * 3 + 0] = acc[0];
        a[k * 3 + 1] = acc[1];
        a[k * 3 + 2] = acc[2];
    }
}

void orc_massless(double *a, int n, const double *x, const double *m, double G, double eps)
{
    const double eps2 = eps * eps;
    for (int k = 0; k < n; k++)
    {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int l = 0; l < n; l++)
        {
            if (l == k || m[l] == 0.0)
                continue;
            add_term(acc, &x[k * 3], &x[l * 3], m[l], G, eps2);
        }
        a[k * 3 + 0] = acc[0];
        a[k * 3 + 1] = acc[1];
        a[k * 3 + 2] = acc[2];
    }
}

/* ------------------------------------------------------------------ */
/* Octree pipeline                                                     */
/* ------------------------------------------------------------------ */

void orc_bbox(double center[3], double *width, int n, const double *x)
{
    double lo[3] = {x[0], x[1], x[2]};
    double hi[3] = {x[0], x[1], x[2]};
    for (int i = 1; i < n; i++)
        for (int d = 0; d < 3; d++)
        {
            lo[d] = fmin(lo[d], x[i * 3 + d]);
            hi[d] = fmax(hi[d], x[i * 3 + d]);
        }
    for (int d = 0; d < 3; d++)
        center[d] = (hi[d] + lo[d]) / 2.0;
    *width = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
}

void orc_morton(int64_t *keys, int n, const double *x, const double center[3], double width)
{
    for (int i = 0; i < n; i++)
    {
        int64_t q[3];
        for (int d = 0; d < 3; d++)
        {
            /* src/linear_octree.c:170-179: normalise, scale by 2^21, truncate, mask */
            const double u = (x[i * 3 + d] - center[d]) / width + 0.5;
            int64_t t = (int64_t)(u * (double)(1 << 21));
            q[d] = t & 0x1fffff;
        }
        int64_t key = 0;
        for (int b = 0; b < MAX_LEVEL; b++)
            for (int d = 0; d < 3; d++)
                key |= ((q[d] >> b) & 1LL) << (3 * b + d);
        keys[i] = key;
    }
}

void orc_sort(int64_t *keys, int32_t *idx, int n)
{
    /* bottom-up merge sort, stable */
    int64_t *k2 = malloc((size_t)n * sizeof(int64_t));
    int32_t *i2 = malloc((size_t)n * sizeof(int32_t));
    for (int i = 0; i < n; i++)
        idx[i] = i;
    int64_t *ks = keys, *kd = k2;
    int32_t *is = idx, *id = i2;
    for (int w = 1; w < n; w *= 2)
    {
        for (int lo = 0; lo < n; lo += 2 * w)
        {
            int mid = lo + w < n ? lo + w : n;
            int hi = lo + 2 * w < n ? lo + 2 * w : n;
            int p = lo, q = mid, o = lo;
            while (p < mid && q < hi)
            {
                if (ks[q] < ks[p])
                {
                    kd[o] = ks[q];
                    id[o++] = is[q++];
                }
                else
                {
                    kd[o] = ks[p];
                    id[o++] = is[p++];
                }
            }
            while (p < mid)
            {
                kd[o] = ks[p];
                id[o++] = is[p++];
            }
            while (q < hi)
            {
                kd[o] = ks[q];
                id[o++] = is[q++];
            }
        }
        int64_t *tk = ks;
        ks = kd;
        kd = tk;
        int32_t *ti = is;
        is = id;
        id = ti;
    }
    if (ks != keys)
    {
        memcpy(keys, ks, (size_t)n * sizeof(int64_t));
        memcpy(idx, is, (size_t)n * sizeof(int32_t));
    }
    free(k2);
    free(i2);
}

int orc_octree(int max_nodes, int n, const double *x, const double *m,
               const int64_t *sorted_keys, const int32_t *sorted_idx, int leaf_cap,
               int *depth, int64_t *prefix, int *count, int *first, int *nchild, int *child0,
               double *mass, double *mx, double *my, double *mz,
               double *comx, double *comy, double *comz)
{
    if (max_nodes < 1)
        return -1;
    /* root */
    depth[0] = 0;
    prefix[0] = 0;
    count[0] = n;
    first[0] = 0;
    int num = 1;
    /* level-order expansion: every node in [level_begin, level_end) is visited
     * in order, and its children (runs of equal prefix at depth+1) appended */
    for (int node = 0; node < num; node++)
    {
        const int d = depth[node];
        const int internal = (node == 0) || (count[node] > leaf_cap && d < MAX_LEVEL);
        nchild[node] = 0;
        child0[node] = -1;
        if (!internal)
            continue;
        const int cd = d + 1;
        const int shift = 3 * (MAX_LEVEL - cd);
        const int s = first[node], e = s + count[node];
        int i = s;
        child0[node] = num;
        while (i < e)
        {
            const int64_t p = sorted_keys[i] >> shift;
            int j = i + 1;
            while (j < e && (sorted_keys[j] >> shift) == p)
                j++;
            if (num >= max_nodes)
                return -1;
            depth[num] = cd;
            prefix[num] = p;
            count[num] = j - i;
            first[num] = i;
            nchild[node]++;
            num++;
            i = j;
        }
    }
    /* bottom-up multipoles: level order means children have larger indices */
    for (int node = num - 1; node >= 0; node--)
    {
        double M = 0.0, X = 0.0, Y = 0.0, Z = 0.0;
        if (nchild[node] == 0)
        {
            for (int t = first[node]; t < first[node] + count[node]; t++)
            {
                const int p = sorted_idx[t];
                M += m[p];
                X += m[p] * x[p * 3 + 0];
                Y += m[p] * x[p * 3 + 1];
                Z += m[p] * x[p * 3 + 2];
            }
        }
        else
        {
            for (int c = child0[node]; c < child0[node] + nchild[node]; c++)
            {
                M += mass[c];
                X += mx[c];
                Y += my[c];
                Z += mz[c];
            }
        }
        mass[node] = M;
        mx[node] = X;
        my[node] = Y;
        mz[node] = Z;
        comx[node] = X / M;
        comy[node] = Y / M;
        comz[node] = Z / M;
    }
    return num;
}

typedef struct
{
    const double *x, *m;
    double G, eps2, theta2, width;
    const int64_t *keys;
    const int32_t *sidx;
    const int *depth, *count, *first, *nchild, *child0;
    const int64_t *prefix;
    const double *mass, *comx, *comy, *comz;
    const double *quad; /* 6 per node (xx yy zz xy xz yz) or NULL: monopole only */
    const double *oct;  /* 10 per node (xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz) or NULL */
    int features;       /* nonzero: quad and oct feed the features, not the acceleration */
} Walk;

/* Walk features of one target (the learned correction's inputs, without G):
 * q  = sum of the quadrupole accelerations of accepted nodes,
 * v  = sum (M l^3 / r^6) R, s2 = sum M l^2 / r^4, s3 = sum M l^3 / r^5 over
 *      accepted nodes (l: node side, R = x_i - com, r^2 = |R|^2 + eps^2),
 * nacc / npp = accepted nodes / particle interactions, where a leaf holding a
 * single particle always counts as a particle interaction. */
typedef struct
{
    double q[3], v[3], s2, s3, nacc, npp, o[3];
} WalkFeat;

/* Quadrupole term of an accepted node at R = x_i - com (r^2 = |R|^2 + eps^2):
 *   a += G [ 3 Q R / r^5 - (15/2) (R.Q.R) R / r^7 + (3/2) tr(Q) R / r^5 ]
 * the gradient of the second-order term of the potential
 *   phi = -G [ M / r + (3 R.Q.R - tr(Q) r^2) / (2 r^5) ],
 * Q_ab = sum m (x - com)_a (x - com)_b. */
static void quad_accel(double acc[3], const double *q, double R0, double R1, double R2, double r2,
                       double G)
{
    const double QR0 = q[0] * R0 + q[3] * R1 + q[4] * R2;
    const double QR1 = q[3] * R0 + q[1] * R1 + q[5] * R2;
    const double QR2 = q[4] * R0 + q[5] * R1 + q[2] * R2;
    const double RQR = R0 * QR0 + R1 * QR1 + R2 * QR2;
    const double tr = q[0] + q[1] + q[2];
    const double ir2 = 1.0 / r2;
    const double ir5 = ir2 * ir2 / sqrt(r2);
    const double s = 1.5 * tr * ir5 - 7.5 * RQR * ir5 * ir2;
    acc[0] += G * (3.0 * QR0 * ir5 + s * R0);
    acc[1] += G * (3.0 * QR1 * ir5 + s * R1);
    acc[2] += G * (3.0 * QR2 * ir5 + s * R2);
}

/* Octupole term of an accepted node at R = x_i - com (r^2 = |R|^2 + eps^2),
 * the next order of the expansion of -G sum m (R - s) / |R - s|^3 in the
 * offsets s from the centre of mass:
 *   a_d += G [ -(35/2) u R_d / r^9 + (15/2) ((t.R) R_d + w_d) / r^7 - (3/2) t_d / r^5 ]
 * with T_abc = sum m s_a s_b s_c, w_d = T_abd R_a R_b, u = w.R, t_d = T_aad
 * (-(1/6) T_abc d_a d_b d_c d_d (1/r) written out). */
static void oct_accel(double acc[3], const double *T, double R0, double R1, double R2, double r2,
                      double G)
{
    const double Txxx = T[0], Tyyy = T[1], Tzzz = T[2], Txxy = T[3], Txxz = T[4], Txyy = T[5],
                 Tyyz = T[6], Txzz = T[7], Tyzz = T[8], Txyz = T[9];
    const double w0 = Txxx * R0 * R0 + Txyy * R1 * R1 + Txzz * R2 * R2 +
                      2.0 * (Txxy * R0 * R1 + Txxz * R0 * R2 + Txyz * R1 * R2);
    const double w1 = Txxy * R0 * R0 + Tyyy * R1 * R1 + Tyzz * R2 * R2 +
                      2.0 * (Txyy * R0 * R1 + Txyz * R0 * R2 + Tyyz * R1 * R2);
    const double w2 = Txxz * R0 * R0 + Tyyz * R1 * R1 + Tzzz * R2 * R2 +
                      2.0 * (Txyz * R0 * R1 + Txzz * R0 * R2 + Tyzz * R1 * R2);
    const double t0 = Txxx + Txyy + Txzz, t1 = Txxy + Tyyy + Tyzz, t2 = Txxz + Tyyz + Tzzz;
    const double u = w0 * R0 + w1 * R1 + w2 * R2;
    const double tR = t0 * R0 + t1 * R1 + t2 * R2;
    const double ir2 = 1.0 / r2;
    const double ir5 = ir2 * ir2 / sqrt(r2);
    const double ir7 = ir5 * ir2, ir9 = ir7 * ir2;
    const double su = -17.5 * u * ir9 + 7.5 * tR * ir7;
    acc[0] += G * (su * R0 + 7.5 * w0 * ir7 - 1.5 * t0 * ir5);
    acc[1] += G * (su * R1 + 7.5 * w1 * ir7 - 1.5 * t1 * ir5);
    acc[2] += G * (su * R2 + 7.5 * w2 * ir7 - 1.5 * t2 * ir5);
}

static void walk_node(const Walk *w, int node, int i, int64_t key_i, const double xi[3],
                      double acc[3], int64_t *nint, WalkFeat *f)
{
    const int d = w->depth[node];
    const int single = w->nchild[node] == 0 && w->count[node] == 1;
    const int included = (key_i >> (3 * (MAX_LEVEL - d))) == w->prefix[node];
    if (!included)
    {
        /* src/acceleration_barnes_hut.c:152-175 */
        const double R0 = xi[0] - w->comx[node];
        const double R1 = xi[1] - w->comy[node];
        const double R2 = xi[2] - w->comz[node];
        const double l = (w->width * 2.0) / (double)(2LL << d);
        const double d2 = R0 * R0 + R1 * R1 + R2 * R2;
        if ((l * l) < w->theta2 * d2)
        {
            const double r = sqrt(d2 + w->eps2);
            const double t = w->G * w->mass[node] / (r * r * r);
            acc[0] -= t * R0;
            acc[1] -= t * R1;
            acc[2] -= t * R2;
            if (w->quad && !w->features)
                quad_accel(acc, &w->quad[6 * node], R0, R1, R2, d2 + w->eps2, w->G);
            if (w->oct && !w->features)
                oct_accel(acc, &w->oct[10 * node], R0, R1, R2, d2 + w->eps2, w->G);
            if (f && single)
                f->npp += 1.0;
            else if (f)
            {
                const double r2 = d2 + w->eps2;
                const double M = w->mass[node];
                quad_accel(f->q, &w->quad[6 * node], R0, R1, R2, r2, 1.0);
                if (w->oct)
                    oct_accel(f->o, &w->oct[10 * node], R0, R1, R2, r2, 1.0);
                f->s2 += M * l * l / (r2 * r2);
                f->s3 += M * l * l * l / (r2 * r2 * r);
                const double k3 = M * l * l * l / (r2 * r2 * r2);
                f->v[0] += k3 * R0;
                f->v[1] += k3 * R1;
                f->v[2] += k3 * R2;
                f->nacc += 1.0;
            }
            (*nint)++;
            return;
        }
    }
    if (w->nchild[node] == 0)
    {
        /* src/acceleration_barnes_hut.c:181-214 */
        for (int t = w->first[node]; t < w->first[node] + w->count[node]; t++)
        {
            const int j = w->sidx[t];
            if (j == i)
                continue;
            const double R0 = xi[0] - w->x[j * 3 + 0];
            const double R1 = xi[1] - w->x[j * 3 + 1];
            const double R2 = xi[2] - w->x[j * 3 + 2];
            const double r = sqrt(R0 * R0 + R1 * R1 + R2 * R2 + w->eps2);
            const double tv = w->G * w->m[j] / (r * r * r);
            acc[0] -= tv * R0;
            acc[1] -= tv * R1;
            acc[2] -= tv * R2;
            (*nint)++;
            if (f)
                f->npp += 1.0;
        }
        return;
    }
    for (int c = w->child0[node]; c < w->child0[node] + w->nchild[node]; c++)
        walk_node(w, c, i, key_i, xi, acc, nint, f);
}

void orc_bh_walk(double *a, int64_t *n_int, int n, const double *x, const double *m, double G,
                 double eps, double theta, double width, const int64_t *sorted_keys,
                 const int32_t *sorted_idx, const int *depth, const int64_t *prefix,
                 const int *count, const int *first, const int *nchild, const int *child0,
                 const double *mass, const double *comx, const double *comy, const double *comz,
                 const double *quad, const double *oct, double *feat)
{
    Walk w = {x,     m,     G,      eps * eps, theta * theta, width,  sorted_keys, sorted_idx,
              depth, count, first,  nchild,    child0,        prefix, mass,        comx,
              comy,  comz,  quad,   oct,       feat != NULL};
#pragma omp parallel for schedule(dynamic, 64)
    for (int s = 0; s < n; s++)
    {
        const int i = sorted_idx[s];
        const double xi[3] = {x[i * 3 + 0], x[i * 3 + 1], x[i * 3 + 2]};
        double acc[3] = {0.0, 0.0, 0.0};
        int64_t cnt = 0;
        WalkFeat f = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, 0.0, 0.0, 0.0, 0.0, {0.0, 0.0, 0.0}};
        /* the root is always internal: start from its children */
        for (int c = child0[0]; c < child0[0] + nchild[0]; c++)
            walk_node(&w, c, i, sorted_keys[s], xi, acc, &cnt, feat ? &f : NULL);
        if (feat)
        {
            double *o = &feat[(size_t)i * 16];
            o[0] = f.q[0], o[1] = f.q[1], o[2] = f.q[2];
            o[3] = f.v[0], o[4] = f.v[1], o[5] = f.v[2];
            o[6] = f.s2, o[7] = f.s3, o[8] = f.nacc, o[9] = f.npp;
            o[10] = f.o[0], o[11] = f.o[1], o[12] = f.o[2];
            o[13] = 0.0, o[14] = 0.0, o[15] = 0.0;
        }
        a[i * 3 + 0] = acc[0];
        a[i * 3 + 1] = acc[1];
        a[i * 3 + 2] = acc[2];
        if (n_int)
            n_int[i] = cnt;
    }
}

/* Barnes-Hut walk of groups of at most `group` consecutive particles in
 * sorted order, formed from the tree as the GPU forms them (below): a node
 * is accepted for the whole group when
 * none of the group's particles lies in it and (l)^2 < theta^2 d^2 with d the
 * distance from the node's centre of mass to the group's bounding box, else
 * opened (leaves: particle-particle terms, skipping self).  Every particle
 * of the group then has d_i >= d, so each accepted node would also be
 * accepted by the per-particle test of src/acceleration_barnes_hut.c:152-175;
 * with group = 1 the box is the particle and the walk is orc_bh_walk's. */
/* Appends the run [f, f + c) of sorted positions to the group list, or its
 * particles one by one when its bounding box is wider than cap in some axis. */
static int emit_group(int (*gl)[2], int ng, const double *x, const int32_t *sorted_idx, int f, int c,
                      double cap)
{
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++)
        lo[k] = hi[k] = x[sorted_idx[f] * 3 + k];
    for (int q = f + 1; q < f + c; q++)
        for (int k = 0; k < 3; k++)
        {
            const double v = x[sorted_idx[q] * 3 + k];
            lo[k] = fmin(lo[k], v);
            hi[k] = fmax(hi[k], v);
        }
    if (hi[0] - lo[0] > cap || hi[1] - lo[1] > cap || hi[2] - lo[2] > cap)
        for (int q = f; q < f + c; q++)
            gl[ng][0] = q, gl[ng][1] = 1, ng++;
    else
        gl[ng][0] = f, gl[ng][1] = c, ng++;
    return ng;
}

typedef struct
{
    const Walk *w;
    int s0, s1; /* sorted positions [s0, s1) */
    double lo[3], hi[3];
    double (*xs)[3];
    double (*acc)[3];
    int64_t *cnt;
    const int *idx;
    /* local expansions (phi2 > 0): box centre, squared half-diagonal, the
     * far nodes' squared opening angle and quadrupole moments, their moments
     * and their number */
    double phi2, c[3], rg2, thf2;
    const double *quad_far;
    double *F;
    int64_t *nfar;
} GroupWalk;

/* Second-order local expansion of far monopole fields about a centre c (no
 * counterpart in the reference: the group walk with local expansions is this
 * library's own method).  A source of mass M at y = c - com, s2 = |y|^2 +
 * eps^2, has the Plummer-softened field a(c + d) = -M (y + d) / (|y + d|^2 +
 * eps^2)^(3/2), whose Taylor terms about d = 0 are
 *   a0_k = -q1 y_k,  J_kl = q2 y_k y_l - q1 delta_kl,
 *   H_klm = q2 (delta_kl y_m + delta_km y_l + delta_lm y_k) - q3 y_k y_l y_m
 * with q1 = M s2^-3/2, q2 = 3 M s2^-5/2, q3 = 15 M s2^-7/2.  F holds the sums
 * over sources of Q = q1, A = -q1 y (3), S = q2 y (3), P = q2 y y (xx, yy,
 * zz, xy, xz, yz) and T = q3 y y y (xxx, yyy, zzz, xxy, xxz, xyy, yyz, xzz,
 * yzz, xyz): 23 numbers. */
void orc_far_add(double *F, double y0, double y1, double y2, double M, double eps2)
{
    const double s2 = y0 * y0 + y1 * y1 + y2 * y2 + eps2;
    const double s = sqrt(s2);
    const double q1 = M / (s2 * s);
    const double q2 = 3.0 * q1 / s2;
    const double q3 = 5.0 * q2 / s2;
    const double y[3] = {y0, y1, y2};
    F[0] += q1;
    for (int k = 0; k < 3; k++)
    {
        F[1 + k] -= q1 * y[k];
        F[4 + k] += q2 * y[k];
    }
    static const int P2[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    for (int j = 0; j < 6; j++)
        F[7 + j] += q2 * y[P2[j][0]] * y[P2[j][1]];
    static const int T3[10][3] = {{0, 0, 0}, {1, 1, 1}, {2, 2, 2}, {0, 0, 1}, {0, 0, 2},
                                  {0, 1, 1}, {1, 1, 2}, {0, 2, 2}, {1, 2, 2}, {0, 1, 2}};
    for (int j = 0; j < 10; j++)
        F[13 + j] += q3 * y[T3[j][0]] * y[T3[j][1]] * y[T3[j][2]];
}

/* a(c + d) = a0 + J d + (1/2) H[d, d] from the sums of orc_far_add, with the
 * symmetric tensors expanded to full index form. */
void orc_far_eval(const double *F, double d0, double d1, double d2, double out[3])
{
    const double d[3] = {d0, d1, d2};
    double P[3][3], S[3], T[3][3][3];
    static const int P2[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    static const int T3[10][3] = {{0, 0, 0}, {1, 1, 1}, {2, 2, 2}, {0, 0, 1}, {0, 0, 2},
                                  {0, 1, 1}, {1, 1, 2}, {0, 2, 2}, {1, 2, 2}, {0, 1, 2}};
    for (int j = 0; j < 6; j++)
        P[P2[j][0]][P2[j][1]] = P[P2[j][1]][P2[j][0]] = F[7 + j];
    for (int j = 0; j < 10; j++)
    {
        const int a = T3[j][0], b = T3[j][1], c = T3[j][2];
        T[a][b][c] = T[a][c][b] = T[b][a][c] = T[b][c][a] = T[c][a][b] = T[c][b][a] = F[13 + j];
    }
    for (int k = 0; k < 3; k++)
        S[k] = F[4 + k];
    for (int k = 0; k < 3; k++)
    {
        double v = F[1 + k];
        for (int l = 0; l < 3; l++)
        {
            const double J = P[k][l] - (k == l ? F[0] : 0.0);
            v += J * d[l];
            for (int m = 0; m < 3; m++)
            {
                const double H = S[m] * (k == l) + S[l] * (k == m) + S[k] * (l == m) - T[k][l][m];
                v += 0.5 * H * d[l] * d[m];
            }
        }
        out[k] = v;
    }
}

static void group_node(const GroupWalk *g, int node)
{
    const Walk *w = g->w;
    const int d = w->depth[node];
    const int f = w->first[node], c = w->count[node];
    const int included = f < g->s1 && f + c > g->s0;
    const int gn = g->s1 - g->s0;
    if (!included)
    {
        double dd2 = 0.0;
        const double com[3] = {w->comx[node], w->comy[node], w->comz[node]};
        for (int k = 0; k < 3; k++)
        {
            const double e = com[k] < g->lo[k] ? g->lo[k] - com[k] : (com[k] > g->hi[k] ? com[k] - g->hi[k] : 0.0);
            dd2 += e * e;
        }
        const double l = (w->width * 2.0) / (double)(2LL << d);
        /* far: rg < phi |com - c| (the same FP64 expressions as the GPU's
         * test), accepted at the far opening angle */
        const double y0 = g->c[0] - com[0], y1 = g->c[1] - com[1], y2 = g->c[2] - com[2];
        const int far = g->phi2 > 0.0 && g->rg2 < g->phi2 * (y0 * y0 + y1 * y1 + y2 * y2);
        if ((l * l) < w->theta2 * dd2 || (far && (l * l) < g->thf2 * dd2))
        {
            if (far)
            {
                orc_far_add(g->F, y0, y1, y2, w->mass[node], w->eps2);
                if (g->quad_far)
                    quad_accel(&g->F[1], &g->quad_far[6 * node], y0, y1, y2,
                               y0 * y0 + y1 * y1 + y2 * y2 + w->eps2, 1.0);
                (*g->nfar)++;
                for (int q = 0; q < gn; q++)
                    g->cnt[q]++;
                return;
            }
            for (int q = 0; q < gn; q++)
            {
                const double R0 = g->xs[q][0] - com[0], R1 = g->xs[q][1] - com[1], R2 = g->xs[q][2] - com[2];
                const double d2 = R0 * R0 + R1 * R1 + R2 * R2;
                const double r = sqrt(d2 + w->eps2);
                const double t = w->G * w->mass[node] / (r * r * r);
                g->acc[q][0] -= t * R0;
                g->acc[q][1] -= t * R1;
                g->acc[q][2] -= t * R2;
                if (w->quad)
                    quad_accel(g->acc[q], &w->quad[6 * node], R0, R1, R2, d2 + w->eps2, w->G);
                g->cnt[q]++;
            }
            return;
        }
    }
    if (w->nchild[node] == 0)
    {
        for (int t = f; t < f + c; t++)
        {
            const int j = w->sidx[t];
            for (int q = 0; q < gn; q++)
            {
                if (g->s0 + q == t)
                    continue;
                const double R0 = g->xs[q][0] - w->x[j * 3 + 0];
                const double R1 = g->xs[q][1] - w->x[j * 3 + 1];
                const double R2 = g->xs[q][2] - w->x[j * 3 + 2];
                const double r = sqrt(R0 * R0 + R1 * R1 + R2 * R2 + w->eps2);
                const double tv = w->G * w->m[j] / (r * r * r);
                g->acc[q][0] -= tv * R0;
                g->acc[q][1] -= tv * R1;
                g->acc[q][2] -= tv * R2;
                g->cnt[q]++;
            }
        }
        return;
    }
    for (int ch = w->child0[node]; ch < w->child0[node] + w->nchild[node]; ch++)
        group_node(g, ch);
}

void orc_bh_walk_group(double *a, int64_t *n_int, int n, const double *x, const double *m,
                       double G, double eps, double theta, double width,
                       const int64_t *sorted_keys, const int32_t *sorted_idx, const int *depth,
                       const int64_t *prefix, const int *count, const int *first,
                       const int *nchild, const int *child0, const double *mass,
                       const double *comx, const double *comy, const double *comz,
                       const double *quad, int group)
{
    orc_bh_walk_group_le(a, n_int, n, x, m, G, eps, theta, width, sorted_keys, sorted_idx, depth,
                         prefix, count, first, nchild, child0, mass, comx, comy, comz, quad, group,
                         0.0, 0.0, NULL, NULL);
}

void orc_bh_walk_group_le(double *a, int64_t *n_int, int n, const double *x, const double *m,
                          double G, double eps, double theta, double width,
                          const int64_t *sorted_keys, const int32_t *sorted_idx, const int *depth,
                          const int64_t *prefix, const int *count, const int *first,
                          const int *nchild, const int *child0, const double *mass,
                          const double *comx, const double *comy, const double *comz,
                          const double *quad, int group, double phi, double theta_far,
                          const double *quad_far, int64_t *far_pairs)
{
    int64_t far_total = 0;
    const double thf2 = theta_far > theta ? theta_far * theta_far : theta * theta;
    Walk w = {x,     m,     G,      eps * eps, theta * theta, width,  sorted_keys, sorted_idx,
              depth, count, first,  nchild,    child0,        prefix, mass,        comx,
              comy,  comz,  quad,   NULL,      0};
    /* groups from the tree: children of at most `group` particles of each node
     * with more are packed in order into runs of at most `group` particles; a
     * root within the size is one group; larger leaves give runs of `group`.
     * A run wider than `cap` in some axis gives one group per particle. */
    const double cap = 4.0 * width * cbrt(64.0 / (double)n);
    int (*gl)[2] = malloc(sizeof(int[2]) * (size_t)(n > 0 ? n : 1));
    int ng = 0;
    if (count[0] <= group)
        gl[ng][0] = 0, gl[ng][1] = n, ng++;
    else
    {
        /* nodes are visited depth-first from the root (the group order does not
         * change any particle's result); at most 8 pending nodes per level */
        int *stack = malloc(sizeof(int) * 8 * (MAX_LEVEL + 2));
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0)
        {
            const int p = stack[--sp];
            if (count[p] <= group)
                continue;
            if (nchild[p] == 0)
            {
                for (int f = first[p]; f < first[p] + count[p]; f += group)
                    ng = emit_group(gl, ng, x, sorted_idx, f,
                                    (first[p] + count[p] - f) < group ? (first[p] + count[p] - f) : group, cap);
                continue;
            }
            int start = -1, cnt = 0;
            for (int c = child0[p]; c < child0[p] + nchild[p]; c++)
            {
                if (count[c] > group || cnt + count[c] > group)
                {
                    if (cnt > 0)
                        ng = emit_group(gl, ng, x, sorted_idx, start, cnt, cap);
                    start = -1, cnt = 0;
                }
                if (count[c] <= group)
                {
                    if (cnt == 0)
                        start = first[c];
                    cnt += count[c];
                }
                else
                    stack[sp++] = c;
            }
            if (cnt > 0)
                ng = emit_group(gl, ng, x, sorted_idx, start, cnt, cap);
        }
        free(stack);
    }
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : far_total)
    for (int gi = 0; gi < ng; gi++)
    {
        GroupWalk g;
        double F[23] = {0.0};
        int64_t nfar = 0;
        g.w = &w;
        g.F = F;
        g.nfar = &nfar;
        g.phi2 = phi > 0.0 ? phi * phi : 0.0;
        g.thf2 = thf2;
        g.quad_far = quad_far;
        g.s0 = gl[gi][0];
        g.s1 = gl[gi][0] + gl[gi][1];
        const int gn = g.s1 - g.s0;
        double (*xs)[3] = malloc(sizeof(double[3]) * gn);
        double (*acc)[3] = calloc(gn, sizeof(double[3]));
        int64_t *cnt = calloc(gn, sizeof(int64_t));
        for (int k = 0; k < 3; k++)
            g.lo[k] = INFINITY, g.hi[k] = -INFINITY;
        for (int q = 0; q < gn; q++)
        {
            const int i = sorted_idx[g.s0 + q];
            for (int k = 0; k < 3; k++)
            {
                xs[q][k] = x[i * 3 + k];
                if (xs[q][k] < g.lo[k]) g.lo[k] = xs[q][k];
                if (xs[q][k] > g.hi[k]) g.hi[k] = xs[q][k];
            }
        }
        g.xs = xs;
        g.acc = acc;
        g.cnt = cnt;
        {
            const double h0 = 0.5 * (g.hi[0] - g.lo[0]), h1 = 0.5 * (g.hi[1] - g.lo[1]),
                         h2 = 0.5 * (g.hi[2] - g.lo[2]);
            g.rg2 = h0 * h0 + h1 * h1 + h2 * h2;
            for (int k = 0; k < 3; k++)
                g.c[k] = 0.5 * (g.lo[k] + g.hi[k]);
        }
        for (int c = child0[0]; c < child0[0] + nchild[0]; c++)
            group_node(&g, c);
        if (nfar > 0)
            for (int q = 0; q < gn; q++)
            {
                double e[3];
                orc_far_eval(F, xs[q][0] - g.c[0], xs[q][1] - g.c[1], xs[q][2] - g.c[2], e);
                for (int k = 0; k < 3; k++)
                    acc[q][k] += w.G * e[k];
            }
        far_total += nfar;
        for (int q = 0; q < gn; q++)
        {
            const int i = sorted_idx[g.s0 + q];
            for (int k = 0; k < 3; k++)
                a[i * 3 + k] = acc[q][k];
            if (n_int)
                n_int[i] = cnt[q];
        }
        free(xs);
        free(acc);
        free(cnt);
    }
    free(gl);
    if (far_pairs)
        *far_pairs = far_total;
}

void orc_octree_quad(int num, const double *x, const double *m, const int32_t *sorted_idx,
                     const int *count, const int *first, const int *nchild, const int *child0,
                     const double *mass, const double *comx, const double *comy,
                     const double *comz, double *quad)
{
    for (int node = num - 1; node >= 0; node--)
    {
        double q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const double c[3] = {comx[node], comy[node], comz[node]};
        if (nchild[node] == 0)
        {
            for (int t = first[node]; t < first[node] + count[node]; t++)
            {
                const int p = sorted_idx[t];
                /* a massless particle adds nothing; skipped, since a massless
                 * node's centre of mass is 0/0 as in the reference */
                if (m[p] == 0.0)
                    continue;
                const double d0 = x[p * 3 + 0] - c[0], d1 = x[p * 3 + 1] - c[1],
                             d2 = x[p * 3 + 2] - c[2];
                q[0] += m[p] * d0 * d0;
                q[1] += m[p] * d1 * d1;
                q[2] += m[p] * d2 * d2;
                q[3] += m[p] * d0 * d1;
                q[4] += m[p] * d0 * d2;
                q[5] += m[p] * d1 * d2;
            }
        }
        else
        {
            for (int ch = child0[node]; ch < child0[node] + nchild[node]; ch++)
            {
                if (mass[ch] == 0.0)
                    continue; /* massless child: no moments, centre 0/0 */
                const double d0 = comx[ch] - c[0], d1 = comy[ch] - c[1], d2 = comz[ch] - c[2];
                const double M = mass[ch];
                const double *qc = &quad[6 * ch];
                q[0] += qc[0] + M * d0 * d0;
                q[1] += qc[1] + M * d1 * d1;
                q[2] += qc[2] + M * d2 * d2;
                q[3] += qc[3] + M * d0 * d1;
                q[4] += qc[4] + M * d0 * d2;
                q[5] += qc[5] + M * d1 * d2;
            }
        }
        for (int k = 0; k < 6; k++)
            quad[6 * node + k] = q[k];
    }
}

/* Third central moments T_abc = sum m s_a s_b s_c (s = x - com), 10 per node
 * (xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz): leaves over their particles,
 * internal nodes from their children shifted to this centre of mass,
 *   T_abc += T^c_abc + Q^c_ab d_c + Q^c_ac d_b + Q^c_bc d_a + M_c d_a d_b d_c
 * (d = com_child - com; the children's first moments about their own centres
 * vanish).  quad: the second moments of orc_octree_quad. */
void orc_octree_oct(int num, const double *x, const double *m, const int32_t *sorted_idx,
                    const int *count, const int *first, const int *nchild, const int *child0,
                    const double *mass, const double *comx, const double *comy,
                    const double *comz, const double *quad, double *oct)
{
    for (int node = num - 1; node >= 0; node--)
    {
        double T[10] = {0.0};
        const double c[3] = {comx[node], comy[node], comz[node]};
        if (nchild[node] == 0)
        {
            for (int t = first[node]; t < first[node] + count[node]; t++)
            {
                const int p = sorted_idx[t];
                if (m[p] == 0.0)
                    continue; /* as in orc_octree_quad */
                const double d0 = x[p * 3 + 0] - c[0], d1 = x[p * 3 + 1] - c[1],
                             d2 = x[p * 3 + 2] - c[2];
                const double mp = m[p];
                T[0] += mp * d0 * d0 * d0;
                T[1] += mp * d1 * d1 * d1;
                T[2] += mp * d2 * d2 * d2;
                T[3] += mp * d0 * d0 * d1;
                T[4] += mp * d0 * d0 * d2;
                T[5] += mp * d0 * d1 * d1;
                T[6] += mp * d1 * d1 * d2;
                T[7] += mp * d0 * d2 * d2;
                T[8] += mp * d1 * d2 * d2;
                T[9] += mp * d0 * d1 * d2;
            }
        }
        else
        {
            for (int ch = child0[node]; ch < child0[node] + nchild[node]; ch++)
            {
                if (mass[ch] == 0.0)
                    continue; /* as in orc_octree_quad */
                const double d0 = comx[ch] - c[0], d1 = comy[ch] - c[1], d2 = comz[ch] - c[2];
                const double M = mass[ch];
                const double *q = &quad[6 * ch]; /* xx yy zz xy xz yz */
                const double *tc = &oct[10 * ch];
                T[0] += tc[0] + 3.0 * q[0] * d0 + M * d0 * d0 * d0;
                T[1] += tc[1] + 3.0 * q[1] * d1 + M * d1 * d1 * d1;
                T[2] += tc[2] + 3.0 * q[2] * d2 + M * d2 * d2 * d2;
                T[3] += tc[3] + (q[0] * d1 + 2.0 * q[3] * d0) + M * d0 * d0 * d1;
                T[4] += tc[4] + (q[0] * d2 + 2.0 * q[4] * d0) + M * d0 * d0 * d2;
                T[5] += tc[5] + (2.0 * q[3] * d1 + q[1] * d0) + M * d0 * d1 * d1;
                T[6] += tc[6] + (q[1] * d2 + 2.0 * q[5] * d1) + M * d1 * d1 * d2;
                T[7] += tc[7] + (2.0 * q[4] * d2 + q[2] * d0) + M * d0 * d2 * d2;
                T[8] += tc[8] + (2.0 * q[5] * d2 + q[2] * d1) + M * d1 * d2 * d2;
                T[9] += tc[9] + (q[3] * d2 + q[4] * d1 + q[5] * d0) + M * d0 * d1 * d2;
            }
        }
        for (int k = 0; k < 10; k++)
            oct[10 * node + k] = T[k];
    }
}

/* ------------------------------------------------------------------ */
/* Leapfrog and energy                                                 */
/* ------------------------------------------------------------------ */

static void accel(double *a, int n, const double *x, const double *m, double G, int method,
                  double eps)
{
    if (method == 2)
        orc_massless(a, n, x, m, G, eps);
    else
        orc_pairwise(a, n, x, m, G, eps);
}

int64_t orc_leapfrog(double *x, double *v, int n, const double *m, double G, int method,
                     double eps, double dt, double tf)
{
    const size_t sz = (size_t)n * 3;
    double *a = malloc(sz * sizeof(double));
    double *xe = calloc(sz, sizeof(double));
    double *ve = calloc(sz, sizeof(double));

    /* src/integrator.c:962-982: a(x0), half kick with compensated sum */
    accel(a, n, x, m, G, method, eps);
    for (size_t q = 0; q < sz; q++)
    {
        const double v0 = v[q];
        ve[q] += 0.5 * a[q] * dt;
        v[q] = v0 + ve[q];
        ve[q] += v0 - v[q];
    }
    /* src/integrator.c:985-1085 */
    const int64_t total = (int64_t)ceil(tf / dt);
    double t = 0.0;
    int64_t steps = 0;
    while (steps < total)
    {
        if (t + dt > tf)
            dt = tf - t;
        for (size_t q = 0; q < sz; q++)
        {
            const double x0 = x[q];
            xe[q] += v[q] * dt;
            x[q] = x0 + xe[q];
            xe[q] += x0 - x[q];
        }
        accel(a, n, x, m, G, method, eps);
        for (size_t q = 0; q < sz; q++)
        {
            const double v0 = v[q];
            ve[q] += a[q] * dt;
            v[q] = v0 + ve[q];
            ve[q] += v0 - v[q];
        }
        steps++;
        t = steps * dt;
    }
    /* src/integrator.c:1087-1094 */
    for (size_t q = 0; q < sz; q++)
        v[q] -= 0.5 * a[q] * dt;
    free(a);
    free(xe);
    free(ve);
    return steps;
}

double orc_energy(int n, const double *x, const double *v, const double *m, double G)
{
    double e = 0.0;
    for (int i = 0; i < n; i++)
    {
        const double vn = norm3(&v[i * 3]);
        e += (0.5 * m[i] * vn * vn);
        for (int j = i + 1; j < n; j++)
        {
            const double r[3] = {x[i * 3 + 0] - x[j * 3 + 0], x[i * 3 + 1] - x[j * 3 + 1],
                                 x[i * 3 + 2] - x[j * 3 + 2]};
            e -= (G * m[i] * m[j] / norm3(r));
        }
    }
    return e;
}

/* ------------------------------------------------------------------ */
/* WHFast                                                              */
/* ------------------------------------------------------------------ */

void orc_stumpff(double z, double c[4])
{
    /* src/integrator_whfast.c:782-814 */
    int k = 0;
    while (fabs(z) > 0.1)
    {
        z /= 4.0;
        k++;
    }
    double c3 = (1.0 - z / 20.0 *
                           (1.0 - z / 42.0 *
                                      (1.0 - z / 72.0 *
                                                 (1.0 - z / 110.0 *
                                                            (1.0 - z / 156.0 * (1.0 - z / 210.0)))))) /
                6.0;
    double c2 = (1.0 - z / 12.0 *
                           (1.0 - z / 30.0 *
                                      (1.0 - z / 56.0 *
                                                 (1.0 - z / 90.0 *
                                                            (1.0 - z / 132.0 * (1.0 - z / 182.0)))))) /
                2.0;
    double c1 = 1.0 - z * c3;
    double c0 = 1.0 - z * c2;
    while (k > 0)
    {
        c3 = (c2 + c0 * c3) / 4.0;
        c2 = (c1 * c1) / 2.0;
        c1 = c0 * c1;
        c0 = (2.0 * c0 * c0) - 1.0;
        k--;
    }
    c[0] = c0;
    c[1] = c1;
    c[2] = c2;
    c[3] = c3;
}

typedef struct
{
    double d;
    int i;
} DistIdx;

static int cmp_dist(const void *pa, const void *pb)
{
    const DistIdx *a = pa, *b = pb;
    if (a->d < b->d)
        return -1;
    if (a->d > b->d)
        return 1;
    return (a->i > b->i) - (a->i < b->i); /* stable for ties */
}

/* src/system.c:1210-1335 applied to (pos, vel, m, ids); distance from the
 * particle whose id is 0, which is forced to distance 0. */
static void sort_by_distance(int n, int32_t *ids, double *pos, double *vel, double *m)
{
    int primary = -1;
    if (0 < n && ids[0] == 0)
        primary = 0;
    else
        for (int i = 0; i < n; i++)
            if (ids[i] == 0)
            {
                primary = i;
                break;
            }
    if (primary < 0)
        return;
    DistIdx *h = malloc((size_t)n * sizeof(DistIdx));
    for (int i = 0; i < n; i++)
    {
        const double dv[3] = {pos[i * 3 + 0] - pos[primary * 3 + 0],
                              pos[i * 3 + 1] - pos[primary * 3 + 1],
                              pos[i * 3 + 2] - pos[primary * 3 + 2]};
        h[i].d = norm3(dv);
        h[i].i = i;
    }
    h[primary].d = 0.0;
    qsort(h, (size_t)n, sizeof(DistIdx), cmp_dist);
    int32_t *ni = malloc((size_t)n * sizeof(int32_t));
    double *np = malloc((size_t)n * 3 * sizeof(double));
    double *nv = malloc((size_t)n * 3 * sizeof(double));
    double *nm = malloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; i++)
    {
        const int s = h[i].i;
        ni[i] = ids[s];
        nm[i] = m[s];
        for (int d = 0; d < 3; d++)
        {
            np[i * 3 + d] = pos[s * 3 + d];
            nv[i * 3 + d] = vel[s * 3 + d];
        }
    }
    memcpy(ids, ni, (size_t)n * sizeof(int32_t));
    memcpy(pos, np, (size_t)n * 3 * sizeof(double));
    memcpy(vel, nv, (size_t)n * 3 * sizeof(double));
    memcpy(m, nm, (size_t)n * sizeof(double));
    free(h);
    free(ni);
    free(np);
    free(nv);
    free(nm);
}

static void compute_eta(double *eta, int n, const double *m)
{
    /* src/integrator_whfast.c:1266-1279 */
    eta[0] = m[0];
    for (int i = 1; i < n; i++)
        eta[i] = eta[i - 1] + m[i];
}

static void cart_to_jacobi(double *jx, double *jv, int n, const double *x, const double *v,
                           const double *m, const double *eta)
{
    /* src/integrator_whfast.c:682-724 */
    double xc[3], vc[3];
    for (int d = 0; d < 3; d++)
    {
        xc[d] = m[0] * x[d];
        vc[d] = m[0] * v[d];
    }
    for (int i = 1; i < n; i++)
        for (int d = 0; d < 3; d++)
        {
            jx[i * 3 + d] = x[i * 3 + d] - xc[d] / eta[i - 1];
            jv[i * 3 + d] = v[i * 3 + d] - vc[d] / eta[i - 1];
            xc[d] = xc[d] * (1.0 + m[i] / eta[i - 1]) + m[i] * jx[i * 3 + d];
            vc[d] = vc[d] * (1.0 + m[i] / eta[i - 1]) + m[i] * jv[i * 3 + d];
        }
    for (int d = 0; d < 3; d++)
    {
        jx[d] = xc[d] / eta[n - 1];
        jv[d] = vc[d] / eta[n - 1];
    }
}

static void jacobi_to_cart(double *x, double *v, int n, const double *jx, const double *jv,
                           const double *m, const double *eta)
{
    /* src/integrator_whfast.c:726-772 */
    double xc[3], vc[3];
    for (int d = 0; d < 3; d++)
    {
        xc[d] = eta[n - 1] * jx[d];
        vc[d] = eta[n - 1] * jv[d];
    }
    for (int i = n - 1; i > 0; i--)
        for (int d = 0; d < 3; d++)
        {
            xc[d] = (xc[d] - m[i] * jx[i * 3 + d]) / eta[i];
            vc[d] = (vc[d] - m[i] * jv[i * 3 + d]) / eta[i];
            x[i * 3 + d] = jx[i * 3 + d] + xc[d];
            v[i * 3 + d] = jv[i * 3 + d] + vc[d];
            xc[d] = eta[i - 1] * xc[d];
            vc[d] = eta[i - 1] * vc[d];
        }
    for (int d = 0; d < 3; d++)
    {
        x[d] = xc[d] / m[0];
        v[d] = vc[d] / m[0];
    }
}

/* Kepler drift of one Jacobi particle (src/integrator_whfast.c:459-578).
 * Returns 1 if the particle must be removed. */
static int drift_one(double *xo, double *vo, double gm, double dt)
{
    const double x[3] = {xo[0], xo[1], xo[2]};
    const double v[3] = {vo[0], vo[1], vo[2]};
    const double xn = norm3(x);
    const double vn = norm3(v);
    const double rv = (x[0] * v[0] + x[1] * v[1] + x[2] * v[2]) / xn;
    const double alpha = 2.0 * gm / xn - (vn * vn);
    double s = dt / xn;
    double c[4] = {0.0, 0.0, 0.0, 0.0};
    int converged = 0, zbad = 0;
    for (int it = 0; it < 500; it++)
    {
        const double z = alpha * (s * s);
        if (!isfinite(z))
        {
            zbad = 1;
            break;
        }
        orc_stumpff(z, c);
        const double F = (xn * s * c[1] + xn * rv * (s * s) * c[2] + gm * (s * s * s) * c[3] - dt);
        const double dF = (xn * c[0] + xn * rv * s * c[1] + gm * (s * s) * c[2]);
        const double ds = -F / dF;
        s += ds;
        if (fabs(ds) < 1e-12)
        {
            converged = 1;
            break;
        }
    }
    const double r = xn * c[0] + xn * rv * s * c[1] + gm * (s * s) * c[2];
    int remove = 0;
    if (!converged)
    {
        const double err = (xn * s * c[1] + xn * rv * (s * s) * c[2] + gm * (s * s * s) * c[3] - dt) / r;
        if (err > 1e-5 || zbad)
            remove = 1;
    }
    const double f = 1.0 - gm * (s * s) * c[2] / xn;
    const double g = dt - gm * (s * s * s) * c[3];
    const double df = -gm * s * c[1] / (r * xn);
    const double dg = 1.0 - gm * (s * s) * c[2] / r;
    for (int d = 0; d < 3; d++)
    {
        xo[d] = f * x[d] + g * v[d];
        vo[d] = df * x[d] + dg * v[d];
    }
    return remove;
}

/* One interaction term of the WHFast accelerations: G*m*(R)/(|R|^3 + eps^3). */
static double cube_plus(const double *tv, double eps3)
{
    const double nrm = norm3(tv);
    return (nrm * nrm * nrm) + eps3;
}

static void whfast_accel(double *a, int n, int method, const double *x, const double *jx,
                         const double *m, const double *eta, double G, double eps)
{
    const double eps3 = eps * eps * eps;
    int *mi = malloc((size_t)n * sizeof(int));
    int *li = malloc((size_t)n * sizeof(int));
    int nm = 0, nl = 0;
    for (int i = 0; i < n; i++)
    {
        if (method == 1 || m[i] != 0.0)
            mi[nm++] = i;
        else
            li[nl++] = i;
    }
    a[0] = a[1] = a[2] = 0.0; /* particle 0 is never written by the reference */
    /* massive (or all, for the pairwise form): src/integrator_whfast.c:859-954, 1022-1127 */
    for (int p = 1; p < nm; p++)
    {
        const int i = mi[p];
        double tv[3], aux[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < 3; d++)
            tv[d] = x[i * 3 + d] - x[d];
        const double c0 = cube_plus(tv, eps3);
        const double cj = cube_plus(&jx[i * 3], eps3);
        for (int d = 0; d < 3; d++)
            a[i * 3 + d] = G * m[0] * eta[i] / eta[i - 1] * (jx[i * 3 + d] / cj - tv[d] / c0);
        for (int q = 1; q < p; q++)
        {
            const int j = mi[q];
            for (int d = 0; d < 3; d++)
                tv[d] = x[i * 3 + d] - x[j * 3 + d];
            const double cc = cube_plus(tv, eps3);
            for (int d = 0; d < 3; d++)
                aux[d] += G * m[j] * tv[d] / cc;
        }
        for (int d = 0; d < 3; d++)
            a[i * 3 + d] -= aux[d] * eta[i] / eta[i - 1];
        aux[0] = aux[1] = aux[2] = 0.0;
        for (int q = p + 1; q < nm; q++)
        {
            const int j = mi[q];
            for (int d = 0; d < 3; d++)
                tv[d] = x[j * 3 + d] - x[i * 3 + d];
            const double cc = cube_plus(tv, eps3);
            for (int d = 0; d < 3; d++)
                aux[d] += G * m[j] * tv[d] / cc;
        }
        for (int d = 0; d < 3; d++)
            a[i * 3 + d] += aux[d];
        aux[0] = aux[1] = aux[2] = 0.0;
        for (int q = 0; q < p; q++)
        {
            const int j = mi[q];
            for (int r = p + 1; r < nm; r++)
            {
                const int k = mi[r];
                for (int d = 0; d < 3; d++)
                    tv[d] = x[k * 3 + d] - x[j * 3 + d];
                const double cc = cube_plus(tv, eps3);
                for (int d = 0; d < 3; d++)
                    aux[d] += G * m[j] * m[k] * tv[d] / cc;
            }
        }
        for (int d = 0; d < 3; d++)
            a[i * 3 + d] -= aux[d] / eta[i - 1];
    }
    /* massless: src/integrator_whfast.c:1133-1258 */
    for (int p = 0; p < nl; p++)
    {
        const int i = li[p];
        if (i == 0)
            continue;
        double tv[3], aux[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < 3; d++)
            tv[d] = x[i * 3 + d] - x[d];
        const double c0 = cube_plus(tv, eps3);
        const double cj = cube_plus(&jx[i * 3], eps3);
        for (int d = 0; d < 3; d++)
            a[i * 3 + d] = G * m[0] * (jx[i * 3 + d] / cj - tv[d] / c0);
        for (int q = 1; q < nm; q++)
        {
            const int j = mi[q];
            if (j >= i)
                break;
            for (int d = 0; d < 3; d++)
                tv[d] = x[i * 3 + d] - x[j * 3 + d];
            const double cc = cube_plus(tv, eps3);
            for (int d = 0; d < 3; d++)
                aux[d] += G * m[j] * tv[d] / cc;
        }
        for (int d = 0; d < 3; d++)
            a[i * 3 + d] -= aux[d];
        aux[0] = aux[1] = aux[2] = 0.0;
        for (int q = 1; q < nm; q++)
        {
            const int j = mi[q];
            if (j <= i)
                continue;
            for (int d = 0; d < 3; d++)
                tv[d] = x[j * 3 + d] - x[i * 3 + d];
            const double cc = cube_plus(tv, eps3);
            for (int d = 0; d < 3; d++)
                aux[d] += G * m[j] * tv[d] / cc;
        }
        for (int d = 0; d < 3; d++)
            a[i * 3 + d] += aux[d];
        aux[0] = aux[1] = aux[2] = 0.0;
        for (int q = 0; q < nm; q++)
        {
            const int j = mi[q];
            if (j >= i)
                break;
            for (int r = q + 1; r < nm; r++)
            {
                const int k = mi[r];
                if (k <= i)
                    continue;
                for (int d = 0; d < 3; d++)
                    tv[d] = x[k * 3 + d] - x[j * 3 + d];
                const double cc = cube_plus(tv, eps3);
                for (int d = 0; d < 3; d++)
                    aux[d] += G * m[j] * m[k] * tv[d] / cc;
            }
        }
        for (int d = 0; d < 3; d++)
            a[i * 3 + d] -= aux[d] / eta[i - 1];
    }
    free(mi);
    free(li);
}

/* compaction in ascending index order (src/system.c:444-528, 530-577) */
static int compact(int n, const int *rm, int nrm, int32_t *ids, double *x, double *v, double *m,
                   double *jx, double *jv)
{
    int o = 0, r = 0;
    for (int i = 0; i < n; i++)
    {
        if (r < nrm && rm[r] == i)
        {
            r++;
            continue;
        }
        ids[o] = ids[i];
        m[o] = m[i];
        for (int d = 0; d < 3; d++)
        {
            x[o * 3 + d] = x[i * 3 + d];
            v[o * 3 + d] = v[i * 3 + d];
            jx[o * 3 + d] = jx[i * 3 + d];
            jv[o * 3 + d] = jv[i * 3 + d];
        }
        o++;
    }
    return o;
}

int orc_whfast(int *np, int32_t *ids, double *x, double *v, double *m, double G, int method,
               double eps, double dt, double tf, int remove_invalid)
{
    int n = *np;
    const size_t sz = (size_t)n * 3;
    double *jx = calloc(sz, sizeof(double));
    double *jv = malloc(sz * sizeof(double));
    double *a = calloc(sz, sizeof(double));
    double *eta = malloc((size_t)n * sizeof(double));
    int *rm = malloc((size_t)n * sizeof(int));

    sort_by_distance(n, ids, x, v, m);
    compute_eta(eta, n, m);
    cart_to_jacobi(jx, jv, n, x, v, m, eta);
    whfast_accel(a, n, method, x, jx, m, eta, G, eps);
    for (int q = 0; q < 3 * n; q++)
        jv[q] += a[q] * (0.5 * dt);

    const int64_t total = (int64_t)ceil(tf / dt);
    double t = 0.0;
    int64_t steps = 0;
    while (steps < total)
    {
        if (t + dt > tf)
            dt = tf - t;
        sort_by_distance(n, ids, jx, jv, m);
        compute_eta(eta, n, m);
        int nrm = 0;
        for (int i = 1; i < n; i++)
        {
            const double gm = G * m[0] * eta[i] / eta[i - 1];
            if (drift_one(&jx[i * 3], &jv[i * 3], gm, dt) && remove_invalid)
                rm[nrm++] = i;
        }
        if (nrm > 0)
        {
            n = compact(n, rm, nrm, ids, x, v, m, jx, jv);
            compute_eta(eta, n, m);
        }
        jacobi_to_cart(x, v, n, jx, jv, m, eta);
        whfast_accel(a, n, method, x, jx, m, eta, G, eps);
        for (int q = 0; q < 3 * n; q++)
            jv[q] += a[q] * dt;
        steps++;
        t = steps * dt;
    }
    *np = n;
    free(jx);
    free(jv);
    free(a);
    free(eta);
    free(rm);
    return 0;
}

/* ------------------------------------------------------------------ */
/* IAS15 (src/integrator_ias15.c:229-615 loop, :702-1382 pieces)       */
/* ------------------------------------------------------------------ */

/* state arrays are [slot][n*3] blocks */
#define IAS_AT(arr, slot, q, n3) ((arr)[(size_t)(slot) * (n3) + (q)])

static double orc_abs_max(const double *a, size_t len)
{
    double mx = fabs(a[0]);
    for (size_t q = 1; q < len; q++)
        mx = fmax(mx, fabs(a[q]));
    return mx;
}

/* position and velocity at the spacing h (src/integrator_ias15.c:808-919):
 * Horner forms of x0 + dt h (v0 + dt h (a0 + h (b0/3 + ... )) / 2), with
 * the compensation term added before the increment */
static void ias_predict(double *x1, double *v1, size_t n3, const double *x, const double *v,
                        const double *a, const double *b, double h, double dt,
                        const double *xc, const double *vc)
{
    for (size_t q = 0; q < n3; q++)
    {
        const double b0 = IAS_AT(b, 0, q, n3), b1 = IAS_AT(b, 1, q, n3), b2 = IAS_AT(b, 2, q, n3),
                     b3 = IAS_AT(b, 3, q, n3), b4 = IAS_AT(b, 4, q, n3), b5 = IAS_AT(b, 5, q, n3),
                     b6 = IAS_AT(b, 6, q, n3);
        double px = b5 / 28.0 + h * b6 / 36.0;
        px = b4 / 21.0 + h * px;
        px = b3 / 15.0 + h * px;
        px = b2 / 10.0 + h * px;
        px = b1 / 6.0 + h * px;
        px = b0 / 3.0 + h * px;
        px = a[q] + h * px;
        x1[q] = x[q];
        x1[q] += xc[q] + dt * h * (v[q] + dt * h * px / 2.0);

        double pv = b5 / 7.0 + h * b6 / 8.0;
        pv = b4 / 6.0 + h * pv;
        pv = b3 / 5.0 + h * pv;
        pv = b2 / 4.0 + h * pv;
        pv = b1 / 3.0 + h * pv;
        pv = b0 / 2.0 + h * pv;
        v1[q] = v[q];
        v1[q] += vc[q] + dt * h * (a[q] + h * pv);
    }
}

/* g and b after the force at spacing i (src/integrator_ias15.c:1117-1274
 * and :985-1115): g_k for k < i by the divided-difference recurrence, then
 * b_k for k < i from all seven g */
static void ias_update_gb(double *g, double *b, size_t n3, const double *F, int i,
                          const Ias15Coeffs *k)
{
    for (size_t q = 0; q < n3; q++)
    {
        for (int j = 0; j < i && j < 7; j++)
        {
            double t = (IAS_AT(F, j + 1, q, n3) - IAS_AT(F, 0, q, n3)) * k->r[j + 1][0];
            for (int l = 1; l <= j; l++)
                t = (t - IAS_AT(g, l - 1, q, n3)) * k->r[j + 1][l];
            IAS_AT(g, j, q, n3) = t;
        }
        for (int j = 0; j < i && j < 7; j++)
        {
            double s = k->c[j][j] * IAS_AT(g, j, q, n3);
            for (int l = j + 1; l < 7; l++)
                s = s + k->c[l][j] * IAS_AT(g, l, q, n3);
            IAS_AT(b, j, q, n3) = s;
        }
    }
}

/* the step's end point with compensated sums (src/integrator_ias15.c:921-983) */
static void ias_step(double *x1, double *v1, double *txc, double *tvc, size_t n3, const double *x,
                     const double *v, const double *a, const double *b, double dt)
{
    for (size_t q = 0; q < n3; q++)
    {
        const double b0 = IAS_AT(b, 0, q, n3), b1 = IAS_AT(b, 1, q, n3), b2 = IAS_AT(b, 2, q, n3),
                     b3 = IAS_AT(b, 3, q, n3), b4 = IAS_AT(b, 4, q, n3), b5 = IAS_AT(b, 5, q, n3),
                     b6 = IAS_AT(b, 6, q, n3);
        const double sx = a[q] + b0 / 3.0 + b1 / 6.0 + b2 / 10.0 + b3 / 15.0 + b4 / 21.0 +
                          b5 / 28.0 + b6 / 36.0;
        txc[q] += dt * (v[q] + dt * sx / 2.0);
        x1[q] = x[q] + txc[q];
        txc[q] += (x[q] - x1[q]);

        const double sv = a[q] + b0 / 2.0 + b1 / 3.0 + b2 / 4.0 + b3 / 5.0 + b4 / 6.0 +
                          b5 / 7.0 + b6 / 8.0;
        tvc[q] += dt * sv;
        v1[q] = v[q] + tvc[q];
        tvc[q] += (v[q] - v1[q]);
    }
}

/* b predicted for the next step from the ratio of step sizes, plus the
 * correction learned from the last prediction (src/integrator_ias15.c:1276-1382) */
static void ias_refine(double *b, double *e, size_t n3, double dt, double dt_new, int refine)
{
    const double q1 = dt_new / dt;
    const double q2 = q1 * q1, q3 = q2 * q1, q4 = q3 * q1, q5 = q4 * q1, q6 = q5 * q1,
                 q7 = q6 * q1;
    for (size_t q = 0; q < n3; q++)
    {
        double db[7];
        for (int j = 0; j < 7; j++)
            db[j] = refine ? IAS_AT(b, j, q, n3) - IAS_AT(e, j, q, n3) : 0.0;
        const double b0 = IAS_AT(b, 0, q, n3), b1 = IAS_AT(b, 1, q, n3), b2 = IAS_AT(b, 2, q, n3),
                     b3 = IAS_AT(b, 3, q, n3), b4 = IAS_AT(b, 4, q, n3), b5 = IAS_AT(b, 5, q, n3),
                     b6 = IAS_AT(b, 6, q, n3);
        double en[7];
        en[0] = q1 * (b6 * 7.0 + b5 * 6.0 + b4 * 5.0 + b3 * 4.0 + b2 * 3.0 + b1 * 2.0 + b0);
        en[1] = q2 * (b6 * 21.0 + b5 * 15.0 + b4 * 10.0 + b3 * 6.0 + b2 * 3.0 + b1);
        en[2] = q3 * (b6 * 35.0 + b5 * 20.0 + b4 * 10.0 + b3 * 4.0 + b2);
        en[3] = q4 * (b6 * 35.0 + b5 * 15.0 + b4 * 5.0 + b3);
        en[4] = q5 * (b6 * 21.0 + b5 * 6.0 + b4);
        en[5] = q6 * (b6 * 7.0 + b5);
        en[6] = q7 * b6;
        for (int j = 0; j < 7; j++)
        {
            IAS_AT(e, j, q, n3) = en[j];
            IAS_AT(b, j, q, n3) = en[j] + db[j];
        }
    }
}

/* Initial step size (src/integrator_ias15.c:702-806 ias15_initial_dt) */
static double ias_initial_dt(int n, const double *x, const double *v, const double *m, double G,
                             int method, double eps, const double *a)
{
    const size_t n3 = (size_t)n * 3;
    double *x1 = malloc(n3 * sizeof(double));
    double *a1 = malloc(n3 * sizeof(double));
    const double d0 = orc_abs_max(x, n3), d1 = orc_abs_max(a, n3);
    const double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
    for (size_t q = 0; q < n3; q++)
        x1[q] = x[q] + dt0 * v[q];
    accel(a1, n, x1, m, G, method, eps);
    for (size_t q = 0; q < n3; q++)
        a1[q] = a1[q] - a[q];
    const double d2 = orc_abs_max(a1, n3) / dt0;
    double dt1;
    if (fmax(d1, d2) <= 1e-15)
        dt1 = fmax(1e-6, dt0 * 1e-3);
    else
        dt1 = pow((0.01 / fmax(d1, d2)), 1.0 / (1.0 + (double)15));
    free(x1);
    free(a1);
    return fmin(100.0 * dt0, dt1);
}

int64_t orc_ias15(double *x, double *v, int n, const double *m, double G, int method, double eps,
                  double tol, double initial_dt, double tf)
{
    Ias15Coeffs K;
    ias15_coeffs(&K);
    const size_t n3 = (size_t)n * 3;
    double *a = malloc(n3 * sizeof(double));
    double *x1 = calloc(n3, sizeof(double)), *v1 = calloc(n3, sizeof(double));
    double *a1 = calloc(n3, sizeof(double));
    double *F = calloc(8 * n3, sizeof(double));
    double *g = calloc(7 * n3, sizeof(double)), *b = calloc(7 * n3, sizeof(double));
    double *b0 = calloc(7 * n3, sizeof(double)), *e = calloc(7 * n3, sizeof(double));
    double *xc = calloc(n3, sizeof(double)), *vc = calloc(n3, sizeof(double));
    double *txc = calloc(n3, sizeof(double)), *tvc = calloc(n3, sizeof(double));
    double *d7 = calloc(n3, sizeof(double));
    const double safety = 0.25, exponent = 1.0 / 7.0, tol_pc = 1e-16;

    accel(a, n, x, m, G, method, eps);
    double dt;
    if (initial_dt > 0.0)
        dt = initial_dt;
    else
    {
        dt = ias_initial_dt(n, x, v, m, G, method, eps, a);
        if (dt > tf)
            dt = tf;
    }
    double t = 0.0;
    int64_t steps = 0;
    int refine = 0;
    while (t < tf)
    {
        for (int it = 0; it < 12; it++)
        {
            for (int i = 0; i < 8; i++)
            {
                ias_predict(x1, v1, n3, x, v, a, b, K.h[i], dt, xc, vc);
                accel(F + (size_t)i * n3, n, x1, m, G, method, eps);
                ias_update_gb(g, b, n3, F, i, &K);
            }
            for (size_t q = 0; q < n3; q++)
                d7[q] = IAS_AT(b, 6, q, n3) - IAS_AT(b0, 6, q, n3);
            memcpy(b0, b, 7 * n3 * sizeof(double));
            if (orc_abs_max(d7, n3) / orc_abs_max(F + 7 * n3, n3) < tol_pc)
                break;
        }
        memcpy(txc, xc, n3 * sizeof(double));
        memcpy(tvc, vc, n3 * sizeof(double));
        ias_step(x1, v1, txc, tvc, n3, x, v, a, b, dt);
        accel(a1, n, x1, m, G, method, eps);

        const double err_b7 = orc_abs_max(b + 6 * n3, n3) / orc_abs_max(a1, n3);
        double err = pow((err_b7 / tol), exponent);
        if (err < 1e-10)
            err = 1e-10;
        const double dt_new = dt / err;
        if (err <= 1.0 || dt == tf * 1e-12)
        {
            steps++;
            t += dt;
            ias_refine(b, e, n3, dt, dt_new, refine);
            refine = 1;
            memcpy(xc, txc, n3 * sizeof(double));
            memcpy(vc, tvc, n3 * sizeof(double));
            memcpy(x, x1, n3 * sizeof(double));
            memcpy(v, v1, n3 * sizeof(double));
            memcpy(a, a1, n3 * sizeof(double));
        }
        if (dt_new > (dt / safety))
            dt /= safety;
        else if (dt_new < dt * safety)
            dt *= safety;
        else
            dt = dt_new;
        if (dt_new < tf * 1e-12)
            dt = tf * 1e-12;
        if ((t < tf) && (t + dt > tf))
            dt = tf - t;
    }
    free(a);
    free(x1);
    free(v1);
    free(a1);
    free(F);
    free(g);
    free(b);
    free(b0);
    free(e);
    free(xc);
    free(vc);
    free(txc);
    free(tvc);
    free(d7);
    return steps;
}

/* ------------------------------------------------------------------ */
/* Euler, Euler-Cromer, RK4 (src/integrator.c:276-892)                */
/* ------------------------------------------------------------------ */

/* src/integrator.c:276-512 (cromer = 0) and :514-630 (cromer = 1): fixed dt
 * with the last-step clamp, compensated sums for x and v */
int64_t orc_euler(double *x, double *v, int n, const double *m, double G, int method, double eps,
                  double dt, double tf, int cromer)
{
    const size_t sz = (size_t)n * 3;
    double *a = malloc(sz * sizeof(double));
    double *xe = calloc(sz, sizeof(double));
    double *ve = calloc(sz, sizeof(double));
    const int64_t total = (int64_t)ceil(tf / dt);
    double t = 0.0;
    int64_t steps = 0;
    while (steps < total)
    {
        if (t + dt > tf)
            dt = tf - t;
        accel(a, n, x, m, G, method, eps);
        for (size_t q = 0; q < sz; q++)
        {
            const double x0 = x[q], v0 = v[q];
            if (!cromer)
            {
                xe[q] += v[q] * dt;
                ve[q] += a[q] * dt;
                x[q] = x0 + xe[q];
                v[q] = v0 + ve[q];
                xe[q] += x0 - x[q];
                ve[q] += v0 - v[q];
            }
            else
            {
                ve[q] += a[q] * dt;
                v[q] = v0 + ve[q];
                ve[q] += v0 - v[q];
                xe[q] += v[q] * dt;
                x[q] = x0 + xe[q];
                xe[q] += x0 - x[q];
            }
        }
        steps++;
        t = steps * dt;
    }
    free(a);
    free(xe);
    free(ve);
    return steps;
}

/* src/integrator.c:632-892: classic RK4 (no last-step clamp), compensated
 * sums in the final update */
int64_t orc_rk4(double *x, double *v, int n, const double *m, double G, int method, double eps,
                double dt, double tf)
{
    const size_t sz = (size_t)n * 3;
    double *x0 = malloc(sz * sizeof(double)), *v0 = malloc(sz * sizeof(double));
    double *vk = malloc(4 * sz * sizeof(double)), *xk = malloc(4 * sz * sizeof(double));
    double *xe = calloc(sz, sizeof(double)), *ve = calloc(sz, sizeof(double));
    const int64_t total = (int64_t)ceil(tf / dt);
    int64_t steps = 0;
    while (steps < total)
    {
        memcpy(x0, x, sz * sizeof(double));
        memcpy(v0, v, sz * sizeof(double));
        for (int k = 0; k < 4; k++)
        {
            if (k > 0)
            {
                /* stages 2, 3: half steps from k-1; stage 4: a full step */
                const double *xp = xk + (size_t)(k - 1) * sz, *vp = vk + (size_t)(k - 1) * sz;
                for (size_t q = 0; q < sz; q++)
                {
                    if (k < 3)
                    {
                        x[q] = x0[q] + 0.5 * xp[q] * dt;
                        v[q] = v0[q] + 0.5 * vp[q] * dt;
                    }
                    else
                    {
                        x[q] = x0[q] + xp[q] * dt;
                        v[q] = v0[q] + vp[q] * dt;
                    }
                }
            }
            accel(vk + (size_t)k * sz, n, x, m, G, method, eps);
            memcpy(xk + (size_t)k * sz, v, sz * sizeof(double));
        }
        for (size_t q = 0; q < sz; q++)
        {
            const double *V = vk, *X = xk;
            ve[q] += (V[q] + 2 * V[sz + q] + 2 * V[2 * sz + q] + V[3 * sz + q]) * dt / 6.0;
            xe[q] += (X[q] + 2 * X[sz + q] + 2 * X[2 * sz + q] + X[3 * sz + q]) * dt / 6.0;
            v[q] = v0[q] + ve[q];
            x[q] = x0[q] + xe[q];
            ve[q] += v0[q] - v[q];
            xe[q] += x0[q] - x[q];
        }
        steps++;
    }
    free(x0);
    free(v0);
    free(vk);
    free(xk);
    free(xe);
    free(ve);
    return steps;
}

/* ------------------------------------------------------------------ */
/* Embedded Runge-Kutta (src/integrator_rk_embedded.c:346-987)        */
/* ------------------------------------------------------------------ */

/* src/integrator_rk_embedded.c:346-520 (the else-less branch at :495-502
 * makes the power law the only formula) */
static double rke_initial_dt(int n, const double *x, const double *v, const double *m, double G,
                             int method, double eps, double tol, int power)
{
    const size_t sz = (size_t)n * 3;
    double *a = malloc(sz * sizeof(double)), *a1 = malloc(sz * sizeof(double));
    double *x1 = malloc(sz * sizeof(double)), *v1 = malloc(sz * sizeof(double));
    double *tx = malloc(sz * sizeof(double)), *tv = malloc(sz * sizeof(double));
    accel(a, n, x, m, G, method, eps);
    for (size_t q = 0; q < sz; q++)
    {
        tx[q] = tol + tol * fabs(x[q]);
        tv[q] = tol + tol * fabs(v[q]);
    }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (size_t q = 0; q < sz; q++)
    {
        s0 += (x[q] / tx[q]) * (x[q] / tx[q]);
        s0 += (v[q] / tv[q]) * (v[q] / tv[q]);
        s1 += (v[q] / tx[q]) * (v[q] / tx[q]);
        s1 += (a[q] / tv[q]) * (a[q] / tv[q]);
    }
    const double d0 = sqrt(s0 / (n * 6)), d1 = sqrt(s1 / (n * 6));
    const double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-4 : d0 / d1;
    for (size_t q = 0; q < sz; q++)
    {
        x1[q] = x[q] + (dt0 / 100.0) * v[q];
        v1[q] = v[q] + (dt0 / 100.0) * a[q];
    }
    accel(a1, n, x1, m, G, method, eps);
    for (size_t q = 0; q < sz; q++)
    {
        s2 += ((v1[q] - v[q]) / tx[q]) * ((v1[q] - v[q]) / tx[q]);
        s2 += ((a1[q] - a[q]) / tv[q]) * ((a1[q] - a[q]) / tv[q]);
    }
    const double d2 = sqrt(s2 / (n * 6)) / dt0;
    const double dt1 = pow((double)(0.01L / fmax(d1, d2)), (double)(1.0L / (1 + power)));
    const double dt = fmin((double)(100.0L * dt0), dt1);
    free(a);
    free(a1);
    free(x1);
    free(v1);
    free(tx);
    free(tv);
    return dt * 1e-2;
}

int64_t orc_rk_embedded(double *x, double *v, int n, const double *m, double G, int method,
                        double eps, int integrator, double tol, double initial_dt, double tf)
{
    RkTableau T;
    if (rk_tableau(integrator, &T) != 0)
        return -1;
    const int S = T.stages;
    const int min_power = T.power < T.power_test ? T.power : T.power_test;
    double dw[RK_MAX_STAGES];
    for (int s = 0; s < S; s++)
        dw[s] = T.weights[s] - T.weights_test[s];
    const double safety_max = 6.0, safety_min = 0.33;
    const double safety = pow(0.38, (1.0 / (1.0 + (double)min_power)));
    const size_t sz = (size_t)n * 3;
    double *x1 = malloc(sz * sizeof(double)), *v1 = malloc(sz * sizeof(double));
    double *vk = malloc((size_t)S * sz * sizeof(double)), *xk = malloc((size_t)S * sz * sizeof(double));
    double *tv = malloc(sz * sizeof(double)), *tx = malloc(sz * sizeof(double));
    double *dv = malloc(sz * sizeof(double)), *dx = malloc(sz * sizeof(double));
    double *xe = calloc(sz, sizeof(double)), *ve = calloc(sz, sizeof(double));
    double *txe = calloc(sz, sizeof(double)), *tve = calloc(sz, sizeof(double));

    double dt;
    if (initial_dt > 0.0)
        dt = initial_dt;
    else
    {
        dt = rke_initial_dt(n, x, v, m, G, method, eps, tol, T.power);
        if (dt > tf)
            dt = tf;
    }
    double t = 0.0;
    int64_t steps = 0;
    while (t < tf)
    {
        /* stages (src/integrator_rk_embedded.c:737-783) */
        accel(vk, n, x, m, G, method, eps);
        memcpy(xk, v, sz * sizeof(double));
        for (int s = 1; s < S; s++)
        {
            for (size_t q = 0; q < sz; q++)
            {
                double sv = 0.0, sx = 0.0;
                for (int i = 0; i < s; i++)
                {
                    sv += T.coeff[(s - 1) * (S - 1) + i] * vk[(size_t)i * sz + q];
                    sx += T.coeff[(s - 1) * (S - 1) + i] * xk[(size_t)i * sz + q];
                }
                tv[q] = v[q] + dt * sv + ve[q];
                tx[q] = x[q] + dt * sx + xe[q];
            }
            accel(vk + (size_t)s * sz, n, tx, m, G, method, eps);
            memcpy(xk + (size_t)s * sz, tv, sz * sizeof(double));
        }
        /* weights, error estimate and compensated update (:800-829) */
        for (size_t q = 0; q < sz; q++)
        {
            double sv = 0.0, sx = 0.0, ev = 0.0, ex = 0.0;
            for (int s = 0; s < S; s++)
            {
                sv += T.weights[s] * vk[(size_t)s * sz + q];
                sx += T.weights[s] * xk[(size_t)s * sz + q];
                ev += dt * dw[s] * vk[(size_t)s * sz + q];
                ex += dt * dw[s] * xk[(size_t)s * sz + q];
            }
            dv[q] = ev;
            dx[q] = ex;
            tve[q] = ve[q] + dt * sv;
            txe[q] = xe[q] + dt * sx;
            v1[q] = v[q] + tve[q];
            x1[q] = x[q] + txe[q];
            tve[q] += v[q] - v1[q];
            txe[q] += x[q] - x1[q];
        }
        /* error norm (:831-853), summed in index order */
        double sum = 0.0;
        for (size_t q = 0; q < sz; q++)
        {
            const double sv = tol + fmax(fabs(v[q]), fabs(v1[q])) * tol;
            const double sx = tol + fmax(fabs(x[q]), fabs(x1[q])) * tol;
            double e = dv[q] / sv;
            sum += e * e;
            e = dx[q] / sx;
            sum += e * e;
        }
        double error = sqrt(sum / (n * 3 * 2));
        if (error <= 1 || dt <= tf * 1e-12)
        {
            steps++;
            t += dt;
            memcpy(x, x1, sz * sizeof(double));
            memcpy(v, v1, sz * sizeof(double));
            memcpy(xe, txe, sz * sizeof(double));
            memcpy(ve, tve, sz * sizeof(double));
        }
        /* step size (:887-914) */
        if (error < 1e-10)
            error = 1e-10;
        const double dt_new = dt * safety / pow(error, (1.0 / (1.0 + (double)min_power)));
        if (dt_new > safety_max * dt)
            dt *= safety_max;
        else if (dt_new < safety_min * dt)
            dt *= safety_min;
        else
            dt = dt_new;
        if (dt_new / tf < 1e-12)
            dt = tf * 1e-12;
        if (t < tf && t + dt > tf)
            dt = tf - t;
    }
    free(x1);
    free(v1);
    free(vk);
    free(xk);
    free(tv);
    free(tx);
    free(dv);
    free(dx);
    free(xe);
    free(ve);
    free(txe);
    free(tve);
    return steps;
}
