// python_iface.cpp -- the remaining ctypes entry points the grav_sim Python
// package binds at start-up (grav_sim/utils.py:38-58,
// src/python_interface.c:8-92,245-304).  Host-side and off the hot path;
// load_built_in_system_python is in systems.cpp.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "gs_internal.h"

extern "C" {

void free_memory_int32(int32 *ptr) { free(ptr); }

void free_memory_double(double *ptr) { free(ptr); }

// Orbital elements -> Cartesian state (the formulae of src/utils.c:82-141).
void keplerian_to_cartesian(double *x, double *v, const double a, const double e,
                            const double inc, const double argp, const double node,
                            const double f, const double total_mass, const double G)
{
    const double ci = cos(inc), si = sin(inc);
    const double cw = cos(argp), sw = sin(argp);
    const double cn = cos(node), sn = sin(node);
    const double cf = cos(f), sf = sin(f);
    const double p[3] = {cn * cw - sn * sw * ci, sn * cw + cn * sw * ci, sw * si};
    const double q[3] = {-cn * sw - sn * cw * ci, -sn * sw + cn * cw * ci, cw * si};
    const double r = a * (1.0 - e * e);
    const double h = sqrt(G * total_mass / (a * (1.0 - e * e)));
    for (int d = 0; d < 3; d++) {
        x[d] = r * (cf * p[d] + sf * q[d]) / (1.0 + e * cf);
        v[d] = h * (-sf * p[d] + (e + cf) * q[d]);
    }
}

void keplerian_to_cartesian_python(double *x, double *y, double *z, double *v_x, double *v_y,
                                   double *v_z, const double a, const double e, const double inc,
                                   const double argp, const double node, const double f,
                                   const double total_mass, const double G)
{
    double xx[3], vv[3];
    keplerian_to_cartesian(xx, vv, a, e, inc, argp, node, f, total_mass, G);
    *x = xx[0];
    *y = xx[1];
    *z = xx[2];
    *v_x = vv[0];
    *v_y = vv[1];
    *v_z = vv[2];
}

// Snapshot rows are (m, x, y, z, vx, vy, vz).
void compute_linear_momentum_python(double *out, const double *sol, const int32 ns,
                                    const int32 n)
{
    for (int s = 0; s < ns; s++) {
        const double *r = sol + static_cast<size_t>(s) * n * 7;
        double p = 0.0;
        for (int i = 0; i < n; i++) p += r[7 * i] * (r[7 * i + 4] + r[7 * i + 5] + r[7 * i + 6]);
        out[s] = p;
    }
}

void compute_angular_momentum_python(double *out, const double *sol, const int32 ns,
                                     const int32 n)
{
    for (int s = 0; s < ns; s++) {
        const double *r = sol + static_cast<size_t>(s) * n * 7;
        double L = 0.0;
        for (int i = 0; i < n; i++) {
            const double *q = r + 7 * i;
            const double lx = q[2] * q[6] - q[3] * q[5];
            const double ly = q[3] * q[4] - q[1] * q[6];
            const double lz = q[1] * q[5] - q[2] * q[4];
            L += q[0] * (lx + ly + lz);
        }
        out[s] = L;
    }
}

}  // extern "C"
