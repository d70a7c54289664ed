// systems.cpp -- the named initial conditions of the grav_sim library
// (initialize_built_in_system, src/system.c:579-1095), which its Python
// package loads through load_built_in_system_python
// (src/python_interface.c:18-53).
//
// The data are physical constants, not code: GM values of DE440 (Park et al.
// 2021, AJ 161, 105) and the barycentric states of the Sun and planets at
// 2024-01-01 00:00 TDB in AU and AU/day (JPL Horizons), plus the classic
// three-body configurations (figure-8 of Chenciner & Montgomery, the
// Pythagorean problem of Burrau).  They are given with the reference's
// digits and go through the same double arithmetic (long-double literals
// rounded to double, masses GM / GM_sun, positions and velocities relative
// to the centre of mass summed in body order), so every double produced is
// the reference's.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "gs_internal.h"

namespace {

struct Body {
    long double gm;  // km^3 s^-2
    long double x[3], v[3];
};

enum { SUN, MERCURY, VENUS, EARTH, MARS, JUPITER, SATURN, URANUS, NEPTUNE, MOON, PLUTO, CERES,
       VESTA, NUM_BODIES };

const Body kBodies[NUM_BODIES] = {
    {132712440041.279419L,
     {-7.967955691533730e-03L, -2.906227441573178e-03L, 2.103054301547123e-04L},
     {4.875094764261564e-06L, -7.057133213976680e-06L, -4.573453713094512e-08L}},
    {22031.868551L,
     {-2.825983269538632e-01L, 1.974559795958082e-01L, 4.177433558063677e-02L},
     {-2.232165900189702e-02L, -2.157207103176252e-02L, 2.855193410495743e-04L}},
    {324858.592000L,
     {-7.232103701666379e-01L, -7.948302026312400e-02L, 4.042871428174315e-02L},
     {2.034068201002341e-03L, -2.020828626592994e-02L, -3.945639843855159e-04L}},
    {398600.435507L,
     {-1.738192017257054e-01L, 9.663245550235138e-01L, 1.553901854897183e-04L},
     {-1.723001232538228e-02L, -2.967721342618870e-03L, 6.382125383116755e-07L}},
    {42828.375816L,
     {-3.013262392582653e-01L, -1.454029331393295e00L, -2.300531433991428e-02L},
     {1.424832259345280e-02L, -1.579236181580905e-03L, -3.823722796161561e-04L}},
    {126712764.100000L,
     {3.485202469657674e00L, 3.552136904413157e00L, -9.271035442798399e-02L},
     {-5.470970658852281e-03L, 5.642487338479145e-03L, 9.896190602066252e-05L}},
    {37940584.841800L,
     {8.988104223143450e00L, -3.719064854634689e00L, -2.931937777323593e-01L},
     {1.822013845554067e-03L, 5.143470425888054e-03L, -1.617235904887937e-04L}},
    {5794556.400000L,
     {1.226302417897505e01L, 1.529738792480545e01L, -1.020549026883563e-01L},
     {-3.097615358317413e-03L, 2.276781932345769e-03L, 4.860433222241686e-05L}},
    {6836527.100580L,
     {2.983501460984741e01L, -1.793812957956852e00L, -6.506401132254588e-01L},
     {1.676536611817232e-04L, 3.152098732861913e-03L, -6.877501095688201e-05L}},
    {4902.800118L,
     {-1.762788124769829e-01L, 9.674377513177153e-01L, 3.236901585768862e-04L},
     {-1.746667306153906e-02L, -3.473438277358121e-03L, -3.359028758606074e-05L}},
    {975.500000L,
     {1.720200478843485e01L, -3.034155683573043e01L, -1.729127607100611e00L},
     {2.802810313667557e-03L, 8.492056438614633e-04L, -9.060790113327894e-04L}},
    {62.62890L,
     {-1.103880510367569e00L, -2.533340440444230e00L, 1.220283937721780e-01L},
     {8.978653480111301e-03L, -4.873256528198994e-03L, -1.807162046049230e-03L}},
    {17.288245L,
     {-8.092549658731499e-02L, 2.558381434460076e00L, -6.695836142398572e-02L},
     {-1.017876585480054e-02L, -5.452367109338154e-04L, 1.255870551153315e-03L}},
};

// Bodies from the table, masses in solar masses, shifted to their own
// centre-of-mass frame.
void ephemeris_system(System *s, const int *bodies, int k)
{
    const double gm_sun = static_cast<double>(kBodies[SUN].gm);
    double M = 0.0;
    for (int b = 0; b < k; b++) {
        // MASS_SUN is the literal 1.0, the others GM / GM_sun
        s->m[b] = bodies[b] == SUN ? 1.0 : static_cast<double>(kBodies[bodies[b]].gm) / gm_sun;
        M += s->m[b];
    }
    for (int d = 0; d < 3; d++) {
        double rx = 0.0, rv = 0.0;
        for (int b = 0; b < k; b++) {
            rx += s->m[b] * static_cast<double>(kBodies[bodies[b]].x[d]);
            rv += s->m[b] * static_cast<double>(kBodies[bodies[b]].v[d]);
        }
        const double R = 1 / M * rx, V = 1 / M * rv;
        for (int b = 0; b < k; b++) {
            s->x[3 * b + d] = static_cast<double>(kBodies[bodies[b]].x[d]) - R;
            s->v[3 * b + d] = static_cast<double>(kBodies[bodies[b]].v[d]) - V;
        }
    }
}

struct Named {
    const char *name;
    int n;
};

const Named kSystems[] = {
    {"circular_binary_orbit", 2}, {"eccentric_binary_orbit", 2}, {"3d_helix", 3},
    {"sun_earth_moon", 3},        {"figure-8", 3},               {"pyth-3-body", 3},
    {"solar_system", 9},          {"solar_system_plus", 12},
};

void set3(double *a, double p, double q, double r)
{
    a[0] = p;
    a[1] = q;
    a[2] = r;
}

void fill(System *s, int which)
{
    const double G = s->G;
    double *x = s->x, *v = s->v, *m = s->m;
    switch (which) {
        case 0:  // equal masses on a circle of radius 1
            set3(x, 1.0, 0.0, 0.0), set3(x + 3, -1.0, 0.0, 0.0);
            set3(v, 0.0, 0.5, 0.0), set3(v + 3, 0.0, -0.5, 0.0);
            m[0] = 1.0 / G, m[1] = 1.0 / G;
            break;
        case 1:  // mass ratio 0.8
            set3(x, 1.0, 0.0, 0.0), set3(x + 3, -1.25, 0.0, 0.0);
            set3(v, 0.0, 0.5, 0.0), set3(v + 3, 0.0, -0.625, 0.0);
            m[0] = 1.0 / G, m[1] = 0.8 / G;
            break;
        case 2: {  // three equal masses on a helix
            const double h = sqrt(3.0) / 2.0, v0 = sqrt(1.0 / sqrt(3));
            set3(x, 0.0, 0.0, -1.0), set3(x + 3, -h, 0.0, 0.5), set3(x + 6, h, 0.0, 0.5);
            set3(v, -v0, 0.5, 0.0), set3(v + 3, 0.5 * v0, 0.5, h * v0);
            set3(v + 6, 0.5 * v0, 0.5, -h * v0);
            m[0] = m[1] = m[2] = 1.0 / G;
            break;
        }
        case 3: {
            const int b[3] = {SUN, EARTH, MOON};
            ephemeris_system(s, b, 3);
            break;
        }
        case 4:  // figure-8 choreography
            set3(x, 0.970043, -0.24308753, 0.0), set3(x + 3, -0.970043, 0.24308753, 0.0);
            set3(x + 6, 0.0, 0.0, 0.0);
            set3(v, 0.466203685, 0.43236573, 0.0), set3(v + 3, 0.466203685, 0.43236573, 0.0);
            set3(v + 6, -0.93240737, -0.86473146, 0.0);
            m[0] = m[1] = m[2] = 1.0 / G;
            break;
        case 5:  // Pythagorean problem: masses 3, 4, 5 at rest
            set3(x, 1.0, 3.0, 0.0), set3(x + 3, -2.0, -1.0, 0.0), set3(x + 6, 1.0, -1.0, 0.0);
            set3(v, 0.0, 0.0, 0.0), set3(v + 3, 0.0, 0.0, 0.0), set3(v + 6, 0.0, 0.0, 0.0);
            m[0] = 3.0 / G, m[1] = 4.0 / G, m[2] = 5.0 / G;
            break;
        case 6: {
            const int b[9] = {SUN, MERCURY, VENUS, EARTH, MARS, JUPITER, SATURN, URANUS, NEPTUNE};
            ephemeris_system(s, b, 9);
            break;
        }
        case 7: {
            const int b[12] = {SUN,    MERCURY, VENUS,   EARTH, MARS,  JUPITER,
                               SATURN, URANUS,  NEPTUNE, PLUTO, CERES, VESTA};
            ephemeris_system(s, b, 12);
            break;
        }
    }
}

}  // namespace

extern "C" {

// src/system.c:579-1095 (same names, sizes, errors; memory from
// get_initialized_system unless the caller's system is already allocated)
ErrorStatus initialize_built_in_system(System *system, const char *system_name,
                                       const bool is_memory_initialized)
{
    if (!system) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System is NULL");
    if (!system_name) return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System name is NULL");
    int which = -1;
    for (int i = 0; i < static_cast<int>(sizeof(kSystems) / sizeof(kSystems[0])); i++)
        if (strcmp(system_name, kSystems[i].name) == 0) which = i;
    if (which < 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "System name not recognized. Got: \"%s\".",
                                    system_name);
    const int n = kSystems[which].n;
    if (!is_memory_initialized) {
        GS_TRY(get_initialized_system(system, n));
    } else if (system->num_particles < n) {
        return WRAP_RAISE_ERROR(GRAV_VALUE_ERROR,
                                "Initialized system is not big enough for the built-in system");
    }
    fill(system, which);
    return make_success_error_status();
}

// src/python_interface.c:18-53
int32 load_built_in_system_python(const char *system_name, int *num_particles_ptr,
                                  int32 **particle_ids_ptr, double **x_ptr, double **v_ptr,
                                  double **m_ptr, double *G_ptr)
{
    System s = get_new_system();
    ErrorStatus st = initialize_built_in_system(&s, system_name, false);
    if (st.return_code != GRAV_SUCCESS) {
        st = WRAP_TRACEBACK(st);
        print_and_free_traceback(&st);
        return st.return_code;
    }
    *num_particles_ptr = s.num_particles;
    *particle_ids_ptr = s.particle_ids;
    *x_ptr = s.x;
    *v_ptr = s.v;
    *m_ptr = s.m;
    *G_ptr = s.G;
    return GRAV_SUCCESS;
}

}  // extern "C"
