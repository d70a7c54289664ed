// params.cpp -- parameter structs, defaults and validation with the
// reference's defaults, messages and error codes.
#include <sys/stat.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <ctime>

#include "gs_internal.h"

extern "C" {

// src/acceleration.c:62-71
AccelerationParam get_new_acceleration_param(void)
{
    AccelerationParam p;
    p.method = ACCELERATION_METHOD_PAIRWISE;
    p.opening_angle = 1.0;
    p.softening_length = 0.0;
    p.max_num_particles_per_leaf = -1;
    return p;
}

static bool is_bh(int method)
{
    return method == ACCELERATION_METHOD_BARNES_HUT ||
           method == ACCELERATION_METHOD_BARNES_HUT_LEARNED ||
           method == ACCELERATION_METHOD_BARNES_HUT_QUADRUPOLE ||
           method == ACCELERATION_METHOD_BARNES_HUT_GROUP ||
           method == ACCELERATION_METHOD_BARNES_HUT_GROUP_QUADRUPOLE ||
           method == ACCELERATION_METHOD_BARNES_HUT_GROUP_EXPANSION;
}

static bool known_method(int method)
{
    switch (method) {
        case ACCELERATION_METHOD_PAIRWISE:
        case ACCELERATION_METHOD_MASSLESS:
        case ACCELERATION_METHOD_BARNES_HUT:
        case ACCELERATION_METHOD_PAIRWISE_FP32:
        case ACCELERATION_METHOD_MASSLESS_FP32:
        case ACCELERATION_METHOD_BARNES_HUT_LEARNED:
        case ACCELERATION_METHOD_BARNES_HUT_QUADRUPOLE:
        case ACCELERATION_METHOD_BARNES_HUT_GROUP:
        case ACCELERATION_METHOD_BARNES_HUT_GROUP_QUADRUPOLE:
        case ACCELERATION_METHOD_BARNES_HUT_GROUP_EXPANSION:
            return true;
        default:
            return false;
    }
}

// src/acceleration.c:73-129 (the additional GPU methods follow the same rules;
// Barnes-Hut variants share the opening-angle and leaf checks)
ErrorStatus finalize_acceleration_param(AccelerationParam *p)
{
    if (!p)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "Acceleration parameter pointer is NULL");
    if (!known_method(p->method))
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Unknown acceleration method. Got: %d",
                                    p->method);
    if (p->softening_length < 0.0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Softening length is negative. Got: %.3g",
                                    p->softening_length);
    if (is_bh(p->method) && p->opening_angle < 0.0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Opening angle is negative. Got: %.3g",
                                    p->opening_angle);
    if (is_bh(p->method)) {
        if (p->max_num_particles_per_leaf == -1)
            p->max_num_particles_per_leaf = 1;
        else if (p->max_num_particles_per_leaf < 1)
            return WRAP_RAISE_ERROR_FMT(
                GRAV_VALUE_ERROR,
                "Maximum number of particles per leaf must be positive. Got: %d",
                p->max_num_particles_per_leaf);
    }
    return make_success_error_status();
}

// src/integrator.c:119-129
IntegratorParam get_new_integrator_param(void)
{
    IntegratorParam p;
    p.integrator = -1;
    p.dt = -1.0;
    p.tolerance = -1.0;
    p.initial_dt = -1.0;
    p.whfast_remove_invalid_particles = true;
    return p;
}

// src/integrator.c:131-189
ErrorStatus finalize_integration_param(IntegratorParam *p)
{
    if (!p)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "integration_param is NULL");
    const int it = p->integrator;
    if (it < INTEGRATOR_EULER || it > INTEGRATOR_WHFAST)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Unknown integrator. Got: %d", it);
    const bool fixed = it == INTEGRATOR_EULER || it == INTEGRATOR_EULER_CROMER ||
                       it == INTEGRATOR_RK4 || it == INTEGRATOR_LEAPFROG ||
                       it == INTEGRATOR_WHFAST;
    if (fixed) {
        if (p->dt <= 0.0)
            return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "dt must be positive. Got: %g", p->dt);
    } else if (p->tolerance <= 0.0) {
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "tolerance must be positive. Got: %g",
                                    p->tolerance);
    }
    return make_success_error_status();
}

// src/output.c:85-98
OutputParam get_new_output_param(void)
{
    OutputParam p;
    p.method = OUTPUT_METHOD_DISABLED;
    p.output_dir = nullptr;
    p.output_initial = false;
    p.output_interval = -1.0;
    p.coordinate_output_dtype = OUTPUT_DTYPE_DOUBLE;
    p.velocity_output_dtype = OUTPUT_DTYPE_DOUBLE;
    p.mass_output_dtype = OUTPUT_DTYPE_DOUBLE;
    p.output_count_ = 0;
    return p;
}

// src/output.c:100-245: method check (HDF5 is not built in, as in a
// reference build without HDF5), positive interval, a directory path that
// ends in '/' (snapshots_YYYYMMDD_HHMMSS/ when none is given), created here;
// an existing directory gives a warning.
ErrorStatus finalize_output_param(OutputParam *p, const Settings *settings)
{
    if (!p)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "Output parameter pointer is NULL");
    switch (p->method) {
        case OUTPUT_METHOD_DISABLED:
            return make_success_error_status();
        case OUTPUT_METHOD_CSV:
            break;
        case OUTPUT_METHOD_HDF5:
            return WRAP_RAISE_ERROR(GRAV_MEMORY_ERROR, "HDF5 output method is not available");
        default:
            return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Unknown output method. Got: %d",
                                        p->method);
    }
    if (p->output_interval <= 0.0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR, "Output interval must be positive. Got: %.17g",
                                    p->output_interval);
    if (!p->output_dir) {
        const size_t len = strlen("snapshots_YYYYMMDD_HHMMSS/") + 1;
        char *dir = static_cast<char *>(malloc(len));
        if (!dir)
            return WRAP_RAISE_ERROR(GRAV_MEMORY_ERROR,
                                    "Failed to allocate memory for directory path.");
        const time_t now = time(nullptr);
        strftime(dir, len, "snapshots_%Y%m%d_%H%M%S/", localtime(&now));
        p->output_dir = dir;
    } else {
        const size_t len = strlen(p->output_dir);
        if (len == 0 || p->output_dir[len - 1] != '/')
            return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                        "Directory path for storing snapshots must end with a "
                                        "trailing slash (\"/\"). Got: \"%s\".",
                                        p->output_dir);
    }
    if (mkdir(p->output_dir, 0777) == -1) {
        struct stat st;
        if (stat(p->output_dir, &st) != 0)
            return WRAP_RAISE_ERROR_FMT(GRAV_OS_ERROR,
                                        "Failed to access path for storing snapshots: \"%s\".",
                                        p->output_dir);
        if ((st.st_mode & S_IFDIR) && settings && settings->verbose >= GRAV_VERBOSITY_IGNORE_INFO)
            GS_TRY(WRAP_RAISE_WARNING_FMT("Directory for storing snapshots already exists. The "
                                          "files will be overwritten. Directory: \"%s\".",
                                          p->output_dir));
    }
    return make_success_error_status();
}

// src/settings.c:5-14: every Settings shares one static exit flag
Settings get_new_settings(void)
{
    static bool is_exit = false;
    Settings s;
    s.verbose = GRAV_VERBOSITY_NORMAL;
    s.enable_progress_bar = true;
    s.is_exit_ptr = &is_exit;
    return s;
}

// src/system.c:26-36
System get_new_system(void)
{
    System s;
    s.num_particles = 0;
    s.particle_ids = nullptr;
    s.x = nullptr;
    s.v = nullptr;
    s.m = nullptr;
    s.G = 0.000295912208284119496676630;  // AU^3 day^-2 M_sun^-1
    return s;
}

void free_system(System *s)
{
    free(s->particle_ids);
    free(s->x);
    free(s->v);
    free(s->m);
}

// src/system.c:38-68
ErrorStatus get_initialized_system(System *s, const int n)
{
    if (!s)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System is NULL");
    *s = get_new_system();
    s->num_particles = n;
    s->particle_ids = static_cast<int *>(malloc(static_cast<size_t>(n) * sizeof(int)));
    s->x = static_cast<double *>(calloc(static_cast<size_t>(n) * 3, sizeof(double)));
    s->v = static_cast<double *>(calloc(static_cast<size_t>(n) * 3, sizeof(double)));
    s->m = static_cast<double *>(calloc(static_cast<size_t>(n), sizeof(double)));
    if (!s->particle_ids || !s->x || !s->v || !s->m) {
        free_system(s);
        return WRAP_RAISE_ERROR(GRAV_MEMORY_ERROR, "Failed to allocate memory for system");
    }
    for (int i = 0; i < n; i++)
        s->particle_ids[i] = i;
    return make_success_error_status();
}

// src/system.c:70-111
ErrorStatus finalize_system(System *s)
{
    if (!s)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System is NULL");
    if (s->num_particles <= 0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Number of particles must be positive. Got: %d",
                                    s->num_particles);
    if (!s->particle_ids)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System array particle_ids is NULL");
    if (!s->x)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System array x is NULL");
    if (!s->v)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System array v is NULL");
    if (!s->m)
        return WRAP_RAISE_ERROR(GRAV_POINTER_ERROR, "System array m is NULL");
    if (s->G <= 0.0)
        return WRAP_RAISE_ERROR_FMT(GRAV_VALUE_ERROR,
                                    "Gravitational constant G must be positive. Got: %g", s->G);
    return make_success_error_status();
}

// src/linear_octree.c:90-103, with every pointer NULL
LinearOctree get_new_linear_octree(void)
{
    LinearOctree t;
    t.box_width = 0.0;
    t.num_internal_nodes = 0;
    t.particle_morton_indices_deepest_level = nullptr;
    t.sorted_indices = nullptr;
    t.tree_num_particles = nullptr;
    t.tree_num_internal_children = nullptr;
    t.tree_first_particle_sorted_idx = nullptr;
    t.tree_first_internal_children_idx = nullptr;
    t.tree_mass = nullptr;
    t.tree_center_of_mass_x = nullptr;
    t.tree_center_of_mass_y = nullptr;
    t.tree_center_of_mass_z = nullptr;
    return t;
}

// src/linear_octree.c:964-976
void free_linear_octree(LinearOctree *t)
{
    free(t->particle_morton_indices_deepest_level);
    free(t->sorted_indices);
    free(t->tree_num_particles);
    free(t->tree_num_internal_children);
    free(t->tree_first_particle_sorted_idx);
    free(t->tree_first_internal_children_idx);
    free(t->tree_mass);
    free(t->tree_center_of_mass_x);
    free(t->tree_center_of_mass_y);
    free(t->tree_center_of_mass_z);
    *t = get_new_linear_octree();
}

// src/linear_octree.c:978-985
bool linear_octree_check_if_included(const int64 ki, const int64 kj, const int level)
{
    const int s = 3 * (MORTON_MAX_LEVEL - level);
    return (ki >> s) == (kj >> s);
}

// src/utils.c:20-25
double grav_get_current_time(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return static_cast<double>(ts.tv_sec) + static_cast<double>(ts.tv_nsec) / 1.0e9;
}

}  // extern "C"
