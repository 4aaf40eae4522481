/*
 * config.c -- cal.txt / params.txt reader.
 *
 * Same keys, defaults, path resolution and derived radii as the reference
 * (mps_fluid3D_parallel/src/sim_config.c:20-79 defaults -- the 2-D
 * program's differ in max_neighbors and domain_max, :86-109 path
 * resolution, :120-159 cal.txt, :166-270 params.txt, :272-274 radii), so an
 * existing case directory runs unchanged.  The parser is table-driven: each
 * key maps to a typed field offset.  Keys added by this library (accepted
 * only here): cg_tolerance_mode (abs|rel), precision (fp64|mixed),
 * inner_tolerance -- they set the library options of mps_hip_set_option.
 */
#include <ctype.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mps_hip.h"

SimConfig *g_config = NULL;

void config_set_defaults(SimConfig *c)
{
    memset(c, 0, sizeof(*c));
    c->particle_distance = 0.025;
    c->influence_ratio_lap = 2.1;
    c->influence_radius_lap = 2.1 * 0.025;
    c->influence_ratio_n = 2.1;
    c->influence_radius_n = 2.1 * 0.025;
#if MPS_DIM == 3
    c->max_neighbors = 512;
#else
    c->max_neighbors = 256;     /* mps_fluid2D/src/sim_config.c defaults */
#endif
    c->wall_layers = 2;
    c->dummy_layers = 2;
    c->density = 1000.0;
    c->viscosity = 1.0e-6;
    c->gravity[1] = -9.81;
    c->dt = 5.0e-4;
    c->t_end = 2.0;
    c->output_interval = 100;
    c->solver_type = 0;
    c->cg_max_iter = 10000;
    c->cg_tolerance = 1.0e-8;
    c->relaxation_coeff = 0.2;
    c->clamp_negative_pressure = 0;
    c->ppe_type = 0;
    c->c_ppe = 1.01;
    c->gamma_ppe = 0.01;
    c->surface_threshold = 0.97;
    c->surface_detection_method = 0;
    c->surface_count_threshold = 0.85;
    c->restitution_coeff = 0.2;
    c->collision_distance_ratio = 0.5;
    c->surface_tension_enabled = 0;
    c->surface_tension_coeff = 0.0728;
    c->surface_tension_re_ratio = 3.2;
    c->influence_radius_st = 3.2 * 0.025;
    c->wetting_angle_SL = 90.0;
#if MPS_DIM == 3
    c->domain_max[0] = 0.4;
    c->domain_max[1] = 0.3;
    c->domain_max[2] = 0.2;
#else
    c->domain_max[0] = 1.0;
    c->domain_max[1] = 0.6;
#endif
#ifdef MPS_FRICTION_FIELDS
    c->friction_delta_theta = 30.0;   /* mps_fluid3D/src/sim_config.c:76-79 */
    c->dynamic_friction_alpha = 600.0;
    c->friction_velocity_threshold = 0.01;
#endif
    c->cmps_gradient = 0;
    c->use_analytical_lambda = 0;
    c->hs_mode = 0;
    strcpy(c->output_dir, "output");
}

/* strips leading and trailing white space in place */
static char *strip(char *s)
{
    while (*s && isspace((unsigned char)*s)) s++;
    size_t len = strlen(s);
    while (len > 0 && isspace((unsigned char)s[len - 1])) s[--len] = '\0';
    return s;
}

/* value relative to the directory holding cal_path unless absolute */
static void join_relative(const char *cal_path, const char *value, char *out, size_t cap)
{
    const char *slash = strrchr(cal_path, '/');
    if (value[0] == '/' || slash == NULL) {
        snprintf(out, cap, "%s", value);
        return;
    }
    char dir[256];
    size_t dlen = (size_t)(slash - cal_path) + 1;
    if (dlen >= sizeof(dir)) dlen = sizeof(dir) - 1;
    memcpy(dir, cal_path, dlen);
    dir[dlen] = '\0';
    snprintf(out, cap, "%s%s", dir, value);
}

/* Reads "key value" pairs; calls fn for every non-comment line. */
typedef void (*kv_fn)(const char *key, const char *value, void *user);
static int read_pairs(const char *path, const char *what, kv_fn fn, void *user)
{
    FILE *fp = fopen(path, "r");
    if (!fp) {
        fprintf(stderr, "Error: cannot open %s file '%s'\n", what, path);
        return -1;
    }
    char line[512];
    while (fgets(line, sizeof line, fp)) {
        char *s = strip(line);
        if (*s == '\0' || *s == '#') continue;
        char key[128], value[256];
        if (sscanf(s, "%127s %255s", key, value) != 2) continue;
        fn(key, value, user);
    }
    fclose(fp);
    return 0;
}

struct cal_ctx { const char *cal_path; SimConfig *cfg; };

static void cal_pair(const char *key, const char *value, void *user)
{
    struct cal_ctx *cx = (struct cal_ctx *)user;
    if (strcmp(key, "particle_file") == 0)
        join_relative(cx->cal_path, value, cx->cfg->particle_file, sizeof cx->cfg->particle_file);
    else if (strcmp(key, "param_file") == 0)
        join_relative(cx->cal_path, value, cx->cfg->param_file, sizeof cx->cfg->param_file);
    else
        fprintf(stderr, "Warning: unknown key in cal file: '%s'\n", key);
}

int config_load_cal(const char *cal_path, SimConfig *cfg)
{
    struct cal_ctx cx = { cal_path, cfg };
    if (read_pairs(cal_path, "cal", cal_pair, &cx) != 0) return -1;
    if (cfg->particle_file[0] == '\0') {
        fprintf(stderr, "Error: particle_file not specified in '%s'\n", cal_path);
        return -1;
    }
    if (cfg->param_file[0] == '\0') {
        fprintf(stderr, "Error: param_file not specified in '%s'\n", cal_path);
        return -1;
    }
    return 0;
}

enum field_kind { F_DOUBLE, F_INT, F_STRING };
struct field { const char *key; enum field_kind kind; size_t offset; };

#define FD(k, m) { k, F_DOUBLE, offsetof(SimConfig, m) }
#define FI(k, m) { k, F_INT, offsetof(SimConfig, m) }
static const struct field k_fields[] = {
    FD("particle_distance", particle_distance),
    FD("influence_ratio_lap", influence_ratio_lap),
    FD("influence_ratio_n", influence_ratio_n),
    FI("max_neighbors", max_neighbors),
    FI("wall_layers", wall_layers),
    FI("dummy_layers", dummy_layers),
    FD("density", density),
    FD("viscosity", viscosity),
    FD("gravity_x", gravity[0]),
    FD("gravity_y", gravity[1]),
#if MPS_DIM == 3
    FD("gravity_z", gravity[2]),
#endif
    FD("dt", dt),
    FD("t_end", t_end),
    FI("output_interval", output_interval),
    FI("solver_type", solver_type),
    FI("cg_max_iter", cg_max_iter),
    FD("cg_tolerance", cg_tolerance),
    FD("relaxation_coeff", relaxation_coeff),
    FI("clamp_negative_pressure", clamp_negative_pressure),
    FI("ppe_type", ppe_type),
    FD("c_ppe", c_ppe),
    FD("gamma_ppe", gamma_ppe),
    FD("surface_threshold", surface_threshold),
    FI("surface_detection_method", surface_detection_method),
    FD("surface_count_threshold", surface_count_threshold),
    FD("restitution_coeff", restitution_coeff),
    FD("collision_distance_ratio", collision_distance_ratio),
    FI("surface_tension_enabled", surface_tension_enabled),
    FD("surface_tension_coeff", surface_tension_coeff),
    FD("surface_tension_re_ratio", surface_tension_re_ratio),
    FD("wetting_angle_SL", wetting_angle_SL),
#ifdef MPS_FRICTION_FIELDS
    FI("friction_enabled", friction_enabled),
    FD("friction_delta_theta", friction_delta_theta),
    FD("dynamic_friction_alpha", dynamic_friction_alpha),
    FD("friction_velocity_threshold", friction_velocity_threshold),
#endif
    FD("domain_x_min", domain_min[0]),
    FD("domain_x_max", domain_max[0]),
    FD("domain_y_min", domain_min[1]),
    FD("domain_y_max", domain_max[1]),
#if MPS_DIM == 3
    FD("domain_z_min", domain_min[2]),
    FD("domain_z_max", domain_max[2]),
#endif
    FI("cmps_gradient", cmps_gradient),
    FI("use_analytical_lambda", use_analytical_lambda),
    FI("hs_mode", hs_mode),
    { "output_dir", F_STRING, offsetof(SimConfig, output_dir) },
};
#undef FD
#undef FI

static void param_pair(const char *key, const char *value, void *user)
{
    SimConfig *cfg = (SimConfig *)user;
    for (size_t f = 0; f < sizeof k_fields / sizeof k_fields[0]; f++) {
        if (strcmp(key, k_fields[f].key) != 0) continue;
        char *dst = (char *)cfg + k_fields[f].offset;
        switch (k_fields[f].kind) {
        case F_DOUBLE: *(double *)dst = atof(value); break;
        case F_INT:    *(int *)dst = atoi(value); break;
        case F_STRING: strncpy(dst, value, 255); break;
        }
        return;
    }
    /* library options carried in params.txt */
    if (strcmp(key, "cg_tolerance_mode") == 0) {
        mps_hip_set_option("tol_mode", (strcmp(value, "rel") == 0 || strcmp(value, "1") == 0) ? 1.0 : 0.0);
        return;
    }
    if (strcmp(key, "precision") == 0) {
        mps_hip_set_option("precision", (strcmp(value, "mixed") == 0 || strcmp(value, "1") == 0) ? 1.0 : 0.0);
        return;
    }
    if (strcmp(key, "inner_tolerance") == 0) {
        mps_hip_set_option("inner_tol", atof(value));
        return;
    }
    fprintf(stderr, "Warning: unknown parameter '%s'\n", key);
}

int config_load_params(const char *param_path, SimConfig *cfg)
{
    if (read_pairs(param_path, "param", param_pair, cfg) != 0) return -1;
    cfg->influence_radius_lap = cfg->influence_ratio_lap * cfg->particle_distance;
    cfg->influence_radius_n = cfg->influence_ratio_n * cfg->particle_distance;
    cfg->influence_radius_st = cfg->surface_tension_re_ratio * cfg->particle_distance;
    return 0;
}

void config_print(const SimConfig *c)
{
    printf("=== Simulation Configuration ===\n");
    printf("particle_distance:    %.6f m\n", c->particle_distance);
    printf("influence_radius_lap: %.6f m  (ratio = %.2f)  [Laplacian]\n",
           c->influence_radius_lap, c->influence_ratio_lap);
    printf("influence_radius_n:   %.6f m  (ratio = %.2f)  [number density]\n",
           c->influence_radius_n, c->influence_ratio_n);
    printf("max_neighbors:        %d\n", c->max_neighbors);
    printf("wall_layers:          %d\n", c->wall_layers);
    printf("dummy_layers:          %d\n", c->dummy_layers);
    printf("density:              %.1f kg/m^3\n", c->density);
    printf("viscosity:            %.2e m^2/s\n", c->viscosity);
#if MPS_DIM == 3
    printf("gravity:              (%.4f, %.4f, %.4f) m/s^2\n",
           c->gravity[0], c->gravity[1], c->gravity[2]);
#else
    printf("gravity:              (%.4f, %.4f) m/s^2\n", c->gravity[0], c->gravity[1]);
#endif
    printf("dt:                   %.2e s\n", c->dt);
    printf("t_end:                %.4f s\n", c->t_end);
    printf("output_interval:      %d steps\n", c->output_interval);
    printf("solver_type:          %s\n", c->solver_type == 1 ? "ICCG" : "CG");
    printf("cg_max_iter:          %d\n", c->cg_max_iter);
    printf("cg_tolerance:         %.2e (%s)\n", c->cg_tolerance,
           mps_hip_get_option("tol_mode") > 0.5 ? "relative" : "absolute");
    printf("relaxation_coeff:     %.4f\n", c->relaxation_coeff);
    printf("clamp_negative_pressure: %s\n", c->clamp_negative_pressure ? "ON" : "OFF");
    printf("ppe_type:             %d  (%s)\n", c->ppe_type,
           c->ppe_type == 1 ? "Natsui weak-compressible" : "standard density");
    printf("surface_threshold:    %.4f\n", c->surface_threshold);
    printf("surface_detection:    %s\n",
           c->surface_detection_method == 1 ? "neighbor count (Natsui)" : "number density");
    printf("restitution_coeff:         %.4f\n", c->restitution_coeff);
    printf("collision_distance_ratio:  %.4f  (col_dist = %.6f m)\n",
           c->collision_distance_ratio, c->collision_distance_ratio * c->particle_distance);
#if MPS_DIM == 3
    printf("domain:               [%.3f, %.3f] x [%.3f, %.3f] x [%.3f, %.3f]\n",
           c->domain_min[0], c->domain_max[0], c->domain_min[1], c->domain_max[1],
           c->domain_min[2], c->domain_max[2]);
#else
    printf("domain:               [%.3f, %.3f] x [%.3f, %.3f]\n",
           c->domain_min[0], c->domain_max[0], c->domain_min[1], c->domain_max[1]);
#endif
    printf("cmps_gradient:        %d\n", c->cmps_gradient);
    printf("use_analytical_lambda: %d\n", c->use_analytical_lambda);
    printf("hs_mode:              %d\n", c->hs_mode);
    printf("precision:            %s\n", mps_hip_get_option("precision") > 0.5 ? "mixed FP32/FP64" : "FP64");
    printf("output_dir:           %s\n", c->output_dir);
    printf("particle_file:        %s\n", c->particle_file);
    printf("param_file:           %s\n", c->param_file);
    printf("================================\n\n");
}
