/*
 * main.c -- mps_sim driver with the reference's command line
 * (mps_fluid3D_parallel/src/main.c:87-136): mps_sim <cal.txt>.
 * Loads cal.txt/params.txt and particles.txt, computes n0/lambda, runs the
 * simulation on the GPU and writes CSV/VTK snapshots every output_interval
 * steps.  The particle container is sized from the file's header, so scenes
 * larger than the reference's fixed 100,000 load.
 */
#include <stdio.h>
#include <stdlib.h>
#include <sys/stat.h>

#include "mps_hip.h"

int main(int argc, char *argv[])
{
    if (argc < 2) {
        fprintf(stderr, "Usage: %s <cal_file>\n", argv[0]);
        fprintf(stderr, "  cal_file: calculation file (e.g., cal.txt)\n");
        return 1;
    }
    SimConfig config;
    config_set_defaults(&config);
    if (config_load_cal(argv[1], &config) != 0) return 1;
    if (config_load_params(config.param_file, &config) != 0) return 1;
    g_config = &config;
    printf("=== MPS %dD Fluid Simulation (HIP) ===\n\n", MPS_DIM);
    config_print(&config);
    mkdir(config.output_dir, 0755);

    int n = mps_count_particles(config.particle_file);
    if (n <= 0) {
        fprintf(stderr, "Error: invalid particle count in '%s'\n", config.particle_file);
        return 1;
    }
    ParticleSystem *ps = particle_system_create(n);
    if (!ps) {
        fprintf(stderr, "Error: failed to create particle system\n");
        return 1;
    }
    if (mps_load_particles(config.particle_file, ps) != 0) {
        particle_system_free(ps);
        return 1;
    }
    particle_system_calc_initial_params(ps);
    simulation_run(ps);
    const int rc = mps_hip_last_error();   /* the run's own error, if any */
    if (rc != 0) fprintf(stderr, "Error: %s\n", mps_hip_error_string(rc));
    particle_system_free(ps);
    mps_hip_finalize();
    return rc != 0;
}
