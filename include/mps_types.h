/*
 * mps_types.h -- in-memory types shared by the host layer and the HIP
 * library.  The layouts are byte-for-byte those of the reference programs
 * so that a caller compiled against the reference headers can link this
 * library instead:
 *
 *   Particle / ParticleSystem   mps_fluid3D_parallel/include/particle.h:7-28
 *   NeighborList / CellList     mps_fluid3D_parallel/include/neighbor_search.h:8-30
 *   SimConfig                   mps_fluid3D_parallel/include/sim_config.h:10-80
 *   particle type codes, DIM    mps_fluid3D_parallel/include/config.h:5-11
 *
 * The spatial dimension is fixed when a caller (and the library) is
 * compiled, as in the reference: define MPS_DIM=2 for the mps_fluid2D
 * layout, leave it unset (3) for mps_fluid3D_parallel.  MPS_FRICTION_FIELDS
 * inserts the four hybrid-friction parameters that mps_fluid3D's SimConfig
 * carries (mps_fluid3D/include/sim_config.h:72-77); the library ignores
 * them but keeps the layout.
 */
#ifndef MPS_TYPES_H
#define MPS_TYPES_H

#ifndef MPS_DIM
#define MPS_DIM 3
#endif
#ifndef DIM
#define DIM MPS_DIM
#endif

#define FLUID_PARTICLE 0
#define WALL_PARTICLE  1
#define GHOST_PARTICLE 2
#define DUMMY_PARTICLE 3

#ifdef __cplusplus
extern "C" {
#endif

/* One particle (104 bytes in 3-D, 80 bytes in 2-D). */
typedef struct {
    double pos[MPS_DIM];
    double vel[MPS_DIM];
    double acc[MPS_DIM];
    double pressure;
    double n;               /* particle number density */
    int    neighbor_count;  /* count inside re_n (count-based surface test) */
    int    type;            /* FLUID / WALL / GHOST / DUMMY */
    int    on_surface;      /* 1 = Dirichlet free-surface particle */
} Particle;

typedef struct {
    Particle *particles;
    int       num;
    int       capacity;
    double    n0;           /* reference number density */
    double    lambda;       /* Laplacian-model constant */
    double    C_LL;         /* liquid-liquid surface-tension coefficient */
    double    C_SL;         /* solid-liquid surface-tension coefficient */
    int       n0_count;     /* reference neighbour count */
} ParticleSystem;

/* Padded neighbour list: neighbors[i*max_neighbors + k], k < count[i]. */
typedef struct {
    int *neighbors;
    int *count;
    int  max_neighbors;
} NeighborList;

/* Uniform grid of cubic cells of edge cell_size with prepend-linked lists. */
typedef struct {
    int   *head;
    int   *next;
    int    nx;
    int    ny;
#if MPS_DIM == 3
    int    nz;
#endif
    int    total_cells;
    double cell_size;
    double origin[MPS_DIM];
} CellList;

typedef struct {
    double particle_distance;      /* l0 */
    double influence_ratio_lap;
    double influence_radius_lap;   /* re_lap = ratio * l0 (derived) */
    double influence_ratio_n;
    double influence_radius_n;     /* re_n = ratio * l0 (derived) */
    int    max_neighbors;
    int    wall_layers;
    int    dummy_layers;

    double density;
    double viscosity;
    double gravity[MPS_DIM];

    double dt;
    double t_end;
    int    output_interval;

    int    solver_type;            /* 0 CG, 1 ICCG */
    int    cg_max_iter;
    double cg_tolerance;
    double relaxation_coeff;
    int    clamp_negative_pressure;
    int    ppe_type;               /* 0 density-deviation, 1 weakly compressible */
    double c_ppe;
    double gamma_ppe;

    double surface_threshold;
    int    surface_detection_method;  /* 0 number density, 1 neighbour count */
    double surface_count_threshold;

    double restitution_coeff;
    double collision_distance_ratio;

    int    surface_tension_enabled;
    double surface_tension_coeff;
    double surface_tension_re_ratio;
    double influence_radius_st;    /* derived */

    double wetting_angle_SL;

    double domain_min[MPS_DIM];
    double domain_max[MPS_DIM];

    int    cmps_gradient;          /* 0 P_j - P_min, 1 CMPS, 2 Oochi */
    int    use_analytical_lambda;
    int    hs_mode;                /* 0 standard source, 1 high-order source */

#ifdef MPS_FRICTION_FIELDS
    int    friction_enabled;
    double friction_delta_theta;
    double dynamic_friction_alpha;
    double friction_velocity_threshold;
#endif

    char   output_dir[256];
    char   particle_file[256];
    char   param_file[256];
} SimConfig;

/* Process-wide configuration read by every step function, as in the
 * reference (mps_fluid3D_parallel/include/sim_config.h:83). */
extern __attribute__((visibility("default"))) SimConfig *g_config;

#ifdef __cplusplus
}
#endif

#endif /* MPS_TYPES_H */
