/*
 * mps_oracle.h -- CPU restatement of the reference MPS timestep, used ONLY
 * as test infrastructure: by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker or the timed CPU
 * comparator, never inside the product library.
 *
 * It restates the arithmetic of TaroOmura/mps (mps_fluid3D_parallel for
 * MPS_DIM=3, mps_fluid2D for MPS_DIM=2) operation by operation, in the same
 * summation order and without FMA contraction, so that on x86-64 (gcc -O2,
 * no -mfma) its results are bit-identical to the reference's own build.
 * Data structures are its own: neighbour lists are CSR, cells are built by
 * a counting sort, the configuration is passed explicitly.
 *
 * Pinning (what checks each function; see tests/):
 *   orc_search, orc_density, orc_surface, orc_visc_predict, orc_collision,
 *   orc_assemble, orc_solve (CG, both IC(0) forms), orc_gradient_correct,
 *   orc_boundary, orc_step, orc_init_params: bit-compared against the
 *   reference compiled from its sources (oracle/_ref) on the shipped
 *   dam-break inputs and generated scenes; neighbour sets also against
 *   O(N^2) brute force; n0/lambda against closed-form lattice sums and the
 *   logged known answers (mps_fluid2D/USAGE.txt:66,
 *   mps_fluid3D_parallel/output/rolling_run.log:37-38); CG/ICCG against
 *   numpy.linalg.solve; IC(0) against a dense pattern-restricted Cholesky.
 */
#ifndef MPS_ORACLE_H
#define MPS_ORACLE_H

#include "mps_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Neighbour lists in CSR form: neighbours of i are idx[start[i]..start[i+1]). */
typedef struct {
    int  n;
    long cap;
    long *start;
    int  *idx;
} orc_nlist;

/* Grid geometry of cell_list_create (neighbor_search.c:92-122). */
typedef struct {
    int    nc[3];
    long   total;
    double size;
    double inv;
    double origin[3];
} orc_grid;

/* Solver options that the reference cannot express in params.txt. */
typedef struct {
    int tol_mode;   /* 0: sqrt(r.r) < tol (reference); 1: < tol*sqrt(b.b) */
    int dot_order;  /* 0: forward serial sum (reference); 1: backward serial
                       sum -- used to measure the spread a different
                       reduction order alone causes */
} orc_opts;

/* PPE of the last step in the reference's layout. */
typedef struct {
    int     n_eq;
    long    nnz;
    int    *eq_idx;     /* [num] */
    long   *row_ptr;    /* [n_eq+1] */
    int    *col;        /* [nnz] sorted per row */
    double *val;        /* [nnz] */
    double *diag;       /* [n_eq] */
    double *b;          /* [n_eq] */
    double *x;          /* [n_eq] */
    /* IC(0) factor (sparse form: strictly lower CSR + 1/sqrt(d);
       dense 2-D form: strictly lower CSR + sqrt(d) in l_diag) */
    long   *l_ptr;
    int    *l_col;
    double *l_val;
    double *l_diag;
    int     iters;
    double  res;        /* sqrt(r.r) at exit */
    double  bnorm;
    int     breakdown;  /* 1 if |pAp| < 1e-30 stopped the loop */
} orc_ppe;

typedef struct {
    orc_nlist nl;       /* current list = stale list of the next step */
    orc_ppe   ppe;
    orc_opts  opts;
    int       overflow; /* particle index that exceeded max_neighbors, or -1 */
} orc_sim;

/* kernel_weight of kernel.c:12-20 */
double orc_weight(double r, double re, double l0);

void orc_grid_make(orc_grid *g, double re, const double *dmin, const double *dmax);
/* cell index of a position, or -1 outside the grid (neighbor_search.c:171-178) */
long orc_cell_of(const orc_grid *g, const double *pos);

int  orc_nlist_reserve(orc_nlist *nl, int n, long cap);
void orc_nlist_free(orc_nlist *nl);

/* neighbor_search_cell_linked_list (neighbor_search.c:146-240): returns 0,
 * or the 1-based index of the first particle whose list overflowed
 * max_neighbors (the reference exits there). */
int  orc_search(const ParticleSystem *ps, const SimConfig *c, double re, orc_nlist *out);
/* O(N^2) set of the reference's neighbor_search_brute_force (:44-81) */
int  orc_search_brute(const ParticleSystem *ps, double re, orc_nlist *out);

void orc_density(ParticleSystem *ps, const SimConfig *c, const orc_nlist *nl);
void orc_surface(ParticleSystem *ps, double threshold);
void orc_neighbor_count(ParticleSystem *ps, const SimConfig *c, const orc_nlist *nl);
void orc_surface_by_count(ParticleSystem *ps, double beta);
/* acc = g (+ viscosity), u* = u + dt acc, r* = r + dt u*  (simulation.c:31-67) */
void orc_visc_predict(ParticleSystem *ps, const SimConfig *c, const orc_nlist *stale);
/* collision (3-D: operators.c:371-446 over the stale list;
 * 2-D: mps_fluid2D/src/operators.c:324-392 over all pairs) */
void orc_collision(ParticleSystem *ps, const SimConfig *c, const orc_nlist *stale);
/* solve_pressure (pressure_solver.c:368-570): assembly + CG/ICCG + scatter */
int  orc_solve_pressure(orc_sim *s, ParticleSystem *ps, const SimConfig *c);
void orc_clamp(ParticleSystem *ps);
/* calc_pressure_gradient + corrector (operators.c:102-181, simulation.c:100-108) */
void orc_gradient_correct(ParticleSystem *ps, const SimConfig *c, const orc_nlist *nl);
/* apply_wall_boundary + remove_out_of_bounds (boundary.c:12-61) */
void orc_boundary(ParticleSystem *ps, const SimConfig *c);

orc_sim *orc_sim_create(void);
void     orc_sim_free(orc_sim *s);
/* the initial search of simulation_run (simulation.c:153) */
int  orc_sim_begin(orc_sim *s, ParticleSystem *ps, const SimConfig *c);
/* one simulation_step (simulation.c:19-118); returns 0 or overflow index */
int  orc_step(orc_sim *s, ParticleSystem *ps, const SimConfig *c);

/* particle_system_calc_initial_params (particle.c:69-221) with the serial
 * tie rule (first strict maximum in index order), O(N) via a grid. */
void orc_init_params(ParticleSystem *ps, const SimConfig *c);

/* Stand-alone linear solve of a CSR system with the reference CG
 * (pressure_solver.c:90-131) or ICCG (:156-322; dense_ic selects the 2-D
 * formulas of mps_fluid2D/src/pressure_solver.c:86-219). */
int  orc_solve_csr(int n, const long *row_ptr, const int *col, const double *val,
                   const double *b, double *x, int solver, int dense_ic,
                   int max_iter, double tol, const orc_opts *o, double *res_out);

/* The particles.txt loader of main.c:20-85, line by line; messages as
 * the reference prints them.  Returns 0, or -1 for a missing file or count. */
int  orc_load_particles(const char *path, ParticleSystem *ps);

/* accessors for language bindings */
int  orc_sim_nlist(const orc_sim *s, const long **start, const int **idx);
const orc_ppe *orc_sim_ppe(const orc_sim *s);
void orc_sim_set_opts(orc_sim *s, int tol_mode, int dot_order);

#ifdef __cplusplus
}
#endif

#endif /* MPS_ORACLE_H */
