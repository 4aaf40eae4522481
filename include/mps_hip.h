/*
 * mps_hip.h -- C-ABI of libmps_hip: the MPS per-timestep hot path on one
 * MI355X (gfx950).  Plain C types and pointers only.
 *
 * Part 1 declares the reference's own entry points with the reference's
 * prototypes; a program written against mps_fluid3D_parallel (or, built with
 * MPS_DIM=2, mps_fluid2D) links this library in place of the reference's
 * object files.  Each declaration names the reference interface it replaces.
 * Part 2 declares what the device-resident design adds: explicit
 * upload/download, a resident multi-step call, options, statistics, exports
 * used by parity tests, kernel timing, the batched ensemble and the adjoint.
 *
 * Semantics of the drop-in entry points: every call leaves the caller's
 * host structures exactly as the reference would (positions, velocities,
 * accelerations, pressure, number density, flags, neighbour lists), unless
 * the ParticleSystem has been made resident with mps_hip_set_resident(ps,1);
 * then state stays in HBM between calls and the host copy is refreshed only
 * by mps_hip_download().  Errors the reference reports with exit() (a
 * neighbour list overflow) still exit with the reference's message unless
 * mps_hip_set_option("exit_on_error", 0) is set, in which case the call
 * returns and mps_hip_last_error() reports the code.
 */
#ifndef MPS_HIP_H
#define MPS_HIP_H

#include "mps_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Only the declarations below are exported; the library is compiled with
 * hidden default visibility. */
#pragma GCC visibility push(default)

/* ===================================================================== */
/* Part 1: reference entry points                                         */
/* ===================================================================== */

/* --- configuration (host C, same keys, defaults and derivations) ------ */
/* mps_fluid3D_parallel/include/sim_config.h:84  config_set_defaults      */
void config_set_defaults(SimConfig *config);
/* mps_fluid3D_parallel/include/sim_config.h:87  config_load_cal          */
int  config_load_cal(const char *cal_path, SimConfig *config);
/* mps_fluid3D_parallel/include/sim_config.h:90  config_load_params       */
int  config_load_params(const char *param_path, SimConfig *config);
/* mps_fluid3D_parallel/include/sim_config.h:93  config_print             */
void config_print(const SimConfig *config);

/* --- particle container ------------------------------------------------ */
/* mps_fluid3D_parallel/include/particle.h:31-32 */
ParticleSystem *particle_system_create(int capacity);
void            particle_system_free(ParticleSystem *ps);
/* mps_fluid3D_parallel/include/particle.h:35 */
int  particle_system_add(ParticleSystem *ps, double pos[], double vel[], int type);
/* mps_fluid3D_parallel/include/particle.h:38 -- n0, lambda (+ N0, C_LL,
 * C_SL) computed on the GPU with a cell list, summing in the reference's
 * ascending-index order so the values are bit-identical to its O(N^2) pass */
void particle_system_calc_initial_params(ParticleSystem *ps);

/* --- kernel function --------------------------------------------------- */
/* mps_fluid3D_parallel/include/kernel.h:5,8 */
double kernel_weight(double r, double re);
double kernel_weight_derivative(double r, double re);

/* --- neighbour search -------------------------------------------------- */
/* mps_fluid3D_parallel/include/neighbor_search.h:33-34 */
NeighborList *neighbor_list_create(int num_particles, int max_neighbors);
void          neighbor_list_free(NeighborList *nl);
/* mps_fluid3D_parallel/include/neighbor_search.h:37-40 */
CellList *cell_list_create(int num_particles, double re,
                           const double domain_min[MPS_DIM],
                           const double domain_max[MPS_DIM]);
void      cell_list_free(CellList *cl);
/* mps_fluid3D_parallel/include/neighbor_search.h:43 (O(N^2) list, same set
 * and index order as the reference's unused brute-force search) */
void neighbor_search_brute_force(NeighborList *nl, ParticleSystem *ps, double re);
/* mps_fluid3D_parallel/include/neighbor_search.h:46-47 -- counting sort by
 * cell on the GPU; list order is the reference's (cells z->y->x, descending
 * particle index inside a cell).  cl->head/next are filled as the
 * reference's linked lists would be. */
void neighbor_search_cell_linked_list(NeighborList *nl, ParticleSystem *ps,
                                      CellList *cl, double re);
/* mps_fluid3D_parallel/include/neighbor_search.h:50 */
int  neighbor_get(NeighborList *nl, int i, int j);

/* --- particle operators (mps_fluid3D_parallel/include/operators.h:8-37;
 *     mps_fluid2D/include/operators.h:8-35 for the 2-D collision) ------- */
void calc_particle_number_density(ParticleSystem *ps, NeighborList *nl);
void calc_viscosity_term(ParticleSystem *ps, NeighborList *nl);
void calc_pressure_gradient(ParticleSystem *ps, NeighborList *nl);
void judge_free_surface(ParticleSystem *ps, double threshold);
void calc_neighbor_count(ParticleSystem *ps, NeighborList *nl);
void judge_free_surface_by_count(ParticleSystem *ps, double beta);
void clamp_pressure(ParticleSystem *ps);
#if MPS_DIM == 3
void collision(ParticleSystem *ps, NeighborList *nl);
#else
void collision(ParticleSystem *ps);
#endif
void calc_surface_tension(ParticleSystem *ps, NeighborList *nl);
void calc_surface_tension_SL(ParticleSystem *ps, NeighborList *nl);
#ifdef MPS_FRICTION_FIELDS
/* mps_fluid3D/include/operators.h:42 -- hybrid wall friction, in the 3-D
 * serial program's build (libmps_hip3ds.so) */
void apply_friction(ParticleSystem *ps, NeighborList *nl);
#endif

/* --- boundary (mps_fluid3D_parallel/include/boundary.h:8,12-14) -------- */
void apply_wall_boundary(ParticleSystem *ps);
#if MPS_DIM == 3
void remove_out_of_bounds(ParticleSystem *ps, double xmin, double xmax,
                          double ymin, double ymax, double zmin, double zmax);
#else
void remove_out_of_bounds(ParticleSystem *ps, double xmin, double xmax,
                          double ymin, double ymax);
#endif

/* --- pressure (mps_fluid3D_parallel/include/pressure_solver.h:8) ------- */
void solve_pressure(ParticleSystem *ps, NeighborList *nl);

/* --- driver (mps_fluid3D_parallel/include/simulation.h:8,17) ----------- */
void simulation_run(ParticleSystem *ps);
void simulation_step(ParticleSystem *ps, NeighborList *nl, CellList *cl, int step);

/* --- output (mps_fluid3D_parallel/include/io.h:8,11) -------------------- */
void output_csv(const ParticleSystem *ps, int step, const char *output_dir);
void output_vtk(const ParticleSystem *ps, int step, const char *output_dir);

/* ===================================================================== */
/* Part 2: additions of the device-resident design (no reference          */
/* counterpart; each is needed because state lives in HBM)               */
/* ===================================================================== */

/* Error codes returned by the mps_hip_* calls and mps_hip_last_error(). */
#define MPS_OK                 0
#define MPS_ERR_HIP            1   /* a HIP runtime call failed            */
#define MPS_ERR_NO_DEVICE      2   /* no usable gfx950 device               */
#define MPS_ERR_ALLOC          3   /* device allocation failed              */
#define MPS_ERR_NEIGHBOR_OVERFLOW 4 /* a particle exceeded max_neighbors    */
#define MPS_ERR_ARG            5   /* invalid argument / state              */
#define MPS_ERR_NO_CONFIG      6   /* g_config is NULL                      */
#define MPS_ERR_UNSUPPORTED    7   /* option combination not built          */
#define MPS_ERR_SOLVER         8   /* solver breakdown / level overflow     */

/* Reads the loader format of mps_fluid3D_parallel/src/main.c:20-85
 * (mps_fluid2D/src/main.c for 2-D): a count line, then "x y [z] vx vy [vz]
 * type" rows; comment lines start with '#'.  Returns 0 or -1. */
int  mps_load_particles(const char *path, ParticleSystem *ps);
/* Counts the particle rows announced by the file's header (for sizing). */
int  mps_count_particles(const char *path);

/* Writes a dam-break scene on the lattice rules of
 * mps_fluid3D_parallel/tools/gen_dambreak.py:99-166 (2-D:
 * mps_fluid2D/tools/gen_dambreak.py:107-150): particles in i->j->k order,
 * wall then dummy layers, fluid inside the column.  If ps is non-NULL the
 * particles are appended to it (with the %.8e rounding the text file would
 * carry); if path is non-NULL the particles.txt text is written there.
 * dims = {domain_x, domain_y, domain_z, water_x, water_y, water_z}.
 * Returns the particle count, or -1. */
int  mps_gen_dambreak(const double dims[6], double l0, int wall_layers,
                      int dummy_layers, ParticleSystem *ps, const char *path);

int  mps_hip_init(int device);          /* optional; first call does it   */
void mps_hip_finalize(void);            /* frees every device context     */
/* The error of the last failed call.  Each void entry point of Part 1
 * clears it when it starts, so after such a call it reports that call. */
int  mps_hip_last_error(void);
/* Text of an error code: the detailed message (with the particle index of
 * a neighbour-list overflow) when the last error had this code. */
const char *mps_hip_error_string(int code);
/* gfx arch name of the active device, e.g. "gfx950" ("" if none). */
const char *mps_hip_device_arch(void);

/* Host <-> device state.  upload copies ps (and nl, if given, as the stale
 * list the next step's viscosity and collision read) into the device
 * context of ps; download copies device state back into ps (and nl). */
int  mps_hip_upload(ParticleSystem *ps, const NeighborList *nl);
int  mps_hip_download(ParticleSystem *ps, NeighborList *nl);
int  mps_hip_set_resident(ParticleSystem *ps, int resident);
void mps_hip_release(ParticleSystem *ps);

/* Runs nsteps timesteps on the resident device state of ps (uploading
 * first if ps has no device state).  No host copies inside. */
int  mps_hip_run_steps(ParticleSystem *ps, int nsteps);

/* Ensemble of independent scenes (no counterpart in the reference, which
 * runs one scene per process): nsteps timesteps of every ps[i] under its
 * own configuration cfgs[i] (viscosity, geometry, ... may differ; same
 * dimensionality), state resident on the GPU.  In the 3-D library with
 * option "batch_flat" 1 (the default) every particle operator, sort and
 * packing stage is one launch over all scenes' particles; otherwise they
 * run per scene.  The CG (FP64, or the mixed form of option "precision")
 * runs batched -- one launch per CG phase for all scenes, per-scene
 * reductions and stop tests.  With option "graph" the phases before and
 * after the solve and a chunk of CG iterations replay captured HIP graphs.
 * Every scene gets the bits of its own single-scene run.  Scenes that use
 * ICCG are stepped one by one.  ps[i] stays resident: read it back with
 * mps_hip_download.  g_config is left as it was. */
int  mps_hip_run_batch(ParticleSystem *const *ps, SimConfig *const *cfgs, int nscenes, int nsteps);

/* Gradient of L = (1/N_f) sum_fluid |r_i(T) - target_i|^2 after nsteps
 * steps from the state in ps (no counterpart in the reference, which has
 * no derivatives).  target: N x DIM positions in particle order.  The
 * forward keeps a checkpoint (particles and neighbour list, host memory)
 * every checkpoint_every steps; the reverse sweep recomputes each segment
 * on the GPU, records its intermediate states, and runs the adjoint of
 * every operator with the forward's discrete decisions frozen (lists,
 * flags, clamp, p_min choice, collisions), the pressure adjoint solved by
 * CG on the same matrix.  Outputs: *loss, *grad_nu (dL/d viscosity),
 * grad_g[DIM] (dL/d gravity), grad_u0[N x DIM] (dL/d initial velocities;
 * 0 for non-fluid particles).  ps is left in the final forward state
 * (device copy released).  Scheme variants without a reverse here
 * (surface tension, hs_mode, ppe_type 1, cmps_gradient, count-based
 * surface detection) return MPS_ERR_UNSUPPORTED. */
int  mps_hip_adjoint(ParticleSystem *ps, int nsteps, int checkpoint_every, const double *target,
                     double *loss, double *grad_nu, double *grad_g, double *grad_u0);

/* The loss of mps_hip_adjoint for nsets parameter sets (viscosity nus[q],
 * gravity gs[q*DIM..], initial velocities u0s[q*N*DIM..] -- or those of ps
 * when u0s is NULL), each computed by replaying the nsteps steps recorded
 * from ps with every discrete decision of that recording frozen (lists,
 * unknown set, clamp, p_min neighbour, collisions, ghosting): the function
 * the adjoint differentiates.  Its finite differences check the adjoint at
 * horizons where those of the free forward cross flipped decisions.
 * Keeps nsteps recorded steps in device memory; ps is left in the recorded
 * run's final state (device copy released). */
int  mps_hip_replay_losses(ParticleSystem *ps, int nsteps, const double *target, int nsets,
                           const double *nus, const double *gs, const double *u0s, double *losses);

/* Backward of one timestep (no counterpart in the reference): runs
 * simulation_step's step from ps under g_config -- predictor on the stale
 * list nl as simulation_step (3DP/include/simulation.h:17) receives it, or
 * on a search of ps's positions when nl is NULL -- then its reverse, with
 * the decisions of that step frozen, for the adjoints of the step's output
 * positions and velocities rbar_out, ubar_out (N x DIM each).  Outputs
 * (each may be NULL): rbar_in, ubar_in (N x DIM, adjoints of the input
 * positions and velocities; 0 for non-fluid particles, whose state the
 * step holds fixed), *grad_nu and grad_g[DIM] (this step's share of the
 * viscosity and gravity gradients).  ps and nl are left as simulation_step
 * leaves them (device copy released).  Chained backwards over steps, with
 * each step's input state, it gives mps_hip_adjoint's gradient. */
int  mps_hip_step_backward(ParticleSystem *ps, NeighborList *nl, const double *rbar_out,
                           const double *ubar_out, double *rbar_in, double *ubar_in, double *grad_nu,
                           double *grad_g);

/* Options (process-wide): keys and values
 *   "tol_mode"       0 absolute sqrt(r.r) < tol (reference), 1 relative
 *                      sqrt(r.r) < tol*sqrt(b.b)
 *   "precision"      0 FP64 throughout; mixed forms iterate on an FP32 copy
 *                      of the PPE matrix with FP64 dot products and FP64
 *                      solution, and make reliable updates (true residual
 *                      b - A x recomputed with the FP64 matrix), testing
 *                      convergence on that FP64 residual: 1 FP64 CG
 *                      vectors, 2 FP32 CG vectors and FP32 SpMV sums,
 *                      3 as 2 with each off-diagonal applied from a 20-bit
 *                      code of its pair distance packed with a 12-bit halo
 *                      index (4 bytes per nonzero instead of 6)
 *   "inner_tol"      residual reduction between reliable updates (1e-3)
 *   "exit_on_error"  1 (reference behaviour) or 0
 *   "check_every"    CG iterations between host convergence polls (16)
 *   "spmv"           SpMV kernel of the CG (same bits in all): 0 eight
 *                      entries per load batch, 1 sixteen, matrix loaded
 *                      non-temporally in both; 2 sixteen, matrix loaded
 *                      through the caches
 *   "fused"          CG iteration, same results bit for bit in all
 *                      forms: 2 (default) three kernels, each finaliser's
 *                      reduction repeated by the blocks of the next vector
 *                      kernel (FP64 1.3 %, mixed 0.8 % faster than 0);
 *                      0 five kernels (SpMV, one-block finaliser, axpy,
 *                      finaliser, p update); 1 (FP64 only; the mixed forms
 *                      run five) two kernels, the previous iteration closed
 *                      inside the SpMV, p formed where it is read (slower)
 *   "graph"          1: the particle phases of each step (predictor to
 *                      assembly) are captured once as a HIP graph and
 *                      replayed, one graph per buffer parity; same results
 *                      bit for bit; not used while kernels are timed or a
 *                      tape records (0, default: plain launches)
 *   "profile"        1 times the main kernels with HIP events, 2 only the
 *                      CG SpMV, 3 only the ICCG forward solve
 *   "pblk"           block size of the per-particle kernels (collision,
 *                      predictor, search, assembly, gradient): 128, 256
 *                      (default) or 512; same results bit for bit
 *   "async_output"   simulation_run's snapshots: 1 (default) each output
 *                      step's state is copied to pinned host memory and a
 *                      host thread writes the CSV/VTK files while the
 *                      device runs on; 0 written before the next step, as
 *                      the reference does; the same bytes either way
 * Returns 0, or MPS_ERR_ARG for an unknown key. */
int    mps_hip_set_option(const char *key, double value);
double mps_hip_get_option(const char *key);

typedef struct {
    int    steps;           /* steps run on this context                  */
    int    last_iters;      /* CG/ICCG iterations of the last solve       */
    long long total_iters;  /* summed over all solves                     */
    int    last_outer;      /* reliable updates of the last solve (mixed) */
    int    n_eq;            /* unknowns of the last PPE                   */
    long long nnz;          /* nonzeros of the last PPE matrix (with diag)*/
    long long n_pairs;      /* directed neighbour pairs of the last search*/
    double last_residual;   /* sqrt(r.r) at exit (recursive residual)     */
    double last_bnorm;      /* sqrt(b.b)                                  */
    int    ic_levels;       /* dependency levels of the last IC(0) factor */
    int    max_count;       /* largest neighbour count of the last search */
} mps_step_stats;
int  mps_hip_get_stats(ParticleSystem *ps, mps_step_stats *out);

/* Autotuner (no counterpart in the reference): times the choices that
 * leave every result bit unchanged -- the CG SpMV kernel (option "spmv"),
 * the host's polling interval (option "check_every"), the block size of
 * the per-particle kernels (option "pblk") and the search's read batch
 * (option "search" 1 or 2) -- on a private device copy of
 * ps's host particles under g_config (one step's PPE, solved once per
 * candidate; the particle phases three times per block size).  ps and its
 * device state are not touched.  With apply != 0 the fastest choices
 * become the options.  Returns 0 or an error code. */
typedef struct {
    int    spmv;            /* faster of "spmv" variants 0 and 1            */
    int    check_every;     /* fastest "check_every"                       */
    double spmv_ms[3];      /* mean SpMV launch, variants 0 (8 entries per
                               batch, non-temporal), 1 (16, non-temporal),
                               2 (16, through the caches)                  */
    double solve_ms[4];     /* one solve polling every 8, 16, 32, 64 iters */
    int    iters;           /* CG iterations of that solve                 */
    int    pblk;            /* fastest "pblk"                              */
    double pblk_ms[3];      /* particle phases of a step at block sizes
                               128, 256, 512 (search option 1)             */
    int    search;          /* faster "search" form, 1 or 2 (staged
                               candidates read 4 or 2 at a time)           */
    double search_ms[2];    /* particle phases at the chosen block size
                               with search 1 and 2                         */
} mps_tune_result;
int  mps_hip_autotune(ParticleSystem *ps, int apply, mps_tune_result *out);
/* Batch autotuner (no counterpart in the reference): the block size of the
 * per-particle kernels (option "pblk") and the search's read batch (option
 * "search") for mps_hip_run_batch's launches over all scenes -- the
 * batch's particle phases timed at 128, 256 and 512, then with search 2 at
 * the best size, on private device copies of the scenes' host particles
 * (pblk, pblk_ms, search and search_ms of *out; the other fields 0).  With
 * apply != 0 the fastest choices become the options.  3-D library;
 * MPS_ERR_UNSUPPORTED otherwise. */
int  mps_hip_autotune_batch(ParticleSystem *const *ps, SimConfig *const *cfgs, int nscenes, int apply,
                            mps_tune_result *out);

/* Exports of the last assembled PPE in the reference's layout
 * (mps_fluid3D_parallel/src/pressure_solver.c:411-551): rows in eq order
 * (ascending particle index over unknowns), columns sorted ascending.
 * eq_idx has ps->num entries; row_ptr n_eq+1; col/val nnz; diag and b
 * n_eq.  Pass NULL for arrays not wanted.  Counts via mps_hip_get_stats. */
int  mps_hip_export_ppe(ParticleSystem *ps, int *eq_idx, int *row_ptr,
                        int *col, double *val, double *diag, double *b);
/* Last PPE solution x in eq order (n_eq entries). */
int  mps_hip_export_solution(ParticleSystem *ps, double *x);
/* IC(0) factor of the last ICCG solve in the reference's layout
 * (pressure_solver.c:156-232): strictly-lower CSR + 1/sqrt(d). */
int  mps_hip_export_ic(ParticleSystem *ps, int *row_ptr, int *col,
                       double *val, double *diag_inv);

/* Kernel timing ("profile" option): accumulated milliseconds and launch
 * count of a named kernel class ("spmv", "search", ...).  Returns -1 if the
 * name is unknown. */
int  mps_hip_profile_get(const char *name, double *total_ms, long long *count);
void mps_hip_profile_reset(void);

/* Solves A x = b on the GPU for a caller-supplied CSR matrix (rows sorted
 * by column) with the library's CG/ICCG -- the linear-algebra half of
 * solve_pressure, exposed for testing against textbook solvers. */
int  mps_hip_solve_csr(int n, const int *row_ptr, const int *col,
                       const double *val, const double *b, double *x,
                       int solver_type, int max_iter, double tol,
                       int *iters_out);

/* mps_hip_solve_csr without allocation: the same arguments, kernels and
 * results (bit for bit), with every device buffer carved from one static
 * device array of MPS_NOALLOC_SCRATCH_BYTES that each GPU holds for the life
 * of the process.  It calls no hipMalloc, hipFree or other allocating
 * routine, for callers that cannot allocate (serving loops, harnesses that
 * forbid it).  No counterpart in the reference.  Solves that do not fit the
 * array (about 36 bytes per entry plus 12 per ELL slot and 170 per unknown,
 * 194 for ICCG:
 * a 3-D PPE of 15,000 unknowns with rows of up to 123 entries takes 71 MB)
 * fail with MPS_ERR_ARG and a
 * message naming the need; nothing is written to x then.  Every call on one
 * GPU uses that GPU's one array: calls that overlap on a GPU race. */
#define MPS_NOALLOC_SCRATCH_BYTES (96u << 20)
int  mps_hip_solve_csr_noalloc(int n, const int *row_ptr, const int *col,
                               const double *val, const double *b, double *x,
                               int solver_type, int max_iter, double tol,
                               int *iters_out);

/* Diagnostics: computes sqrt(a[i]) and a[i]/b[i] on the GPU with the FP64
 * operations the kernels use (to check they round exactly as the host). */
int  mps_hip_fp64_probe(int n, const double *a, const double *b, double *sqrt_out,
                        double *div_out);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif

#endif /* MPS_HIP_H */
