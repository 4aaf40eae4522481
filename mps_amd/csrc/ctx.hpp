// ctx.hpp -- device-resident state of one ParticleSystem and the host-side
// launchers of the step kernels.
//
// HBM layout (slot s = position in the current cell-sorted order):
//   X[2], V[2]   dvec per slot, ping-ponged by the kernels that move
//                particles (predictor, collision, corrector)
//   T[2], O[2]   particle type and original index (ps->particles[] index),
//                ping-ponged by the counting sort
//   P[2]         pressure (ping-ponged by the corrector, which reads the
//                pressures of neighbours while it ghosts particles)
//   ACC, ND, SURF, NCOUNT [2]  per-slot fields, permuted with T/O
//   unk          1 = PPE unknown (interior fluid or wall)
//   nbr          neighbour list, sliced-ELL (64-slot slices, stride =
//                max_neighbors), entries are slots in reference order
//   ncnt         neighbour counts
//   Aval, Acol, Alen   off-diagonal PPE entries per unknown slot, same
//                sliced-ELL layout, plus diag[] and b[]
//   x, r, pv, Ap       CG vectors over all slots (0 at non-unknown slots)
#pragma once

#include <hip/hip_runtime.h>

#include "dev_common.hpp"
#include "mps_hip.h"

#define HIP_OK(call)                                                               \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(e_)); \
            return MPS_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

void mps_set_error(int code, const char *file, int line, const char *msg);
void mps_clear_error(void);
int particle_block();                       // block size of the per-particle kernels (runtime.hip)
int search_form();                          // 1: candidates staged in LDS, 0: gathered from HBM (runtime.hip)

// Step constants computed on the host in the reference's association order
// (host and device FP64 arithmetic are both IEEE, so precomputed products
// are bit-identical to the reference computing them inline).
struct StepConsts {
    Grid grid;
    double l0, r_min;
    double re_lap, re_n, re2;        // re2 = search radius squared
    double re_n2;                    // re_n*re_n (count-based surface test)
    double re_n2_hi;                 // r2 above this => w(re_n) = 0
    double n0, coeff;                // coeff = 2*DIM/(n0*lambda)
    double neg_coeff;                // -coeff
    double visc_coeff;               // viscosity*coeff
    double diag_coeff;               // coeff or c_ppe*coeff
    double g[3], dt;
    double cd2, col_tiny, e_half;    // collision: cd^2, 1e-12*l0, (1+e)*0.5
    double thr_n0;                   // surface_threshold*n0
    double beta_n0c;                 // surface_count_threshold*n0_count
    double src_k;                    // relax*(rho/dt^2) or gamma*(rho/dt^2)
    double neg_rho_dt;               // -rho/dt (weakly compressible PPE)
    double hs_k;                     // -(rho/n0/dt)
    double dim_n0;                   // DIM/n0
    double grad_min_r2;              // l0*l0*1e-12
    double neg_gc, rho;              // -(DIM/n0), density
    double bnd_lo[3], bnd_hi[3];     // domain -/+ l0*(wall_layers+1)
    double re_st, st_ll, st_sl, st_tiny;   // surface tension (operators.c:281-354)
    int stride;                      // max_neighbors
    int surf_method, ppe_type, hs_mode, cmps, clamp, collide;
    int st_enabled, st_sl_on;
    // hybrid wall friction of the 3-D serial program (built with
    // MPS_FRICTION_FIELDS; mps_fluid3D/src/operators.c:337-463)
    int fric;
    double fr_theta_s, fr_dth, fr_vc, fr_dyn;   // radians, radians, m/s, alpha*nu*2*DIM/(n0*lambda)
};

// Counters written by the search and assembly kernels.
struct DevCounters {
    int overflow_orig;               // smallest original index that overflowed, or INT_MAX
    int n_eq;                        // unknowns of the last search
    int max_count;
    int pad;
    unsigned long long n_pairs;
    unsigned long long nnz;
};

// The CG's copy of the PPE (k_pack.hip), rebuilt after each assembly.
// Rows are the unknowns in slot order (row k <-> slot slot_of_row[k]);
// columns are row numbers.  Rows are sliced 64 at a time (entry q of the
// rows of slice s at 64 * (smeta[s].x + q) + lane, a slice as wide as its
// longest row) and grouped 256 at a time into blocks, one per SpMV
// workgroup.  A block's halo is the sorted set of rows its entries
// reference (halo[bmeta[b].x ...], bmeta[b].y of them); in halo mode
// (smeta[s].z = 1) an entry's column is a 12-bit index into that halo;
// entries 8g..8g+7 of a row fill the 96 bits (entry u at bit 12u, word
// 0 lowest) of words smeta[s].y + 3g, +1, +2 (word w at w * 64 + lane of
// chw); else a 32-bit row number (c32[(smeta[s].y + q) * 64 + lane]).  Entries of
// a row keep the assembly's order, so every row sums as before.
#define PK_ROWS 256                  // rows per block = SpMV workgroup size
#define PK_HALO_CAP 4096             // largest halo staged in LDS (12-bit indices)
struct Packed {
    int n_eq = 0, n_slices = 0, n_blocks = 0, hmax = 0;
    int cap_rows = 0, cap_slots = 0;
    size_t cap_val = 0, cap_hw = 0, cap_c32 = 0, cap_fval = 0, cap_halo = 0;
    size_t tot_val = 0, tot_hw = 0, tot_c32 = 0, tot_halo = 0;
    int *row_of_slot = nullptr;      // exclusive scan of unk (n + 1)
    int *slot_of_row = nullptr;
    int *len = nullptr;              // off-diagonal entries of the row
    double *diag = nullptr, *b = nullptr, *x = nullptr;
    int *sw = nullptr, *whw = nullptr, *w32 = nullptr;       // per slice
    int *bh = nullptr, *bmode = nullptr;                     // per block
    int *voff = nullptr, *chwoff = nullptr, *c32off = nullptr, *hoff = nullptr;   // scans
    int4 *smeta = nullptr;           // per slice {value offset, column offset, halo mode, 0}
    int2 *bmeta = nullptr;           // per block {halo offset, halo length}
    int *halo = nullptr;
    double *val = nullptr;
    unsigned *chw = nullptr;
    int *c32 = nullptr;
    float *fval = nullptr;           // mixed precision: FP32 values
    double *ddir = nullptr;          // mixed precision: Dirichlet part of each row
    unsigned *cq = nullptr;          // precision 3: (distance code << 12) | halo index, per entry
    size_t cap_cq = 0;
    int *dmax = nullptr;             // device, 8 ints: [0] largest halo, [1..4] scan totals
    int *htot = nullptr;             // pinned: [0] rows, [1..5] a copy of dmax[0..4]
};

struct CgCtrl;   // device control block of the solvers (k_cg.hip)
struct IcData;   // IC(0) factor and schedules (k_ic.hip)
struct MixedBufs;

struct Ctx {
    ParticleSystem *ps = nullptr;
    int n = 0, cap = 0, stride = 0, grid_cap = 0;
    hipStream_t stream = nullptr;
    int cx = 0, ct = 0, cp = 0;      // current index of X/V, T/O/ACC/ND/SURF/NCOUNT, P
    dvec *X[2] = {nullptr, nullptr}, *V[2] = {nullptr, nullptr};
    dvec *ACC[2] = {nullptr, nullptr};
    int *T[2] = {nullptr, nullptr}, *O[2] = {nullptr, nullptr};
    int *SURF[2] = {nullptr, nullptr}, *NCOUNT[2] = {nullptr, nullptr};
    double *P[2] = {nullptr, nullptr}, *ND[2] = {nullptr, nullptr};
    int *unk = nullptr;
    double *pmin = nullptr;
    // counting sort
    int *cell = nullptr, *ccount = nullptr, *cstart = nullptr, *cfill = nullptr;
    int *tmp_key = nullptr, *sflag = nullptr, *sscan = nullptr;
    int *scan_blk = nullptr;
    // neighbour list
    int *nbr = nullptr, *ncnt = nullptr;
    // PPE + CG
    double *diag = nullptr, *b = nullptr, *Aval = nullptr;
    int *Acol = nullptr, *Alen = nullptr;
    double *x = nullptr, *r = nullptr, *pv = nullptr, *Ap = nullptr;
    double *pv2 = nullptr;           // second search-direction buffer (fused CG)
    double *partA = nullptr, *partB = nullptr;
    float *fAval = nullptr, *fr = nullptr, *fp = nullptr, *fAp = nullptr;
    double *ddir = nullptr;          // mixed solve: Dirichlet part of each row
    CgCtrl *ctrl = nullptr;          // device
    CgCtrl *hctrl = nullptr;         // pinned host mirror
    DevCounters *counters = nullptr, *hcounters = nullptr;
    // staging for host <-> device copies
    void *stage = nullptr;
    size_t stage_bytes = 0;
    int *nl_stage = nullptr;         // compact list staging (device)
    size_t nl_stage_elems = 0;
    int *cnt_orig = nullptr, *off_orig = nullptr;   // per original index
    // pinned host staging of the per-call path (the caller's particles,
    // the compact list, its counts and offsets), grown as needed
    Particle *h_part = nullptr;
    int *h_flat = nullptr, *h_idx = nullptr;   // h_idx: counts [0, n], then offsets [n + 1, 2n + 2)
    size_t h_part_cap = 0, h_flat_cap = 0, h_idx_cap = 0;
    // state flags
    bool has_state = false, resident = false, list_valid = false, ppe_valid = false;
    Grid grid{};
    mps_step_stats stats{};
    IcData *ic = nullptr;
    MixedBufs *mixed = nullptr;
    Packed *pk = nullptr;
    // The particle phases of a step (predictor to assembly) captured as a
    // HIP graph, one per buffer parity at entry (every step flips the
    // X/V, T/O/... and P indices); valid while the buffers and the step
    // constants stay those of the capture (runtime.hip).
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        int in[4] = {0, 0, 0, 0};     // cx, ct, cp, n at entry
        int out[3] = {0, 0, 0};       // cx, ct, cp after
        StepConsts k{};
    } sg[2];
};

// ---- batches of scenes stepped by one launch per operator ---------------
// One scene of such a batch: its device buffers (both halves of each
// ping-pong pair), step constants and sizes.  The scenes of a batch flip
// their buffers together, so the one parity (cx, ct, cp) passed to each
// launch holds for all of them.  A launch covers the concatenated block
// ranges of every scene; first[s] is the first block of scene s (first[ns]
// the total) and a block finds its scene by bisection (flat_scene).
struct FlatScene {
    int n, total;                    // particles, grid cells
    dvec *X[2], *V[2], *ACC[2];
    int *T[2], *O[2], *SURF[2], *NCOUNT[2];
    double *P[2], *ND[2];
    int *unk;
    double *pmin;
    int *cell, *ccount, *cstart, *cfill, *tmp_key, *sflag, *sscan;
    int *nbr, *ncnt;
    double *diag, *b, *Aval;
    int *Acol, *Alen;
    double *x;
    DevCounters *ctr;
    StepConsts k;
};

// Segmented exclusive scan: segment g scans in[0..n) into out[0..n]
// (out[n] = its total, also written to *tot when tot is set).
struct ScanSeg {
    const int *in;
    int *out;
    int n;
    int *tot;
};
__device__ __forceinline__ ScanSeg seg_global(const ScanSeg &s)
{
    ScanSeg g = s;
    g.in = gptr(s.in); g.out = gptr(s.out); g.tot = s.tot ? gptr(s.tot) : nullptr;
    return g;
}
// A set of segments on the device: segs[0..nseg), tile0[0..nseg] (the
// segments' first 2048-element tiles; tile0[nseg] = all tiles).
struct ScanSegs {
    ScanSeg *segs = nullptr;
    int *tile0 = nullptr;
    int nseg = 0, ntiles = 0;
};

// Block offsets of one launch shape over the scenes of a batch (device).
struct FlatGrid {
    int *first = nullptr;
    int blocks = 0, cap = 0;
};

struct FlatLaunch {
    const FlatScene *tab;            // device
    int ns;
    FlatGrid gp;                     // per-particle kernels at the block size of option "pblk"
    FlatGrid gb;                     // 256-thread kernels over the particles
    FlatGrid gc;                     // 256-thread kernels over the cells (total + 1)
    ScanSegs sc_cells, sc_sflag;     // the counting sort's two scans
    int *scan_blk;                   // tile sums of the segmented scans
    int any_cmps1;                   // some scene takes p_min from its neighbours' minimum
    int cx, ct, cp;
    hipStream_t s;
};

// ---- launchers (k_particles.hip) -------------------------------------
int  ctx_alloc(Ctx *c, int n, int stride, int grid_cells);
void ctx_free(Ctx *c);
int  launch_fill_int(int *p, int v, size_t n, hipStream_t s);
int  launch_fill_double(double *p, double v, size_t n, hipStream_t s);
int  launch_scan_excl(const int *in, int *out, int n, int *blk, hipStream_t s);
// up to four independent exclusive scans in three launches; blk needs
// 4 x 2050 ints, tot[k] receives array k's total (device memory)
struct ScanSet {
    const int *in[4];
    int *out[4];
    int n[4];
    int count;
};
int  launch_scan_excl_multi(const ScanSet &ss, int *blk, int *tot, hipStream_t s);
int  launch_unpack(Ctx *c, const void *stage_dev, int with_list);
int  launch_pack(Ctx *c, void *stage_dev);
int  launch_visc_predict(Ctx *c, const StepConsts &k);
int  launch_friction(Ctx *c, const StepConsts &k);   // the friction alone, on the current velocities
int  launch_collide_bin(Ctx *c, const StepConsts &k, int collide);
int  launch_bin(Ctx *c, const StepConsts &k);
int  launch_sort(Ctx *c, const StepConsts &k, int full);
int  launch_search(Ctx *c, const StepConsts &k, int with_density, DevCounters *ctr);
int  launch_assemble(Ctx *c, const StepConsts &k, DevCounters *ctr);
int  launch_pressure_out(Ctx *c, const StepConsts &k, int n_eq_zero);
int  launch_grad_correct(Ctx *c, const StepConsts &k);
int  launch_density_only(Ctx *c, const StepConsts &k);
// neighbour list <-> compact CSR over original indices (host interchange)
int  launch_list_export(Ctx *c, long long *total_out);
int  launch_list_import(Ctx *c, const int *counts_dev, const int *offsets_dev);

// one launch per operator over every scene of a batch (k_particles.hip)
// host block counts -> device offsets (copied on stream s, which is then
// synchronised)
int  flat_build_grid(FlatGrid *g, const int *blocks, int ns, hipStream_t s);
int  launch_scan_seg(const ScanSegs &ss, int *blk, hipStream_t s);
int  scan_segs_set(ScanSegs *ss, const ScanSeg *h, int nseg, hipStream_t s);   // host table -> device
int  flat_step_pre(FlatLaunch &f);      // predictor .. assembly (3-D)
int  flat_step_post(FlatLaunch &f);     // pressure, gradient, corrector, boundary
int  flat_counters(FlatLaunch &f, DevCounters *out_dev);        // every scene's counters, in scene order

// ---- solvers (k_cg.hip, k_ic.hip) ------------------------------------
struct SolveOpts {
    int solver;        // 0 CG, 1 ICCG
    int max_iter;
    double tol;
    int tol_mode;      // 0 absolute, 1 relative
    int precision;     // 0 FP64; 1 FP32 matrix, FP64 vectors; 2 FP32 matrix and vectors;
                       // 3 FP32 vectors, off-diagonals from 20-bit distance codes
    double inner_tol;
    int check_every;
    int dense_ic;      // 2-D dense-formula IC(0)
    int spmv_variant;  // entries per load batch in the SpMV
    int fused;         // FP64 CG in two kernels per iteration
    int graph;         // batched CG: replay one captured chunk of iterations
    // the PPE's off-diagonal A_ij = neg_coeff w(r_ij) (w of re_lap, r
    // clamped at r_min), for the distance-coded operator (precision 3)
    double neg_coeff, re_lap, r_min;
};
struct SolveResult {
    int iters, outer, breakdown;
    double res, bnorm;
};
int cg_alloc_ctrl(Ctx *c);
// FP64 CG over the packed systems of several contexts at once (k_cg.hip)
int batch_cg(Ctx *const *cs, int ns, const SolveOpts *opts, SolveResult *res, hipStream_t s);
int batch_cg_mixed(Ctx *const *cs, int ns, const SolveOpts *opts, SolveResult *res, hipStream_t s);
void cg_batch_free();
int pack_ppe(Ctx *c, int with_f32);        // k_pack.hip
int pack_stage_a(Ctx *c);                   // pack_ppe = a, wait, b, wait, c
int pack_stage_b(Ctx *c);
int pack_stage_c(Ctx *c, int with_f32);
int pack_wait(Ctx *c);
int unpack_solution(Ctx *c);               // packed x -> c->x (slot order)
void packed_free(Packed *p);
int solve_ppe(Ctx *c, const SolveOpts &o, SolveResult *out);
int cg_time_spmv(Ctx *c, int variant, int reps, double *ms);   // k_cg.hip

// ---- runtime accessors (runtime.hip) -----------------------------------
Ctx *mps_ctx_for(ParticleSystem *ps, int create);
int mps_ctx_upload(Ctx *c, ParticleSystem *ps, const NeighborList *nl, const Grid &g);
// particles only, device rows of `stride` entries (a caller's list capacity)
int mps_ctx_upload_rows(Ctx *c, ParticleSystem *ps, int stride, const Grid &g);
int mps_ctx_download(Ctx *c, ParticleSystem *ps, NeighborList *nl);
int mps_ctx_search(Ctx *c, const StepConsts &k);
StepConsts mps_consts(const SimConfig *cf, const ParticleSystem *ps, const Grid &g, double re);
Grid mps_grid_for(double re, const double *dmin, const double *dmax);
Grid mps_grid_from_cells(const CellList *cl);
double mps_search_radius(const SimConfig *cf);

// ---- adjoint tape (k_adjoint.hip) -------------------------------------
// Stages of a step at which a recording adjoint copies the state.
bool tape_recording();   // k_adjoint.hip: a tape is armed
enum TapeStage { TAPE_START, TAPE_PREDICTED, TAPE_COLLIDED, TAPE_SEARCHED, TAPE_SOLVED };
void tape_hook(Ctx *c, int stage);
int mps_ctx_step(Ctx *c, const SimConfig *cf);    // one resident timestep (runtime.hip)

// ---- timing (runtime.hip) ----------------------------------------------
enum ProfKernel { PK_VISC, PK_COLLIDE, PK_SORT, PK_SEARCH, PK_ASSEMBLE, PK_SPMV, PK_AXPY,
                  PK_PUPDATE, PK_GRAD, PK_IC, PK_PACK, PK_ICF, PK_ICB, PK_ICL, PK_COUNT };
void prof_begin(int k, hipStream_t s);
void prof_end(int k, hipStream_t s);
