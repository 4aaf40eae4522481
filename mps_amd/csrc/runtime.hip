// runtime.hip -- device contexts, host<->device transfers, the timestep
// orchestration and the C-ABI entry points of libmps_hip.
//
// One HIP stream per ParticleSystem.  State stays in HBM between calls when
// the system is resident (simulation_run, mps_hip_run_steps); otherwise each
// drop-in call uploads the caller's structures, runs, and downloads, so the
// caller sees exactly what the reference would have left behind.
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <map>
#include <string>
#include <thread>
#include <vector>

#include <chrono>

#include "ctx.hpp"

// ---------------------------------------------------------------------
// errors, options, profiling
// ---------------------------------------------------------------------

static int g_last_error = MPS_OK;
static char g_last_msg[512] = "";

static int g_msg_code = MPS_OK;   // the code g_last_msg was written for

void mps_set_error(int code, const char *file, int line, const char *msg)
{
    g_last_error = code;
    g_msg_code = code;
    snprintf(g_last_msg, sizeof g_last_msg, "%s:%d: %s", file, line, msg ? msg : "");
}

// Each void entry point of the reference's API starts with this, so that
// mps_hip_last_error() afterwards reports that call alone.
void mps_clear_error(void) { g_last_error = MPS_OK; }

struct Options {
    double tol_mode = 0, precision = 0, inner_tol = 1e-3, exit_on_error = 1, check_every = 16;
    double spmv = 0, profile = 0, verbose = 0, fused = 2, graph = 0, pblk = 256;
    double async_output = 1;
    double search = 1;                 // neighbour search: 1 (2) candidates staged in LDS, read 4 (2) at a time; 0 gathered from HBM
    double batch_flat = 1;             // mps_hip_run_batch: 1 one launch per operator for all scenes, 0 per scene
    double ic_wait_limit = 1048576;   // unsuccessful polls before an IC(0) wait gives up
    double ic_sleep = 1;               // longest backoff of an idle IC(0) wave (0-3)
    double ic_bucket = -1;             // log2 of the levels per bucket of the IC(0) solve order (-1: automatic)
    double ic_chunk = 64;              // positions per workgroup chunk of the IC(0) solves (32, 64, 128)
    double ic_waves = 16;              // wavefronts per workgroup of the IC(0) solves (4, 8, 16)
    double ic_factor_form = 2;         // IC(0) factor: 2 two-phase (rows of at most 64 entries), 1 one-phase
    double ic_grid = 0;                // workgroups of the IC(0) solves (0: automatic)
    double ic_fgrid = 0;               // workgroups of the IC(0) factor (0: automatic)
};
static Options g_opt;

// Snapshot output of simulation_run.  With option async_output (the
// default) an output step's state is copied into one of two pinned host
// buffers, and a host thread formats and writes it while the device runs
// the following steps.  The thread reads nothing but its snapshot and the
// output directory, so the files hold the same bytes as a synchronous
// write.  At most one writer runs: each output step waits for the previous
// writer before it starts the next, and a snapshot is downloaded into the
// buffer that writer is not reading.
struct SnapshotWriter {
    Particle *buf[2] = {nullptr, nullptr};
    size_t cap = 0;
    int next = 0;
    std::thread th;

    // a writer still running at exit finishes its files (the pinned
    // buffers are left to the process teardown)
    ~SnapshotWriter() { join(); }
    void join()
    {
        if (th.joinable()) th.join();
    }
    int reserve(size_t n)
    {
        if (n < 1) n = 1;
        if (n <= cap) return MPS_OK;
        release();
        for (int b = 0; b < 2; b++)
            if (hipHostMalloc((void **)&buf[b], sizeof(Particle) * n, hipHostMallocDefault) != hipSuccess) {
                buf[b] = nullptr;
                release();
                return MPS_ERR_ALLOC;
            }
        cap = n;
        return MPS_OK;
    }
    void release()
    {
        join();
        for (int b = 0; b < 2; b++) {
            if (buf[b]) hipHostFree(buf[b]);
            buf[b] = nullptr;
        }
        cap = 0;
        next = 0;
    }
};
static SnapshotWriter g_writer;

static void write_snapshot(ParticleSystem snap, int step, std::string dir)
{
    output_csv(&snap, step, dir.c_str());
    output_vtk(&snap, step, dir.c_str());
}

// Starts the writer thread for snap; writes synchronously where no thread
// can be created.  The previous writer must have been joined.
static void start_snapshot(const ParticleSystem &snap, int step, const char *dir)
{
    try {
        g_writer.th = std::thread(write_snapshot, snap, step, std::string(dir));
    } catch (...) {
        write_snapshot(snap, step, std::string(dir));
    }
}

int particle_block() { return (int)g_opt.pblk; }
int search_form() { return (int)g_opt.search; }

// Kernel timing: HIP events recorded on the launching stream around each
// launch of a kernel class, drawn from a pool so recording costs no
// allocation; pairs are folded into totals after the stream synchronises.
struct ProfSlot {
    const char *name;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0;
    long long count = 0;
};
static ProfSlot g_prof[PK_COUNT] = {{"visc_predict"}, {"collide_bin"}, {"sort"}, {"search"},
                                    {"assemble"}, {"spmv"}, {"axpy"}, {"pupdate"},
                                    {"grad_correct"}, {"ic_fwd"}, {"pack"}, {"ic_factor"}, {"ic_bwd"}, {"ic_levels"}};
static std::vector<hipEvent_t> g_event_pool;
static hipEvent_t g_open[PK_COUNT];

static hipEvent_t pool_get()
{
    if (g_event_pool.empty()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
}

// profile 1 times every kernel class; 2 only the SpMV (two events per CG
// iteration instead of six, which keeps the timed loop closer to its
// unprofiled speed).
static bool prof_on(int k)
{
    // 1 every kernel class, 2 the SpMV alone, 3 the IC forward solve alone
    if (g_opt.profile > 2.5) return k == PK_IC;
    return g_opt.profile > 1.5 ? k == PK_SPMV : g_opt.profile > 0.5;
}

void prof_begin(int k, hipStream_t s)
{
    if (!prof_on(k)) return;
    hipEvent_t e = pool_get();
    if (!e) return;
    hipEventRecord(e, s);
    g_open[k] = e;
}

void prof_end(int k, hipStream_t s)
{
    if (!prof_on(k) || !g_open[k]) return;
    hipEvent_t e = pool_get();
    if (!e) return;
    hipEventRecord(e, s);
    g_prof[k].pending.push_back({g_open[k], e});
    g_open[k] = nullptr;
}

// Folds completed event pairs into the totals (called after a sync).
static void prof_collect()
{
    for (int k = 0; k < PK_COUNT; k++) {
        ProfSlot &p = g_prof[k];
        for (auto &pr : p.pending) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
                p.total_ms += ms;
                p.count++;
            }
            g_event_pool.push_back(pr.first);
            g_event_pool.push_back(pr.second);
        }
        p.pending.clear();
    }
}

// ---------------------------------------------------------------------
// device and contexts
// ---------------------------------------------------------------------

static int g_device = -1;
int ic_noalloc_setup(int dev);   // k_ic.hip
static char g_arch[64] = "";
static std::map<ParticleSystem *, Ctx *> g_ctx;

static int ensure_device()
{
    if (g_device >= 0) return MPS_OK;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        mps_set_error(MPS_ERR_NO_DEVICE, __FILE__, __LINE__, "no HIP device visible");
        return MPS_ERR_NO_DEVICE;
    }
    HIP_OK(hipSetDevice(0));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, 0));
    snprintf(g_arch, sizeof g_arch, "%s", prop.gcnArchName);
    if (strncmp(g_arch, "gfx950", 6) != 0) {
        mps_set_error(MPS_ERR_NO_DEVICE, __FILE__, __LINE__, "library is built for gfx950 only");
        return MPS_ERR_NO_DEVICE;
    }
    g_device = 0;
    (void)ic_noalloc_setup(0);   // its status is kept for mps_hip_solve_csr_noalloc
    return MPS_OK;
}

template <typename T>
static int dalloc(T **p, size_t n)
{
    if (*p) return 0;
    HIP_OK(hipMalloc((void **)p, sizeof(T) * (n > 0 ? n : 1)));
    return 0;
}

#define ALLOC(p, n)                   \
    do {                              \
        int rc_ = dalloc(&(p), (n));  \
        if (rc_) return rc_;          \
    } while (0)

#define DFREE(p)                      \
    do {                              \
        if (p) hipFree(p);            \
        (p) = nullptr;                \
    } while (0)

void ic_free(IcData *d);   // k_ic.hip
void mixed_free(MixedBufs *m);   // k_cg.hip

// Drops the captured step graphs (their kernel arguments hold the
// context's buffer addresses).
static void graph_reset(Ctx *c)
{
    for (auto &g : c->sg) {
        if (g.exec) hipGraphExecDestroy(g.exec);
        g = Ctx::StepGraph();
    }
}

void ctx_free(Ctx *c)
{
    graph_reset(c);
    for (int q = 0; q < 2; q++) {
        DFREE(c->X[q]); DFREE(c->V[q]); DFREE(c->ACC[q]); DFREE(c->T[q]); DFREE(c->O[q]);
        DFREE(c->SURF[q]); DFREE(c->NCOUNT[q]); DFREE(c->P[q]); DFREE(c->ND[q]);
    }
    DFREE(c->unk); DFREE(c->pmin); DFREE(c->cell); DFREE(c->ccount); DFREE(c->cstart);
    DFREE(c->cfill); DFREE(c->tmp_key); DFREE(c->sflag); DFREE(c->sscan); DFREE(c->scan_blk);
    DFREE(c->nbr); DFREE(c->ncnt); DFREE(c->diag); DFREE(c->b); DFREE(c->Aval); DFREE(c->Acol);
    DFREE(c->Alen); DFREE(c->x); DFREE(c->r); DFREE(c->pv); DFREE(c->pv2); DFREE(c->Ap); DFREE(c->partA);
    DFREE(c->partB); DFREE(c->fAval); DFREE(c->ddir); DFREE(c->fr); DFREE(c->fp); DFREE(c->fAp);
    DFREE(c->stage); DFREE(c->nl_stage); DFREE(c->cnt_orig); DFREE(c->off_orig);
    DFREE(c->counters);
    if (c->ctrl) { hipFree(c->ctrl); c->ctrl = nullptr; }
    if (c->hctrl) { hipHostFree(c->hctrl); c->hctrl = nullptr; }
    if (c->hcounters) { hipHostFree(c->hcounters); c->hcounters = nullptr; }
    for (void *h : {(void *)c->h_part, (void *)c->h_flat, (void *)c->h_idx})
        if (h) hipHostFree(h);
    c->h_part = nullptr;
    c->h_flat = c->h_idx = nullptr;
    c->h_part_cap = c->h_flat_cap = c->h_idx_cap = 0;
    if (c->ic) { ic_free(c->ic); c->ic = nullptr; }
    if (c->mixed) { mixed_free(c->mixed); c->mixed = nullptr; }
    if (c->pk) { packed_free(c->pk); c->pk = nullptr; }
    c->n = c->cap = c->stride = c->grid_cap = 0;
    c->stage_bytes = c->nl_stage_elems = 0;
    c->has_state = c->list_valid = false;
}

int ctx_alloc(Ctx *c, int n, int stride, int grid_cells)
{
    if (n > c->cap || stride != c->stride) {
        // particle-sized buffers
        Grid keep = c->grid;
        int gcap = c->grid_cap;
        int *cc = c->ccount, *cs = c->cstart, *cf = c->cfill;
        c->ccount = c->cstart = c->cfill = nullptr;
        ctx_free(c);
        c->grid = keep;
        c->ccount = cc; c->cstart = cs; c->cfill = cf; c->grid_cap = gcap;
        const size_t N = (size_t)(n > 0 ? n : 1);
        const size_t NS = ((N + 63) / 64) * 64;   // whole slices
        for (int q = 0; q < 2; q++) {
            ALLOC(c->X[q], N); ALLOC(c->V[q], N); ALLOC(c->ACC[q], N); ALLOC(c->T[q], N);
            ALLOC(c->O[q], N); ALLOC(c->SURF[q], N); ALLOC(c->NCOUNT[q], N); ALLOC(c->P[q], N);
            ALLOC(c->ND[q], N);
        }
        ALLOC(c->unk, N); ALLOC(c->pmin, N); ALLOC(c->cell, N); ALLOC(c->tmp_key, N);
        ALLOC(c->sflag, N + 1); ALLOC(c->sscan, N + 1); ALLOC(c->scan_blk, 8200);
        ALLOC(c->nbr, NS * (size_t)stride); ALLOC(c->ncnt, N);
        ALLOC(c->diag, N); ALLOC(c->b, N); ALLOC(c->Aval, NS * (size_t)stride);
        ALLOC(c->Acol, NS * (size_t)stride); ALLOC(c->Alen, N);
        ALLOC(c->x, N); ALLOC(c->r, N); ALLOC(c->pv, N); ALLOC(c->Ap, N);
        ALLOC(c->partA, 65536); ALLOC(c->partB, 65536);
        ALLOC(c->cnt_orig, N + 1); ALLOC(c->off_orig, N + 1);
        ALLOC(c->counters, 1);
        HIP_OK(hipHostMalloc(&c->hcounters, sizeof(DevCounters), hipHostMallocDefault));
        c->stage_bytes = sizeof(Particle) * N;
        HIP_OK(hipMalloc(&c->stage, c->stage_bytes));
        c->cap = n;
        c->stride = stride;
        // zero the matrix so unused entries of the last slice are benign
        HIP_OK(hipMemsetAsync(c->Alen, 0, sizeof(int) * N, c->stream));
        HIP_OK(hipMemsetAsync(c->ncnt, 0, sizeof(int) * N, c->stream));
    }
    if (grid_cells + 4 > c->grid_cap) {
        graph_reset(c);
        DFREE(c->ccount); DFREE(c->cstart); DFREE(c->cfill);
        ALLOC(c->ccount, (size_t)grid_cells + 4);
        ALLOC(c->cstart, (size_t)grid_cells + 4);
        ALLOC(c->cfill, (size_t)grid_cells + 4);
        c->grid_cap = grid_cells + 4;
    }
    c->n = n;
    return 0;
}

static Ctx *find_ctx(ParticleSystem *ps)
{
    auto it = g_ctx.find(ps);
    return it == g_ctx.end() ? nullptr : it->second;
}

static Ctx *get_ctx(ParticleSystem *ps)
{
    Ctx *c = find_ctx(ps);
    if (c) return c;
    if (ensure_device() != MPS_OK) return nullptr;
    c = new Ctx();
    c->ps = ps;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, "stream creation failed");
        delete c;
        return nullptr;
    }
    g_ctx[ps] = c;
    return c;
}

// ---------------------------------------------------------------------
// constants and geometry
// ---------------------------------------------------------------------

static double search_radius(const SimConfig *cf)
{
    // simulation.c:22-24
    double re = fmax(cf->influence_radius_lap, cf->influence_radius_n);
    if (cf->surface_tension_enabled) re = fmax(re, cf->influence_radius_st);
    return re;
}

// cell_list_create (neighbor_search.c:92-122)
static Grid grid_for(double re, const double *dmin, const double *dmax)
{
    Grid g{};
    const double margin = 4.0 * re;
    g.total = 1;
    for (int d = 0; d < 3; d++) {
        if (d < D_) {
            g.origin[d] = dmin[d] - margin;
            const double span = (dmax[d] + margin) - g.origin[d];
            g.nc[d] = (int)(span / re) + 2;
        } else {
            g.origin[d] = 0.0;
            g.nc[d] = 1;
        }
        g.total *= g.nc[d];
    }
    g.inv = 1.0 / re;
    return g;
}

static Grid grid_from_cells(const CellList *cl)
{
    Grid g{};
    g.nc[0] = cl->nx;
    g.nc[1] = cl->ny;
#if MPS_DIM == 3
    g.nc[2] = cl->nz;
#else
    g.nc[2] = 1;
#endif
    g.total = g.nc[0] * g.nc[1] * g.nc[2];
    for (int d = 0; d < 3; d++) g.origin[d] = d < D_ ? cl->origin[d] : 0.0;
    g.inv = 1.0 / cl->cell_size;
    return g;
}

static StepConsts make_consts(const SimConfig *cf, const ParticleSystem *ps, const Grid &g, double re)
{
    StepConsts k{};
    k.grid = g;
    k.l0 = cf->particle_distance;
    k.r_min = 0.01 * cf->particle_distance;
    k.re_lap = cf->influence_radius_lap;
    k.re_n = cf->influence_radius_n;
    k.re2 = re * re;
    k.re_n2 = cf->influence_radius_n * cf->influence_radius_n;
    k.re_n2_hi = k.re_n2 * (1.0 + 1.0e-14);
    k.n0 = ps->n0;
    k.coeff = 2.0 * D_ / (ps->n0 * ps->lambda);
    k.neg_coeff = -k.coeff;
    k.visc_coeff = cf->viscosity * k.coeff;
    k.diag_coeff = (cf->ppe_type == 1) ? cf->c_ppe * k.coeff : k.coeff;
    for (int d = 0; d < 3; d++) k.g[d] = d < D_ ? cf->gravity[d] : 0.0;
    k.dt = cf->dt;
    const double cd = cf->collision_distance_ratio * cf->particle_distance;
    k.cd2 = cd * cd;
    k.col_tiny = 1.0e-12 * cf->particle_distance;
    k.e_half = (1.0 + cf->restitution_coeff) * 0.5;
    k.thr_n0 = cf->surface_threshold * ps->n0;
    k.beta_n0c = cf->surface_count_threshold * ps->n0_count;
    const double dt2 = cf->dt * cf->dt;
    k.src_k = (cf->ppe_type == 1) ? cf->gamma_ppe * (cf->density / dt2)
                                  : cf->relaxation_coeff * (cf->density / dt2);
    k.neg_rho_dt = -cf->density / cf->dt;
    k.hs_k = -(cf->density / ps->n0 / cf->dt);
    k.dim_n0 = (double)D_ / ps->n0;
    k.grad_min_r2 = cf->particle_distance * cf->particle_distance * 1.0e-12;
    k.neg_gc = -((double)D_ / ps->n0);
    k.rho = cf->density;
    const double margin = cf->particle_distance * (cf->wall_layers + 1);
    for (int d = 0; d < 3; d++) {
        k.bnd_lo[d] = d < D_ ? cf->domain_min[d] - margin : -INFINITY;
        k.bnd_hi[d] = d < D_ ? cf->domain_max[d] + margin : INFINITY;
    }
    k.stride = 0;
    k.surf_method = cf->surface_detection_method;
    k.ppe_type = cf->ppe_type;
    k.hs_mode = cf->hs_mode;
    k.cmps = cf->cmps_gradient;
    k.clamp = cf->clamp_negative_pressure;
    k.collide = (D_ == 2) ? 1 : (cf->collision_distance_ratio > 0.0);
    k.st_enabled = cf->surface_tension_enabled;
    k.re_st = cf->influence_radius_st;
#if MPS_DIM == 3
    k.st_ll = ps->C_LL / (cf->density * cf->particle_distance * cf->particle_distance *
                          cf->particle_distance);
    k.st_sl = ps->C_SL / (cf->density * cf->particle_distance * cf->particle_distance *
                          cf->particle_distance);
#else
    // the 2-D program divides by rho*l0^2 (mps_fluid2D/src/operators.c)
    k.st_ll = ps->C_LL / (cf->density * cf->particle_distance * cf->particle_distance);
    k.st_sl = ps->C_SL / (cf->density * cf->particle_distance * cf->particle_distance);
#endif
    k.st_sl_on = ps->C_SL != 0.0;
    k.st_tiny = 1.0e-9 * cf->particle_distance;
#ifdef MPS_FRICTION_FIELDS
    // mps_fluid3D/src/operators.c:339-349, same products
    k.fric = cf->friction_enabled;
    k.fr_theta_s = cf->wetting_angle_SL * M_PI / 180.0;
    k.fr_dth = cf->friction_delta_theta * M_PI / 180.0;
    k.fr_vc = cf->friction_velocity_threshold;
    k.fr_dyn = cf->dynamic_friction_alpha * cf->viscosity * 2.0 * D_ / (ps->n0 * ps->lambda);
#endif
    return k;
}

// ---------------------------------------------------------------------
// host <-> device
// ---------------------------------------------------------------------

// Capacity of a device neighbour row: the caller's list when there is one
// (the reference's overflow test is count >= nl->max_neighbors,
// neighbor_search.c:221-227), else the configuration's.
static int stride_for(const SimConfig *cf, const NeighborList *nl)
{
    int s = nl ? nl->max_neighbors : (cf ? cf->max_neighbors : 512);
    if (s < 1) s = 1;
    return s;
}

// Copies ps (and nl's lists, when nl is given) to the device; rows hold
// `stride` entries (0: as stride_for).
// f(lo, hi) over pieces of [0, n) on up to 16 host threads: the host side
// of the per-call copies (gathering and scattering the caller's padded
// list and particle array), which one thread would make the bottleneck.
template <typename F>
static void host_for(long long n, F f)
{
    unsigned nt = std::thread::hardware_concurrency();
    if (nt > 16) nt = 16;
    if (nt < 2 || n < (1 << 16)) {
        f(0LL, n);
        return;
    }
    const long long per = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    long long done = per < n ? per : n;
    try {
        for (unsigned t = 1; t < nt && done < n; t++) {
            const long long lo = done, hi = lo + per < n ? lo + per : n;
            th.emplace_back(f, lo, hi);
            done = hi;
        }
    } catch (...) {
    }
    f(0LL, per < n ? per : n);
    for (auto &x : th) x.join();
    if (done < n) f(done, n);   // pieces no thread could be started for
}

// Pinned host memory of at least `bytes` in *p (capacity *cap), kept.
template <typename T>
static int pinned_grow(T **p, size_t *cap, size_t bytes)
{
    if (*p && *cap >= bytes) return 0;
    if (*p) hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    const size_t want = bytes + bytes / 8 + 64;
    HIP_OK(hipHostMalloc((void **)p, want, hipHostMallocDefault));
    *cap = want;
    return 0;
}

static void host_copy(void *dst, const void *src, size_t bytes)
{
    host_for((long long)((bytes + 4095) / 4096), [&](long long lo, long long hi) {
        const size_t a = (size_t)lo * 4096, b = (size_t)hi * 4096 < bytes ? (size_t)hi * 4096 : bytes;
        if (a < b) memcpy((char *)dst + a, (const char *)src + a, b - a);
    });
}

static int upload_rows(Ctx *c, ParticleSystem *ps, const NeighborList *nl, int stride, const Grid &g)
{
    int rc = ctx_alloc(c, ps->num, stride > 0 ? stride : stride_for(g_config, nl), g.total);
    if (rc) return rc;
    c->grid = g;
    const int n = ps->num;
    if (n > 0) {
        // the caller's particles through pinned staging
        rc = pinned_grow(&c->h_part, &c->h_part_cap, sizeof(Particle) * (size_t)n);
        if (rc) return rc;
        host_copy(c->h_part, ps->particles, sizeof(Particle) * (size_t)n);
        HIP_OK(hipMemcpyAsync(c->stage, c->h_part, sizeof(Particle) * (size_t)n, hipMemcpyHostToDevice, c->stream));
        rc = launch_unpack(c, c->stage, 0);
        if (rc) return rc;
    } else {
        c->cx = c->ct = c->cp = 0;
    }
    c->list_valid = false;
    if (nl && nl->count && nl->neighbors && n > 0) {
        // the caller's padded list as a compact CSR (count[i] entries of
        // each row), expanded on the device
        rc = pinned_grow(&c->h_idx, &c->h_idx_cap, sizeof(int) * 2 * ((size_t)n + 1));
        if (rc) return rc;
        int *cnt = c->h_idx, *off = c->h_idx + n + 1;
        long long tot = 0;
        for (int i = 0; i < n; i++) {
            int m = nl->count[i];
            if (m < 0) m = 0;
            if (m > nl->max_neighbors) m = nl->max_neighbors;
            cnt[i] = m;
            off[i] = (int)tot;
            tot += m;
        }
        off[n] = (int)tot;
        rc = pinned_grow(&c->h_flat, &c->h_flat_cap, sizeof(int) * (size_t)(tot > 0 ? tot : 1));
        if (rc) return rc;
        int *flat = c->h_flat;
        host_for(n, [&](long long lo, long long hi) {
            for (long long i = lo; i < hi; i++)
                memcpy(flat + off[i], nl->neighbors + (size_t)i * nl->max_neighbors, sizeof(int) * cnt[i]);
        });
        if ((size_t)tot > c->nl_stage_elems) {
            DFREE(c->nl_stage);
            HIP_OK(hipMalloc(&c->nl_stage, sizeof(int) * (size_t)(tot > 0 ? tot : 1)));
            c->nl_stage_elems = (size_t)(tot > 0 ? tot : 1);
        }
        HIP_OK(hipMemcpyAsync(c->nl_stage, flat, sizeof(int) * (size_t)tot, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(c->cnt_orig, cnt, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(c->off_orig, off, sizeof(int) * (n + 1), hipMemcpyHostToDevice, c->stream));
        rc = launch_list_import(c, c->cnt_orig, c->off_orig);
        if (rc) return rc;
        c->list_valid = true;
    }
    // the staging is rewritten by the next call
    HIP_OK(hipStreamSynchronize(c->stream));
    c->has_state = true;
    return 0;
}

static int upload(Ctx *c, ParticleSystem *ps, const NeighborList *nl, const Grid &g)
{
    return upload_rows(c, ps, nl, 0, g);
}

static int download(Ctx *c, ParticleSystem *ps, NeighborList *nl)
{
    if (!c->has_state || ps->num != c->n) {
        mps_set_error(MPS_ERR_ARG, __FILE__, __LINE__, "no device state matching this ParticleSystem");
        return MPS_ERR_ARG;
    }
    const int n = ps->num;
    if (n > 0) {
        int rc = launch_pack(c, c->stage);
        if (!rc) rc = pinned_grow(&c->h_part, &c->h_part_cap, sizeof(Particle) * (size_t)n);
        if (rc) return rc;
        HIP_OK(hipMemcpyAsync(c->h_part, c->stage, sizeof(Particle) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    }
    const bool list = nl && nl->count && nl->neighbors && c->list_valid && n > 0;
    long long tot = 0;
    if (list) {
        // the device list as a compact CSR by original index
        int rc = launch_list_export(c, &tot);
        if (!rc) rc = pinned_grow(&c->h_idx, &c->h_idx_cap, sizeof(int) * 2 * ((size_t)n + 1));
        if (!rc) rc = pinned_grow(&c->h_flat, &c->h_flat_cap, sizeof(int) * (size_t)(tot > 0 ? tot : 1));
        if (rc) return rc;
        HIP_OK(hipMemcpyAsync(c->h_flat, c->nl_stage, sizeof(int) * (size_t)tot, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(c->h_idx, c->cnt_orig, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(c->h_idx + n + 1, c->off_orig, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    if (n > 0) host_copy(ps->particles, c->h_part, sizeof(Particle) * (size_t)n);
    if (list) {
        const int *flat = c->h_flat, *cnt = c->h_idx, *off = c->h_idx + n + 1;
        host_for(n, [&](long long lo, long long hi) {
            for (long long i = lo; i < hi; i++) {
                const int m = cnt[i] < nl->max_neighbors ? cnt[i] : nl->max_neighbors;
                memcpy(nl->neighbors + (size_t)i * nl->max_neighbors, flat + off[i], sizeof(int) * m);
                nl->count[i] = m;
            }
        });
    }
    return 0;
}

// ---------------------------------------------------------------------
// the timestep
// ---------------------------------------------------------------------

__global__ void k_reset_counters(DevCounters *c)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    DevCounters z{};
    z.overflow_orig = INT_MAX;
    *c = z;
}

// A kernel rather than a copy from host memory, so that a captured step
// replays it with the same effect.
static int reset_counters(Ctx *c)
{
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, c->stream, c->counters);
    HIP_OK(hipGetLastError());
    return 0;
}

// Neighbour-overflow handling: the reference prints and exits
// (neighbor_search.c:221-227).
static int check_overflow(Ctx *c, int stride)
{
    if (c->hcounters->overflow_orig == INT_MAX) return 0;
    if (g_opt.exit_on_error > 0.5) {
        fprintf(stderr, "Error: neighbor count exceeded max_neighbors (%d) for particle %d. "
                        "Increase max_neighbors in params.txt.\n",
                stride, c->hcounters->overflow_orig);
        g_writer.join();   // the snapshot being written is completed first
        exit(EXIT_FAILURE);
    }
    char msg[128];
    snprintf(msg, sizeof msg, "neighbour list overflow at particle %d", c->hcounters->overflow_orig);
    mps_set_error(MPS_ERR_NEIGHBOR_OVERFLOW, __FILE__, __LINE__, msg);
    return MPS_ERR_NEIGHBOR_OVERFLOW;
}

static int fetch_counters(Ctx *c)
{
    HIP_OK(hipMemcpyAsync(c->hcounters, c->counters, sizeof(DevCounters), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return 0;
}

// Sorts by cell and builds the list (the initial search of simulation_run,
// simulation.c:153, and the neighbor_search_cell_linked_list drop-in).
static int search_only(Ctx *c, const StepConsts &k)
{
    int rc = launch_bin(c, k);
    if (!rc) rc = launch_sort(c, k, 1);
    if (!rc) rc = reset_counters(c);
    if (!rc) rc = launch_search(c, k, 0, c->counters);
    if (!rc) rc = fetch_counters(c);
    if (rc) return rc;
    c->list_valid = true;
    c->stats.n_pairs = (long long)c->hcounters->n_pairs;
    c->stats.max_count = c->hcounters->max_count;
    return check_overflow(c, k.stride);
}

int collide_2d(Ctx *c, const StepConsts &k);   // k_particles.hip

// The timestep up to the assembled PPE (simulation.c:31-96 before
// solve_pressure): predictor, collision, sort, search with density and
// flags, assembly.
static int step_pre_launch(Ctx *c, const StepConsts &k)
{
    tape_hook(c, TAPE_START);
    int rc = launch_visc_predict(c, k);
    if (rc) return rc;
    tape_hook(c, TAPE_PREDICTED);
#if MPS_DIM == 2
    rc = collide_2d(c, k);
    if (!rc) tape_hook(c, TAPE_COLLIDED);
    if (!rc) rc = launch_collide_bin(c, k, 0);
#else
    rc = launch_collide_bin(c, k, k.collide);
    if (!rc) tape_hook(c, TAPE_COLLIDED);
#endif
    if (!rc) rc = launch_sort(c, k, 0);
    if (!rc) rc = reset_counters(c);
    if (!rc) rc = launch_search(c, k, 1, c->counters);
    if (!rc) tape_hook(c, TAPE_SEARCHED);
    if (!rc) rc = launch_assemble(c, k, c->counters);
    return rc;
}

// The particle phases through a captured graph: replays the capture made
// from the same entry state (buffer indices, particle count, constants),
// or captures one; then moves the buffer indices as the launches did.
static int step_pre_graph(Ctx *c, const StepConsts &k)
{
    const int in[4] = {c->cx, c->ct, c->cp, c->n};
    Ctx::StepGraph *hit = nullptr;
    for (auto &g : c->sg)
        if (g.exec && memcmp(g.in, in, sizeof in) == 0 && memcmp(&g.k, &k, sizeof k) == 0) hit = &g;
    if (!hit) {
        Ctx::StepGraph &g = c->sg[c->cx & 1];
        if (g.exec) hipGraphExecDestroy(g.exec);
        g = Ctx::StepGraph();
        HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = step_pre_launch(c, k);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &gr);
        const int out[3] = {c->cx, c->ct, c->cp};
        c->cx = in[0];
        c->ct = in[1];
        c->cp = in[2];
        if (!rc && e != hipSuccess) {
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(e));
            rc = MPS_ERR_HIP;
        }
        if (!rc && hipGraphInstantiate(&g.exec, gr, nullptr, nullptr, 0) != hipSuccess) {
            g.exec = nullptr;
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, "hipGraphInstantiate failed");
            rc = MPS_ERR_HIP;
        }
        if (gr) hipGraphDestroy(gr);
        if (rc) return rc;
        memcpy(g.in, in, sizeof in);
        memcpy(g.out, out, sizeof out);
        g.k = k;
        hit = &g;
    }
    HIP_OK(hipGraphLaunch(hit->exec, c->stream));
    c->cx = hit->out[0];
    c->ct = hit->out[1];
    c->cp = hit->out[2];
    return 0;
}

// The step up to the solve (simulation.c:31-95): predictor, collision,
// sort, search with density and flags, assembly.  With the option "graph"
// the launches are replayed from a capture, except while a tape records
// the step or kernels are timed.
static int step_pre(Ctx *c, const SimConfig *cf, StepConsts *kout)
{
    StepConsts k = make_consts(cf, c->ps, c->grid, search_radius(cf));
    k.stride = c->stride;
    const bool graph = g_opt.graph > 0.5 && !tape_recording() && g_opt.profile < 0.5;
    const int rc = graph ? step_pre_graph(c, k) : step_pre_launch(c, k);
    if (rc) return rc;
    c->list_valid = true;
    *kout = k;
    return 0;
}

static SolveOpts solve_opts(const SimConfig *cf, const StepConsts &k)
{
    SolveOpts so{};
    so.solver = cf->solver_type;
    so.max_iter = cf->cg_max_iter;
    so.tol = cf->cg_tolerance;
    so.tol_mode = (int)g_opt.tol_mode;
    so.precision = (int)g_opt.precision;
    so.inner_tol = g_opt.inner_tol;
    so.check_every = (int)g_opt.check_every;
    so.dense_ic = (D_ == 2);
    so.spmv_variant = (int)g_opt.spmv;
    so.fused = (int)g_opt.fused;
    so.graph = g_opt.graph > 0.5 && g_opt.profile < 0.5 && !tape_recording();
    so.neg_coeff = k.neg_coeff;
    so.re_lap = k.re_lap;
    so.r_min = k.r_min;
    return so;
}

// The timestep after the solve (simulation.c:96-118): pressure, gradient
// corrector, boundary, counters and statistics.
// In two halves: the launches with an asynchronous copy of the counters,
// then the wait and the bookkeeping, so that an ensemble can launch every
// scene's half before it waits.
static int step_post_launch(Ctx *c, const StepConsts &k)
{
    tape_hook(c, TAPE_SOLVED);
    int rc = launch_pressure_out(c, k, 0);
    if (!rc) rc = launch_grad_correct(c, k);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(c->hcounters, c->counters, sizeof(DevCounters), hipMemcpyDeviceToHost, c->stream));
    return 0;
}

// The statistics of a finished step (its counters already in hcounters)
// and the neighbour-overflow check.
static int step_bookkeep(Ctx *c, const StepConsts &k, const SolveResult &sr)
{
    c->ppe_valid = true;
    c->stats.steps++;
    c->stats.last_iters = sr.iters;
    c->stats.total_iters += sr.iters;
    c->stats.last_outer = sr.outer;
    c->stats.last_residual = sr.res;
    c->stats.last_bnorm = sr.bnorm;
    c->stats.n_eq = c->hcounters->n_eq;
    c->stats.nnz = (long long)c->hcounters->nnz;
    c->stats.n_pairs = (long long)c->hcounters->n_pairs;
    c->stats.max_count = c->hcounters->max_count;
    if (g_opt.profile > 0.5) prof_collect();
    return check_overflow(c, k.stride);
}

static int step_post_finish(Ctx *c, const StepConsts &k, const SolveResult &sr)
{
    HIP_OK(hipStreamSynchronize(c->stream));
    return step_bookkeep(c, k, sr);
}

static int step_post(Ctx *c, const StepConsts &k, const SolveResult &sr)
{
    const int rc = step_post_launch(c, k);
    return rc ? rc : step_post_finish(c, k, sr);
}

static int run_step(Ctx *c, const SimConfig *cf)
{
    StepConsts k;
    int rc = step_pre(c, cf, &k);
    if (rc) return rc;
    const SolveOpts so = solve_opts(cf, k);
    SolveResult sr{};
    rc = solve_ppe(c, so, &sr);
    if (rc) return rc;
    return step_post(c, k, sr);
}

// Resident device state of ps under configuration cf (uploaded on first
// use), with a valid neighbour list.
static int ensure_resident(Ctx *c, ParticleSystem *ps, SimConfig *cf)
{
    SimConfig *saved = g_config;
    g_config = cf;
    int rc = 0;
    const double re = search_radius(cf);
    if (!c->has_state || c->n != ps->num) rc = upload(c, ps, nullptr, grid_for(re, cf->domain_min, cf->domain_max));
    if (!rc) {
        c->resident = true;
        if (!c->list_valid) {
            StepConsts k = make_consts(cf, ps, c->grid, re);
            k.stride = c->stride;
            rc = search_only(c, k);
        }
    }
    g_config = saved;
    return rc;
}

// ---------------------------------------------------------------------
// Batches stepped by one launch per operator for all scenes (3-D, CG)
// ---------------------------------------------------------------------

int flat_pack(Ctx *const *cs, int ns, const FlatGrid &gb, int with_f32, hipStream_t s);   // k_pack.hip
int flat_unpack(Ctx *const *cs, int ns, const FlatGrid &gb, hipStream_t s);
void flat_pack_free();

// A captured phase of the flat step, valid from the buffer parity it was
// captured at (every step flips the parities, so two captures alternate).
struct FlatGraph {
    hipGraphExec_t exec = nullptr;
    int in[3] = {0, 0, 0}, out[3] = {0, 0, 0};
};

struct FlatBatchState {
    FlatGraph pre[2], post[2];
    std::vector<FlatScene> h;
    FlatScene *tab = nullptr;
    int cap = 0;
    FlatGrid gp, gb, gc;
    ScanSegs sc_cells, sc_sflag;
    int *scan_blk = nullptr;
    size_t cap_blk = 0;
    DevCounters *ctr = nullptr, *hctr = nullptr;
    int cap_ctr = 0;
};
static FlatBatchState g_flat;

static void flat_graphs_drop()
{
    for (FlatGraph *gs : {g_flat.pre, g_flat.post})
        for (int q = 0; q < 2; q++) {
            if (gs[q].exec) hipGraphExecDestroy(gs[q].exec);
            gs[q] = FlatGraph();
        }
}

static void flat_free()
{
    flat_graphs_drop();
    FlatBatchState &F = g_flat;
    void *d[] = {F.tab, F.gp.first, F.gb.first, F.gc.first, F.sc_cells.segs, F.sc_cells.tile0, F.sc_sflag.segs,
                 F.sc_sflag.tile0, F.scan_blk, F.ctr};
    for (void *p : d)
        if (p) hipFree(p);
    if (F.hctr) hipHostFree(F.hctr);
    F = FlatBatchState();
    flat_pack_free();
}

// Makes buffer index 0 current for X/V, T/O/... and P by swapping the
// halves of each ping-pong pair (no data moves), so that scenes with
// different histories share one parity.
static void ctx_parity_zero(Ctx *c)
{
    if (c->cx) {
        std::swap(c->X[0], c->X[1]);
        std::swap(c->V[0], c->V[1]);
        c->cx = 0;
        graph_reset(c);
    }
    if (c->ct) {
        std::swap(c->T[0], c->T[1]);
        std::swap(c->O[0], c->O[1]);
        std::swap(c->ACC[0], c->ACC[1]);
        std::swap(c->ND[0], c->ND[1]);
        std::swap(c->SURF[0], c->SURF[1]);
        std::swap(c->NCOUNT[0], c->NCOUNT[1]);
        c->ct = 0;
        graph_reset(c);
    }
    if (c->cp) {
        std::swap(c->P[0], c->P[1]);
        c->cp = 0;
        graph_reset(c);
    }
}

// The scene table, block offsets, scan segments and counters of a batch.
static int flat_setup(Ctx *const *cs, SimConfig *const *cfgs, int ns, hipStream_t s, FlatLaunch *f,
                      std::vector<StepConsts> &ks)
{
    FlatBatchState &F = g_flat;
    flat_graphs_drop();   // their launches hold this call's tables and sizes
    if (F.cap < ns) {
        if (F.tab) hipFree(F.tab);
        if (F.ctr) hipFree(F.ctr);
        if (F.hctr) hipHostFree(F.hctr);
        F.tab = nullptr;
        F.ctr = F.hctr = nullptr;
        HIP_OK(hipMalloc(&F.tab, sizeof(FlatScene) * ns));
        HIP_OK(hipMalloc(&F.ctr, sizeof(DevCounters) * ns));
        HIP_OK(hipHostMalloc(&F.hctr, sizeof(DevCounters) * ns, hipHostMallocDefault));
        F.cap = ns;
    }
    F.h.resize(ns);
    std::vector<int> bp(ns), bb(ns), bc(ns);
    std::vector<ScanSeg> seg_c(ns), seg_f(ns);
    const int pb = particle_block();
    int any_cmps1 = 0;
    for (int q = 0; q < ns; q++) {
        Ctx *c = cs[q];
        ctx_parity_zero(c);
        StepConsts k = make_consts(cfgs[q], c->ps, c->grid, search_radius(cfgs[q]));
        k.stride = c->stride;
        ks[q] = k;
        any_cmps1 |= k.cmps == 1;
        FlatScene &t = F.h[q];
        t.n = c->n;
        t.total = k.grid.total;
        for (int h = 0; h < 2; h++) {
            t.X[h] = c->X[h]; t.V[h] = c->V[h]; t.ACC[h] = c->ACC[h];
            t.T[h] = c->T[h]; t.O[h] = c->O[h]; t.SURF[h] = c->SURF[h]; t.NCOUNT[h] = c->NCOUNT[h];
            t.P[h] = c->P[h]; t.ND[h] = c->ND[h];
        }
        t.unk = c->unk; t.pmin = c->pmin;
        t.cell = c->cell; t.ccount = c->ccount; t.cstart = c->cstart; t.cfill = c->cfill;
        t.tmp_key = c->tmp_key; t.sflag = c->sflag; t.sscan = c->sscan;
        t.nbr = c->nbr; t.ncnt = c->ncnt;
        t.diag = c->diag; t.b = c->b; t.Aval = c->Aval; t.Acol = c->Acol; t.Alen = c->Alen;
        t.x = c->x;
        t.ctr = c->counters;
        t.k = k;
        bp[q] = (c->n + pb - 1) / pb;
        bb[q] = (c->n + 255) / 256;
        bc[q] = (k.grid.total + 1 + 255) / 256;
        seg_c[q] = ScanSeg{c->ccount, c->cstart, k.grid.total + 1, nullptr};
        seg_f[q] = ScanSeg{c->sflag, c->sscan, c->n, nullptr};
    }
    HIP_OK(hipMemcpyAsync(F.tab, F.h.data(), sizeof(FlatScene) * ns, hipMemcpyHostToDevice, s));
    int rc = flat_build_grid(&F.gp, bp.data(), ns, s);
    if (!rc) rc = flat_build_grid(&F.gb, bb.data(), ns, s);
    if (!rc) rc = flat_build_grid(&F.gc, bc.data(), ns, s);
    if (!rc) rc = scan_segs_set(&F.sc_cells, seg_c.data(), ns, s);
    if (!rc) rc = scan_segs_set(&F.sc_sflag, seg_f.data(), ns, s);
    if (rc) return rc;
    const size_t need = (size_t)std::max(F.sc_cells.ntiles, F.sc_sflag.ntiles) + ns + 64;
    if (F.cap_blk < need) {
        if (F.scan_blk) hipFree(F.scan_blk);
        F.scan_blk = nullptr;
        HIP_OK(hipMalloc(&F.scan_blk, sizeof(int) * need));
        F.cap_blk = need;
    }
    *f = FlatLaunch{F.tab, ns, F.gp, F.gb, F.gc, F.sc_cells, F.sc_sflag, F.scan_blk, any_cmps1, 0, 0, 0, s};
    return 0;
}

// The phase after the solve: unpacked solution, pressure, gradient,
// corrector and boundary, and the counters gathered in scene order.
static int flat_post_all(FlatLaunch &f)
{
    int rc = flat_unpack(nullptr, f.ns, f.gb, f.s);
    if (!rc) rc = flat_step_post(f);
    if (!rc) rc = flat_counters(f, g_flat.ctr);
    return rc;
}

// Runs phase fn from f's parity: replays the capture made from the same
// parity, or captures it into the slot of that parity and replays it; f's
// parity then moves as fn's launches move it.
static int flat_graph(FlatGraph *slots, FlatLaunch &f, int (*fn)(FlatLaunch &))
{
    const int in[3] = {f.cx, f.ct, f.cp};
    FlatGraph *hit = nullptr;
    for (int q = 0; q < 2; q++)
        if (slots[q].exec && memcmp(slots[q].in, in, sizeof in) == 0) hit = &slots[q];
    if (!hit) {
        // the two parities a phase starts from in turn differ in ct
        FlatGraph &g = slots[f.ct & 1];
        if (g.exec) hipGraphExecDestroy(g.exec);
        g = FlatGraph();
        HIP_OK(hipStreamBeginCapture(f.s, hipStreamCaptureModeThreadLocal));
        int rc = fn(f);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(f.s, &gr);
        const int out[3] = {f.cx, f.ct, f.cp};
        f.cx = in[0];
        f.ct = in[1];
        f.cp = in[2];
        if (!rc && e != hipSuccess) {
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(e));
            rc = MPS_ERR_HIP;
        }
        if (!rc && hipGraphInstantiate(&g.exec, gr, nullptr, nullptr, 0) != hipSuccess) {
            g.exec = nullptr;
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, "hipGraphInstantiate failed");
            rc = MPS_ERR_HIP;
        }
        if (gr) hipGraphDestroy(gr);
        if (rc) return rc;
        memcpy(g.in, in, sizeof in);
        memcpy(g.out, out, sizeof out);
        hit = &g;
    }
    HIP_OK(hipGraphLaunch(hit->exec, f.s));
    f.cx = hit->out[0];
    f.ct = hit->out[1];
    f.cp = hit->out[2];
    return 0;
}

// nsteps timesteps of every scene, each operator one launch for all
// scenes; the CG (FP64 or mixed) batched as in the per-scene path.  With
// option "graph" the phases before and after the solve replay captured
// graphs and the CG replays one captured chunk of iterations; only the
// packing's two waits for sizes (which decide allocations) and the CG's
// polls of its stop counter return to the host.
static int run_batch_flat(Ctx *const *cs, SimConfig *const *cfgs, int ns, int nsteps)
{
    hipStream_t s = cs[0]->stream;
    // work a caller queued on the scenes' own streams comes first
    for (int q = 0; q < ns; q++) HIP_OK(hipStreamSynchronize(cs[q]->stream));
    std::vector<StepConsts> ks(ns);
    std::vector<SolveOpts> so(ns);
    std::vector<SolveResult> sr(ns);
    FlatLaunch f;
    int rc = flat_setup(cs, cfgs, ns, s, &f, ks);
    if (rc) return rc;
    for (int q = 0; q < ns; q++) so[q] = solve_opts(cfgs[q], ks[q]);
    const int mixed = so[0].precision >= 1 && so[0].precision <= 3;
    const bool graph = g_opt.graph > 0.5 && g_opt.profile < 0.5;
    for (int step = 0; step < nsteps && !rc; step++) {
        rc = graph ? flat_graph(g_flat.pre, f, flat_step_pre) : flat_step_pre(f);
        for (int q = 0; q < ns && !rc; q++) rc = cg_alloc_ctrl(cs[q]);
        if (!rc) rc = flat_pack(cs, ns, f.gb, mixed, s);
        if (!rc)
            rc = mixed ? batch_cg_mixed(cs, ns, so.data(), sr.data(), s) : batch_cg(cs, ns, so.data(), sr.data(), s);
        if (!rc) rc = graph ? flat_graph(g_flat.post, f, flat_post_all) : flat_post_all(f);
        // the scenes' buffer indices follow the launches made, also when a
        // later part of the step failed
        for (int q = 0; q < ns; q++) {
            cs[q]->cx = f.cx;
            cs[q]->ct = f.ct;
            cs[q]->cp = f.cp;
        }
        if (rc) break;
        HIP_OK(hipMemcpyAsync(g_flat.hctr, g_flat.ctr, sizeof(DevCounters) * ns, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        for (int q = 0; q < ns && !rc; q++) {
            Ctx *c = cs[q];
            c->list_valid = true;
            *c->hcounters = g_flat.hctr[q];
            rc = step_bookkeep(c, ks[q], sr[q]);
        }
    }
    return rc;
}

// ---------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------

extern "C" {

int mps_hip_init(int device)
{
    (void)device;
    return ensure_device();
}

void mps_hip_finalize(void)
{
    for (auto &kv : g_ctx) {
        ctx_free(kv.second);
        if (kv.second->stream) hipStreamDestroy(kv.second->stream);
        delete kv.second;
    }
    g_ctx.clear();
    cg_batch_free();
    flat_free();
    g_writer.release();
}

int mps_hip_last_error(void) { return g_last_error; }

const char *mps_hip_error_string(int code)
{
    // the detailed message of the last error, when it was written for this code
    if (code != MPS_OK && code == g_msg_code && g_last_msg[0]) return g_last_msg;
    switch (code) {
    case MPS_OK: return "ok";
    case MPS_ERR_HIP: return "HIP runtime error";
    case MPS_ERR_NO_DEVICE: return "no gfx950 device";
    case MPS_ERR_ALLOC: return "device allocation failed";
    case MPS_ERR_NEIGHBOR_OVERFLOW: return "neighbour list overflow";
    case MPS_ERR_ARG: return "invalid argument";
    case MPS_ERR_NO_CONFIG: return "g_config is NULL";
    case MPS_ERR_UNSUPPORTED: return "unsupported";
    case MPS_ERR_SOLVER: return "solver failure";
    default: return "unknown error";
    }
}

const char *mps_hip_device_arch(void)
{
    if (ensure_device() != MPS_OK) return "";
    return g_arch;
}

int mps_hip_set_option(const char *key, double value)
{
    if (!key) return MPS_ERR_ARG;
    std::string k(key);
    if (k == "tol_mode") g_opt.tol_mode = value;
    else if (k == "precision") g_opt.precision = value;
    else if (k == "inner_tol") g_opt.inner_tol = value;
    else if (k == "exit_on_error") g_opt.exit_on_error = value;
    else if (k == "check_every") g_opt.check_every = value;
    else if (k == "spmv") g_opt.spmv = value;
    else if (k == "fused") g_opt.fused = value;
    else if (k == "graph") g_opt.graph = value;
    else if (k == "profile") g_opt.profile = value;
    else if (k == "verbose") g_opt.verbose = value;
    else if (k == "async_output") g_opt.async_output = value;
    else if (k == "ic_sleep") g_opt.ic_sleep = value;
    else if (k == "ic_grid") g_opt.ic_grid = value;
    else if (k == "ic_bucket") g_opt.ic_bucket = value;
    else if (k == "ic_chunk") {
        if (value != 32 && value != 64 && value != 128) return MPS_ERR_ARG;
        g_opt.ic_chunk = value;
    } else if (k == "ic_factor_form") {
        if (value != 1 && value != 2) return MPS_ERR_ARG;
        g_opt.ic_factor_form = value;
    } else if (k == "ic_waves") {
        if (value != 4 && value != 8 && value != 16) return MPS_ERR_ARG;
        g_opt.ic_waves = value;
    }
    else if (k == "ic_fgrid") g_opt.ic_fgrid = value;
    else if (k == "batch_flat") {
        if (value != 0 && value != 1) return MPS_ERR_ARG;
        g_opt.batch_flat = value;
    }
    else if (k == "search") {
        if (value != 0 && value != 1 && value != 2) return MPS_ERR_ARG;
        if (value != g_opt.search)
            for (auto &kv : g_ctx) graph_reset(kv.second);   // captures hold the kernel
        g_opt.search = value;
    }
    else if (k == "ic_wait_limit") {
        if (!(value >= 1.0)) return MPS_ERR_ARG;
        g_opt.ic_wait_limit = value;
    }
    else if (k == "pblk") {
        if (value != 128 && value != 256 && value != 512) return MPS_ERR_ARG;
        if (value != g_opt.pblk)
            for (auto &kv : g_ctx) graph_reset(kv.second);   // captures hold the launch sizes
        g_opt.pblk = value;
    } else return MPS_ERR_ARG;
    return MPS_OK;
}

double mps_hip_get_option(const char *key)
{
    if (!key) return NAN;
    std::string k(key);
    if (k == "tol_mode") return g_opt.tol_mode;
    if (k == "precision") return g_opt.precision;
    if (k == "inner_tol") return g_opt.inner_tol;
    if (k == "exit_on_error") return g_opt.exit_on_error;
    if (k == "check_every") return g_opt.check_every;
    if (k == "spmv") return g_opt.spmv;
    if (k == "fused") return g_opt.fused;
    if (k == "graph") return g_opt.graph;
    if (k == "profile") return g_opt.profile;
    if (k == "verbose") return g_opt.verbose;
    if (k == "pblk") return g_opt.pblk;
    if (k == "async_output") return g_opt.async_output;
    if (k == "ic_wait_limit") return g_opt.ic_wait_limit;
    if (k == "ic_sleep") return g_opt.ic_sleep;
    if (k == "ic_grid") return g_opt.ic_grid;
    if (k == "ic_bucket") return g_opt.ic_bucket;
    if (k == "ic_chunk") return g_opt.ic_chunk;
    if (k == "ic_waves") return g_opt.ic_waves;
    if (k == "ic_factor_form") return g_opt.ic_factor_form;
    if (k == "ic_fgrid") return g_opt.ic_fgrid;
    if (k == "search") return g_opt.search;
    if (k == "batch_flat") return g_opt.batch_flat;
    return NAN;
}

int mps_hip_upload(ParticleSystem *ps, const NeighborList *nl)
{
    if (!ps) return MPS_ERR_ARG;
    if (!g_config) return g_last_error = MPS_ERR_NO_CONFIG;
    Ctx *c = get_ctx(ps);
    if (!c) return g_last_error;
    const double re = search_radius(g_config);
    return upload(c, ps, nl, grid_for(re, g_config->domain_min, g_config->domain_max));
}

int mps_hip_download(ParticleSystem *ps, NeighborList *nl)
{
    Ctx *c = ps ? find_ctx(ps) : nullptr;
    if (!c) return MPS_ERR_ARG;
    return download(c, ps, nl);
}

int mps_hip_set_resident(ParticleSystem *ps, int resident)
{
    Ctx *c = ps ? get_ctx(ps) : nullptr;
    if (!c) return MPS_ERR_ARG;
    c->resident = resident != 0;
    return MPS_OK;
}

void mps_hip_release(ParticleSystem *ps)
{
    auto it = g_ctx.find(ps);
    if (it == g_ctx.end()) return;
    ctx_free(it->second);
    if (it->second->stream) hipStreamDestroy(it->second->stream);
    delete it->second;
    g_ctx.erase(it);
}

// Resident multi-step path: upload once (with the initial search when no
// list is held), then steps with no host copies.
int mps_hip_run_steps(ParticleSystem *ps, int nsteps)
{
    if (!ps) return MPS_ERR_ARG;
    if (!g_config) return g_last_error = MPS_ERR_NO_CONFIG;
    Ctx *c = get_ctx(ps);
    if (!c) return g_last_error;
    int rc;
    const double re = search_radius(g_config);
    if (!c->has_state || c->n != ps->num) {
        rc = upload(c, ps, nullptr, grid_for(re, g_config->domain_min, g_config->domain_max));
        if (rc) return rc;
    }
    c->resident = true;
    if (!c->list_valid) {
        StepConsts k = make_consts(g_config, ps, c->grid, re);
        k.stride = c->stride;
        rc = search_only(c, k);
        if (rc) return rc;
    }
    for (int s = 0; s < nsteps; s++) {
        rc = run_step(c, g_config);
        if (rc) return rc;
    }
    return MPS_OK;
}

int mps_hip_run_batch(ParticleSystem *const *ps, SimConfig *const *cfgs, int nscenes, int nsteps)
{
    if (!ps || !cfgs || nscenes <= 0 || nsteps < 0) return g_last_error = MPS_ERR_ARG;
    int rc = ensure_device();
    if (rc) return rc;
    std::vector<Ctx *> cs(nscenes);
    for (int q = 0; q < nscenes; q++) {
        if (!ps[q] || !cfgs[q]) return g_last_error = MPS_ERR_ARG;
        cs[q] = get_ctx(ps[q]);
        if (!cs[q]) return g_last_error;
        rc = ensure_resident(cs[q], ps[q], cfgs[q]);
        if (rc) return rc;
    }
    // CG (FP64 or mixed) is solved batched; ICCG runs per scene
    bool batched = true;
    for (int q = 0; q < nscenes; q++) batched = batched && cfgs[q]->solver_type == 0;
#if MPS_DIM == 3 && !defined(MPS_FRICTION_FIELDS)
    if (batched && g_opt.batch_flat > 0.5 && !tape_recording()) {
        SimConfig *saved = g_config;
        rc = run_batch_flat(cs.data(), cfgs, nscenes, nsteps);
        g_config = saved;
        if (rc) g_last_error = rc;
        return rc;
    }
#endif
    std::vector<StepConsts> ks(nscenes);
    std::vector<SolveOpts> so(nscenes);
    std::vector<SolveResult> sr(nscenes);
    hipStream_t bstream = cs[0]->stream;
    SimConfig *saved = g_config;
    for (int step = 0; step < nsteps && !rc; step++) {
        for (int q = 0; q < nscenes && !rc; q++) {
            g_config = cfgs[q];
            rc = step_pre(cs[q], cfgs[q], &ks[q]);
            so[q] = solve_opts(cfgs[q], ks[q]);
        }
        if (rc) break;
        if (batched) {
            const int mixed = so[0].precision >= 1 && so[0].precision <= 3;
            // each packing stage for every scene before its waits, so the
            // scenes' packing kernels overlap
            for (int q = 0; q < nscenes && !rc; q++) {
                rc = cg_alloc_ctrl(cs[q]);
                if (!rc) rc = pack_stage_a(cs[q]);
            }
            for (int q = 0; q < nscenes && !rc; q++) {
                rc = pack_wait(cs[q]);
                if (!rc) rc = pack_stage_b(cs[q]);
            }
            for (int q = 0; q < nscenes && !rc; q++) {
                rc = pack_wait(cs[q]);
                if (!rc) rc = pack_stage_c(cs[q], mixed);
            }
            for (int q = 0; q < nscenes && !rc; q++) rc = pack_wait(cs[q]);
            if (!rc)
                rc = mixed ? batch_cg_mixed(cs.data(), nscenes, so.data(), sr.data(), bstream)
                           : batch_cg(cs.data(), nscenes, so.data(), sr.data(), bstream);
            for (int q = 0; q < nscenes && !rc; q++) rc = unpack_solution(cs[q]);
        } else {
            for (int q = 0; q < nscenes && !rc; q++) {
                g_config = cfgs[q];
                rc = solve_ppe(cs[q], so[q], &sr[q]);
            }
        }
        for (int q = 0; q < nscenes && !rc; q++) {
            g_config = cfgs[q];
            rc = step_post_launch(cs[q], ks[q]);
        }
        for (int q = 0; q < nscenes && !rc; q++) {
            g_config = cfgs[q];
            rc = step_post_finish(cs[q], ks[q], sr[q]);
        }
    }
    g_config = saved;
    if (rc) g_last_error = rc;
    return rc;
}

// Times the bit-preserving choices of the FP64 solve on a private device
// copy of ps (its host particles, under g_config): one step up to the
// assembly, then each SpMV variant on the packed system, then whole solves
// of that system with each host polling interval.  Neither choice changes
// a result bit: the variants sum every row in the same order, and the
// polling interval only decides how far past convergence the host has
// queued iterations that exit at once.
int mps_hip_autotune(ParticleSystem *ps, int apply, mps_tune_result *out)
{
    if (!ps || !out) return g_last_error = MPS_ERR_ARG;
    if (!g_config) return g_last_error = MPS_ERR_NO_CONFIG;
    int rc = ensure_device();
    if (rc) return rc;
    memset(out, 0, sizeof *out);
    ParticleSystem shadow = *ps;     // same host particles, a context of its own
    Ctx *c = get_ctx(&shadow);
    if (!c) return g_last_error;
    const Options saved = g_opt;
    g_opt.profile = 0;
    g_opt.graph = 0;
    g_opt.precision = 0;
    g_opt.fused = 0;
    rc = ensure_resident(c, &shadow, g_config);
    StepConsts k;
    if (!rc) rc = step_pre(c, g_config, &k);
    SolveOpts so{};
    SolveResult sr{};
    if (!rc) {
        so = solve_opts(g_config, k);
        so.solver = 0;
        rc = solve_ppe(c, so, &sr);
    }
    // variant 2 (matrix through the caches) is timed for the record but not
    // chosen: it ties where a system's matrix fits the 256 MB Infinity
    // Cache and is 14 % slower where it does not (1.08M particles), and the
    // choice carries over to batches of scenes, whose matrices never fit
    int best_v = 0;
    for (int v = 0; v < 3 && !rc; v++) {
        rc = cg_time_spmv(c, v, 20, &out->spmv_ms[v]);
        if (!rc && v < 2 && (v == 0 || out->spmv_ms[v] < out->spmv_ms[best_v])) best_v = v;
    }
    const int ce[4] = {8, 16, 32, 64};
    int best_c = 0;
    for (int q = 0; q < 4 && !rc; q++) {
        so.spmv_variant = best_v;
        so.check_every = ce[q];
        if (hipStreamSynchronize(c->stream) != hipSuccess) rc = MPS_ERR_HIP;
        const auto t0 = std::chrono::steady_clock::now();
        if (!rc) rc = solve_ppe(c, so, &sr);
        out->solve_ms[q] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (!rc && (q == 0 || out->solve_ms[q] < out->solve_ms[best_c])) best_c = q;
    }
    // the particle phases (predictor to assembly) at each block size of the
    // per-particle kernels: one untimed pass, then three timed ones
    const int pb[3] = {128, 256, 512};
    int best_p = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = MPS_ERR_HIP;
    g_opt.search = 1;
    for (int q = 0; q < 3 && !rc; q++) {
        g_opt.pblk = pb[q];
        rc = step_pre_launch(c, k);
        if (!rc && hipEventRecord(e0, c->stream) != hipSuccess) rc = MPS_ERR_HIP;
        for (int r = 0; r < 3 && !rc; r++) rc = step_pre_launch(c, k);
        float ms = 0.0f;
        if (!rc && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = MPS_ERR_HIP;
        out->pblk_ms[q] = ms / 3.0;
        if (!rc && (q == 0 || out->pblk_ms[q] < out->pblk_ms[best_p])) best_p = q;
    }
    // at that block size, the search reading 4 (option "search" 1) or 2
    // (2) staged candidates per batch
    int best_s = 1;
    out->search_ms[0] = out->pblk_ms[best_p];
    if (!rc) {
        g_opt.pblk = pb[best_p];
        g_opt.search = 2;
        rc = step_pre_launch(c, k);
        if (!rc && hipEventRecord(e0, c->stream) != hipSuccess) rc = MPS_ERR_HIP;
        for (int r = 0; r < 3 && !rc; r++) rc = step_pre_launch(c, k);
        float ms = 0.0f;
        if (!rc && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = MPS_ERR_HIP;
        out->search_ms[1] = ms / 3.0;
        if (!rc && out->search_ms[1] < out->search_ms[0]) best_s = 2;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    g_opt = saved;
    mps_hip_release(&shadow);
    if (rc) return g_last_error = rc;
    out->spmv = best_v;
    out->check_every = ce[best_c];
    out->iters = sr.iters;
    out->pblk = pb[best_p];
    out->search = best_s;
    if (apply) {
        g_opt.spmv = best_v;
        g_opt.check_every = ce[best_c];
        mps_hip_set_option("pblk", pb[best_p]);
        mps_hip_set_option("search", best_s);
    }
    return MPS_OK;
}

// The block size of the per-particle kernels for a batch stepped with one
// launch per operator: the batch's particle phases (predictor to assembly,
// all scenes) timed three times at each size after an untimed pass, on private
// device copies of the scenes' host particles.  The caller's scenes and
// their device state are not touched.
int mps_hip_autotune_batch(ParticleSystem *const *ps, SimConfig *const *cfgs, int nscenes, int apply,
                           mps_tune_result *out)
{
    if (!ps || !cfgs || nscenes <= 0 || !out) return g_last_error = MPS_ERR_ARG;
    int rc = ensure_device();
    if (rc) return rc;
    memset(out, 0, sizeof *out);
#if MPS_DIM == 3 && !defined(MPS_FRICTION_FIELDS)
    for (int q = 0; q < nscenes; q++)
        if (!ps[q] || !cfgs[q]) return g_last_error = MPS_ERR_ARG;
    std::vector<ParticleSystem> shadow(nscenes);
    std::vector<Ctx *> cs(nscenes, nullptr);
    const Options saved = g_opt;
    SimConfig *saved_cfg = g_config;
    g_opt.profile = 0;
    g_opt.graph = 0;
    for (int q = 0; q < nscenes && !rc; q++) {
        shadow[q] = *ps[q];   // same host particles, a context of its own
        cs[q] = get_ctx(&shadow[q]);
        rc = cs[q] ? ensure_resident(cs[q], &shadow[q], cfgs[q]) : g_last_error;
    }
    const int pb[3] = {128, 256, 512};
    int best = 0, best_s = 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = MPS_ERR_HIP;
    // block sizes with search form 1, then form 2 at the best size
    g_opt.search = 1;
    for (int k = 0; k < 4 && !rc; k++) {
        if (k == 3) {
            g_opt.search = 2;
            g_opt.pblk = pb[best];
        } else {
            g_opt.pblk = pb[k];
        }
        std::vector<StepConsts> ks(nscenes);
        FlatLaunch f;
        hipStream_t s = cs[0]->stream;
        rc = flat_setup(cs.data(), cfgs, nscenes, s, &f, ks);
        if (!rc) rc = flat_step_pre(f);
        if (!rc && hipEventRecord(e0, s) != hipSuccess) rc = MPS_ERR_HIP;
        for (int r = 0; r < 3 && !rc; r++) rc = flat_step_pre(f);
        float ms = 0.0f;
        if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = MPS_ERR_HIP;
        for (int q = 0; q < nscenes; q++) {
            cs[q]->cx = f.cx;
            cs[q]->ct = f.ct;
            cs[q]->cp = f.cp;
        }
        if (k < 3) {
            out->pblk_ms[k] = ms / 3.0;
            if (!rc && k > 0 && out->pblk_ms[k] < out->pblk_ms[best]) best = k;
        } else {
            out->search_ms[0] = out->pblk_ms[best];
            out->search_ms[1] = ms / 3.0;
            if (!rc && out->search_ms[1] < out->search_ms[0]) best_s = 2;
        }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    g_opt = saved;
    g_config = saved_cfg;
    for (int q = 0; q < nscenes; q++)
        if (cs[q]) mps_hip_release(&shadow[q]);
    if (rc) return g_last_error = rc;
    out->pblk = pb[best];
    out->search = best_s;
    if (apply) {
        mps_hip_set_option("pblk", pb[best]);
        mps_hip_set_option("search", best_s);
    }
    return MPS_OK;
#else
    (void)apply;
    mps_set_error(MPS_ERR_UNSUPPORTED, __FILE__, __LINE__, "batch tuning is for the 3-D library");
    return g_last_error = MPS_ERR_UNSUPPORTED;
#endif
}

int mps_hip_get_stats(ParticleSystem *ps, mps_step_stats *out)
{
    Ctx *c = ps ? find_ctx(ps) : nullptr;
    if (!c || !out) return MPS_ERR_ARG;
    *out = c->stats;
    return MPS_OK;
}

int mps_hip_profile_get(const char *name, double *total_ms, long long *count)
{
    prof_collect();
    for (int k = 0; k < PK_COUNT; k++)
        if (name && strcmp(name, g_prof[k].name) == 0) {
            if (total_ms) *total_ms = g_prof[k].total_ms;
            if (count) *count = g_prof[k].count;
            return 0;
        }
    return -1;
}

void mps_hip_profile_reset(void)
{
    prof_collect();
    for (int k = 0; k < PK_COUNT; k++) {
        g_prof[k].total_ms = 0;
        g_prof[k].count = 0;
    }
}

// --- reference drop-ins: driver ----------------------------------------

// simulation_step (simulation.c:19-118).  The CellList supplies the grid
// geometry; nl is the stale list this step's viscosity and collision read
// and receives the new list.
void simulation_step(ParticleSystem *ps, NeighborList *nl, CellList *cl, int step)
{
    (void)step;
    mps_clear_error();
    if (!g_config) { g_last_error = MPS_ERR_NO_CONFIG; return; }
    Ctx *c = get_ctx(ps);
    if (!c) return;
    const double re = search_radius(g_config);
    Grid g = cl ? grid_from_cells(cl) : grid_for(re, g_config->domain_min, g_config->domain_max);
    int rc = 0;
    if (!c->resident || !c->has_state || c->n != ps->num) {
        rc = upload(c, ps, nl, g);
        if (rc) return;
    }
    c->grid = g;
    rc = run_step(c, g_config);
    if (rc) return;
    if (!c->resident) download(c, ps, nl);
}

// simulation_run (simulation.c:123-181): resident state, one download per
// output step.
void simulation_run(ParticleSystem *ps)
{
    mps_clear_error();
    if (!g_config) { g_last_error = MPS_ERR_NO_CONFIG; return; }
    const double re = search_radius(g_config);
    const double dt = g_config->dt;
    const int total_steps = (int)(g_config->t_end / dt);
    const int out_interval = g_config->output_interval;
    const char *out_dir = g_config->output_dir;
    printf("HIP device: %s\n", mps_hip_device_arch());
    NeighborList *nl = neighbor_list_create(ps->num, g_config->max_neighbors);
    CellList *cl = cell_list_create(ps->num, re, g_config->domain_min, g_config->domain_max);
    if (!nl || !cl) {
        fprintf(stderr, "Error: failed to create neighbor/cell list\n");
        neighbor_list_free(nl);
        cell_list_free(cl);
        return;
    }
    // every early return frees the lists and leaves the error code set
    auto fail = [&](const char *what) {
        fprintf(stderr, "Error: %s%s\n", what, mps_hip_error_string(g_last_error));
        cell_list_free(cl);
        neighbor_list_free(nl);
    };
    Ctx *c = get_ctx(ps);
    if (!c) { fail(""); return; }
    Grid g = grid_from_cells(cl);
    if (upload(c, ps, nullptr, g)) { fail("upload failed: "); return; }
    c->resident = true;
    StepConsts k = make_consts(g_config, ps, g, re);
    k.stride = c->stride;
    if (search_only(c, k)) { c->resident = false; fail(""); return; }
    // The step-0 writer reads ps->particles itself: nothing writes them
    // before the final download, which waits for every writer.
    bool async_out = g_opt.async_output > 0.5;
    if (async_out && out_interval > 0 && total_steps >= out_interval &&
        g_writer.reserve((size_t)ps->num) != MPS_OK)
        async_out = false;
    g_writer.join();
    if (async_out) {
        start_snapshot(*ps, 0, out_dir);
    } else {
        output_csv(ps, 0, out_dir);
        output_vtk(ps, 0, out_dir);
    }
    printf("Starting simulation (%dD, HIP %s): %d steps, dt = %.2e\n", D_, g_arch, total_steps, dt);
    for (int step = 1; step <= total_steps; step++) {
        if (run_step(c, g_config)) {
            fprintf(stderr, "Error at step %d: %s\n", step, g_last_msg);
            break;
        }
        if (out_interval > 0 && step % out_interval == 0) {
            ParticleSystem snap = *ps;
            if (async_out) snap.particles = g_writer.buf[g_writer.next];
            if (download(c, &snap, nullptr)) {
                fprintf(stderr, "Error at step %d: %s\n", step, g_last_msg);
                break;
            }
            int fluid = 0;
            for (int i = 0; i < snap.num; i++) fluid += snap.particles[i].type == FLUID_PARTICLE;
            printf("Step %6d / %d  (t = %.4f s)  fluid particles: %d\n", step, total_steps,
                   step * dt, fluid);
            if (async_out) {
                g_writer.join();
                start_snapshot(snap, step, out_dir);
                g_writer.next ^= 1;
            } else {
                output_csv(ps, step, out_dir);
                output_vtk(ps, step, out_dir);
            }
        }
    }
    g_writer.release();
    download(c, ps, nl);
    c->resident = false;
    printf("Simulation complete.\n");
    cell_list_free(cl);
    neighbor_list_free(nl);
}

}  // extern "C"

// accessors for the other translation units
Ctx *mps_ctx_for(ParticleSystem *ps, int create) { return create ? get_ctx(ps) : find_ctx(ps); }
int mps_ctx_step(Ctx *c, const SimConfig *cf) { return run_step(c, cf); }
int mps_ctx_upload(Ctx *c, ParticleSystem *ps, const NeighborList *nl, const Grid &g) { return upload(c, ps, nl, g); }
int mps_ctx_upload_rows(Ctx *c, ParticleSystem *ps, int stride, const Grid &g) { return upload_rows(c, ps, nullptr, stride, g); }
int mps_ctx_download(Ctx *c, ParticleSystem *ps, NeighborList *nl) { return download(c, ps, nl); }
int mps_ctx_search(Ctx *c, const StepConsts &k) { return search_only(c, k); }
StepConsts mps_consts(const SimConfig *cf, const ParticleSystem *ps, const Grid &g, double re)
{
    return make_consts(cf, ps, g, re);
}
Grid mps_grid_for(double re, const double *dmin, const double *dmax) { return grid_for(re, dmin, dmax); }
Grid mps_grid_from_cells(const CellList *cl) { return grid_from_cells(cl); }
double mps_search_radius(const SimConfig *cf) { return search_radius(cf); }
