// dropins.hip -- the reference's per-function entry points, the host-side
// containers they use, and the PPE/solution exports.
//
// Each operator call runs on the GPU.  If the ParticleSystem is resident
// the device state is used as is; otherwise the call uploads the caller's
// particles (and list), runs its kernel(s), and downloads the fields the
// reference function writes, so a caller composing its own timestep from
// these functions sees the reference's results.
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ctx.hpp"

int extra_surface(Ctx *c, const StepConsts &k);
int extra_count(Ctx *c, const StepConsts &k);
int extra_surface_count(Ctx *c, const StepConsts &k);
int extra_clamp(Ctx *c);
int extra_wall(Ctx *c);
int extra_oob(Ctx *c, const StepConsts &k);
int extra_viscosity(Ctx *c, const StepConsts &k);
int extra_surface_tension(Ctx *c, const StepConsts &k, int solid);
int extra_gradient(Ctx *c, const StepConsts &k);
int extra_ppe_rows(Ctx *c, const StepConsts &k);
int extra_brute(Ctx *c, double re2);
int extra_init_params(Ctx *c, const StepConsts &k, const Grid &g, double reach2, double *out_host);
int extra_export_rows(Ctx *c, std::vector<int> &off, std::vector<int> &col, std::vector<double> &val);
int collide_2d(Ctx *c, const StepConsts &k);

// ---------------------------------------------------------------------
// host containers (neighbor_search.c:10-35, :92-131), kernel function
// ---------------------------------------------------------------------

extern "C" {

NeighborList *neighbor_list_create(int num_particles, int max_neighbors)
{
    NeighborList *nl = (NeighborList *)calloc(1, sizeof(NeighborList));
    if (!nl) return NULL;
    size_t cells = (size_t)(num_particles > 0 ? num_particles : 1) * (size_t)(max_neighbors > 0 ? max_neighbors : 1);
    nl->neighbors = (int *)malloc(sizeof(int) * cells);
    nl->count = (int *)calloc((size_t)(num_particles > 0 ? num_particles : 1), sizeof(int));
    nl->max_neighbors = max_neighbors;
    if (!nl->neighbors || !nl->count) {
        free(nl->neighbors);
        free(nl->count);
        free(nl);
        return NULL;
    }
    return nl;
}

void neighbor_list_free(NeighborList *nl)
{
    if (!nl) return;
    free(nl->neighbors);
    free(nl->count);
    free(nl);
}

CellList *cell_list_create(int num_particles, double re, const double domain_min[MPS_DIM],
                           const double domain_max[MPS_DIM])
{
    CellList *cl = (CellList *)calloc(1, sizeof(CellList));
    if (!cl) return NULL;
    Grid g = mps_grid_for(re, domain_min, domain_max);
    cl->cell_size = re;
    cl->nx = g.nc[0];
    cl->ny = g.nc[1];
#if MPS_DIM == 3
    cl->nz = g.nc[2];
#endif
    cl->total_cells = g.total;
    for (int d = 0; d < MPS_DIM; d++) cl->origin[d] = g.origin[d];
    cl->head = (int *)malloc(sizeof(int) * (size_t)(g.total > 0 ? g.total : 1));
    cl->next = (int *)malloc(sizeof(int) * (size_t)(num_particles > 0 ? num_particles : 1));
    if (!cl->head || !cl->next) {
        cell_list_free(cl);
        return NULL;
    }
    return cl;
}

void cell_list_free(CellList *cl)
{
    if (!cl) return;
    free(cl->head);
    free(cl->next);
    free(cl);
}

int neighbor_get(NeighborList *nl, int i, int j) { return nl->neighbors[(size_t)i * nl->max_neighbors + j]; }

// kernel.c:12-30 -- scalar helpers for host callers; the GPU kernels use
// the same formula (wfun in dev_common.hpp).
double kernel_weight(double r, double re)
{
    if (r >= re) return 0.0;
    double r_min = 0.01 * g_config->particle_distance;
    if (r < r_min) r = r_min;
    return re / r - 1.0;
}

double kernel_weight_derivative(double r, double re)
{
    if (r >= re) return 0.0;
    double r_min = 0.01 * g_config->particle_distance;
    if (r < r_min) r = r_min;
    return -re / (r * r);
}

}  // extern "C"

// ---------------------------------------------------------------------
// call scaffolding
// ---------------------------------------------------------------------

struct Call {
    Ctx *c = nullptr;
    bool synced = false;   // state was uploaded by this call
    StepConsts k{};
};

static bool begin_call(Call &cl, ParticleSystem *ps, const NeighborList *nl, bool need_list)
{
    mps_clear_error();
    if (!ps || !g_config) return false;
    cl.c = mps_ctx_for(ps, 1);
    if (!cl.c) return false;
    const double re = mps_search_radius(g_config);
    Grid g = mps_grid_for(re, g_config->domain_min, g_config->domain_max);
    if (!cl.c->resident || !cl.c->has_state || cl.c->n != ps->num) {
        if (mps_ctx_upload(cl.c, ps, need_list ? nl : nullptr, g)) return false;
        cl.synced = true;
    }
    cl.k = mps_consts(g_config, ps, cl.c->grid, re);
    cl.k.stride = cl.c->stride;
    return true;
}

static void end_call(Call &cl, ParticleSystem *ps)
{
    if (cl.synced) mps_ctx_download(cl.c, ps, nullptr);
}

// ---------------------------------------------------------------------
// reference entry points
// ---------------------------------------------------------------------

extern "C" {

// neighbor_search_cell_linked_list (neighbor_search.c:146-240)
void neighbor_search_cell_linked_list(NeighborList *nl, ParticleSystem *ps, CellList *cl, double re)
{
    mps_clear_error();
    if (!ps || !nl || !g_config) return;
    Ctx *c = mps_ctx_for(ps, 1);
    if (!c) return;
    Grid g = cl ? mps_grid_from_cells(cl) : mps_grid_for(re, g_config->domain_min, g_config->domain_max);
    bool synced = false;
    if (!c->resident || !c->has_state || c->n != ps->num) {
        // rows of nl's capacity: an overflow is found against nl->max_neighbors
        if (mps_ctx_upload_rows(c, ps, nl->max_neighbors, g)) return;
        synced = true;
    }
    c->grid = g;
    StepConsts k = mps_consts(g_config, ps, g, re);
    k.stride = c->stride;
    if (mps_ctx_search(c, k)) return;
    if (synced) {
        // list into nl; the particle data itself is unchanged
        mps_ctx_download(c, ps, nl);
    }
    if (cl && cl->head && cl->next && c->n > 0) {
        // the linked-list view of the same cells: head = largest index in a
        // cell, next = the next smaller index (neighbor_search.c:180-182)
        std::vector<int> O(c->n), cs((size_t)g.total + 2);
        if (hipMemcpy(O.data(), c->O[c->ct], sizeof(int) * c->n, hipMemcpyDeviceToHost) != hipSuccess) return;
        if (hipMemcpy(cs.data(), c->cstart, sizeof(int) * ((size_t)g.total + 2), hipMemcpyDeviceToHost) != hipSuccess) return;
        // head[] is indexed in the reference's cell order (z slowest); the
        // device stores cells in cell_key order
        for (int s = 0; s < c->n; s++) cl->next[O[s]] = -1;
        for (int cz = 0; cz < g.nc[2]; cz++)
            for (int cy = 0; cy < g.nc[1]; cy++)
                for (int cx = 0; cx < g.nc[0]; cx++) {
                    const int q = (cz * g.nc[1] + cy) * g.nc[0] + cx;
                    const int key = cell_key(g, cx, cy, cz);
                    cl->head[q] = cs[key] < cs[key + 1] ? O[cs[key]] : -1;
                    for (int s = cs[key]; s + 1 < cs[key + 1]; s++) cl->next[O[s]] = O[s + 1];
                }
    }
}

// neighbor_search_brute_force (neighbor_search.c:44-81)
void neighbor_search_brute_force(NeighborList *nl, ParticleSystem *ps, double re)
{
    mps_clear_error();
    if (!ps || !nl || !g_config) return;
    Ctx *c = mps_ctx_for(ps, 1);
    if (!c) return;
    if (c->resident) {
        // brute force lists indices in ascending original order, which the
        // identity order of a fresh upload provides
        mps_ctx_download(c, ps, nullptr);
        c->resident = false;
    }
    const double rs = mps_search_radius(g_config);
    if (mps_ctx_upload_rows(c, ps, nl->max_neighbors, mps_grid_for(rs, g_config->domain_min, g_config->domain_max)))
        return;
    DevCounters z{};
    z.overflow_orig = INT_MAX;
    if (hipMemcpy(c->counters, &z, sizeof z, hipMemcpyHostToDevice) != hipSuccess) return;
    if (extra_brute(c, re * re)) return;
    DevCounters h;
    if (hipMemcpy(&h, c->counters, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) return;
    if (h.overflow_orig != INT_MAX) {
        fprintf(stderr, "Error: neighbor count exceeded max_neighbors (%d) for particle %d. "
                        "Increase max_neighbors in params.txt.\n", c->stride, h.overflow_orig);
        if (mps_hip_get_option("exit_on_error") > 0.5) exit(EXIT_FAILURE);
        char msg[128];
        snprintf(msg, sizeof msg, "neighbour list overflow at particle %d", h.overflow_orig);
        mps_set_error(MPS_ERR_NEIGHBOR_OVERFLOW, __FILE__, __LINE__, msg);
        return;
    }
    c->list_valid = true;
    mps_ctx_download(c, ps, nl);
}

void calc_particle_number_density(ParticleSystem *ps, NeighborList *nl)
{
    Call cl;
    if (!begin_call(cl, ps, nl, true)) return;
    launch_density_only(cl.c, cl.k);
    end_call(cl, ps);
}

void calc_viscosity_term(ParticleSystem *ps, NeighborList *nl)
{
    Call cl;
    if (!begin_call(cl, ps, nl, true)) return;
    extra_viscosity(cl.c, cl.k);
    end_call(cl, ps);
}

#ifdef MPS_FRICTION_FIELDS
// apply_friction (mps_fluid3D/src/operators.c:337-463): the hybrid wall
// friction on the current velocities, over nl
void apply_friction(ParticleSystem *ps, NeighborList *nl)
{
    if (!g_config || !g_config->friction_enabled) return;
    Call cl;
    if (!begin_call(cl, ps, nl, true)) return;
    launch_friction(cl.c, cl.k);
    end_call(cl, ps);
}
#endif

void calc_pressure_gradient(ParticleSystem *ps, NeighborList *nl)
{
    Call cl;
    if (!begin_call(cl, ps, nl, true)) return;
    extra_gradient(cl.c, cl.k);
    end_call(cl, ps);
}

void judge_free_surface(ParticleSystem *ps, double threshold)
{
    Call cl;
    if (!begin_call(cl, ps, nullptr, false)) return;
    cl.k.thr_n0 = threshold * ps->n0;
    extra_surface(cl.c, cl.k);
    end_call(cl, ps);
}

void calc_neighbor_count(ParticleSystem *ps, NeighborList *nl)
{
    Call cl;
    if (!begin_call(cl, ps, nl, true)) return;
    extra_count(cl.c, cl.k);
    end_call(cl, ps);
}

void judge_free_surface_by_count(ParticleSystem *ps, double beta)
{
    Call cl;
    if (!begin_call(cl, ps, nullptr, false)) return;
    cl.k.beta_n0c = beta * ps->n0_count;
    extra_surface_count(cl.c, cl.k);
    end_call(cl, ps);
}

void clamp_pressure(ParticleSystem *ps)
{
    Call cl;
    if (!begin_call(cl, ps, nullptr, false)) return;
    extra_clamp(cl.c);
    end_call(cl, ps);
}

#if MPS_DIM == 3
// collision (operators.c:371-446) over the stale list nl
void collision(ParticleSystem *ps, NeighborList *nl)
{
    Call cl;
    if (!begin_call(cl, ps, nl, true)) return;
    launch_collide_bin(cl.c, cl.k, 1);
    end_call(cl, ps);
}
#else
// collision (mps_fluid2D/src/operators.c:324-392), all pairs
void collision(ParticleSystem *ps)
{
    Call cl;
    if (!begin_call(cl, ps, nullptr, false)) return;
    collide_2d(cl.c, cl.k);
    end_call(cl, ps);
}
#endif

void calc_surface_tension(ParticleSystem *ps, NeighborList *nl)
{
    Call cl;
    if (!begin_call(cl, ps, nl, true)) return;
    extra_surface_tension(cl.c, cl.k, 0);
    end_call(cl, ps);
}

void calc_surface_tension_SL(ParticleSystem *ps, NeighborList *nl)
{
    if (ps && ps->C_SL == 0.0) return;
    Call cl;
    if (!begin_call(cl, ps, nl, true)) return;
    extra_surface_tension(cl.c, cl.k, 1);
    end_call(cl, ps);
}

void apply_wall_boundary(ParticleSystem *ps)
{
    Call cl;
    if (!begin_call(cl, ps, nullptr, false)) return;
    extra_wall(cl.c);
    end_call(cl, ps);
}

#if MPS_DIM == 3
void remove_out_of_bounds(ParticleSystem *ps, double xmin, double xmax, double ymin, double ymax,
                          double zmin, double zmax)
#else
void remove_out_of_bounds(ParticleSystem *ps, double xmin, double xmax, double ymin, double ymax)
#endif
{
    Call cl;
    if (!begin_call(cl, ps, nullptr, false)) return;
    const double margin = g_config->particle_distance * (g_config->wall_layers + 1);
    cl.k.bnd_lo[0] = xmin - margin;
    cl.k.bnd_hi[0] = xmax + margin;
    cl.k.bnd_lo[1] = ymin - margin;
    cl.k.bnd_hi[1] = ymax + margin;
#if MPS_DIM == 3
    cl.k.bnd_lo[2] = zmin - margin;
    cl.k.bnd_hi[2] = zmax + margin;
#endif
    extra_oob(cl.c, cl.k);
    end_call(cl, ps);
}

// solve_pressure (pressure_solver.c:368-570): unknowns, assembly, CG/ICCG,
// scatter -- from the caller's list, densities and flags.
void solve_pressure(ParticleSystem *ps, NeighborList *nl)
{
    mps_clear_error();
    Call cl;
    if (!begin_call(cl, ps, nl, true)) return;
    Ctx *c = cl.c;
    DevCounters z{};
    z.overflow_orig = INT_MAX;
    if (hipMemcpy(c->counters, &z, sizeof z, hipMemcpyHostToDevice) != hipSuccess) return;
    if (extra_ppe_rows(c, cl.k) || launch_assemble(c, cl.k, c->counters)) return;
    SolveOpts so{};
    so.solver = g_config->solver_type;
    so.max_iter = g_config->cg_max_iter;
    so.tol = g_config->cg_tolerance;
    so.tol_mode = (int)mps_hip_get_option("tol_mode");
    so.precision = (int)mps_hip_get_option("precision");
    so.inner_tol = mps_hip_get_option("inner_tol");
    so.check_every = (int)mps_hip_get_option("check_every");
    so.dense_ic = MPS_DIM == 2;
    so.spmv_variant = (int)mps_hip_get_option("spmv");
    so.fused = (int)mps_hip_get_option("fused");
    SolveResult sr{};
    if (solve_ppe(c, so, &sr)) return;
    StepConsts k = cl.k;
    k.clamp = 0;
    launch_pressure_out(c, k, 0);
    DevCounters h;
    if (hipMemcpy(&h, c->counters, sizeof h, hipMemcpyDeviceToHost) == hipSuccess) {
        c->stats.n_eq = h.n_eq;
        c->stats.nnz = (long long)h.nnz;
    }
    c->stats.last_iters = sr.iters;
    c->stats.last_residual = sr.res;
    c->stats.last_bnorm = sr.bnorm;
    c->ppe_valid = true;
    end_call(cl, ps);
}

// particle_system_calc_initial_params (particle.c:69-221) on the GPU
void particle_system_calc_initial_params(ParticleSystem *ps)
{
    mps_clear_error();
    if (!ps || !g_config) return;
    const SimConfig *cf = g_config;
    const double re_n = cf->influence_radius_n, re_lap = cf->influence_radius_lap;
    const double re_st = cf->influence_radius_st, l0 = cf->particle_distance;
    const double reach = fmax(re_n, re_lap);
    // bounding box of the finite positions (host: geometry set-up only)
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool any = false;
    for (int i = 0; i < ps->num; i++) {
        const double *x = ps->particles[i].pos;
        bool fin = true;
        for (int d = 0; d < MPS_DIM; d++) fin = fin && isfinite(x[d]);
        if (!fin) continue;
        for (int d = 0; d < MPS_DIM; d++) {
            if (!any || x[d] < lo[d]) lo[d] = x[d];
            if (!any || x[d] > hi[d]) hi[d] = x[d];
        }
        any = true;
    }
    Grid g{};
    const double cell = reach * 1.0001;
    g.inv = 1.0 / cell;
    g.total = 1;
    for (int d = 0; d < 3; d++) {
        if (d < MPS_DIM) {
            g.origin[d] = lo[d] - cell;
            g.nc[d] = (int)((hi[d] - lo[d]) / cell) + 3;
        } else {
            g.origin[d] = 0.0;
            g.nc[d] = 1;
        }
        g.total *= g.nc[d];
    }
    Ctx *c = mps_ctx_for(ps, 1);
    if (!c) return;
    if (c->resident && c->has_state) {
        mps_ctx_download(c, ps, nullptr);
        c->resident = false;
    }
    if (mps_ctx_upload(c, ps, nullptr, g)) {
        fprintf(stderr, "Error: %s\n", mps_hip_error_string(mps_hip_last_error()));
        return;
    }
    StepConsts k = mps_consts(cf, ps, g, reach);
    k.stride = c->stride;
    DevCounters z{};
    z.overflow_orig = INT_MAX;
    if (hipMemcpy(c->counters, &z, sizeof z, hipMemcpyHostToDevice) != hipSuccess) return;
    double out[6];
    if (extra_init_params(c, k, g, reach * reach * 1.0001, out)) {
        fprintf(stderr, "Error: %s\n", mps_hip_error_string(mps_hip_last_error()));
        return;
    }
    DevCounters h;
    if (hipMemcpy(&h, c->counters, sizeof h, hipMemcpyDeviceToHost) == hipSuccess &&
        h.overflow_orig != INT_MAX) {
        fprintf(stderr, "Error: more than max_neighbors particles within the influence radius of "
                        "particle %d\n", h.overflow_orig);
        char msg[128];
        snprintf(msg, sizeof msg, "neighbour list overflow at particle %d", h.overflow_orig);
        mps_set_error(MPS_ERR_NEIGHBOR_OVERFLOW, __FILE__, __LINE__, msg);
        return;
    }
    c->has_state = false;   // the pass used the state buffers as scratch
    if (out[3] < 0.0) {
        fprintf(stderr, "Error: no fluid particle found for initial params\n");
        ps->n0 = 1.0;
        ps->lambda = 1.0;
        return;
    }
    ps->n0 = out[0];
    if (cf->surface_detection_method == 1) ps->n0_count = (int)out[4];
    if (cf->use_analytical_lambda)
        ps->lambda = re_lap * re_lap * (double)MPS_DIM * (double)(MPS_DIM - 1) /
                     ((double)(MPS_DIM + 1) * (double)(MPS_DIM + 2));
    else
        ps->lambda = out[5];
    if (cf->surface_tension_enabled) {
        // lattice constant of the interface sum (particle.c:188-212); no
        // particle data involved
        int nmax = (int)ceil(re_st / l0);
        double sum = 0.0;
        for (int ax = 1; ax <= nmax; ax++)
            for (int bx = -(nmax - 1); bx <= 0; bx++)
                for (int by = -nmax; by <= nmax; by++)
#if MPS_DIM == 3
                    for (int bz = -nmax; bz <= nmax; bz++)
#endif
                    {
                        double rx = (double)(ax - bx) * l0, ry = (double)by * l0;
#if MPS_DIM == 3
                        double rz = (double)bz * l0;
                        double rab = sqrt(rx * rx + ry * ry + rz * rz);
#else
                        double rab = sqrt(rx * rx + ry * ry);
#endif
                        if (rab > 0.0 && rab < re_st)
                            sum += (1.0 / 3.0) * (rab - 1.5 * l0 + 0.5 * re_st) * (rab - re_st) * (rab - re_st);
                    }
        if (sum > 0.0) {
#if MPS_DIM == 3
            ps->C_LL = cf->surface_tension_coeff * l0 * l0 / sum;
#else
            ps->C_LL = cf->surface_tension_coeff * l0 / sum;
#endif
            double theta = cf->wetting_angle_SL * M_PI / 180.0;
            ps->C_SL = 0.5 * (1.0 + cos(theta)) * ps->C_LL;
        }
    }
    printf("Initial params: n0 = %.6f (re_n=%.4f)  N0 = %d  lambda = %.6f (re_lap=%.4f)%s\n", ps->n0, re_n,
           ps->n0_count, ps->lambda, re_lap, cf->use_analytical_lambda ? "  [analytical]" : "");
}

// ---------------------------------------------------------------------
// exports
// ---------------------------------------------------------------------

static int eq_numbering(Ctx *c, std::vector<int> &O, std::vector<int> &unk, std::vector<int> &eq_of_orig)
{
    const int n = c->n;
    O.resize(n);
    unk.resize(n);
    HIP_OK(hipMemcpy(O.data(), c->O[c->ct], sizeof(int) * n, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(unk.data(), c->unk, sizeof(int) * n, hipMemcpyDeviceToHost));
    std::vector<int> flag(n, 0);
    for (int s = 0; s < n; s++) flag[O[s]] = unk[s];
    eq_of_orig.assign(n, -1);
    int e = 0;
    for (int i = 0; i < n; i++)
        if (flag[i]) eq_of_orig[i] = e++;
    return e;
}

int mps_hip_export_ppe(ParticleSystem *ps, int *eq_idx, int *row_ptr, int *col, double *val,
                       double *diag, double *b)
{
    Ctx *c = ps ? mps_ctx_for(ps, 0) : nullptr;
    if (!c || !c->ppe_valid) return MPS_ERR_ARG;
    std::vector<int> O, unk, eq;
    int n_eq = eq_numbering(c, O, unk, eq);
    if (n_eq < 0) return MPS_ERR_HIP;
    if (eq_idx) memcpy(eq_idx, eq.data(), sizeof(int) * c->n);
    std::vector<int> off, scol;
    std::vector<double> sval;
    int rc = extra_export_rows(c, off, scol, sval);
    if (rc) return rc;
    std::vector<double> hb(c->n);
    HIP_OK(hipMemcpy(hb.data(), c->b, sizeof(double) * c->n, hipMemcpyDeviceToHost));
    // rows in eq order
    std::vector<int> slot_of_eq(n_eq);
    for (int s = 0; s < c->n; s++)
        if (unk[s]) slot_of_eq[eq[O[s]]] = s;
    std::vector<std::pair<int, double>> row;
    int at = 0;
    if (row_ptr) row_ptr[0] = 0;
    for (int e = 0; e < n_eq; e++) {
        const int s = slot_of_eq[e];
        row.clear();
        for (int q = off[s]; q < off[s + 1]; q++) row.push_back({eq[O[scol[q]]], sval[q]});
        std::sort(row.begin(), row.end(),
                  [](const std::pair<int, double> &a, const std::pair<int, double> &bb) { return a.first < bb.first; });
        for (auto &pr : row) {
            if (col) col[at] = pr.first;
            if (val) val[at] = pr.second;
            if (diag && pr.first == e) diag[e] = pr.second;
            at++;
        }
        if (row_ptr) row_ptr[e + 1] = at;
        if (b) b[e] = hb[s];
    }
    return MPS_OK;
}

int mps_hip_export_solution(ParticleSystem *ps, double *x)
{
    Ctx *c = ps ? mps_ctx_for(ps, 0) : nullptr;
    if (!c || !c->ppe_valid || !x) return MPS_ERR_ARG;
    std::vector<int> O, unk, eq;
    eq_numbering(c, O, unk, eq);
    std::vector<double> hx(c->n);
    HIP_OK(hipMemcpy(hx.data(), c->x, sizeof(double) * c->n, hipMemcpyDeviceToHost));
    for (int s = 0; s < c->n; s++)
        if (unk[s]) x[eq[O[s]]] = hx[s];
    return MPS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------

__global__ void k_fp64_probe(int n, const double *a, const double *b, double *s, double *q)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    s[i] = sqrt(a[i]);
    q[i] = a[i] / b[i];
}

extern "C" int mps_hip_fp64_probe(int n, const double *a, const double *b, double *sqrt_out, double *div_out)
{
    if (n <= 0) return MPS_OK;
    int rc = mps_hip_init(0);
    if (rc) return rc;
    double *d = nullptr;
    HIP_OK(hipMalloc(&d, sizeof(double) * 4 * (size_t)n));
    HIP_OK(hipMemcpy(d, a, sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d + n, b, sizeof(double) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_fp64_probe, dim3((n + 255) / 256), dim3(256), 0, 0, n, d, d + n, d + 2 * n, d + 3 * n);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(sqrt_out, d + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(div_out, d + 3 * n, sizeof(double) * n, hipMemcpyDeviceToHost));
    HIP_OK(hipFree(d));
    return MPS_OK;
}
