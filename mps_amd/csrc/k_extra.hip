// k_extra.hip -- kernels behind the per-function drop-in entry points, the
// initial-parameter pass and the PPE export.
#include <limits.h>

#include "ctx.hpp"

#define BLK 256

static inline unsigned nblk(size_t n, unsigned b = BLK) { return (unsigned)((n + b - 1) / b); }

#define LAUNCH_CHECK()                                                       \
    do {                                                                     \
        hipError_t e_ = hipGetLastError();                                   \
        if (e_ != hipSuccess) {                                              \
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(e_)); \
            return MPS_ERR_HIP;                                              \
        }                                                                    \
    } while (0)

// ---------------------------------------------------------------------
// elementwise operators
// ---------------------------------------------------------------------

// judge_free_surface (operators.c:206-219)
__global__ void k_surface(int n, const double *ND, const int *T, int *S, double thr_n0)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i < n) S[i] = (T[i] == FLUID_PARTICLE) ? (ND[i] < thr_n0) : 0;
}

// calc_neighbor_count (operators.c:226-246)
__global__ void k_count_n(int n, const dvec *X, const int *nbr, const int *ncnt, int *NC, double re_n2,
                          int stride)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const dvec xi = X[i];
    int c = 0;
    for (int q = 0; q < ncnt[i]; q++) {
        double dr[D_];
        c += dist2(xi, X[nbr[sell_off(i, q, stride)]], dr) < re_n2;
    }
    NC[i] = c;
}

// judge_free_surface_by_count (operators.c:253-267)
__global__ void k_surface_count(int n, const int *NC, const int *T, int *S, double beta_n0c)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i < n) S[i] = (T[i] == FLUID_PARTICLE) ? ((double)NC[i] < beta_n0c) : 0;
}

// clamp_pressure (operators.c:188-199)
__global__ void k_clamp(int n, double *P, const int *T)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i < n && T[i] == FLUID_PARTICLE && P[i] < 0.0) P[i] = 0.0;
}

// apply_wall_boundary (boundary.c:12-26)
__global__ void k_wall(int n, const int *T, dvec *V, dvec *A)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    if (T[i] == WALL_PARTICLE || T[i] == GHOST_PARTICLE) {
        dvec v = V[i], a = A[i];
        for (int d = 0; d < D_; d++) v.c[d] = a.c[d] = 0.0;
        V[i] = v;
        A[i] = a;
    }
}

// remove_out_of_bounds (boundary.c:33-61)
__global__ void k_oob(int n, const dvec *X, int *T, dvec *V, dvec *A, double *P, double *ND, StepConsts k)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n || T[i] != FLUID_PARTICLE) return;
    const dvec x = X[i];
    bool out = false;
    for (int d = 0; d < D_; d++)
        if (x.c[d] < k.bnd_lo[d] || x.c[d] > k.bnd_hi[d] || isnan(x.c[d])) out = true;
    if (!out) return;
    T[i] = GHOST_PARTICLE;
    P[i] = 0.0;
    ND[i] = 0.0;
    dvec v = V[i], a = A[i];
    for (int d = 0; d < D_; d++) v.c[d] = a.c[d] = 0.0;
    V[i] = v;
    A[i] = a;
}

// calc_viscosity_term (operators.c:50-85): acc += nu*coeff*lap
__global__ void k_viscosity(int n, const dvec *X, const dvec *V, const int *T, const int *nbr,
                            const int *ncnt, dvec *A, StepConsts k)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n || T[i] != FLUID_PARTICLE) return;
    const dvec xi = X[i], vi = V[i];
    double lap[D_];
    for (int d = 0; d < D_; d++) lap[d] = 0.0;
    for (int q = 0; q < ncnt[i]; q++) {
        const int j = nbr[sell_off(i, q, k.stride)];
        double dr[D_];
        const double w = wfun(sqrt(dist2(xi, X[j], dr)), k.re_lap, k.r_min);
        const dvec vj = V[j];
        for (int d = 0; d < D_; d++) lap[d] += (vj.c[d] - vi.c[d]) * w;
    }
    dvec a = A[i];
    for (int d = 0; d < D_; d++) a.c[d] += k.visc_coeff * lap[d];
    A[i] = a;
}

// calc_surface_tension / _SL (operators.c:281-354)
__global__ void k_surface_tension(int n, const dvec *X, const int *T, const int *nbr, const int *ncnt,
                                  dvec *A, StepConsts k, int partner, double cst)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n || T[i] != FLUID_PARTICLE) return;
    const dvec xi = X[i];
    dvec a = A[i];
    for (int q = 0; q < ncnt[i]; q++) {
        const int j = nbr[sell_off(i, q, k.stride)];
        if (T[j] != partner) continue;
        double dr[D_];
        const double r = sqrt(dist2(xi, X[j], dr));
        if (r < k.st_tiny || r >= k.re_st) continue;
        const double f = -(r - k.l0) * (r - k.re_st);
        for (int d = 0; d < D_; d++) a.c[d] += cst * f * (dr[d] / r);
    }
    A[i] = a;
}

// calc_pressure_gradient without the corrector (operators.c:102-181)
__global__ void k_gradient(int n, const dvec *X, const double *P, const double *pmin, const int *T,
                           const int *nbr, const int *ncnt, dvec *A, StepConsts k)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n || T[i] != FLUID_PARTICLE) return;
    const dvec xi = X[i];
    const double pi = P[i];
    const double pimin = (k.cmps == 2) ? 0.0 : pmin[i];
    double grad[D_];
    for (int d = 0; d < D_; d++) grad[d] = 0.0;
    for (int q = 0; q < ncnt[i]; q++) {
        const int j = nbr[sell_off(i, q, k.stride)];
        double dr[D_];
        const double r2 = dist2(xi, X[j], dr);
        if (r2 < k.grad_min_r2) continue;
        const double w = wfun(sqrt(r2), k.re_lap, k.r_min);
        const double pj = P[j];
        double dp;
        if (k.cmps == 1) dp = pi + pj - pimin - pmin[j];
        else if (k.cmps == 2) dp = pi + pj;
        else dp = pj - pimin;
        for (int d = 0; d < D_; d++) grad[d] += dp / r2 * dr[d] * w;
    }
    dvec a = A[i];
    for (int d = 0; d < D_; d++) a.c[d] = k.neg_gc * grad[d] / k.rho;
    A[i] = a;
}

__global__ void k_pmin_all(int n, const double *P, const int *nbr, const int *ncnt, double *pmin, int stride)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    double m = P[i];
    for (int q = 0; q < ncnt[i]; q++) {
        const double pj = P[nbr[sell_off(i, q, stride)]];
        if (pj < m) m = pj;
    }
    pmin[i] = m;
}

// Unknown flags, diagonal and standard source from a given list, number
// density and flags (pressure_solver.c:411-529) -- the first half of the
// solve_pressure drop-in; k_assemble then fills the off-diagonals.
__global__ void k_ppe_rows(int n, const dvec *X, const int *T, const double *ND, const int *S,
                           const int *nbr, const int *ncnt, int *unk, double *diag, double *b,
                           DevCounters *ctr, StepConsts k)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    int u = 0;
    if (i < n) {
        const int t = T[i];
        u = (t == FLUID_PARTICLE && !S[i]) || t == WALL_PARTICLE;
        unk[i] = u;
        double sumw = 0.0;
        if (u) {
            const dvec xi = X[i];
            for (int q = 0; q < ncnt[i]; q++) {
                const int j = nbr[sell_off(i, q, k.stride)];
                if (T[j] == DUMMY_PARTICLE) continue;
                double dr[D_];
                sumw += wfun(sqrt(dist2(xi, X[j], dr)), k.re_lap, k.r_min);
            }
        }
        diag[i] = u ? k.diag_coeff * sumw : 0.0;
        double bi = 0.0;
        if (u && t == FLUID_PARTICLE && k.hs_mode == 0 && k.ppe_type == 0)
            bi = k.src_k * (ND[i] - k.n0) / k.n0;
        b[i] = bi;
    }
    const long long nu = block_sum<long long, BLK>((long long)u);
    if (threadIdx.x == 0) atomicAdd(&ctr->n_eq, (int)nu);
}

// neighbor_search_brute_force (neighbor_search.c:44-81) in identity order
__global__ void k_brute(int n, const dvec *X, const int *T, int *nbr, int *ncnt, double re2, int stride,
                        DevCounters *ctr, const int *O)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    int cnt = 0;
    if (T[i] != GHOST_PARTICLE) {
        const dvec xi = X[i];
        for (int j = 0; j < n; j++) {
            if (j == i || T[j] == GHOST_PARTICLE) continue;
            double dr[D_];
            if (dist2(xi, X[j], dr) < re2) {
                if (cnt < stride) nbr[sell_off(i, cnt, stride)] = j;
                cnt++;
            }
        }
    }
    ncnt[i] = cnt < stride ? cnt : stride;
    if (cnt > stride) atomicMin(&ctr->overflow_orig, O[i]);
}

int extra_surface(Ctx *c, const StepConsts &k)
{
    hipLaunchKernelGGL(k_surface, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->ND[c->ct],
                       c->T[c->ct], c->SURF[c->ct], k.thr_n0);
    LAUNCH_CHECK();
    return 0;
}

int extra_count(Ctx *c, const StepConsts &k)
{
    hipLaunchKernelGGL(k_count_n, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->X[c->cx], c->nbr,
                       c->ncnt, c->NCOUNT[c->ct], k.re_n2, c->stride);
    LAUNCH_CHECK();
    return 0;
}

int extra_surface_count(Ctx *c, const StepConsts &k)
{
    hipLaunchKernelGGL(k_surface_count, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n,
                       c->NCOUNT[c->ct], c->T[c->ct], c->SURF[c->ct], k.beta_n0c);
    LAUNCH_CHECK();
    return 0;
}

int extra_clamp(Ctx *c)
{
    hipLaunchKernelGGL(k_clamp, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->P[c->cp], c->T[c->ct]);
    LAUNCH_CHECK();
    return 0;
}

int extra_wall(Ctx *c)
{
    hipLaunchKernelGGL(k_wall, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->T[c->ct], c->V[c->cx],
                       c->ACC[c->ct]);
    LAUNCH_CHECK();
    return 0;
}

int extra_oob(Ctx *c, const StepConsts &k)
{
    hipLaunchKernelGGL(k_oob, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->X[c->cx], c->T[c->ct],
                       c->V[c->cx], c->ACC[c->ct], c->P[c->cp], c->ND[c->ct], k);
    LAUNCH_CHECK();
    return 0;
}

int extra_viscosity(Ctx *c, const StepConsts &k)
{
    hipLaunchKernelGGL(k_viscosity, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->X[c->cx],
                       c->V[c->cx], c->T[c->ct], c->nbr, c->ncnt, c->ACC[c->ct], k);
    LAUNCH_CHECK();
    return 0;
}

int extra_surface_tension(Ctx *c, const StepConsts &k, int solid)
{
    hipLaunchKernelGGL(k_surface_tension, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->X[c->cx],
                       c->T[c->ct], c->nbr, c->ncnt, c->ACC[c->ct], k,
                       solid ? WALL_PARTICLE : FLUID_PARTICLE, solid ? k.st_sl : k.st_ll);
    LAUNCH_CHECK();
    return 0;
}

int extra_gradient(Ctx *c, const StepConsts &k)
{
    if (k.cmps != 2) {
        hipLaunchKernelGGL(k_pmin_all, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->P[c->cp],
                           c->nbr, c->ncnt, c->pmin, c->stride);
        LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_gradient, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->X[c->cx],
                       c->P[c->cp], c->pmin, c->T[c->ct], c->nbr, c->ncnt, c->ACC[c->ct], k);
    LAUNCH_CHECK();
    return 0;
}

int extra_ppe_rows(Ctx *c, const StepConsts &k)
{
    hipLaunchKernelGGL(k_ppe_rows, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->X[c->cx],
                       c->T[c->ct], c->ND[c->ct], c->SURF[c->ct], c->nbr, c->ncnt, c->unk, c->diag, c->b,
                       c->counters, k);
    LAUNCH_CHECK();
    return 0;
}

int extra_brute(Ctx *c, double re2)
{
    hipLaunchKernelGGL(k_brute, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->X[c->cx], c->T[c->ct],
                       c->nbr, c->ncnt, re2, c->stride, c->counters, c->O[c->ct]);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------
// initial parameters (particle.c:69-221)
//
// For every fluid particle the reference sums over all j != i in ascending
// index the positive weights w(r, re_n) (n0_i) and, for w(r, re_lap) > 0,
// r^2 w and w (lambda numerator/denominator).  Here the candidates come from
// a grid of cell edge 1.0001*max(re_n, re_lap) built over every particle with
// a finite position (ghosts included, as in the reference), the hits of each
// particle are written to its neighbour-list row, sorted by index, and
// summed in that order: the same terms in the same order.  Then n0 is the
// largest n0_i over fluid particles and lambda comes from the first particle
// (lowest index) that attains it -- the serial reference's rule.
// ---------------------------------------------------------------------

struct InitGrid {
    Grid g;
    double reach2;
};

__global__ void k_init_bin(int n, const dvec *X, int *cell, int *ccount, Grid g)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    int cc[3];
    int ci = cell_coords(g, X[i], cc);
    if (ci < 0) ci = g.total;      // non-finite positions
    cell[i] = ci;
    atomicAdd(&ccount[ci], 1);
}

__global__ void k_init_sums(int n, const dvec *X, const int *T, const int *cstart, const int *bucket,
                            int *hits, int stride, InitGrid ig, StepConsts k, double *n0i,
                            double *numi, double *deni, int *cnti, DevCounters *ctr)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    n0i[i] = numi[i] = deni[i] = 0.0;
    cnti[i] = 0;
    if (T[i] != FLUID_PARTICLE) return;
    const dvec xi = X[i];
    int cc[3];
    if (cell_coords(ig.g, xi, cc) < 0) return;
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = cc[d] > 0 ? cc[d] - 1 : 0;
        hi[d] = cc[d] < ig.g.nc[d] - 1 ? cc[d] + 1 : ig.g.nc[d] - 1;
    }
    int m = 0;
    for (int cz = lo[2]; cz <= hi[2]; cz++)
        for (int cy = lo[1]; cy <= hi[1]; cy++)
            for (int cxx = lo[0]; cxx <= hi[0]; cxx++) {
                const int cid = cell_key(ig.g, cxx, cy, cz);
                for (int u = cstart[cid]; u < cstart[cid + 1]; u++) {
                    const int j = bucket[u];
                    if (j == i) continue;
                    double dr[D_];
                    if (!(dist2(xi, X[j], dr) < ig.reach2)) continue;
                    if (m < stride) {
                        // insertion into the sorted row
                        int q = m;
                        while (q > 0 && hits[sell_off(i, q - 1, stride)] > j) {
                            hits[sell_off(i, q, stride)] = hits[sell_off(i, q - 1, stride)];
                            q--;
                        }
                        hits[sell_off(i, q, stride)] = j;
                    }
                    m++;
                }
            }
    if (m > stride) {
        atomicMin(&ctr->overflow_orig, i);
        return;
    }
    double sn = 0.0, snum = 0.0, sden = 0.0;
    int cnt = 0;
    for (int q = 0; q < m; q++) {
        double dr[D_];
        const double r2 = dist2(xi, X[hits[sell_off(i, q, stride)]], dr);
        const double r = sqrt(r2);
        const double wn = wfun(r, k.re_n, k.r_min);
        if (wn > 0.0) sn += wn;
        const double wl = wfun(r, k.re_lap, k.r_min);
        if (wl > 0.0) {
            snum += r2 * wl;
            sden += wl;
        }
        cnt += (r2 < k.re_n2);
    }
    n0i[i] = sn;
    numi[i] = snum;
    deni[i] = sden;
    cnti[i] = cnt;
}

// first strict maximum over fluid particles: largest value, lowest index
struct ArgMax {
    double v;
    int i;
    int count_max;
};

__global__ void __launch_bounds__(BLK) k_init_argmax(int n, const int *T, const double *n0i,
                                                     const int *cnti, ArgMax *part)
{
    __shared__ ArgMax sh[BLK];
    const int i = blockIdx.x * BLK + threadIdx.x;
    ArgMax a = {0.0, INT_MAX, 0};
    if (i < n && T[i] == FLUID_PARTICLE) {
        if (n0i[i] > 0.0) a = {n0i[i], i, cnti[i]};
        else a.count_max = cnti[i];
    }
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int s = BLK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            ArgMax x = sh[threadIdx.x], y = sh[threadIdx.x + s];
            ArgMax z = x;
            if (y.v > x.v || (y.v == x.v && y.i < x.i)) z = y;
            z.count_max = x.count_max > y.count_max ? x.count_max : y.count_max;
            sh[threadIdx.x] = z;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(BLK) k_init_final(int nparts, const ArgMax *part, const double *numi,
                                                    const double *deni, double *out)
{
    __shared__ ArgMax sh[BLK];
    ArgMax a = {0.0, INT_MAX, 0};
    for (int q = threadIdx.x; q < nparts; q += BLK) {
        ArgMax y = part[q];
        if (y.v > a.v || (y.v == a.v && y.i < a.i)) {
            int cm = a.count_max;
            a = y;
            a.count_max = cm > y.count_max ? cm : y.count_max;
        } else if (y.count_max > a.count_max) {
            a.count_max = y.count_max;
        }
    }
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int s = BLK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            ArgMax x = sh[threadIdx.x], y = sh[threadIdx.x + s];
            ArgMax z = x;
            if (y.v > x.v || (y.v == x.v && y.i < x.i)) z = y;
            z.count_max = x.count_max > y.count_max ? x.count_max : y.count_max;
            sh[threadIdx.x] = z;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const ArgMax w = sh[0];
        out[0] = w.v;                                   // n0 (0 if none)
        out[1] = w.i == INT_MAX ? 0.0 : numi[w.i];      // lambda numerator
        out[2] = w.i == INT_MAX ? 0.0 : deni[w.i];      // lambda denominator
        out[3] = w.i == INT_MAX ? -1.0 : (double)w.i;   // argmax
        out[4] = (double)w.count_max;                   // N0
        out[5] = (w.i == INT_MAX) ? 0.0 : (out[2] > 1.0e-10 ? out[1] / out[2] : 1.0);
    }
}

// Runs the pass on the identity-ordered state in c.  out[6] as above.
int extra_init_params(Ctx *c, const StepConsts &k, const Grid &g, double reach2, double *out_host)
{
    const int n = c->n;
    int rc = launch_fill_int(c->ccount, 0, (size_t)g.total + 1, c->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_init_bin, dim3(nblk(n)), dim3(BLK), 0, c->stream, n, c->X[c->cx], c->cell,
                       c->ccount, g);
    LAUNCH_CHECK();
    rc = launch_scan_excl(c->ccount, c->cstart, g.total + 1, c->scan_blk, c->stream);
    if (!rc) rc = launch_fill_int(c->cfill, 0, (size_t)g.total + 1, c->stream);
    if (rc) return rc;
    // bucket members (order inside a cell does not matter here)
    extern int bucket_slots(Ctx * c);
    rc = bucket_slots(c);
    if (rc) return rc;
    double *n0i = c->x, *numi = c->r, *deni = c->pv;
    int *cnti = c->Alen;
    InitGrid ig;
    ig.g = g;
    ig.reach2 = reach2;
    hipLaunchKernelGGL(k_init_sums, dim3(nblk(n)), dim3(BLK), 0, c->stream, n, c->X[c->cx], c->T[c->ct],
                       c->cstart, c->tmp_key, c->nbr, c->stride, ig, k, n0i, numi, deni, cnti,
                       c->counters);
    LAUNCH_CHECK();
    const unsigned nparts = nblk(n > 0 ? n : 1);
    ArgMax *part = (ArgMax *)c->partA;   // 65536 doubles = room for 21845 ArgMax records
    if (nparts * sizeof(ArgMax) > 65536 * sizeof(double)) {
        mps_set_error(MPS_ERR_UNSUPPORTED, __FILE__, __LINE__, "too many particles for init reduction");
        return MPS_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(k_init_argmax, dim3(nparts), dim3(BLK), 0, c->stream, n, c->T[c->ct], n0i, cnti, part);
    hipLaunchKernelGGL(k_init_final, dim3(1), dim3(BLK), 0, c->stream, (int)nparts, part, numi, deni, c->partB);
    LAUNCH_CHECK();
    HIP_OK(hipMemcpyAsync(out_host, c->partB, sizeof(double) * 6, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return 0;
}

__global__ void k_bucket_all(int n, const int *cell, const int *cstart, int *cfill, int *bucket)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const int ci = cell[i];
    bucket[cstart[ci] + atomicAdd(&cfill[ci], 1)] = i;
}

int bucket_slots(Ctx *c)
{
    hipLaunchKernelGGL(k_bucket_all, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->cell, c->cstart,
                       c->cfill, c->tmp_key);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------
// PPE export: slot-ordered CSR with the diagonal as an entry
// ---------------------------------------------------------------------

__global__ void k_row_len(int n, const int *unk, const int *Alen, int *len)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i < n) len[i] = unk[i] ? Alen[i] + 1 : 0;
}

__global__ void k_compact_rows(int n, int stride, const int *unk, const int *Alen, const double *Aval,
                               const int *Acol, const double *diag, const int *off, int *col, double *val)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n || !unk[i]) return;
    int at = off[i];
    const int m = Alen[i];
    for (int q = 0; q < m; q++) {
        col[at] = Acol[sell_off(i, q, stride)];
        val[at] = Aval[sell_off(i, q, stride)];
        at++;
    }
    col[at] = i;
    val[at] = diag[i];
}

// Fills host arrays: per-slot row offsets (n+1), cols (slots), vals.
int extra_export_rows(Ctx *c, std::vector<int> &off, std::vector<int> &col, std::vector<double> &val)
{
    const int n = c->n;
    int *len = c->cnt_orig, *doff = c->off_orig;
    hipLaunchKernelGGL(k_row_len, dim3(nblk(n)), dim3(BLK), 0, c->stream, n, c->unk, c->Alen, len);
    LAUNCH_CHECK();
    int rc = launch_scan_excl(len, doff, n, c->scan_blk, c->stream);
    if (rc) return rc;
    off.resize(n + 1);
    HIP_OK(hipMemcpyAsync(off.data(), doff, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    const size_t nnz = (size_t)off[n];
    int *dcol = nullptr;
    double *dval = nullptr;
    HIP_OK(hipMalloc(&dcol, sizeof(int) * (nnz ? nnz : 1)));
    HIP_OK(hipMalloc(&dval, sizeof(double) * (nnz ? nnz : 1)));
    hipLaunchKernelGGL(k_compact_rows, dim3(nblk(n)), dim3(BLK), 0, c->stream, n, c->stride, c->unk,
                       c->Alen, c->Aval, c->Acol, c->diag, doff, dcol, dval);
    hipError_t e = hipGetLastError();
    col.resize(nnz);
    val.resize(nnz);
    if (e == hipSuccess) e = hipMemcpyAsync(col.data(), dcol, sizeof(int) * nnz, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(val.data(), dval, sizeof(double) * nnz, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(dcol);
    hipFree(dval);
    if (e != hipSuccess) {
        mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(e));
        return MPS_ERR_HIP;
    }
    return 0;
}
