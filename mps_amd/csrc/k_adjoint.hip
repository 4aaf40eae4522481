// k_adjoint.hip -- reverse-mode gradient of the MPS timestep on gfx950.
//
// Loss: L = (1/N_f) sum over fluid particles of |r_i(T) - r_i^target|^2
// after T steps.  Gradients: dL/dnu, dL/dg and dL/du_i(0).
//
// The forward is the production step (runtime.hip).  While a segment is
// recomputed, hooks copy its intermediate states into a tape, in particle
// (original index) order:
//   r, u      at the start of the step
//   r*, u*    after the viscosity/gravity predictor (simulation.c:31-67)
//   r**, u**  after the collision (operators.c:371-446; 2-D all pairs)
//   L         the neighbour list searched on r** (CSR over original indices)
//   x, unk    the PPE solution and unknown flags
// and the reverse sweep runs the adjoint of each operator with the
// forward's discrete decisions frozen: the lists, the surface flags and
// unknown set, the negative-pressure clamp, the neighbour that gives p_min,
// the set and order of the collisions that fired, and ghosted particles.
// Only fluid particles move, so only their positions and velocities carry
// adjoints; the pressure adjoint p-bar covers every particle.
//
// The reverse operators are written in pull form wherever the lists allow
// (j in L(i) <=> i in L(j)): each particle gathers every term that lands on
// it, so the sums are deterministic.  The collision's partner terms are the
// exception (scattered with atomics; few particles collide).
//
// Per step, in reverse:
//   corrector   a_i = G sum_j (p_j - p_mu(i)) f(d_ij), f(d) = d w(|d|)/|d|^2,
//               G = -(D/n0)/rho; u' = u** + dt a; r' = r** + dt^2 a
//               -> p-bar, r**-bar
//   clamp       p-bar = 0 where the fluid pressure was clamped
//   PPE         lambda = A^-1 x-bar (A symmetric; CG on the same matrix),
//               b-bar = lambda -> n-bar = (relax rho/dt^2)/n0 lambda (fluid
//               rows), A-bar = -lambda x^T through A_ij = -kappa w(r_ij)
//               and A_ii = kappa sum_(non-dummy j) w(r_ij) -> r**-bar
//   density     n_i = sum_j w_n(r_ij) -> r**-bar
//   collision   the firing sequence of each fluid particle, reversed
//   predictor   a_i = g + nu*coeff sum_j (u_j - u_i) w(r_ij) over the
//               previous step's list, u* = u + dt a, r* = r + dt u*
//               -> r-bar, u-bar, nu-bar, g-bar
#include <limits.h>
#include <string.h>

#include <vector>

#include "ctx.hpp"

#define AB 256
#define COLL_MAX 16      // collisions of one particle in one step that the reverse sweep can hold

static inline unsigned ab(size_t n) { return (unsigned)((n + AB - 1) / AB); }

#define LAUNCH_CHECK()                                                       \
    do {                                                                     \
        hipError_t e_ = hipGetLastError();                                   \
        if (e_ != hipSuccess) {                                              \
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(e_)); \
            return MPS_ERR_HIP;                                              \
        }                                                                    \
    } while (0)

// w(r) of kernel.c:12-20 and its derivative (0 where r is clamped to r_min
// or r >= re)
__device__ __forceinline__ double wdiff(double r, double re, double r_min)
{
    if (r >= re || r < r_min) return 0.0;
    return -re / (r * r);
}

// w at the replayed distance r with its two branches (cut-off at re, clamp
// at r_min) taken at the recorded distance tr: the function whose
// derivative wdiff(tr) is, smooth in r along a replay
__device__ __forceinline__ double wfrozen(double r, double tr, double re, double r_min)
{
    if (tr >= re) return 0.0;
    if (tr < r_min) return re / r_min - 1.0;
    return re / r - 1.0;
}

__device__ __forceinline__ double dist(const dvec &a, const dvec &b)
{
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < D_; k++) s += (a.c[k] - b.c[k]) * (a.c[k] - b.c[k]);
    return sqrt(s);
}

__device__ __forceinline__ dvec dsub(const dvec &a, const dvec &b)
{
    dvec d;
#pragma unroll
    for (int k = 0; k < (D_ == 3 ? 4 : 2); k++) d.c[k] = 0.0;
#pragma unroll
    for (int k = 0; k < D_; k++) d.c[k] = a.c[k] - b.c[k];
    return d;
}

__device__ __forceinline__ double ddot(const dvec &a, const dvec &b)
{
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < D_; k++) s += a.c[k] * b.c[k];
    return s;
}

__device__ __forceinline__ void daxpy(dvec &y, double a, const dvec &x)
{
#pragma unroll
    for (int k = 0; k < D_; k++) y.c[k] += a * x.c[k];
}

static inline dvec dzero_host()
{
    dvec d;
    for (int k = 0; k < (D_ == 3 ? 4 : 2); k++) d.c[k] = 0.0;
    return d;
}

__device__ __forceinline__ dvec dzero()
{
    dvec d;
#pragma unroll
    for (int k = 0; k < (D_ == 3 ? 4 : 2); k++) d.c[k] = 0.0;
    return d;
}

// ---------------------------------------------------------------------
// tape
// ---------------------------------------------------------------------

struct TapeStep {
    dvec *r = nullptr, *u = nullptr, *rs = nullptr, *us = nullptr, *rss = nullptr, *uss = nullptr;
    double *x = nullptr;
    int *unk = nullptr, *typ = nullptr;
    int *lst = nullptr, *lidx = nullptr;     // CSR of L over original indices
    size_t lcap = 0;
};

static TapeStep *g_rec = nullptr;            // the step being recorded, or null
static int g_rec_err = 0;

__global__ void k_snap_state(int n, const int *__restrict__ O, const dvec *__restrict__ X, const dvec *__restrict__ V,
                             const int *__restrict__ T, dvec *__restrict__ r, dvec *__restrict__ u, int *__restrict__ typ)
{
    const int s = blockIdx.x * AB + threadIdx.x;
    if (s >= n) return;
    const int o = O[s];
    r[o] = X[s];
    u[o] = V[s];
    if (typ) typ[o] = T[s];
}

__global__ void k_snap_x(int n, const int *__restrict__ O, const double *__restrict__ x, const int *__restrict__ unk,
                         double *__restrict__ xo, int *__restrict__ unko)
{
    const int s = blockIdx.x * AB + threadIdx.x;
    if (s >= n) return;
    const int o = O[s];
    xo[o] = unk[s] ? x[s] : 0.0;
    unko[o] = unk[s];
}

static int rec_list(Ctx *c, TapeStep *t)
{
    long long total = 0;
    int rc = launch_list_export(c, &total);
    if (rc) return rc;
    if ((size_t)total > t->lcap) {
        if (t->lidx) hipFree(t->lidx);
        t->lidx = nullptr;
        const size_t cap = (size_t)total + (size_t)total / 4 + 64;
        HIP_OK(hipMalloc(&t->lidx, sizeof(int) * cap));
        t->lcap = cap;
    }
    HIP_OK(hipMemcpyAsync(t->lst, c->off_orig, sizeof(int) * ((size_t)c->n + 1), hipMemcpyDeviceToDevice, c->stream));
    if (total > 0)
        HIP_OK(hipMemcpyAsync(t->lidx, c->nl_stage, sizeof(int) * (size_t)total, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

bool tape_recording() { return g_rec != nullptr; }

// Called by the step at its stages (runtime.hip) while a tape is recorded.
void tape_hook(Ctx *c, int stage)
{
    TapeStep *t = g_rec;
    if (!t || g_rec_err) return;
    const int n = c->n;
    const unsigned g = ab(n);
    switch (stage) {
    case TAPE_START:
        hipLaunchKernelGGL(k_snap_state, dim3(g), dim3(AB), 0, c->stream, n, c->O[c->ct], c->X[c->cx], c->V[c->cx],
                           c->T[c->ct], t->r, t->u, t->typ);
        break;
    case TAPE_PREDICTED:
        hipLaunchKernelGGL(k_snap_state, dim3(g), dim3(AB), 0, c->stream, n, c->O[c->ct], c->X[c->cx], c->V[c->cx],
                           c->T[c->ct], t->rs, t->us, nullptr);
        break;
    case TAPE_COLLIDED:
        hipLaunchKernelGGL(k_snap_state, dim3(g), dim3(AB), 0, c->stream, n, c->O[c->ct], c->X[c->cx], c->V[c->cx],
                           c->T[c->ct], t->rss, t->uss, nullptr);
        break;
    case TAPE_SEARCHED:
        if (rec_list(c, t)) g_rec_err = MPS_ERR_HIP;
        break;
    case TAPE_SOLVED:
        hipLaunchKernelGGL(k_snap_x, dim3(g), dim3(AB), 0, c->stream, n, c->O[c->ct], c->x, c->unk, t->x, t->unk);
        break;
    default:
        break;
    }
    if (hipGetLastError() != hipSuccess) g_rec_err = MPS_ERR_HIP;
}

static int tape_alloc(TapeStep *t, int n)
{
    HIP_OK(hipMalloc(&t->r, sizeof(dvec) * n));
    HIP_OK(hipMalloc(&t->u, sizeof(dvec) * n));
    HIP_OK(hipMalloc(&t->rs, sizeof(dvec) * n));
    HIP_OK(hipMalloc(&t->us, sizeof(dvec) * n));
    HIP_OK(hipMalloc(&t->rss, sizeof(dvec) * n));
    HIP_OK(hipMalloc(&t->uss, sizeof(dvec) * n));
    HIP_OK(hipMalloc(&t->x, sizeof(double) * n));
    HIP_OK(hipMalloc(&t->unk, sizeof(int) * n));
    HIP_OK(hipMalloc(&t->typ, sizeof(int) * n));
    HIP_OK(hipMalloc(&t->lst, sizeof(int) * ((size_t)n + 1)));
    return 0;
}

static void tape_free(TapeStep *t)
{
    void *b[] = {t->r, t->u, t->rs, t->us, t->rss, t->uss, t->x, t->unk, t->typ, t->lst, t->lidx};
    for (void *p : b)
        if (p) hipFree(p);
    *t = TapeStep();
}

// ---------------------------------------------------------------------
// reverse kernels (particle order; one thread per particle)
// ---------------------------------------------------------------------

// The step's pressures (operators.c:188-199 clamp on fluid), p-bar's
// neighbour-minimum index mu(i), a-bar and F_i = sum_j f(d_ij).
__global__ void k_adj_corr_pre(int n, const int *__restrict__ typ, const double *__restrict__ x,
                               const int *__restrict__ unk, const dvec *__restrict__ rss,
                               const int *__restrict__ lst, const int *__restrict__ lidx,
                               const dvec *__restrict__ rb, const dvec *__restrict__ ub, StepConsts k, int clamp,
                               dvec *__restrict__ abar, dvec *__restrict__ F, int *__restrict__ mu)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    if (typ[i] != FLUID_PARTICLE) {
        abar[i] = dzero();
        F[i] = dzero();
        mu[i] = i;
        return;
    }
    auto pres = [&](int j) {
        double p = unk[j] ? x[j] : 0.0;
        if (clamp && typ[j] == FLUID_PARTICLE && p < 0.0) p = 0.0;
        return p;
    };
    double pm = pres(i);
    int m = i;
    dvec f = dzero();
    for (int q = lst[i]; q < lst[i + 1]; q++) {
        const int j = lidx[q];
        const double pj = pres(j);
        if (pj < pm) {
            pm = pj;
            m = j;
        }
        const dvec d = dsub(rss[j], rss[i]);
        const double r2 = ddot(d, d);
        if (r2 < k.grad_min_r2) continue;
        const double w = wfun(sqrt(r2), k.re_lap, k.r_min);
        daxpy(f, w / r2, d);
    }
    mu[i] = m;
    F[i] = f;
    dvec a = dzero();
    daxpy(a, k.dt, ub[i]);
    daxpy(a, k.dt * k.dt, rb[i]);
    abar[i] = a;
}

// Pull form of the corrector's reverse: p-bar of every particle m and the
// r**-bar terms of fluid m; x-bar = p-bar through the clamp (0 for
// non-unknowns).  rbss, ubss start from r-bar', u-bar'.
__global__ void k_adj_corr(int n, const int *__restrict__ typ, const double *__restrict__ x, const int *__restrict__ unk,
                           const dvec *__restrict__ rss, const int *__restrict__ lst, const int *__restrict__ lidx,
                           const dvec *__restrict__ rb, const dvec *__restrict__ ub, const dvec *__restrict__ abar,
                           const dvec *__restrict__ F, const int *__restrict__ mu, StepConsts k, int clamp,
                           double *__restrict__ xbar, dvec *__restrict__ rbss, dvec *__restrict__ ubss)
{
    const int m = blockIdx.x * AB + threadIdx.x;
    if (m >= n) return;
    const double G = k.neg_gc / k.rho;
    auto pres = [&](int j) {
        double p = unk[j] ? x[j] : 0.0;
        if (clamp && typ[j] == FLUID_PARTICLE && p < 0.0) p = 0.0;
        return p;
    };
    const bool mf = typ[m] == FLUID_PARTICLE;
    const double pm = pres(m);
    double pb = 0.0;
    dvec rbm = mf ? rb[m] : dzero();
    // m's own p_min role
    if (mf && mu[m] == m) pb -= G * ddot(abar[m], F[m]);
    for (int q = lst[m]; q < lst[m + 1]; q++) {
        const int i = lidx[q];
        const dvec dim = dsub(rss[m], rss[i]);      // d_im = r_m - r_i
        const double r2 = ddot(dim, dim);
        const bool pair_ok = !(r2 < k.grad_min_r2);
        const double rr = sqrt(r2);
        const double w = wfun(rr, k.re_lap, k.r_min);
        const double wp = wdiff(rr, k.re_lap, k.r_min);
        if (typ[i] == FLUID_PARTICLE) {
            const dvec ai = abar[i];
            if (pair_ok) {
                // p_m enters a_i through the pair (i, m)
                pb += G * (w / r2) * ddot(ai, dim);
                // r_m enters a_i through d_im
                if (mf) {
                    const double qim = pm - pres(mu[i]);
                    const double gg = w / r2, gp = wp / r2 - 2.0 * w / (r2 * rr);
                    dvec h = dzero();
                    daxpy(h, gg, ai);
                    daxpy(h, (gp / rr) * ddot(dim, ai), dim);
                    daxpy(rbm, G * qim, h);
                }
            }
            if (mu[i] == m) pb -= G * ddot(ai, F[i]);
        }
        if (mf && pair_ok) {
            // r_m enters a_m through d_mi = r_i - r_m
            const dvec dmi = dsub(rss[i], rss[m]);
            const double qmi = pres(i) - pres(mu[m]);
            const double gg = w / r2, gp = wp / r2 - 2.0 * w / (r2 * rr);
            const dvec am = abar[m];
            dvec h = dzero();
            daxpy(h, gg, am);
            daxpy(h, (gp / rr) * ddot(dmi, am), dmi);
            daxpy(rbm, -G * qmi, h);
        }
    }
    double xb = 0.0;
    if (unk[m]) {
        const bool clamped = clamp && mf && x[m] < 0.0;
        xb = clamped ? 0.0 : pb;
    }
    xbar[m] = xb;
    rbss[m] = rbm;
    ubss[m] = mf ? ub[m] : dzero();
}

// eq-order CSR of the step's PPE (pressure_solver.c:437-551): row e is the
// e-th unknown in particle order; the diagonal first, then the unknown
// neighbours in list order.
__global__ void k_adj_rowlen(int n, const int *__restrict__ unk, const int *__restrict__ lst,
                             const int *__restrict__ lidx, int *__restrict__ len)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    int c = 0;
    if (unk[i]) {
        c = 1;
        for (int q = lst[i]; q < lst[i + 1]; q++) c += unk[lidx[q]] ? 1 : 0;
    }
    len[i] = c;
}

__global__ void k_adj_fill(int n, const int *__restrict__ typ, const int *__restrict__ unk,
                           const int *__restrict__ eq, const int *__restrict__ roff, const dvec *__restrict__ rss,
                           const int *__restrict__ lst, const int *__restrict__ lidx, const double *__restrict__ xbar,
                           StepConsts k, int *__restrict__ col, double *__restrict__ val, double *__restrict__ rhs)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n || !unk[i]) return;
    const int e = eq[i];
    int at = roff[i];
    const int dpos = at++;
    double sumw = 0.0;
    for (int q = lst[i]; q < lst[i + 1]; q++) {
        const int j = lidx[q];
        if (typ[j] == DUMMY_PARTICLE) continue;
        const dvec d = dsub(rss[j], rss[i]);
        const double w = wfun(sqrt(ddot(d, d)), k.re_lap, k.r_min);
        sumw += w;
        if (unk[j]) {
            col[at] = eq[j];
            val[at] = k.neg_coeff * w;
            at++;
        }
    }
    col[dpos] = e;
    val[dpos] = k.diag_coeff * sumw;
    rhs[e] = xbar[i];
}

__global__ void k_adj_lambda(int n, const int *__restrict__ unk, const int *__restrict__ eq,
                             const double *__restrict__ sol, double *__restrict__ lam)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    lam[i] = unk[i] ? sol[eq[i]] : 0.0;
}

// Pull form of the PPE's reverse: A-bar = -lambda x^T through the matrix
// entries, and n-bar = src/n0 lambda (fluid rows) through the density;
// adds to r**-bar of fluid m.
__global__ void k_adj_ppe_pos(int n, const int *__restrict__ typ, const int *__restrict__ unk,
                              const double *__restrict__ x, const double *__restrict__ lam,
                              const dvec *__restrict__ rss, const int *__restrict__ lst, const int *__restrict__ lidx,
                              StepConsts k, dvec *__restrict__ rbss)
{
    const int m = blockIdx.x * AB + threadIdx.x;
    if (m >= n || typ[m] != FLUID_PARTICLE) return;
    const double kap = k.diag_coeff, kof = -k.neg_coeff;   // A_ii = kap sum w, A_ij = -kof w
    auto nbar = [&](int i) { return (unk[i] && typ[i] == FLUID_PARTICLE) ? k.src_k / k.n0 * lam[i] : 0.0; };
    dvec acc = dzero();
    const double lm = lam[m], xm = x[m], nbm = nbar(m);
    const bool mu_ = unk[m] != 0;
    for (int q = lst[m]; q < lst[m + 1]; q++) {
        const int i = lidx[q];
        const dvec dim = dsub(rss[m], rss[i]);     // r_m - r_i
        const double rr = sqrt(ddot(dim, dim));
        if (!(rr > 0.0)) continue;
        const double wl = wdiff(rr, k.re_lap, k.r_min), wn = wdiff(rr, k.re_n, k.r_min);
        // dL/d rho_im for the pair seen from row i (i unknown) and from row m
        double c = 0.0;
        if (unk[i]) {
            if (mu_) c += kof * wl * lam[i] * xm;                      // A_im = -kof w
            if (typ[m] != DUMMY_PARTICLE) c -= kap * wl * lam[i] * x[i];   // A_ii += kap w
        }
        if (mu_) {
            if (unk[i]) c += kof * wl * lm * x[i];                     // A_mi
            if (typ[i] != DUMMY_PARTICLE) c -= kap * wl * lm * xm;     // A_mm += kap w
        }
        c += (nbar(i) + nbm) * wn;                                     // n_i and n_m
        // rho_im = |r_m - r_i|: d rho / d r_m = dim / rho
        daxpy(acc, c / rr, dim);
    }
    daxpy(rbss[m], 1.0, acc);
}

// Collision reverse, own terms: r*-bar = r**-bar; u*-bar = -dt r**-bar
// (r** = r* + dt (v_after - u*)); the chain adds the rest.
__global__ void k_adj_coll_init(int n, const int *__restrict__ typ, const dvec *__restrict__ rbss, StepConsts k,
                                dvec *__restrict__ rbs, dvec *__restrict__ ubs)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    if (typ[i] != FLUID_PARTICLE) {
        rbs[i] = dzero();
        ubs[i] = dzero();
        return;
    }
    rbs[i] = rbss[i];
    dvec u = dzero();
    daxpy(u, -k.dt, rbss[i]);
    ubs[i] = u;
}

// Collision reverse, the firing sequence of fluid i recomputed from r*, u*
// (candidates: the stale list, every partner closer than the collision
// distance; 2-D in ascending index as mps_fluid2D/src/operators.c:341-379,
// 3-D in list order as operators.c:395-432), then run backwards.  The own
// terms go to i's adjoints; the partner terms are left as records
// (partner, r-bar, u-bar) in i's COLL_MAX slots for k_adj_coll_pull, so
// that every sum runs in a fixed order.
__global__ void k_adj_coll(int n, const int *__restrict__ typ, const dvec *__restrict__ rs, const dvec *__restrict__ us,
                           const int *__restrict__ lst, const int *__restrict__ lidx, const dvec *__restrict__ rbss,
                           const dvec *__restrict__ ubss, StepConsts k, dvec *__restrict__ rbs, dvec *__restrict__ ubs,
                           int *__restrict__ ck, int *__restrict__ cj, dvec *__restrict__ cr, dvec *__restrict__ cu,
                           unsigned char *__restrict__ ctk, int *err)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    if (typ[i] != FLUID_PARTICLE) {
        ck[i] = 0;
        return;
    }
    int js[COLL_MAX];
    dvec vbefore[COLL_MAX];
    int K = 0;
    dvec vi = us[i];
    const dvec ri = rs[i];
    if (!k.collide) goto reverse;
    {
#if MPS_DIM == 2
    // ascending index among the partners closer than cd
    int last = -1;
    for (;;) {
        int best = INT_MAX;
        for (int q = lst[i]; q < lst[i + 1]; q++) {
            const int j = lidx[q];
            if (j <= last || j >= best || typ[j] == GHOST_PARTICLE) continue;
            const dvec d = dsub(rs[j], ri);
            if (ddot(d, d) >= k.cd2) continue;
            best = j;
        }
        if (best == INT_MAX) break;
        last = best;
        const int j = best;
#else
    for (int q = lst[i]; q < lst[i + 1]; q++) {
        const int j = lidx[q];
        if (typ[j] == GHOST_PARTICLE) continue;
#endif
        const dvec d = dsub(rs[j], ri);
        const double r2 = ddot(d, d);
        if (r2 >= k.cd2) continue;
        const double r = sqrt(r2);
        if (r < k.col_tiny) continue;
        double vn = 0.0;
#pragma unroll
        for (int dd = 0; dd < D_; dd++) vn += (vi.c[dd] - us[j].c[dd]) * (d.c[dd] / r);
        if (vn <= 0.0) continue;
        if (K == COLL_MAX) {
            ck[i] = 0;
            atomicOr(err, 1);
            return;
        }
        js[K] = j;
        vbefore[K] = vi;
        K++;
        const double imp = k.e_half * vn;
#pragma unroll
        for (int dd = 0; dd < D_; dd++) vi.c[dd] -= imp * (d.c[dd] / r);
    }
    }
reverse:
    ck[i] = K;
    {
    // v-bar of the final velocity: u** = v_f, r** = r* + dt (v_f - u*)
    dvec vb = ubss[i];
    daxpy(vb, k.dt, rbss[i]);
    dvec rown = dzero();
    const size_t base = (size_t)i * COLL_MAX;
    for (int s = K - 1; s >= 0; s--) {
        const int j = js[s];
        const dvec d = dsub(rs[j], ri);
        const double r = sqrt(ddot(d, d));
        dvec nh = dzero();
        daxpy(nh, 1.0 / r, d);
        const dvec delta = dsub(vbefore[s], us[j]);
        const double vn = ddot(delta, nh);
        const double nv = ddot(nh, vb);
        // v_after = v_before - c vn nh
        // d/d nh: -c (nv delta + vn vb); through nh = d / r
        dvec nbar = dzero();
        daxpy(nbar, -k.e_half * nv, delta);
        daxpy(nbar, -k.e_half * vn, vb);
        const double nn = ddot(nh, nbar);
        dvec dbar = nbar;
        daxpy(dbar, -nn, nh);
#pragma unroll
        for (int dd = 0; dd < D_; dd++) dbar.c[dd] /= r;
        // d = r*_j - r*_i: the partner's share, when it is fluid
        dvec uj = dzero();
        daxpy(uj, k.e_half * nv, nh);
        cj[base + s] = typ[j] == FLUID_PARTICLE ? j : -1;
        cr[base + s] = dbar;
        cu[base + s] = uj;
        ctk[base + s] = 0;
        daxpy(rown, -1.0, dbar);
        // propagate to v_before: (I - c nh nh^T) vb
        daxpy(vb, -k.e_half * nv, nh);
    }
    dvec a = rbs[i];
    daxpy(a, 1.0, rown);
    rbs[i] = a;
    a = ubs[i];
    daxpy(a, 1.0, vb);
    ubs[i] = a;
    }
}

// Collision reverse, partner terms: fluid j adds the records that name it,
// pulled from the particles of its stale list in list order (the list is
// symmetric, so every particle whose sequence holds j is in it; a pair
// fires at most once per sequence).  Each consumed record is marked.
__global__ void k_adj_coll_pull(int n, const int *__restrict__ typ, const int *__restrict__ lst,
                                const int *__restrict__ lidx, const int *__restrict__ ck, const int *__restrict__ cj,
                                const dvec *__restrict__ cr, const dvec *__restrict__ cu,
                                unsigned char *__restrict__ ctk, dvec *__restrict__ rbs, dvec *__restrict__ ubs)
{
    const int j = blockIdx.x * AB + threadIdx.x;
    if (j >= n || typ[j] != FLUID_PARTICLE) return;
    dvec ra = rbs[j], ua = ubs[j];
    for (int q = lst[j]; q < lst[j + 1]; q++) {
        const int i = lidx[q];
        const int K = ck[i];
        const size_t base = (size_t)i * COLL_MAX;
        for (int s = 0; s < K; s++) {
            if (cj[base + s] != j) continue;
            daxpy(ra, 1.0, cr[base + s]);
            daxpy(ua, 1.0, cu[base + s]);
            ctk[base + s] = 1;
            break;
        }
    }
    rbs[j] = ra;
    ubs[j] = ua;
}

// Every record with a fluid partner was consumed: a record left over means
// the partner's list does not hold the particle (an asymmetric list).
__global__ void k_adj_coll_check(int n, const int *__restrict__ ck, const int *__restrict__ cj,
                                 const unsigned char *__restrict__ ctk, int *err)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    const size_t base = (size_t)i * COLL_MAX;
    for (int s = 0; s < ck[i]; s++)
        if (cj[base + s] >= 0 && !ctk[base + s]) {
            atomicOr(err, 2);
            return;
        }
}

// Predictor reverse, step 1: a-bar of the predictor for every fluid i.
__global__ void k_adj_pred_pre(int n, const int *__restrict__ typ, const dvec *__restrict__ rbs,
                               const dvec *__restrict__ ubs, StepConsts k, dvec *__restrict__ apbar,
                               dvec *__restrict__ ustar_total)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    if (typ[i] != FLUID_PARTICLE) {
        apbar[i] = dzero();
        ustar_total[i] = dzero();
        return;
    }
    // r* = r + dt u*: u*-bar += dt r*-bar; u* = u + dt a: a-bar = dt u*-bar
    dvec ut = ubs[i];
    daxpy(ut, k.dt, rbs[i]);
    ustar_total[i] = ut;
    dvec a = dzero();
    daxpy(a, k.dt, ut);
    apbar[i] = a;
}

// Predictor reverse, step 2 (pull over the stale list): r-bar, u-bar of
// fluid m, and per-block partials of nu-bar (coeff a-bar . lap) and g-bar.
__global__ void __launch_bounds__(AB)
k_adj_pred(int n, const int *__restrict__ typ, const dvec *__restrict__ r, const dvec *__restrict__ u,
           const int *__restrict__ lst, const int *__restrict__ lidx, const dvec *__restrict__ rbs,
           const dvec *__restrict__ apbar, const dvec *__restrict__ ustar_total, StepConsts k,
           dvec *__restrict__ rb_out, dvec *__restrict__ ub_out, double *part)
{
    const int m = blockIdx.x * AB + threadIdx.x;
    double pnu = 0.0, pg[3] = {0.0, 0.0, 0.0};
    if (m < n) {
        if (typ[m] != FLUID_PARTICLE) {
            rb_out[m] = dzero();
            ub_out[m] = dzero();
        } else {
            const dvec am = apbar[m];
            dvec rbm = rbs[m];
            dvec ubm = ustar_total[m];
            dvec lap = dzero();
            double wsum = 0.0;
            for (int q = lst[m]; q < lst[m + 1]; q++) {
                const int i = lidx[q];
                const dvec dim = dsub(r[m], r[i]);
                const double rr = sqrt(ddot(dim, dim));
                const double w = wfun(rr, k.re_lap, k.r_min);
                const double wp = rr > 0.0 ? wdiff(rr, k.re_lap, k.r_min) : 0.0;
                const dvec du = dsub(u[i], u[m]);           // u_i - u_m
                daxpy(lap, w, du);
                wsum += w;
                // a_m = ... + K sum_i (u_i - u_m) w(rho_mi)
                double e = k.visc_coeff * ddot(am, du) * wp;
                if (typ[i] == FLUID_PARTICLE) {
                    // a_i = ... + K (u_m - u_i) w(rho_im): u_m's share, r_m through rho_im
                    const dvec ai = apbar[i];
                    daxpy(ubm, k.visc_coeff * w, ai);
                    const dvec dumi = dsub(u[m], u[i]);
                    e += k.visc_coeff * ddot(ai, dumi) * wp;
                }
                if (rr > 0.0) daxpy(rbm, e / rr, dim);
            }
            daxpy(ubm, -k.visc_coeff * wsum, am);
            rb_out[m] = rbm;
            ub_out[m] = ubm;
            pnu = k.coeff * ddot(am, lap);
#pragma unroll
            for (int d = 0; d < D_; d++) pg[d] = am.c[d];
        }
    }
    pnu = block_sum<double, AB>(pnu);
    double s0 = block_sum<double, AB>(pg[0]);
    double s1 = block_sum<double, AB>(pg[1]);
    double s2 = block_sum<double, AB>(pg[2]);
    if (threadIdx.x == 0) {
        part[blockIdx.x * 4 + 0] = pnu;
        part[blockIdx.x * 4 + 1] = s0;
        part[blockIdx.x * 4 + 2] = s1;
        part[blockIdx.x * 4 + 3] = s2;
    }
}

// nu-bar, g-bar += the block partials, in block order (one block).
__global__ void k_adj_param_fin(const double *part, int nblocks, double *acc)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < nblocks; b++)
        for (int q = 0; q < 4; q++) s[q] += part[b * 4 + q];
    for (int q = 0; q < 4; q++) acc[q] += s[q];
}

// ---------------------------------------------------------------------
// Replay: the forward of recorded steps re-run with other parameters
// (viscosity, gravity, initial velocities) but with every discrete
// decision of the recording -- lists, unknown set, clamp mask, the
// neighbour giving p_min, the collisions that fired and their order,
// the pair cut-offs, ghosting.  This is the function the adjoint
// differentiates; its finite differences check the adjoint where those of
// the free forward cross a flipped decision (long horizons).
// ---------------------------------------------------------------------

// predictor over the stale list, on the replayed state (r, u); non-fluid
// particles keep the recording's state
__global__ void k_rep_pred(int n, const int *__restrict__ typ, const dvec *__restrict__ r, const dvec *__restrict__ u,
                           const int *__restrict__ lst, const int *__restrict__ lidx, StepConsts k, dvec g,
                           const dvec *__restrict__ tr, const dvec *__restrict__ tu, dvec *__restrict__ rs,
                           dvec *__restrict__ us)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    if (typ[i] != FLUID_PARTICLE) {
        rs[i] = tr[i];
        us[i] = tu[i];
        return;
    }
    dvec lap = dzero();
    for (int q = lst[i]; q < lst[i + 1]; q++) {
        const int j = lidx[q];
        const dvec d = dsub(r[j], r[i]);
        const double w = wfrozen(sqrt(ddot(d, d)), dist(tr[j], tr[i]), k.re_lap, k.r_min);
        daxpy(lap, w, dsub(u[j], u[i]));
    }
    dvec a = g;
    daxpy(a, k.visc_coeff, lap);
    dvec v = u[i];
    daxpy(v, k.dt, a);
    dvec x = r[i];
    daxpy(x, k.dt, v);
    rs[i] = x;
    us[i] = v;
}

// collisions: the recording's sequence (decided on the recorded r*, u*),
// applied to the replayed r*, u*
__global__ void k_rep_coll(int n, const int *__restrict__ typ, const dvec *__restrict__ trs,
                           const dvec *__restrict__ tus, const dvec *__restrict__ rs, const dvec *__restrict__ us,
                           const int *__restrict__ lst, const int *__restrict__ lidx, StepConsts k,
                           dvec *__restrict__ rss, dvec *__restrict__ uss)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    if (typ[i] != FLUID_PARTICLE) {
        rss[i] = rs[i];
        uss[i] = us[i];
        return;
    }
    dvec vt = tus[i], v = us[i];
    const dvec ti = trs[i];
    if (k.collide) {
#if MPS_DIM == 2
        int last = -1;
        for (;;) {
            int best = INT_MAX;
            for (int q = lst[i]; q < lst[i + 1]; q++) {
                const int j = lidx[q];
                if (j <= last || j >= best || typ[j] == GHOST_PARTICLE) continue;
                const dvec d = dsub(trs[j], ti);
                if (ddot(d, d) >= k.cd2) continue;
                best = j;
            }
            if (best == INT_MAX) break;
            last = best;
            const int j = best;
#else
        for (int q = lst[i]; q < lst[i + 1]; q++) {
            const int j = lidx[q];
            if (typ[j] == GHOST_PARTICLE) continue;
#endif
            // the decision, on the recording
            const dvec d = dsub(trs[j], ti);
            const double r2 = ddot(d, d);
            if (r2 >= k.cd2) continue;
            const double r = sqrt(r2);
            if (r < k.col_tiny) continue;
            double vn = 0.0;
#pragma unroll
            for (int dd = 0; dd < D_; dd++) vn += (vt.c[dd] - tus[j].c[dd]) * (d.c[dd] / r);
            if (vn <= 0.0) continue;
#pragma unroll
            for (int dd = 0; dd < D_; dd++) vt.c[dd] -= k.e_half * vn * (d.c[dd] / r);
            // the same update on the replayed state
            const dvec e = dsub(rs[j], rs[i]);
            const double re_ = sqrt(ddot(e, e));
            double wn = 0.0;
#pragma unroll
            for (int dd = 0; dd < D_; dd++) wn += (v.c[dd] - us[j].c[dd]) * (e.c[dd] / re_);
#pragma unroll
            for (int dd = 0; dd < D_; dd++) v.c[dd] -= k.e_half * wn * (e.c[dd] / re_);
        }
    }
    dvec x = rs[i];
#pragma unroll
    for (int dd = 0; dd < D_; dd++) x.c[dd] += k.dt * (v.c[dd] - us[i].c[dd]);
    rss[i] = x;
    uss[i] = v;
}

// PPE of the replayed r** with the recording's unknown set: the eq-order
// rows of k_adj_fill, and b from the replayed density
__global__ void k_rep_fill(int n, const int *__restrict__ typ, const int *__restrict__ unk, const int *__restrict__ eq,
                           const int *__restrict__ roff, const dvec *__restrict__ rss, const dvec *__restrict__ trss,
                           const int *__restrict__ lst, const int *__restrict__ lidx, StepConsts k,
                           int *__restrict__ col, double *__restrict__ val, double *__restrict__ rhs)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n || !unk[i]) return;
    const int e = eq[i];
    int at = roff[i];
    const int dpos = at++;
    double sumw = 0.0, nd = 0.0;
    for (int q = lst[i]; q < lst[i + 1]; q++) {
        const int j = lidx[q];
        const dvec d = dsub(rss[j], rss[i]);
        const double rr = sqrt(ddot(d, d));
        const double tr = dist(trss[j], trss[i]);
        nd += wfrozen(rr, tr, k.re_n, k.r_min);
        if (typ[j] == DUMMY_PARTICLE) continue;
        const double w = wfrozen(rr, tr, k.re_lap, k.r_min);
        sumw += w;
        if (unk[j]) {
            col[at] = eq[j];
            val[at] = k.neg_coeff * w;
            at++;
        }
    }
    col[dpos] = e;
    val[dpos] = k.diag_coeff * sumw;
    rhs[e] = typ[i] == FLUID_PARTICLE ? k.src_k * (nd - k.n0) / k.n0 : 0.0;
}

// pressure gradient with the recording's clamp mask and p_min neighbour
__global__ void k_rep_corr(int n, const int *__restrict__ typ, const int *__restrict__ unk,
                           const double *__restrict__ tx, const double *__restrict__ x, const dvec *__restrict__ trss,
                           const dvec *__restrict__ rss, const dvec *__restrict__ uss, const int *__restrict__ lst,
                           const int *__restrict__ lidx, StepConsts k, int clamp, dvec *__restrict__ r,
                           dvec *__restrict__ u)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    if (typ[i] != FLUID_PARTICLE) {
        r[i] = rss[i];
        u[i] = uss[i];
        return;
    }
    // recorded pressures decide the clamp and p_min
    auto tp = [&](int j) {
        double p = unk[j] ? tx[j] : 0.0;
        if (clamp && typ[j] == FLUID_PARTICLE && p < 0.0) p = 0.0;
        return p;
    };
    auto pp = [&](int j) {
        if (!unk[j]) return 0.0;
        if (clamp && typ[j] == FLUID_PARTICLE && tx[j] < 0.0) return 0.0;
        return x[j];
    };
    double pm = tp(i);
    int mu = i;
    for (int q = lst[i]; q < lst[i + 1]; q++) {
        const int j = lidx[q];
        if (tp(j) < pm) {
            pm = tp(j);
            mu = j;
        }
    }
    const double pmin = pp(mu);
    dvec grad = dzero();
    for (int q = lst[i]; q < lst[i + 1]; q++) {
        const int j = lidx[q];
        const dvec td = dsub(trss[j], trss[i]);
        if (ddot(td, td) < k.grad_min_r2) continue;
        const dvec d = dsub(rss[j], rss[i]);
        const double r2 = ddot(d, d);
        const double w = wfrozen(sqrt(r2), sqrt(ddot(td, td)), k.re_lap, k.r_min);
        daxpy(grad, (pp(j) - pmin) / r2 * w, d);
    }
    dvec a = dzero();
    daxpy(a, k.neg_gc / k.rho, grad);
    dvec v = uss[i];
    daxpy(v, k.dt, a);
    dvec xx = rss[i];
    daxpy(xx, k.dt * k.dt, a);
    u[i] = v;
    r[i] = xx;
}

__global__ void k_rep_set_u0(int n, const int *__restrict__ typ, const double *__restrict__ u0, dvec *__restrict__ u)
{
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n || typ[i] != FLUID_PARTICLE) return;
#pragma unroll
    for (int d = 0; d < D_; d++) u[i].c[d] = u0[(size_t)i * D_ + d];
}

// ---------------------------------------------------------------------
// driver
// ---------------------------------------------------------------------

int mps_ctx_step(Ctx *c, const SimConfig *cf);                     // runtime.hip
int csr_cg_device(int n, long long nnz, const int *rp, const int *col, const double *val, const double *b,
                  double *x, const SolveOpts &o, SolveResult *res, hipStream_t s);   // k_ic.hip

// The collision reverse's error bits: 1, a particle collided more often in
// one step than its COLL_MAX record slots hold; 2, a record's partner did
// not find the particle in its own list.
static int coll_error(int bits)
{
    if (bits & 1)
        mps_set_error(MPS_ERR_UNSUPPORTED, __FILE__, __LINE__,
                      "more collisions of one particle in one step than the reverse sweep holds");
    else
        mps_set_error(MPS_ERR_ARG, __FILE__, __LINE__,
                      "a collision partner's neighbour list does not hold the particle (asymmetric list)");
    return (bits & 1) ? MPS_ERR_UNSUPPORTED : MPS_ERR_ARG;
}

struct AdjWork {
    int n = 0;
    dvec *rb = nullptr, *ub = nullptr, *rbss = nullptr, *ubss = nullptr, *rbs = nullptr, *ubs = nullptr;
    dvec *abar = nullptr, *F = nullptr, *apbar = nullptr, *ust = nullptr;
    int *mu = nullptr, *len = nullptr, *roff = nullptr, *eq = nullptr, *col = nullptr, *err = nullptr, *scan_blk = nullptr;
    // collision records: count per particle, COLL_MAX (partner, r-bar,
    // u-bar, consumed) slots per particle
    int *ck = nullptr, *cj = nullptr;
    dvec *cr = nullptr, *cu = nullptr;
    unsigned char *ctk = nullptr;
    double *xbar = nullptr, *lam = nullptr, *val = nullptr, *rhs = nullptr, *sol = nullptr, *part = nullptr,
           *params = nullptr;
    size_t cap_nnz = 0;
};

static void adj_free(AdjWork &w)
{
    void *b[] = {w.rb, w.ub, w.rbss, w.ubss, w.rbs, w.ubs, w.abar, w.F, w.apbar, w.ust, w.mu, w.len, w.roff, w.eq,
                 w.col, w.err, w.scan_blk, w.xbar, w.lam, w.val, w.rhs, w.sol, w.part, w.params,
                 w.ck, w.cj, w.cr, w.cu, w.ctk};
    for (void *p : b)
        if (p) hipFree(p);
    w = AdjWork();
}

static int adj_alloc(AdjWork &w, int n, hipStream_t s)
{
    w.n = n;
    dvec **dv[] = {&w.rb, &w.ub, &w.rbss, &w.ubss, &w.rbs, &w.ubs, &w.abar, &w.F, &w.apbar, &w.ust};
    for (dvec **p : dv) HIP_OK(hipMalloc((void **)p, sizeof(dvec) * n));
    HIP_OK(hipMalloc(&w.mu, sizeof(int) * n));
    HIP_OK(hipMalloc(&w.len, sizeof(int) * n));
    HIP_OK(hipMalloc(&w.roff, sizeof(int) * ((size_t)n + 1)));
    HIP_OK(hipMalloc(&w.eq, sizeof(int) * ((size_t)n + 1)));
    HIP_OK(hipMalloc(&w.err, sizeof(int)));
    HIP_OK(hipMalloc(&w.scan_blk, sizeof(int) * 8200));
    HIP_OK(hipMalloc(&w.xbar, sizeof(double) * n));
    HIP_OK(hipMalloc(&w.lam, sizeof(double) * n));
    HIP_OK(hipMalloc(&w.rhs, sizeof(double) * n));
    HIP_OK(hipMalloc(&w.sol, sizeof(double) * n));
    HIP_OK(hipMalloc(&w.part, sizeof(double) * 4 * ((size_t)ab(n) + 1)));
    HIP_OK(hipMalloc(&w.params, sizeof(double) * 4));
    HIP_OK(hipMalloc(&w.ck, sizeof(int) * n));
    HIP_OK(hipMalloc(&w.cj, sizeof(int) * COLL_MAX * (size_t)n));
    HIP_OK(hipMalloc(&w.cr, sizeof(dvec) * COLL_MAX * (size_t)n));
    HIP_OK(hipMalloc(&w.cu, sizeof(dvec) * COLL_MAX * (size_t)n));
    HIP_OK(hipMalloc(&w.ctk, COLL_MAX * (size_t)n));
    // on the stream the reverse kernels run on (a non-blocking stream does
    // not wait for null-stream work)
    HIP_OK(hipMemsetAsync(w.params, 0, sizeof(double) * 4, s));
    HIP_OK(hipMemsetAsync(w.err, 0, sizeof(int), s));
    return 0;
}

// The step's PPE in eq order (the unknowns of the recording t, rows of
// k_adj_fill or, with replay, k_rep_fill on positions rss) solved by CG;
// right-hand side w.xbar (adjoint) or the replayed source; the solution,
// per eq row, in w.sol.
static int eq_csr_solve(AdjWork &w, const TapeStep &t, bool replay, const dvec *rss, const StepConsts &k,
                        const SolveOpts &so, hipStream_t s)
{
    const int n = w.n;
    const unsigned g = ab(n);
    hipLaunchKernelGGL(k_adj_rowlen, dim3(g), dim3(AB), 0, s, n, t.unk, t.lst, t.lidx, w.len);
    LAUNCH_CHECK();
    int rc = launch_scan_excl(w.len, w.roff, n, w.scan_blk, s);
    if (!rc) rc = launch_scan_excl(t.unk, w.eq, n, w.scan_blk, s);
    if (rc) return rc;
    int tot[2];
    HIP_OK(hipMemcpyAsync(&tot[0], w.roff + n, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(&tot[1], w.eq + n, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const long long nnz = tot[0];
    const int neq = tot[1];
    if ((size_t)nnz > w.cap_nnz) {
        if (w.col) hipFree(w.col);
        if (w.val) hipFree(w.val);
        w.col = nullptr;
        w.val = nullptr;
        w.cap_nnz = (size_t)nnz + (size_t)nnz / 4 + 64;
        HIP_OK(hipMalloc(&w.col, sizeof(int) * w.cap_nnz));
        HIP_OK(hipMalloc(&w.val, sizeof(double) * w.cap_nnz));
    }
    if (replay)
        hipLaunchKernelGGL(k_rep_fill, dim3(g), dim3(AB), 0, s, n, t.typ, t.unk, w.eq, w.roff, rss, t.rss, t.lst,
                           t.lidx, k, w.col, w.val, w.rhs);
    else
        hipLaunchKernelGGL(k_adj_fill, dim3(g), dim3(AB), 0, s, n, t.typ, t.unk, w.eq, w.roff, rss, t.lst, t.lidx,
                           w.xbar, k, w.col, w.val, w.rhs);
    LAUNCH_CHECK();
    if (neq == 0) return 0;
    // row pointers of the eq rows: the unknowns' offsets, in order
    std::vector<int> unk_h(n), roff_h(n + 1);
    HIP_OK(hipMemcpyAsync(roff_h.data(), w.roff, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(unk_h.data(), t.unk, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    std::vector<int> rp(neq + 1);
    int e = 0;
    for (int i = 0; i < n; i++)
        if (unk_h[i]) rp[e++] = roff_h[i];
    rp[neq] = (int)nnz;
    int *rp_d = w.len;    // free again; n >= neq + 1 unless every particle is an unknown
    int *tmp = nullptr;
    if (neq + 1 > n) {
        HIP_OK(hipMalloc(&tmp, sizeof(int) * (neq + 1)));
        rp_d = tmp;
    }
    SolveResult sr{};
    if (hipMemcpyAsync(rp_d, rp.data(), sizeof(int) * (neq + 1), hipMemcpyHostToDevice, s) != hipSuccess)
        rc = MPS_ERR_HIP;
    if (!rc) rc = csr_cg_device(neq, nnz, rp_d, w.col, w.val, w.rhs, w.sol, so, &sr, s);
    if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = MPS_ERR_HIP;
    if (tmp) hipFree(tmp);
    return rc;
}

// The reverse of one recorded step: (rb, ub) = adjoint of the step's
// outputs in, adjoint of its inputs out.  ls/lsi: the list the predictor
// used (the previous step's search).
static int adj_step(AdjWork &w, const TapeStep &t, const int *ls, const int *lsi, const StepConsts &k, int clamp,
                    const SolveOpts &so, hipStream_t s)
{
    const int n = w.n;
    const unsigned g = ab(n);
    hipLaunchKernelGGL(k_adj_corr_pre, dim3(g), dim3(AB), 0, s, n, t.typ, t.x, t.unk, t.rss, t.lst, t.lidx, w.rb, w.ub,
                       k, clamp, w.abar, w.F, w.mu);
    hipLaunchKernelGGL(k_adj_corr, dim3(g), dim3(AB), 0, s, n, t.typ, t.x, t.unk, t.rss, t.lst, t.lidx, w.rb, w.ub,
                       w.abar, w.F, w.mu, k, clamp, w.xbar, w.rbss, w.ubss);
    // adjoint PPE on the step's matrix
    int rc = eq_csr_solve(w, t, false, t.rss, k, so, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_adj_lambda, dim3(g), dim3(AB), 0, s, n, t.unk, w.eq, w.sol, w.lam);
    hipLaunchKernelGGL(k_adj_ppe_pos, dim3(g), dim3(AB), 0, s, n, t.typ, t.unk, t.x, w.lam, t.rss, t.lst, t.lidx, k,
                       w.rbss);
    hipLaunchKernelGGL(k_adj_coll_init, dim3(g), dim3(AB), 0, s, n, t.typ, w.rbss, k, w.rbs, w.ubs);
    hipLaunchKernelGGL(k_adj_coll, dim3(g), dim3(AB), 0, s, n, t.typ, t.rs, t.us, ls, lsi, w.rbss, w.ubss, k, w.rbs,
                       w.ubs, w.ck, w.cj, w.cr, w.cu, w.ctk, w.err);
    hipLaunchKernelGGL(k_adj_coll_pull, dim3(g), dim3(AB), 0, s, n, t.typ, ls, lsi, w.ck, w.cj, w.cr, w.cu, w.ctk,
                       w.rbs, w.ubs);
    hipLaunchKernelGGL(k_adj_coll_check, dim3(g), dim3(AB), 0, s, n, w.ck, w.cj, w.ctk, w.err);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_adj_pred_pre, dim3(g), dim3(AB), 0, s, n, t.typ, w.rbs, w.ubs, k, w.apbar, w.ust);
    hipLaunchKernelGGL(k_adj_pred, dim3(g), dim3(AB), 0, s, n, t.typ, t.r, t.u, ls, lsi, w.rbs, w.apbar, w.ust, k, w.rb,
                       w.ub, w.part);
    hipLaunchKernelGGL(k_adj_param_fin, dim3(1), dim3(64), 0, s, w.part, (int)g, w.params);
    LAUNCH_CHECK();
    return 0;
}

// Unsupported scheme variants of the adjoint (their operators have no
// reverse here): returns 0 or a message.
static const char *adj_unsupported(const SimConfig *cf)
{
    if (cf->surface_tension_enabled) return "surface tension";
    if (cf->hs_mode != 0) return "hs_mode";
    if (cf->ppe_type != 0) return "ppe_type";
    if (cf->cmps_gradient != 0) return "cmps_gradient";
    if (cf->surface_detection_method != 0) return "surface_detection_method";
#ifdef MPS_FRICTION_FIELDS
    if (cf->friction_enabled) return "hybrid wall friction";
#endif
    return nullptr;
}

extern "C" int mps_hip_adjoint(ParticleSystem *ps, int nsteps, int checkpoint_every, const double *target,
                               double *loss_out, double *grad_nu, double *grad_g, double *grad_u0)
{
    if (!ps || nsteps < 1 || !target) return MPS_ERR_ARG;
    if (!g_config) return MPS_ERR_NO_CONFIG;
    SimConfig *cf = g_config;
    if (const char *u = adj_unsupported(cf)) {
        mps_set_error(MPS_ERR_UNSUPPORTED, __FILE__, __LINE__, u);
        return MPS_ERR_UNSUPPORTED;
    }
    const int n = ps->num;
    const int K = checkpoint_every > 0 ? checkpoint_every : nsteps;
    const int nseg = (nsteps + K - 1) / K;
    Ctx *c = mps_ctx_for(ps, 1);
    if (!c) return mps_hip_last_error();
    hipStream_t s = c->stream;
    const double re = mps_search_radius(cf);
    const Grid grid = mps_grid_for(re, cf->domain_min, cf->domain_max);
    NeighborList *nl = neighbor_list_create(n, cf->max_neighbors);
    if (!nl) return MPS_ERR_ALLOC;
    int rc = mps_ctx_upload(c, ps, nullptr, grid);
    StepConsts k = mps_consts(cf, ps, grid, re);
    k.stride = c->stride;
    if (!rc) rc = mps_ctx_search(c, k);
    // forward with checkpoints (particles and the stale list, on the host)
    std::vector<std::vector<Particle>> ck_p(nseg);
    std::vector<std::vector<int>> ck_cnt(nseg), ck_nbr(nseg);
    const size_t nbr_len = (size_t)n * (size_t)cf->max_neighbors;
    for (int sg = 0; sg < nseg && !rc; sg++) {
        rc = mps_ctx_download(c, ps, nl);
        if (rc) break;
        ck_p[sg].assign(ps->particles, ps->particles + n);
        ck_cnt[sg].assign(nl->count, nl->count + n);
        ck_nbr[sg].assign(nl->neighbors, nl->neighbors + nbr_len);
        const int m = (sg == nseg - 1) ? nsteps - sg * K : K;
        for (int q = 0; q < m && !rc; q++) rc = mps_ctx_step(c, cf);
    }
    if (!rc) rc = mps_ctx_download(c, ps, nullptr);
    if (rc) {
        neighbor_list_free(nl);
        return rc;
    }
    const std::vector<Particle> final_p(ps->particles, ps->particles + n);
    // loss and its gradient in the final positions
    std::vector<dvec> rb_h(n), zero_h(n);
    int nf = 0;
    for (int i = 0; i < n; i++)
        if (final_p[i].type == FLUID_PARTICLE) nf++;
    double loss = 0.0;
    for (int i = 0; i < n; i++) {
        dvec z;
        memset(&z, 0, sizeof z);
        zero_h[i] = z;
        rb_h[i] = z;
        if (final_p[i].type != FLUID_PARTICLE) continue;
        for (int d = 0; d < D_; d++) {
            const double e = final_p[i].pos[d] - target[(size_t)i * D_ + d];
            loss += e * e;
            rb_h[i].c[d] = 2.0 * e / nf;
        }
    }
    loss = nf > 0 ? loss / nf : 0.0;
    AdjWork w;
    std::vector<TapeStep> tapes(K);
    TapeStep ls0;
    rc = adj_alloc(w, n, s);
    for (int q = 0; q < K && !rc; q++) rc = tape_alloc(&tapes[q], n);
    if (!rc) {
        hipError_t e = hipMalloc(&ls0.lst, sizeof(int) * ((size_t)n + 1));
        if (e != hipSuccess) rc = MPS_ERR_HIP;
    }
    if (!rc) {
        if (hipMemcpyAsync(w.rb, rb_h.data(), sizeof(dvec) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(w.ub, zero_h.data(), sizeof(dvec) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = MPS_ERR_HIP;
    }
    SolveOpts so{};
    so.solver = 0;
    so.max_iter = cf->cg_max_iter;
    so.tol = cf->cg_tolerance;
    so.tol_mode = (int)mps_hip_get_option("tol_mode");
    so.check_every = (int)mps_hip_get_option("check_every");
    for (int sg = nseg - 1; sg >= 0 && !rc; sg--) {
        // recompute the segment from its checkpoint, recording it
        memcpy(ps->particles, ck_p[sg].data(), sizeof(Particle) * n);
        memcpy(nl->count, ck_cnt[sg].data(), sizeof(int) * n);
        memcpy(nl->neighbors, ck_nbr[sg].data(), sizeof(int) * nbr_len);
        rc = mps_ctx_upload(c, ps, nl, grid);
        if (!rc) rc = rec_list(c, &ls0);
        const int m = (sg == nseg - 1) ? nsteps - sg * K : K;
        for (int q = 0; q < m && !rc; q++) {
            g_rec = &tapes[q];
            g_rec_err = 0;
            rc = mps_ctx_step(c, cf);
            g_rec = nullptr;
            if (!rc) rc = g_rec_err;
        }
        if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = MPS_ERR_HIP;
        for (int q = m - 1; q >= 0 && !rc; q--) {
            const TapeStep &lsrc = q == 0 ? ls0 : tapes[q - 1];
            rc = adj_step(w, tapes[q], lsrc.lst, lsrc.lidx, k, cf->clamp_negative_pressure, so, s);
        }
    }
    int herr = 0;
    double params[4] = {0.0, 0.0, 0.0, 0.0};
    std::vector<dvec> ub_h(n);
    if (!rc) {
        if (hipMemcpyAsync(&herr, w.err, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(params, w.params, sizeof params, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(ub_h.data(), w.ub, sizeof(dvec) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = MPS_ERR_HIP;
    }
    if (!rc && herr) rc = coll_error(herr);
    if (!rc) {
        if (loss_out) *loss_out = loss;
        if (grad_nu) *grad_nu = params[0];
        if (grad_g)
            for (int d = 0; d < D_; d++) grad_g[d] = params[1 + d];
        if (grad_u0)
            for (int i = 0; i < n; i++)
                for (int d = 0; d < D_; d++) grad_u0[(size_t)i * D_ + d] = ub_h[i].c[d];
    }
    // ps holds the final forward state; its device copy is dropped
    memcpy(ps->particles, final_p.data(), sizeof(Particle) * n);
    mps_hip_release(ps);
    for (auto &t : tapes) tape_free(&t);
    if (ls0.lst) hipFree(ls0.lst);
    if (ls0.lidx) hipFree(ls0.lidx);
    adj_free(w);
    neighbor_list_free(nl);
    return rc;
}

// Loss of the frozen-decision replay (see above) for several parameter
// sets: the steps are recorded once from ps under g_config, then each set
// (nus[q], gs[q*DIM..], u0s[q*N*DIM..] or ps's velocities when u0s is null)
// replays them.  Checks the adjoint by finite differences that cannot
// cross a decision.  ps is left in the recorded run's final state.
extern "C" int mps_hip_replay_losses(ParticleSystem *ps, int nsteps, const double *target, int nsets,
                                     const double *nus, const double *gs, const double *u0s, double *losses)
{
    if (!ps || nsteps < 1 || !target || nsets < 1 || !nus || !gs || !losses) return MPS_ERR_ARG;
    if (!g_config) return MPS_ERR_NO_CONFIG;
    SimConfig *cf = g_config;
    if (const char *u = adj_unsupported(cf)) {
        mps_set_error(MPS_ERR_UNSUPPORTED, __FILE__, __LINE__, u);
        return MPS_ERR_UNSUPPORTED;
    }
    const int n = ps->num;
    Ctx *c = mps_ctx_for(ps, 1);
    if (!c) return mps_hip_last_error();
    hipStream_t s = c->stream;
    const double re = mps_search_radius(cf);
    const Grid grid = mps_grid_for(re, cf->domain_min, cf->domain_max);
    int rc = mps_ctx_upload(c, ps, nullptr, grid);
    StepConsts k = mps_consts(cf, ps, grid, re);
    k.stride = c->stride;
    if (!rc) rc = mps_ctx_search(c, k);
    std::vector<TapeStep> tapes(nsteps);
    TapeStep ls0;
    for (int t = 0; t < nsteps && !rc; t++) rc = tape_alloc(&tapes[t], n);
    if (!rc && hipMalloc(&ls0.lst, sizeof(int) * ((size_t)n + 1)) != hipSuccess) rc = MPS_ERR_HIP;
    if (!rc) rc = rec_list(c, &ls0);
    for (int t = 0; t < nsteps && !rc; t++) {
        g_rec = &tapes[t];
        g_rec_err = 0;
        rc = mps_ctx_step(c, cf);
        g_rec = nullptr;
        if (!rc) rc = g_rec_err;
    }
    if (!rc) rc = mps_ctx_download(c, ps, nullptr);
    AdjWork w;
    dvec *r = nullptr, *u = nullptr, *rs = nullptr, *us = nullptr, *rss = nullptr, *uss = nullptr;
    double *xr = nullptr, *u0d = nullptr;
    if (!rc) rc = adj_alloc(w, n, s);
    if (!rc) {
        dvec **bufs[] = {&r, &u, &rs, &us, &rss, &uss};
        for (dvec **b : bufs)
            if (hipMalloc((void **)b, sizeof(dvec) * n) != hipSuccess) rc = MPS_ERR_HIP;
        if (hipMalloc(&xr, sizeof(double) * n) != hipSuccess) rc = MPS_ERR_HIP;
        if (u0s && hipMalloc(&u0d, sizeof(double) * (size_t)n * D_) != hipSuccess) rc = MPS_ERR_HIP;
    }
    SolveOpts so{};
    so.solver = 0;
    so.max_iter = cf->cg_max_iter;
    so.tol = cf->cg_tolerance;
    so.tol_mode = (int)mps_hip_get_option("tol_mode");
    so.check_every = (int)mps_hip_get_option("check_every");
    const unsigned g = ab(n);
    std::vector<dvec> rh(n);
    for (int q = 0; q < nsets && !rc; q++) {
        StepConsts kq = k;
        kq.visc_coeff = nus[q] * k.coeff;
        dvec gq = dzero_host();
        for (int d = 0; d < D_; d++) gq.c[d] = gs[(size_t)q * D_ + d];
        // errors end the loop and fall through to the frees below
        if (hipMemcpyAsync(r, tapes[0].r, sizeof(dvec) * n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
            hipMemcpyAsync(u, tapes[0].u, sizeof(dvec) * n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
            (u0s && hipMemcpyAsync(u0d, u0s + (size_t)q * n * D_, sizeof(double) * (size_t)n * D_,
                                   hipMemcpyHostToDevice, s) != hipSuccess)) {
            rc = MPS_ERR_HIP;
            break;
        }
        if (u0s) hipLaunchKernelGGL(k_rep_set_u0, dim3(g), dim3(AB), 0, s, n, tapes[0].typ, u0d, u);
        for (int t = 0; t < nsteps && !rc; t++) {
            const TapeStep &T = tapes[t];
            const TapeStep &L = t == 0 ? ls0 : tapes[t - 1];
            hipLaunchKernelGGL(k_rep_pred, dim3(g), dim3(AB), 0, s, n, T.typ, r, u, L.lst, L.lidx, kq, gq, T.r, T.u, rs,
                               us);
            hipLaunchKernelGGL(k_rep_coll, dim3(g), dim3(AB), 0, s, n, T.typ, T.rs, T.us, rs, us, L.lst, L.lidx, kq, rss,
                               uss);
            if (hipGetLastError() != hipSuccess) {
                rc = MPS_ERR_HIP;
                break;
            }
            rc = eq_csr_solve(w, T, true, rss, kq, so, s);
            if (rc) break;
            hipLaunchKernelGGL(k_adj_lambda, dim3(g), dim3(AB), 0, s, n, T.unk, w.eq, w.sol, xr);
            hipLaunchKernelGGL(k_rep_corr, dim3(g), dim3(AB), 0, s, n, T.typ, T.unk, T.x, xr, T.rss, rss, uss, T.lst,
                               T.lidx, kq, cf->clamp_negative_pressure, r, u);
            if (hipGetLastError() != hipSuccess) rc = MPS_ERR_HIP;
        }
        if (rc) break;
        if (hipMemcpyAsync(rh.data(), r, sizeof(dvec) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = MPS_ERR_HIP;
            break;
        }
        double l = 0.0;
        int nf = 0;
        for (int i = 0; i < n; i++) {
            if (ps->particles[i].type != FLUID_PARTICLE) continue;
            nf++;
            for (int d = 0; d < D_; d++) {
                const double e = rh[i].c[d] - target[(size_t)i * D_ + d];
                l += e * e;
            }
        }
        losses[q] = nf > 0 ? l / nf : 0.0;
    }
    mps_hip_release(ps);
    for (auto &t : tapes) tape_free(&t);
    if (ls0.lst) hipFree(ls0.lst);
    if (ls0.lidx) hipFree(ls0.lidx);
    void *bufs[] = {r, u, rs, us, rss, uss, xr, u0d};
    for (void *b : bufs)
        if (b) hipFree(b);
    adj_free(w);
    return rc;
}

// One step forward from ps (the stale list nl as simulation_step receives
// it, or a search of ps's positions when nl is null), recorded, then its
// reverse for the output adjoints (rbar_out, ubar_out): a vector-Jacobian
// product of the step.  ps and nl advance as simulation_step leaves them.
extern "C" int mps_hip_step_backward(ParticleSystem *ps, NeighborList *nl, const double *rbar_out,
                                     const double *ubar_out, double *rbar_in, double *ubar_in, double *grad_nu,
                                     double *grad_g)
{
    if (!ps || !rbar_out || !ubar_out) return MPS_ERR_ARG;
    if (!g_config) return MPS_ERR_NO_CONFIG;
    SimConfig *cf = g_config;
    if (const char *u = adj_unsupported(cf)) {
        mps_set_error(MPS_ERR_UNSUPPORTED, __FILE__, __LINE__, u);
        return MPS_ERR_UNSUPPORTED;
    }
    const int n = ps->num;
    Ctx *c = mps_ctx_for(ps, 1);
    if (!c) return mps_hip_last_error();
    hipStream_t s = c->stream;
    const double re = mps_search_radius(cf);
    const Grid grid = mps_grid_for(re, cf->domain_min, cf->domain_max);
    int rc = mps_ctx_upload(c, ps, nl, grid);
    StepConsts k = mps_consts(cf, ps, grid, re);
    k.stride = c->stride;
    if (!rc && !nl) rc = mps_ctx_search(c, k);
    AdjWork w;
    TapeStep t, ls0;
    if (!rc) rc = adj_alloc(w, n, s);
    if (!rc) rc = tape_alloc(&t, n);
    if (!rc && hipMalloc(&ls0.lst, sizeof(int) * ((size_t)n + 1)) != hipSuccess) rc = MPS_ERR_HIP;
    if (!rc) rc = rec_list(c, &ls0);
    if (!rc) {
        g_rec = &t;
        g_rec_err = 0;
        rc = mps_ctx_step(c, cf);
        g_rec = nullptr;
        if (!rc) rc = g_rec_err;
    }
    std::vector<dvec> rb_h(n), ub_h(n);
    for (int i = 0; i < n; i++) {
        rb_h[i] = dzero_host();
        ub_h[i] = dzero_host();
        for (int d = 0; d < D_; d++) {
            rb_h[i].c[d] = rbar_out[(size_t)i * D_ + d];
            ub_h[i].c[d] = ubar_out[(size_t)i * D_ + d];
        }
    }
    if (!rc) {
        if (hipMemcpyAsync(w.rb, rb_h.data(), sizeof(dvec) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(w.ub, ub_h.data(), sizeof(dvec) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = MPS_ERR_HIP;
    }
    SolveOpts so{};
    so.solver = 0;
    so.max_iter = cf->cg_max_iter;
    so.tol = cf->cg_tolerance;
    so.tol_mode = (int)mps_hip_get_option("tol_mode");
    so.check_every = (int)mps_hip_get_option("check_every");
    if (!rc) rc = adj_step(w, t, ls0.lst, ls0.lidx, k, cf->clamp_negative_pressure, so, s);
    int herr = 0;
    double params[4] = {0.0, 0.0, 0.0, 0.0};
    if (!rc) {
        if (hipMemcpyAsync(&herr, w.err, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(params, w.params, sizeof params, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(rb_h.data(), w.rb, sizeof(dvec) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(ub_h.data(), w.ub, sizeof(dvec) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = MPS_ERR_HIP;
    }
    if (!rc && herr) rc = coll_error(herr);
    if (!rc) rc = mps_ctx_download(c, ps, nl);
    if (!rc) {
        if (grad_nu) *grad_nu = params[0];
        if (grad_g)
            for (int d = 0; d < D_; d++) grad_g[d] = params[1 + d];
        for (int i = 0; i < n; i++)
            for (int d = 0; d < D_; d++) {
                if (rbar_in) rbar_in[(size_t)i * D_ + d] = rb_h[i].c[d];
                if (ubar_in) ubar_in[(size_t)i * D_ + d] = ub_h[i].c[d];
            }
    }
    mps_hip_release(ps);
    tape_free(&t);
    if (ls0.lst) hipFree(ls0.lst);
    if (ls0.lidx) hipFree(ls0.lidx);
    adj_free(w);
    return rc;
}
