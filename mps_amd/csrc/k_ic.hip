// k_ic.hip -- ICCG with the reference's exact IC(0) factor, on gfx950.
//
// The reference factors A in its own unknown order ("eq order": ascending
// particle index over unknowns, pressure_solver.c:411-422) and its factor
// values depend on that order, so this path rebuilds A in eq order with
// columns sorted (the reference's CSR, :437-551), then:
//   factor    IC(0) row by row (pressure_solver.c:156-232 restated up-
//             looking; same operations in the same order): entry (e,k) =
//             (A_ek - sum_j L_ej L_kj) * dinv_k, d_e = A_ee - sum L_ep^2,
//             d <= 0 -> A_ee (if > 0) or 1e-30, dinv_e = 1/sqrt(d)
//   forward   y_e = (r_e - sum_p L_ep y_col) * dinv_e         (:235-243)
//   backward  z_j = (y_j - sum_{i>j, descending} L_ij z_i) * dinv_j, the
//             reference's column scatter (:246-255) as a gather over the
//             upper half of row j (the pattern is symmetric)
// The 2-D program's dense formulas (mps_fluid2D/src/pressure_solver.c:
// 86-141: sums first, divide by L_ii) are selected by `dense`.
//
// All three are dependency-bound (about 9*nx + 3*ny + nz levels for the
// reference's lattice order).  They run "sync-free": persistent waves claim
// 64-row chunks in row order from an atomic ticket, so every row a wave
// waits on belongs to a wave that is already running; each lane publishes
// its row's value with an agent-scope atomic store (the value itself, or
// 1/sqrt(d) for the factor, replaces a sentinel NaN pattern) and polls the
// values it needs with agent-scope loads, which bypass the per-XCD L2.
// Every polling loop is bounded: past the bound the row is abandoned and an
// error flag set, so the kernel always drains.
#include <limits.h>
#include <stdint.h>

#include <vector>

#include "ctx.hpp"

#define IC_BLK 256
#define SPIN_LIMIT (1LL << 22)

static inline unsigned nblk(size_t n, unsigned b = IC_BLK) { return (unsigned)((n + b - 1) / b); }

#define LAUNCH_CHECK()                                                       \
    do {                                                                     \
        hipError_t e_ = hipGetLastError();                                   \
        if (e_ != hipSuccess) {                                              \
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(e_)); \
            return MPS_ERR_HIP;                                              \
        }                                                                    \
    } while (0)

// A signalling-NaN payload no arithmetic produces: "not yet computed".
#define SENT_BITS 0x7FF4DEADBEEF1234ULL

__device__ __forceinline__ bool is_sent(double v) { return __double_as_longlong(v) == (long long)SENT_BITS; }
__device__ __forceinline__ double sent() { return __longlong_as_double((long long)SENT_BITS); }

__device__ __forceinline__ double ld_acq(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_rlx(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct IcData {
    int n_eq = 0, cap_n = 0;
    long long nnz = 0, cap_nnz = 0;
    int *rp = nullptr, *col = nullptr, *dpos = nullptr, *tpos = nullptr;
    int *eq_of_slot = nullptr, *slot_of_eq = nullptr, *flag = nullptr, *eq_of_orig = nullptr, *len = nullptr;
    double *aval = nullptr, *lval = nullptr, *dinv = nullptr;
    double *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *Ap = nullptr, *y = nullptr, *b = nullptr;
    unsigned *ticket = nullptr;   // [4]
    int *err = nullptr;
    double *part = nullptr;
    int levels = 0;
};

struct IcCtrl {
    double rz[2], rr_new, pap, alpha, beta, bnorm, stop, res;
    int done, iters, breakdown, pad;
};

// ---------------------------------------------------------------------
// eq-order CSR of A
// ---------------------------------------------------------------------

__global__ void k_ic_flag(int n, const int *O, const int *unk, int *flag)
{
    const int s = blockIdx.x * IC_BLK + threadIdx.x;
    if (s < n) flag[O[s]] = unk[s];
}

__global__ void k_ic_maps(int n, const int *O, const int *unk, const int *Alen, const int *eq_of_orig,
                          int *eq_of_slot, int *slot_of_eq, int *len)
{
    const int s = blockIdx.x * IC_BLK + threadIdx.x;
    if (s >= n) return;
    const int e = unk[s] ? eq_of_orig[O[s]] : -1;
    eq_of_slot[s] = e;
    if (e >= 0) {
        slot_of_eq[e] = s;
        len[e] = Alen[s] + 1;
    }
}

// Row e of the eq-order CSR: off-diagonals then the diagonal, sorted by
// column in place (insertion sort; columns of a row are distinct).
__global__ void k_ic_fill(int n_eq, int stride, const int *slot_of_eq, const int *eq_of_slot,
                          const int *Alen, const double *Aval, const int *Acol, const double *diag,
                          const double *bslot, const int *rp, int *col, double *aval, int *dpos, double *beq)
{
    const int e = blockIdx.x * IC_BLK + threadIdx.x;
    if (e >= n_eq) return;
    const int s = slot_of_eq[e];
    const int m = Alen[s];
    const int base = rp[e];
    for (int q = 0; q < m; q++) {
        const int cq = eq_of_slot[Acol[sell_off(s, q, stride)]];
        const double vq = Aval[sell_off(s, q, stride)];
        int at = base + q;
        while (at > base && col[at - 1] > cq) {
            col[at] = col[at - 1];
            aval[at] = aval[at - 1];
            at--;
        }
        col[at] = cq;
        aval[at] = vq;
    }
    // diagonal
    int at = base + m;
    while (at > base && col[at - 1] > e) {
        col[at] = col[at - 1];
        aval[at] = aval[at - 1];
        at--;
    }
    col[at] = e;
    aval[at] = diag[s];
    dpos[e] = at - base;
    beq[e] = bslot[s];
}

// Position of entry (i, j) in row i for every upper entry (j, i), i > j.
__global__ void k_ic_tpos(int n_eq, const int *rp, const int *col, const int *dpos, int *tpos)
{
    const int j = blockIdx.x * IC_BLK + threadIdx.x;
    if (j >= n_eq) return;
    for (int p = rp[j] + dpos[j] + 1; p < rp[j + 1]; p++) {
        const int i = col[p];
        int lo = rp[i], hi = rp[i] + dpos[i] - 1, at = -1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            const int cm = col[mid];
            if (cm == j) { at = mid; break; }
            if (cm < j) lo = mid + 1; else hi = mid - 1;
        }
        tpos[p] = at;
    }
}

// ---------------------------------------------------------------------
// sync-free factor and triangular solves
// ---------------------------------------------------------------------

// Claims the next 64-row chunk for the calling wave (lane 0 takes the
// ticket, the others read it from the wave).
__device__ __forceinline__ int claim_chunk(unsigned *ticket)
{
    int base = 0;
    if ((threadIdx.x & (WAVE - 1)) == 0) base = (int)atomicAdd(ticket, (unsigned)WAVE);
    return __shfl(base, 0, WAVE);
}

__global__ void __launch_bounds__(IC_BLK)
k_ic_factor(int n, const int *__restrict__ rp, const int *__restrict__ col, const int *__restrict__ dpos,
            const double *__restrict__ aval, double *lval, double *dinv, unsigned *ticket, int *err, int dense)
{
    for (;;) {
        const int base = claim_chunk(ticket);
        if (base >= n) return;
        const int e = base + (threadIdx.x & (WAVE - 1));
        const bool mine = e < n;
        int p = 0, pend = 0, row0 = 0;
        if (mine) {
            row0 = rp[e];
            p = row0;
            pend = row0 + dpos[e];
        }
        bool done = !mine;
        long long spins = 0;
        while (__any(!done)) {
            if (!done) {
                if (p < pend) {
                    const int k = col[p];
                    const double dk = ld_acq(&dinv[k]);
                    if (!is_sent(dk)) {
                        const int kb = rp[k], ke = rp[k] + dpos[k];
                        if (!dense) {
                            double s = aval[p];
                            int a = row0, bq = kb;
                            while (a < p && bq < ke) {
                                const int ja = col[a], jb = col[bq];
                                if (ja < jb) a++;
                                else if (ja > jb) bq++;
                                else { s -= lval[a] * lval[bq]; a++; bq++; }
                            }
                            lval[p] = s * dk;
                        } else {
                            const double aik = aval[p];
                            if (fabs(aik) < 1.0e-30) {
                                lval[p] = 0.0;
                            } else {
                                double s = 0.0;
                                int a = row0, bq = kb;
                                while (a < p && bq < ke) {
                                    const int ja = col[a], jb = col[bq];
                                    if (ja < jb) a++;
                                    else if (ja > jb) bq++;
                                    else { s += lval[a] * lval[bq]; a++; bq++; }
                                }
                                lval[p] = (aik - s) * (1.0 / dk);   // dk holds sqrt(d) here
                            }
                        }
                        p++;
                        spins = 0;
                    } else if (++spins > SPIN_LIMIT) {
                        atomicExch(err, 1);
                        p = pend;   // abandon: finish the row with what is there
                    }
                } else {
                    const double akk = aval[pend];
                    double d;
                    if (!dense) {
                        d = akk;
                        for (int q = row0; q < pend; q++) d -= lval[q] * lval[q];
                    } else {
                        double sq = 0.0;
                        for (int q = row0; q < pend; q++) sq += lval[q] * lval[q];
                        d = akk - sq;
                    }
                    if (d <= 0.0) d = (akk > 0.0) ? akk : 1.0e-30;
                    st_rel(&dinv[e], dense ? sqrt(d) : 1.0 / sqrt(d));
                    done = true;
                }
            }
        }
    }
}

// y = L^-1 r (sparse: (r - sum) * dinv; dense: (r - sum) / Lkk)
__global__ void __launch_bounds__(IC_BLK)
k_ic_forward(int n, const int *__restrict__ rp, const int *__restrict__ col, const int *__restrict__ dpos,
             const double *__restrict__ lval, const double *__restrict__ dinv, const double *__restrict__ rin,
             double *y, unsigned *ticket, int *err, int dense, const IcCtrl *ctrl)
{
    if (ctrl && ctrl->done) return;
    for (;;) {
        const int base = claim_chunk(ticket);
        if (base >= n) return;
        const int e = base + (threadIdx.x & (WAVE - 1));
        const bool mine = e < n;
        int p = 0, pend = 0;
        double s = 0.0;
        if (mine) {
            p = rp[e];
            pend = p + dpos[e];
            s = dense ? 0.0 : rin[e];
        }
        bool done = !mine;
        long long spins = 0;
        while (__any(!done)) {
            if (!done) {
                while (p < pend) {
                    const double yc = ld_rlx(&y[col[p]]);
                    if (is_sent(yc)) break;
                    if (dense) s += lval[p] * yc;
                    else s -= lval[p] * yc;
                    p++;
                    spins = 0;
                }
                if (p >= pend) {
                    st_rlx(&y[e], dense ? (rin[e] - s) / dinv[e] : s * dinv[e]);
                    done = true;
                } else if (++spins > SPIN_LIMIT) {
                    atomicExch(err, 1);
                    st_rlx(&y[e], 0.0);
                    done = true;
                }
            }
        }
    }
}

// z = L^-T y as a gather over the upper half of row j, rows claimed from
// the bottom; sparse: subtract in descending i then * dinv; dense: sum in
// ascending i then (y - sum) / Ljj.
__global__ void __launch_bounds__(IC_BLK)
k_ic_backward(int n, const int *__restrict__ rp, const int *__restrict__ col, const int *__restrict__ dpos,
              const int *__restrict__ tpos, const double *__restrict__ lval, const double *__restrict__ dinv,
              const double *__restrict__ yin, double *z, unsigned *ticket, int *err, int dense,
              const IcCtrl *ctrl)
{
    if (ctrl && ctrl->done) return;
    for (;;) {
        const int claimed = claim_chunk(ticket);
        if (claimed >= n) return;
        const int j = n - 1 - (claimed + (threadIdx.x & (WAVE - 1)));
        const bool mine = j >= 0;
        int p = 0, pstop = 0, step = 0;
        double s = 0.0;
        if (mine) {
            const int up0 = rp[j] + dpos[j] + 1, up1 = rp[j + 1];
            if (!dense) { p = up1 - 1; pstop = up0 - 1; step = -1; s = yin[j]; }
            else { p = up0; pstop = up1; step = 1; s = 0.0; }
        }
        bool done = !mine;
        long long spins = 0;
        while (__any(!done)) {
            if (!done) {
                while (p != pstop) {
                    const double zc = ld_rlx(&z[col[p]]);
                    if (is_sent(zc)) break;
                    const int tp = tpos[p];
                    const double l = tp >= 0 ? lval[tp] : 0.0;   // asymmetric pattern: no (i,j)
                    if (dense) s += l * zc;
                    else s -= l * zc;
                    p += step;
                    spins = 0;
                }
                if (p == pstop) {
                    st_rlx(&z[j], dense ? (yin[j] - s) / dinv[j] : s * dinv[j]);
                    done = true;
                } else if (++spins > SPIN_LIMIT) {
                    atomicExch(err, 1);
                    st_rlx(&z[j], 0.0);
                    done = true;
                }
            }
        }
    }
}

__global__ void k_ic_sentinel(int n, double *a, double *b)
{
    for (int i = blockIdx.x * IC_BLK + threadIdx.x; i < n; i += gridDim.x * IC_BLK) {
        if (a) a[i] = sent();
        if (b) b[i] = sent();
    }
}

// ---------------------------------------------------------------------
// PCG in eq order (pressure_solver.c:261-322); with precond = 0 it is the
// plain CG of :90-131 (z = r)
// ---------------------------------------------------------------------

__device__ __forceinline__ double fin_sum(const double *part, unsigned nparts)
{
    double s = 0.0;
    for (unsigned q = threadIdx.x; q < nparts; q += IC_BLK) s += part[q];
    return block_sum<double, IC_BLK>(s);
}

static inline unsigned vgrid(int n)
{
    unsigned g = (unsigned)((n + IC_BLK * 4 - 1) / (IC_BLK * 4));
    return g < 1 ? 1 : (g > 2048 ? 2048 : g);
}

// CSR y = A x, rows summed in column order (csr_mat_vec, :57-68), + partial x.y
__global__ void __launch_bounds__(IC_BLK)
k_csr_spmv_dot(int n, const int *__restrict__ rp, const int *__restrict__ col, const double *__restrict__ val,
               const double *__restrict__ x, double *__restrict__ y, double *part, const IcCtrl *ctrl)
{
    if (ctrl->done) return;
    const int i = blockIdx.x * IC_BLK + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        double s = 0.0;
        for (int q = rp[i]; q < rp[i + 1]; q++) s += val[q] * x[col[q]];
        y[i] = s;
        v = x[i] * s;
    }
    v = block_sum<double, IC_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(IC_BLK)
k_pcg_fin_pap(const double *part, unsigned np, IcCtrl *c, int t)
{
    if (c->done) return;
    const double pap = fin_sum(part, np);
    if (threadIdx.x == 0) {
        c->pap = pap;
        if (fabs(pap) < 1.0e-30) {
            c->done = 1;
            c->iters = t;
            c->breakdown = 1;
        } else {
            c->alpha = c->rz[t & 1] / pap;
        }
    }
}

// x += alpha p, r -= alpha Ap, partial r.r; z (and y) reset to "unknown"
__global__ void __launch_bounds__(IC_BLK)
k_pcg_axpy(int n, const double *p, const double *Ap, double *x, double *r, double *z, double *y,
           double *part, const IcCtrl *c, int precond)
{
    if (c->done) return;
    const double alpha = c->alpha;
    double v = 0.0;
    for (int i = blockIdx.x * IC_BLK + threadIdx.x; i < n; i += gridDim.x * IC_BLK) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * Ap[i];
        r[i] = ri;
        v += ri * ri;
        if (precond) {
            z[i] = sent();
            y[i] = sent();
        }
    }
    v = block_sum<double, IC_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// r.r of the new residual: the stop tests of the next loop top, and the
// |rz| < 1e-60 break of this iteration (decided before z is computed,
// which only feeds p)
__global__ void __launch_bounds__(IC_BLK)
k_pcg_fin_rr(const double *part, unsigned np, IcCtrl *c, int t, int max_iter)
{
    if (c->done) return;
    const double rr = fin_sum(part, np);
    if (threadIdx.x == 0) {
        c->rr_new = rr;
        c->res = sqrt(rr);
        if (fabs(c->rz[t & 1]) < 1.0e-60) {
            c->done = 1;
            c->iters = t;
        } else if (t + 1 >= max_iter) {
            c->done = 1;
            c->iters = max_iter;
        } else if (sqrt(rr) < c->stop) {
            c->done = 1;
            c->iters = t + 1;
        }
    }
}

__global__ void __launch_bounds__(IC_BLK)
k_pcg_dot(int n, const double *a, const double *b, double *part, const IcCtrl *c)
{
    if (c->done) return;
    double v = 0.0;
    for (int i = blockIdx.x * IC_BLK + threadIdx.x; i < n; i += gridDim.x * IC_BLK) v += a[i] * b[i];
    v = block_sum<double, IC_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(IC_BLK)
k_pcg_fin_rz(const double *part, unsigned np, IcCtrl *c, int t)
{
    if (c->done) return;
    const double rz = fin_sum(part, np);
    if (threadIdx.x == 0) {
        c->beta = rz / c->rz[t & 1];
        c->rz[(t + 1) & 1] = rz;
    }
}

__global__ void __launch_bounds__(IC_BLK)
k_pcg_pupdate(int n, const double *z, double *p, const IcCtrl *c)
{
    if (c->done) return;
    const double beta = c->beta;
    for (int i = blockIdx.x * IC_BLK + threadIdx.x; i < n; i += gridDim.x * IC_BLK) p[i] = z[i] + beta * p[i];
}

__global__ void __launch_bounds__(IC_BLK)
k_pcg_init(int n, const double *b, double *x, double *r, double *part_bb)
{
    double v = 0.0;
    for (int i = blockIdx.x * IC_BLK + threadIdx.x; i < n; i += gridDim.x * IC_BLK) {
        x[i] = 0.0;
        r[i] = b[i];
        v += b[i] * b[i];
    }
    v = block_sum<double, IC_BLK>(v);
    if (threadIdx.x == 0) part_bb[blockIdx.x] = v;
}

__global__ void __launch_bounds__(IC_BLK)
k_pcg_init2(int n, const double *z, double *p, const double *r, double *part_rz)
{
    double v = 0.0;
    for (int i = blockIdx.x * IC_BLK + threadIdx.x; i < n; i += gridDim.x * IC_BLK) {
        p[i] = z[i];
        v += r[i] * z[i];
    }
    v = block_sum<double, IC_BLK>(v);
    if (threadIdx.x == 0) part_rz[blockIdx.x] = v;
}

__global__ void __launch_bounds__(IC_BLK)
k_pcg_init_fin(const double *part_bb, const double *part_rz, unsigned np, IcCtrl *c, double tol, int tol_mode,
               int max_iter)
{
    const double bb = fin_sum(part_bb, np);
    const double rz = fin_sum(part_rz, np);
    if (threadIdx.x == 0) {
        c->rz[0] = rz;
        c->bnorm = sqrt(bb);
        c->stop = tol_mode ? tol * sqrt(bb) : tol;
        c->done = 0;
        c->iters = 0;
        c->breakdown = 0;
        c->res = sqrt(bb);
        if (max_iter <= 0 || sqrt(bb) < c->stop) c->done = 1;
    }
}

__global__ void k_ic_scatter_x(int n_slots, const int *eq_of_slot, const double *xeq, double *xslot)
{
    const int s = blockIdx.x * IC_BLK + threadIdx.x;
    if (s >= n_slots) return;
    const int e = eq_of_slot[s];
    xslot[s] = e >= 0 ? xeq[e] : 0.0;
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------

static int ic_reserve(IcData *d, int n_slots, int n_eq, long long nnz)
{
    if (n_slots + 1 > d->cap_n) {
        int *pi[] = {d->rp, d->dpos, d->eq_of_slot, d->slot_of_eq, d->flag, d->eq_of_orig, d->len};
        for (int *q : pi) if (q) hipFree(q);
        double *pd[] = {d->dinv, d->x, d->r, d->z, d->p, d->Ap, d->y, d->b};
        for (double *q : pd) if (q) hipFree(q);
        const size_t N = (size_t)n_slots + 2;
        HIP_OK(hipMalloc(&d->rp, sizeof(int) * N));
        HIP_OK(hipMalloc(&d->dpos, sizeof(int) * N));
        HIP_OK(hipMalloc(&d->eq_of_slot, sizeof(int) * N));
        HIP_OK(hipMalloc(&d->slot_of_eq, sizeof(int) * N));
        HIP_OK(hipMalloc(&d->flag, sizeof(int) * N));
        HIP_OK(hipMalloc(&d->eq_of_orig, sizeof(int) * N));
        HIP_OK(hipMalloc(&d->len, sizeof(int) * N));
        HIP_OK(hipMalloc(&d->dinv, sizeof(double) * N));
        HIP_OK(hipMalloc(&d->x, sizeof(double) * N));
        HIP_OK(hipMalloc(&d->r, sizeof(double) * N));
        HIP_OK(hipMalloc(&d->z, sizeof(double) * N));
        HIP_OK(hipMalloc(&d->p, sizeof(double) * N));
        HIP_OK(hipMalloc(&d->Ap, sizeof(double) * N));
        HIP_OK(hipMalloc(&d->y, sizeof(double) * N));
        HIP_OK(hipMalloc(&d->b, sizeof(double) * N));
        d->cap_n = n_slots + 1;
    }
    if (nnz > d->cap_nnz) {
        if (d->col) hipFree(d->col);
        if (d->tpos) hipFree(d->tpos);
        if (d->aval) hipFree(d->aval);
        if (d->lval) hipFree(d->lval);
        const size_t Z = (size_t)(nnz > 0 ? nnz : 1);
        HIP_OK(hipMalloc(&d->col, sizeof(int) * Z));
        HIP_OK(hipMalloc(&d->tpos, sizeof(int) * Z));
        HIP_OK(hipMalloc(&d->aval, sizeof(double) * Z));
        HIP_OK(hipMalloc(&d->lval, sizeof(double) * Z));
        d->cap_nnz = nnz;
    }
    if (!d->ticket) {
        HIP_OK(hipMalloc(&d->ticket, sizeof(unsigned) * 4));
        HIP_OK(hipMalloc(&d->err, sizeof(int)));
        HIP_OK(hipMalloc(&d->part, sizeof(double) * 3 * 65536));
    }
    (void)n_eq;
    return 0;
}

void ic_free(IcData *d)
{
    if (!d) return;
    void *ps[] = {d->rp, d->col, d->dpos, d->tpos, d->eq_of_slot, d->slot_of_eq, d->flag, d->eq_of_orig, d->len,
                  d->aval, d->lval, d->dinv, d->x, d->r, d->z, d->p, d->Ap, d->y, d->b, d->ticket, d->err, d->part};
    for (void *q : ps) if (q) hipFree(q);
    delete d;
}

static int reset_ticket(IcData *d, hipStream_t s)
{
    HIP_OK(hipMemsetAsync(d->ticket, 0, sizeof(unsigned) * 4, s));
    return 0;
}

static unsigned persistent_grid()
{
    static unsigned g = 0;
    if (!g) {
        int dev = 0, cus = 256;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        g = (unsigned)(cus > 0 ? cus : 256) * 2;
    }
    return g;
}

static int ic_factor(IcData *d, hipStream_t s, int dense)
{
    if (d->n_eq == 0) return 0;
    hipLaunchKernelGGL(k_ic_sentinel, dim3(vgrid(d->n_eq)), dim3(IC_BLK), 0, s, d->n_eq, d->dinv, (double *)nullptr);
    int rc = reset_ticket(d, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ic_factor, dim3(persistent_grid()), dim3(IC_BLK), 0, s, d->n_eq, d->rp, d->col, d->dpos,
                       d->aval, d->lval, d->dinv, d->ticket, d->err, dense);
    LAUNCH_CHECK();
    return 0;
}

// z = M^-1 r with y as scratch; y and z must hold the sentinel already.
// With ctrl given, both solves return at once once the solver has stopped.
static int ic_apply(IcData *d, const double *r, hipStream_t s, int dense, const IcCtrl *ctrl)
{
    int rc = reset_ticket(d, s);
    if (rc) return rc;
    prof_begin(PK_IC, s);
    hipLaunchKernelGGL(k_ic_forward, dim3(persistent_grid()), dim3(IC_BLK), 0, s, d->n_eq, d->rp, d->col, d->dpos,
                       d->lval, d->dinv, r, d->y, d->ticket + 0, d->err, dense, ctrl);
    hipLaunchKernelGGL(k_ic_backward, dim3(persistent_grid()), dim3(IC_BLK), 0, s, d->n_eq, d->rp, d->col,
                       d->dpos, d->tpos, d->lval, d->dinv, d->y, d->z, d->ticket + 1, d->err, dense, ctrl);
    prof_end(PK_IC, s);
    LAUNCH_CHECK();
    return 0;
}

// PCG (precond 1) or CG (precond 0) on the eq-order CSR held in d.
static int pcg_run(IcData *d, hipStream_t s, IcCtrl *dctrl, IcCtrl *hctrl, const SolveOpts &o, int precond,
                   SolveResult *res)
{
    const int n = d->n_eq;
    const unsigned gv = vgrid(n), gm = nblk(n);
    double *pa = d->part, *pb = d->part + 65536, *pc = d->part + 2 * 65536;
    const int dense = o.dense_ic;
    HIP_OK(hipMemsetAsync(d->err, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_pcg_init, dim3(gv), dim3(IC_BLK), 0, s, n, d->b, d->x, d->r, pa);
    if (precond) {
        hipLaunchKernelGGL(k_ic_sentinel, dim3(gv), dim3(IC_BLK), 0, s, n, d->y, d->z);
        int rc = ic_apply(d, d->r, s, dense, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL(k_pcg_init2, dim3(gv), dim3(IC_BLK), 0, s, n, d->z, d->p, d->r, pb);
    } else {
        hipLaunchKernelGGL(k_pcg_init2, dim3(gv), dim3(IC_BLK), 0, s, n, d->r, d->p, d->r, pb);
    }
    hipLaunchKernelGGL(k_pcg_init_fin, dim3(1), dim3(IC_BLK), 0, s, pa, pb, gv, dctrl, o.tol, o.tol_mode, o.max_iter);
    LAUNCH_CHECK();
    int chunk = o.check_every > 0 ? o.check_every : 16;
    if (precond && chunk > 4) chunk = 4;
    hipEvent_t ev[2];
    HIP_OK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    volatile int *hdone = &hctrl->done;
    int pending = -1, t = 0, k = 0;
    const double *zz = precond ? d->z : d->r;
    while (t < o.max_iter) {
        const int t_end = (t + chunk < o.max_iter) ? t + chunk : o.max_iter;
        for (; t < t_end; t++) {
            prof_begin(PK_SPMV, s);
            hipLaunchKernelGGL(k_csr_spmv_dot, dim3(gm), dim3(IC_BLK), 0, s, n, d->rp, d->col, d->aval, d->p, d->Ap,
                               pa, dctrl);
            prof_end(PK_SPMV, s);
            hipLaunchKernelGGL(k_pcg_fin_pap, dim3(1), dim3(IC_BLK), 0, s, pa, gm, dctrl, t);
            hipLaunchKernelGGL(k_pcg_axpy, dim3(gv), dim3(IC_BLK), 0, s, n, d->p, d->Ap, d->x, d->r, d->z, d->y, pb,
                               dctrl, precond);
            hipLaunchKernelGGL(k_pcg_fin_rr, dim3(1), dim3(IC_BLK), 0, s, pb, gv, dctrl, t, o.max_iter);
            if (precond) {
                // the solves return at once after the stop flag is set
                int rc = ic_apply(d, d->r, s, dense, dctrl);
                if (rc) return rc;
            }
            hipLaunchKernelGGL(k_pcg_dot, dim3(gv), dim3(IC_BLK), 0, s, n, d->r, zz, pc, dctrl);
            hipLaunchKernelGGL(k_pcg_fin_rz, dim3(1), dim3(IC_BLK), 0, s, pc, gv, dctrl, t);
            hipLaunchKernelGGL(k_pcg_pupdate, dim3(gv), dim3(IC_BLK), 0, s, n, zz, d->p, dctrl);
        }
        LAUNCH_CHECK();
        HIP_OK(hipMemcpyAsync((void *)hdone, &dctrl->done, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_OK(hipEventRecord(ev[k & 1], s));
        if (pending >= 0) {
            HIP_OK(hipEventSynchronize(ev[pending & 1]));
            if (*hdone) break;
        }
        pending = k;
        k++;
    }
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipMemcpy(hctrl, dctrl, sizeof(IcCtrl), hipMemcpyDeviceToHost));
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    int herr = 0;
    HIP_OK(hipMemcpy(&herr, d->err, sizeof(int), hipMemcpyDeviceToHost));
    res->iters = hctrl->done ? hctrl->iters : o.max_iter;
    res->breakdown = hctrl->breakdown;
    res->res = hctrl->res;
    res->bnorm = hctrl->bnorm;
    res->outer = 1;
    if (herr) {
        mps_set_error(MPS_ERR_SOLVER, __FILE__, __LINE__, "IC(0) dependency wait exceeded its bound");
        return MPS_ERR_SOLVER;
    }
    return 0;
}

static IcCtrl *g_dctrl = nullptr, *g_hctrl = nullptr;

static int ctrl_ready()
{
    if (!g_dctrl) HIP_OK(hipMalloc(&g_dctrl, sizeof(IcCtrl)));
    if (!g_hctrl) HIP_OK(hipHostMalloc(&g_hctrl, sizeof(IcCtrl), hipHostMallocDefault));
    return 0;
}

// ICCG on the PPE held slot-ordered in c; the solution goes to c->x.
int iccg_solve(Ctx *c, const SolveOpts &o, SolveResult *res)
{
    int rc = ctrl_ready();
    if (rc) return rc;
    if (!c->ic) c->ic = new IcData();
    IcData *d = c->ic;
    const int n = c->n;
    hipStream_t s = c->stream;
    rc = ic_reserve(d, n, 0, 0);
    if (rc) return rc;
    // eq numbering
    hipLaunchKernelGGL(k_ic_flag, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, c->O[c->ct], c->unk, d->flag);
    LAUNCH_CHECK();
    rc = launch_scan_excl(d->flag, d->eq_of_orig, n, c->scan_blk, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ic_maps, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, c->O[c->ct], c->unk, c->Alen, d->eq_of_orig,
                       d->eq_of_slot, d->slot_of_eq, d->len);
    LAUNCH_CHECK();
    int n_eq = 0;
    HIP_OK(hipMemcpyAsync(&n_eq, d->eq_of_orig + n, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    d->n_eq = n_eq;
    if (n_eq == 0) {
        rc = launch_fill_double(c->x, 0.0, (size_t)n, s);
        res->iters = 0; res->breakdown = 0; res->res = 0; res->bnorm = 0; res->outer = 1;
        return rc;
    }
    rc = launch_scan_excl(d->len, d->rp, n_eq, c->scan_blk, s);
    if (rc) return rc;
    int nnz = 0;
    HIP_OK(hipMemcpyAsync(&nnz, d->rp + n_eq, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    rc = ic_reserve(d, n, n_eq, nnz);
    if (rc) return rc;
    d->nnz = nnz;
    hipLaunchKernelGGL(k_ic_fill, dim3(nblk(n_eq)), dim3(IC_BLK), 0, s, n_eq, c->stride, d->slot_of_eq, d->eq_of_slot,
                       c->Alen, c->Aval, c->Acol, c->diag, c->b, d->rp, d->col, d->aval, d->dpos, d->b);
    hipLaunchKernelGGL(k_ic_tpos, dim3(nblk(n_eq)), dim3(IC_BLK), 0, s, n_eq, d->rp, d->col, d->dpos, d->tpos);
    LAUNCH_CHECK();
    HIP_OK(hipMemsetAsync(d->err, 0, sizeof(int), s));
    rc = ic_factor(d, s, o.dense_ic);
    if (rc) return rc;
    rc = pcg_run(d, s, g_dctrl, g_hctrl, o, 1, res);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ic_scatter_x, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, d->eq_of_slot, d->x, c->x);
    LAUNCH_CHECK();
    return 0;
}

// CG (no preconditioner) on a CSR held in device memory, rows with their
// diagonal included in any column order; used by the adjoint's second
// solve on the step's matrix.  x receives the solution.
int csr_cg_device(int n, long long nnz, const int *rp, const int *col, const double *val, const double *b,
                  double *x, const SolveOpts &o, SolveResult *res, hipStream_t s)
{
    static IcData *d = nullptr;
    int rc = ctrl_ready();
    if (rc) return rc;
    if (!d) d = new IcData();
    rc = ic_reserve(d, n, n, nnz);
    if (rc) return rc;
    d->n_eq = n;
    d->nnz = nnz;
    HIP_OK(hipMemcpyAsync(d->rp, rp, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(d->col, col, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(d->aval, val, sizeof(double) * (size_t)nnz, hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(d->b, b, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    SolveOpts oo = o;
    oo.solver = 0;
    rc = pcg_run(d, s, g_dctrl, g_hctrl, oo, 0, res);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(x, d->x, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    return 0;
}

// Stand-alone CSR solve (rows with sorted columns), CG or ICCG.
extern "C" int mps_hip_solve_csr(int n, const int *row_ptr, const int *col, const double *val, const double *b,
                                 double *x, int solver_type, int max_iter, double tol, int *iters_out)
{
    int rc = mps_hip_init(0);
    if (rc) return rc;
    rc = ctrl_ready();
    if (rc) return rc;
    if (n <= 0) {
        if (iters_out) *iters_out = 0;
        return MPS_OK;
    }
    IcData *d = new IcData();
    const long long nnz = row_ptr[n];
    rc = ic_reserve(d, n, n, nnz);
    if (rc) { ic_free(d); return rc; }
    d->n_eq = n;
    d->nnz = nnz;
    std::vector<int> dpos(n, 0);
    for (int i = 0; i < n; i++) {
        int q = row_ptr[i];
        while (q < row_ptr[i + 1] && col[q] < i) q++;
        dpos[i] = q - row_ptr[i];
    }
    hipStream_t s = nullptr;
    HIP_OK(hipMemcpy(d->rp, row_ptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d->col, col, sizeof(int) * nnz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d->aval, val, sizeof(double) * nnz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d->b, b, sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d->dpos, dpos.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_ic_tpos, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, d->rp, d->col, d->dpos, d->tpos);
    SolveOpts o{};
    o.solver = solver_type;
    o.max_iter = max_iter;
    o.tol = tol;
    o.tol_mode = (int)mps_hip_get_option("tol_mode");
    o.check_every = (int)mps_hip_get_option("check_every");
    o.dense_ic = MPS_DIM == 2;
    SolveResult sr{};
    HIP_OK(hipMemsetAsync(d->err, 0, sizeof(int), s));
    if (solver_type == 1) {
        rc = ic_factor(d, s, o.dense_ic);
        if (!rc) rc = pcg_run(d, s, g_dctrl, g_hctrl, o, 1, &sr);
    } else {
        rc = pcg_run(d, s, g_dctrl, g_hctrl, o, 0, &sr);
    }
    if (!rc) {
        hipError_t e = hipMemcpy(x, d->x, sizeof(double) * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = MPS_ERR_HIP;
    }
    if (iters_out) *iters_out = sr.iters;
    ic_free(d);
    return rc;
}

// IC(0) factor of the last ICCG solve in the reference's layout.
extern "C" int mps_hip_export_ic(ParticleSystem *ps, int *row_ptr, int *col, double *val, double *diag_inv)
{
    Ctx *c = ps ? mps_ctx_for(ps, 0) : nullptr;
    if (!c || !c->ic || c->ic->n_eq == 0) return MPS_ERR_ARG;
    IcData *d = c->ic;
    const int n = d->n_eq;
    std::vector<int> rp(n + 1), dp(n), cl((size_t)d->nnz);
    std::vector<double> lv((size_t)d->nnz), di(n);
    HIP_OK(hipMemcpy(rp.data(), d->rp, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(dp.data(), d->dpos, sizeof(int) * n, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(cl.data(), d->col, sizeof(int) * d->nnz, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(lv.data(), d->lval, sizeof(double) * d->nnz, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(di.data(), d->dinv, sizeof(double) * n, hipMemcpyDeviceToHost));
    int at = 0;
    if (row_ptr) row_ptr[0] = 0;
    for (int e = 0; e < n; e++) {
        for (int q = rp[e]; q < rp[e] + dp[e]; q++) {
            if (col) col[at] = cl[q];
            if (val) val[at] = lv[q];
            at++;
        }
        if (row_ptr) row_ptr[e + 1] = at;
        if (diag_inv) diag_inv[e] = di[e];
    }
    return MPS_OK;
}

int ic_levels_unused;
