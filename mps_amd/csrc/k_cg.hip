// k_cg.hip -- conjugate-gradient solve of the PPE on gfx950.
//
// The recurrences are the reference's (pressure_solver.c:90-131): x0 = 0,
// r = p = b; per iteration: stop if sqrt(rr) < tol; Ap; pAp; stop if
// |pAp| < 1e-30; alpha = rr/pAp; x += alpha p; r -= alpha Ap; rr' = r.r;
// stop if rr < 1e-60; beta = rr'/rr; p = r + beta p.  Only the summation
// order of the dot products differs from the reference's serial loop.
//
// Per iteration:
//   spmv_dot   Ap = A p (one thread per row, sliced-ELL rows), partial p.Ap
//   spmv_fin   one block: p.Ap, alpha (or breakdown)
//   axpy_dot   x += alpha p, r -= alpha Ap, partial r.r
//   axpy_fin   one block: r.r, stop tests, beta
//   pupdate    p = r + beta p
// A dot product is a fixed-tree block sum per block plus a block-order sum
// of the partials, so the results are deterministic run to run.  Scalars and the stop
// flag live in a device control block; the host enqueues chunks of
// iterations and polls the flag one chunk behind, so the GPU never idles
// waiting for the host.
#include "ctx.hpp"

#define CG_BLK 256

struct CgCtrl {
    double rr[2];          // r.r entering iteration t is rr[t & 1]
    double pap, alpha, beta;
    double bnorm, stop, res;
    int done, iters, breakdown, pad;
};

static inline unsigned cg_grid(int n) { return (unsigned)((n + CG_BLK - 1) / CG_BLK); }

#define LAUNCH_CHECK()                                                       \
    do {                                                                     \
        hipError_t e_ = hipGetLastError();                                   \
        if (e_ != hipSuccess) {                                              \
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(e_)); \
            return MPS_ERR_HIP;                                              \
        }                                                                    \
    } while (0)

// One partial per block, summed in block order by a one-block finalize
// kernel: deterministic, and free of the single-address atomic ticket that
// serialises thousands of blocks.
__device__ __forceinline__ double fin_sum(const double *part, unsigned nparts)
{
    double s = 0.0;
    for (unsigned q = threadIdx.x; q < nparts; q += CG_BLK) s += part[q];
    return block_sum<double, CG_BLK>(s);
}

// Grid for the streaming vector kernels (grid-stride, 4 waves per block).
static inline unsigned vec_grid(int n)
{
    unsigned g = (unsigned)((n + CG_BLK * 4 - 1) / (CG_BLK * 4));
    return g < 1 ? 1 : (g > 2048 ? 2048 : g);
}

__global__ void __launch_bounds__(CG_BLK)
k_cg_init(int n, const double *__restrict__ b, double *__restrict__ x, double *__restrict__ r,
          double *__restrict__ p, double *part)
{
    double v = 0.0;
    for (int i = blockIdx.x * CG_BLK + threadIdx.x; i < n; i += gridDim.x * CG_BLK) {
        const double bi = b[i];
        x[i] = 0.0;
        r[i] = bi;
        p[i] = bi;
        v += bi * bi;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(CG_BLK)
k_cg_init_fin(const double *part, unsigned nparts, CgCtrl *ctrl, double tol, int tol_mode, int max_iter)
{
    const double bb = fin_sum(part, nparts);
    if (threadIdx.x == 0) {
        ctrl->rr[0] = bb;
        ctrl->bnorm = sqrt(bb);
        ctrl->stop = tol_mode ? tol * sqrt(bb) : tol;
        ctrl->done = 0;
        ctrl->iters = 0;
        ctrl->breakdown = 0;
        ctrl->res = sqrt(bb);
        if (max_iter <= 0 || sqrt(bb) < ctrl->stop) ctrl->done = 1;
    }
}

// Ap = A p over sliced-ELL rows, one thread per row; the entries of a row
// are fetched U at a time (column indices and values, then the U gathers
// of p) so each lane keeps several independent loads in flight, and summed
// in row order.  NT streams the matrix with non-temporal loads so the
// caches keep the gathered vector; XCD gives each XCD a contiguous row
// range (xcd_block).  Partial sums are indexed by logical block, so every
// variant sums in the same order and gives the same bits.
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p, bool nt)
{
    return nt ? __builtin_nontemporal_load(p) : *p;
}

template <int U, bool NT, bool XCD>
__global__ void __launch_bounds__(CG_BLK)
k_spmv_dot(int n, int stride, const double *__restrict__ diag, const double *__restrict__ Aval,
           const int *__restrict__ Acol, const int *__restrict__ Alen, const double *__restrict__ p,
           double *__restrict__ Ap, double *part, const CgCtrl *ctrl)
{
    if (ctrl->done) return;
    const int bl = XCD ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int i = bl * CG_BLK + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const int m = Alen[i];
        const double pi = p[i];
        double s = diag[i] * pi;
        const size_t base = ((size_t)(i >> 6) * (size_t)stride) * 64u + (size_t)(i & 63);
        for (int q0 = 0; q0 < m; q0 += U) {
            int c[U];
            double a[U], x[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool ok = q0 + u < m;
                const size_t o = base + (size_t)(q0 + u) * 64u;
                c[u] = ok ? ld_stream(Acol + o, NT) : i;
                a[u] = ok ? ld_stream(Aval + o, NT) : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; u++) x[u] = p[c[u]];
#pragma unroll
            for (int u = 0; u < U; u++)
                if (q0 + u < m) s += a[u] * x[u];
        }
        Ap[i] = s;
        v = pi * s;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[bl] = v;
}

typedef void (*spmv_fn)(int, int, const double *, const double *, const int *, const int *, const double *,
                        double *, double *, const CgCtrl *);

static spmv_fn pick_spmv(int variant)
{
    switch (variant) {
    case 1: return k_spmv_dot<16, false, false>;
    case 2: return k_spmv_dot<8, true, false>;
    case 3: return k_spmv_dot<16, true, false>;
    case 4: return k_spmv_dot<4, false, false>;
    case 5: return k_spmv_dot<8, false, true>;
    case 6: return k_spmv_dot<16, false, true>;
    case 7: return k_spmv_dot<16, true, true>;
    default: return k_spmv_dot<8, false, false>;
    }
}

__global__ void __launch_bounds__(CG_BLK)
k_spmv_fin(const double *part, unsigned nparts, CgCtrl *ctrl, int t)
{
    if (ctrl->done) return;
    const double pap = fin_sum(part, nparts);
    if (threadIdx.x == 0) {
        ctrl->pap = pap;
        if (fabs(pap) < 1.0e-30) {
            ctrl->done = 1;
            ctrl->iters = t;
            ctrl->breakdown = 1;
            ctrl->res = sqrt(ctrl->rr[t & 1]);
        } else {
            ctrl->alpha = ctrl->rr[t & 1] / pap;
        }
    }
}

__global__ void __launch_bounds__(CG_BLK)
k_axpy_dot(int n, const double *__restrict__ p, const double *__restrict__ Ap, double *__restrict__ x,
           double *__restrict__ r, double *part, const CgCtrl *ctrl)
{
    if (ctrl->done) return;
    const double alpha = ctrl->alpha;
    double v = 0.0;
    for (int i = blockIdx.x * CG_BLK + threadIdx.x; i < n; i += gridDim.x * CG_BLK) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * Ap[i];
        r[i] = ri;
        v += ri * ri;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(CG_BLK)
k_axpy_fin(const double *part, unsigned nparts, CgCtrl *ctrl, int t, int max_iter)
{
    if (ctrl->done) return;
    const double rr_new = fin_sum(part, nparts);
    if (threadIdx.x == 0) {
        const double rr = ctrl->rr[t & 1];
        ctrl->rr[(t + 1) & 1] = rr_new;
        ctrl->res = sqrt(rr_new);
        if (rr < 1.0e-60) {
            ctrl->done = 1;
            ctrl->iters = t;
        } else {
            ctrl->beta = rr_new / rr;
            if (t + 1 >= max_iter) {
                ctrl->done = 1;
                ctrl->iters = max_iter;
            } else if (sqrt(rr_new) < ctrl->stop) {
                ctrl->done = 1;
                ctrl->iters = t + 1;
            }
        }
    }
}

__global__ void __launch_bounds__(CG_BLK)
k_pupdate(int n, const double *__restrict__ r, double *__restrict__ p, const CgCtrl *ctrl)
{
    if (ctrl->done) return;
    const double beta = ctrl->beta;
    for (int i = blockIdx.x * CG_BLK + threadIdx.x; i < n; i += gridDim.x * CG_BLK)
        p[i] = r[i] + beta * p[i];
}

int cg_alloc_ctrl(Ctx *c)
{
    if (!c->ctrl) {
        HIP_OK(hipMalloc(&c->ctrl, sizeof(CgCtrl)));
        HIP_OK(hipMemset(c->ctrl, 0, sizeof(CgCtrl)));
    }
    if (!c->hctrl) HIP_OK(hipHostMalloc(&c->hctrl, sizeof(CgCtrl), hipHostMallocDefault));
    return 0;
}

// FP64 CG on the slot-ordered system held in c (diag, Aval/Acol/Alen, b).
static int cg_fp64(Ctx *c, const SolveOpts &o, SolveResult *res)
{
    const int n = c->n;
    const unsigned g = cg_grid(n), gv = vec_grid(n);
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_cg_init, dim3(gv), dim3(CG_BLK), 0, s, n, c->b, c->x, c->r, c->pv, c->partB);
    hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(CG_BLK), 0, s, c->partB, gv, c->ctrl, o.tol,
                       o.tol_mode, o.max_iter);
    LAUNCH_CHECK();
    const int chunk = o.check_every > 0 ? o.check_every : 16;
    const spmv_fn spmv = pick_spmv(o.spmv_variant);
    hipEvent_t ev[2];
    HIP_OK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    volatile int *hdone = &c->hctrl->done;
    int pending = -1;     // chunk whose flag copy is in flight
    int t = 0, k = 0;
    while (t < o.max_iter) {
        const int t_end = (t + chunk < o.max_iter) ? t + chunk : o.max_iter;
        for (; t < t_end; t++) {
            prof_begin(PK_SPMV, s);
            hipLaunchKernelGGL(spmv, dim3(g), dim3(CG_BLK), 0, s, n, c->stride, c->diag, c->Aval,
                               c->Acol, c->Alen, c->pv, c->Ap, c->partA, c->ctrl);
            prof_end(PK_SPMV, s);
            hipLaunchKernelGGL(k_spmv_fin, dim3(1), dim3(CG_BLK), 0, s, c->partA, g, c->ctrl, t);
            prof_begin(PK_AXPY, s);
            hipLaunchKernelGGL(k_axpy_dot, dim3(gv), dim3(CG_BLK), 0, s, n, c->pv, c->Ap, c->x, c->r,
                               c->partB, c->ctrl);
            prof_end(PK_AXPY, s);
            hipLaunchKernelGGL(k_axpy_fin, dim3(1), dim3(CG_BLK), 0, s, c->partB, gv, c->ctrl, t,
                               o.max_iter);
            prof_begin(PK_PUPDATE, s);
            hipLaunchKernelGGL(k_pupdate, dim3(gv), dim3(CG_BLK), 0, s, n, c->r, c->pv, c->ctrl);
            prof_end(PK_PUPDATE, s);
        }
        LAUNCH_CHECK();
        // this chunk's stop flag -> pinned host; wait for the previous one
        HIP_OK(hipMemcpyAsync((void *)(hdone + 0), &c->ctrl->done, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_OK(hipEventRecord(ev[k & 1], s));
        if (pending >= 0) {
            HIP_OK(hipEventSynchronize(ev[pending & 1]));
            if (*hdone) break;
        }
        pending = k;
        k++;
    }
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipMemcpy(c->hctrl, c->ctrl, sizeof(CgCtrl), hipMemcpyDeviceToHost));
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    res->iters = c->hctrl->done ? c->hctrl->iters : o.max_iter;
    res->breakdown = c->hctrl->breakdown;
    res->res = c->hctrl->res;
    res->bnorm = c->hctrl->bnorm;
    res->outer = 1;
    return 0;
}

// ---------------------------------------------------------------------
// Mixed precision with reliable updates.  The CG iterates on A', the
// matrix with FP32 off-diagonals in row-sum-preserving form (k_to_f32;
// symmetric, since A_ij and A_ji round alike); dot products and the
// solution x stay FP64.  The
// vectors r, p, Ap are FP64 (precision 1) or FP32 (precision 2, FP32
// SpMV accumulation).  When the recursive residual has fallen by
// `inner_tol` since the last update, or reads as converged, the true
// residual r = b - A x is recomputed with the FP64 matrix and replaces it
// (the search direction is kept).  Convergence is only accepted on that
// FP64 residual, so the answer solves the FP64 system to the tolerance.
// ---------------------------------------------------------------------

struct MixedCtrl {
    double rr, rr_ru, pap, alpha, beta, bnorm, stop, res, delta2;
    int done, iters, breakdown, ru, n_ru, pad;
};

// FP32 copy of the off-diagonals, and each row's Dirichlet part
// d_i = A_ii + sum_j A_ij in FP64 (the coupling to neighbours whose
// pressure is fixed at 0).  The mixed SpMV applies
//     (A' p)_i = d_i p_i + sum_j A'_ij (p_j - p_i),
// which is A with its off-diagonals rounded to FP32 and its diagonal
// adjusted so every row sum is kept: the rounding error E = A' - A has zero
// row sums, so v.E v = -1/2 sum E_ij (v_i - v_j)^2 and every eigenvalue,
// the small ones of the smooth modes included, moves by at most the FP32
// rounding of the entries (2^-24 relative).  Rounding the diagonal on its
// own instead would move the smallest eigenvalues of this Laplacian-like
// matrix by far more and slow CG down.  The differences p_j - p_i of a
// smooth vector are exact or nearly so in either precision.
__global__ void k_to_f32(int n, int stride, const int *__restrict__ Alen, const double *__restrict__ Aval,
                         const double *__restrict__ diag, float *__restrict__ fAval, double *__restrict__ ddir)
{
    const int i = blockIdx.x * CG_BLK + threadIdx.x;
    if (i >= n) return;
    const size_t base = ((size_t)(i >> 6) * (size_t)stride) * 64u + (size_t)(i & 63);
    const int m = Alen[i];
    double d = diag[i];
    for (int q = 0; q < m; q++) {
        const double a = Aval[base + (size_t)q * 64u];
        fAval[base + (size_t)q * 64u] = (float)a;
        d += a;
    }
    ddir[i] = d;
}

template <typename V>
__global__ void __launch_bounds__(CG_BLK)
k_mx_init(int n, const double *__restrict__ b, double *__restrict__ x, V *__restrict__ r, V *__restrict__ p,
          double *part)
{
    double v = 0.0;
    for (int i = blockIdx.x * CG_BLK + threadIdx.x; i < n; i += gridDim.x * CG_BLK) {
        const double bi = b[i];
        x[i] = 0.0;
        const V f = (V)bi;
        r[i] = f;
        p[i] = f;
        v += (double)f * (double)f;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(CG_BLK)
k_mx_init_fin(const double *part, unsigned np, MixedCtrl *c, double tol, int tol_mode, int max_iter,
              double delta, const double *bpart, unsigned nb)
{
    const double rr = fin_sum(part, np);
    const double bb = fin_sum(bpart, nb);
    if (threadIdx.x == 0) {
        c->rr = rr;
        c->rr_ru = rr;
        c->bnorm = sqrt(bb);
        c->stop = tol_mode ? tol * sqrt(bb) : tol;
        c->delta2 = delta * delta;
        c->done = 0;
        c->iters = 0;
        c->breakdown = 0;
        c->ru = 0;
        c->n_ru = 0;
        c->res = sqrt(bb);
        if (max_iter <= 0 || sqrt(bb) < c->stop) c->done = 1;
    }
}

__global__ void __launch_bounds__(CG_BLK)
k_bb(int n, const double *__restrict__ b, double *part)
{
    double v = 0.0;
    for (int i = blockIdx.x * CG_BLK + threadIdx.x; i < n; i += gridDim.x * CG_BLK) v += b[i] * b[i];
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// Ap = A32 p; accumulation in V (FP32 for V = float, FP64 for V = double)
template <int U, typename V>
__global__ void __launch_bounds__(CG_BLK)
k_mx_spmv(int n, int stride, const double *__restrict__ ddir, const float *__restrict__ fAval,
          const int *__restrict__ Acol, const int *__restrict__ Alen, const V *__restrict__ p,
          V *__restrict__ Ap, double *part, const MixedCtrl *c)
{
    if (c->done) return;
    const int bl = xcd_block(blockIdx.x, gridDim.x);
    const int i = bl * CG_BLK + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const int m = Alen[i];
        const V pi = p[i];
        V s = (V)ddir[i] * pi;
        const size_t base = ((size_t)(i >> 6) * (size_t)stride) * 64u + (size_t)(i & 63);
        for (int q0 = 0; q0 < m; q0 += U) {
            int cc[U];
            float a[U];
            V x[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool ok = q0 + u < m;
                const size_t o = base + (size_t)(q0 + u) * 64u;
                cc[u] = ok ? Acol[o] : i;
                a[u] = ok ? fAval[o] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < U; u++) x[u] = p[cc[u]];
#pragma unroll
            for (int u = 0; u < U; u++)
                if (q0 + u < m) s += (V)a[u] * (x[u] - pi);
        }
        Ap[i] = s;
        v = (double)pi * (double)s;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[bl] = v;
}

// alpha, after folding in the true residual norm if the previous iteration
// made a reliable update
__global__ void __launch_bounds__(CG_BLK)
k_mx_fin1(const double *part, unsigned np, const double *part_ru, unsigned np_ru, MixedCtrl *c, int t)
{
    if (c->done) return;
    const double pap = fin_sum(part, np);
    const int ru = c->ru;
    double rr_true = 0.0;
    if (ru) rr_true = fin_sum(part_ru, np_ru);
    if (threadIdx.x == 0) {
        if (ru) {
            c->ru = 0;
            c->n_ru++;
            c->rr = rr_true;
            c->rr_ru = rr_true;
            c->res = sqrt(rr_true);
            if (sqrt(rr_true) < c->stop) {
                c->done = 1;
                c->iters = t;
                return;
            }
        }
        c->pap = pap;
        if (fabs(pap) < 1.0e-30) {
            c->done = 1;
            c->iters = t;
            c->breakdown = 1;
        } else {
            c->alpha = c->rr / pap;
        }
    }
}

template <typename V>
__global__ void __launch_bounds__(CG_BLK)
k_mx_axpy(int n, const V *__restrict__ p, const V *__restrict__ Ap, double *__restrict__ x, V *__restrict__ r,
          double *part, const MixedCtrl *c)
{
    if (c->done) return;
    const double alpha = c->alpha;
    const V av = (V)alpha;
    double v = 0.0;
    for (int i = blockIdx.x * CG_BLK + threadIdx.x; i < n; i += gridDim.x * CG_BLK) {
        x[i] += alpha * (double)p[i];
        const V ri = r[i] - av * Ap[i];
        r[i] = ri;
        v += (double)ri * (double)ri;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(CG_BLK)
k_mx_fin2(const double *part, unsigned np, MixedCtrl *c, int t, int max_iter)
{
    if (c->done) return;
    const double rr_new = fin_sum(part, np);
    if (threadIdx.x == 0) {
        const double rr = c->rr;
        if (rr < 1.0e-60 || t + 1 >= max_iter) {
            c->done = 1;
            c->iters = t + 1;
            c->res = sqrt(rr_new);
            return;
        }
        c->beta = rr_new / rr;
        c->rr = rr_new;
        // reliable update once the residual has fallen by delta, or when it
        // reads as converged (the FP64 residual decides)
        c->ru = (rr_new < c->delta2 * c->rr_ru) || (sqrt(rr_new) < c->stop);
    }
}

// p = r + beta p; on a reliable update first r = b - A x in FP64 for the
// own row (partial |r|^2 for the next fin1)
template <int U, typename V>
__global__ void __launch_bounds__(CG_BLK)
k_mx_pupdate(int n, int stride, const double *__restrict__ b, const double *__restrict__ diag,
             const double *__restrict__ Aval, const int *__restrict__ Acol, const int *__restrict__ Alen,
             const double *__restrict__ x, V *__restrict__ r, V *__restrict__ p, double *part_ru,
             const MixedCtrl *c)
{
    if (c->done) return;
    const V beta = (V)c->beta;
    const int ru = c->ru;
    const int bl = ru ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int i = bl * CG_BLK + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        V ri = r[i];
        if (ru) {
            const int m = Alen[i];
            double s = diag[i] * x[i];
            const size_t base = ((size_t)(i >> 6) * (size_t)stride) * 64u + (size_t)(i & 63);
            for (int q0 = 0; q0 < m; q0 += U) {
                int cc[U];
                double a[U], xv[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const bool ok = q0 + u < m;
                    const size_t o = base + (size_t)(q0 + u) * 64u;
                    cc[u] = ok ? Acol[o] : i;
                    a[u] = ok ? Aval[o] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < U; u++) xv[u] = x[cc[u]];
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (q0 + u < m) s += a[u] * xv[u];
            }
            const double r64 = b[i] - s;
            ri = (V)r64;
            v = (double)ri * (double)ri;
            r[i] = ri;
        }
        p[i] = ri + beta * p[i];
    }
    if (ru) {
        v = block_sum<double, CG_BLK>(v);
        if (threadIdx.x == 0) part_ru[bl] = v;
    }
}

struct MixedBufs {
    MixedCtrl *d = nullptr, *h = nullptr;
    double *part_ru = nullptr;
};

template <typename V>
static int cg_mixed_run(Ctx *c, const SolveOpts &o, SolveResult *res, V *r, V *p, V *Ap)
{
    const int n = c->n;
    const unsigned g = cg_grid(n), gv = vec_grid(n);
    hipStream_t s = c->stream;
    MixedBufs *mb = c->mixed;
    hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(CG_BLK), 0, s, n, c->stride, c->Alen, c->Aval, c->diag, c->fAval,
                       c->ddir);
    hipLaunchKernelGGL(k_mx_init<V>, dim3(gv), dim3(CG_BLK), 0, s, n, c->b, c->x, r, p, c->partB);
    hipLaunchKernelGGL(k_bb, dim3(gv), dim3(CG_BLK), 0, s, n, c->b, c->partA);
    const double delta = o.inner_tol > 0 ? o.inner_tol : 1e-3;
    hipLaunchKernelGGL(k_mx_init_fin, dim3(1), dim3(CG_BLK), 0, s, c->partB, gv, mb->d, o.tol, o.tol_mode,
                       o.max_iter, delta, c->partA, gv);
    LAUNCH_CHECK();
    const int chunk = o.check_every > 0 ? o.check_every : 16;
    hipEvent_t ev[2];
    HIP_OK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    volatile int *hdone = &mb->h->done;
    int pending = -1, t = 0, k = 0;
    while (t < o.max_iter) {
        const int t_end = (t + chunk < o.max_iter) ? t + chunk : o.max_iter;
        for (; t < t_end; t++) {
            prof_begin(PK_SPMV, s);
            hipLaunchKernelGGL((k_mx_spmv<8, V>), dim3(g), dim3(CG_BLK), 0, s, n, c->stride, c->ddir, c->fAval,
                               c->Acol, c->Alen, p, Ap, c->partA, mb->d);
            prof_end(PK_SPMV, s);
            hipLaunchKernelGGL(k_mx_fin1, dim3(1), dim3(CG_BLK), 0, s, c->partA, g, mb->part_ru, g, mb->d, t);
            prof_begin(PK_AXPY, s);
            hipLaunchKernelGGL(k_mx_axpy<V>, dim3(gv), dim3(CG_BLK), 0, s, n, p, Ap, c->x, r, c->partB, mb->d);
            prof_end(PK_AXPY, s);
            hipLaunchKernelGGL(k_mx_fin2, dim3(1), dim3(CG_BLK), 0, s, c->partB, gv, mb->d, t, o.max_iter);
            prof_begin(PK_PUPDATE, s);
            hipLaunchKernelGGL((k_mx_pupdate<8, V>), dim3(g), dim3(CG_BLK), 0, s, n, c->stride, c->b, c->diag,
                               c->Aval, c->Acol, c->Alen, c->x, r, p, mb->part_ru, mb->d);
            prof_end(PK_PUPDATE, s);
        }
        LAUNCH_CHECK();
        HIP_OK(hipMemcpyAsync((void *)hdone, &mb->d->done, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_OK(hipEventRecord(ev[k & 1], s));
        if (pending >= 0) {
            HIP_OK(hipEventSynchronize(ev[pending & 1]));
            if (*hdone) break;
        }
        pending = k;
        k++;
    }
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipMemcpy(mb->h, mb->d, sizeof(MixedCtrl), hipMemcpyDeviceToHost));
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    res->iters = mb->h->done ? mb->h->iters : o.max_iter;
    res->breakdown = mb->h->breakdown;
    res->res = mb->h->res;
    res->bnorm = mb->h->bnorm;
    res->outer = mb->h->n_ru;
    return 0;
}

static int cg_mixed(Ctx *c, const SolveOpts &o, SolveResult *res)
{
    if (!c->mixed) c->mixed = new MixedBufs();
    MixedBufs *mb = c->mixed;
    const size_t NS = ((size_t)(c->cap > 0 ? c->cap : 1) + 63) / 64 * 64;
    if (!c->fAval) {
        HIP_OK(hipMalloc(&c->fAval, sizeof(float) * NS * (size_t)c->stride));
        HIP_OK(hipMalloc(&c->ddir, sizeof(double) * NS));
    }
    if (o.precision == 2 && !c->fr) {
        HIP_OK(hipMalloc(&c->fr, sizeof(float) * NS));
        HIP_OK(hipMalloc(&c->fp, sizeof(float) * NS));
        HIP_OK(hipMalloc(&c->fAp, sizeof(float) * NS));
    }
    if (!mb->d) {
        HIP_OK(hipMalloc(&mb->d, sizeof(MixedCtrl)));
        HIP_OK(hipHostMalloc(&mb->h, sizeof(MixedCtrl), hipHostMallocDefault));
        HIP_OK(hipMalloc(&mb->part_ru, sizeof(double) * 65536));
    }
    if (o.precision == 2) return cg_mixed_run<float>(c, o, res, c->fr, c->fp, c->fAp);
    return cg_mixed_run<double>(c, o, res, c->r, c->pv, c->Ap);
}

void mixed_free(MixedBufs *m)
{
    if (!m) return;
    if (m->d) hipFree(m->d);
    if (m->h) hipHostFree(m->h);
    if (m->part_ru) hipFree(m->part_ru);
    delete m;
}

int iccg_solve(Ctx *c, const SolveOpts &o, SolveResult *res);   // k_ic.hip

int solve_ppe(Ctx *c, const SolveOpts &o, SolveResult *res)
{
    int rc = cg_alloc_ctrl(c);
    if (rc) return rc;
    if (cg_grid(c->n) > 65536u) {   // partial-sum buffers hold 65536 blocks
        mps_set_error(MPS_ERR_ARG, __FILE__, __LINE__, "too many particles for one solve (max 16.7M)");
        return MPS_ERR_ARG;
    }
    if (o.solver == 1) return iccg_solve(c, o, res);
    if (o.precision == 1 || o.precision == 2) return cg_mixed(c, o, res);
    return cg_fp64(c, o, res);
}
