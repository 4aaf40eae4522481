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
    unsigned int ticket[4];
    // mixed precision (iterative refinement)
    float frr[2];
    float fpap, falpha, fbeta, fstop;
    int fdone, fiters;
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
// are fetched 8 at a time (column indices and values, then the 8 gathers
// of p) so each lane keeps several independent loads in flight, and summed
// in row order.
#define SPMV_UNROLL 8

__global__ void __launch_bounds__(CG_BLK)
k_spmv_dot(int n, int stride, const double *__restrict__ diag, const double *__restrict__ Aval,
           const int *__restrict__ Acol, const int *__restrict__ Alen, const double *__restrict__ p,
           double *__restrict__ Ap, double *part, const CgCtrl *ctrl)
{
    if (ctrl->done) return;
    const int i = blockIdx.x * CG_BLK + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const int m = Alen[i];
        const double pi = p[i];
        double s = diag[i] * pi;
        const size_t base = ((size_t)(i >> 6) * (size_t)stride) * 64u + (size_t)(i & 63);
        for (int q0 = 0; q0 < m; q0 += SPMV_UNROLL) {
            int c[SPMV_UNROLL];
            double a[SPMV_UNROLL], x[SPMV_UNROLL];
#pragma unroll
            for (int u = 0; u < SPMV_UNROLL; u++) {
                const bool ok = q0 + u < m;
                const size_t o = base + (size_t)(q0 + u) * 64u;
                c[u] = ok ? Acol[o] : i;
                a[u] = ok ? Aval[o] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < SPMV_UNROLL; u++) x[u] = p[c[u]];
#pragma unroll
            for (int u = 0; u < SPMV_UNROLL; u++)
                if (q0 + u < m) s += a[u] * x[u];
        }
        Ap[i] = s;
        v = pi * s;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
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
            hipLaunchKernelGGL(k_spmv_dot, dim3(g), dim3(CG_BLK), 0, s, n, c->stride, c->diag, c->Aval,
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

int iccg_solve(Ctx *c, const SolveOpts &o, SolveResult *res);   // k_ic.hip

int solve_ppe(Ctx *c, const SolveOpts &o, SolveResult *res)
{
    int rc = cg_alloc_ctrl(c);
    if (rc) return rc;
    if (o.solver == 1) return iccg_solve(c, o, res);
    return cg_fp64(c, o, res);
}
