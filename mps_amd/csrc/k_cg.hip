// k_cg.hip -- conjugate-gradient solve of the PPE on gfx950.
//
// The recurrences are the reference's (pressure_solver.c:90-131): x0 = 0,
// r = p = b; per iteration: stop if sqrt(rr) < tol; Ap; pAp; stop if
// |pAp| < 1e-30; alpha = rr/pAp; x += alpha p; r -= alpha Ap; rr' = r.r;
// stop if rr < 1e-60; beta = rr'/rr; p = r + beta p.  Only the summation
// order of the dot products differs from the reference's serial loop.
//
// The system is the packed copy of the PPE (k_pack.hip): one row per
// unknown, 64-row slices, 12-bit halo indices where they fit.
// Per iteration:
//   spmv       Ap = A p (one thread per row, entries U at a time), partial p.Ap
//   spmv_fin   one block: p.Ap, alpha (or breakdown)
//   axpy_dot   x += alpha p, r -= alpha Ap, partial r.r
//   axpy_fin   one block: r.r, stop tests, beta
//   pupdate    p = r + beta p
// A dot product is a fixed-tree block sum per block plus a block-order sum
// of the partials, so the results are deterministic run to run.  Scalars
// and the stop flag live in a device control block; the host enqueues
// chunks of iterations and polls the flag one chunk behind, so the GPU
// never idles waiting for the host.
#include "ctx.hpp"

#define CG_BLK 256
#define VEC_U 4     // rows per thread in flight in the grid-stride vector kernels

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
// kernel: deterministic, and free of a single-address atomic ticket that
// would serialise thousands of blocks.  Each thread issues its loads 8 at a
// time before adding, so the partials arrive in a few L2 round trips.
// The stop flag as the whole block sees it: read once by thread 0 and
// shared, so that every wave of a block takes the same branch even when
// block 0 of the same kernel sets the flag while the block starts.
__device__ __forceinline__ int block_flag(const int *flag)
{
    __shared__ int sh;
    if (threadIdx.x == 0) sh = *(const volatile int *)flag;
    __syncthreads();
    return sh;
}

__device__ __forceinline__ double fin_sum(const double *part, unsigned nparts)
{
    double s = 0.0;
    for (unsigned q0 = 0; q0 < nparts; q0 += 8 * CG_BLK) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const unsigned q = q0 + u * CG_BLK + threadIdx.x;
            v[u] = q < nparts ? part[q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) s += v[u];
    }
    return block_sum<double, CG_BLK>(s);
}

// Grid for the streaming vector kernels (grid-stride, at most 512 blocks
// of 4 waves: enough to stream the vectors, and few partials for the
// blocks of the next kernel to sum).
static inline unsigned vec_grid(int n)
{
    unsigned g = (unsigned)((n + CG_BLK * 4 - 1) / (CG_BLK * 4));
    return g < 1 ? 1 : (g > 512 ? 512 : g);
}

// ---------------------------------------------------------------------
// row sums over the packed matrix
// ---------------------------------------------------------------------

struct PkView {
    const int *len;
    const int4 *smeta;
    const int2 *bmeta;
    const int *halo;
    const double *diag, *val, *ddir;
    const float *fval;
    const unsigned *chw;
    const int *c32;
};

// The view with its buffers as global pointers (gptr), for views read out
// of a per-scene table.
__device__ __forceinline__ PkView pk_global(const PkView &s)
{
    PkView g;
    g.len = gptr(s.len); g.smeta = gptr(s.smeta); g.bmeta = gptr(s.bmeta); g.halo = gptr(s.halo);
    g.diag = gptr(s.diag); g.val = gptr(s.val); g.ddir = gptr(s.ddir); g.fval = gptr(s.fval);
    g.chw = gptr(s.chw); g.c32 = gptr(s.c32);
    return g;
}

static PkView pk_view(const Packed *p)
{
    PkView v;
    v.len = p->len;
    v.smeta = p->smeta;
    v.bmeta = p->bmeta;
    v.halo = p->halo;
    v.diag = p->diag;
    v.val = p->val;
    v.ddir = p->ddir;
    v.fval = p->fval;
    v.chw = p->chw;
    v.c32 = p->c32;
    return v;
}

// Vector accessors: the value at row c of the vector a row is applied to.
template <typename T>
struct XPlain {
    const T *x;
    __device__ __forceinline__ T operator()(int c) const { return x[c]; }
};
// p_t = r + beta p_{t-1}, formed where it is read (fused CG iteration)
struct XFused {
    const double *r, *p;
    double beta;
    __device__ __forceinline__ double operator()(int c) const { return r[c] + beta * p[c]; }
};

// a matrix-stream load; NT marks it non-temporal (read once per SpMV, so
// it should not displace the gathered vector from the caches)
template <bool NT, typename T>
__device__ __forceinline__ T ldm(const T *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <bool DIFF, typename ACC, typename VT, typename XT>
__device__ __forceinline__ ACC madd(ACC s, VT a, XT xj, XT xi)
{
    if (DIFF) return s + (ACC)a * (ACC)(xj - xi);
    return s + (ACC)a * (ACC)xj;
}

// s + sum over the row's m entries, in order, of a_q * x[c_q] (or, with
// DIFF, a_q * (x[c_q] - xi)); 32-bit row numbers, x gathered from global
// memory.  U column and value loads, then U gathers, per round.
template <int U, bool DIFF, typename ACC, bool NT = false, typename VT, typename XT, typename XA>
__device__ __forceinline__ ACC row_sum_global(int m, const VT *__restrict__ vrow, const int *__restrict__ crow,
                                              XA x, XT xi, int self, ACC s)
{
    for (int q0 = 0; q0 < m; q0 += U) {
        int c[U];
        VT a[U];
        XT xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const bool ok = q0 + u < m;
            const size_t o = (size_t)(q0 + u) * 64u;
            c[u] = ok ? ldm<NT>(crow + o) : self;
            a[u] = ok ? ldm<NT>(vrow + o) : VT(0);
        }
#pragma unroll
        for (int u = 0; u < U; u++) xv[u] = x(c[u]);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (q0 + u < m) s = madd<DIFF>(s, a[u], xv[u], xi);
    }
    return s;
}

// 12-bit halo indices: entries 8g..8g+7 of a row fill the 96 bits of three
// consecutive words (entry u at bit 12u, word 0 lowest).  Word j of a group
// holds bits of entries from IDX_FIRST[j] on, so it is read only when that
// entry exists.
#define IDX_GROUP_WORDS 3
static_assert(PK_HALO_CAP <= 4096, "halo indices have 12 bits");
__device__ __forceinline__ int idx_first(int j) { return j == 0 ? 0 : (j == 1 ? 2 : 5); }

// entry u (0-7) of a group whose words are g[0..2]
__device__ __forceinline__ unsigned idx_of(const unsigned *g, int u)
{
    const int bit = 12 * u, wd = bit >> 5, sh = bit & 31;
    unsigned v = g[wd] >> sh;
    if (sh > 20) v |= g[wd + 1] << (32 - sh);
    return v & 0xfffu;
}

// The same sum with 12-bit halo indices: x is read from `hx`, either the
// block's halo staged in LDS or, with INDIRECT, global x at
// halo[hbase + index].  U is a multiple of 8 (whole index groups).
template <int U, bool DIFF, bool INDIRECT, typename ACC, bool NT = false, typename VT, typename XT>
__device__ __forceinline__ ACC row_sum_halo(int m, const VT *__restrict__ vrow, const unsigned *__restrict__ wrow,
                                            const XT *hx, const int *__restrict__ hidx, XT xi, ACC s)
{
    static_assert(U % 8 == 0, "whole groups of eight indices");
    constexpr int NW = U / 8 * IDX_GROUP_WORDS;
    for (int q0 = 0; q0 < m; q0 += U) {
        unsigned w[NW];
        VT a[U];
        XT xv[U];
#pragma unroll
        for (int j = 0; j < NW; j++) {
            const int first = (j / IDX_GROUP_WORDS) * 8 + idx_first(j % IDX_GROUP_WORDS);
            w[j] = (q0 + first < m) ? ldm<NT>(wrow + (size_t)(q0 / 8 * IDX_GROUP_WORDS + j) * 64u) : 0u;
        }
#pragma unroll
        for (int u = 0; u < U; u++) a[u] = (q0 + u < m) ? ldm<NT>(vrow + (size_t)(q0 + u) * 64u) : VT(0);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const unsigned li = (q0 + u < m) ? idx_of(w + (u / 8) * IDX_GROUP_WORDS, u % 8) : 0u;
            xv[u] = INDIRECT ? hx[hidx[li]] : hx[li];
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (q0 + u < m) s = madd<DIFF>(s, a[u], xv[u], xi);
    }
    return s;
}

// Row i of the packed matrix applied to a vector: `vals` is the FP64 or
// FP32 copy of the values; in halo mode the vector comes from `hx` (the
// block's halo in LDS), otherwise from global `x`.
template <int U, bool DIFF, typename ACC, bool NT = false, typename VT, typename XT, typename XA>
__device__ __forceinline__ ACC packed_row(const PkView &P, const VT *vals, int i, XA x, const XT *hx, XT xi, ACC s)
{
    const int4 mt = P.smeta[i >> 6];
    const int lane = i & 63;
    const int m = P.len[i];
    const VT *vrow = vals + (size_t)mt.x * 64u + lane;
    if (mt.z)
        return row_sum_halo<U, DIFF, false, ACC, NT>(m, vrow, P.chw + (size_t)mt.y * 64u + lane, hx, nullptr, xi, s);
    return row_sum_global<U, DIFF, ACC, NT>(m, vrow, P.c32 + (size_t)mt.y * 64u + lane, x, xi, i, s);
}

// Row i applied to a global vector (halo indices resolved through halo[]).
template <int U, typename VT>
__device__ __forceinline__ double packed_row_global(const PkView &P, const VT *vals, int i, const double *x,
                                                    double s)
{
    const int4 mt = P.smeta[i >> 6];
    const int lane = i & 63;
    const int m = P.len[i];
    const VT *vrow = vals + (size_t)mt.x * 64u + lane;
    if (mt.z) {
        const int *hidx = P.halo + P.bmeta[i / PK_ROWS].x;
        return row_sum_halo<U, false, true, double>(m, vrow, P.chw + (size_t)mt.y * 64u + lane, x, hidx, 0.0, s);
    }
    return row_sum_global<U, false, double>(m, vrow, P.c32 + (size_t)mt.y * 64u + lane, XPlain<double>{x}, 0.0,
                                            i, s);
}

// Stages the block's halo of x in LDS (hx); every thread of the block
// takes part and the values are visible to all on return.
template <typename XT, typename XA>
__device__ __forceinline__ void stage_halo_of(const PkView &P, int blk, XA x, XT *hx)
{
    const int2 bm = P.bmeta[blk];
    int h = threadIdx.x;
    for (; h + 3 * CG_BLK < bm.y; h += 4 * CG_BLK) {   // four index loads, then four gathers
        int id[4];
        XT v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) id[u] = P.halo[bm.x + h + u * CG_BLK];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = x(id[u]);
#pragma unroll
        for (int u = 0; u < 4; u++) hx[h + u * CG_BLK] = v[u];
    }
    for (; h < bm.y; h += CG_BLK) hx[h] = x(P.halo[bm.x + h]);
    __syncthreads();
}

template <typename XT, typename XA>
__device__ __forceinline__ void stage_halo(const PkView &P, XA x, XT *hx)
{
    stage_halo_of(P, (int)blockIdx.x, x, hx);
}

// ---------------------------------------------------------------------
// FP64 CG
// ---------------------------------------------------------------------

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

// Ap = A p, partial p.Ap per block.  Each row sums diag*p_i, then its
// entries in the assembly's order (the gathered values are the same
// whether they come from the LDS halo or from global memory).
template <int U, bool NT = false>
__global__ void __launch_bounds__(CG_BLK)
k_spmv_pk(int n, PkView P, const double *__restrict__ p, double *__restrict__ Ap, double *part, const CgCtrl *ctrl)
{
    extern __shared__ unsigned char smem_raw[];
    double *hx = reinterpret_cast<double *>(smem_raw);
    if (ctrl->done) return;
    stage_halo(P, XPlain<double>{p}, hx);
    const int i = blockIdx.x * CG_BLK + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const double pi = p[i];
        const double s = packed_row<U, false, double, NT>(P, P.val, i, XPlain<double>{p}, hx, pi, P.diag[i] * pi);
        Ap[i] = s;
        v = pi * s;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

typedef void (*spmv_fn)(int, PkView, const double *, double *, double *, const CgCtrl *);

// SpMV variants (option "spmv"), the same sums in the same order in all:
// 0 eight entries per load batch, 1 sixteen, both loading the matrix
// stream non-temporally; 2 sixteen, matrix loaded through the caches.
// Non-temporal loads keep the once-per-iteration matrix stream from
// evicting the gathered vector: 12 % off the FP64 SpMV and 16 % off the
// coded one on the 1.08M-particle dam break.
static spmv_fn pick_spmv(int variant)
{
    switch (variant) {
    case 1: return k_spmv_pk<16, true>;
    case 2: return k_spmv_pk<16, false>;
    default: return k_spmv_pk<8, true>;
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

// ---- fused form: two kernels per iteration ---------------------------
// k_cg_iter_a(t): closes iteration t-1 -- every block sums the axpy
// partials in the order k_axpy_fin does and takes the same decisions;
// block 0 records them -- then forms p_t = r + beta p_{t-1} where it is
// read (the block's halo as it is staged in LDS, its own rows, which it
// stores in the other p buffer, and the gathers of 32-bit blocks) and
// computes Ap_t = A p_t with the partial p.Ap.
// k_cg_iter_b(t): every block sums the SpMV partials as k_spmv_fin does
// (breakdown test, alpha; block 0 records), then x += alpha p_t,
// r -= alpha Ap_t, partial r.r.
// Every element is computed with the arithmetic of the five-kernel form,
// so the two forms give the same bits.
template <int U, bool NT>
__global__ void __launch_bounds__(CG_BLK)
k_cg_iter_a(int n, PkView P, int t, int max_iter, const double *__restrict__ r, const double *__restrict__ p_prev,
            double *__restrict__ p_cur, double *__restrict__ Ap, const double *partB, unsigned nb, double *partA,
            CgCtrl *ctrl)
{
    extern __shared__ unsigned char smem_raw[];
    double *hx = reinterpret_cast<double *>(smem_raw);
    __shared__ double sh_beta;
    __shared__ int sh_stop;
    if (block_flag(&ctrl->done)) return;
    double beta = 0.0;
    if (t > 0) {
        const double rr_new = fin_sum(partB, nb);
        if (threadIdx.x == 0) {
            const double rr = ctrl->rr[(t - 1) & 1];
            int stop = 0, iters = 0;
            double bt = 0.0;
            if (rr < 1.0e-60) {
                stop = 1;
                iters = t - 1;
            } else {
                bt = rr_new / rr;
                if (t >= max_iter) {
                    stop = 1;
                    iters = max_iter;
                } else if (sqrt(rr_new) < ctrl->stop) {
                    stop = 1;
                    iters = t;
                }
            }
            if (blockIdx.x == 0) {
                ctrl->rr[t & 1] = rr_new;
                ctrl->res = sqrt(rr_new);
                ctrl->beta = bt;
                if (stop) {
                    ctrl->iters = iters;
                    ctrl->done = 1;
                }
            }
            sh_beta = bt;
            sh_stop = stop;
        }
        __syncthreads();
        if (sh_stop) return;
        beta = sh_beta;
    }
    const XFused pt{r, p_prev, beta};
    const XPlain<double> p0{p_prev};
    if (t > 0)
        stage_halo(P, pt, hx);
    else
        stage_halo(P, p0, hx);
    const int i = blockIdx.x * CG_BLK + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const double pi = t > 0 ? pt(i) : p0(i);
        p_cur[i] = pi;
        const double s = t > 0 ? packed_row<U, false, double, NT>(P, P.val, i, pt, hx, pi, P.diag[i] * pi)
                               : packed_row<U, false, double, NT>(P, P.val, i, p0, hx, pi, P.diag[i] * pi);
        Ap[i] = s;
        v = pi * s;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) partA[blockIdx.x] = v;
}

__global__ void __launch_bounds__(CG_BLK)
k_cg_iter_b(int n, int t, const double *__restrict__ p, const double *__restrict__ Ap, double *__restrict__ x,
            double *__restrict__ r, const double *partA, unsigned na, double *partB, CgCtrl *ctrl)
{
    __shared__ double sh_alpha;
    __shared__ int sh_stop;
    if (block_flag(&ctrl->done)) return;
    // the first VEC_U rows' operands, loaded before the partials are summed
    // so that the two latencies overlap
    const int st = gridDim.x * CG_BLK;
    const int i0 = blockIdx.x * CG_BLK + threadIdx.x;
    const bool pre = i0 + (VEC_U - 1) * st < n;
    double pv0[VEC_U], av0[VEC_U], xv0[VEC_U], rv0[VEC_U];
    if (pre) {
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            pv0[u] = p[i0 + u * st];
            av0[u] = Ap[i0 + u * st];
            xv0[u] = x[i0 + u * st];
            rv0[u] = r[i0 + u * st];
        }
    }
    const double pap = fin_sum(partA, na);
    if (threadIdx.x == 0) {
        const double rr = ctrl->rr[t & 1];
        const int stop = fabs(pap) < 1.0e-30;
        const double alpha = stop ? 0.0 : rr / pap;
        if (blockIdx.x == 0) {
            ctrl->pap = pap;
            if (stop) {
                ctrl->iters = t;
                ctrl->breakdown = 1;
                ctrl->res = sqrt(rr);
                ctrl->done = 1;
            } else {
                ctrl->alpha = alpha;
            }
        }
        sh_alpha = alpha;
        sh_stop = stop;
    }
    __syncthreads();
    if (sh_stop) return;
    const double alpha = sh_alpha;
    double v = 0.0;
    int i = i0;
    if (pre) {
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            x[i + u * st] = xv0[u] + alpha * pv0[u];
            const double ri = rv0[u] - alpha * av0[u];
            r[i + u * st] = ri;
            v += ri * ri;
        }
        i += VEC_U * st;
    }
    for (; i + (VEC_U - 1) * st < n; i += VEC_U * st) {   // VEC_U rows in flight, same order
        double pv[VEC_U], av[VEC_U], xv[VEC_U], rv[VEC_U];
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            pv[u] = p[i + u * st];
            av[u] = Ap[i + u * st];
            xv[u] = x[i + u * st];
            rv[u] = r[i + u * st];
        }
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            x[i + u * st] = xv[u] + alpha * pv[u];
            const double ri = rv[u] - alpha * av[u];
            r[i + u * st] = ri;
            v += ri * ri;
        }
    }
    for (; i < n; i += st) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * Ap[i];
        r[i] = ri;
        v += ri * ri;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) partB[blockIdx.x] = v;
}

// ---- three-kernel form (option fused = 2) ----------------------------
// The SpMV of the five-kernel form; then k_cg_iter_b (every block sums the
// SpMV partials as k_spmv_fin does, then x, r and r.r); then k_cg3_pupdate:
// every block sums the r.r partials as k_axpy_fin does, takes its stop
// tests (block 0 records them), and forms p = r + beta p.  Only the
// 512-block vector kernels repeat a reduction, over 3,366 and 512 partials;
// the arithmetic is the five-kernel form's, so the bits are too.
__global__ void __launch_bounds__(CG_BLK)
k_cg3_pupdate(int n, int t, int max_iter, const double *__restrict__ r, double *__restrict__ p, const double *partB,
              unsigned nb, CgCtrl *ctrl)
{
    __shared__ double sh_beta;
    __shared__ int sh_stop;
    if (block_flag(&ctrl->done)) return;
    // the first VEC_U rows, loaded before the partials are summed
    const int st = gridDim.x * CG_BLK;
    const int i0 = blockIdx.x * CG_BLK + threadIdx.x;
    const bool pre = i0 + (VEC_U - 1) * st < n;
    double rv0[VEC_U], pv0[VEC_U];
    if (pre) {
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            rv0[u] = r[i0 + u * st];
            pv0[u] = p[i0 + u * st];
        }
    }
    const double rr_new = fin_sum(partB, nb);
    if (threadIdx.x == 0) {
        const double rr = ctrl->rr[t & 1];
        int stop = 0, iters = 0;
        double bt = 0.0;
        if (rr < 1.0e-60) {
            stop = 1;
            iters = t;
        } else {
            bt = rr_new / rr;
            if (t + 1 >= max_iter) {
                stop = 1;
                iters = max_iter;
            } else if (sqrt(rr_new) < ctrl->stop) {
                stop = 1;
                iters = t + 1;
            }
        }
        if (blockIdx.x == 0) {
            ctrl->rr[(t + 1) & 1] = rr_new;
            ctrl->res = sqrt(rr_new);
            ctrl->beta = bt;
            if (stop) {
                ctrl->iters = iters;
                ctrl->done = 1;
            }
        }
        sh_beta = bt;
        sh_stop = stop;
    }
    __syncthreads();
    if (sh_stop) return;
    const double beta = sh_beta;
    int i = i0;
    if (pre) {
#pragma unroll
        for (int u = 0; u < VEC_U; u++) p[i + u * st] = rv0[u] + beta * pv0[u];
        i += VEC_U * st;
    }
    for (; i + (VEC_U - 1) * st < n; i += VEC_U * st) {
        double rv[VEC_U], pv[VEC_U];
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            rv[u] = r[i + u * st];
            pv[u] = p[i + u * st];
        }
#pragma unroll
        for (int u = 0; u < VEC_U; u++) p[i + u * st] = rv[u] + beta * pv[u];
    }
    for (; i < n; i += st) p[i] = r[i] + beta * p[i];
}

int cg_alloc_ctrl(Ctx *c)
{
    if (!c->ctrl) {
        HIP_OK(hipMalloc(&c->ctrl, sizeof(CgCtrl)));
        HIP_OK(hipMemsetAsync(c->ctrl, 0, sizeof(CgCtrl), c->stream));
    }
    if (!c->hctrl) HIP_OK(hipHostMalloc(&c->hctrl, sizeof(CgCtrl), hipHostMallocDefault));
    return 0;
}

// Enqueues `chunk` iterations at a time; after each chunk the stop flag is
// copied to pinned memory and the previous chunk's copy is waited for, so
// at most one chunk runs past convergence (its kernels exit at once).
template <typename LaunchIter>
static int run_chunks(hipStream_t s, int max_iter, int check_every, const int *dflag, volatile int *hflag,
                      LaunchIter launch_iter)
{
    const int chunk = check_every > 0 ? check_every : 16;
    hipEvent_t ev[2];
    HIP_OK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int pending = -1;     // chunk whose flag copy is in flight
    int t = 0, k = 0, rc = 0;
    while (t < max_iter) {
        const int t_end = (t + chunk < max_iter) ? t + chunk : max_iter;
        for (; t < t_end; t++) launch_iter(t);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync((void *)hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipEventRecord(ev[k & 1], s);
        if (e == hipSuccess && pending >= 0) {
            e = hipEventSynchronize(ev[pending & 1]);
            if (e == hipSuccess && *hflag) break;
        }
        if (e != hipSuccess) {
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(e));
            rc = MPS_ERR_HIP;
            break;
        }
        pending = k;
        k++;
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    return rc;
}

// Mean launch time (ms) of FP64 SpMV variant `variant` on c's packed
// system, over `reps` launches after two warm-up launches.  The control
// block is re-armed first (the kernel skips its work once a solve has
// stopped); the next solve initialises it again.  Ap and the partials are
// scratch between solves.
int cg_time_spmv(Ctx *c, int variant, int reps, double *ms)
{
    const Packed *pk = c->pk;
    if (!pk || pk->n_eq == 0 || !c->ctrl || reps < 1) return MPS_ERR_ARG;
    hipStream_t s = c->stream;
    const PkView P = pk_view(pk);
    const int n = pk->n_eq;
    const unsigned g = cg_grid(n);
    const size_t lds = sizeof(double) * (size_t)pk->hmax;
    const spmv_fn f = pick_spmv(variant);
    HIP_OK(hipMemsetAsync(&c->ctrl->done, 0, sizeof(int), s));
    for (int q = 0; q < 2; q++) hipLaunchKernelGGL(f, dim3(g), dim3(CG_BLK), lds, s, n, P, c->pv, c->Ap, c->partA, c->ctrl);
    LAUNCH_CHECK();
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, s));
    for (int q = 0; q < reps; q++) hipLaunchKernelGGL(f, dim3(g), dim3(CG_BLK), lds, s, n, P, c->pv, c->Ap, c->partA, c->ctrl);
    HIP_OK(hipEventRecord(e1, s));
    HIP_OK(hipEventSynchronize(e1));
    float t = 0.0f;
    HIP_OK(hipEventElapsedTime(&t, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms = (double)t / reps;
    return 0;
}

// FP64 CG on the packed system; the solution is left in the packed x.
static int cg_fp64(Ctx *c, const SolveOpts &o, SolveResult *res)
{
    const Packed *pk = c->pk;
    const int n = pk->n_eq;
    const unsigned g = cg_grid(n), gv = vec_grid(n);
    hipStream_t s = c->stream;
    const PkView P = pk_view(pk);
    double *x = pk->x, *r = c->r, *p = c->pv, *Ap = c->Ap;
    hipLaunchKernelGGL(k_cg_init, dim3(gv), dim3(CG_BLK), 0, s, n, pk->b, x, r, p, c->partB);
    hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(CG_BLK), 0, s, c->partB, gv, c->ctrl, o.tol,
                       o.tol_mode, o.max_iter);
    LAUNCH_CHECK();
    const size_t lds = sizeof(double) * (size_t)pk->hmax;
    int rc = 0;
    if (o.fused == 2) {
        const spmv_fn spmv = pick_spmv(o.spmv_variant);
        rc = run_chunks(s, o.max_iter, o.check_every, &c->ctrl->done, &c->hctrl->done, [&](int t) {
            prof_begin(PK_SPMV, s);
            hipLaunchKernelGGL(spmv, dim3(g), dim3(CG_BLK), lds, s, n, P, p, Ap, c->partA, c->ctrl);
            prof_end(PK_SPMV, s);
            prof_begin(PK_AXPY, s);
            hipLaunchKernelGGL(k_cg_iter_b, dim3(gv), dim3(CG_BLK), 0, s, n, t, p, Ap, x, r, c->partA, g, c->partB,
                               c->ctrl);
            prof_end(PK_AXPY, s);
            prof_begin(PK_PUPDATE, s);
            hipLaunchKernelGGL(k_cg3_pupdate, dim3(gv), dim3(CG_BLK), 0, s, n, t, o.max_iter, r, p, c->partB, gv,
                               c->ctrl);
            prof_end(PK_PUPDATE, s);
        });
        if (rc) return rc;
    } else if (o.fused) {
        // p ping-pongs: iteration t reads pb[t & 1] (p_{t-1}; b at t = 0)
        // and writes p_t to pb[(t + 1) & 1]
        if (!c->pv2) HIP_OK(hipMalloc(&c->pv2, sizeof(double) * (((size_t)c->cap + 63) / 64 * 64)));
        double *pb[2] = {p, c->pv2};
        auto iter_a = o.spmv_variant == 1 ? k_cg_iter_a<16, true>
                                          : (o.spmv_variant == 2 ? k_cg_iter_a<16, false> : k_cg_iter_a<8, true>);
        rc = run_chunks(s, o.max_iter, o.check_every, &c->ctrl->done, &c->hctrl->done, [&](int t) {
            double *prev = pb[t & 1], *cur = pb[(t + 1) & 1];
            prof_begin(PK_SPMV, s);
            hipLaunchKernelGGL(iter_a, dim3(g), dim3(CG_BLK), lds, s, n, P, t, o.max_iter, r, prev, cur, Ap,
                               c->partB, gv, c->partA, c->ctrl);
            prof_end(PK_SPMV, s);
            prof_begin(PK_AXPY, s);
            hipLaunchKernelGGL(k_cg_iter_b, dim3(gv), dim3(CG_BLK), 0, s, n, t, cur, Ap, x, r, c->partA, g,
                               c->partB, c->ctrl);
            prof_end(PK_AXPY, s);
        });
        if (rc) return rc;
        // the last iteration's stop tests when the loop ran to max_iter
        hipLaunchKernelGGL(k_axpy_fin, dim3(1), dim3(CG_BLK), 0, s, c->partB, gv, c->ctrl, o.max_iter - 1,
                           o.max_iter);
        LAUNCH_CHECK();
    } else {
    const spmv_fn spmv = pick_spmv(o.spmv_variant);
    rc = run_chunks(s, o.max_iter, o.check_every, &c->ctrl->done, &c->hctrl->done, [&](int t) {
        prof_begin(PK_SPMV, s);
        hipLaunchKernelGGL(spmv, dim3(g), dim3(CG_BLK), lds, s, n, P, p, Ap, c->partA, c->ctrl);
        prof_end(PK_SPMV, s);
        hipLaunchKernelGGL(k_spmv_fin, dim3(1), dim3(CG_BLK), 0, s, c->partA, g, c->ctrl, t);
        prof_begin(PK_AXPY, s);
        hipLaunchKernelGGL(k_axpy_dot, dim3(gv), dim3(CG_BLK), 0, s, n, p, Ap, x, r, c->partB, c->ctrl);
        prof_end(PK_AXPY, s);
        hipLaunchKernelGGL(k_axpy_fin, dim3(1), dim3(CG_BLK), 0, s, c->partB, gv, c->ctrl, t, o.max_iter);
        prof_begin(PK_PUPDATE, s);
        hipLaunchKernelGGL(k_pupdate, dim3(gv), dim3(CG_BLK), 0, s, n, r, p, c->ctrl);
        prof_end(PK_PUPDATE, s);
    });
    if (rc) return rc;
    }
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipMemcpy(c->hctrl, c->ctrl, sizeof(CgCtrl), hipMemcpyDeviceToHost));
    res->iters = c->hctrl->done ? c->hctrl->iters : o.max_iter;
    res->breakdown = c->hctrl->breakdown;
    res->res = c->hctrl->res;
    res->bnorm = c->hctrl->bnorm;
    res->outer = 1;
    return 0;
}

// ---------------------------------------------------------------------
// Mixed precision with reliable updates.  The CG iterates on A', the PPE
// with its off-diagonals rounded to FP32 and applied in row-sum-preserving
// form (below); dot products and the solution x stay FP64.  The vectors
// r, p, Ap are FP64 (precision 1) or FP32 (precision 2, FP32 SpMV sums).
// When the recursive residual has fallen by `inner_tol` since the last
// update, or reads as converged, the true residual r = b - A x is
// recomputed with the FP64 matrix and replaces it (the search direction is
// kept).  Convergence is only accepted on that FP64 residual, so the
// answer solves the FP64 system to the tolerance.
//
// Row-sum-preserving form: with d_i = A_ii + sum_j A_ij (the coupling to
// neighbours whose pressure is fixed at 0, kept in FP64),
//     (A' p)_i = d_i p_i + sum_j fl32(A_ij) (p_j - p_i).
// The rounding error E = A' - A then has zero row sums, so
// v.E v = -1/2 sum E_ij (v_i - v_j)^2 and every eigenvalue -- the small ones
// of the smooth modes included -- moves by at most the FP32 rounding of the
// entries (2^-24 relative).  Rounding the diagonal on its own instead moves
// the smallest eigenvalues of this Laplacian-like matrix by far more and
// slows CG down (measured: 786 against 568 iterations per step on the
// 1.08M-particle dam break).  The differences p_j - p_i of a smooth vector
// are exact or nearly so in either precision.
// ---------------------------------------------------------------------

struct MixedCtrl {
    double rr, rr_ru, pap, alpha, beta, bnorm, stop, res, delta2;
    int done, iters, breakdown, ru, n_ru, pad;
    // three-kernel form: r.r, its value at the last reliable update and the
    // reliable-update flag entering iteration t sit in slot t & 1 (written
    // by the previous p update, read by every block of the next axpy);
    // m_rr, m_rr_ru are the axpy's values after folding in the true
    // residual, read by every block of the p update
    double f_rr[2], f_rr_ru[2], m_rr, m_rr_ru;
    int f_ru[2];
};

template <typename V>
__global__ void __launch_bounds__(CG_BLK)
k_mx_init(int n, const double *__restrict__ b, double *__restrict__ x, V *__restrict__ r, V *__restrict__ p,
          double *part, double *bpart)
{
    double v = 0.0, w = 0.0;
    for (int i = blockIdx.x * CG_BLK + threadIdx.x; i < n; i += gridDim.x * CG_BLK) {
        const double bi = b[i];
        x[i] = 0.0;
        const V f = (V)bi;
        r[i] = f;
        p[i] = f;
        v += (double)f * (double)f;
        w += bi * bi;
    }
    v = block_sum<double, CG_BLK>(v);
    w = block_sum<double, CG_BLK>(w);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = v;
        bpart[blockIdx.x] = w;
    }
}

__global__ void __launch_bounds__(CG_BLK)
k_mx_init_fin(const double *part, const double *bpart, unsigned np, MixedCtrl *c, double tol, int tol_mode,
              int max_iter, double delta)
{
    const double rr = fin_sum(part, np);
    const double bb = fin_sum(bpart, np);
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
        c->f_rr[0] = rr;
        c->f_rr_ru[0] = rr;
        c->f_ru[0] = 0;
        if (max_iter <= 0 || sqrt(bb) < c->stop) c->done = 1;
    }
}

// Ap = A' p; sums in V (FP32 for V = float, FP64 for V = double)
template <int U, typename V, bool NT>
__global__ void __launch_bounds__(CG_BLK)
k_mx_spmv_pk(int n, PkView P, const V *__restrict__ p, V *__restrict__ Ap, double *part, const MixedCtrl *c)
{
    extern __shared__ unsigned char smem_raw[];
    V *hx = reinterpret_cast<V *>(smem_raw);
    if (c->done) return;
    stage_halo(P, XPlain<V>{p}, hx);
    const int i = blockIdx.x * CG_BLK + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const V pi = p[i];
        const V s = packed_row<U, true, V, NT>(P, P.fval, i, XPlain<V>{p}, hx, pi, (V)P.ddir[i] * pi);
        Ap[i] = s;
        v = (double)pi * (double)s;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// ---- precision 3: off-diagonals from 20-bit distance codes ------------
// Every off-diagonal is A_ij = neg_coeff w(r_ij), w(r) = re / max(r, r_min)
// - 1 for r < re (pressure_solver.c:490-531, kernel.c:12-20).  In halo-mode
// blocks one 32-bit word per entry carries the entry's halo index (12 bits:
// a staged halo holds at most PK_HALO_CAP = 4096 rows) and a 20-bit code of
// its distance, r = r_min + code * dr with dr = (re - r_min) / (2^20 - 1),
// recovered from the FP64 value when the system is packed
// (r = re / (A_ij / neg_coeff + 1)).  The SpMV applies
// fl32(neg_coeff) (re / r' - 1) in the row-sum-preserving form.  Rounding r
// to half a code step (1.5e-6 l0 at re = 3.1 l0) moves w by about 2e-6
// relative at r = l0 and 3e-5 at r = 0.98 re, by under 1e-6 absolute
// nearer re, and the operator error keeps zero row sums.  The SpMV streams
// 4 B per nonzero instead of 5.5 (FP32 value and 12-bit index).  Blocks
// stored with 32-bit columns apply their FP32 values.
#define Q_CODE_MAX 1048575.0
struct QConsts {
    float neg_coeff, re, r_lo, dr;
};

// the coded words of row k (halo-mode slices only)
__device__ __forceinline__ void q16_build_row(int k, const PkView &P, double neg_coeff, double re, double r_min,
                                              unsigned *__restrict__ cq)
{
    const int4 mt = P.smeta[k >> 6];
    if (!mt.z) return;
    const int lane = k & 63;
    const int m = P.len[k];
    const size_t vb = (size_t)mt.x * 64u + lane;
    const unsigned *wrow = P.chw + (size_t)mt.y * 64u + lane;
    const double dr = (re - r_min) / Q_CODE_MAX;
    // a group of eight entries at a time: values and index words first
    for (int q0 = 0; q0 < m; q0 += 8) {
        double a[8];
        unsigned g[IDX_GROUP_WORDS];
#pragma unroll
        for (int u = 0; u < 8; u++) a[u] = (q0 + u < m) ? P.val[vb + (size_t)(q0 + u) * 64u] : 0.0;
#pragma unroll
        for (int j = 0; j < IDX_GROUP_WORDS; j++)
            g[j] = (q0 + idx_first(j) < m) ? wrow[(size_t)(q0 / 8 * IDX_GROUP_WORDS + j) * 64u] : 0u;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (q0 + u >= m) break;
            const unsigned idx = idx_of(g, u);
            const double w = a[u] / neg_coeff;
            double r = re / (w + 1.0);
            if (r < r_min) r = r_min;
            double code = rint((r - r_min) / dr);
            code = code < 0.0 ? 0.0 : (code > Q_CODE_MAX ? Q_CODE_MAX : code);
            cq[vb + (size_t)(q0 + u) * 64u] = ((unsigned)code << 12) | (idx & 0xfffu);
        }
    }
}

__global__ void __launch_bounds__(CG_BLK)
k_mx_q16_build(int n_eq, PkView P, double neg_coeff, double re, double r_min, unsigned *__restrict__ cq)
{
    const int k = blockIdx.x * CG_BLK + threadIdx.x;
    if (k < n_eq) q16_build_row(k, P, neg_coeff, re, r_min, cq);
}

// What a row of the coded operator needs before its block's halo is
// staged: its slice descriptor, length, p_i, Dirichlet part and first U
// words.  Loading them ahead of the staging overlaps the two latencies.
template <int U>
struct CodedRow {
    int4 mt;
    int m;
    float pi, d;
    unsigned w[U];
};

template <int U, bool NT>
__device__ __forceinline__ void coded_pre(const PkView &P, const unsigned *__restrict__ cq,
                                          const float *__restrict__ p, int i, CodedRow<U> &c)
{
    c.mt = P.smeta[i >> 6];
    c.m = P.len[i];
    c.pi = p[i];
    c.d = (float)P.ddir[i];
    const unsigned *qrow = cq + (size_t)c.mt.x * 64u + (i & 63);
#pragma unroll
    for (int u = 0; u < U; u++) c.w[u] = (c.mt.z && u < c.m) ? ldm<NT>(qrow + (size_t)u * 64u) : 0u;
}

// (A' p)_i of the coded operator, the halo of p staged in hx; the next
// batch of words is loaded before the current one is gathered and summed
template <int U, bool NT = false>
__device__ __forceinline__ float coded_row_from(const PkView &P, const unsigned *__restrict__ cq, const QConsts &qk,
                                                const float *__restrict__ p, const float *hx, int i,
                                                const CodedRow<U> &c)
{
    const float pi = c.pi;
    float s = c.d * pi;
    const int m = c.m;
    if (c.mt.z) {
        const unsigned *qrow = cq + (size_t)c.mt.x * 64u + (i & 63);
        for (int q0 = 0; q0 < m; q0 += U) {
            unsigned w[U];
            if (q0 == 0) {
#pragma unroll
                for (int u = 0; u < U; u++) w[u] = c.w[u];
            } else {
#pragma unroll
                for (int u = 0; u < U; u++) w[u] = (q0 + u < m) ? ldm<NT>(qrow + (size_t)(q0 + u) * 64u) : 0u;
            }
            float h[U];
#pragma unroll
            for (int u = 0; u < U; u++) h[u] = hx[w[u] & 0xfffu];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const float r = qk.r_lo + (float)(w[u] >> 12) * qk.dr;
                const float a = qk.neg_coeff * (qk.re * __builtin_amdgcn_rcpf(r) - 1.0f);
                if (q0 + u < m) s += a * (h[u] - pi);
            }
        }
    } else {
        s = row_sum_global<U, true, float, NT>(m, P.fval + (size_t)c.mt.x * 64u + (i & 63),
                                               P.c32 + (size_t)c.mt.y * 64u + (i & 63), XPlain<float>{p}, pi, i, s);
    }
    return s;
}

template <int U, bool NT = false>
__device__ __forceinline__ float coded_row(const PkView &P, const unsigned *__restrict__ cq, const QConsts &qk,
                                           const float *__restrict__ p, const float *hx, int i, float pi)
{
    CodedRow<U> c;
    coded_pre<U, NT>(P, cq, p, i, c);
    c.pi = pi;
    return coded_row_from<U, NT>(P, cq, qk, p, hx, i, c);
}

template <int U, bool NT = false>
__global__ void __launch_bounds__(CG_BLK)
k_mx_spmv_q(int n, PkView P, const unsigned *__restrict__ cq, QConsts qk, const float *__restrict__ p,
            float *__restrict__ Ap, double *part, const MixedCtrl *c)
{
    extern __shared__ unsigned char smem_raw[];
    float *hx = reinterpret_cast<float *>(smem_raw);
    if (c->done) return;
    const int i = blockIdx.x * CG_BLK + threadIdx.x;
    CodedRow<U> row;
    if (i < n) coded_pre<U, NT>(P, cq, p, i, row);
    stage_halo(P, XPlain<float>{p}, hx);
    double v = 0.0;
    if (i < n) {
        const float s = coded_row_from<U, NT>(P, cq, qk, p, hx, i, row);
        Ap[i] = s;
        v = (double)row.pi * (double)s;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// alpha, after folding in the true residual norm if the previous iteration
// made a reliable update
__global__ void __launch_bounds__(CG_BLK)
k_mx_fin1(const double *part, unsigned np, const double *part_ru, MixedCtrl *c, int t)
{
    if (c->done) return;
    const double pap = fin_sum(part, np);
    const int ru = c->ru;
    double rr_true = 0.0;
    if (ru) rr_true = fin_sum(part_ru, np);
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

// p = r + beta p over PU_R consecutive 256-row chunks per block (c0 is
// the first; nch of them exist).  On a reliable update first r = b - A x
// with the FP64 matrix for each row, and part_ru[chunk] = the chunk's
// |r|^2, summed with the thread-to-row map of one 256-thread block per
// chunk -- the partition of the SpMV's blocks, which the next fin1 sums.
// One chunk per block measured fastest (1, 2, 4 on one box: 63.7, 64.2,
// 64.2 ms per mixed step at 1.08M particles; 173.6, 176.1, 179.3 ms per
// step of the 64-scene ensemble): the reliable-update SpMV wants rows
// spread over threads more than the plain update wants rows per thread.
#define PU_R 1       // 256-row chunks per p-update block, one scene
#define PU_RB 1      // the same in the batched solve (more chunks per block serialise the
                     // reliable updates' FP64 residual rows: 4 measured 3 % slower per step)
template <int U, typename V, int R>
__device__ __forceinline__ void pupdate_chunks(const PkView &P, const double *__restrict__ b,
                                               const double *__restrict__ x, V *__restrict__ r, V *__restrict__ p,
                                               int n, int c0, int nch, V beta, int ru, double *part_ru)
{
    if (!ru) {
        V rv[R], pv[R];
#pragma unroll
        for (int c = 0; c < R; c++) {
            const int i = (c0 + c) * CG_BLK + threadIdx.x;
            if (c < nch && i < n) {
                rv[c] = r[i];
                pv[c] = p[i];
            }
        }
#pragma unroll
        for (int c = 0; c < R; c++) {
            const int i = (c0 + c) * CG_BLK + threadIdx.x;
            if (c < nch && i < n) p[i] = rv[c] + beta * pv[c];
        }
        return;
    }
    // each chunk's sum as block_sum forms it: wave sums, then the waves'
    // sums in wave order; one barrier for all chunks
    __shared__ double wpart[R][CG_BLK / WAVE];
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    for (int c = 0; c < nch; c++) {
        const int i = (c0 + c) * CG_BLK + threadIdx.x;
        double v = 0.0;
        if (i < n) {
            const double xi = x[i];
            const double s = packed_row_global<U>(P, P.val, i, x, P.diag[i] * xi);
            const V ri = (V)(b[i] - s);
            v = (double)ri * (double)ri;
            r[i] = ri;
            p[i] = ri + beta * p[i];
        }
        v = wave_sum(v);
        if (lane == 0) wpart[c][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < nch) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < CG_BLK / WAVE; w++) s += wpart[threadIdx.x][w];
        part_ru[c0 + threadIdx.x] = s;
    }
}

static inline unsigned pu_grid(unsigned nblk) { return (nblk + PU_R - 1) / PU_R; }
static inline unsigned pu_grid_b(unsigned nblk) { return (nblk + PU_RB - 1) / PU_RB; }

// the p update of the mixed iteration over nblk = cg_grid(n) chunks
template <int U, typename V>
__global__ void __launch_bounds__(CG_BLK)
k_mx_pupdate_pk(int n, unsigned nblk, PkView P, const double *__restrict__ b, const double *__restrict__ x,
                V *__restrict__ r, V *__restrict__ p, double *part_ru, const MixedCtrl *c)
{
    if (c->done) return;
    const int c0 = blockIdx.x * PU_R;
    const int nch = min(PU_R, (int)nblk - c0);
    pupdate_chunks<U, V, PU_R>(P, b, x, r, p, n, c0, nch, (V)c->beta, c->ru, part_ru);
}

// ---- the mixed iteration in three kernels (option fused = 2) ----------
// SpMV, then k_mx3_axpy (every block repeats k_mx_fin1's sums and
// decisions; block 0 records them), then k_mx3_pupdate (every block
// repeats k_mx_fin2's; block 0 records).  The arithmetic is the five-kernel
// form's, so the iterates have its bits.  A block that finds the solve
// stopped, or reads the stop flag block 0 of the same kernel has just set,
// does no update -- the same outcome either way.
template <typename V>
__global__ void __launch_bounds__(CG_BLK)
k_mx3_axpy(int n, int t, const V *__restrict__ p, const V *__restrict__ Ap, double *__restrict__ x,
           V *__restrict__ r, const double *part, const double *part_ru, unsigned np, double *part2, MixedCtrl *c)
{
    __shared__ double sh_alpha;
    __shared__ int sh_go;
    if (block_flag(&c->done)) return;
    // the first VEC_U rows' operands, loaded before the partials are summed
    const int st = gridDim.x * CG_BLK;
    const int i0 = blockIdx.x * CG_BLK + threadIdx.x;
    const bool pre = i0 + (VEC_U - 1) * st < n;
    V pv0[VEC_U], apv0[VEC_U], rv0[VEC_U];
    double xv0[VEC_U];
    if (pre) {
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            pv0[u] = p[i0 + u * st];
            apv0[u] = Ap[i0 + u * st];
            xv0[u] = x[i0 + u * st];
            rv0[u] = r[i0 + u * st];
        }
    }
    const double pap = fin_sum(part, np);
    const int ru = c->f_ru[t & 1];
    double rr_true = 0.0;
    if (ru) rr_true = fin_sum(part_ru, np);
    if (threadIdx.x == 0) {
        double rr = c->f_rr[t & 1], rr_ru = c->f_rr_ru[t & 1];
        int go = 1, conv = 0, brk = 0;
        if (ru) {
            rr = rr_true;
            rr_ru = rr_true;
            conv = sqrt(rr_true) < c->stop;
        }
        if (!conv) brk = fabs(pap) < 1.0e-30;
        go = !conv && !brk;
        const double alpha = go ? rr / pap : 0.0;
        if (blockIdx.x == 0) {
            if (ru) {
                c->n_ru++;
                c->rr = rr_true;
                c->rr_ru = rr_true;
                c->res = sqrt(rr_true);
            }
            if (!conv) c->pap = pap;
            if (go) {
                c->alpha = alpha;
                c->m_rr = rr;
                c->m_rr_ru = rr_ru;
            } else {
                c->iters = t;
                c->breakdown = brk;
                c->done = 1;
            }
        }
        sh_alpha = alpha;
        sh_go = go;
    }
    __syncthreads();
    if (!sh_go) return;
    const double alpha = sh_alpha;
    const V av = (V)alpha;
    double v = 0.0;
    int i = i0;
    if (pre) {
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            x[i + u * st] = xv0[u] + alpha * (double)pv0[u];
            const V ri = rv0[u] - av * apv0[u];
            r[i + u * st] = ri;
            v += (double)ri * (double)ri;
        }
        i += VEC_U * st;
    }
    for (; i + (VEC_U - 1) * st < n; i += VEC_U * st) {   // VEC_U rows in flight, same order
        V pv[VEC_U], apv[VEC_U], rv[VEC_U];
        double xv[VEC_U];
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            pv[u] = p[i + u * st];
            apv[u] = Ap[i + u * st];
            xv[u] = x[i + u * st];
            rv[u] = r[i + u * st];
        }
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            x[i + u * st] = xv[u] + alpha * (double)pv[u];
            const V ri = rv[u] - av * apv[u];
            r[i + u * st] = ri;
            v += (double)ri * (double)ri;
        }
    }
    for (; i < n; i += st) {
        x[i] += alpha * (double)p[i];
        const V ri = r[i] - av * Ap[i];
        r[i] = ri;
        v += (double)ri * (double)ri;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part2[blockIdx.x] = v;
}

template <int U, typename V>
__global__ void __launch_bounds__(CG_BLK)
k_mx3_pupdate(int n, unsigned nblk, int t, int max_iter, PkView P, const double *__restrict__ b,
              const double *__restrict__ x, V *__restrict__ r, V *__restrict__ p, const double *part2,
              unsigned np2, double *part_ru, MixedCtrl *c)
{
    __shared__ double sh_beta;
    __shared__ int sh_go, sh_ru;
    if (block_flag(&c->done)) return;
    // this block's row of r and p, loaded before the partials are summed
    // (used unless the iteration takes a reliable update)
    static_assert(PU_R == 1, "one chunk per p-update block");
    const int i0 = blockIdx.x * CG_BLK + threadIdx.x;
    V rpre = (V)0, ppre = (V)0;
    if (i0 < n) {
        rpre = r[i0];
        ppre = p[i0];
    }
    const double rr_new = fin_sum(part2, np2);
    if (threadIdx.x == 0) {
        const double rr = c->m_rr, rr_ru = c->m_rr_ru;
        const int go = !(rr < 1.0e-60 || t + 1 >= max_iter);
        double beta = 0.0;
        int ru = 0;
        if (go) {
            beta = rr_new / rr;
            ru = (rr_new < c->delta2 * rr_ru) || (sqrt(rr_new) < c->stop);
        }
        if (blockIdx.x == 0) {
            if (go) {
                c->beta = beta;
                c->rr = rr_new;
                c->ru = ru;
                c->f_rr[(t + 1) & 1] = rr_new;
                c->f_rr_ru[(t + 1) & 1] = rr_ru;
                c->f_ru[(t + 1) & 1] = ru;
            } else {
                c->iters = t + 1;
                c->res = sqrt(rr_new);
                c->done = 1;
            }
        }
        sh_beta = beta;
        sh_go = go;
        sh_ru = ru;
    }
    __syncthreads();
    if (!sh_go) return;
    if (!sh_ru) {
        if (i0 < n) p[i0] = rpre + (V)sh_beta * ppre;
        return;
    }
    const int c0 = blockIdx.x * PU_R;
    const int nch = min(PU_R, (int)nblk - c0);
    pupdate_chunks<U, V, PU_R>(P, b, x, r, p, n, c0, nch, (V)sh_beta, sh_ru, part_ru);
}

struct MixedBufs {
    MixedCtrl *d = nullptr, *h = nullptr;
    double *part_ru = nullptr;
};

template <typename V>
static int cg_mixed_run(Ctx *c, const SolveOpts &o, SolveResult *res, V *r, V *p, V *Ap)
{
    const Packed *pk = c->pk;
    const int n = pk->n_eq;
    const unsigned g = cg_grid(n), gv = vec_grid(n);
    hipStream_t s = c->stream;
    MixedBufs *mb = c->mixed;
    const PkView P = pk_view(pk);
    double *x = pk->x;
    hipLaunchKernelGGL(k_mx_init<V>, dim3(gv), dim3(CG_BLK), 0, s, n, pk->b, x, r, p, c->partB, c->partA);
    const double delta = o.inner_tol > 0 ? o.inner_tol : 1e-3;
    hipLaunchKernelGGL(k_mx_init_fin, dim3(1), dim3(CG_BLK), 0, s, c->partB, c->partA, gv, mb->d, o.tol,
                       o.tol_mode, o.max_iter, delta);
    LAUNCH_CHECK();
    typedef void (*mx_fn)(int, PkView, const V *, V *, double *, const MixedCtrl *);
    const mx_fn spmv = o.spmv_variant == 1 ? k_mx_spmv_pk<16, V, true>
                                           : (o.spmv_variant == 2 ? k_mx_spmv_pk<16, V, false> : k_mx_spmv_pk<8, V, true>);
    const size_t lds = sizeof(V) * (size_t)pk->hmax;
    const bool coded = sizeof(V) == sizeof(float) && o.precision == 3;
    QConsts qk{};
    if (coded) {
        hipLaunchKernelGGL(k_mx_q16_build, dim3(g), dim3(CG_BLK), 0, s, n, P, o.neg_coeff, o.re_lap, o.r_min, c->pk->cq);
        LAUNCH_CHECK();
        qk.neg_coeff = (float)o.neg_coeff;
        qk.re = (float)o.re_lap;
        qk.r_lo = (float)o.r_min;
        qk.dr = (float)((o.re_lap - o.r_min) / Q_CODE_MAX);
    }
    const unsigned *cq = c->pk->cq;
    typedef void (*q_fn)(int, PkView, const unsigned *, QConsts, const float *, float *, double *, const MixedCtrl *);
    const q_fn spmv_q = o.spmv_variant == 1 ? k_mx_spmv_q<16, true>
                                            : (o.spmv_variant == 2 ? k_mx_spmv_q<16, false> : k_mx_spmv_q<8, true>);
    const bool three = o.fused == 2;
    int rc = run_chunks(s, o.max_iter, o.check_every, &mb->d->done, &mb->h->done, [&](int t) {
        prof_begin(PK_SPMV, s);
        if constexpr (sizeof(V) == sizeof(float)) {
            if (coded)
                hipLaunchKernelGGL(spmv_q, dim3(g), dim3(CG_BLK), lds, s, n, P, cq, qk, p, Ap, c->partA, mb->d);
            else
                hipLaunchKernelGGL(spmv, dim3(g), dim3(CG_BLK), lds, s, n, P, p, Ap, c->partA, mb->d);
        } else {
            hipLaunchKernelGGL(spmv, dim3(g), dim3(CG_BLK), lds, s, n, P, p, Ap, c->partA, mb->d);
        }
        prof_end(PK_SPMV, s);
        if (three) {
            prof_begin(PK_AXPY, s);
            hipLaunchKernelGGL(k_mx3_axpy<V>, dim3(gv), dim3(CG_BLK), 0, s, n, t, p, Ap, x, r, c->partA, mb->part_ru,
                               g, c->partB, mb->d);
            prof_end(PK_AXPY, s);
            prof_begin(PK_PUPDATE, s);
            hipLaunchKernelGGL((k_mx3_pupdate<8, V>), dim3(pu_grid(g)), dim3(CG_BLK), 0, s, n, g, t, o.max_iter, P,
                               pk->b, x, r, p, c->partB, gv, mb->part_ru, mb->d);
            prof_end(PK_PUPDATE, s);
            return;
        }
        hipLaunchKernelGGL(k_mx_fin1, dim3(1), dim3(CG_BLK), 0, s, c->partA, g, mb->part_ru, mb->d, t);
        prof_begin(PK_AXPY, s);
        hipLaunchKernelGGL(k_mx_axpy<V>, dim3(gv), dim3(CG_BLK), 0, s, n, p, Ap, x, r, c->partB, mb->d);
        prof_end(PK_AXPY, s);
        hipLaunchKernelGGL(k_mx_fin2, dim3(1), dim3(CG_BLK), 0, s, c->partB, gv, mb->d, t, o.max_iter);
        prof_begin(PK_PUPDATE, s);
        hipLaunchKernelGGL((k_mx_pupdate_pk<8, V>), dim3(pu_grid(g)), dim3(CG_BLK), 0, s, n, g, P, pk->b, x, r, p,
                           mb->part_ru, mb->d);
        prof_end(PK_PUPDATE, s);
    });
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipMemcpy(mb->h, mb->d, sizeof(MixedCtrl), hipMemcpyDeviceToHost));
    res->iters = mb->h->done ? mb->h->iters : o.max_iter;
    res->breakdown = mb->h->breakdown;
    res->res = mb->h->res;
    res->bnorm = mb->h->bnorm;
    res->outer = mb->h->n_ru;
    return 0;
}

// the FP32 vectors (precisions 2 and 3) and the control block of c's mixed solve
static int mixed_reserve(Ctx *c, int precision)
{
    if (!c->mixed) c->mixed = new MixedBufs();
    MixedBufs *mb = c->mixed;
    const size_t NS = ((size_t)(c->cap > 0 ? c->cap : 1) + 63) / 64 * 64;
    if (precision >= 2 && !c->fr) {
        HIP_OK(hipMalloc(&c->fr, sizeof(float) * NS));
        HIP_OK(hipMalloc(&c->fp, sizeof(float) * NS));
        HIP_OK(hipMalloc(&c->fAp, sizeof(float) * NS));
    }
    if (!mb->d) {
        HIP_OK(hipMalloc(&mb->d, sizeof(MixedCtrl)));
        HIP_OK(hipHostMalloc(&mb->h, sizeof(MixedCtrl), hipHostMallocDefault));
        HIP_OK(hipMalloc(&mb->part_ru, sizeof(double) * 65536));
    }
    return 0;
}

// room for the coded words of c's packed system (one per stored entry)
static int coded_reserve(Ctx *c)
{
    Packed *pw = c->pk;
    const size_t need = pw->tot_val * 64;
    if (pw->cap_cq < need) {
        if (pw->cq) hipFree(pw->cq);
        pw->cq = nullptr;
        pw->cap_cq = 0;
        HIP_OK(hipMalloc(&pw->cq, sizeof(unsigned) * (need + need / 8 + 64)));
        pw->cap_cq = need + need / 8 + 64;
    }
    return 0;
}

static int cg_mixed(Ctx *c, const SolveOpts &o, SolveResult *res)
{
    int rc = mixed_reserve(c, o.precision);
    if (!rc && o.precision == 3) rc = coded_reserve(c);
    if (rc) return rc;
    if (o.precision >= 2) return cg_mixed_run<float>(c, o, res, c->fr, c->fp, c->fAp);
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

// Solves the PPE assembled in c; the solution goes to c->x in slot order
// (0 at slots that are not unknowns).
int solve_ppe(Ctx *c, const SolveOpts &o, SolveResult *res)
{
    int rc = cg_alloc_ctrl(c);
    if (rc) return rc;
    if (cg_grid(c->n) > 65536u) {   // partial-sum buffers hold 65536 blocks
        mps_set_error(MPS_ERR_ARG, __FILE__, __LINE__, "too many particles for one solve (max 16.7M)");
        return MPS_ERR_ARG;
    }
    if (o.solver == 1) return iccg_solve(c, o, res);
    const int mixed = (o.precision >= 1 && o.precision <= 3);
    rc = pack_ppe(c, mixed);
    if (rc) return rc;
    if (c->pk->n_eq == 0) {
        res->iters = 0;
        res->breakdown = 0;
        res->res = 0;
        res->bnorm = 0;
        res->outer = 0;
        return unpack_solution(c);
    }
    rc = mixed ? cg_mixed(c, o, res) : cg_fp64(c, o, res);
    if (rc) return rc;
    return unpack_solution(c);
}

// ---------------------------------------------------------------------
// Batched FP64 CG: many independent systems (one per scene of an
// ensemble) iterated together, one launch per phase for all of them.
// Block b of a phase belongs to scene bs[b] and is that scene's local
// block b - off; each scene keeps exactly the block partition of its own
// single-scene solve (SpMV: one block per 256 rows; vector kernels:
// vec_grid(n_eq) blocks, grid-stride) and its finalisers sum the same
// partials in the same order, so every scene's iterates are the bits of
// its single-scene CG.  A scene whose solve has stopped masks its blocks
// out; ndone counts the stopped scenes for the host.
// ---------------------------------------------------------------------

struct BScene;
typedef void (*spmv_fn_b)(const BScene *, const int *, double *, const CgCtrl *);

__global__ void k_fill_zero_int(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

struct BScene {
    PkView P;
    const double *b;
    double *x, *r, *p, *Ap;
    int n_eq;
    int blk_off, nblk;       // SpMV blocks
    int ax_off, nax;         // vector-kernel blocks
    int max_iter, tol_mode;
    double tol;
};

// A scene of the table with its buffers as global pointers (gptr).
__device__ __forceinline__ BScene bscene_global(const BScene &s)
{
    BScene g = s;
    g.P = pk_global(s.P);
    g.b = gptr(s.b); g.x = gptr(s.x); g.r = gptr(s.r); g.p = gptr(s.p); g.Ap = gptr(s.Ap);
    return g;
}

__global__ void __launch_bounds__(CG_BLK)
k_b_init(const BScene *__restrict__ sc, const int *__restrict__ axs, double *part2)
{
    const BScene S = bscene_global(sc[axs[blockIdx.x]]);
    const int j = blockIdx.x - S.ax_off;
    double v = 0.0;
    for (int i = j * CG_BLK + threadIdx.x; i < S.n_eq; i += S.nax * CG_BLK) {
        const double bi = S.b[i];
        S.x[i] = 0.0;
        S.r[i] = bi;
        S.p[i] = bi;
        v += bi * bi;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part2[blockIdx.x] = v;
}

__global__ void __launch_bounds__(CG_BLK)
k_b_init_fin(const BScene *__restrict__ sc, const double *part2, CgCtrl *ctrl, int *ndone)
{
    const BScene S = bscene_global(sc[blockIdx.x]);
    CgCtrl *c = ctrl + blockIdx.x;
    const double bb = fin_sum(part2 + S.ax_off, S.nax);
    if (threadIdx.x == 0) {
        c->rr[0] = bb;
        c->bnorm = sqrt(bb);
        c->stop = S.tol_mode ? S.tol * sqrt(bb) : S.tol;
        c->done = 0;
        c->iters = 0;
        c->breakdown = 0;
        c->res = sqrt(bb);
        if (S.n_eq == 0 || S.max_iter <= 0 || sqrt(bb) < c->stop) {
            c->done = 1;
            atomicAdd(ndone, 1);
        }
    }
}

template <int U, bool NT>
__global__ void __launch_bounds__(CG_BLK)
k_b_spmv(const BScene *__restrict__ sc, const int *__restrict__ bs, double *part, const CgCtrl *ctrl)
{
    extern __shared__ unsigned char smem_raw[];
    double *hx = reinterpret_cast<double *>(smem_raw);
    const int si = bs[blockIdx.x];
    if (ctrl[si].done) return;
    const BScene S = bscene_global(sc[si]);
    const int lb = blockIdx.x - S.blk_off;
    const double *p = S.p;
    stage_halo_of(S.P, lb, XPlain<double>{p}, hx);
    const int i = lb * CG_BLK + threadIdx.x;
    double v = 0.0;
    if (i < S.n_eq) {
        const double pi = p[i];
        const double s = packed_row<U, false, double, NT>(S.P, S.P.val, i, XPlain<double>{p}, hx, pi, S.P.diag[i] * pi);
        S.Ap[i] = s;
        v = pi * s;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(CG_BLK)
k_b_spmv_fin(const BScene *__restrict__ sc, const double *part, CgCtrl *ctrl, int *ndone)
{
    CgCtrl *c = ctrl + blockIdx.x;
    if (c->done) return;
    const int t = ndone[1];
    const BScene S = bscene_global(sc[blockIdx.x]);
    const double pap = fin_sum(part + S.blk_off, S.nblk);
    if (threadIdx.x == 0) {
        c->pap = pap;
        if (fabs(pap) < 1.0e-30) {
            c->done = 1;
            c->iters = t;
            c->breakdown = 1;
            c->res = sqrt(c->rr[t & 1]);
            atomicAdd(ndone, 1);
        } else {
            c->alpha = c->rr[t & 1] / pap;
        }
    }
}

__global__ void __launch_bounds__(CG_BLK)
k_b_axpy(const BScene *__restrict__ sc, const int *__restrict__ axs, double *part2, const CgCtrl *ctrl)
{
    const int si = axs[blockIdx.x];
    if (ctrl[si].done) return;
    const BScene S = bscene_global(sc[si]);
    const int j = blockIdx.x - S.ax_off;
    const double alpha = ctrl[si].alpha;
    double v = 0.0;
    const int st = S.nax * CG_BLK;
    int i = j * CG_BLK + threadIdx.x;
    for (; i + (VEC_U - 1) * st < S.n_eq; i += VEC_U * st) {
        double pv[VEC_U], av[VEC_U], xv[VEC_U], rv[VEC_U];
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            pv[u] = S.p[i + u * st];
            av[u] = S.Ap[i + u * st];
            xv[u] = S.x[i + u * st];
            rv[u] = S.r[i + u * st];
        }
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            S.x[i + u * st] = xv[u] + alpha * pv[u];
            const double ri = rv[u] - alpha * av[u];
            S.r[i + u * st] = ri;
            v += ri * ri;
        }
    }
    for (; i < S.n_eq; i += st) {
        S.x[i] += alpha * S.p[i];
        const double ri = S.r[i] - alpha * S.Ap[i];
        S.r[i] = ri;
        v += ri * ri;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part2[blockIdx.x] = v;
}

__global__ void __launch_bounds__(CG_BLK)
k_b_axpy_fin(const BScene *__restrict__ sc, const double *part2, CgCtrl *ctrl, int *ndone)
{
    CgCtrl *c = ctrl + blockIdx.x;
    if (c->done) return;
    const int t = ndone[1];
    const BScene S = bscene_global(sc[blockIdx.x]);
    const double rr_new = fin_sum(part2 + S.ax_off, S.nax);
    if (threadIdx.x == 0) {
        const double rr = c->rr[t & 1];
        c->rr[(t + 1) & 1] = rr_new;
        c->res = sqrt(rr_new);
        int stop = 0;
        if (rr < 1.0e-60) {
            stop = 1;
            c->iters = t;
        } else {
            c->beta = rr_new / rr;
            if (t + 1 >= S.max_iter) {
                stop = 1;
                c->iters = S.max_iter;
            } else if (sqrt(rr_new) < c->stop) {
                stop = 1;
                c->iters = t + 1;
            }
        }
        if (stop) {
            c->done = 1;
            atomicAdd(ndone, 1);
        }
    }
}

__global__ void __launch_bounds__(CG_BLK)
k_b_pupdate(const BScene *__restrict__ sc, const int *__restrict__ axs, const CgCtrl *ctrl, int *ndone)
{
    // the iteration's last kernel advances the iteration index (no other
    // kernel of the iteration reads it while this one runs)
    if (blockIdx.x == 0 && threadIdx.x == 0) ndone[1] += 1;
    const int si = axs[blockIdx.x];
    if (ctrl[si].done) return;
    const BScene S = bscene_global(sc[si]);
    const int j = blockIdx.x - S.ax_off;
    const double beta = ctrl[si].beta;
    const int st = S.nax * CG_BLK;
    int i = j * CG_BLK + threadIdx.x;
    for (; i + (VEC_U - 1) * st < S.n_eq; i += VEC_U * st) {
        double rv[VEC_U], pv[VEC_U];
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            rv[u] = S.r[i + u * st];
            pv[u] = S.p[i + u * st];
        }
#pragma unroll
        for (int u = 0; u < VEC_U; u++) S.p[i + u * st] = rv[u] + beta * pv[u];
    }
    for (; i < S.n_eq; i += st) S.p[i] = S.r[i] + beta * S.p[i];
}

struct BatchBufs {
    int cap_sc = 0, cap_blk = 0, cap_ax = 0;
    int cap_ctrl = 0, cap_part = 0, cap_part2 = 0;
    BScene *sc = nullptr, *hsc = nullptr;
    int *bs = nullptr, *hbs = nullptr, *axs = nullptr, *haxs = nullptr;
    double *part = nullptr, *part2 = nullptr;
    CgCtrl *ctrl = nullptr, *hctrl = nullptr;
    int *ndone = nullptr, *hndone = nullptr;
};
static BatchBufs g_batch;

template <typename T>
static int reserve_pair(T **d, T **h, int *cap, int need)
{
    if (*cap >= need) return 0;
    if (*d) hipFree(*d);
    if (*h) hipHostFree(*h);
    *d = nullptr;
    *h = nullptr;
    const int n = need + need / 4 + 16;
    HIP_OK(hipMalloc((void **)d, sizeof(T) * n));
    HIP_OK(hipHostMalloc((void **)h, sizeof(T) * n, hipHostMallocDefault));
    *cap = n;
    return 0;
}

// The iteration loop of a batched solve: iterations in chunks of `chunk`
// (iter() enqueues one), with the count of stopped scenes (dndone[0])
// copied back after each chunk and read one chunk later, until all ns
// scenes have stopped.  The kernels take the iteration index from
// dndone[1], which each iteration's last kernel advances, so with `graph`
// one chunk is captured once and the same graph is replayed for every
// chunk; scenes that have stopped return at once in either form, so both
// give the same bits.
template <typename F>
static int batch_loop(hipStream_t s, int max_iter, int chunk, int ns, const int *dndone, int *hndone, bool graph,
                      F iter)
{
    hipGraphExec_t ge = nullptr;
    if (graph) {
        HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int u = 0; u < chunk; u++) iter();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (e != hipSuccess || hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
            if (g) hipGraphDestroy(g);
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, "capture of the batched CG iterations failed");
            return MPS_ERR_HIP;
        }
        hipGraphDestroy(g);
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = 0;
    if (hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess)
        rc = MPS_ERR_HIP;
    int pending = -1, t = 0, k = 0;
    while (!rc && t < max_iter) {
        if (ge) {
            if (hipGraphLaunch(ge, s) != hipSuccess) rc = MPS_ERR_HIP;
            t += chunk;
        } else {
            const int t_end = (t + chunk < max_iter) ? t + chunk : max_iter;
            for (; t < t_end; t++) iter();
        }
        if (!rc && (hipGetLastError() != hipSuccess ||
                    hipMemcpyAsync(hndone, dndone, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                    hipEventRecord(ev[k & 1], s) != hipSuccess))
            rc = MPS_ERR_HIP;
        if (!rc && pending >= 0) {
            if (hipEventSynchronize(ev[pending & 1]) != hipSuccess) rc = MPS_ERR_HIP;
            else if (*(volatile int *)hndone >= ns) break;
        }
        pending = k;
        k++;
    }
    for (hipEvent_t e : ev)
        if (e) hipEventDestroy(e);
    if (ge) hipGraphExecDestroy(ge);
    if (rc) mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, "batched CG iterations");
    return rc;
}

// Solves the packed systems of cs[0..ns) (each already packed by pack_ppe)
// with FP64 CG on stream s; solutions stay in each scene's packed x.
int batch_cg(Ctx *const *cs, int ns, const SolveOpts *opts, SolveResult *res, hipStream_t s)
{
    BatchBufs &B = g_batch;
    int tot_blk = 0, tot_ax = 0, hmax = 0, max_iter = 0, check_every = 16;
    for (int k = 0; k < ns; k++) {
        const int n = cs[k]->pk->n_eq;
        tot_blk += (int)cg_grid(n);
        tot_ax += n > 0 ? (int)vec_grid(n) : 0;
        if (cs[k]->pk->hmax > hmax) hmax = cs[k]->pk->hmax;
        if (opts[k].max_iter > max_iter) max_iter = opts[k].max_iter;
        if (opts[k].check_every > 0) check_every = opts[k].check_every;
    }
    int rc = reserve_pair(&B.sc, &B.hsc, &B.cap_sc, ns);
    if (!rc) rc = reserve_pair(&B.bs, &B.hbs, &B.cap_blk, tot_blk + 1);
    if (!rc) rc = reserve_pair(&B.axs, &B.haxs, &B.cap_ax, tot_ax + 1);
    if (rc) return rc;
    if (!rc) rc = reserve_pair(&B.ctrl, &B.hctrl, &B.cap_ctrl, ns);
    if (rc) return rc;
    if (B.cap_part < B.cap_blk) {
        if (B.part) hipFree(B.part);
        B.part = nullptr;
        HIP_OK(hipMalloc(&B.part, sizeof(double) * B.cap_blk));
        B.cap_part = B.cap_blk;
    }
    if (B.cap_part2 < B.cap_ax) {
        if (B.part2) hipFree(B.part2);
        B.part2 = nullptr;
        HIP_OK(hipMalloc(&B.part2, sizeof(double) * B.cap_ax));
        B.cap_part2 = B.cap_ax;
    }
    if (!B.ndone) {
        HIP_OK(hipMalloc(&B.ndone, 2 * sizeof(int)));   // [0] stopped scenes, [1] iteration
        HIP_OK(hipHostMalloc(&B.hndone, sizeof(int), hipHostMallocDefault));
    }
    int bo = 0, ao = 0;
    for (int k = 0; k < ns; k++) {
        Ctx *c = cs[k];
        const Packed *pk = c->pk;
        BScene &S = B.hsc[k];
        S.P = pk_view(pk);
        S.b = pk->b;
        S.x = pk->x;
        S.r = c->r;
        S.p = c->pv;
        S.Ap = c->Ap;
        S.n_eq = pk->n_eq;
        S.blk_off = bo;
        S.nblk = (int)cg_grid(pk->n_eq);
        S.ax_off = ao;
        S.nax = pk->n_eq > 0 ? (int)vec_grid(pk->n_eq) : 0;
        S.max_iter = opts[k].max_iter;
        S.tol_mode = opts[k].tol_mode;
        S.tol = opts[k].tol;
        for (int q = 0; q < S.nblk; q++) B.hbs[bo + q] = k;
        for (int q = 0; q < S.nax; q++) B.haxs[ao + q] = k;
        bo += S.nblk;
        ao += S.nax;
    }
    HIP_OK(hipMemcpyAsync(B.sc, B.hsc, sizeof(BScene) * ns, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(B.bs, B.hbs, sizeof(int) * (tot_blk + 1), hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(B.axs, B.haxs, sizeof(int) * (tot_ax + 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_zero_int, dim3(1), dim3(64), 0, s, B.ndone, 2);
    if (tot_ax > 0) hipLaunchKernelGGL(k_b_init, dim3(tot_ax), dim3(CG_BLK), 0, s, B.sc, B.axs, B.part2);
    hipLaunchKernelGGL(k_b_init_fin, dim3(ns), dim3(CG_BLK), 0, s, B.sc, B.part2, B.ctrl, B.ndone);
    LAUNCH_CHECK();
    const size_t lds = sizeof(double) * (size_t)hmax;
    const spmv_fn_b spmv = opts[0].spmv_variant == 1 ? k_b_spmv<16, true>
                                                     : (opts[0].spmv_variant == 2 ? k_b_spmv<16, false> : k_b_spmv<8, true>);
    // the host stops when every scene has stopped: watch ndone == ns
    int rc2 = batch_loop(s, max_iter, check_every, ns, B.ndone, B.hndone, opts[0].graph, [&]() {
        prof_begin(PK_SPMV, s);
        if (tot_blk > 0) hipLaunchKernelGGL(spmv, dim3(tot_blk), dim3(CG_BLK), lds, s, B.sc, B.bs, B.part, B.ctrl);
        prof_end(PK_SPMV, s);
        hipLaunchKernelGGL(k_b_spmv_fin, dim3(ns), dim3(CG_BLK), 0, s, B.sc, B.part, B.ctrl, B.ndone);
        if (tot_ax > 0) hipLaunchKernelGGL(k_b_axpy, dim3(tot_ax), dim3(CG_BLK), 0, s, B.sc, B.axs, B.part2, B.ctrl);
        hipLaunchKernelGGL(k_b_axpy_fin, dim3(ns), dim3(CG_BLK), 0, s, B.sc, B.part2, B.ctrl, B.ndone);
        if (tot_ax > 0)
            hipLaunchKernelGGL(k_b_pupdate, dim3(tot_ax), dim3(CG_BLK), 0, s, B.sc, B.axs, B.ctrl, B.ndone);
    });
    if (rc2) return rc2;
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipMemcpy(B.hctrl, B.ctrl, sizeof(CgCtrl) * ns, hipMemcpyDeviceToHost));
    for (int q = 0; q < ns; q++) {
        const CgCtrl &c = B.hctrl[q];
        res[q].iters = c.done ? c.iters : max_iter;
        res[q].breakdown = c.breakdown;
        res[q].res = c.res;
        res[q].bnorm = c.bnorm;
        res[q].outer = 1;
    }
    return 0;
}

// ---------------------------------------------------------------------
// Batched mixed CG (precisions 1-3): the mixed iteration above over many
// scenes, with the block tables of the batched FP64 CG.  Every scene keeps
// the block partition of its own single-scene mixed solve (SpMV and
// p-update: one block per 256 rows; init and axpy: vec_grid(n_eq) blocks,
// grid-stride), the finalisers sum the same partials in the same order,
// and the reliable updates fire on the same iterations, so each scene ends
// with the bits of its single-scene solve.
// ---------------------------------------------------------------------

template <typename V>
struct BMScene {
    PkView P;
    const double *b;
    double *x;
    V *r, *p, *Ap;
    unsigned *cq;
    QConsts qk;
    int n_eq;
    int blk_off, nblk;       // SpMV blocks (256-row chunks)
    int pb_off, npb;         // p-update blocks (PU_R chunks each)
    int ax_off, nax;         // init / axpy blocks
    int max_iter, tol_mode;
    double tol, delta, neg_coeff, re, r_min;
};

template <typename V>
__device__ __forceinline__ BMScene<V> bmscene_global(const BMScene<V> &s)
{
    BMScene<V> g = s;
    g.P = pk_global(s.P);
    g.b = gptr(s.b); g.x = gptr(s.x); g.r = gptr(s.r); g.p = gptr(s.p); g.Ap = gptr(s.Ap);
    g.cq = gptr(s.cq);
    return g;
}

__global__ void __launch_bounds__(CG_BLK)
k_bm_q16_build(const BMScene<float> *__restrict__ sc, const int *__restrict__ bs)
{
    const BMScene<float> S = bmscene_global(sc[bs[blockIdx.x]]);
    const int k = (blockIdx.x - S.blk_off) * CG_BLK + threadIdx.x;
    if (k < S.n_eq) q16_build_row(k, S.P, S.neg_coeff, S.re, S.r_min, S.cq);
}

template <typename V>
__global__ void __launch_bounds__(CG_BLK)
k_bm_init(const BMScene<V> *__restrict__ sc, const int *__restrict__ axs, double *part2, double *part3)
{
    const BMScene<V> S = bmscene_global(sc[axs[blockIdx.x]]);
    const int j = blockIdx.x - S.ax_off;
    double v = 0.0, w = 0.0;
    for (int i = j * CG_BLK + threadIdx.x; i < S.n_eq; i += S.nax * CG_BLK) {
        const double bi = S.b[i];
        S.x[i] = 0.0;
        const V f = (V)bi;
        S.r[i] = f;
        S.p[i] = f;
        v += (double)f * (double)f;
        w += bi * bi;
    }
    v = block_sum<double, CG_BLK>(v);
    w = block_sum<double, CG_BLK>(w);
    if (threadIdx.x == 0) {
        part2[blockIdx.x] = v;
        part3[blockIdx.x] = w;
    }
}

template <typename V>
__global__ void __launch_bounds__(CG_BLK)
k_bm_init_fin(const BMScene<V> *__restrict__ sc, const double *part2, const double *part3, MixedCtrl *ctrl,
              int *ndone)
{
    const BMScene<V> S = bmscene_global(sc[blockIdx.x]);
    MixedCtrl *c = ctrl + blockIdx.x;
    const double rr = fin_sum(part2 + S.ax_off, S.nax);
    const double bb = fin_sum(part3 + S.ax_off, S.nax);
    if (threadIdx.x == 0) {
        c->rr = rr;
        c->rr_ru = rr;
        c->bnorm = sqrt(bb);
        c->stop = S.tol_mode ? S.tol * sqrt(bb) : S.tol;
        c->delta2 = S.delta * S.delta;
        c->done = 0;
        c->iters = 0;
        c->breakdown = 0;
        c->ru = 0;
        c->n_ru = 0;
        c->res = sqrt(bb);
        // an empty system is solved (the single-scene solve skips it)
        if (S.n_eq == 0 || S.max_iter <= 0 || sqrt(bb) < c->stop) {
            c->done = 1;
            atomicAdd(ndone, 1);
        }
    }
}

template <int U, typename V, bool CODED, bool NT>
__global__ void __launch_bounds__(CG_BLK)
k_bm_spmv(const BMScene<V> *__restrict__ sc, const int *__restrict__ bs, double *part, const MixedCtrl *ctrl)
{
    extern __shared__ unsigned char smem_raw[];
    V *hx = reinterpret_cast<V *>(smem_raw);
    const int si = bs[blockIdx.x];
    if (ctrl[si].done) return;
    const BMScene<V> S = bmscene_global(sc[si]);
    const int lb = blockIdx.x - S.blk_off;
    const V *p = S.p;
    stage_halo_of(S.P, lb, XPlain<V>{p}, hx);
    const int i = lb * CG_BLK + threadIdx.x;
    double v = 0.0;
    if (i < S.n_eq) {
        const V pi = p[i];
        V s;
        if constexpr (CODED)
            s = coded_row<U, NT>(S.P, S.cq, S.qk, p, hx, i, pi);
        else
            s = packed_row<U, true, V, NT>(S.P, S.P.fval, i, XPlain<V>{p}, hx, pi, (V)S.P.ddir[i] * pi);
        S.Ap[i] = s;
        v = (double)pi * (double)s;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

template <typename V>
__global__ void __launch_bounds__(CG_BLK)
k_bm_fin1(const BMScene<V> *__restrict__ sc, const double *part, const double *part_ru, MixedCtrl *ctrl,
          int *ndone)
{
    MixedCtrl *c = ctrl + blockIdx.x;
    if (c->done) return;
    const int t = ndone[1];
    const BMScene<V> S = bmscene_global(sc[blockIdx.x]);
    const double pap = fin_sum(part + S.blk_off, S.nblk);
    const int ru = c->ru;
    double rr_true = 0.0;
    if (ru) rr_true = fin_sum(part_ru + S.blk_off, S.nblk);
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
                atomicAdd(ndone, 1);
                return;
            }
        }
        c->pap = pap;
        if (fabs(pap) < 1.0e-30) {
            c->done = 1;
            c->iters = t;
            c->breakdown = 1;
            atomicAdd(ndone, 1);
        } else {
            c->alpha = c->rr / pap;
        }
    }
}

template <typename V>
__global__ void __launch_bounds__(CG_BLK)
k_bm_axpy(const BMScene<V> *__restrict__ sc, const int *__restrict__ axs, double *part2, const MixedCtrl *ctrl)
{
    const int si = axs[blockIdx.x];
    if (ctrl[si].done) return;
    const BMScene<V> S = bmscene_global(sc[si]);
    const int j = blockIdx.x - S.ax_off;
    const double alpha = ctrl[si].alpha;
    const V av = (V)alpha;
    double v = 0.0;
    const int st = S.nax * CG_BLK;
    int i = j * CG_BLK + threadIdx.x;
    for (; i + (VEC_U - 1) * st < S.n_eq; i += VEC_U * st) {
        V pv[VEC_U], apv[VEC_U], rv[VEC_U];
        double xv[VEC_U];
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            pv[u] = S.p[i + u * st];
            apv[u] = S.Ap[i + u * st];
            xv[u] = S.x[i + u * st];
            rv[u] = S.r[i + u * st];
        }
#pragma unroll
        for (int u = 0; u < VEC_U; u++) {
            S.x[i + u * st] = xv[u] + alpha * (double)pv[u];
            const V ri = rv[u] - av * apv[u];
            S.r[i + u * st] = ri;
            v += (double)ri * (double)ri;
        }
    }
    for (; i < S.n_eq; i += st) {
        S.x[i] += alpha * (double)S.p[i];
        const V ri = S.r[i] - av * S.Ap[i];
        S.r[i] = ri;
        v += (double)ri * (double)ri;
    }
    v = block_sum<double, CG_BLK>(v);
    if (threadIdx.x == 0) part2[blockIdx.x] = v;
}

template <typename V>
__global__ void __launch_bounds__(CG_BLK)
k_bm_fin2(const BMScene<V> *__restrict__ sc, const double *part2, MixedCtrl *ctrl, int *ndone)
{
    MixedCtrl *c = ctrl + blockIdx.x;
    if (c->done) return;
    const int t = ndone[1];
    const BMScene<V> S = bmscene_global(sc[blockIdx.x]);
    const double rr_new = fin_sum(part2 + S.ax_off, S.nax);
    if (threadIdx.x == 0) {
        const double rr = c->rr;
        if (rr < 1.0e-60 || t + 1 >= S.max_iter) {
            c->done = 1;
            c->iters = t + 1;
            c->res = sqrt(rr_new);
            atomicAdd(ndone, 1);
            return;
        }
        c->beta = rr_new / rr;
        c->rr = rr_new;
        c->ru = (rr_new < c->delta2 * c->rr_ru) || (sqrt(rr_new) < c->stop);
    }
}

template <int U, typename V>
__global__ void __launch_bounds__(CG_BLK)
k_bm_pupdate(const BMScene<V> *__restrict__ sc, const int *__restrict__ pbs, double *part_ru, const MixedCtrl *ctrl,
             int *ndone)
{
    // the iteration's last kernel advances the iteration index (no other
    // kernel of the iteration reads it while this one runs)
    if (blockIdx.x == 0 && threadIdx.x == 0) ndone[1] += 1;
    const int si = pbs[blockIdx.x];
    if (ctrl[si].done) return;
    const BMScene<V> S = bmscene_global(sc[si]);
    const int c0 = (blockIdx.x - S.pb_off) * PU_RB;
    const int nch = min(PU_RB, S.nblk - c0);
    pupdate_chunks<U, V, PU_RB>(S.P, S.b, S.x, S.r, S.p, S.n_eq, c0, nch, (V)ctrl[si].beta, ctrl[si].ru,
                                part_ru + S.blk_off);
}

struct BatchMixedBufs {
    int cap_sc = 0, cap_ctrl = 0, cap_part3 = 0, cap_ru = 0, cap_pb = 0;
    unsigned char *sc = nullptr, *hsc = nullptr;
    int *pbs = nullptr, *hpbs = nullptr;
    MixedCtrl *ctrl = nullptr, *hctrl = nullptr;
    double *part3 = nullptr, *part_ru = nullptr;
};
static BatchMixedBufs g_bmixed;

static int reserve_dev(double **d, int *cap, int need)
{
    if (*cap >= need) return 0;
    if (*d) hipFree(*d);
    *d = nullptr;
    const int n = need + need / 4 + 16;
    HIP_OK(hipMalloc((void **)d, sizeof(double) * n));
    *cap = n;
    return 0;
}

template <typename V>
static int batch_mixed_run(Ctx *const *cs, int ns, const SolveOpts *opts, SolveResult *res, hipStream_t s)
{
    BatchBufs &B = g_batch;
    BatchMixedBufs &M = g_bmixed;
    const bool coded = sizeof(V) == sizeof(float) && opts[0].precision == 3;
    int tot_blk = 0, tot_pb = 0, tot_ax = 0, hmax = 0, max_iter = 0, check_every = 16;
    for (int k = 0; k < ns; k++) {
        const int n = cs[k]->pk->n_eq;
        tot_blk += (int)cg_grid(n);
        tot_pb += (int)pu_grid_b(cg_grid(n));
        tot_ax += n > 0 ? (int)vec_grid(n) : 0;
        if (cs[k]->pk->hmax > hmax) hmax = cs[k]->pk->hmax;
        if (opts[k].max_iter > max_iter) max_iter = opts[k].max_iter;
        if (opts[k].check_every > 0) check_every = opts[k].check_every;
    }
    int rc = reserve_pair(&M.sc, &M.hsc, &M.cap_sc, (int)(sizeof(BMScene<V>) * ns));
    if (!rc) rc = reserve_pair(&M.pbs, &M.hpbs, &M.cap_pb, tot_pb + 1);
    if (!rc) rc = reserve_pair(&B.bs, &B.hbs, &B.cap_blk, tot_blk + 1);
    if (!rc) rc = reserve_pair(&B.axs, &B.haxs, &B.cap_ax, tot_ax + 1);
    if (!rc) rc = reserve_pair(&M.ctrl, &M.hctrl, &M.cap_ctrl, ns);
    if (!rc) rc = reserve_dev(&B.part, &B.cap_part, tot_blk + 1);
    if (!rc) rc = reserve_dev(&M.part_ru, &M.cap_ru, tot_blk + 1);
    if (!rc) rc = reserve_dev(&B.part2, &B.cap_part2, tot_ax + 1);
    if (!rc) rc = reserve_dev(&M.part3, &M.cap_part3, tot_ax + 1);
    if (rc) return rc;
    if (!B.ndone) {
        HIP_OK(hipMalloc(&B.ndone, 2 * sizeof(int)));   // [0] stopped scenes, [1] iteration
        HIP_OK(hipHostMalloc(&B.hndone, sizeof(int), hipHostMallocDefault));
    }
    BMScene<V> *hsc = reinterpret_cast<BMScene<V> *>(M.hsc);
    const BMScene<V> *dsc = reinterpret_cast<const BMScene<V> *>(M.sc);
    int bo = 0, po = 0, ao = 0;
    for (int k = 0; k < ns; k++) {
        Ctx *c = cs[k];
        const Packed *pk = c->pk;
        const SolveOpts &o = opts[k];
        BMScene<V> &S = hsc[k];
        S.P = pk_view(pk);
        S.b = pk->b;
        S.x = pk->x;
        if constexpr (sizeof(V) == sizeof(float)) {
            S.r = c->fr;
            S.p = c->fp;
            S.Ap = c->fAp;
        } else {
            S.r = c->r;
            S.p = c->pv;
            S.Ap = c->Ap;
        }
        S.cq = pk->cq;
        S.qk.neg_coeff = (float)o.neg_coeff;
        S.qk.re = (float)o.re_lap;
        S.qk.r_lo = (float)o.r_min;
        S.qk.dr = (float)((o.re_lap - o.r_min) / Q_CODE_MAX);
        S.n_eq = pk->n_eq;
        S.blk_off = bo;
        S.nblk = (int)cg_grid(pk->n_eq);
        S.pb_off = po;
        S.npb = (int)pu_grid_b(cg_grid(pk->n_eq));
        S.ax_off = ao;
        S.nax = pk->n_eq > 0 ? (int)vec_grid(pk->n_eq) : 0;
        S.max_iter = o.max_iter;
        S.tol_mode = o.tol_mode;
        S.tol = o.tol;
        S.delta = o.inner_tol > 0 ? o.inner_tol : 1e-3;
        S.neg_coeff = o.neg_coeff;
        S.re = o.re_lap;
        S.r_min = o.r_min;
        for (int q = 0; q < S.nblk; q++) B.hbs[bo + q] = k;
        for (int q = 0; q < S.npb; q++) M.hpbs[po + q] = k;
        for (int q = 0; q < S.nax; q++) B.haxs[ao + q] = k;
        bo += S.nblk;
        po += S.npb;
        ao += S.nax;
    }
    HIP_OK(hipMemcpyAsync(M.sc, M.hsc, sizeof(BMScene<V>) * ns, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(M.pbs, M.hpbs, sizeof(int) * (tot_pb + 1), hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(B.bs, B.hbs, sizeof(int) * (tot_blk + 1), hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(B.axs, B.haxs, sizeof(int) * (tot_ax + 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_zero_int, dim3(1), dim3(64), 0, s, B.ndone, 2);
    if (coded && tot_blk > 0)
        hipLaunchKernelGGL(k_bm_q16_build, dim3(tot_blk), dim3(CG_BLK), 0, s,
                           reinterpret_cast<const BMScene<float> *>(dsc), B.bs);
    if (tot_ax > 0) hipLaunchKernelGGL(k_bm_init<V>, dim3(tot_ax), dim3(CG_BLK), 0, s, dsc, B.axs, B.part2, M.part3);
    hipLaunchKernelGGL(k_bm_init_fin<V>, dim3(ns), dim3(CG_BLK), 0, s, dsc, B.part2, M.part3, M.ctrl, B.ndone);
    LAUNCH_CHECK();
    typedef void (*bm_fn)(const BMScene<V> *, const int *, double *, const MixedCtrl *);
    const int sv = opts[0].spmv_variant;
    bm_fn spmv = sv == 1 ? k_bm_spmv<16, V, false, true> : (sv == 2 ? k_bm_spmv<16, V, false, false> : k_bm_spmv<8, V, false, true>);
    if constexpr (sizeof(V) == sizeof(float))
        if (coded)
            spmv = sv == 1 ? k_bm_spmv<16, V, true, true> : (sv == 2 ? k_bm_spmv<16, V, true, false> : k_bm_spmv<8, V, true, true>);
    const size_t lds = sizeof(V) * (size_t)hmax;
    // the host stops when every scene has stopped: watch ndone == ns
    rc = batch_loop(s, max_iter, check_every, ns, B.ndone, B.hndone, opts[0].graph, [&]() {
        prof_begin(PK_SPMV, s);
        if (tot_blk > 0) hipLaunchKernelGGL(spmv, dim3(tot_blk), dim3(CG_BLK), lds, s, dsc, B.bs, B.part, M.ctrl);
        prof_end(PK_SPMV, s);
        hipLaunchKernelGGL(k_bm_fin1<V>, dim3(ns), dim3(CG_BLK), 0, s, dsc, B.part, M.part_ru, M.ctrl, B.ndone);
        prof_begin(PK_AXPY, s);
        if (tot_ax > 0) hipLaunchKernelGGL(k_bm_axpy<V>, dim3(tot_ax), dim3(CG_BLK), 0, s, dsc, B.axs, B.part2, M.ctrl);
        prof_end(PK_AXPY, s);
        hipLaunchKernelGGL(k_bm_fin2<V>, dim3(ns), dim3(CG_BLK), 0, s, dsc, B.part2, M.ctrl, B.ndone);
        prof_begin(PK_PUPDATE, s);
        if (tot_pb > 0)
            hipLaunchKernelGGL((k_bm_pupdate<8, V>), dim3(tot_pb), dim3(CG_BLK), 0, s, dsc, M.pbs, M.part_ru, M.ctrl,
                               B.ndone);
        prof_end(PK_PUPDATE, s);
    });
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(s));
    HIP_OK(hipMemcpy(M.hctrl, M.ctrl, sizeof(MixedCtrl) * ns, hipMemcpyDeviceToHost));
    for (int q = 0; q < ns; q++) {
        const MixedCtrl &c = M.hctrl[q];
        res[q].iters = c.done ? c.iters : opts[q].max_iter;
        res[q].breakdown = c.breakdown;
        res[q].res = c.res;
        res[q].bnorm = c.bnorm;
        res[q].outer = c.n_ru;
    }
    return 0;
}

// Mixed-precision counterpart of batch_cg (opts[*].precision 1-3, all
// alike): each scene's system packed by pack_ppe(c, 1); solutions stay in
// each scene's packed x.
int batch_cg_mixed(Ctx *const *cs, int ns, const SolveOpts *opts, SolveResult *res, hipStream_t s)
{
    const int prec = opts[0].precision;
    for (int k = 0; k < ns; k++) {
        if (opts[k].precision != prec || prec < 1 || prec > 3) {
            mps_set_error(MPS_ERR_ARG, __FILE__, __LINE__, "batched mixed solve: one precision (1-3) for all scenes");
            return MPS_ERR_ARG;
        }
        int rc = mixed_reserve(cs[k], prec);
        if (!rc && prec == 3) rc = coded_reserve(cs[k]);
        if (rc) return rc;
    }
    if (prec >= 2) return batch_mixed_run<float>(cs, ns, opts, res, s);
    return batch_mixed_run<double>(cs, ns, opts, res, s);
}

// Frees the batched solvers' block tables, partials and control blocks
// (mps_hip_finalize); the next batch allocates them again.
void cg_batch_free()
{
    BatchBufs &B = g_batch;
    void *dev[] = {B.sc, B.bs, B.axs, B.part, B.part2, B.ctrl, B.ndone};
    void *host[] = {B.hsc, B.hbs, B.haxs, B.hctrl, B.hndone};
    for (void *p : dev)
        if (p) hipFree(p);
    for (void *p : host)
        if (p) hipHostFree(p);
    B = BatchBufs();
    BatchMixedBufs &M = g_bmixed;
    void *mdev[] = {M.sc, M.pbs, M.ctrl, M.part3, M.part_ru};
    void *mhost[] = {M.hsc, M.hpbs, M.hctrl};
    for (void *p : mdev)
        if (p) hipFree(p);
    for (void *p : mhost)
        if (p) hipHostFree(p);
    M = BatchMixedBufs();
}
