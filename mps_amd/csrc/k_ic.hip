// k_ic.hip -- ICCG with the reference's exact IC(0) factor, on gfx950.
//
// The reference factors A in its own unknown order ("eq order": ascending
// particle index over unknowns, pressure_solver.c:411-422) and its factor
// values depend on that order, so this path rebuilds A in eq order with
// columns sorted (the reference's CSR, :437-551), then:
//   factor    IC(0) row by row (pressure_solver.c:156-232 restated up-
//             looking; same operations in the same order): entry (e,k) =
//             (A_ek - sum_j L_ej L_kj) * dinv_k over the common columns j in
//             ascending order, d_e = A_ee - sum L_ep^2, d <= 0 -> A_ee (if
//             > 0) or 1e-30, dinv_e = 1/sqrt(d)
//   forward   y_e = (r_e - sum_p L_ep y_col) * dinv_e         (:235-243)
//   backward  z_j = (y_j - sum_{i>j, descending} L_ij z_i) * dinv_j, the
//             reference's column scatter (:246-255) as a gather over the
//             columns of L (the pattern is symmetric)
// The 2-D program's dense formulas (mps_fluid2D/src/pressure_solver.c:
// 86-141: sums first, divide by L_ii) are selected by `dense`.
//
// All three are dependency chains: a row's level is 1 + the largest level of
// its strictly-lower columns, about 9*nx + 3*ny + nz levels in the
// reference's lattice order (2,307 for the 1.08M-particle dam break).  A
// level pass computes them, and every kernel claims rows in level order, so
// the rows in flight sit on the dependency frontier.  All run sync-free,
// with every wait bounded:
//
//   factor   one wavefront per row, rows claimed in level order from an
//            atomic ticket (so a row only waits on rows of waves already
//            running).  Lane t owns entry t of the row, (e, k).  Two forms:
//            one-phase (any row length: lane t streams row k once it is
//            published and folds the products in ascending column order)
//            and two-phase (rows of at most 64 lower entries: the wave
//            loads each published row k whole, keeps every shared column
//            in lane t's match list in LDS, then lane t folds its list in
//            order).  The row's values live in LDS, so the in-row chain
//            costs LDS latency, not a memory round trip.
//   solves   chunks of consecutive positions of the solve order (buckets
//            of levels) claimed by workgroups in order (the backward
//            solve from the end); inside a chunk the rows go by level,
//            one wavefront per row.  Lane j loads entry j and the value it
//            multiplies (from LDS for rows of the chunk) in one round
//            trip, and the products are folded across the lanes in entry
//            order up to the first value not yet published, which its
//            lane polls for.
// Values published to other waves use agent-scope stores and loads (they
// bypass the per-XCD L2) over a sentinel NaN pattern: a published value is
// its own ready flag, so no fence orders a flag after its data.  Every wait
// counts its unsuccessful polls; past the option "ic_wait_limit" the wait
// gives up (the term is dropped), an error flag is set and the solve reports
// MPS_ERR_SOLVER, so every kernel always drains (in the solves the flag
// makes every other wait give up as well).
#include <limits.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "ctx.hpp"

#define IC_BLK 256
#define FB 16                 // row-k entries per batch in the factor
#define SB 16                 // row entries per batch in the solves
#define LDS_ROW_MAX 5400      // longest strictly-lower row the factor holds in LDS

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

__device__ __forceinline__ double ld_rlx(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The error flag another wave may have raised (a wait that gave up): then
// every other wait gives up too, so the kernel drains at once.
__device__ __forceinline__ bool err_raised(const int *err)
{
    return __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// Orders LDS accesses of the lanes of one wavefront around this point (a
// compiler barrier plus the wave's LDS counter).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned long long ballot64(bool p) { return (unsigned long long)__ballot(p); }

// LDS values another wave (or lane) writes: workgroup-scope atomic accesses,
// so the compiler keeps them as LDS operations and never caches them.
template <typename T>
__device__ __forceinline__ T lds_ld(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void lds_st(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Sleep of a wave that made no progress in `idle` iterations in a row: about
// 64 cycles at first, up to about 2,000 (1 us) for a long wait, so that
// waves far from the dependency frontier poll rarely.
__device__ __forceinline__ void backoff(int idle, int cap)
{
    if (idle < 4 || cap <= 0) __builtin_amdgcn_s_sleep(1);
    else if (idle < 16 || cap == 1) __builtin_amdgcn_s_sleep(4);
    else if (idle < 64 || cap == 2) __builtin_amdgcn_s_sleep(12);
    else __builtin_amdgcn_s_sleep(32);
}

struct IcData {
    int n_eq = 0, cap_n = 0;
    long long nnz = 0, cap_nnz = 0;
    // eq-order CSR of A (sorted columns, diagonal included) and the factor
    int *rp = nullptr, *col = nullptr, *dpos = nullptr, *tpos = nullptr;
    int *eq_of_slot = nullptr, *slot_of_eq = nullptr, *flag = nullptr, *eq_of_orig = nullptr, *len = nullptr;
    double *aval = nullptr, *lval = nullptr, *dinv = nullptr, *beq = nullptr, *xeq = nullptr;
    // the level order (rows by dependency level) and its inverse; the
    // rows' levels
    int *order = nullptr, *opos = nullptr;
    double *lev = nullptr;
    int *lhist = nullptr, *lstart = nullptr;
    // the solve order (rows by bucket of levels, index order inside) and
    // its inverse
    int *sperm = nullptr, *spos = nullptr;
    size_t cap_bh = 0;
    int *bh = nullptr, *boff = nullptr;
    const int *perm = nullptr;       // the order the vectors below follow (NULL: eq order)
    // PCG vectors, by position of that order
    double *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *Ap = nullptr, *y = nullptr, *b = nullptr;
    // A in sliced ELL by position (64-position slices, each row's entries in
    // CSR order, columns renamed to positions)
    int cap_sl = 0;
    int *wA = nullptr, *oA = nullptr;
    size_t cap_ea = 0;
    int *acol = nullptr;
    double *aell = nullptr;
    // L and its columns (U) by position, rows contiguous (loff, uoff)
    int *cntl = nullptr, *cntu = nullptr, *loff = nullptr, *uoff = nullptr;
    size_t cap_lu = 0;
    int *lcol = nullptr, *ucol = nullptr;
    double *lv = nullptr, *uv = nullptr;
    // device: [0] A's ELL total, [3] longest lower row, [4] largest level
    int *hdr = nullptr;
    int *hhdr = nullptr;                                 // pinned host copy
    int *scan_blk = nullptr;
    unsigned *ticket = nullptr;   // [4]
    int *err = nullptr;
    double *part = nullptr;
    int max_lower = 0, levels = 0, bucket = 0;
    // arena mode: every buffer above is carved from [arena, arena +
    // arena_cap) instead of hipMalloc'd, the host header is hhdr_host and
    // the PCG poll uses the events ev (created beforehand)
    char *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    int hhdr_host[8] = {0};
    hipEvent_t ev[2] = {nullptr, nullptr};
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

// Row e of the eq-order CSR: the off-diagonals of its unknown slot and the
// diagonal, each placed at its rank among the row's columns (columns of a
// row are distinct), so the row comes out sorted.  One wavefront per row;
// the row's columns are staged in LDS 512 at a time and every lane counts
// the smaller ones for its own entries.
#define FILL_KCHUNK 512
__global__ void __launch_bounds__(IC_BLK)
k_ic_fill_sorted(int n_eq, int stride, const int *slot_of_eq, const int *eq_of_slot, const int *Alen,
                 const double *Aval, const int *Acol, const double *diag, const double *bslot, const int *rp,
                 int *col, double *aval, int *dpos, double *beq)
{
    __shared__ int skey_all[IC_BLK / WAVE][FILL_KCHUNK];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
    const int e = blockIdx.x * (IC_BLK / WAVE) + w;
    if (e >= n_eq) return;   // whole wavefront
    int *skey = skey_all[w];
    const int s = slot_of_eq[e];
    const int m = Alen[s];   // off-diagonals; entry m is the diagonal
    const int tot = m + 1;
    const int base = rp[e];
    for (int q0 = 0; q0 < tot; q0 += WAVE) {
        const int q = q0 + lane;
        int key = INT_MAX;
        double v = 0.0;
        if (q < m) {
            const size_t o = sell_off(s, q, stride);
            key = eq_of_slot[Acol[o]];
            v = Aval[o];
        } else if (q == m) {
            key = e;
            v = diag[s];
        }
        int rank = 0;
        for (int c0 = 0; c0 < tot; c0 += FILL_KCHUNK) {
            const int cn = min(FILL_KCHUNK, tot - c0);
            wave_sync();
            for (int t = lane; t < cn; t += WAVE) {
                const int qq = c0 + t;
                skey[t] = qq < m ? eq_of_slot[Acol[sell_off(s, qq, stride)]] : e;
            }
            wave_sync();
            if (q < tot)
                for (int t = 0; t < cn; t++) rank += skey[t] < key;
        }
        if (q < tot) {
            col[base + rank] = key;
            aval[base + rank] = v;
            if (q == m) dpos[e] = rank;
        }
    }
    if (lane == 0) beq[e] = bslot[s];
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

// Diagonal position of each row for a caller's CSR (columns sorted).
__global__ void k_ic_dpos(int n, const int *rp, const int *col, int *dpos)
{
    const int i = blockIdx.x * IC_BLK + threadIdx.x;
    if (i >= n) return;
    int q = rp[i];
    while (q < rp[i + 1] && col[q] < i) q++;
    dpos[i] = q - rp[i];
}

// ---------------------------------------------------------------------
// sliced-ELL copies
// ---------------------------------------------------------------------

__global__ void k_ic_hdr_reset(int *hdr, int count)
{
    if ((int)threadIdx.x < count) hdr[threadIdx.x] = 0;
}

// Rows are processed in a permuted order: position t holds row perm[t]
// (perm NULL: the identity), and pos[r] is the position of row r.  The
// sliced-ELL copies are laid out by position, with columns renamed to
// positions; each row keeps its entries in the CSR's order, so every sum
// runs as before.
__device__ __forceinline__ int row_at(const int *perm, int t) { return perm ? perm[t] : t; }
__device__ __forceinline__ int pos_of(const int *pos, int r) { return pos ? pos[r] : r; }

// Width of each 64-position slice of A's sliced-ELL copy; hdr[3] = longest
// strictly-lower row.
__global__ void __launch_bounds__(IC_BLK)
k_ic_slice_w(int n, const int *perm, const int *rp, const int *dpos, int *wA, int *hdr)
{
    const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
    const int sl = blockIdx.x * (IC_BLK / WAVE) + w;
    if (sl * WAVE >= n) return;
    const int t = sl * WAVE + lane;
    int la = 0, ll = 0;
    if (t < n) {
        const int r = row_at(perm, t);
        la = rp[r + 1] - rp[r];
        ll = dpos[r];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        la = max(la, __shfl_xor(la, o, WAVE));
        ll = max(ll, __shfl_xor(ll, o, WAVE));
    }
    if (lane == 0) {
        wA[sl] = la;
        atomicMax(&hdr[3], ll);
    }
}

// A in sliced ELL by position (for the SpMV).
__global__ void k_ic_fill_a(int n, const int *perm, const int *pos, const int *rp, const int *col,
                            const double *aval, const int *oA, int *acol, double *aell)
{
    const int t = blockIdx.x * IC_BLK + threadIdx.x;
    if (t >= n) return;
    const int r = row_at(perm, t);
    const int r0 = rp[r], m = rp[r + 1] - r0;
    const size_t b = (size_t)oA[t >> 6] * WAVE + (t & (WAVE - 1));
    for (int q = 0; q < m; q++) {
        acol[b + (size_t)q * WAVE] = pos_of(pos, col[r0 + q]);
        aell[b + (size_t)q * WAVE] = aval[r0 + q];
    }
}

// The factor's values (strictly lower entries, 1/sqrt(d)) reset to "not
// computed".
__global__ void k_ic_reset_factor(int n, const int *rp, const int *dpos, double *lval, double *dinv)
{
    const int e = blockIdx.x * IC_BLK + threadIdx.x;
    if (e >= n) return;
    for (int q = 0; q < dpos[e]; q++) lval[rp[e] + q] = sent();
    dinv[e] = sent();
}

// L and its columns (U) by position from the finished factor, each row's
// entries contiguous (a wavefront reads a row at once): row t of L at
// loff[t], its columns in ascending eq order; row t of U (L_ij, i > j) at
// uoff[t], descending i for the sparse formulas, ascending for the dense
// ones; columns renamed to positions.
__global__ void k_ic_fill_lu(int n, const int *perm, const int *pos, const int *rp, const int *col,
                             const int *dpos, const int *tpos, const double *lval, const int *loff, const int *uoff,
                             int *lcol, double *lv, int *ucol, double *uv, int dense)
{
    const int t = blockIdx.x * IC_BLK + threadIdx.x;
    if (t >= n) return;
    const int j = row_at(perm, t);
    const int r0 = rp[j], m = dpos[j], nu = rp[j + 1] - r0 - m - 1;
    const int bl = loff[t], bu = uoff[t];
    for (int q = 0; q < m; q++) {
        lcol[bl + q] = pos_of(pos, col[r0 + q]);
        lv[bl + q] = lval[r0 + q];
    }
    for (int u = 0; u < nu; u++) {
        const int p = r0 + m + 1 + u;   // (j, i), i ascending
        const int tp = tpos[p];
        const int slot = dense ? u : nu - 1 - u;
        ucol[bu + slot] = pos_of(pos, col[p]);
        uv[bu + slot] = tp >= 0 ? lval[tp] : 0.0;   // asymmetric pattern: no (i, j)
    }
}

__global__ void k_ic_gather(int n, const int *perm, const double *src, double *dst)
{
    const int t = blockIdx.x * IC_BLK + threadIdx.x;
    if (t < n) dst[t] = src[row_at(perm, t)];
}

__global__ void k_ic_scatter(int n, const int *perm, const double *src, double *dst)
{
    const int t = blockIdx.x * IC_BLK + threadIdx.x;
    if (t < n) dst[row_at(perm, t)] = src[t];
}

// ---------------------------------------------------------------------
// the level order
// ---------------------------------------------------------------------

__global__ void k_ic_lev_hist(int n, const double *lev, int *hist, int *hdr)
{
    const int e = blockIdx.x * IC_BLK + threadIdx.x;
    if (e >= n) return;
    const int L = (int)lev[e];
    atomicAdd(&hist[L], 1);
    atomicMax(&hdr[4], L);
}

__global__ void k_ic_lev_scatter(int n, const double *lev, int *cursor, int *fperm)
{
    const int e = blockIdx.x * IC_BLK + threadIdx.x;
    if (e >= n) return;
    fperm[atomicAdd(&cursor[(int)lev[e]], 1)] = e;
}

// Inverse of the level order: the position of each row.
__global__ void k_ic_inverse(int n, const int *perm, int *pos)
{
    const int t = blockIdx.x * IC_BLK + threadIdx.x;
    if (t < n) pos[perm[t]] = t;
}

// The solve order: rows by bucket of 2^gshift consecutive levels, ascending
// index inside a bucket.  Dependencies point to lower indices and lower
// levels, so this is a topological order for any bucket size, and inside a
// bucket a lattice column contributes a run of consecutive rows, whose
// dependencies along the column then stay inside one chunk of the solves.
// Per block of 256 rows (index order): the block's rows in each bucket,
// bh[b * nb + block].
__global__ void __launch_bounds__(IC_BLK)
k_ic_bucket_hist(int n, const double *lev, int gshift, int nbk, int *bh)
{
    extern __shared__ int h[];   // nbk counters
    const int t = threadIdx.x, nb = gridDim.x;
    for (int q = t; q < nbk; q += IC_BLK) h[q] = 0;
    __syncthreads();
    const int e = blockIdx.x * IC_BLK + t;
    if (e < n) atomicAdd(&h[(int)lev[e] >> gshift], 1);
    __syncthreads();
    for (int q = t; q < nbk; q += IC_BLK) bh[(size_t)q * nb + blockIdx.x] = h[q];
}

// Stable scatter: row e goes to its bucket's offset for this block plus its
// rank among the block's rows of the same bucket.
__global__ void __launch_bounds__(IC_BLK)
k_ic_bucket_scatter(int n, const double *lev, int gshift, const int *boff, int *sperm, int *spos)
{
    __shared__ int bk[IC_BLK];
    const int t = threadIdx.x, nb = gridDim.x;
    const int e = blockIdx.x * IC_BLK + t;
    const int b = e < n ? ((int)lev[e] >> gshift) : -1;
    bk[t] = b;
    __syncthreads();
    if (e >= n) return;
    int rank = 0;
    for (int u = 0; u < t; u++) rank += bk[u] == b;
    const int p = boff[(size_t)b * nb + blockIdx.x] + rank;
    sperm[p] = e;
    spos[e] = p;
}

// Entry counts by position: strictly lower (L) and strictly upper (L's
// columns) of the row at each position.
__global__ void k_ic_counts(int n, const int *perm, const int *rp, const int *dpos, int *cl, int *cu)
{
    const int t = blockIdx.x * IC_BLK + threadIdx.x;
    if (t >= n) return;
    const int r = row_at(perm, t);
    cl[t] = dpos[r];
    cu[t] = rp[r + 1] - rp[r] - dpos[r] - 1;
}

// ---------------------------------------------------------------------
// factor: one wavefront per row
// ---------------------------------------------------------------------

// d_e of a finished row from its values in LDS, in column order (every lane
// computes it: d = A_ee - sum L^2, d <= 0 -> A_ee (if > 0) or 1e-30), and
// its 1/sqrt(d) (dense formulas: sqrt(d)) published after the row's values
// (readers check each value anyway).
__device__ __forceinline__ void finish_row(int e, int m, double akk, const double *sval, double *dinv, int dense,
                                           int lane)
{
    double d;
    if (!dense) {
        d = akk;
        for (int q = 0; q < m; q++) {
            const double l = lds_ld(&sval[q]);
            d -= l * l;
        }
    } else {
        double sq = 0.0;
        for (int q = 0; q < m; q++) {
            const double l = lds_ld(&sval[q]);
            sq += l * l;
        }
        d = akk - sq;
    }
    if (d <= 0.0) d = (akk > 0.0) ? akk : 1.0e-30;
    const double out = dense ? sqrt(d) : 1.0 / sqrt(d);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (lane == 0) st_rlx(&dinv[e], out);
    wave_sync();
}

// ---------------------------------------------------------------------
// factor, one-phase form (any row length)
// ---------------------------------------------------------------------

// One row e of the factor (one-phase form, any row length), claimed by the
// calling wave: lane T owns entries T, T + 64, ... in groups of 64; row e's
// columns and finished values sit in scol/sval (LDS, m entries each).
__device__ __forceinline__ void factor_row_any(int e, int lane, const int *__restrict__ rp, const int *__restrict__ col,
                                               const int *__restrict__ dpos, const double *__restrict__ aval,
                                               double *lval, double *dinv, int dense, long long wait_limit,
                                               int sleep_cap, int *scol, double *sval, bool &failed)
{
    const int r0 = rp[e], m = dpos[e];
    wave_sync();
    for (int t = lane; t < m; t += WAVE) scol[t] = col[r0 + t];
    wave_sync();
    const int pstep = m > 0 ? (1 << (31 - __clz(m))) : 0;
    for (int g0 = 0; g0 < m; g0 += WAVE) {
        const int T = g0 + lane;
        const bool act = T < m;
        int k = 0, kb = 0, mk = 0;
        double a = 0.0;
        if (act) {
            k = scol[T];
            a = aval[r0 + T];
            kb = rp[k];
            mk = dpos[k];
        }
        int stage = act ? 0 : 2;   // 0: waiting for row k, 1: folding, 2: done
        double dk = 0.0, s = dense ? 0.0 : a;
        int u = 0, bn = 0, bi = 0;
        int bpos[FB];
        double bval[FB];
#pragma unroll
        for (int j = 0; j < FB; j++) { bpos[j] = -1; bval[j] = 0.0; }
        long long spins = 0;
        bool poll_all = true;
        int idle = 0;
        for (;;) {
            if (ballot64(stage != 2) == 0ull) break;
            const unsigned long long donem = ballot64(stage == 2);
            const unsigned long long waitm = ballot64(stage == 0);
            bool prog = false;
            if (waitm) {
                const int lo = __ffsll((long long)waitm) - 1;
                if (stage == 0 && (poll_all || lane == lo)) {
                    const double d = ld_rlx(&dinv[k]);
                    if (!is_sent(d)) {
                        dk = d;
                        stage = 1;
                        prog = true;
                        spins = 0;
                    } else if (++spins > wait_limit) {
                        failed = true;   // give up on row k: its terms read as 0
                        stage = 1;
                    }
                }
                poll_all = __builtin_amdgcn_readlane((int)(stage != 0), lo) != 0;
            }
            if (stage == 1 && bi == bn) {
                if (u >= mk) {
                    double L;
                    if (!dense) L = s * dk;
                    else L = (fabs(a) < 1.0e-30) ? 0.0 : (a - s) * (1.0 / dk);   // dk holds sqrt(d) here
                    lds_st(&sval[T], L);
                    st_rlx(&lval[r0 + T], L);
                    stage = 2;
                    prog = true;
                } else {
                    bn = min(FB, mk - u);
                    int bc[FB];
#pragma unroll
                    for (int j = 0; j < FB; j++) {
                        bc[j] = INT_MAX;
                        if (j < bn) bc[j] = col[kb + u + j];
                    }
#pragma unroll
                    for (int j = 0; j < FB; j++) bval[j] = j < bn ? lval[kb + u + j] : 0.0;   // cached; see k_ic_factor2
                    // row k's columns are all below k = scol[T]: find each
                    // among this row's columns 0..T-1 (last index < c)
                    int ps[FB];
#pragma unroll
                    for (int j = 0; j < FB; j++) ps[j] = -1;
                    for (int st = pstep; st > 0; st >>= 1) {
#pragma unroll
                        for (int j = 0; j < FB; j++) {
                            const int cand = ps[j] + st;
                            if (cand < T && scol[cand] < bc[j]) ps[j] = cand;
                        }
                    }
#pragma unroll
                    for (int j = 0; j < FB; j++) {
                        const int cand = ps[j] + 1;
                        bpos[j] = (j < bn && cand < T && scol[cand] == bc[j]) ? cand : -1;
                    }
                    bi = 0;
                    u += bn;
                    prog = true;
                }
            }
            if (stage == 1 && bi < bn) {
                bool stop = false;
#pragma unroll
                for (int j = 0; j < FB; j++) {
                    if (!stop && j == bi && j < bn) {
                        const int v = bpos[j];
                        if (v < 0) {
                            bi++;
                            prog = true;
                        } else {
                            double lk = bval[j];
                            if (is_sent(lk)) {
                                lk = ld_rlx(&lval[kb + u - bn + j]);
                                bval[j] = lk;
                            }
                            const bool ready = !is_sent(lk) && (v < g0 || ((donem >> (v - g0)) & 1ull));
                            if (ready) {
                                const double lv = lds_ld(&sval[v]);
                                if (!dense) s -= lv * lk;
                                else s += lv * lk;
                                bi++;
                                prog = true;
                                spins = 0;
                            } else if (++spins > wait_limit) {
                                failed = true;   // give up on this term
                                bi++;
                            } else {
                                stop = true;
                            }
                        }
                    }
                }
            }
            if (__any(prog)) idle = 0;
            else backoff(idle++, sleep_cap);
        }
        wave_sync();
    }
    finish_row(e, m, aval[r0 + m], sval, dinv, dense, lane);
}

__global__ void __launch_bounds__(IC_BLK)
k_ic_factor(int n, const int *__restrict__ order, const int *__restrict__ rp, const int *__restrict__ col,
            const int *__restrict__ dpos, const double *__restrict__ aval, double *lval, double *dinv,
            unsigned *ticket, int *err, int dense, int lcap, long long wait_limit, int sleep_cap)
{
    extern __shared__ int f_smem[];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
    int *scol = f_smem + (size_t)w * lcap * 3;
    double *sval = (double *)(f_smem + (size_t)w * lcap * 3 + lcap);
    bool failed = false;
    for (;;) {
        int tk = 0;
        if (lane == 0) tk = (int)atomicAdd(ticket, 1u);
        tk = __builtin_amdgcn_readlane(tk, 0);
        if (tk >= n) break;
        const int e = order[tk];   // rows in level order: every row waited on is claimed earlier
        factor_row_any(e, lane, rp, col, dpos, aval, lval, dinv, dense, wait_limit, sleep_cap, scol, sval, failed);
    }
    if (ballot64(failed) && lane == 0) atomicOr(err, 1);
}

// Two-phase form of the factor for rows of at most 64 strictly-lower
// entries (lcap <= 64, its LDS layout), one wavefront per workgroup; a
// longer row (up to lcap_all entries) takes the one-phase form's row in the
// same wave, with the same arithmetic, its columns and values in the same
// LDS.  Lane t owns entry t,
// (e, k).  Once row k is published the whole wave loads row k (its at most
// 64 lower entries, entry j in lane j: two coalesced loads), up to FG2 such
// rows in flight, and keeps each match -- the position v in this row of a
// column the two rows share, and L_kj -- in lane t's list in LDS, in row k's
// entry order (lane t has at most t matches: list t starts at t (t - 1) / 2),
// without waiting on its own row.  Lane t then folds its list in order, each
// product as soon as lane v of this row is done, all from LDS.  A row's
// in-lane chain costs LDS iterations instead of a memory round trip per
// link, and a gathered row costs a few cache lines instead of one per entry
// and lane; the arithmetic is the one-phase kernel's, term for term.
#define FG2 4
__global__ void __launch_bounds__(WAVE)
k_ic_factor2(int n, const int *__restrict__ order, const int *__restrict__ rp, const int *__restrict__ col,
             const int *__restrict__ dpos, const double *__restrict__ aval, double *lval, double *dinv,
             unsigned *ticket, int *err, int dense, int lcap, int lcap_all, long long wait_limit, int sleep_cap)
{
    extern __shared__ double f2_smem[];
    const int lane = threadIdx.x;
    const int tri = lcap * (lcap - 1) / 2;
    double *sval = f2_smem;                               // lcap
    double *mval = f2_smem + lcap;                        // tri
    int *scol = (int *)(f2_smem + lcap + tri);            // lcap
    unsigned short *mpos = (unsigned short *)(scol + lcap);   // tri
    bool failed = false;
    for (;;) {
        int tk = 0;
        if (lane == 0) tk = (int)atomicAdd(ticket, 1u);
        tk = __builtin_amdgcn_readlane(tk, 0);
        if (tk >= n) break;
        const int e = order[tk];   // rows in level order: every row waited on is claimed earlier
        const int r0 = rp[e], m = dpos[e];
        if (m > lcap) {
            int *gcol = (int *)f2_smem;
            factor_row_any(e, lane, rp, col, dpos, aval, lval, dinv, dense, wait_limit, sleep_cap, gcol,
                           (double *)(gcol + lcap_all), failed);
            continue;
        }
        wave_sync();
        if (lane < m) scol[lane] = col[r0 + lane];
        wave_sync();
        const int pstep = m > 0 ? (1 << (31 - __clz(m))) : 0;
        const int T = lane;
        const bool act = T < m;
        int k = 0, kb = 0, mk = 0;
        double a = 0.0;
        if (act) {
            k = scol[T];
            a = aval[r0 + T];
            kb = rp[k];
            mk = dpos[k];
        }
        const int lb = T * (T - 1) / 2;   // this lane's match list
        int stage = act ? 0 : 3;          // 0 waiting for row k, 1 gathering, 2 folding, 3 done
        double dk = 0.0, s = dense ? 0.0 : a;
        int cnt = 0, mi = 0;
        long long spins = 0;
        bool poll_all = true;
        int idle = 0;
        for (;;) {
            if (ballot64(stage != 3) == 0ull) break;
            const unsigned long long waitm = ballot64(stage == 0);
            bool prog = false;
            if (waitm) {
                const int lo = __ffsll((long long)waitm) - 1;
                if (stage == 0 && (poll_all || lane == lo)) {
                    // the first poll through the caches (a non-sentinel is final)
                    const double d = spins == 0 ? dinv[k] : ld_rlx(&dinv[k]);
                    if (!is_sent(d)) {
                        dk = d;
                        stage = 1;
                        prog = true;
                        spins = 0;
                    } else if (++spins > wait_limit) {
                        failed = true;   // give up on row k: its terms read as 0
                        stage = 1;
                    }
                }
                poll_all = __builtin_amdgcn_readlane((int)(stage != 0), lo) != 0;
            }
            // gather: row k of each lane in stage 1, FG2 rows at a time,
            // the whole wave loading one row (entry j in lane j)
            unsigned long long readym = ballot64(stage == 1);
            while (readym) {
                int tg[FG2], kbg[FG2], mkg[FG2], cj[FG2];
                double vj[FG2];
#pragma unroll
                for (int g = 0; g < FG2; g++) {
                    tg[g] = readym ? __ffsll((long long)readym) - 1 : -1;
                    readym &= readym ? readym - 1 : 0ull;
                }
#pragma unroll
                for (int g = 0; g < FG2; g++) {
                    cj[g] = INT_MAX;
                    vj[g] = 0.0;
                    kbg[g] = 0;
                    mkg[g] = 0;
                    if (tg[g] >= 0) {
                        kbg[g] = __builtin_amdgcn_readlane(kb, tg[g]);
                        mkg[g] = __builtin_amdgcn_readlane(mk, tg[g]);
                        const bool again = __builtin_amdgcn_readlane((int)(spins != 0), tg[g]) != 0;
                        if (lane < mkg[g]) {
                            cj[g] = col[kbg[g] + lane];
                            // The first load goes through the caches: a value
                            // is written once per factor (the sentinel before
                            // it by an earlier kernel, whose end made it
                            // visible to every cache), so a non-sentinel read
                            // from any cache is final.  Only a sentinel can be
                            // stale: that row is loaded again past the caches.
                            vj[g] = again ? ld_rlx(&lval[kbg[g] + lane]) : lval[kbg[g] + lane];
                        }
                    }
                }
#pragma unroll
                for (int g = 0; g < FG2; g++) {
                    const int Tg = tg[g];
                    if (Tg < 0) continue;
                    // row k in pieces of 64 entries (a row k past 64 entries
                    // is one of the long rows the one-phase form takes; its
                    // later pieces are loaded here, one at a time); each
                    // piece's matches follow the earlier ones in lane Tg's
                    // list, so the list stays in row k's entry order
                    const int mkk = mkg[g];
                    const bool again = mkk > WAVE && __builtin_amdgcn_readlane((int)(spins != 0), Tg) != 0;
                    bool missing = false;
                    int base = 0;
                    for (int c0 = 0; c0 < mkk; c0 += WAVE) {
                        int cjc = cj[g];
                        double vjc = vj[g];
                        if (c0 > 0) {
                            cjc = INT_MAX;
                            vjc = 0.0;
                            if (c0 + lane < mkk) {
                                cjc = col[kbg[g] + c0 + lane];
                                vjc = again ? ld_rlx(&lval[kbg[g] + c0 + lane]) : lval[kbg[g] + c0 + lane];
                            }
                        }
                        if (ballot64(c0 + lane < mkk && is_sent(vjc))) missing = true;
                        int ps = -1;
                        for (int st = pstep; st > 0; st >>= 1) {
                            const int cand = ps + st;
                            if (cand < Tg && scol[cand] < cjc) ps = cand;
                        }
                        ps++;
                        const bool hit = c0 + lane < mkk && ps < Tg && scol[ps] == cjc;
                        const unsigned long long hm = ballot64(hit);
                        if (hit) {
                            const int q = Tg * (Tg - 1) / 2 + base + __popcll(hm & ((1ull << lane) - 1ull));
                            mpos[q] = (unsigned short)ps;
                            mval[q] = is_sent(vjc) ? 0.0 : vjc;
                        }
                        base += __popcll(hm);
                    }
                    if (missing) {
                        // row k's 1/sqrt(d) came before some of its values:
                        // load it again next time (the list is rewritten
                        // then), or give up at the bound
                        int quit = 0;
                        if (lane == Tg) quit = ++spins > wait_limit;
                        if (!__builtin_amdgcn_readlane(quit, Tg)) continue;
                        if (lane == Tg) failed = true;   // missing values read as 0
                    }
                    if (lane == Tg) {
                        cnt = base;
                        stage = 2;
                        spins = 0;
                    }
                    prog = true;
                }
            }
            wave_sync();   // the lists before their owners read them
            // the fold chain from LDS, as many links as are ready, before
            // the next poll or gather
            for (;;) {
                const unsigned long long donem = ballot64(stage == 3);
                bool fprog = false;
                if (stage == 2) {
                    while (mi < cnt) {
                        const int v = mpos[lb + mi];
                        if (!((donem >> v) & 1ull)) break;
                        const double lv = lds_ld(&sval[v]);
                        if (!dense) s -= lv * mval[lb + mi];
                        else s += lv * mval[lb + mi];
                        mi++;
                        fprog = true;
                    }
                    if (mi == cnt) {
                        double L;
                        if (!dense) L = s * dk;
                        else L = (fabs(a) < 1.0e-30) ? 0.0 : (a - s) * (1.0 / dk);   // dk holds sqrt(d) here
                        lds_st(&sval[T], L);
                        st_rlx(&lval[r0 + T], L);
                        stage = 3;
                        fprog = true;
                    }
                }
                if (!__any(fprog)) break;
                prog = true;
            }
            if (__any(prog)) idle = 0;
            else backoff(idle++, sleep_cap);
        }
        wave_sync();
        finish_row(e, m, aval[r0 + m], sval, dinv, dense, lane);
    }
    if (ballot64(failed) && lane == 0) atomicOr(err, 1);
}

// ---------------------------------------------------------------------
// triangular solves: one workgroup per chunk of 256 rows
// ---------------------------------------------------------------------

// Dependency levels, sync-free over chunks of 256 consecutive rows (index
// order, chunks claimed in order): lev_e = 1 + the largest level among row
// e's strictly-lower columns (0 without any).  Levels of rows of the same
// chunk pass through LDS, so a lattice column's chain stays on chip; other
// levels are read from HBM, where each row publishes its level over the
// sentinel.  A row's columns are read from the eq-order CSR in batches whose
// outside levels are all requested at once.
__global__ void __launch_bounds__(IC_BLK)
k_ic_levels(int n, const int *__restrict__ rp, const int *__restrict__ dpos, const int *__restrict__ col,
            double *lev, unsigned *ticket, int *err, long long wait_limit, int sleep_cap)
{
    __shared__ double sy[IC_BLK];
    __shared__ int sflag[IC_BLK];
    __shared__ int sclaim;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    lds_st(&sflag[tid], -1);
    const int nch = (n + IC_BLK - 1) / IC_BLK;
    bool failed = false;
    for (;;) {
        __syncthreads();
        if (tid == 0) sclaim = (int)atomicAdd(ticket, 1u);
        __syncthreads();
        const int tk = sclaim;
        if (tk >= nch) break;
        const int base = tk * IC_BLK;
        const int e = base + tid;
        bool done = e >= n;
        int m = 0;
        size_t eb = 0;
        double s = -1.0;
        if (!done) {
            m = dpos[e];
            eb = (size_t)rp[e];
        }
        int q = 0, bn = 0, bi = 0;
        int bc[SB];
        double bv[SB];
#pragma unroll
        for (int j = 0; j < SB; j++) { bc[j] = 0; bv[j] = 0.0; }
        long long spins = 0;
        int idle = 0;
        for (;;) {
            if (ballot64(!done) == 0ull) break;
            bool prog = false;
            if (!done && bi == bn) {
                if (q >= m) {
                    const double out = s + 1.0;
                    lds_st(&sy[tid], out);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");   // value before flag
                    lds_st(&sflag[tid], tk);
                    st_rlx(&lev[e], out);
                    done = true;
                    prog = true;
                } else {
                    bn = min(SB, m - q);
#pragma unroll
                    for (int j = 0; j < SB; j++)
                        if (j < bn) bc[j] = col[eb + (size_t)(q + j)];
#pragma unroll
                    // first loads through the caches: lev held the
                    // sentinel before this kernel and each level is written
                    // once, so only a sentinel can be stale (polled again
                    // past the caches below)
                    for (int j = 0; j < SB; j++) bv[j] = (j < bn && bc[j] < base) ? lev[bc[j]] : 0.0;
                    q += bn;
                    bi = 0;
                    prog = true;
                }
            }
            bool blocked = false;   // waits on a level outside the chunk
            if (!done && bi < bn) {
                bool stop = false;
#pragma unroll
                for (int j = 0; j < SB; j++) {
                    if (!stop && j == bi && j < bn) {
                        const int c = bc[j];
                        const bool ext = c < base;
                        if (ext && is_sent(bv[j])) bv[j] = ld_rlx(&lev[c]);
                        bool ready;
                        double v;
                        if (ext) {
                            v = bv[j];
                            ready = !is_sent(v);
                        } else {
                            ready = lds_ld(&sflag[c - base]) == tk;
                            v = 0.0;
                            if (ready) {
                                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                                v = lds_ld(&sy[c - base]);
                            }
                        }
                        if (ready) {
                            s = fmax(s, v);
                            bi++;
                            prog = true;
                            spins = 0;
                        } else if (++spins > wait_limit) {
                            failed = true;   // give up on this column
                            bi++;
                        } else {
                            stop = true;
                            blocked = ext;
                        }
                    }
                }
            }
            if (__any(prog)) idle = 0;
            else backoff(idle++, sleep_cap);
            (void)blocked;
        }
    }
    if (ballot64(failed) && lane == 0) atomicOr(err, 1);
}

// Both halves of a 64-bit value from lane k (k uniform).
__device__ __forceinline__ double readlane_d(double v, int k)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), k);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), k);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Substitution over chunks of CS_CHUNK consecutive positions of the solve
// order, one workgroup per chunk (chunks claimed in order; BACK: from the
// end).  The workgroup sorts its chunk's rows by dependency level (BACK:
// descending), and its wavefronts claim rows one at a time from that list,
// so the lattice columns of a chunk advance together, level by level, as
// the DAG lets them.  Position t holds row perm[t] (level lev[perm[t]]); its
// entries sit at eoff[t] .. eoff[t+1] (columns are positions).  Lane j
// loads entry j and the value it multiplies -- from LDS when that row is in
// the chunk, from HBM otherwise -- all in one round trip; the products are
// then folded in entry order across the lanes (every lane holds the running
// sum) up to the first value not yet published, which its lane polls for.
// MODE 0 sparse: s = r, s -= l v, out = s * dinv; MODE 1 dense: s = 0,
// s += l v, out = (r - s) / dinv (dinv holds L_ii).  rin and y are by
// position.
template <int MODE, bool BACK, int CS_CHUNK, int CS_WAVES>
__global__ void __launch_bounds__(CS_WAVES * WAVE)
k_ic_csolve(int n, const int *__restrict__ perm, const double *__restrict__ lev, const int *__restrict__ eoff,
            const int *__restrict__ ecol, const double *__restrict__ eval, const double *__restrict__ dinv,
            const double *__restrict__ rin, double *y, unsigned *ticket, int *err, long long wait_limit,
            int sleep_cap, const IcCtrl *ctrl)
{
    // the chunk's entries (columns and values, one contiguous range),
    // right-hand side, 1/sqrt(d) and entry offsets, staged in one pass when
    // the chunk is claimed (chunks of at most ECAP entries; a larger chunk
    // reads its entries from HBM row by row), so that a row's loads are off
    // the chain of rows that wait on one another
    constexpr int ECAP = CS_CHUNK <= 32 ? 2048 : CS_CHUNK <= 64 ? 4096 : 8192;
    __shared__ int sec[ECAP];
    __shared__ double sev[ECAP];
    __shared__ double srin[CS_CHUNK], sdr[CS_CHUNK];
    __shared__ int seo[CS_CHUNK + 1];
    __shared__ double sy[CS_CHUNK];
    __shared__ int sflag[CS_CHUNK];
    __shared__ int skey[CS_CHUNK];
    // by rank in dependency order: the row (chunk offset), its entry range,
    // right-hand side and 1/sqrt(d), read together when a wave claims it
    struct RowInfo {
        int me, o0, o1, pad;
        double rv, dr;
    };
    __shared__ RowInfo sri[CS_CHUNK];
    __shared__ int sclaim, snext;
    if (ctrl && ctrl->done) return;
    const int lane = threadIdx.x & (WAVE - 1);
    if (threadIdx.x < CS_CHUNK) lds_st(&sflag[threadIdx.x], -1);
    const int nch = (n + CS_CHUNK - 1) / CS_CHUNK;
    bool failed = false;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) {
            sclaim = (int)atomicAdd(ticket, 1u);
            snext = 0;
        }
        __syncthreads();
        // the claim, the row and its entry range are uniform over the wave;
        // read into scalar registers, the term loops below run on scalar
        // control instead of exec masks
        const int tk = __builtin_amdgcn_readfirstlane(sclaim);
        if (tk >= nch) break;
        const int ch = BACK ? nch - 1 - tk : tk;
        const int cbase = ch * CS_CHUNK, cnt_rows = min(CS_CHUNK, n - cbase);
        // the chunk's rows in dependency order: by level, then position
        // (BACK: both descending); ranks are distinct, so sri is filled
        // once for each row of the chunk
        if (threadIdx.x < cnt_rows) {
            const int t = cbase + threadIdx.x;
            const int r = perm ? perm[t] : t;
            const int L = (int)lev[r];
            skey[threadIdx.x] = BACK ? -L : L;
            srin[threadIdx.x] = rin[t];
            sdr[threadIdx.x] = dinv[r];
        }
        if (threadIdx.x <= cnt_rows) seo[threadIdx.x] = eoff[cbase + threadIdx.x];
        const int e0 = eoff[cbase], ne = eoff[cbase + cnt_rows] - e0;
        const bool staged = ne <= ECAP;   // uniform over the block
        if (staged) {
            for (int q = threadIdx.x; q < ne; q += CS_WAVES * WAVE) {
                const int c = ecol[e0 + q];
                sec[q] = c;
                sev[q] = eval[e0 + q];
            }
        }
        __syncthreads();
        if (threadIdx.x < cnt_rows) {
            const int me = threadIdx.x, km = skey[me];
            int rank = 0;
            for (int u = 0; u < cnt_rows; u++) {
                const int ku = skey[u];
                rank += ku < km || (ku == km && (BACK ? u > me : u < me));
            }
            sri[rank] = RowInfo{me, seo[me], seo[me + 1], 0, srin[me], sdr[me]};
        }
        __syncthreads();
        for (;;) {
            int i = 0;
            if (lane == 0) i = atomicAdd(&snext, 1);
            i = __builtin_amdgcn_readlane(i, 0);
            if (i >= cnt_rows) break;
            const RowInfo ri = sri[i];
            const int t = cbase + __builtin_amdgcn_readfirstlane(ri.me);
            const int o0 = __builtin_amdgcn_readfirstlane(ri.o0);
            const int m = __builtin_amdgcn_readfirstlane(ri.o1) - o0;
            const double rv = ri.rv;
            const double dr = ri.dr;
            double s = MODE == 0 ? rv : 0.0;
            for (int g0 = 0; g0 < m; g0 += WAVE) {
                const int cnt = min(WAVE, m - g0);
                const bool act = lane < cnt;
                int c = 0;
                bool in = false;
                double l = 0.0;
                // a value from HBM and one from LDS in separate registers,
                // so the LDS read need not wait for the HBM load that would
                // otherwise write the same register
                double vg = 0.0, vl = sent();
                if (act) {
                    const int q = o0 + g0 + lane;
                    // LDS reads of their own kind, so the two sources do not
                    // merge into flat loads (which wait for every memory
                    // access in flight, the last row's store included)
                    if (staged) {
                        c = lds_ld(&sec[q - e0]);
                        l = lds_ld(&sev[q - e0]);
                    } else {
                        c = ecol[q];
                        l = eval[q];
                    }
                    in = c >= cbase && c < cbase + cnt_rows;
                    // first load through the caches: y held the sentinel
                    // before this kernel and each value is written once, so
                    // only a sentinel can be stale (polled again past them)
                    if (!in) vg = y[c];
                }
                if (act && in && lds_ld(&sflag[c - cbase]) == tk) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    vl = lds_ld(&sy[c - cbase]);
                }
                double v = in ? vl : vg;
                int k = 0;
                long long spins = 0;
                int idle = 0;
                while (k < cnt) {
                    const unsigned long long miss = ballot64(act && is_sent(v)) & (~0ull << k);
                    const int stop = miss ? min(cnt, __ffsll((long long)miss) - 1) : cnt;
                    const double pr = l * v;
                    // in entry order; four products read out ahead of
                    // their adds
                    for (; k + 4 <= stop; k += 4) {
                        const double p0 = readlane_d(pr, k), p1 = readlane_d(pr, k + 1);
                        const double p2 = readlane_d(pr, k + 2), p3 = readlane_d(pr, k + 3);
                        if (MODE == 0) {
                            s -= p0; s -= p1; s -= p2; s -= p3;
                        } else {
                            s += p0; s += p1; s += p2; s += p3;
                        }
                    }
                    for (; k < stop; k++) {
                        const double pk = readlane_d(pr, k);
                        if (MODE == 0) s -= pk;
                        else s += pk;
                    }
                    if (k >= cnt) break;
                    // lane k's value (and maybe later ones) not yet there
                    {
                        const bool poll = act && is_sent(v);
                        double pg = v, pl = v;
                        if (poll && !in) pg = ld_rlx(&y[c]);
                        if (poll && in && lds_ld(&sflag[c - cbase]) == tk) {
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                            pl = lds_ld(&sy[c - cbase]);
                        }
                        v = in ? pl : pg;
                    }
                    if (!is_sent(readlane_d(v, k))) {
                        spins = 0;
                        idle = 0;
                    } else if (++spins > wait_limit || ((spins & 31) == 0 && err_raised(err))) {
                        // give up on this term, and make every other wait
                        // of the kernel give up too
                        if (!failed && lane == 0) atomicOr(err, 1);
                        failed = true;
                        if (lane == k) v = 0.0;
                    } else {
                        backoff(idle++, sleep_cap);
                    }
                }
            }
            const double out = MODE == 0 ? s * dr : (rv - s) / dr;
            if (lane == 0) {
                st_rlx(&y[t], out);   // first: other chunks wait on it longest
                lds_st(&sy[t - cbase], out);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");   // value before flag
                lds_st(&sflag[t - cbase], tk);
            }
        }
    }
    if (ballot64(failed) && lane == 0) atomicOr(err, 1);
}

__global__ void k_ic_sentinel(int n, double *a, double *b)
{
    for (int i = blockIdx.x * IC_BLK + threadIdx.x; i < n; i += gridDim.x * IC_BLK) {
        if (a) a[i] = sent();
        if (b) b[i] = sent();
    }
}

__global__ void k_ic_ticket_reset(unsigned *t)
{
    if (threadIdx.x < 4) t[threadIdx.x] = 0u;
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

// y = A x from the sliced-ELL copy (by position), each row summed in its
// CSR column order (csr_mat_vec, :57-68), + the block partials of x.y
__global__ void __launch_bounds__(IC_BLK)
k_ic_spmv(int n, const int *__restrict__ perm, const int *__restrict__ rp, const int *__restrict__ oA,
          const int *__restrict__ acol, const double *__restrict__ aell, const double *__restrict__ x,
          double *__restrict__ y, double *part, const IcCtrl *ctrl)
{
    if (ctrl->done) return;
    const int i = blockIdx.x * IC_BLK + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const int ri = row_at(perm, i);
        const int m = rp[ri + 1] - rp[ri];
        const size_t b = (size_t)oA[i >> 6] * WAVE + (i & (WAVE - 1));
        double s = 0.0;
        int q = 0;
        for (; q + 8 <= m; q += 8) {
            int c[8];
            double a[8], xv[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                c[j] = acol[b + (size_t)(q + j) * WAVE];
                a[j] = aell[b + (size_t)(q + j) * WAVE];
            }
#pragma unroll
            for (int j = 0; j < 8; j++) xv[j] = x[c[j]];
#pragma unroll
            for (int j = 0; j < 8; j++) s += a[j] * xv[j];
        }
        for (; q < m; q++) s += aell[b + (size_t)q * WAVE] * x[acol[b + (size_t)q * WAVE]];
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
// which only feeds p).  Without a preconditioner (plain) z is r, and
// k_pcg_dot(r, r) would form the same partials as k_pcg_axpy (same grid,
// rows and block sums): r.z is this sum, and its finaliser's work (beta,
// the next r.z) is done here instead.
__global__ void __launch_bounds__(IC_BLK)
k_pcg_fin_rr(const double *part, unsigned np, IcCtrl *c, int t, int max_iter, int plain)
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
        } else if (plain) {
            c->beta = rr / c->rz[t & 1];
            c->rz[(t + 1) & 1] = rr;
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

// Bytes one arena carve of n elements of size sz takes (256-byte aligned).
static size_t carve_bytes(size_t sz, size_t n)
{
    return (sz * (n > 0 ? n : 1) + 255) & ~(size_t)255;
}

template <typename T>
static int regrow(IcData *d, T **p, size_t n)
{
    if (d->arena) {
        // arena buffers are never freed: a regrow takes fresh room
        const size_t need = carve_bytes(sizeof(T), n);
        if (d->arena_used + need > d->arena_cap) {
            mps_set_error(MPS_ERR_ARG, __FILE__, __LINE__, "solve does not fit the no-allocation scratch");
            return MPS_ERR_ARG;
        }
        *p = (T *)(d->arena + d->arena_used);
        d->arena_used += need;
        return 0;
    }
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    HIP_OK(hipMalloc(p, sizeof(T) * (n > 0 ? n : 1)));
    return 0;
}

#define REGROW(p, n)                          \
    do {                                      \
        int rc_ = regrow(d, &(p), (size_t)(n)); \
        if (rc_) return rc_;                  \
    } while (0)

static int ic_reserve(IcData *d, int n_slots, long long nnz)
{
    if (n_slots + 1 > d->cap_n) {
        const size_t N = (size_t)n_slots + 2;
        REGROW(d->rp, N); REGROW(d->dpos, N); REGROW(d->eq_of_slot, N); REGROW(d->slot_of_eq, N);
        REGROW(d->flag, N); REGROW(d->eq_of_orig, N); REGROW(d->len, N);
        REGROW(d->dinv, N); REGROW(d->beq, N); REGROW(d->xeq, N);
        REGROW(d->order, N); REGROW(d->opos, N); REGROW(d->lev, N);
        REGROW(d->sperm, N); REGROW(d->spos, N);
        REGROW(d->lhist, N + 1); REGROW(d->lstart, N + 1);
        REGROW(d->cntl, N); REGROW(d->cntu, N); REGROW(d->loff, N + 1); REGROW(d->uoff, N + 1);
        REGROW(d->x, N); REGROW(d->r, N); REGROW(d->z, N); REGROW(d->p, N);
        REGROW(d->Ap, N); REGROW(d->y, N); REGROW(d->b, N);
        d->cap_n = n_slots + 1;
    }
    const int nsl = (n_slots + WAVE - 1) / WAVE + 1;
    if (nsl > d->cap_sl) {
        REGROW(d->wA, nsl);
        REGROW(d->oA, nsl + 1);
        d->cap_sl = nsl;
    }
    if (nnz > d->cap_nnz) {
        const size_t Z = (size_t)(nnz > 0 ? nnz : 1);
        REGROW(d->col, Z); REGROW(d->tpos, Z); REGROW(d->aval, Z); REGROW(d->lval, Z);
        d->cap_nnz = nnz;
    }
    if (!d->ticket) {
        REGROW(d->ticket, 4);
        REGROW(d->err, 1);
        REGROW(d->part, 3 * 65536);
        REGROW(d->hdr, 8);
        REGROW(d->scan_blk, 4 * 2050);
        if (d->arena) d->hhdr = d->hhdr_host;
        else HIP_OK(hipHostMalloc(&d->hhdr, sizeof(int) * 8, hipHostMallocDefault));
    }
    return 0;
}

void ic_free(IcData *d)
{
    if (!d) return;
    void *ps[] = {d->rp, d->col, d->dpos, d->tpos, d->eq_of_slot, d->slot_of_eq, d->flag, d->eq_of_orig, d->len,
                  d->aval, d->lval, d->dinv, d->beq, d->xeq, d->order, d->opos, d->lev, d->lhist, d->lstart,
                  d->sperm, d->spos, d->bh, d->boff,
                  d->x, d->r, d->z, d->p, d->Ap, d->y, d->b, d->wA, d->oA, d->acol, d->aell, d->cntl, d->cntu,
                  d->loff, d->uoff, d->lcol, d->ucol, d->lv, d->uv, d->hdr, d->scan_blk, d->ticket, d->err,
                  d->part};
    if (!d->arena) {
        for (void *q : ps) if (q) (void)hipFree(q);
        if (d->hhdr) (void)hipHostFree(d->hhdr);
    }
    delete d;
}

static long long wait_limit()
{
    const double v = mps_hip_get_option("ic_wait_limit");
    return v >= 1.0 ? (long long)v : 1;
}

static int sleep_cap()
{
    const double v = mps_hip_get_option("ic_sleep");
    return v < 0 ? 0 : (v > 3 ? 3 : (int)v);
}

static int reset_ticket(IcData *d, hipStream_t s)
{
    hipLaunchKernelGGL(k_ic_ticket_reset, dim3(1), dim3(64), 0, s, d->ticket);
    LAUNCH_CHECK();
    return 0;
}

static unsigned cu_count()
{
    static unsigned g = 0;
    if (!g) {
        int dev = 0, cus = 256;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        g = (unsigned)(cus > 0 ? cus : 256);
    }
    return g;
}

// Dependency levels of the rows (needs rp, col, dpos), then the level order
// (rows sorted by level; within a level in any order, since a level's rows
// are independent) and its inverse.  The factor and both solves claim rows
// in this order, which keeps the rows in flight on the dependency frontier:
// claimed in index order, only the few slabs of the lattice that fit the
// resident waves overlap, and the chain they leave is far longer than the
// DAG's depth.
// levels per bucket of the solve order (at least): with 16 a chunk of the
// solves (64 positions, 16 waves by default) holds four neighbouring
// lattice columns, so the dependencies on the next columns stay in the
// chunk as well as the ones along a column (forward solve at 1.08M
// particles 3.42-3.47 ms; 32-position chunks of 8 waves 3.56-3.62 ms;
// 32-level buckets 3.83-3.90 ms, 8-level 4.50-4.58 ms)
#define LEVEL_BUCKET 16
#define MAX_BUCKETS 1024  // buckets of the solve order (at most; their histograms are sized for it)
#define SCAN_MAX ((size_t)8192 * 2048)   // longest exclusive scan launch_scan_excl takes

static int ic_order(IcData *d, hipStream_t s)
{
    const int n = d->n_eq;
    int rc = reset_ticket(d, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ic_hdr_reset, dim3(1), dim3(64), 0, s, d->hdr, 8);
    hipLaunchKernelGGL(k_ic_sentinel, dim3(vgrid(n)), dim3(IC_BLK), 0, s, n, d->lev, (double *)nullptr);
    prof_begin(PK_ICL, s);
    hipLaunchKernelGGL(k_ic_levels, dim3(cu_count() * 2), dim3(IC_BLK), 0, s, n, d->rp, d->dpos, d->col, d->lev,
                       d->ticket, d->err, wait_limit(), sleep_cap());
    prof_end(PK_ICL, s);
    LAUNCH_CHECK();
    rc = launch_fill_int(d->lhist, 0, (size_t)n + 1, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ic_lev_hist, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, d->lev, d->lhist, d->hdr);
    LAUNCH_CHECK();
    rc = launch_scan_excl(d->lhist, d->lstart, n + 1, d->scan_blk, s);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(d->lhist, d->lstart, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_ic_lev_scatter, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, d->lev, d->lhist, d->order);
    hipLaunchKernelGGL(k_ic_inverse, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, d->order, d->opos);
    LAUNCH_CHECK();
    HIP_OK(hipMemcpyAsync(d->hhdr, d->hdr, sizeof(int) * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    d->levels = d->hhdr[4] + 1;
    // the solve order: buckets of 2^gshift levels, at most 256 of them
    int gshift = 0;
    const double want = mps_hip_get_option("ic_bucket");
    if (want >= 0) gshift = (int)want;
    else
        while ((1 << gshift) < LEVEL_BUCKET) gshift++;
    const unsigned nb = nblk(n);
    while (((d->levels - 1) >> gshift) >= MAX_BUCKETS ||
           ((size_t)(((d->levels - 1) >> gshift) + 1) * nb + 1 > SCAN_MAX))
        gshift++;
    const int nbk = ((d->levels - 1) >> gshift) + 1;
    d->bucket = 1 << gshift;
    const size_t nbh = (size_t)nbk * nb;
    if (nbh + 1 > d->cap_bh) {
        REGROW(d->bh, nbh + 1);
        REGROW(d->boff, nbh + 1);
        d->cap_bh = nbh + 1;
    }
    hipLaunchKernelGGL(k_ic_bucket_hist, dim3(nb), dim3(IC_BLK), sizeof(int) * (size_t)nbk, s, n, d->lev, gshift,
                       nbk, d->bh);
    LAUNCH_CHECK();
    rc = launch_scan_excl(d->bh, d->boff, (int)nbh, d->scan_blk, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ic_bucket_scatter, dim3(nb), dim3(IC_BLK), 0, s, n, d->lev, gshift, d->boff, d->sperm,
                       d->spos);
    LAUNCH_CHECK();
    return 0;
}

// A's sliced-ELL copy and the right-hand side in the order perm (NULL: eq
// order); needs rp, col, aval, dpos, beq.
static int build_ell(IcData *d, hipStream_t s, const int *perm, const int *pos)
{
    const int n = d->n_eq;
    const int nsl = (n + WAVE - 1) / WAVE;
    d->perm = perm;
    hipLaunchKernelGGL(k_ic_hdr_reset, dim3(1), dim3(64), 0, s, d->hdr, 4);
    hipLaunchKernelGGL(k_ic_slice_w, dim3(nblk((size_t)nsl * WAVE)), dim3(IC_BLK), 0, s, n, perm, d->rp, d->dpos,
                       d->wA, d->hdr);
    LAUNCH_CHECK();
    int rc = launch_scan_excl(d->wA, d->oA, nsl, d->scan_blk, s);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(d->hhdr, d->hdr, sizeof(int) * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(d->hhdr + 5, d->oA + nsl, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const size_t ea = (size_t)d->hhdr[5] * WAVE;
    d->max_lower = d->hhdr[3];
    if (ea > d->cap_ea) {
        REGROW(d->acol, ea); REGROW(d->aell, ea);
        d->cap_ea = ea;
    }
    hipLaunchKernelGGL(k_ic_fill_a, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, perm, pos, d->rp, d->col, d->aval, d->oA,
                       d->acol, d->aell);
    // the right-hand side by position
    hipLaunchKernelGGL(k_ic_gather, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, perm, d->beq, d->b);
    LAUNCH_CHECK();
    return 0;
}

// The factor in level order, then L and its columns by position of the
// solve order.  Needs rp, col, aval, dpos, tpos, ic_order and build_ell
// with the solve order.
static int ic_factor(IcData *d, hipStream_t s, int dense)
{
    const int n = d->n_eq;
    if (n == 0) return 0;
    int lcap = d->max_lower > 0 ? d->max_lower : 1;
    lcap += lcap & 1;   // even: the values that follow the columns stay 8-byte aligned
    if (lcap > LDS_ROW_MAX) {
        mps_set_error(MPS_ERR_UNSUPPORTED, __FILE__, __LINE__,
                      "IC(0): a PPE row has more strictly-lower entries than the factor holds in LDS");
        return MPS_ERR_UNSUPPORTED;
    }
    // four waves (rows) per workgroup while their rows fit 64 KB of LDS
    const int wpb = (size_t)lcap * 12 * 4 <= 65536 ? 4 : 1;
    const size_t lds = (size_t)lcap * 12 * wpb;
    hipLaunchKernelGGL(k_ic_reset_factor, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, d->rp, d->dpos, d->lval, d->dinv);
    int rc = reset_ticket(d, s);
    if (rc) return rc;
    const double og = mps_hip_get_option("ic_fgrid");
    const unsigned grid = og >= 1 ? (unsigned)og : cu_count() * (wpb == 4 ? 4 : 8);
    prof_begin(PK_ICF, s);
    if (mps_hip_get_option("ic_factor_form") != 1) {
        // two-phase form: one wave per workgroup, match lists in LDS for
        // rows of up to 64 entries; longer rows in the one-phase form
        const int l2 = lcap < WAVE ? lcap : WAVE;
        size_t lds2 = (size_t)l2 * 12 + (size_t)(l2 * (l2 - 1) / 2) * 10 + 16;
        if ((size_t)lcap * 12 > lds2) lds2 = (size_t)lcap * 12;
        const unsigned grid2 = og >= 1 ? (unsigned)og : cu_count() * 8;
        hipLaunchKernelGGL(k_ic_factor2, dim3(grid2), dim3(WAVE), lds2, s, n, d->order, d->rp, d->col, d->dpos,
                           d->aval, d->lval, d->dinv, d->ticket, d->err, dense, l2, lcap, wait_limit(), sleep_cap());
    } else {
        hipLaunchKernelGGL(k_ic_factor, dim3(grid), dim3(WAVE * wpb), lds, s, n, d->order, d->rp, d->col, d->dpos,
                           d->aval, d->lval, d->dinv, d->ticket, d->err, dense, lcap, wait_limit(), sleep_cap());
    }
    prof_end(PK_ICF, s);
    LAUNCH_CHECK();
    // L and U by position of the solve order, rows contiguous
    hipLaunchKernelGGL(k_ic_counts, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, d->perm, d->rp, d->dpos, d->cntl, d->cntu);
    LAUNCH_CHECK();
    rc = launch_scan_excl(d->cntl, d->loff, n, d->scan_blk, s);
    if (!rc) rc = launch_scan_excl(d->cntu, d->uoff, n, d->scan_blk, s);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(d->hhdr + 6, d->loff + n, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(d->hhdr + 7, d->uoff + n, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const size_t nl = (size_t)(d->hhdr[6] > d->hhdr[7] ? d->hhdr[6] : d->hhdr[7]);
    if (nl > d->cap_lu) {
        REGROW(d->lcol, nl); REGROW(d->lv, nl); REGROW(d->ucol, nl); REGROW(d->uv, nl);
        d->cap_lu = nl;
    }
    hipLaunchKernelGGL(k_ic_fill_lu, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, d->perm, d->spos, d->rp, d->col, d->dpos,
                       d->tpos, d->lval, d->loff, d->uoff, d->lcol, d->lv, d->ucol, d->uv, dense);
    LAUNCH_CHECK();
    return 0;
}

// z = M^-1 r with y as scratch (by position of the solve order); y and z
// must hold the sentinel already.  With ctrl given, both solves return at
// once once the solver has stopped.
static int ic_apply(IcData *d, const double *r, hipStream_t s, int dense, const IcCtrl *ctrl)
{
    int rc = reset_ticket(d, s);
    if (rc) return rc;
    const long long wl = wait_limit();
    const int sc = sleep_cap();
    const double og = mps_hip_get_option("ic_grid");
    const unsigned grid = og >= 1 ? (unsigned)og : cu_count() * 2;
    // chunks of "ic_chunk" positions (32, 64, 128), "ic_waves" wavefronts
    // (4, 8 or 16) per workgroup
    const int ck = (int)mps_hip_get_option("ic_chunk");
    const int ow = (int)mps_hip_get_option("ic_waves");
    const int nw = ow == 4 ? 4 : (ow == 16 ? 16 : 8);
    decltype(&k_ic_csolve<0, false, 64, 8>) fwd, bwd;
#define IC_PICK(C, W)                                                                  \
    do {                                                                               \
        fwd = dense ? k_ic_csolve<1, false, C, W> : k_ic_csolve<0, false, C, W>;       \
        bwd = dense ? k_ic_csolve<1, true, C, W> : k_ic_csolve<0, true, C, W>;         \
    } while (0)
    if (ck == 64) { if (nw == 4) IC_PICK(64, 4); else if (nw == 16) IC_PICK(64, 16); else IC_PICK(64, 8); }
    else if (ck == 128) { if (nw == 4) IC_PICK(128, 4); else if (nw == 16) IC_PICK(128, 16); else IC_PICK(128, 8); }
    else { if (nw == 4) IC_PICK(32, 4); else if (nw == 16) IC_PICK(32, 16); else IC_PICK(32, 8); }
#undef IC_PICK
    prof_begin(PK_IC, s);
    hipLaunchKernelGGL(fwd, dim3(grid), dim3(nw * WAVE), 0, s, d->n_eq, d->perm, d->lev, d->loff, d->lcol,
                       d->lv, d->dinv, r, d->y, d->ticket + 0, d->err, wl, sc, ctrl);
    prof_end(PK_IC, s);
    prof_begin(PK_ICB, s);
    hipLaunchKernelGGL(bwd, dim3(grid), dim3(nw * WAVE), 0, s, d->n_eq, d->perm, d->lev, d->uoff, d->ucol,
                       d->uv, d->dinv, d->y, d->z, d->ticket + 1, d->err, wl, sc, ctrl);
    prof_end(PK_ICB, s);
    LAUNCH_CHECK();
    return 0;
}

static int clear_err(IcData *d, hipStream_t s)
{
    return launch_fill_int(d->err, 0, 1, s);
}

// PCG (precond 1) or CG (precond 0) on the system held in d by position of
// d->perm (build_ell done); the solution is left in d->xeq in eq order.
// The caller clears d->err before the level pass and the factor; it is read
// back here, after the solves.
static int pcg_run(IcData *d, hipStream_t s, IcCtrl *dctrl, IcCtrl *hctrl, const SolveOpts &o, int precond,
                   SolveResult *res)
{
    const int n = d->n_eq;
    const unsigned gv = vgrid(n), gm = nblk(n);
    double *pa = d->part, *pb = d->part + 65536, *pc = d->part + 2 * 65536;
    const int dense = o.dense_ic;
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
    hipEvent_t ev[2] = {d->ev[0], d->ev[1]};
    const bool own_ev = !ev[0];
    if (own_ev) {
        HIP_OK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    }
    volatile int *hdone = &hctrl->done;
    int pending = -1, t = 0, k = 0, rc = 0;
    const double *zz = precond ? d->z : d->r;
    while (t < o.max_iter && !rc) {
        const int t_end = (t + chunk < o.max_iter) ? t + chunk : o.max_iter;
        for (; t < t_end && !rc; t++) {
            prof_begin(PK_SPMV, s);
            hipLaunchKernelGGL(k_ic_spmv, dim3(gm), dim3(IC_BLK), 0, s, n, d->perm, d->rp, d->oA, d->acol, d->aell,
                               d->p, d->Ap, pa, dctrl);
            prof_end(PK_SPMV, s);
            hipLaunchKernelGGL(k_pcg_fin_pap, dim3(1), dim3(IC_BLK), 0, s, pa, gm, dctrl, t);
            hipLaunchKernelGGL(k_pcg_axpy, dim3(gv), dim3(IC_BLK), 0, s, n, d->p, d->Ap, d->x, d->r, d->z, d->y, pb,
                               dctrl, precond);
            hipLaunchKernelGGL(k_pcg_fin_rr, dim3(1), dim3(IC_BLK), 0, s, pb, gv, dctrl, t, o.max_iter, !precond);
            if (precond) {
                rc = ic_apply(d, d->r, s, dense, dctrl);   // return at once after the stop flag
                hipLaunchKernelGGL(k_pcg_dot, dim3(gv), dim3(IC_BLK), 0, s, n, d->r, zz, pc, dctrl);
                hipLaunchKernelGGL(k_pcg_fin_rz, dim3(1), dim3(IC_BLK), 0, s, pc, gv, dctrl, t);
            }
            hipLaunchKernelGGL(k_pcg_pupdate, dim3(gv), dim3(IC_BLK), 0, s, n, zz, d->p, dctrl);
        }
        if (rc) break;
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) {
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(le));
            rc = MPS_ERR_HIP;
            break;
        }
        if (hipMemcpyAsync((void *)hdone, &dctrl->done, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipEventRecord(ev[k & 1], s) != hipSuccess) {
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, "PCG poll");
            rc = MPS_ERR_HIP;
            break;
        }
        if (pending >= 0) {
            if (hipEventSynchronize(ev[pending & 1]) != hipSuccess) {
                mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, "PCG poll");
                rc = MPS_ERR_HIP;
                break;
            }
            if (*hdone) break;
        }
        pending = k;
        k++;
    }
    if (!rc) hipLaunchKernelGGL(k_ic_scatter, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, d->perm, d->x, d->xeq);
    const hipError_t se = hipStreamSynchronize(s);
    if (own_ev) {
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
    }
    if (rc) return rc;
    HIP_OK(se);
    HIP_OK(hipMemcpy(hctrl, dctrl, sizeof(IcCtrl), hipMemcpyDeviceToHost));
    int herr = 0;
    HIP_OK(hipMemcpy(&herr, d->err, sizeof(int), hipMemcpyDeviceToHost));
    res->iters = hctrl->done ? hctrl->iters : o.max_iter;
    res->breakdown = hctrl->breakdown;
    res->res = hctrl->res;
    res->bnorm = hctrl->bnorm;
    res->outer = 1;
    if (herr) {
        mps_set_error(MPS_ERR_SOLVER, __FILE__, __LINE__, "IC(0) dependency wait exceeded its bound (ic_wait_limit)");
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
    rc = ic_reserve(d, n, 0);
    if (rc) return rc;
    // eq numbering
    hipLaunchKernelGGL(k_ic_flag, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, c->O[c->ct], c->unk, d->flag);
    LAUNCH_CHECK();
    rc = launch_scan_excl(d->flag, d->eq_of_orig, n, c->scan_blk, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ic_maps, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, c->O[c->ct], c->unk, c->Alen, d->eq_of_orig,
                       d->eq_of_slot, d->slot_of_eq, d->len);
    LAUNCH_CHECK();
    HIP_OK(hipMemcpyAsync(d->hhdr, d->eq_of_orig + n, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const int n_eq = d->hhdr[0];
    d->n_eq = n_eq;
    if (n_eq == 0) {
        rc = launch_fill_double(c->x, 0.0, (size_t)n, s);
        res->iters = 0; res->breakdown = 0; res->res = 0; res->bnorm = 0; res->outer = 1;
        return rc;
    }
    rc = launch_scan_excl(d->len, d->rp, n_eq, c->scan_blk, s);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(d->hhdr, d->rp + n_eq, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    const long long nnz = d->hhdr[0];
    rc = ic_reserve(d, n, nnz);
    if (rc) return rc;
    d->nnz = nnz;
    hipLaunchKernelGGL(k_ic_fill_sorted, dim3(nblk((size_t)n_eq * WAVE)), dim3(IC_BLK), 0, s, n_eq, c->stride,
                       d->slot_of_eq, d->eq_of_slot, c->Alen, c->Aval, c->Acol, c->diag, c->b, d->rp, d->col,
                       d->aval, d->dpos, d->beq);
    hipLaunchKernelGGL(k_ic_tpos, dim3(nblk(n_eq)), dim3(IC_BLK), 0, s, n_eq, d->rp, d->col, d->dpos, d->tpos);
    LAUNCH_CHECK();
    rc = clear_err(d, s);
    if (!rc) rc = ic_order(d, s);
    if (!rc) rc = build_ell(d, s, d->sperm, d->spos);
    if (!rc) rc = ic_factor(d, s, o.dense_ic);
    if (!rc) rc = pcg_run(d, s, g_dctrl, g_hctrl, o, 1, res);
    if (rc) return rc;
    c->stats.ic_levels = d->levels;
    hipLaunchKernelGGL(k_ic_scatter_x, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, d->eq_of_slot, d->xeq, c->x);
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
    rc = ic_reserve(d, n, nnz);
    if (rc) return rc;
    d->n_eq = n;
    d->nnz = nnz;
    HIP_OK(hipMemcpyAsync(d->rp, rp, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(d->col, col, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(d->aval, val, sizeof(double) * (size_t)nnz, hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(d->beq, b, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    // the slice widths read a diagonal position; CG needs only A's copy
    rc = launch_fill_int(d->dpos, 0, (size_t)n, s);
    if (!rc) rc = build_ell(d, s, nullptr, nullptr);
    if (!rc) rc = clear_err(d, s);
    if (rc) return rc;
    SolveOpts oo = o;
    oo.solver = 0;
    rc = pcg_run(d, s, g_dctrl, g_hctrl, oo, 0, res);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(x, d->xeq, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    return 0;
}

// mps_hip_solve_csr's work on a fresh IcData (freed by the caller), with
// the PCG control block dctrl (device) and its host copy hctrl.
static int solve_csr_on(IcData *d, IcCtrl *dctrl, IcCtrl *hctrl, int n, const int *row_ptr, const int *col,
                        const double *val, const double *b, double *x, int solver_type, int max_iter, double tol,
                        int *iters_out)
{
    const long long nnz = row_ptr[n];
    int rc = ic_reserve(d, n, nnz);
    if (rc) return rc;
    d->n_eq = n;
    d->nnz = nnz;
    hipStream_t s = nullptr;
    HIP_OK(hipMemcpy(d->rp, row_ptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d->col, col, sizeof(int) * nnz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d->aval, val, sizeof(double) * nnz, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d->beq, b, sizeof(double) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_ic_dpos, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, d->rp, d->col, d->dpos);
    hipLaunchKernelGGL(k_ic_tpos, dim3(nblk(n)), dim3(IC_BLK), 0, s, n, d->rp, d->col, d->dpos, d->tpos);
    LAUNCH_CHECK();
    SolveOpts o{};
    o.solver = solver_type;
    o.max_iter = max_iter;
    o.tol = tol;
    o.tol_mode = (int)mps_hip_get_option("tol_mode");
    o.check_every = (int)mps_hip_get_option("check_every");
    o.dense_ic = MPS_DIM == 2;
    SolveResult sr{};
    const bool ic = solver_type == 1;
    rc = clear_err(d, s);
    if (!rc && ic) rc = ic_order(d, s);
    if (!rc) rc = build_ell(d, s, ic ? d->sperm : nullptr, ic ? d->spos : nullptr);
    if (!rc && ic) rc = ic_factor(d, s, o.dense_ic);
    if (!rc) rc = pcg_run(d, s, dctrl, hctrl, o, ic, &sr);
    if (!rc) HIP_OK(hipMemcpy(x, d->xeq, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (iters_out) *iters_out = sr.iters;
    return rc;
}

// Checks mps_hip_solve_csr's arguments: row_ptr from 0 and non-decreasing,
// columns in range and ascending in each row, and for ICCG a symmetric
// pattern.  The reference fills L_ik from the entry (k, i) of row k
// (pressure_solver.c:196-228), the up-looking factor here from (i, k).
static int check_csr(const char *who, int n, const int *row_ptr, const int *col, const double *val,
                     const double *b, const double *x, int solver_type)
{
    char msg[160];
    if (!row_ptr || !col || !val || !b || !x || row_ptr[0] != 0) {
        snprintf(msg, sizeof msg, "%s: invalid CSR", who);
        mps_set_error(MPS_ERR_ARG, __FILE__, __LINE__, msg);
        return MPS_ERR_ARG;
    }
    for (int i = 0; i < n; i++) {
        if (row_ptr[i + 1] < row_ptr[i]) {
            snprintf(msg, sizeof msg, "%s: row_ptr decreases", who);
            mps_set_error(MPS_ERR_ARG, __FILE__, __LINE__, msg);
            return MPS_ERR_ARG;
        }
        for (int q = row_ptr[i]; q < row_ptr[i + 1]; q++) {
            const bool sorted = q == row_ptr[i] || col[q - 1] < col[q];
            if (col[q] < 0 || col[q] >= n || !sorted) {
                snprintf(msg, sizeof msg, "%s: columns out of range or unsorted", who);
                mps_set_error(MPS_ERR_ARG, __FILE__, __LINE__, msg);
                return MPS_ERR_ARG;
            }
        }
    }
    if (solver_type == 1) {
        // pattern symmetry: every (i, j) has its (j, i)
        for (int i = 0; i < n; i++)
            for (int q = row_ptr[i]; q < row_ptr[i + 1]; q++) {
                const int j = col[q];
                const int *lo = col + row_ptr[j], *hi = col + row_ptr[j + 1];
                const int *at = std::lower_bound(lo, hi, i);
                if (at == hi || *at != i) {
                    snprintf(msg, sizeof msg, "%s: ICCG needs a symmetric pattern", who);
                    mps_set_error(MPS_ERR_ARG, __FILE__, __LINE__, msg);
                    return MPS_ERR_ARG;
                }
            }
    }
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
    if (iters_out) *iters_out = 0;
    if (n <= 0) return MPS_OK;
    rc = check_csr("mps_hip_solve_csr", n, row_ptr, col, val, b, x, solver_type);
    if (rc) return rc;
    IcData *d = new IcData();
    rc = solve_csr_on(d, g_dctrl, g_hctrl, n, row_ptr, col, val, b, x, solver_type, max_iter, tol, iters_out);
    ic_free(d);
    return rc;
}

// ---------------------------------------------------------------------
// the CSR solve without allocation
// ---------------------------------------------------------------------

// Every buffer of mps_hip_solve_csr_noalloc lives in this array: the CSR
// staged from the host, the orders, A's sliced ELL, the factor, the PCG
// vectors, partial sums, counters and the control block.  Each device
// holds its own copy for the life of the process; host code uses only the
// address hipGetSymbolAddress returns for the current device.
#define NOALLOC_DEVICES 16
static __device__ __align__(256) unsigned char scratch[MPS_NOALLOC_SCRATCH_BYTES];

static void *g_na_ptr[NOALLOC_DEVICES];            // scratch's device address, per device
static int g_na_rc[NOALLOC_DEVICES];               // set-up status, per device
static hipEvent_t g_na_ev[NOALLOC_DEVICES][2];     // the PCG poll's events, per device
static IcData g_na_ic[NOALLOC_DEVICES];            // the solve's buffer table (pointers into scratch)
static IcCtrl g_na_hctrl[NOALLOC_DEVICES];         // host copy of the control block

// Set-up of the no-allocation solve on device dev (current): scratch's
// address and the two poll events.  Run once per device from the library's
// initialisation; a failure is kept and reported by the entry point.
int ic_noalloc_setup(int dev)
{
    if (dev < 0 || dev >= NOALLOC_DEVICES) return MPS_OK;   // the entry point rejects such a device
    if (g_na_ptr[dev] || g_na_rc[dev]) return g_na_rc[dev];
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess && !ptr) e = hipErrorInvalidSymbol;
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g_na_ev[dev][0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g_na_ev[dev][1], hipEventDisableTiming);
    if (e != hipSuccess) {
        mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(e));
        g_na_rc[dev] = MPS_ERR_HIP;
        return MPS_ERR_HIP;
    }
    g_na_ptr[dev] = ptr;
    return MPS_OK;
}

// Scratch bytes the solve takes, as the carves of ic_reserve, ic_order,
// build_ell and ic_factor take them: n unknowns, nnz entries, rows of at
// most maxlen entries, nl the larger of the strictly-lower and
// strictly-upper counts (ICCG only).
static size_t noalloc_bytes(int n, long long nnz, int maxlen, long long nl, bool ic)
{
    const size_t N = (size_t)n + 2, Z = (size_t)(nnz > 0 ? nnz : 1);
    const size_t nsl = (size_t)(n + WAVE - 1) / WAVE + 1;
    size_t t = 0;
    t += 13 * carve_bytes(sizeof(int), N) + 4 * carve_bytes(sizeof(int), N + 1);
    t += 11 * carve_bytes(sizeof(double), N);
    t += carve_bytes(sizeof(int), nsl) + carve_bytes(sizeof(int), nsl + 1);
    t += 2 * carve_bytes(sizeof(int), Z) + 2 * carve_bytes(sizeof(double), Z);
    t += carve_bytes(sizeof(unsigned), 4) + carve_bytes(sizeof(int), 1) + carve_bytes(sizeof(double), 3 * 65536);
    t += carve_bytes(sizeof(int), 8) + carve_bytes(sizeof(int), 4 * 2050);
    t += carve_bytes(sizeof(IcCtrl), 1);
    const size_t ea = (size_t)((n + WAVE - 1) / WAVE) * WAVE * (size_t)(maxlen > 0 ? maxlen : 1);
    t += carve_bytes(sizeof(int), ea) + carve_bytes(sizeof(double), ea);
    if (ic) {
        const size_t nbh = (size_t)MAX_BUCKETS * nblk((size_t)n) + 1;
        t += 2 * carve_bytes(sizeof(int), nbh);
        const size_t L = (size_t)(nl > 0 ? nl : 1);
        t += 2 * carve_bytes(sizeof(int), L) + 2 * carve_bytes(sizeof(double), L);
    }
    return t;
}

// mps_hip_solve_csr with every device buffer in scratch: same arguments,
// same kernels and launch shapes, so the same bits.  No hipMalloc, hipFree
// or other allocation.  Calls that overlap on one device race on scratch.
extern "C" int mps_hip_solve_csr_noalloc(int n, const int *row_ptr, const int *col, const double *val,
                                         const double *b, double *x, int solver_type, int max_iter, double tol,
                                         int *iters_out)
{
    int rc = mps_hip_init(0);
    if (rc) return rc;
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_DEVICES) {
        mps_set_error(MPS_ERR_UNSUPPORTED, __FILE__, __LINE__, "mps_hip_solve_csr_noalloc: device id above 15");
        return MPS_ERR_UNSUPPORTED;
    }
    rc = ic_noalloc_setup(dev);
    if (rc) return rc;
    if (iters_out) *iters_out = 0;
    if (n <= 0) return MPS_OK;
    rc = check_csr("mps_hip_solve_csr_noalloc", n, row_ptr, col, val, b, x, solver_type);
    if (rc) return rc;
    int maxlen = 0;
    long long lo = 0, up = 0;
    for (int i = 0; i < n; i++) {
        const int len = row_ptr[i + 1] - row_ptr[i];
        if (len > maxlen) maxlen = len;
        for (int q = row_ptr[i]; q < row_ptr[i + 1]; q++) {
            if (col[q] < i) lo++;
            else if (col[q] > i) up++;
        }
    }
    const size_t need = noalloc_bytes(n, row_ptr[n], maxlen, lo > up ? lo : up, solver_type == 1);
    if (need > (size_t)MPS_NOALLOC_SCRATCH_BYTES) {
        char msg[200];
        snprintf(msg, sizeof msg,
                 "mps_hip_solve_csr_noalloc: %d unknowns, %d entries, rows up to %d need %zu bytes of scratch, "
                 "above its %zu", n, row_ptr[n], maxlen, need, (size_t)MPS_NOALLOC_SCRATCH_BYTES);
        mps_set_error(MPS_ERR_ARG, __FILE__, __LINE__, msg);
        return MPS_ERR_ARG;
    }
    IcData *d = &g_na_ic[dev];
    *d = IcData();
    d->arena = (char *)g_na_ptr[dev];
    d->arena_cap = MPS_NOALLOC_SCRATCH_BYTES;
    d->ev[0] = g_na_ev[dev][0];
    d->ev[1] = g_na_ev[dev][1];
    IcCtrl *dctrl = nullptr;
    REGROW(dctrl, 1);
    rc = solve_csr_on(d, dctrl, &g_na_hctrl[dev], n, row_ptr, col, val, b, x, solver_type, max_iter, tol,
                      iters_out);
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
