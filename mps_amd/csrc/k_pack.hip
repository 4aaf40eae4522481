// k_pack.hip -- the CG's copy of the PPE matrix.
//
// The assembly writes each unknown's row into the neighbour-list layout
// (64-slot slices, stride max_neighbors), which suits the assembly but
// leaves idle lanes at every non-unknown slot and gathers the CG vector
// through the texture path once per nonzero.  The CG reads the matrix
// hundreds of times per step, so after each assembly the rows are repacked
// (about the cost of two SpMVs):
//   * rows = the unknowns in slot order, columns = row numbers, so the CG
//     vectors hold only unknowns (n_eq, not n) and every lane has a row;
//   * slices of 64 rows, each as wide as its longest row;
//   * blocks of 256 rows (one SpMV workgroup).  The rows a block's entries
//     reference -- its halo, about 1,900 rows for a 256-row block of the
//     1.08M-particle dam break, each used ~15 times -- are listed once; the
//     SpMV stages the vector at the halo in LDS and gathers from there, and
//     an entry's column is a 12-bit index into the halo, eight per three
//     32-bit words.  A block whose halo is larger than PK_HALO_CAP, or whose
//     columns span more than 2^17 rows, keeps 32-bit row numbers and
//     gathers from global memory.
// Entries keep their order inside a row, so each row's sum is unchanged.
// The halo of a block is found with an LDS bitmap over its column span:
// set a bit per referenced row, then an entry's halo index is the number
// of set bits below its own.  Halo indices are 12-bit (a staged halo holds
// at most PK_HALO_CAP = 4096 rows), eight to three 32-bit words.
#include <limits.h>

#include <vector>

#include "ctx.hpp"

#define PKB 256
#define BM_WORDS 4096     // bitmap of 2^17 rows
#define PK_U 8            // row entries per load batch (gathers overlap)
static_assert(PKB == PK_ROWS, "pack block = SpMV block");

static inline unsigned pk_blocks(size_t n) { return (unsigned)((n + PKB - 1) / PKB); }

#define LAUNCH_CHECK()                                                       \
    do {                                                                     \
        hipError_t e_ = hipGetLastError();                                   \
        if (e_ != hipSuccess) {                                              \
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(e_)); \
            return MPS_ERR_HIP;                                              \
        }                                                                    \
    } while (0)

__device__ __forceinline__ void pk_slots_body(int lb, int n, const int *__restrict__ unk, const int *__restrict__ row_of_slot,
                           int *__restrict__ slot_of_row)
{
    const int s = lb * PKB + threadIdx.x;
    if (s < n && unk[s]) slot_of_row[row_of_slot[s]] = s;
}

__global__ void __launch_bounds__(PKB)
k_pk_slots(int n, const int *__restrict__ unk, const int *__restrict__ row_of_slot,
                           int *__restrict__ slot_of_row)
{
    pk_slots_body(blockIdx.x, n, unk, row_of_slot, slot_of_row);
}

__device__ __forceinline__ int wave_max_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int u = __shfl_xor(v, o, WAVE);
        v = u > v ? u : v;
    }
    return v;
}

__device__ __forceinline__ int wave_min_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int u = __shfl_xor(v, o, WAVE);
        v = u < v ? u : v;
    }
    return v;
}

__device__ __forceinline__ int wave_sum_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// The row of this thread and the column range of the whole block (every
// thread gets the block's min and max).
struct BlockRows {
    int k, slot, m, bmn, bmx;
};

__device__ __forceinline__ BlockRows block_rows(int lb, int n_eq, int stride, const int *__restrict__ slot_of_row,
                                                const int *__restrict__ row_of_slot, const int *__restrict__ Alen,
                                                const int *__restrict__ Acol, int *red)
{
    BlockRows br;
    br.k = lb * PKB + threadIdx.x;
    br.slot = -1;
    br.m = 0;
    int mn = INT_MAX, mx = -1;
    if (br.k < n_eq) {
        br.slot = slot_of_row[br.k];
        br.m = Alen[br.slot];
        for (int q0 = 0; q0 < br.m; q0 += PK_U) {
            int cs[PK_U];
#pragma unroll
            for (int u = 0; u < PK_U; u++) cs[u] = (q0 + u < br.m) ? Acol[sell_off(br.slot, q0 + u, stride)] : -1;
#pragma unroll
            for (int u = 0; u < PK_U; u++) {
                if (cs[u] < 0) continue;
                const int cr = row_of_slot[cs[u]];
                mn = cr < mn ? cr : mn;
                mx = cr > mx ? cr : mx;
            }
        }
    }
    mn = wave_min_i(mn);
    mx = wave_max_i(mx);
    const int w = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        red[w] = mn;
        red[PKB / WAVE + w] = mx;
    }
    __syncthreads();
    br.bmn = INT_MAX;
    br.bmx = -1;
#pragma unroll
    for (int j = 0; j < PKB / WAVE; j++) {
        br.bmn = red[j] < br.bmn ? red[j] : br.bmn;
        br.bmx = red[PKB / WAVE + j] > br.bmx ? red[PKB / WAVE + j] : br.bmx;
    }
    __syncthreads();
    return br;
}

// Sets the bitmap of the block's referenced rows (bit c - bmn).
__device__ __forceinline__ void block_bitmap(const BlockRows &br, int stride, const int *__restrict__ row_of_slot,
                                             const int *__restrict__ Acol, unsigned *bits)
{
    for (int w = threadIdx.x; w < BM_WORDS; w += PKB) bits[w] = 0u;
    __syncthreads();
    for (int q0 = 0; q0 < br.m; q0 += PK_U) {
        int cs[PK_U];
#pragma unroll
        for (int u = 0; u < PK_U; u++) cs[u] = (q0 + u < br.m) ? Acol[sell_off(br.slot, q0 + u, stride)] : -1;
#pragma unroll
        for (int u = 0; u < PK_U; u++) {
            if (cs[u] < 0) continue;
            const int c = row_of_slot[cs[u]] - br.bmn;
            atomicOr(&bits[c >> 5], 1u << (c & 31));
        }
    }
    __syncthreads();
}

// Pass 1, one workgroup per block: slice widths, halo size and mode.
__device__ __forceinline__ void pk_pass1_body(int lb, int n_eq, int n_slices, int stride, const int *__restrict__ slot_of_row,
           const int *__restrict__ row_of_slot, const int *__restrict__ Alen, const int *__restrict__ Acol,
           int *__restrict__ sw, int *__restrict__ whw, int *__restrict__ w32, int *__restrict__ bh,
           int *__restrict__ bmode, int *dmax)
{
    __shared__ unsigned bits[BM_WORDS];
    __shared__ int red[2 * PKB / WAVE];
    const BlockRows br = block_rows(lb, n_eq, stride, slot_of_row, row_of_slot, Alen, Acol, red);
    const bool span_ok = br.bmx < 0 || (br.bmx - br.bmn) < BM_WORDS * 32;
    int H = 0;
    if (span_ok && br.bmx >= 0) {           // uniform over the block
        block_bitmap(br, stride, row_of_slot, Acol, bits);
        int cnt = 0;
        for (int w = threadIdx.x; w < BM_WORDS; w += PKB) cnt += __popc(bits[w]);
        cnt = wave_sum_i(cnt);
        if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = cnt;
        __syncthreads();
        for (int j = 0; j < PKB / WAVE; j++) H += red[j];
    }
    const int mode = span_ok && H <= PK_HALO_CAP;
    const int wmax = wave_max_i(br.m);
    const int sl = lb * (PKB / WAVE) + threadIdx.x / WAVE;
    if (sl < n_slices && (threadIdx.x & (WAVE - 1)) == 0) {
        sw[sl] = wmax;
        whw[sl] = mode ? (wmax + 7) / 8 * 3 : 0;     // three words per 8 entries
        w32[sl] = mode ? 0 : wmax;
    }
    if (threadIdx.x == 0) {
        bh[lb] = mode ? H : 0;
        bmode[lb] = mode;
        if (mode) atomicMax(dmax, H);
    }
}

__global__ void __launch_bounds__(PKB)
k_pk_pass1(int n_eq, int n_slices, int stride, const int *__restrict__ slot_of_row,
           const int *__restrict__ row_of_slot, const int *__restrict__ Alen, const int *__restrict__ Acol,
           int *__restrict__ sw, int *__restrict__ whw, int *__restrict__ w32, int *__restrict__ bh,
           int *__restrict__ bmode, int *dmax)
{
    pk_pass1_body(blockIdx.x, n_eq, n_slices, stride, slot_of_row, row_of_slot, Alen, Acol, sw, whw, w32, bh, bmode, dmax);
}

// Pass 2, one workgroup per block: halo list, values, columns, slice and
// block descriptors.
__device__ __forceinline__ void pk_pass2_body(int lb, int n_eq, int n_slices, int stride, const int *__restrict__ slot_of_row,
           const int *__restrict__ row_of_slot, const int *__restrict__ Alen, const double *__restrict__ Aval,
           const int *__restrict__ Acol, const double *__restrict__ diag, const double *__restrict__ b,
           const int *__restrict__ bmode, const int *__restrict__ bh, const int *__restrict__ hoff,
           const int *__restrict__ voff, const int *__restrict__ chwoff, const int *__restrict__ c32off,
           int *__restrict__ halo, double *__restrict__ val, unsigned *__restrict__ chw,
           int *__restrict__ c32, int *__restrict__ len, double *__restrict__ pdiag, double *__restrict__ pb,
           int4 *__restrict__ smeta, int2 *__restrict__ bmeta)
{
    __shared__ unsigned bits[BM_WORDS];
    __shared__ int pre[BM_WORDS];
    __shared__ int red[2 * PKB / WAVE];
    const BlockRows br = block_rows(lb, n_eq, stride, slot_of_row, row_of_slot, Alen, Acol, red);
    const int mode = bmode[lb];
    if (mode && br.bmx >= 0) {              // uniform over the block
        block_bitmap(br, stride, row_of_slot, Acol, bits);
        const int w0 = threadIdx.x * (BM_WORDS / PKB);
        int cnt[BM_WORDS / PKB];
        int tot = 0;
#pragma unroll
        for (int j = 0; j < BM_WORDS / PKB; j++) {
            cnt[j] = __popc(bits[w0 + j]);
            tot += cnt[j];
        }
        int all;
        int run = block_scan_excl<PKB>(tot, &all);
#pragma unroll
        for (int j = 0; j < BM_WORDS / PKB; j++) {
            pre[w0 + j] = run;
            run += cnt[j];
        }
        const int ho = hoff[lb];
#pragma unroll
        for (int j = 0; j < BM_WORDS / PKB; j++) {
            unsigned word = bits[w0 + j];
            int at = ho + pre[w0 + j];
            while (word) {
                const int bit = __ffs(word) - 1;
                halo[at++] = br.bmn + (w0 + j) * 32 + bit;
                word &= word - 1u;
            }
        }
        __syncthreads();
    }
    const int lane = threadIdx.x & (WAVE - 1);
    const int sl = lb * (PKB / WAVE) + threadIdx.x / WAVE;
    if (sl < n_slices && lane == 0)
        smeta[sl] = make_int4(voff[sl], mode ? chwoff[sl] : c32off[sl], mode, 0);
    if (threadIdx.x == 0) bmeta[lb] = make_int2(mode ? hoff[lb] : 0, bh[lb]);
    if (br.slot < 0) return;
    const int vo = voff[sl];
    const int co = mode ? chwoff[sl] : c32off[sl];
    const int m = br.m;
    unsigned g0 = 0u, g1 = 0u, g2 = 0u;      // the row's current group of eight 12-bit indices
    for (int q0 = 0; q0 < m; q0 += PK_U) {
        int crs[PK_U];
        double avs[PK_U];
#pragma unroll
        for (int u = 0; u < PK_U; u++) {
            const size_t o = sell_off(br.slot, q0 + u, stride);
            crs[u] = (q0 + u < m) ? Acol[o] : 0;
            avs[u] = (q0 + u < m) ? Aval[o] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < PK_U; u++)
            if (q0 + u < m) crs[u] = row_of_slot[crs[u]];
#pragma unroll
        for (int u = 0; u < PK_U; u++) {
            const int q = q0 + u;
            if (q >= m) break;
            const int cr = crs[u];
            val[(size_t)(vo + q) * 64u + lane] = avs[u];
            if (mode) {
                const int c = cr - br.bmn;
                const unsigned below = bits[c >> 5] & ((1u << (c & 31)) - 1u);
                const unsigned li = (unsigned)(pre[c >> 5] + __popc(below));   // < PK_HALO_CAP
                const int bit = 12 * (q & 7), wd = bit >> 5, sh = bit & 31;
                const unsigned lo = li << sh, hi = sh > 20 ? li >> (32 - sh) : 0u;
                if (wd == 0) {
                    g0 |= lo;
                    g1 |= hi;
                } else if (wd == 1) {
                    g1 |= lo;
                    g2 |= hi;
                } else {
                    g2 |= lo;
                }
                if ((q & 7) == 7 || q == m - 1) {
                    const size_t w = (size_t)(co + (q >> 3) * 3) * 64u + lane;
                    chw[w] = g0;
                    chw[w + 64u] = g1;
                    chw[w + 128u] = g2;
                    g0 = g1 = g2 = 0u;
                }
            } else {
                c32[(size_t)(co + q) * 64u + lane] = cr;
            }
        }
    }
    len[br.k] = m;
    pdiag[br.k] = diag[br.slot];
    pb[br.k] = b[br.slot];
}

__global__ void __launch_bounds__(PKB)
k_pk_pass2(int n_eq, int n_slices, int stride, const int *__restrict__ slot_of_row,
           const int *__restrict__ row_of_slot, const int *__restrict__ Alen, const double *__restrict__ Aval,
           const int *__restrict__ Acol, const double *__restrict__ diag, const double *__restrict__ b,
           const int *__restrict__ bmode, const int *__restrict__ bh, const int *__restrict__ hoff,
           const int *__restrict__ voff, const int *__restrict__ chwoff, const int *__restrict__ c32off,
           int *__restrict__ halo, double *__restrict__ val, unsigned *__restrict__ chw,
           int *__restrict__ c32, int *__restrict__ len, double *__restrict__ pdiag, double *__restrict__ pb,
           int4 *__restrict__ smeta, int2 *__restrict__ bmeta)
{
    pk_pass2_body(blockIdx.x, n_eq, n_slices, stride, slot_of_row, row_of_slot, Alen, Aval, Acol, diag, b, bmode, bh, hoff, voff, chwoff, c32off, halo, val, chw, c32, len, pdiag, pb, smeta, bmeta);
}

// FP32 copy of the off-diagonals and each row's Dirichlet part
// d_k = A_kk + sum_j A_kj in FP64 (see the mixed solver in k_cg.hip).
__device__ __forceinline__ void pk_f32_body(int lb, int n_eq, const int *__restrict__ len, const int4 *__restrict__ smeta,
                         const double *__restrict__ val, const double *__restrict__ pdiag,
                         float *__restrict__ fval, double *__restrict__ ddir)
{
    const int k = lb * PKB + threadIdx.x;
    if (k >= n_eq) return;
    const int m = len[k];
    const size_t base = (size_t)smeta[k >> 6].x * 64u + (size_t)(k & 63);
    double d = pdiag[k];
    for (int q0 = 0; q0 < m; q0 += PK_U) {
        double a[PK_U];
#pragma unroll
        for (int u = 0; u < PK_U; u++) a[u] = (q0 + u < m) ? val[base + (size_t)(q0 + u) * 64u] : 0.0;
#pragma unroll
        for (int u = 0; u < PK_U; u++) {
            if (q0 + u >= m) break;
            fval[base + (size_t)(q0 + u) * 64u] = (float)a[u];
            d += a[u];
        }
    }
    ddir[k] = d;
}

__global__ void __launch_bounds__(PKB)
k_pk_f32(int n_eq, const int *__restrict__ len, const int4 *__restrict__ smeta,
                         const double *__restrict__ val, const double *__restrict__ pdiag,
                         float *__restrict__ fval, double *__restrict__ ddir)
{
    pk_f32_body(blockIdx.x, n_eq, len, smeta, val, pdiag, fval, ddir);
}

__device__ __forceinline__ void pk_unpack_x_body(int lb, int n, const int *__restrict__ unk, const int *__restrict__ row_of_slot,
                              const double *__restrict__ xp, double *__restrict__ x)
{
    const int s = lb * PKB + threadIdx.x;
    if (s >= n) return;
    x[s] = unk[s] ? xp[row_of_slot[s]] : 0.0;
}

__global__ void __launch_bounds__(PKB)
k_pk_unpack_x(int n, const int *__restrict__ unk, const int *__restrict__ row_of_slot,
                              const double *__restrict__ xp, double *__restrict__ x)
{
    pk_unpack_x_body(blockIdx.x, n, unk, row_of_slot, xp, x);
}

template <typename T>
static int grow(T **p, size_t *cap, size_t need)
{
    if (*p && *cap >= need) return 0;
    if (*p) hipFree(*p);
    *p = nullptr;
    size_t n = need + need / 8 + 64;
    HIP_OK(hipMalloc((void **)p, sizeof(T) * n));
    *cap = n;
    return 0;
}

#define GROW(p, cap, need)                     \
    do {                                       \
        size_t c_ = (cap);                     \
        int rc_ = grow(&(p), &c_, (need));     \
        if (rc_) return rc_;                   \
        (cap) = c_;                            \
    } while (0)

void packed_free(Packed *p)
{
    if (!p) return;
    void *bufs[] = {p->row_of_slot, p->slot_of_row, p->len, p->diag, p->b, p->x, p->sw, p->whw, p->w32,
                    p->bh, p->bmode, p->voff, p->chwoff, p->c32off, p->hoff, p->smeta, p->bmeta, p->halo,
                    p->val, p->chw, p->c32, p->fval, p->ddir, p->dmax, p->cq};
    for (void *b : bufs)
        if (b) hipFree(b);
    if (p->htot) hipHostFree(p->htot);
    delete p;
}

// Row-, slice- and block-sized buffers for up to n rows.
static int pk_reserve_rows(Packed *p, int n)
{
    if (p->cap_slots < n + 1) {
        if (p->row_of_slot) hipFree(p->row_of_slot);
        p->row_of_slot = nullptr;
        HIP_OK(hipMalloc(&p->row_of_slot, sizeof(int) * ((size_t)n + 1)));
        p->cap_slots = n + 1;
    }
    const int rows = (n + PKB - 1) / PKB * PKB;
    if (p->cap_rows < rows) {
        void *old[] = {p->slot_of_row, p->len, p->diag, p->b, p->x, p->ddir, p->sw, p->whw, p->w32, p->bh,
                       p->bmode, p->voff, p->chwoff, p->c32off, p->hoff, p->smeta, p->bmeta};
        for (void *b : old)
            if (b) hipFree(b);
        const size_t ns = (size_t)rows / 64 + 1, nb = (size_t)rows / PKB + 1;
        HIP_OK(hipMalloc(&p->slot_of_row, sizeof(int) * rows));
        HIP_OK(hipMalloc(&p->len, sizeof(int) * rows));
        HIP_OK(hipMalloc(&p->diag, sizeof(double) * rows));
        HIP_OK(hipMalloc(&p->b, sizeof(double) * rows));
        HIP_OK(hipMalloc(&p->x, sizeof(double) * rows));
        HIP_OK(hipMalloc(&p->ddir, sizeof(double) * rows));
        HIP_OK(hipMalloc(&p->sw, sizeof(int) * ns));
        HIP_OK(hipMalloc(&p->whw, sizeof(int) * ns));
        HIP_OK(hipMalloc(&p->w32, sizeof(int) * ns));
        HIP_OK(hipMalloc(&p->voff, sizeof(int) * ns));
        HIP_OK(hipMalloc(&p->chwoff, sizeof(int) * ns));
        HIP_OK(hipMalloc(&p->c32off, sizeof(int) * ns));
        HIP_OK(hipMalloc(&p->smeta, sizeof(int4) * ns));
        HIP_OK(hipMalloc(&p->bh, sizeof(int) * nb));
        HIP_OK(hipMalloc(&p->bmode, sizeof(int) * nb));
        HIP_OK(hipMalloc(&p->hoff, sizeof(int) * nb));
        HIP_OK(hipMalloc(&p->bmeta, sizeof(int2) * nb));
        p->cap_rows = rows;
    }
    if (!p->dmax) HIP_OK(hipMalloc(&p->dmax, sizeof(int) * 8));
    if (!p->htot) HIP_OK(hipHostMalloc(&p->htot, sizeof(int) * 8, hipHostMallocDefault));
    return 0;
}

// pack_ppe in three stages, each ending with copies of sizes to pinned
// host memory that the next stage reads once the stream has reached them
// (pack_wait).  Callers packing many systems run a stage for all of them
// before waiting, so the systems' kernels overlap.
int pack_stage_a(Ctx *c)
{
    if (!c->pk) c->pk = new Packed();
    Packed *p = c->pk;
    const int n = c->n;
    hipStream_t s = c->stream;
    int rc = pk_reserve_rows(p, n);
    if (rc) return rc;
    prof_begin(PK_PACK, s);
    rc = launch_scan_excl(c->unk, p->row_of_slot, n, c->scan_blk, s);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(&p->htot[0], p->row_of_slot + n, sizeof(int), hipMemcpyDeviceToHost, s));
    return 0;
}

int pack_wait(Ctx *c)
{
    HIP_OK(hipStreamSynchronize(c->stream));
    return 0;
}

int pack_stage_b(Ctx *c)
{
    Packed *p = c->pk;
    const int n = c->n;
    hipStream_t s = c->stream;
    int rc = 0;
    const int n_eq = p->htot[0];
    p->n_eq = n_eq;
    p->n_slices = (n_eq + 63) / 64;
    p->n_blocks = (n_eq + PKB - 1) / PKB;
    p->hmax = 0;
    if (n_eq == 0) {
        prof_end(PK_PACK, s);
        return 0;
    }
    const int ns = p->n_slices, nb = p->n_blocks;
    hipLaunchKernelGGL(k_pk_slots, dim3(pk_blocks(n)), dim3(PKB), 0, s, n, c->unk, p->row_of_slot,
                       p->slot_of_row);
    rc = launch_fill_int(p->dmax, 0, 1, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pk_pass1, dim3(nb), dim3(PKB), 0, s, n_eq, ns, c->stride, p->slot_of_row,
                       p->row_of_slot, c->Alen, c->Acol, p->sw, p->whw, p->w32, p->bh, p->bmode, p->dmax);
    LAUNCH_CHECK();
    // value, halo-word and 32-bit column offsets per slice, halo offsets
    // per block: one scan set, whose totals land beside the largest halo
    // so that one copy brings all five back
    const ScanSet ss = {{p->sw, p->whw, p->w32, p->bh}, {p->voff, p->chwoff, p->c32off, p->hoff},
                        {ns, ns, ns, nb}, 4};
    rc = launch_scan_excl_multi(ss, c->scan_blk, p->dmax + 1, s);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(&p->htot[1], p->dmax, sizeof(int) * 5, hipMemcpyDeviceToHost, s));
    return 0;
}

int pack_stage_c(Ctx *c, int with_f32)
{
    Packed *p = c->pk;
    hipStream_t s = c->stream;
    const int n_eq = p->n_eq;
    if (n_eq == 0) return 0;
    const int ns = p->n_slices, nb = p->n_blocks;
    p->hmax = p->htot[1];
    p->tot_val = (size_t)p->htot[2];
    p->tot_hw = (size_t)p->htot[3];
    p->tot_c32 = (size_t)p->htot[4];
    p->tot_halo = (size_t)p->htot[5];
    GROW(p->val, p->cap_val, p->tot_val * 64);
    GROW(p->chw, p->cap_hw, p->tot_hw * 64);
    GROW(p->c32, p->cap_c32, p->tot_c32 * 64);
    GROW(p->halo, p->cap_halo, p->tot_halo);
    hipLaunchKernelGGL(k_pk_pass2, dim3(nb), dim3(PKB), 0, s, n_eq, ns, c->stride, p->slot_of_row, p->row_of_slot,
                       c->Alen, c->Aval, c->Acol, c->diag, c->b, p->bmode, p->bh, p->hoff, p->voff, p->chwoff,
                       p->c32off, p->halo, p->val, p->chw, p->c32, p->len, p->diag, p->b,
                       p->smeta, p->bmeta);
    LAUNCH_CHECK();
    if (with_f32) {
        GROW(p->fval, p->cap_fval, p->tot_val * 64);
        hipLaunchKernelGGL(k_pk_f32, dim3(pk_blocks(n_eq)), dim3(PKB), 0, s, n_eq, p->len, p->smeta, p->val,
                           p->diag, p->fval, p->ddir);
        LAUNCH_CHECK();
    }
    prof_end(PK_PACK, s);
    return 0;
}

int pack_ppe(Ctx *c, int with_f32)
{
    int rc = pack_stage_a(c);
    if (!rc) rc = pack_wait(c);
    if (!rc) rc = pack_stage_b(c);
    if (!rc) rc = pack_wait(c);
    if (!rc) rc = pack_stage_c(c, with_f32);
    return rc;
}

int unpack_solution(Ctx *c)
{
    Packed *p = c->pk;
    const int n = c->n;
    if (!p || p->n_eq == 0) return launch_fill_double(c->x, 0.0, (size_t)n, c->stream);
    hipLaunchKernelGGL(k_pk_unpack_x, dim3(pk_blocks(n)), dim3(PKB), 0, c->stream, n, c->unk, p->row_of_slot,
                       p->x, c->x);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------
// Packing for a batch of scenes: each stage as one launch over the
// scenes' block ranges (block -> scene by bisection), the scans as one
// segmented scan, and the sizes the host needs (unknowns, then slice,
// halo and block totals) brought back for all scenes in one copy each.
// Every scene's packed system is the one pack_ppe builds alone.
// ---------------------------------------------------------------------

struct FlatPack {
    int n, n_eq, n_slices, stride;
    const int *unk, *Alen, *Acol;
    const double *Aval, *diag, *b;
    int *row_of_slot, *slot_of_row;
    int *sw, *whw, *w32, *bh, *bmode, *dmax, *voff, *chwoff, *c32off, *hoff, *halo;
    double *val;
    unsigned *chw;
    int *c32, *len;
    double *pdiag, *pb;
    int4 *smeta;
    int2 *bmeta;
    float *fval;
    double *ddir, *px, *x;
};

// A scene of the table with its buffers as global pointers (gptr).
__device__ __forceinline__ FlatPack pack_global(const FlatPack &s)
{
    FlatPack g = s;
    g.unk = gptr(s.unk); g.Alen = gptr(s.Alen); g.Acol = gptr(s.Acol);
    g.Aval = gptr(s.Aval); g.diag = gptr(s.diag); g.b = gptr(s.b);
    g.row_of_slot = gptr(s.row_of_slot); g.slot_of_row = gptr(s.slot_of_row);
    g.sw = gptr(s.sw); g.whw = gptr(s.whw); g.w32 = gptr(s.w32); g.bh = gptr(s.bh);
    g.bmode = gptr(s.bmode); g.dmax = gptr(s.dmax); g.voff = gptr(s.voff); g.chwoff = gptr(s.chwoff);
    g.c32off = gptr(s.c32off); g.hoff = gptr(s.hoff); g.halo = gptr(s.halo);
    g.val = gptr(s.val); g.chw = gptr(s.chw); g.c32 = gptr(s.c32); g.len = gptr(s.len);
    g.pdiag = gptr(s.pdiag); g.pb = gptr(s.pb); g.smeta = gptr(s.smeta); g.bmeta = gptr(s.bmeta);
    g.fval = gptr(s.fval); g.ddir = gptr(s.ddir); g.px = gptr(s.px); g.x = gptr(s.x);
    return g;
}

struct FlatPackBufs {
    std::vector<FlatPack> h;
    FlatPack *tab = nullptr;
    int cap = 0;
    ScanSegs sc_unk, sc_set;
    std::vector<ScanSeg> hseg;
    FlatGrid g_rows;
    int cap_rows_first = 0;
    int *tot = nullptr, *htot = nullptr;   // 8 ints per scene: [0] unknowns / largest halo, [1..4] scan totals
    int *blk = nullptr;
    size_t cap_blk = 0;
};
static FlatPackBufs g_fpk;

void flat_pack_free()
{
    FlatPackBufs &F = g_fpk;
    void *d[] = {F.tab, F.sc_unk.segs, F.sc_unk.tile0, F.sc_set.segs, F.sc_set.tile0, F.g_rows.first, F.tot, F.blk};
    for (void *p : d)
        if (p) hipFree(p);
    if (F.htot) hipHostFree(F.htot);
    F = FlatPackBufs();
}

static int fp_reserve(int ns)
{
    FlatPackBufs &F = g_fpk;
    if (F.cap >= ns) return 0;
    if (F.tab) hipFree(F.tab);
    if (F.tot) hipFree(F.tot);
    if (F.htot) hipHostFree(F.htot);
    F.tab = nullptr;
    F.tot = F.htot = nullptr;
    HIP_OK(hipMalloc(&F.tab, sizeof(FlatPack) * ns));
    HIP_OK(hipMalloc(&F.tot, sizeof(int) * 8 * ns));
    HIP_OK(hipHostMalloc(&F.htot, sizeof(int) * 8 * ns, hipHostMallocDefault));
    F.cap = ns;
    return 0;
}

static int fp_blk(size_t need)
{
    FlatPackBufs &F = g_fpk;
    if (F.blk && F.cap_blk >= need) return 0;
    if (F.blk) hipFree(F.blk);
    F.blk = nullptr;
    HIP_OK(hipMalloc(&F.blk, sizeof(int) * (need + 64)));
    F.cap_blk = need + 64;
    return 0;
}

// The table of the scenes' packed buffers, uploaded (stream-ordered).
static int fp_table(Ctx *const *cs, int ns, hipStream_t s)
{
    FlatPackBufs &F = g_fpk;
    F.h.resize(ns);
    for (int q = 0; q < ns; q++) {
        const Ctx *c = cs[q];
        const Packed *p = c->pk;
        FlatPack &t = F.h[q];
        t.n = c->n;
        t.n_eq = p->n_eq;
        t.n_slices = p->n_slices;
        t.stride = c->stride;
        t.unk = c->unk; t.Alen = c->Alen; t.Acol = c->Acol; t.Aval = c->Aval; t.diag = c->diag; t.b = c->b;
        t.row_of_slot = p->row_of_slot; t.slot_of_row = p->slot_of_row;
        t.sw = p->sw; t.whw = p->whw; t.w32 = p->w32; t.bh = p->bh; t.bmode = p->bmode; t.dmax = p->dmax;
        t.voff = p->voff; t.chwoff = p->chwoff; t.c32off = p->c32off; t.hoff = p->hoff; t.halo = p->halo;
        t.val = p->val; t.chw = p->chw; t.c32 = p->c32; t.len = p->len; t.pdiag = p->diag; t.pb = p->b;
        t.smeta = p->smeta; t.bmeta = p->bmeta; t.fval = p->fval; t.ddir = p->ddir; t.px = p->x; t.x = c->x;
    }
    HIP_OK(hipMemcpyAsync(F.tab, F.h.data(), sizeof(FlatPack) * ns, hipMemcpyHostToDevice, s));
    // the host table is rewritten by the next call: wait for the copy
    HIP_OK(hipStreamSynchronize(s));
    return 0;
}

#define FLAT_PACK()                                         \
    const int s_ = flat_scene(first, ns, blockIdx.x);       \
    const FlatPack P = pack_global(tab[s_]);                \
    const int lb = blockIdx.x - first[s_]

__global__ void __launch_bounds__(PKB)
k_pk_slots_flat(const FlatPack *__restrict__ tab, const int *__restrict__ first, int ns)
{
    FLAT_PACK();
    pk_slots_body(lb, P.n, P.unk, P.row_of_slot, P.slot_of_row);
}

__global__ void k_pk_dmax_zero_flat(const FlatPack *__restrict__ tab, int ns)
{
    for (int s = threadIdx.x; s < ns; s += blockDim.x) tab[s].dmax[0] = 0;
}

__global__ void k_pk_dmax_gather_flat(const FlatPack *__restrict__ tab, int ns, int *out)
{
    for (int s = threadIdx.x; s < ns; s += blockDim.x)
        for (int j = 0; j < 5; j++) out[8 * s + j] = tab[s].dmax[j];
}

__global__ void __launch_bounds__(PKB)
k_pk_pass1_flat(const FlatPack *__restrict__ tab, const int *__restrict__ first, int ns)
{
    FLAT_PACK();
    pk_pass1_body(lb, P.n_eq, P.n_slices, P.stride, P.slot_of_row, P.row_of_slot, P.Alen, P.Acol, P.sw, P.whw,
                  P.w32, P.bh, P.bmode, P.dmax);
}

__global__ void __launch_bounds__(PKB)
k_pk_pass2_flat(const FlatPack *__restrict__ tab, const int *__restrict__ first, int ns)
{
    FLAT_PACK();
    pk_pass2_body(lb, P.n_eq, P.n_slices, P.stride, P.slot_of_row, P.row_of_slot, P.Alen, P.Aval, P.Acol,
                  P.diag, P.b, P.bmode, P.bh, P.hoff, P.voff, P.chwoff, P.c32off, P.halo, P.val, P.chw, P.c32,
                  P.len, P.pdiag, P.pb, P.smeta, P.bmeta);
}

__global__ void __launch_bounds__(PKB)
k_pk_f32_flat(const FlatPack *__restrict__ tab, const int *__restrict__ first, int ns)
{
    FLAT_PACK();
    pk_f32_body(lb, P.n_eq, P.len, P.smeta, P.val, P.pdiag, P.fval, P.ddir);
}

__global__ void __launch_bounds__(PKB)
k_pk_unpack_x_flat(const FlatPack *__restrict__ tab, const int *__restrict__ first, int ns)
{
    FLAT_PACK();
    pk_unpack_x_body(lb, P.n, P.unk, P.row_of_slot, P.px, P.x);
}

// Packs the assembled PPE of every scene (as pack_ppe does each alone);
// gb holds the scenes' 256-thread block offsets over their particles.
int flat_pack(Ctx *const *cs, int ns, const FlatGrid &gb, int with_f32, hipStream_t s)
{
    FlatPackBufs &F = g_fpk;
    int rc = fp_reserve(ns);
    if (rc) return rc;
    for (int q = 0; q < ns; q++) {
        if (!cs[q]->pk) cs[q]->pk = new Packed();
        rc = pk_reserve_rows(cs[q]->pk, cs[q]->n);
        if (rc) return rc;
    }
    prof_begin(PK_PACK, s);
    // stage a: the unknowns' row numbers, the counts to the host
    F.hseg.resize(4 * (size_t)ns);
    for (int q = 0; q < ns; q++)
        F.hseg[q] = ScanSeg{cs[q]->unk, cs[q]->pk->row_of_slot, cs[q]->n, F.tot + 8 * q};
    rc = scan_segs_set(&F.sc_unk, F.hseg.data(), ns, s);
    if (!rc) rc = fp_blk((size_t)F.sc_unk.ntiles + ns);
    if (!rc) rc = launch_scan_seg(F.sc_unk, F.blk, s);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(F.htot, F.tot, sizeof(int) * 8 * ns, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    std::vector<int> nb(ns);
    for (int q = 0; q < ns; q++) {
        Packed *p = cs[q]->pk;
        p->n_eq = F.htot[8 * q];
        p->n_slices = (p->n_eq + 63) / 64;
        p->n_blocks = (p->n_eq + PKB - 1) / PKB;
        p->hmax = 0;
        nb[q] = p->n_blocks;
    }
    rc = fp_table(cs, ns, s);
    if (!rc) rc = flat_build_grid(&F.g_rows, nb.data(), ns, s);
    if (rc) return rc;
    // stage b: slice widths, halos and modes, their scans
    if (gb.blocks > 0) hipLaunchKernelGGL(k_pk_slots_flat, dim3(gb.blocks), dim3(PKB), 0, s, F.tab, gb.first, ns);
    hipLaunchKernelGGL(k_pk_dmax_zero_flat, dim3(1), dim3(64), 0, s, F.tab, ns);
    if (F.g_rows.blocks > 0)
        hipLaunchKernelGGL(k_pk_pass1_flat, dim3(F.g_rows.blocks), dim3(PKB), 0, s, F.tab, F.g_rows.first, ns);
    LAUNCH_CHECK();
    for (int q = 0; q < ns; q++) {
        Packed *p = cs[q]->pk;
        F.hseg[4 * q + 0] = ScanSeg{p->sw, p->voff, p->n_slices, p->dmax + 1};
        F.hseg[4 * q + 1] = ScanSeg{p->whw, p->chwoff, p->n_slices, p->dmax + 2};
        F.hseg[4 * q + 2] = ScanSeg{p->w32, p->c32off, p->n_slices, p->dmax + 3};
        F.hseg[4 * q + 3] = ScanSeg{p->bh, p->hoff, p->n_blocks, p->dmax + 4};
    }
    rc = scan_segs_set(&F.sc_set, F.hseg.data(), 4 * ns, s);
    if (!rc) rc = fp_blk((size_t)F.sc_set.ntiles + 4 * ns);
    if (!rc) rc = launch_scan_seg(F.sc_set, F.blk, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pk_dmax_gather_flat, dim3(1), dim3(64), 0, s, F.tab, ns, F.tot);
    LAUNCH_CHECK();
    HIP_OK(hipMemcpyAsync(F.htot, F.tot, sizeof(int) * 8 * ns, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    // stage c: values, indices, halos, descriptors
    for (int q = 0; q < ns; q++) {
        Packed *p = cs[q]->pk;
        if (p->n_eq == 0) continue;
        const int *t = F.htot + 8 * q;
        p->hmax = t[0];
        p->tot_val = (size_t)t[1];
        p->tot_hw = (size_t)t[2];
        p->tot_c32 = (size_t)t[3];
        p->tot_halo = (size_t)t[4];
        GROW(p->val, p->cap_val, p->tot_val * 64);
        GROW(p->chw, p->cap_hw, p->tot_hw * 64);
        GROW(p->c32, p->cap_c32, p->tot_c32 * 64);
        GROW(p->halo, p->cap_halo, p->tot_halo);
        if (with_f32) GROW(p->fval, p->cap_fval, p->tot_val * 64);
    }
    rc = fp_table(cs, ns, s);
    if (rc) return rc;
    if (F.g_rows.blocks > 0) {
        hipLaunchKernelGGL(k_pk_pass2_flat, dim3(F.g_rows.blocks), dim3(PKB), 0, s, F.tab, F.g_rows.first, ns);
        if (with_f32)
            hipLaunchKernelGGL(k_pk_f32_flat, dim3(F.g_rows.blocks), dim3(PKB), 0, s, F.tab, F.g_rows.first, ns);
    }
    LAUNCH_CHECK();
    prof_end(PK_PACK, s);
    return 0;
}

// Every scene's packed solution into its slot-ordered x (unpack_solution).
int flat_unpack(Ctx *const *cs, int ns, const FlatGrid &gb, hipStream_t s)
{
    (void)cs;
    if (gb.blocks > 0)
        hipLaunchKernelGGL(k_pk_unpack_x_flat, dim3(gb.blocks), dim3(PKB), 0, s, g_fpk.tab, gb.first, ns);
    LAUNCH_CHECK();
    return 0;
}
