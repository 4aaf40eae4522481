// k_particles.hip -- particle-side kernels of one MPS timestep on gfx950.
//
// Design (one thread per particle slot, 256-thread blocks = 4 wave64):
//   * particles live in cell-sorted order: a counting sort by cell each step
//     (histogram fused into the collision kernel, exclusive scan, atomic
//     scatter, then a deterministic rank inside each cell by descending
//     original index -- the order the reference's prepend-linked cell lists
//     are walked in, neighbor_search.c:180-182, :211-233);
//   * the neighbour search walks the clamped 3x3(x3) cell block z->y->x and
//     writes each list in exactly the reference's order, so every per-
//     particle sum below accumulates in the reference's order;
//   * number density, free-surface flag, unknown flag, PPE diagonal
//     (coeff*sum of non-dummy weights) and the density-deviation source are
//     fused into the search: they need nothing but the list being built;
//   * lists and matrix rows are sliced-ELL (64-slot slices) so a wavefront's
//     k-th entries are one contiguous load.
// FP64 everywhere here, association order of the reference, no contraction.
#include <limits.h>

#include <vector>

#include "ctx.hpp"

#define BLK 256

static inline unsigned nblk(size_t n, unsigned b = BLK) { return (unsigned)((n + b - 1) / b); }

// The per-particle kernels (collision, predictor, search, assembly,
// gradient) at the block size particle_block() names (128, 256 or 512;
// option "pblk"): one thread per particle and integer block counters, so
// every size gives the same bits.
#define PARTICLE_LAUNCH(kern, c, ...)                                                                   \
    do {                                                                                              \
        const int pb_ = particle_block();                                                             \
        if (pb_ == 128)                                                                               \
            hipLaunchKernelGGL(kern<128>, dim3(nblk((c)->n, 128)), dim3(128), 0, (c)->stream, __VA_ARGS__); \
        else if (pb_ == 512)                                                                          \
            hipLaunchKernelGGL(kern<512>, dim3(nblk((c)->n, 512)), dim3(512), 0, (c)->stream, __VA_ARGS__); \
        else                                                                                          \
            hipLaunchKernelGGL(kern<256>, dim3(nblk((c)->n, 256)), dim3(256), 0, (c)->stream, __VA_ARGS__); \
    } while (0)

#define LAUNCH_CHECK()                                                       \
    do {                                                                     \
        hipError_t e_ = hipGetLastError();                                   \
        if (e_ != hipSuccess) {                                              \
            mps_set_error(MPS_ERR_HIP, __FILE__, __LINE__, hipGetErrorString(e_)); \
            return MPS_ERR_HIP;                                              \
        }                                                                    \
    } while (0)

// ---------------------------------------------------------------------
// fills and scans
// ---------------------------------------------------------------------

__global__ void k_fill_int(int *p, int v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_double(double *p, double v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

int launch_fill_int(int *p, int v, size_t n, hipStream_t s)
{
    if (n == 0) return 0;
    unsigned g = nblk(n);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill_int, dim3(g), dim3(BLK), 0, s, p, v, n);
    LAUNCH_CHECK();
    return 0;
}

int launch_fill_double(double *p, double v, size_t n, hipStream_t s)
{
    if (n == 0) return 0;
    unsigned g = nblk(n);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill_double, dim3(g), dim3(BLK), 0, s, p, v, n);
    LAUNCH_CHECK();
    return 0;
}

// Exclusive scan: out[0..n] with out[n] = total.  Three passes over
// 2048-element tiles (256 threads x 8 items): tile sums, a one-block scan of
// the tile sums (up to 8192 tiles = 16.7M elements), tile scans + offset.
#define SCAN_ITEMS 8
#define SCAN_TILE (BLK * SCAN_ITEMS)

__global__ void __launch_bounds__(BLK) k_scan_tiles(const int *in, int n, int *tile_sum)
{
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++)
        if (base + q < n) s += in[base + q];
    int tot;
    (void)block_scan_excl<BLK>(s, &tot);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_scan_tile_sums(int *tile_sum, int ntiles)
{
    // 1024 threads x 8 tiles each
    const int base = threadIdx.x * 8;
    int v[8];
    int s = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        v[q] = (base + q < ntiles) ? tile_sum[base + q] : 0;
        s += v[q];
    }
    int tot;
    int run = block_scan_excl<1024>(s, &tot);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (base + q < ntiles) tile_sum[base + q] = run;
        run += v[q];
    }
    if (threadIdx.x == 0) tile_sum[ntiles] = tot;
}

__global__ void __launch_bounds__(BLK) k_scan_apply(const int *in, int *out, int n,
                                                    const int *tile_sum, int ntiles)
{
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++) {
        v[q] = (base + q < n) ? in[base + q] : 0;
        s += v[q];
    }
    int tot;
    int run = block_scan_excl<BLK>(s, &tot) + tile_sum[blockIdx.x];
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++) {
        if (base + q < n) out[base + q] = run;
        run += v[q];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = tile_sum[ntiles];
}

int launch_scan_excl(const int *in, int *out, int n, int *blk, hipStream_t s)
{
    const int ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (ntiles > 8192) {
        mps_set_error(MPS_ERR_UNSUPPORTED, __FILE__, __LINE__, "scan longer than 16.7M elements");
        return MPS_ERR_UNSUPPORTED;
    }
    if (n == 0) {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, out, 0, (size_t)1);
        LAUNCH_CHECK();
        return 0;
    }
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(BLK), 0, s, in, n, blk);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(1024), 0, s, blk, ntiles);
    hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(BLK), 0, s, in, out, n, blk, ntiles);
    LAUNCH_CHECK();
    return 0;
}

// Up to four independent exclusive scans in the three launches of one:
// grid.y selects the array, and array k's tile sums use its own segment of
// the tile-sum buffer (MSCAN_SEG ints each: up to MSCAN_TILES tile sums and
// the total).  out[k][0..n[k]] as launch_scan_excl writes it, and tot[k] =
// its total.
#define MSCAN_SEG 2050
#define MSCAN_TILES 2048   // the tile-sum scan's 1024 threads x 2

__global__ void __launch_bounds__(BLK) k_mscan_tiles(ScanSet ss, int *blk)
{
    const int k = blockIdx.y, n = ss.n[k];
    if ((int)blockIdx.x * SCAN_TILE >= n) return;   // whole block: past this array's tiles
    const int *in = ss.in[k];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++)
        if (base + q < n) s += in[base + q];
    int tot;
    (void)block_scan_excl<BLK>(s, &tot);
    if (threadIdx.x == 0) blk[k * MSCAN_SEG + blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_mscan_tile_sums(ScanSet ss, int *blk, int *tot_out)
{
    const int k = blockIdx.x, n = ss.n[k];
    const int ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    int *tile_sum = blk + k * MSCAN_SEG;
    const int base = threadIdx.x * 2;   // 1024 threads x 2 tiles cover MSCAN_TILES
    int v[2];
    int s = 0;
#pragma unroll
    for (int q = 0; q < 2; q++) {
        v[q] = (base + q < ntiles) ? tile_sum[base + q] : 0;
        s += v[q];
    }
    int tot;
    int run = block_scan_excl<1024>(s, &tot);
#pragma unroll
    for (int q = 0; q < 2; q++) {
        if (base + q < ntiles) tile_sum[base + q] = run;
        run += v[q];
    }
    if (threadIdx.x == 0) {
        tile_sum[ntiles] = tot;
        tot_out[k] = tot;
        if (n == 0) ss.out[k][0] = 0;
    }
}

__global__ void __launch_bounds__(BLK) k_mscan_apply(ScanSet ss, const int *blk)
{
    const int k = blockIdx.y, n = ss.n[k];
    if ((int)blockIdx.x * SCAN_TILE >= n) return;
    const int ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    const int *in = ss.in[k], *tile_sum = blk + k * MSCAN_SEG;
    int *out = ss.out[k];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++) {
        v[q] = (base + q < n) ? in[base + q] : 0;
        s += v[q];
    }
    int tot;
    int run = block_scan_excl<BLK>(s, &tot) + tile_sum[blockIdx.x];
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++) {
        if (base + q < n) out[base + q] = run;
        run += v[q];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = tile_sum[ntiles];
}

int launch_scan_excl_multi(const ScanSet &ss, int *blk, int *tot, hipStream_t s)
{
    if (ss.count < 1 || ss.count > 4) {
        mps_set_error(MPS_ERR_ARG, __FILE__, __LINE__, "scan set of 1 to 4 arrays");
        return MPS_ERR_ARG;
    }
    int max_tiles = 0;
    for (int k = 0; k < ss.count; k++) {
        const int t = (ss.n[k] + SCAN_TILE - 1) / SCAN_TILE;
        if (ss.n[k] < 0 || t > MSCAN_TILES) {
            mps_set_error(MPS_ERR_UNSUPPORTED, __FILE__, __LINE__, "scan set array longer than 4.19M elements");
            return MPS_ERR_UNSUPPORTED;
        }
        if (t > max_tiles) max_tiles = t;
    }
    if (max_tiles > 0)
        hipLaunchKernelGGL(k_mscan_tiles, dim3(max_tiles, ss.count), dim3(BLK), 0, s, ss, blk);
    hipLaunchKernelGGL(k_mscan_tile_sums, dim3(ss.count), dim3(1024), 0, s, ss, blk, tot);
    if (max_tiles > 0)
        hipLaunchKernelGGL(k_mscan_apply, dim3(max_tiles, ss.count), dim3(BLK), 0, s, ss, blk);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------
// host AoS <-> device SoA
// ---------------------------------------------------------------------

struct SlotArrays {
    dvec *X, *V, *A;
    double *P, *ND;
    int *T, *O, *S, *NC;
};

static SlotArrays cur_arrays(Ctx *c)
{
    SlotArrays a;
    a.X = c->X[c->cx]; a.V = c->V[c->cx]; a.A = c->ACC[c->ct];
    a.P = c->P[c->cp]; a.ND = c->ND[c->ct];
    a.T = c->T[c->ct]; a.O = c->O[c->ct]; a.S = c->SURF[c->ct]; a.NC = c->NCOUNT[c->ct];
    return a;
}

__global__ void __launch_bounds__(BLK) k_unpack(const Particle *st, int n, SlotArrays a)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const Particle q = st[i];
    dvec x, v, ac;
#pragma unroll
    for (int d = 0; d < (D_ == 3 ? 4 : 2); d++) x.c[d] = v.c[d] = ac.c[d] = 0.0;
#pragma unroll
    for (int d = 0; d < D_; d++) {
        x.c[d] = q.pos[d];
        v.c[d] = q.vel[d];
        ac.c[d] = q.acc[d];
    }
    a.X[i] = x; a.V[i] = v; a.A[i] = ac;
    a.P[i] = q.pressure; a.ND[i] = q.n;
    a.T[i] = q.type; a.O[i] = i; a.S[i] = q.on_surface; a.NC[i] = q.neighbor_count;
}

__global__ void __launch_bounds__(BLK) k_pack(Particle *st, int n, SlotArrays a)
{
    const int s = blockIdx.x * BLK + threadIdx.x;
    if (s >= n) return;
    Particle q;
    const dvec x = a.X[s], v = a.V[s], ac = a.A[s];
#pragma unroll
    for (int d = 0; d < D_; d++) {
        q.pos[d] = x.c[d];
        q.vel[d] = v.c[d];
        q.acc[d] = ac.c[d];
    }
    q.pressure = a.P[s];
    q.n = a.ND[s];
    q.neighbor_count = a.NC[s];
    q.type = a.T[s];
    q.on_surface = a.S[s];
    st[a.O[s]] = q;
}

int launch_unpack(Ctx *c, const void *stage_dev, int)
{
    c->cx = c->ct = c->cp = 0;
    hipLaunchKernelGGL(k_unpack, dim3(nblk(c->n)), dim3(BLK), 0, c->stream,
                       (const Particle *)stage_dev, c->n, cur_arrays(c));
    LAUNCH_CHECK();
    return 0;
}

int launch_pack(Ctx *c, void *stage_dev)
{
    hipLaunchKernelGGL(k_pack, dim3(nblk(c->n)), dim3(BLK), 0, c->stream,
                       (Particle *)stage_dev, c->n, cur_arrays(c));
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------
// neighbour list <-> compact CSR over original indices
// ---------------------------------------------------------------------

__global__ void __launch_bounds__(BLK) k_counts_by_orig(int n, const int *O, const int *ncnt, int *cnt_orig)
{
    const int s = blockIdx.x * BLK + threadIdx.x;
    if (s < n) cnt_orig[O[s]] = ncnt[s];
}

__global__ void __launch_bounds__(BLK) k_list_compact(int n, int stride, const int *O, const int *nbr,
                                                      const int *ncnt, const int *off_orig, int *out)
{
    const int s = blockIdx.x * BLK + threadIdx.x;
    if (s >= n) return;
    const int m = ncnt[s];
    int *dst = out + off_orig[O[s]];
    for (int k = 0; k < m; k++) dst[k] = O[nbr[sell_off(s, k, stride)]];
}

// The imported list refers to original indices; the device order is the
// identity after an upload, so slots equal original indices.
__global__ void __launch_bounds__(BLK) k_list_expand(int n, int stride, const int *src, const int *counts,
                                                     const int *offsets, int *nbr, int *ncnt)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const int m = counts[i] < stride ? counts[i] : stride;
    const int *row = src + offsets[i];
    for (int k = 0; k < m; k++) nbr[sell_off(i, k, stride)] = row[k];
    ncnt[i] = m;
}

int launch_list_export(Ctx *c, long long *total_out)
{
    hipLaunchKernelGGL(k_counts_by_orig, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n,
                       c->O[c->ct], c->ncnt, c->cnt_orig);
    LAUNCH_CHECK();
    int rc = launch_scan_excl(c->cnt_orig, c->off_orig, c->n, c->scan_blk, c->stream);
    if (rc) return rc;
    int total = 0;
    HIP_OK(hipMemcpyAsync(&total, c->off_orig + c->n, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if ((size_t)total > c->nl_stage_elems) {
        if (c->nl_stage) HIP_OK(hipFree(c->nl_stage));
        c->nl_stage = nullptr;
        HIP_OK(hipMalloc(&c->nl_stage, sizeof(int) * (size_t)(total > 0 ? total : 1)));
        c->nl_stage_elems = (size_t)(total > 0 ? total : 1);
    }
    hipLaunchKernelGGL(k_list_compact, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->stride,
                       c->O[c->ct], c->nbr, c->ncnt, c->off_orig, c->nl_stage);
    LAUNCH_CHECK();
    *total_out = total;
    return 0;
}

int launch_list_import(Ctx *c, const int *counts_dev, const int *offsets_dev)
{
    hipLaunchKernelGGL(k_list_expand, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->stride,
                       c->nl_stage, counts_dev, offsets_dev, c->nbr, c->ncnt);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------
// 1. acceleration = gravity + viscosity (stale list), predictor
//    simulation.c:31-67, operators.c:50-85
// ---------------------------------------------------------------------

// Hybrid wall friction of the 3-D serial program (mps_fluid3D/src/
// operators.c:337-463) on the predicted velocity v of fluid slot i, over
// the stale list, in the reference's arithmetic and order.  A particle is
// on the contact line when its fluid-only density is below
// surface_threshold n0 and a wall particle lies within re_st; the nearest
// one (first strict minimum in list order) gives the wall normal from its
// own wall neighbours, the particle's fluid neighbours the free-surface
// normal, their angle the contact angle.  Slow particles (|v| <= v_c) are
// held when that angle is within delta_theta of the wetting angle; faster
// ones lose slip velocity along the wall, at most all of it.  Only v
// changes and only positions and types are read, so particles are
// independent, as in the reference's serial loop.
__device__ void wall_friction(int i, const dvec &xi, const dvec *__restrict__ X, const int *__restrict__ T,
                              const int *__restrict__ nbr, const int *__restrict__ ncnt, const StepConsts &k,
                              double v[D_])
{
    const int m = ncnt[i];
    double nf = 0.0, r2s = 1.0e30, dr[D_];
    int js = -1;
    for (int q = 0; q < m; q++) {
        const int j = nbr[sell_off(i, q, k.stride)];
        const double r2 = dist2(xi, X[j], dr);
        const int tj = T[j];
        if (tj == FLUID_PARTICLE) {
            if (r2 < k.re_n2) nf += wfun(sqrt(r2), k.re_n, k.r_min);
        } else if (tj == WALL_PARTICLE && r2 < k.re_st * k.re_st && r2 < r2s) {
            r2s = r2;
            js = j;
        }
    }
    if (js < 0 || nf >= k.thr_n0) return;
    double nw[D_], nlv[D_], es[D_];
#pragma unroll
    for (int d = 0; d < D_; d++) nw[d] = nlv[d] = 0.0;
    const dvec xw = X[js];
    const int mw = ncnt[js];
    for (int q = 0; q < mw; q++) {
        const int mm = nbr[sell_off(js, q, k.stride)];
        if (T[mm] != WALL_PARTICLE) continue;
        const double r = sqrt(dist2(X[mm], xw, dr));
        if (r < k.st_tiny) continue;
        const double w = wfun(r, k.re_st, k.r_min);
#pragma unroll
        for (int d = 0; d < D_; d++) nw[d] += w * dr[d] / r;
    }
    double m2 = 0.0;
#pragma unroll
    for (int d = 0; d < D_; d++) m2 += nw[d] * nw[d];
    const double magw = sqrt(m2);
    if (magw < 1.0e-9) return;
#pragma unroll
    for (int d = 0; d < D_; d++) nw[d] /= magw;
    int nfl = 0;
    for (int q = 0; q < m; q++) {
        const int j = nbr[sell_off(i, q, k.stride)];
        if (T[j] != FLUID_PARTICLE) continue;
        const double r = sqrt(dist2(xi, X[j], dr));
        if (r < k.st_tiny) continue;
        const double w = wfun(r, k.re_st, k.r_min);
#pragma unroll
        for (int d = 0; d < D_; d++) nlv[d] -= w * dr[d] / r;
        nfl++;
    }
    if (nfl == 0) return;
    m2 = 0.0;
#pragma unroll
    for (int d = 0; d < D_; d++) m2 += nlv[d] * nlv[d];
    const double magl = sqrt(m2);
    if (magl < 1.0e-9) return;
#pragma unroll
    for (int d = 0; d < D_; d++) nlv[d] /= magl;
    double cs = 0.0;
#pragma unroll
    for (int d = 0; d < D_; d++) cs -= nw[d] * nlv[d];
    const double theta = acos(fmax(-1.0, fmin(1.0, cs)));
    double dot = 0.0;
#pragma unroll
    for (int d = 0; d < D_; d++) dot += nlv[d] * nw[d];
    m2 = 0.0;
#pragma unroll
    for (int d = 0; d < D_; d++) {
        es[d] = nlv[d] - dot * nw[d];
        m2 += es[d] * es[d];
    }
    const double mage = sqrt(m2);
    if (mage < 1.0e-9) return;
#pragma unroll
    for (int d = 0; d < D_; d++) es[d] /= mage;
    double v2 = 0.0;
#pragma unroll
    for (int d = 0; d < D_; d++) v2 += v[d] * v[d];
    if (sqrt(v2) <= k.fr_vc) {
        if (fabs(theta - k.fr_theta_s) <= k.fr_dth)
#pragma unroll
            for (int d = 0; d < D_; d++) v[d] = 0.0;
        return;
    }
    double sw = 0.0;
    for (int q = 0; q < m; q++) {
        const int j = nbr[sell_off(i, q, k.stride)];
        if (T[j] != WALL_PARTICLE) continue;
        const double r = sqrt(dist2(xi, X[j], dr));
        if (r < k.st_tiny) continue;
        sw += wfun(r, k.re_st, k.r_min);
    }
    double vn = 0.0;
#pragma unroll
    for (int d = 0; d < D_; d++) vn += v[d] * es[d];
    double dvn = -k.fr_dyn * k.dt * sw * vn;
    if (dvn > vn) dvn = vn;
    if (dvn < -vn) dvn = -vn;
#pragma unroll
    for (int d = 0; d < D_; d++) v[d] += dvn * es[d];
}

// the friction alone on the current velocities (apply_friction drop-in)
__global__ void __launch_bounds__(BLK)
k_friction(int n, const dvec *__restrict__ X, dvec *__restrict__ V, const int *__restrict__ T,
           const int *__restrict__ nbr, const int *__restrict__ ncnt, StepConsts k)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n || T[i] != FLUID_PARTICLE) return;
    dvec v = V[i];
    wall_friction(i, X[i], X, T, nbr, ncnt, k, v.c);
    V[i] = v;
}

int launch_friction(Ctx *c, const StepConsts &k)
{
    hipLaunchKernelGGL(k_friction, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->X[c->cx], c->V[c->cx],
                       c->T[c->ct], c->nbr, c->ncnt, k);
    LAUNCH_CHECK();
    return 0;
}


// Neighbour loops load a batch of list entries, then the neighbours' data,
// before summing in list order: the gathers of a batch overlap.
#define ASM_U 8    // neighbours per load batch of the assembly
#define NB_U 4     // of the viscosity and pressure-gradient sums
#define COL_U 8    // of the collision (4 here and in the assembly, for fewer
                   // registers and more waves, measured 24-32 % slower)

template <int B>
__device__ __forceinline__ void visc_predict_body(int lb, int n, const dvec *__restrict__ X, const dvec *__restrict__ V, const int *__restrict__ T,
               const int *__restrict__ nbr, const int *__restrict__ ncnt, dvec *__restrict__ Xo,
               dvec *__restrict__ Vo, StepConsts k)
{
    const int i = lb * B + threadIdx.x;
    if (i >= n) return;
    const dvec xi = X[i], vi = V[i];
    if (T[i] != FLUID_PARTICLE) {
        Xo[i] = xi;
        Vo[i] = vi;
        return;
    }
    double lap[D_];
#pragma unroll
    for (int d = 0; d < D_; d++) lap[d] = 0.0;
    const int m = ncnt[i];
    // NB_U neighbours at a time: indices, then positions and velocities,
    // loaded together; summed in list order
    for (int q0 = 0; q0 < m; q0 += NB_U) {
        int jj[NB_U];
        dvec xx[NB_U], vv[NB_U];
#pragma unroll
        for (int u = 0; u < NB_U; u++) jj[u] = (q0 + u < m) ? nbr[sell_off(i, q0 + u, k.stride)] : i;
#pragma unroll
        for (int u = 0; u < NB_U; u++) {
            xx[u] = X[jj[u]];
            vv[u] = V[jj[u]];
        }
#pragma unroll
        for (int u = 0; u < NB_U; u++) {
            if (q0 + u >= m) break;
            double dr[D_];
            const double w = wfun(sqrt(dist2(xi, xx[u], dr)), k.re_lap, k.r_min);
#pragma unroll
            for (int d = 0; d < D_; d++) lap[d] += (vv[u].c[d] - vi.c[d]) * w;
        }
    }
    double acc[D_];
#pragma unroll
    for (int d = 0; d < D_; d++) acc[d] = k.g[d] + k.visc_coeff * lap[d];
    if (k.st_enabled) {
        // potential-type surface tension, liquid-liquid then solid-liquid
        // (operators.c:281-354), added partner by partner as the reference does
        for (int pass = 0; pass < 2; pass++) {
            if (pass == 1 && !k.st_sl_on) break;
            const int partner = pass == 0 ? FLUID_PARTICLE : WALL_PARTICLE;
            const double cst = pass == 0 ? k.st_ll : k.st_sl;
            for (int q = 0; q < m; q++) {
                const int j = nbr[sell_off(i, q, k.stride)];
                if (T[j] != partner) continue;
                double dr[D_];
                const double r = sqrt(dist2(xi, X[j], dr));
                if (r < k.st_tiny || r >= k.re_st) continue;
                const double f = -(r - k.l0) * (r - k.re_st);
#pragma unroll
                for (int d = 0; d < D_; d++) acc[d] += cst * f * (dr[d] / r);
            }
        }
    }
    dvec xo = xi, vo = vi;
#pragma unroll
    for (int d = 0; d < D_; d++) vo.c[d] = vi.c[d] + k.dt * acc[d];
    if (k.fric) wall_friction(i, xi, X, T, nbr, ncnt, k, vo.c);
#pragma unroll
    for (int d = 0; d < D_; d++) xo.c[d] = xi.c[d] + k.dt * vo.c[d];
    Xo[i] = xo;
    Vo[i] = vo;
}

template <int B>
__global__ void __launch_bounds__(B)
k_visc_predict(int n, const dvec *__restrict__ X, const dvec *__restrict__ V, const int *__restrict__ T,
               const int *__restrict__ nbr, const int *__restrict__ ncnt, dvec *__restrict__ Xo,
               dvec *__restrict__ Vo, StepConsts k)
{
    visc_predict_body<B>(blockIdx.x, n, X, V, T, nbr, ncnt, Xo, Vo, k);
}

int launch_visc_predict(Ctx *c, const StepConsts &k)
{
    prof_begin(PK_VISC, c->stream);
    PARTICLE_LAUNCH(k_visc_predict, c, c->n,
                       c->X[c->cx], c->V[c->cx], c->T[c->ct], c->nbr, c->ncnt,
                       c->X[1 - c->cx], c->V[1 - c->cx], k);
    prof_end(PK_VISC, c->stream);
    LAUNCH_CHECK();
    c->cx ^= 1;
    return 0;
}

// ---------------------------------------------------------------------
// 2. collision over the stale list (operators.c:371-446), then the cell of
//    every particle and the per-cell histogram of the counting sort
// ---------------------------------------------------------------------

template <int B>
__device__ __forceinline__ void collide_bin_body(int lb, int n, const dvec *__restrict__ X, const dvec *__restrict__ V, const int *__restrict__ T,
              const int *__restrict__ nbr, const int *__restrict__ ncnt, dvec *__restrict__ Xo,
              dvec *__restrict__ Vo, int *__restrict__ cell, int *__restrict__ ccount, StepConsts k,
              int collide)
{
    const int i = lb * B + threadIdx.x;
    if (i >= n) return;
    const int t = T[i];
    const dvec xi = X[i], vi = V[i];
    dvec xo = xi, vo = vi;
    if (collide && t == FLUID_PARTICLE) {
        double w[D_];
#pragma unroll
        for (int d = 0; d < D_; d++) w[d] = vi.c[d];
        const int m = ncnt[i];
        for (int q0 = 0; q0 < m; q0 += COL_U) {
            int jj[COL_U], tt[COL_U];
            dvec xx[COL_U];
#pragma unroll
            for (int u = 0; u < COL_U; u++) jj[u] = (q0 + u < m) ? nbr[sell_off(i, q0 + u, k.stride)] : i;
#pragma unroll
            for (int u = 0; u < COL_U; u++) {
                tt[u] = T[jj[u]];
                xx[u] = X[jj[u]];
            }
#pragma unroll
            for (int u = 0; u < COL_U; u++) {
                if (q0 + u >= m) break;
                const int j = jj[u];
                if (tt[u] == GHOST_PARTICLE) continue;
                double dr[D_];
                const double r2 = dist2(xi, xx[u], dr);
                if (r2 >= k.cd2) continue;
                const double r = sqrt(r2);
                if (r < k.col_tiny) continue;
                const dvec vj = V[j];
                double vn = 0.0;
#pragma unroll
                for (int d = 0; d < D_; d++) vn += (w[d] - vj.c[d]) * (dr[d] / r);
                if (vn <= 0.0) continue;
                const double imp = k.e_half * vn;
#pragma unroll
                for (int d = 0; d < D_; d++) w[d] -= imp * (dr[d] / r);
            }
        }
#pragma unroll
        for (int d = 0; d < D_; d++) {
            const double dv = w[d] - vi.c[d];
            xo.c[d] = xi.c[d] + k.dt * dv;
            vo.c[d] = w[d];
        }
    }
    Xo[i] = xo;
    Vo[i] = vo;
    int cc[3];
    int ci = (t == GHOST_PARTICLE) ? -1 : cell_coords(k.grid, xo, cc);
    if (ci < 0) ci = k.grid.total;
    cell[i] = ci;
    atomicAdd(&ccount[ci], 1);
}

template <int B>
__global__ void __launch_bounds__(B)
k_collide_bin(int n, const dvec *__restrict__ X, const dvec *__restrict__ V, const int *__restrict__ T,
              const int *__restrict__ nbr, const int *__restrict__ ncnt, dvec *__restrict__ Xo,
              dvec *__restrict__ Vo, int *__restrict__ cell, int *__restrict__ ccount, StepConsts k,
              int collide)
{
    collide_bin_body<B>(blockIdx.x, n, X, V, T, nbr, ncnt, Xo, Vo, cell, ccount, k, collide);
}

int launch_collide_bin(Ctx *c, const StepConsts &k, int collide)
{
    int rc = launch_fill_int(c->ccount, 0, (size_t)k.grid.total + 1, c->stream);
    if (rc) return rc;
    prof_begin(PK_COLLIDE, c->stream);
    PARTICLE_LAUNCH(k_collide_bin, c, c->n,
                       c->X[c->cx], c->V[c->cx], c->T[c->ct], c->nbr, c->ncnt,
                       c->X[1 - c->cx], c->V[1 - c->cx], c->cell, c->ccount, k, collide);
    prof_end(PK_COLLIDE, c->stream);
    LAUNCH_CHECK();
    c->cx ^= 1;
    return 0;
}

int launch_bin(Ctx *c, const StepConsts &k) { return launch_collide_bin(c, k, 0); }

// ---------------------------------------------------------------------
// 2-D collision (mps_fluid2D/src/operators.c:324-392): every j in ascending
// index order, not a neighbour list.  Only partners closer than the
// collision distance (< cell size) act, so each fluid particle visits them
// in ascending original index by repeated selection over its 3x3 cell block
// plus every non-ghost particle outside the grid (or over all particles if
// it is itself outside).  The sequence of velocity updates is the
// reference's.
// ---------------------------------------------------------------------

__global__ void __launch_bounds__(BLK)
k_bin_cells(int n, const dvec *__restrict__ X, const int *__restrict__ T, int *__restrict__ cell,
            int *__restrict__ ccount, Grid g)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    int cc[3];
    int ci;
    if (T[i] == GHOST_PARTICLE) ci = g.total + 1;
    else {
        ci = cell_coords(g, X[i], cc);
        if (ci < 0) ci = g.total;
    }
    cell[i] = ci;
    atomicAdd(&ccount[ci], 1);
}

__global__ void __launch_bounds__(BLK)
k_bucket_slots(int n, const int *__restrict__ cell, const int *__restrict__ cstart,
               int *__restrict__ cfill, int *__restrict__ bucket)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const int ci = cell[i];
    bucket[cstart[ci] + atomicAdd(&cfill[ci], 1)] = i;
}

__device__ __forceinline__ void collide_one(const dvec &xi, const dvec &xj, const dvec &vj,
                                            double w[D_], const StepConsts &k)
{
    double dr[D_];
    const double r2 = dist2(xi, xj, dr);
    if (r2 >= k.cd2) return;
    const double r = sqrt(r2);
    if (r < k.col_tiny) return;
    double vn = 0.0;
#pragma unroll
    for (int d = 0; d < D_; d++) vn += (w[d] - vj.c[d]) * (dr[d] / r);
    if (vn <= 0.0) return;
    const double imp = k.e_half * vn;
#pragma unroll
    for (int d = 0; d < D_; d++) w[d] -= imp * (dr[d] / r);
}

__global__ void __launch_bounds__(BLK)
k_collide2d(int n, const dvec *__restrict__ X, const dvec *__restrict__ V, const int *__restrict__ T,
            const int *__restrict__ O, const int *__restrict__ cstart, const int *__restrict__ bucket,
            dvec *__restrict__ Xo, dvec *__restrict__ Vo, StepConsts k)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const dvec xi = X[i], vi = V[i];
    if (T[i] != FLUID_PARTICLE) {
        Xo[i] = xi;
        Vo[i] = vi;
        return;
    }
    const Grid &g = k.grid;
    int cc[3];
    const bool inside = cell_coords(g, xi, cc) >= 0;
    int lo[3] = {0, 0, 0}, hi[3] = {-1, -1, 0};
    if (inside) {
        for (int d = 0; d < 2; d++) {
            lo[d] = cc[d] > 0 ? cc[d] - 1 : 0;
            hi[d] = cc[d] < g.nc[d] - 1 ? cc[d] + 1 : g.nc[d] - 1;
        }
    }
    double w[D_];
#pragma unroll
    for (int d = 0; d < D_; d++) w[d] = vi.c[d];
    int last = -1;
    for (;;) {
        int best = INT_MAX, best_slot = -1;
        // a candidate acts in the reference iff it is not a ghost and
        // !(r^2 >= cd^2) (NaN distances included)
#define CONSIDER(jj)                                                         \
        {                                                                    \
            const int j_ = (jj);                                             \
            if (j_ != i && T[j_] != GHOST_PARTICLE) {                        \
                double dr_[D_];                                              \
                if (!(dist2(xi, X[j_], dr_) >= k.cd2)) {                     \
                    const int o_ = O[j_];                                    \
                    if (o_ > last && o_ < best) { best = o_; best_slot = j_; } \
                }                                                            \
            }                                                                \
        }
        if (inside) {
            for (int cy = lo[1]; cy <= hi[1]; cy++)
                for (int cxx = lo[0]; cxx <= hi[0]; cxx++) {
                    const int cid = cell_key(g, cxx, cy, 0);
                    for (int u = cstart[cid]; u < cstart[cid + 1]; u++) CONSIDER(bucket[u]);
                }
            for (int u = cstart[g.total]; u < cstart[g.total + 1]; u++) CONSIDER(bucket[u]);
        } else {
            for (int j = 0; j < n; j++) CONSIDER(j);
        }
#undef CONSIDER
        if (best_slot < 0) break;
        collide_one(xi, X[best_slot], V[best_slot], w, k);
        last = best;
    }
    dvec xo = xi, vo = vi;
#pragma unroll
    for (int d = 0; d < D_; d++) {
        const double dv = w[d] - vi.c[d];
        xo.c[d] = xi.c[d] + k.dt * dv;
        vo.c[d] = w[d];
    }
    Xo[i] = xo;
    Vo[i] = vo;
}

int collide_2d(Ctx *c, const StepConsts &k)
{
    const int total = k.grid.total;
    int rc = launch_fill_int(c->ccount, 0, (size_t)total + 2, c->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bin_cells, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->X[c->cx],
                       c->T[c->ct], c->cell, c->ccount, k.grid);
    LAUNCH_CHECK();
    rc = launch_scan_excl(c->ccount, c->cstart, total + 2, c->scan_blk, c->stream);
    if (!rc) rc = launch_fill_int(c->cfill, 0, (size_t)total + 2, c->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bucket_slots, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->cell,
                       c->cstart, c->cfill, c->tmp_key);
    prof_begin(PK_COLLIDE, c->stream);
    hipLaunchKernelGGL(k_collide2d, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->X[c->cx],
                       c->V[c->cx], c->T[c->ct], c->O[c->ct], c->cstart, c->tmp_key,
                       c->X[1 - c->cx], c->V[1 - c->cx], k);
    prof_end(PK_COLLIDE, c->stream);
    LAUNCH_CHECK();
    c->cx ^= 1;
    return 0;
}

// ---------------------------------------------------------------------
// 3. counting sort by (cell, descending original index)
// ---------------------------------------------------------------------

__device__ __forceinline__ void scatter_body(int lb, int n, const int *__restrict__ cell, const int *__restrict__ O, const int *__restrict__ cstart,
          int *__restrict__ cfill, int *__restrict__ tmp_key, int *__restrict__ sflag, int total)
{
    const int i = lb * BLK + threadIdx.x;
    if (i >= n) return;
    const int ci = cell[i];
    sflag[i] = (ci == total);
    if (ci == total) return;
    const int at = atomicAdd(&cfill[ci], 1);
    tmp_key[cstart[ci] + at] = O[i];
}

__global__ void __launch_bounds__(BLK)
k_scatter(int n, const int *__restrict__ cell, const int *__restrict__ O, const int *__restrict__ cstart,
          int *__restrict__ cfill, int *__restrict__ tmp_key, int *__restrict__ sflag, int total)
{
    scatter_body(blockIdx.x, n, cell, O, cstart, cfill, tmp_key, sflag, total);
}

struct SortMove {
    SlotArrays src, dst;
    int full;
};

__device__ __forceinline__ void rank_move_body(int lb, int n, const int *__restrict__ cell, const int *__restrict__ cstart,
            const int *__restrict__ tmp_key, const int *__restrict__ sscan, int total, SortMove mv)
{
    const int i = lb * BLK + threadIdx.x;
    if (i >= n) return;
    const int ci = cell[i];
    const int key = mv.src.O[i];
    int dst;
    if (ci < total) {
        const int s0 = cstart[ci], s1 = cstart[ci + 1];
        int rank = 0;
        for (int u = s0; u < s1; u++) rank += (tmp_key[u] > key);
        dst = s0 + rank;
    } else {
        dst = cstart[total] + sscan[i];   // outside the grid: previous slot order
    }
    mv.dst.X[dst] = mv.src.X[i];
    mv.dst.V[dst] = mv.src.V[i];
    mv.dst.T[dst] = mv.src.T[i];
    mv.dst.O[dst] = key;
    mv.dst.NC[dst] = mv.src.NC[i];
    if (mv.full) {
        mv.dst.A[dst] = mv.src.A[i];
        mv.dst.P[dst] = mv.src.P[i];
        mv.dst.ND[dst] = mv.src.ND[i];
        mv.dst.S[dst] = mv.src.S[i];
    }
}

__global__ void __launch_bounds__(BLK)
k_rank_move(int n, const int *__restrict__ cell, const int *__restrict__ cstart,
            const int *__restrict__ tmp_key, const int *__restrict__ sscan, int total, SortMove mv)
{
    rank_move_body(blockIdx.x, n, cell, cstart, tmp_key, sscan, total, mv);
}

int launch_sort(Ctx *c, const StepConsts &k, int full)
{
    const int total = k.grid.total;
    prof_begin(PK_SORT, c->stream);
    int rc = launch_scan_excl(c->ccount, c->cstart, total + 1, c->scan_blk, c->stream);
    if (rc) return rc;
    rc = launch_fill_int(c->cfill, 0, (size_t)total + 1, c->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scatter, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->cell,
                       c->O[c->ct], c->cstart, c->cfill, c->tmp_key, c->sflag, total);
    LAUNCH_CHECK();
    rc = launch_scan_excl(c->sflag, c->sscan, c->n, c->scan_blk, c->stream);
    if (rc) return rc;
    SortMove mv;
    mv.src = cur_arrays(c);
    mv.dst.X = c->X[1 - c->cx]; mv.dst.V = c->V[1 - c->cx];
    mv.dst.A = c->ACC[1 - c->ct]; mv.dst.ND = c->ND[1 - c->ct];
    mv.dst.T = c->T[1 - c->ct]; mv.dst.O = c->O[1 - c->ct];
    mv.dst.S = c->SURF[1 - c->ct]; mv.dst.NC = c->NCOUNT[1 - c->ct];
    mv.dst.P = c->P[1 - c->cp];
    mv.full = full;
    hipLaunchKernelGGL(k_rank_move, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->cell,
                       c->cstart, c->tmp_key, c->sscan, total, mv);
    prof_end(PK_SORT, c->stream);
    LAUNCH_CHECK();
    c->cx ^= 1;
    c->ct ^= 1;
    if (full) c->cp ^= 1;
    return 0;
}

// ---------------------------------------------------------------------
// 4. neighbour search (neighbor_search.c:185-239) fused with number
//    density (operators.c:17-38), free-surface test (:206-219 / :226-267),
//    unknown flag, PPE diagonal and density-deviation source
//    (pressure_solver.c:411-422, :495-529, :328-340)
// ---------------------------------------------------------------------

struct SearchOut {
    int *nbr, *ncnt, *surf, *unk, *ncount;
    double *nd, *diag, *b;
    DevCounters *ctr;
};

// Running state of one particle's search.
struct SearchAcc {
    int cnt, cnt_n;
    double nsum, sumw;
};

// Candidate j of particle i (neighbor_search.c:211-233): listed when
// r^2 < re^2 (strict), with the number density, the diagonal's weight sum
// and the count-based surface test folded in as the pair is listed.
__device__ __forceinline__ void search_pair(int i, int j, const dvec &xi, const dvec &xj, int tj, SearchAcc &a,
                                            const SearchOut &out, const StepConsts &k, int with_density)
{
    if (j == i) return;
    double dr[D_];
    const double r2 = dist2(xi, xj, dr);
    if (!(r2 < k.re2)) return;
    if (a.cnt < k.stride) out.nbr[sell_off(i, a.cnt, k.stride)] = j;
    a.cnt++;
    if (with_density) {
        const bool need_n = r2 <= k.re_n2_hi;
        const bool need_l = tj != DUMMY_PARTICLE;
        if (need_n || need_l) {
            const double r = sqrt(r2);
            if (need_n) a.nsum += wfun(r, k.re_n, k.r_min);
            if (need_l) a.sumw += wfun(r, k.re_lap, k.r_min);
        }
        a.cnt_n += (r2 < k.re_n2);
    }
}

// The per-particle results of the search and the block's counters.
template <int B>
__device__ __forceinline__ void search_finish(int i, int n, int t, const int *__restrict__ O, const SearchAcc &a,
                                              const SearchOut &out, const StepConsts &k, int with_density)
{
    int u = 0;
    if (i < n) {
        const int cnt = a.cnt;
        out.ncnt[i] = cnt < k.stride ? cnt : k.stride;
        if (cnt > k.stride) atomicMin(&out.ctr->overflow_orig, O[i]);
        if (with_density) {
            int s;
            if (k.surf_method == 1) {
                out.ncount[i] = a.cnt_n;
                s = (t == FLUID_PARTICLE) ? ((double)a.cnt_n < k.beta_n0c) : 0;
            } else {
                s = (t == FLUID_PARTICLE) ? (a.nsum < k.thr_n0) : 0;
            }
            u = (t == FLUID_PARTICLE && !s) || t == WALL_PARTICLE;
            out.nd[i] = a.nsum;
            out.surf[i] = s;
            out.unk[i] = u;
            out.diag[i] = u ? k.diag_coeff * a.sumw : 0.0;
            double bi = 0.0;
            if (u && t == FLUID_PARTICLE && k.hs_mode == 0 && k.ppe_type == 0)
                bi = k.src_k * (a.nsum - k.n0) / k.n0;
            out.b[i] = bi;
        }
    }
    // counters: unknowns, directed pairs, largest list
    const long long pairs = block_sum<long long, B>((long long)(i < n ? a.cnt : 0));
    const long long nu = block_sum<long long, B>((long long)u);
    if (threadIdx.x == 0) {
        atomicAdd(&out.ctr->n_pairs, (unsigned long long)pairs);
        atomicAdd(&out.ctr->n_eq, (int)nu);
    }
    if (i < n && a.cnt > 0) atomicMax(&out.ctr->max_count, a.cnt);
}

// One thread per slot; candidate positions gathered from HBM/L2 (option
// "search" 0).
template <int B>
__device__ __forceinline__ void search_hbm_body(int lb, int n, const dvec *__restrict__ X, const int *__restrict__ T, const int *__restrict__ O,
         const int *__restrict__ cstart, SearchOut out, StepConsts k, int with_density)
{
    const int i = lb * B + threadIdx.x;
    SearchAcc a{0, 0, 0.0, 0.0};
    int t = GHOST_PARTICLE;
    if (i < n) {
        t = T[i];
        const dvec xi = X[i];
        int cc[3];
        if (t != GHOST_PARTICLE && cell_coords(k.grid, xi, cc) >= 0) {
            int lo[3], hi[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                lo[d] = cc[d] > 0 ? cc[d] - 1 : 0;
                hi[d] = cc[d] < k.grid.nc[d] - 1 ? cc[d] + 1 : k.grid.nc[d] - 1;
            }
            for (int cz = lo[2]; cz <= hi[2]; cz++)
                for (int cy = lo[1]; cy <= hi[1]; cy++)
                    for (int cxx = lo[0]; cxx <= hi[0]; cxx++) {
                        const int cid = cell_key(k.grid, cxx, cy, cz);
                        const int j1 = cstart[cid + 1];
                        // the cell's candidates NB_U at a time: positions
                        // loaded together, tested and listed in order
                        for (int j0 = cstart[cid]; j0 < j1; j0 += NB_U) {
                            dvec xx[NB_U];
                            int tt[NB_U];
#pragma unroll
                            for (int u = 0; u < NB_U; u++)
                                if (j0 + u < j1) {
                                    xx[u] = X[j0 + u];
                                    tt[u] = with_density ? T[j0 + u] : 0;
                                }
#pragma unroll
                            for (int u = 0; u < NB_U; u++) {
                                if (j0 + u >= j1) break;
                                search_pair(i, j0 + u, xi, xx[u], tt[u], a, out, k, with_density);
                            }
                        }
                    }
        }
    }
    search_finish<B>(i, n, t, O, a, out, k, with_density);
}

template <int B>
__global__ void __launch_bounds__(B)
k_search(int n, const dvec *__restrict__ X, const int *__restrict__ T, const int *__restrict__ O,
         const int *__restrict__ cstart, SearchOut out, StepConsts k, int with_density)
{
    search_hbm_body<B>(blockIdx.x, n, X, T, O, cstart, out, k, with_density);
}

// The same search with the candidates staged in LDS (option "search" 1,
// the default).  Slots are in cell order with x the fastest cell axis, so
// the three cells x-1..x+1 of one (y, z) row of a particle's block are one
// contiguous slot range, and the particle's 27 cells (9 in 2-D) are 9 (3)
// such ranges, visited z -> y in the reference's order.  The workgroup
// walks the 9 row offsets in that order; for each it stages the union of
// its particles' ranges (consecutive slots, so a few cells wider than the
// block) in chunks of SRCH_CAP positions and types, and every particle
// scans its own range from LDS in ascending slot order, SRCH_U candidates
// per batch of reads.  The lists and every sum are those of k_search, bit
// for bit.  A union wider than SRCH_SPAN (a block that straddles a y layer
// of cells) is read from HBM directly for that row offset.
#define SRCH_CAP 512
#define SRCH_SPAN (4 * SRCH_CAP)
#define SRCH_FB 1   // candidates per batch on the HBM path of a wide row offset

__device__ __forceinline__ int wave_min_int(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, WAVE));
    return v;
}
__device__ __forceinline__ int wave_max_int(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, WAVE));
    return v;
}

template <int B, int SRCH_U>
__device__ __forceinline__ void search_lds_body(int lb, int n, const dvec *__restrict__ X, const int *__restrict__ T, const int *__restrict__ O,
             const int *__restrict__ cstart, SearchOut out, StepConsts k, int with_density)
{
    __shared__ double sxc[D_][SRCH_CAP + SRCH_U];   // staged positions, one array per axis
    __shared__ int stp[SRCH_CAP + SRCH_U];
    __shared__ int smin[B / WAVE], smax[B / WAVE];
    const int i = lb * B + threadIdx.x;
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    SearchAcc a{0, 0, 0.0, 0.0};
    int t = GHOST_PARTICLE;
    dvec xi{};
    bool act = false;
    int cc[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (i < n) {
        t = T[i];
        xi = X[i];
        act = t != GHOST_PARTICLE && cell_coords(k.grid, xi, cc) >= 0;
        if (act) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                lo[d] = cc[d] > 0 ? cc[d] - 1 : 0;
                hi[d] = cc[d] < k.grid.nc[d] - 1 ? cc[d] + 1 : k.grid.nc[d] - 1;
            }
        }
    }
    const int npz = D_ == 3 ? 3 : 1;
    for (int pz = 0; pz < npz; pz++) {
        for (int py = 0; py < 3; py++) {
            // this particle's slot range for row offset (pz, py)
            int ra = 0, rb = 0;
            if (act) {
                const int cz = D_ == 3 ? cc[2] + pz - 1 : 0;
                const int cy = cc[1] + py - 1;
                if (cz >= lo[2] && cz <= hi[2] && cy >= lo[1] && cy <= hi[1]) {
                    ra = cstart[cell_key(k.grid, lo[0], cy, cz)];
                    rb = cstart[cell_key(k.grid, hi[0], cy, cz) + 1];
                }
            }
            // the block's union of ranges
            const int wmin = wave_min_int(ra < rb ? ra : INT_MAX);
            const int wmax = wave_max_int(ra < rb ? rb : INT_MIN);
            __syncthreads();   // the previous row offset's readers are done
            if (lane == 0) {
                smin[w] = wmin;
                smax[w] = wmax;
            }
            __syncthreads();
            int bmin = INT_MAX, bmax = INT_MIN;
#pragma unroll
            for (int q = 0; q < B / WAVE; q++) {
                bmin = min(bmin, smin[q]);
                bmax = max(bmax, smax[q]);
            }
            if (bmin >= bmax) continue;   // uniform over the block
            if (bmax - bmin > SRCH_SPAN) {
                for (int j0 = ra; j0 < rb; j0 += SRCH_FB) {
                    dvec xx[SRCH_FB];
                    int tt[SRCH_FB];
#pragma unroll
                    for (int u = 0; u < SRCH_FB; u++)
                        if (j0 + u < rb) {
                            xx[u] = X[j0 + u];
                            tt[u] = with_density ? T[j0 + u] : 0;
                        }
#pragma unroll
                    for (int u = 0; u < SRCH_FB; u++) {
                        if (j0 + u >= rb) break;
                        search_pair(i, j0 + u, xi, xx[u], tt[u], a, out, k, with_density);
                    }
                }
                continue;
            }
            for (int s0 = bmin; s0 < bmax; s0 += SRCH_CAP) {
                const int s1 = min(s0 + SRCH_CAP, bmax);
                __syncthreads();
                for (int q = threadIdx.x; q < s1 - s0; q += B) {
                    const dvec xq = X[s0 + q];
#pragma unroll
                    for (int d = 0; d < D_; d++) sxc[d][q] = xq.c[d];
                    if (with_density) stp[q] = T[s0 + q];
                }
                __syncthreads();
                // reads of a batch may run up to SRCH_U - 1 past the
                // staged entries (into the pad); those candidates are not
                // tested
                const int j1 = min(rb, s1);
                for (int j0 = max(ra, s0); j0 < j1; j0 += SRCH_U) {
                    dvec xx[SRCH_U];
                    int tt[SRCH_U];
#pragma unroll
                    for (int u = 0; u < SRCH_U; u++) {
#pragma unroll
                        for (int d = 0; d < D_; d++) xx[u].c[d] = sxc[d][j0 + u - s0];
                        tt[u] = with_density ? stp[j0 + u - s0] : 0;
                    }
#pragma unroll
                    for (int u = 0; u < SRCH_U; u++) {
                        if (j0 + u >= j1) break;
                        search_pair(i, j0 + u, xi, xx[u], tt[u], a, out, k, with_density);
                    }
                }
            }
        }
    }
    search_finish<B>(i, n, t, O, a, out, k, with_density);
}

template <int B>
__global__ void __launch_bounds__(B)
k_search_lds(int n, const dvec *__restrict__ X, const int *__restrict__ T, const int *__restrict__ O,
             const int *__restrict__ cstart, SearchOut out, StepConsts k, int with_density)
{
    search_lds_body<B, 4>(blockIdx.x, n, X, T, O, cstart, out, k, with_density);
}

// The same with two candidates per batch of LDS reads (option "search" 2):
// fewer registers, so more waves per SIMD.
template <int B>
__global__ void __launch_bounds__(B)
k_search_lds2(int n, const dvec *__restrict__ X, const int *__restrict__ T, const int *__restrict__ O,
              const int *__restrict__ cstart, SearchOut out, StepConsts k, int with_density)
{
    search_lds_body<B, 2>(blockIdx.x, n, X, T, O, cstart, out, k, with_density);
}

int launch_search(Ctx *c, const StepConsts &k, int with_density, DevCounters *ctr)
{
    SearchOut o;
    o.nbr = c->nbr; o.ncnt = c->ncnt; o.surf = c->SURF[c->ct]; o.unk = c->unk;
    o.ncount = c->NCOUNT[c->ct]; o.nd = c->ND[c->ct]; o.diag = c->diag; o.b = c->b; o.ctr = ctr;
    prof_begin(PK_SEARCH, c->stream);
    if (search_form() == 1)
        PARTICLE_LAUNCH(k_search_lds, c, c->n, c->X[c->cx], c->T[c->ct], c->O[c->ct], c->cstart, o, k,
                        with_density);
    else if (search_form() == 2)
        PARTICLE_LAUNCH(k_search_lds2, c, c->n, c->X[c->cx], c->T[c->ct], c->O[c->ct], c->cstart, o, k,
                        with_density);
    else
        PARTICLE_LAUNCH(k_search, c, c->n, c->X[c->cx], c->T[c->ct], c->O[c->ct], c->cstart, o, k,
                        with_density);
    prof_end(PK_SEARCH, c->stream);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------
// 5. PPE off-diagonals (pressure_solver.c:490-531) and the sources that
//    need velocities (HS :342-362, weakly compressible :381-409)
// ---------------------------------------------------------------------

template <int B>
__device__ __forceinline__ void assemble_body(int lb, int n, const dvec *__restrict__ X, const dvec *__restrict__ V, const int *__restrict__ T,
           const double *__restrict__ ND, const int *__restrict__ unk, const int *__restrict__ nbr,
           const int *__restrict__ ncnt, double *__restrict__ Aval, int *__restrict__ Acol,
           int *__restrict__ Alen, double *__restrict__ b, DevCounters *ctr, StepConsts k)
{
    const int i = lb * B + threadIdx.x;
    long long row_nnz = 0;
    if (i < n) {
        if (!unk[i]) {
            Alen[i] = 0;
        } else {
            const dvec xi = X[i];
            const int t = T[i];
            const bool need_v = (t == FLUID_PARTICLE) && (k.hs_mode == 1 || k.ppe_type == 1);
            dvec vi;
            if (need_v) vi = V[i];
            double hs = 0.0, div = 0.0;
            int m = 0;
            const int cnt = ncnt[i];
            if (!need_v) {
                // the common case, eight neighbours at a time: their
                // indices, then their flags and positions, are loaded
                // together (the gathers overlap); the entries are formed
                // and stored in list order as below
                for (int q0 = 0; q0 < cnt; q0 += ASM_U) {
                    int jj[ASM_U], ok[ASM_U];
                    dvec xx[ASM_U];
#pragma unroll
                    for (int u = 0; u < ASM_U; u++) jj[u] = (q0 + u < cnt) ? nbr[sell_off(i, q0 + u, k.stride)] : i;
#pragma unroll
                    for (int u = 0; u < ASM_U; u++) ok[u] = q0 + u < cnt && unk[jj[u]] && T[jj[u]] != DUMMY_PARTICLE;
#pragma unroll
                    for (int u = 0; u < ASM_U; u++)
                        if (ok[u]) xx[u] = X[jj[u]];
#pragma unroll
                    for (int u = 0; u < ASM_U; u++) {
                        if (!ok[u]) continue;
                        double dr[D_];
                        const double r2 = dist2(xi, xx[u], dr);
                        const double w = wfun(sqrt(r2), k.re_lap, k.r_min);
                        Aval[sell_off(i, m, k.stride)] = k.neg_coeff * w;
                        Acol[sell_off(i, m, k.stride)] = jj[u];
                        m++;
                    }
                }
            }
            for (int q = 0; need_v && q < cnt; q++) {
                const int j = nbr[sell_off(i, q, k.stride)];
                const int uj = unk[j];
                if (!uj && !need_v) continue;
                if (T[j] == DUMMY_PARTICLE) continue;
                double dr[D_];
                const double r2 = dist2(xi, X[j], dr);
                if (uj) {
                    const double w = wfun(sqrt(r2), k.re_lap, k.r_min);
                    Aval[sell_off(i, m, k.stride)] = k.neg_coeff * w;
                    Acol[sell_off(i, m, k.stride)] = j;
                    m++;
                }
                if (need_v && !(r2 < 1.0e-24)) {
                    const dvec vj = V[j];
                    if (k.hs_mode == 1) {
                        const double r = sqrt(r2);
                        double dv = 0.0;
#pragma unroll
                        for (int d = 0; d < D_; d++) dv += dr[d] * (vj.c[d] - vi.c[d]);
                        hs += k.re_lap / (r * r * r) * dv;
                    } else {
                        double dv = 0.0;
#pragma unroll
                        for (int d = 0; d < D_; d++) dv += (vj.c[d] - vi.c[d]) * dr[d];
                        div += dv / r2 * wfun(sqrt(r2), k.re_lap, k.r_min);
                    }
                }
            }
            Alen[i] = m;
            row_nnz = m + 1;
            if (need_v) {
                if (k.hs_mode == 1) {
                    b[i] = k.hs_k * hs;
                } else {
                    const double vel_div = k.dim_n0 * div;
                    b[i] = k.neg_rho_dt * vel_div + k.src_k * (ND[i] - k.n0) / k.n0;
                }
            }
        }
    }
    const long long tot = block_sum<long long, B>(row_nnz);
    if (threadIdx.x == 0) atomicAdd(&ctr->nnz, (unsigned long long)tot);
}

template <int B>
__global__ void __launch_bounds__(B)
k_assemble(int n, const dvec *__restrict__ X, const dvec *__restrict__ V, const int *__restrict__ T,
           const double *__restrict__ ND, const int *__restrict__ unk, const int *__restrict__ nbr,
           const int *__restrict__ ncnt, double *__restrict__ Aval, int *__restrict__ Acol,
           int *__restrict__ Alen, double *__restrict__ b, DevCounters *ctr, StepConsts k)
{
    assemble_body<B>(blockIdx.x, n, X, V, T, ND, unk, nbr, ncnt, Aval, Acol, Alen, b, ctr, k);
}

int launch_assemble(Ctx *c, const StepConsts &k, DevCounters *ctr)
{
    prof_begin(PK_ASSEMBLE, c->stream);
    PARTICLE_LAUNCH(k_assemble, c, c->n, c->X[c->cx],
                       c->V[c->cx], c->T[c->ct], c->ND[c->ct], c->unk, c->nbr, c->ncnt, c->Aval,
                       c->Acol, c->Alen, c->b, ctr, k);
    prof_end(PK_ASSEMBLE, c->stream);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------
// 6. solution -> pressure (pressure_solver.c:563-566), negative clamp
//    (operators.c:188-199)
// ---------------------------------------------------------------------

__device__ __forceinline__ void pressure_out_body(int lb, int n, const double *__restrict__ x, const int *__restrict__ unk, const int *__restrict__ T,
               double *__restrict__ P, int clamp)
{
    const int i = lb * BLK + threadIdx.x;
    if (i >= n) return;
    double p = unk[i] ? x[i] : 0.0;
    if (clamp && T[i] == FLUID_PARTICLE && p < 0.0) p = 0.0;
    P[i] = p;
}

__global__ void __launch_bounds__(BLK)
k_pressure_out(int n, const double *__restrict__ x, const int *__restrict__ unk, const int *__restrict__ T,
               double *__restrict__ P, int clamp)
{
    pressure_out_body(blockIdx.x, n, x, unk, T, P, clamp);
}

int launch_pressure_out(Ctx *c, const StepConsts &k, int n_eq_zero)
{
    (void)n_eq_zero;
    hipLaunchKernelGGL(k_pressure_out, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->x,
                       c->unk, c->T[c->ct], c->P[c->cp], k.clamp);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------
// 7. pressure gradient (operators.c:102-181), corrector (simulation.c:
//    100-108), wall re-fix and out-of-domain ghosting (boundary.c:12-61)
// ---------------------------------------------------------------------

__device__ __forceinline__ void pmin_body(int lb, int n, const double *__restrict__ P, const int *__restrict__ nbr, const int *__restrict__ ncnt,
       double *__restrict__ pmin, int stride)
{
    const int i = lb * BLK + threadIdx.x;
    if (i >= n) return;
    double m = P[i];
    const int cnt = ncnt[i];
    for (int q = 0; q < cnt; q++) {
        const double pj = P[nbr[sell_off(i, q, stride)]];
        if (pj < m) m = pj;
    }
    pmin[i] = m;
}

__global__ void __launch_bounds__(BLK)
k_pmin(int n, const double *__restrict__ P, const int *__restrict__ nbr, const int *__restrict__ ncnt,
       double *__restrict__ pmin, int stride)
{
    pmin_body(blockIdx.x, n, P, nbr, ncnt, pmin, stride);
}

struct GradIO {
    const dvec *X, *V;
    const double *P, *pmin;
    dvec *Xo, *Vo, *A;
    double *Po, *ND;
    int *T;
};

template <int B>
__device__ __forceinline__ void grad_correct_body(int lb, int n, GradIO io, const int *__restrict__ nbr, const int *__restrict__ ncnt, StepConsts k)
{
    const int i = lb * B + threadIdx.x;
    if (i >= n) return;
    int t = io.T[i];
    const dvec xi = io.X[i], vi = io.V[i];
    const double pi = io.P[i];
    dvec xo = xi, vo = vi, ao;
#pragma unroll
    for (int d = 0; d < (D_ == 3 ? 4 : 2); d++) ao.c[d] = 0.0;
    double po = pi;
    if (t == FLUID_PARTICLE) {
        const int cnt = ncnt[i];
        double pimin = 0.0;
        if (k.cmps == 1) {
            pimin = io.pmin[i];
        } else if (k.cmps == 0) {
            pimin = pi;
            for (int q0 = 0; q0 < cnt; q0 += 2 * NB_U) {
                double pp[2 * NB_U];
#pragma unroll
                for (int u = 0; u < 2 * NB_U; u++)
                    pp[u] = (q0 + u < cnt) ? io.P[nbr[sell_off(i, q0 + u, k.stride)]] : pimin;
#pragma unroll
                for (int u = 0; u < 2 * NB_U; u++)
                    if (pp[u] < pimin) pimin = pp[u];
            }
        }
        double grad[D_];
#pragma unroll
        for (int d = 0; d < D_; d++) grad[d] = 0.0;
        // NB_U neighbours at a time, summed in list order
        for (int q0 = 0; q0 < cnt; q0 += NB_U) {
            int jj[NB_U];
            dvec xx[NB_U];
            double pp[NB_U];
#pragma unroll
            for (int u = 0; u < NB_U; u++) jj[u] = (q0 + u < cnt) ? nbr[sell_off(i, q0 + u, k.stride)] : i;
#pragma unroll
            for (int u = 0; u < NB_U; u++) {
                xx[u] = io.X[jj[u]];
                pp[u] = io.P[jj[u]];
            }
#pragma unroll
            for (int u = 0; u < NB_U; u++) {
                if (q0 + u >= cnt) break;
                double dr[D_];
                const double r2 = dist2(xi, xx[u], dr);
                if (r2 < k.grad_min_r2) continue;
                const double w = wfun(sqrt(r2), k.re_lap, k.r_min);
                const double pj = pp[u];
                double dp;
                if (k.cmps == 1) dp = pi + pj - pimin - io.pmin[jj[u]];
                else if (k.cmps == 2) dp = pi + pj;
                else dp = pj - pimin;
#pragma unroll
                for (int d = 0; d < D_; d++) grad[d] += dp / r2 * dr[d] * w;
            }
        }
#pragma unroll
        for (int d = 0; d < D_; d++) {
            ao.c[d] = k.neg_gc * grad[d] / k.rho;
            const double du = k.dt * ao.c[d];
            vo.c[d] = vi.c[d] + du;
            xo.c[d] = xi.c[d] + k.dt * du;
        }
        bool out = false;
#pragma unroll
        for (int d = 0; d < D_; d++)
            if (xo.c[d] < k.bnd_lo[d] || xo.c[d] > k.bnd_hi[d] || isnan(xo.c[d])) out = true;
        if (out) {
            t = GHOST_PARTICLE;
            io.T[i] = GHOST_PARTICLE;
            po = 0.0;
            io.ND[i] = 0.0;
#pragma unroll
            for (int d = 0; d < D_; d++) vo.c[d] = ao.c[d] = 0.0;
        }
    } else if (t == WALL_PARTICLE || t == GHOST_PARTICLE) {
#pragma unroll
        for (int d = 0; d < D_; d++) vo.c[d] = 0.0;
    }
    io.Xo[i] = xo;
    io.Vo[i] = vo;
    io.A[i] = ao;
    io.Po[i] = po;
}

template <int B>
__global__ void __launch_bounds__(B)
k_grad_correct(int n, GradIO io, const int *__restrict__ nbr, const int *__restrict__ ncnt, StepConsts k)
{
    grad_correct_body<B>(blockIdx.x, n, io, nbr, ncnt, k);
}

int launch_grad_correct(Ctx *c, const StepConsts &k)
{
    if (k.cmps == 1) {
        hipLaunchKernelGGL(k_pmin, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->P[c->cp],
                           c->nbr, c->ncnt, c->pmin, k.stride);
        LAUNCH_CHECK();
    }
    GradIO io;
    io.X = c->X[c->cx]; io.V = c->V[c->cx]; io.P = c->P[c->cp]; io.pmin = c->pmin;
    io.Xo = c->X[1 - c->cx]; io.Vo = c->V[1 - c->cx]; io.A = c->ACC[c->ct];
    io.Po = c->P[1 - c->cp]; io.ND = c->ND[c->ct]; io.T = c->T[c->ct];
    prof_begin(PK_GRAD, c->stream);
    PARTICLE_LAUNCH(k_grad_correct, c, c->n, io, c->nbr, c->ncnt, k);
    prof_end(PK_GRAD, c->stream);
    LAUNCH_CHECK();
    c->cx ^= 1;
    c->cp ^= 1;
    return 0;
}

// ---------------------------------------------------------------------
// stand-alone operators over the current list (drop-in entry points)
// ---------------------------------------------------------------------

// calc_particle_number_density (operators.c:17-38)
__global__ void __launch_bounds__(BLK)
k_density(int n, const dvec *__restrict__ X, const int *__restrict__ nbr, const int *__restrict__ ncnt,
          double *__restrict__ ND, StepConsts k)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const dvec xi = X[i];
    double s = 0.0;
    const int cnt = ncnt[i];
    for (int q = 0; q < cnt; q++) {
        double dr[D_];
        s += wfun(sqrt(dist2(xi, X[nbr[sell_off(i, q, k.stride)]], dr)), k.re_n, k.r_min);
    }
    ND[i] = s;
}

int launch_density_only(Ctx *c, const StepConsts &k)
{
    hipLaunchKernelGGL(k_density, dim3(nblk(c->n)), dim3(BLK), 0, c->stream, c->n, c->X[c->cx], c->nbr,
                       c->ncnt, c->ND[c->ct], k);
    LAUNCH_CHECK();
    return 0;
}

// ---------------------------------------------------------------------
// Batches of scenes: every operator of the step as one launch over the
// concatenated block ranges of all scenes.  Each block finds its scene by
// bisection over the launch's block offsets and runs the single-scene
// kernel's body on that scene's buffers, with that scene's local block
// index, so every scene gets the bits of a launch of its own.
// ---------------------------------------------------------------------

int flat_build_grid(FlatGrid *g, const int *blocks, int ns, hipStream_t s)
{
    std::vector<int> h((size_t)ns + 1);
    h[0] = 0;
    for (int q = 0; q < ns; q++) h[q + 1] = h[q] + blocks[q];
    if (g->cap < ns + 1) {
        if (g->first) hipFree(g->first);
        g->first = nullptr;
        HIP_OK(hipMalloc(&g->first, sizeof(int) * h.size()));
        g->cap = ns + 1;
    }
    HIP_OK(hipMemcpyAsync(g->first, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, s));
    HIP_OK(hipStreamSynchronize(s));
    g->blocks = h[ns];
    return 0;
}

// Segmented exclusive scans in three launches: the tiles of every segment
// (segment g on tiles tile0[g] .. tile0[g + 1] - 1), one block per segment
// scanning its tile sums (kept at blk[tile0[g] + g ...], the segment's
// total after them), then the tiles again with their offsets.
__global__ void __launch_bounds__(BLK) k_sscan_tiles(const ScanSeg *__restrict__ segs,
                                                     const int *__restrict__ tile0, int nseg, int *blk)
{
    const int g = flat_scene(tile0, nseg, blockIdx.x);
    const ScanSeg S = seg_global(segs[g]);
    const int lt = blockIdx.x - tile0[g];
    const int base = lt * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++)
        if (base + q < S.n) s += S.in[base + q];
    int tot;
    (void)block_scan_excl<BLK>(s, &tot);
    if (threadIdx.x == 0) blk[tile0[g] + g + lt] = tot;
}

__global__ void __launch_bounds__(1024) k_sscan_tile_sums(const ScanSeg *__restrict__ segs,
                                                          const int *__restrict__ tile0, int *blk)
{
    const int g = blockIdx.x;
    const ScanSeg S = seg_global(segs[g]);
    const int ntiles = tile0[g + 1] - tile0[g];
    int *tile_sum = blk + tile0[g] + g;
    const int base = threadIdx.x * 8;   // 1024 threads x 8 tiles each
    int v[8];
    int s = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        v[q] = (base + q < ntiles) ? tile_sum[base + q] : 0;
        s += v[q];
    }
    int tot;
    int run = block_scan_excl<1024>(s, &tot);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (base + q < ntiles) tile_sum[base + q] = run;
        run += v[q];
    }
    if (threadIdx.x == 0) {
        tile_sum[ntiles] = tot;
        if (S.tot) *S.tot = tot;
        if (S.n == 0) S.out[0] = 0;
    }
}

__global__ void __launch_bounds__(BLK) k_sscan_apply(const ScanSeg *__restrict__ segs,
                                                     const int *__restrict__ tile0, int nseg, const int *blk)
{
    const int g = flat_scene(tile0, nseg, blockIdx.x);
    const ScanSeg S = seg_global(segs[g]);
    const int lt = blockIdx.x - tile0[g];
    const int ntiles = tile0[g + 1] - tile0[g];
    const int *tile_sum = blk + tile0[g] + g;
    const int base = lt * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++) {
        v[q] = (base + q < S.n) ? S.in[base + q] : 0;
        s += v[q];
    }
    int tot;
    int run = block_scan_excl<BLK>(s, &tot) + tile_sum[lt];
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++) {
        if (base + q < S.n) S.out[base + q] = run;
        run += v[q];
    }
    if (lt == 0 && threadIdx.x == 0) S.out[S.n] = tile_sum[ntiles];
}

int scan_segs_set(ScanSegs *ss, const ScanSeg *h, int nseg, hipStream_t s)
{
    std::vector<int> t0((size_t)nseg + 1);
    t0[0] = 0;
    for (int g = 0; g < nseg; g++) {
        const int nt = (h[g].n + SCAN_TILE - 1) / SCAN_TILE;
        if (h[g].n < 0 || nt > 8192) {
            mps_set_error(MPS_ERR_UNSUPPORTED, __FILE__, __LINE__, "scan longer than 16.7M elements");
            return MPS_ERR_UNSUPPORTED;
        }
        t0[g + 1] = t0[g] + nt;
    }
    if (nseg > ss->nseg) {
        if (ss->segs) hipFree(ss->segs);
        if (ss->tile0) hipFree(ss->tile0);
        ss->segs = nullptr;
        ss->tile0 = nullptr;
        HIP_OK(hipMalloc(&ss->segs, sizeof(ScanSeg) * nseg));
        HIP_OK(hipMalloc(&ss->tile0, sizeof(int) * (nseg + 1)));
    }
    ss->nseg = nseg;
    ss->ntiles = t0[nseg];
    // synchronous copies: the host tables may be rebuilt at once
    HIP_OK(hipMemcpy(ss->segs, h, sizeof(ScanSeg) * nseg, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(ss->tile0, t0.data(), sizeof(int) * (nseg + 1), hipMemcpyHostToDevice));
    (void)s;
    return 0;
}

// blk needs ntiles + nseg ints.
int launch_scan_seg(const ScanSegs &ss, int *blk, hipStream_t s)
{
    if (ss.nseg <= 0) return 0;
    if (ss.ntiles > 0)
        hipLaunchKernelGGL(k_sscan_tiles, dim3(ss.ntiles), dim3(BLK), 0, s, ss.segs, ss.tile0, ss.nseg, blk);
    hipLaunchKernelGGL(k_sscan_tile_sums, dim3(ss.nseg), dim3(1024), 0, s, ss.segs, ss.tile0, blk);
    if (ss.ntiles > 0)
        hipLaunchKernelGGL(k_sscan_apply, dim3(ss.ntiles), dim3(BLK), 0, s, ss.segs, ss.tile0, ss.nseg, blk);
    LAUNCH_CHECK();
    return 0;
}

// A block's scene S and its block index lb within the scene.  The kernels
// below pass S's buffers on through gptr(): read from the table, they are
// global pointers to the compiler (flat ones otherwise).
#define FLAT_SCENE()                                        \
    const int s_ = flat_scene(first, ns, blockIdx.x);       \
    const FlatScene &S = tab[s_];                           \
    const int lb = blockIdx.x - first[s_]

template <int B>
__global__ void __launch_bounds__(B)
k_visc_predict_flat(const FlatScene *__restrict__ tab, const int *__restrict__ first, int ns, int cx, int ct)
{
    FLAT_SCENE();
    visc_predict_body<B>(lb, S.n, gptr(S.X[cx]), gptr(S.V[cx]), gptr(S.T[ct]), gptr(S.nbr), gptr(S.ncnt),
                         gptr(S.X[1 - cx]), gptr(S.V[1 - cx]), S.k);
}

__global__ void __launch_bounds__(BLK)
k_cells_zero_flat(const FlatScene *__restrict__ tab, const int *__restrict__ first, int ns)
{
    FLAT_SCENE();
    const int c = lb * BLK + threadIdx.x;
    if (c > S.total) return;
    gptr(S.ccount)[c] = 0;
    gptr(S.cfill)[c] = 0;
}

template <int B>
__global__ void __launch_bounds__(B)
k_collide_bin_flat(const FlatScene *__restrict__ tab, const int *__restrict__ first, int ns, int cx, int ct)
{
    FLAT_SCENE();
    collide_bin_body<B>(lb, S.n, gptr(S.X[cx]), gptr(S.V[cx]), gptr(S.T[ct]), gptr(S.nbr), gptr(S.ncnt),
                        gptr(S.X[1 - cx]), gptr(S.V[1 - cx]), gptr(S.cell),
                        gptr(S.ccount), S.k, S.k.collide);
}

__global__ void __launch_bounds__(BLK)
k_scatter_flat(const FlatScene *__restrict__ tab, const int *__restrict__ first, int ns, int ct)
{
    FLAT_SCENE();
    scatter_body(lb, S.n, gptr(S.cell), gptr(S.O[ct]), gptr(S.cstart), gptr(S.cfill), gptr(S.tmp_key), gptr(S.sflag),
                 S.total);
}

__global__ void __launch_bounds__(BLK)
k_rank_move_flat(const FlatScene *__restrict__ tab, const int *__restrict__ first, int ns, int cx, int ct, int cp)
{
    FLAT_SCENE();
    SortMove mv;
    mv.src.X = gptr(S.X[cx]); mv.src.V = gptr(S.V[cx]); mv.src.A = gptr(S.ACC[ct]); mv.src.P = gptr(S.P[cp]);
    mv.src.ND = gptr(S.ND[ct]);
    mv.src.T = gptr(S.T[ct]); mv.src.O = gptr(S.O[ct]); mv.src.S = gptr(S.SURF[ct]); mv.src.NC = gptr(S.NCOUNT[ct]);
    mv.dst.X = gptr(S.X[1 - cx]); mv.dst.V = gptr(S.V[1 - cx]); mv.dst.A = gptr(S.ACC[1 - ct]);
    mv.dst.ND = gptr(S.ND[1 - ct]);
    mv.dst.T = gptr(S.T[1 - ct]); mv.dst.O = gptr(S.O[1 - ct]); mv.dst.S = gptr(S.SURF[1 - ct]);
    mv.dst.NC = gptr(S.NCOUNT[1 - ct]);
    mv.dst.P = gptr(S.P[1 - cp]);
    mv.full = 0;
    rank_move_body(lb, S.n, gptr(S.cell), gptr(S.cstart), gptr(S.tmp_key), gptr(S.sscan), S.total, mv);
}

__global__ void k_counters_reset_flat(const FlatScene *__restrict__ tab, int ns)
{
    for (int s = threadIdx.x; s < ns; s += blockDim.x) {
        DevCounters z{};
        z.overflow_orig = INT_MAX;
        *tab[s].ctr = z;
    }
}

__global__ void k_counters_gather_flat(const FlatScene *__restrict__ tab, int ns, DevCounters *out)
{
    for (int s = threadIdx.x; s < ns; s += blockDim.x) out[s] = *tab[s].ctr;
}

// The search over all scenes, one kernel per form (each holds only the
// registers its form needs).
#define SEARCH_FLAT_ARGS                                                                                  \
    SearchOut o;                                                                                          \
    o.nbr = gptr(S.nbr); o.ncnt = gptr(S.ncnt); o.surf = gptr(S.SURF[ct]); o.unk = gptr(S.unk);           \
    o.ncount = gptr(S.NCOUNT[ct]); o.nd = gptr(S.ND[ct]); o.diag = gptr(S.diag); o.b = gptr(S.b);         \
    o.ctr = gptr(S.ctr)

template <int B>
__global__ void __launch_bounds__(B)
k_search_flat(const FlatScene *__restrict__ tab, const int *__restrict__ first, int ns, int cx, int ct)
{
    FLAT_SCENE();
    SEARCH_FLAT_ARGS;
    search_lds_body<B, 4>(lb, S.n, gptr(S.X[cx]), gptr(S.T[ct]), gptr(S.O[ct]), gptr(S.cstart), o, S.k, 1);
}

template <int B>
__global__ void __launch_bounds__(B)
k_search_flat2(const FlatScene *__restrict__ tab, const int *__restrict__ first, int ns, int cx, int ct)
{
    FLAT_SCENE();
    SEARCH_FLAT_ARGS;
    search_lds_body<B, 2>(lb, S.n, gptr(S.X[cx]), gptr(S.T[ct]), gptr(S.O[ct]), gptr(S.cstart), o, S.k, 1);
}

template <int B>
__global__ void __launch_bounds__(B)
k_search_flat_hbm(const FlatScene *__restrict__ tab, const int *__restrict__ first, int ns, int cx, int ct)
{
    FLAT_SCENE();
    SEARCH_FLAT_ARGS;
    search_hbm_body<B>(lb, S.n, gptr(S.X[cx]), gptr(S.T[ct]), gptr(S.O[ct]), gptr(S.cstart), o, S.k, 1);
}

template <int B>
__global__ void __launch_bounds__(B)
k_assemble_flat(const FlatScene *__restrict__ tab, const int *__restrict__ first, int ns, int cx, int ct)
{
    FLAT_SCENE();
    assemble_body<B>(lb, S.n, gptr(S.X[cx]), gptr(S.V[cx]), gptr(S.T[ct]), gptr(S.ND[ct]), gptr(S.unk), gptr(S.nbr),
                     gptr(S.ncnt), gptr(S.Aval), gptr(S.Acol), gptr(S.Alen), 
                     gptr(S.b), gptr(S.ctr), S.k);
}

__global__ void __launch_bounds__(BLK)
k_pressure_out_flat(const FlatScene *__restrict__ tab, const int *__restrict__ first, int ns, int ct, int cp)
{
    FLAT_SCENE();
    pressure_out_body(lb, S.n, gptr(S.x), gptr(S.unk), gptr(S.T[ct]), gptr(S.P[cp]), S.k.clamp);
}

__global__ void __launch_bounds__(BLK)
k_pmin_flat(const FlatScene *__restrict__ tab, const int *__restrict__ first, int ns, int cp)
{
    FLAT_SCENE();
    pmin_body(lb, S.n, gptr(S.P[cp]), gptr(S.nbr), gptr(S.ncnt), gptr(S.pmin), S.k.stride);
}

template <int B>
__global__ void __launch_bounds__(B)
k_grad_correct_flat(const FlatScene *__restrict__ tab, const int *__restrict__ first, int ns, int cx, int ct,
                    int cp)
{
    FLAT_SCENE();
    GradIO io;
    io.X = gptr(S.X[cx]); io.V = gptr(S.V[cx]); io.P = gptr(S.P[cp]); io.pmin = gptr(S.pmin);
    io.Xo = gptr(S.X[1 - cx]); io.Vo = gptr(S.V[1 - cx]); io.A = gptr(S.ACC[ct]);
    io.Po = gptr(S.P[1 - cp]); io.ND = gptr(S.ND[ct]); io.T = gptr(S.T[ct]);
    grad_correct_body<B>(lb, S.n, io, gptr(S.nbr), gptr(S.ncnt), S.k);
}

// Per-particle kernel at the block size of option "pblk" over the scenes.
#define FLAT_PARTICLE_LAUNCH(kern, f, ...)                                                            \
    do {                                                                                              \
        const int pb_ = particle_block();                                                             \
        if ((f).gp.blocks > 0) {                                                                      \
            if (pb_ == 128)                                                                           \
                hipLaunchKernelGGL(kern<128>, dim3((f).gp.blocks), dim3(128), 0, (f).s, __VA_ARGS__); \
            else if (pb_ == 512)                                                                      \
                hipLaunchKernelGGL(kern<512>, dim3((f).gp.blocks), dim3(512), 0, (f).s, __VA_ARGS__); \
            else                                                                                      \
                hipLaunchKernelGGL(kern<256>, dim3((f).gp.blocks), dim3(256), 0, (f).s, __VA_ARGS__); \
        }                                                                                             \
    } while (0)

// The step of every scene up to the assembled PPE: predictor, collision
// with the cell histogram, counting sort, search with density and flags,
// assembly (launch_visc_predict .. launch_assemble of one scene).
int flat_step_pre(FlatLaunch &f)
{
    prof_begin(PK_VISC, f.s);
    FLAT_PARTICLE_LAUNCH(k_visc_predict_flat, f, f.tab, f.gp.first, f.ns, f.cx, f.ct);
    prof_end(PK_VISC, f.s);
    LAUNCH_CHECK();
    f.cx ^= 1;
    if (f.gc.blocks > 0)
        hipLaunchKernelGGL(k_cells_zero_flat, dim3(f.gc.blocks), dim3(BLK), 0, f.s, f.tab, f.gc.first, f.ns);
    prof_begin(PK_COLLIDE, f.s);
    FLAT_PARTICLE_LAUNCH(k_collide_bin_flat, f, f.tab, f.gp.first, f.ns, f.cx, f.ct);
    prof_end(PK_COLLIDE, f.s);
    LAUNCH_CHECK();
    f.cx ^= 1;
    prof_begin(PK_SORT, f.s);
    int rc = launch_scan_seg(f.sc_cells, f.scan_blk, f.s);
    if (rc) return rc;
    if (f.gb.blocks > 0)
        hipLaunchKernelGGL(k_scatter_flat, dim3(f.gb.blocks), dim3(BLK), 0, f.s, f.tab, f.gb.first, f.ns, f.ct);
    LAUNCH_CHECK();
    rc = launch_scan_seg(f.sc_sflag, f.scan_blk, f.s);
    if (rc) return rc;
    if (f.gb.blocks > 0)
        hipLaunchKernelGGL(k_rank_move_flat, dim3(f.gb.blocks), dim3(BLK), 0, f.s, f.tab, f.gb.first, f.ns, f.cx,
                           f.ct, f.cp);
    prof_end(PK_SORT, f.s);
    LAUNCH_CHECK();
    f.cx ^= 1;
    f.ct ^= 1;
    hipLaunchKernelGGL(k_counters_reset_flat, dim3(1), dim3(64), 0, f.s, f.tab, f.ns);
    prof_begin(PK_SEARCH, f.s);
    if (search_form() == 1) FLAT_PARTICLE_LAUNCH(k_search_flat, f, f.tab, f.gp.first, f.ns, f.cx, f.ct);
    else if (search_form() == 2) FLAT_PARTICLE_LAUNCH(k_search_flat2, f, f.tab, f.gp.first, f.ns, f.cx, f.ct);
    else FLAT_PARTICLE_LAUNCH(k_search_flat_hbm, f, f.tab, f.gp.first, f.ns, f.cx, f.ct);
    prof_end(PK_SEARCH, f.s);
    LAUNCH_CHECK();
    prof_begin(PK_ASSEMBLE, f.s);
    FLAT_PARTICLE_LAUNCH(k_assemble_flat, f, f.tab, f.gp.first, f.ns, f.cx, f.ct);
    prof_end(PK_ASSEMBLE, f.s);
    LAUNCH_CHECK();
    return 0;
}

// The step of every scene after the solve (x already in slot order):
// pressure and clamp, p_min where a scene needs it, gradient, corrector
// and boundary (launch_pressure_out, launch_grad_correct of one scene).
int flat_step_post(FlatLaunch &f)
{
    if (f.gb.blocks > 0)
        hipLaunchKernelGGL(k_pressure_out_flat, dim3(f.gb.blocks), dim3(BLK), 0, f.s, f.tab, f.gb.first, f.ns,
                           f.ct, f.cp);
    LAUNCH_CHECK();
    if (f.any_cmps1 && f.gb.blocks > 0) {
        hipLaunchKernelGGL(k_pmin_flat, dim3(f.gb.blocks), dim3(BLK), 0, f.s, f.tab, f.gb.first, f.ns, f.cp);
        LAUNCH_CHECK();
    }
    prof_begin(PK_GRAD, f.s);
    FLAT_PARTICLE_LAUNCH(k_grad_correct_flat, f, f.tab, f.gp.first, f.ns, f.cx, f.ct, f.cp);
    prof_end(PK_GRAD, f.s);
    LAUNCH_CHECK();
    f.cx ^= 1;
    f.cp ^= 1;
    return 0;
}

int flat_counters(FlatLaunch &f, DevCounters *out_dev)
{
    hipLaunchKernelGGL(k_counters_gather_flat, dim3(1), dim3(64), 0, f.s, f.tab, f.ns, out_dev);
    LAUNCH_CHECK();
    return 0;
}
