// dev_common.hpp -- shared device helpers for the gfx950 MPS kernels.
//
// Everything on the FP64 path is compiled with -ffp-contract=off and written
// in the reference's association order, so that neighbour sets, number
// density, surface flags and the PPE pattern come out bit-identical to the
// reference's x86 build (gcc -O2 emits no FMA there).  FP64 division and
// sqrt on gfx950 are the correctly rounded IEEE sequences LLVM emits by
// default (no fast-math anywhere in this library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mps_types.h"

#define D_ MPS_DIM

// Wavefront width of CDNA; kernels are written for 64.
#define WAVE 64

// Per-particle vector storage: 32-byte (3-D) or 16-byte (2-D) aligned
// records so a neighbour gather is one or two dwordx4 loads.
struct alignas(D_ == 3 ? 32 : 16) dvec {
    double c[D_ == 3 ? 4 : 2];
};

// Sliced-ELL layout shared by the neighbour list and the PPE matrix: slices
// of 64 consecutive slots, `stride` entries per slot; entry k of slot i sits
// at (i/64)*stride*64 + k*64 + i%64, so one wavefront reading entry k of 64
// consecutive slots touches one contiguous 256/512-byte span.
__host__ __device__ __forceinline__ size_t sell_off(int i, int k, int stride)
{
    return ((size_t)(i >> 6) * (size_t)stride + (size_t)k) * 64u + (size_t)(i & 63);
}

// kernel.c:12-20: w = re/r - 1 with r clamped below at r_min = 0.01*l0,
// and 0 for r >= re.  NaN r falls through to re/NaN - 1 = NaN as in C.
__device__ __forceinline__ double wfun(double r, double re, double r_min)
{
    if (r >= re) return 0.0;
    if (r < r_min) r = r_min;
    return re / r - 1.0;
}

// C's (int) cast as x86-64 performs it (cvttsd2si): truncation toward zero
// in range, INT_MIN for NaN or out-of-range values.  The GPU's own
// conversion saturates and maps NaN to 0, so it is not used directly.
__device__ __forceinline__ int to_int_x86(double v)
{
    if (!(v > -2147483649.0 && v < 2147483648.0)) return (int)0x80000000;
    return (int)v;
}

// Squared distance accumulated from 0 in axis order (neighbor_search.c:214-219).
__device__ __forceinline__ double dist2(const dvec &pi, const dvec &pj, double dr[D_])
{
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < D_; d++) {
        dr[d] = pj.c[d] - pi.c[d];
        r2 += dr[d] * dr[d];
    }
    return r2;
}

// Grid geometry of cell_list_create (neighbor_search.c:92-122).
struct Grid {
    double origin[3];
    double inv;         // 1.0 / cell_size
    int nc[3];
    int total;          // nx*ny*nz (the sentinel bucket is index `total`)
};

// Storage order of the cells (the counting sort's key): y, the vertical
// axis of the dam-break scenes, slowest, then z, then x.  Only where cells
// sit in memory follows it; every search still visits the neighbour cells
// z -> y -> x as neighbor_search.c:185-239 does, so the lists and every sum
// over them are unchanged.  With y slowest the slots of three consecutive
// cell layers span few enough positions that the columns of a 64-row slice
// of the PPE matrix mostly fit 16-bit offsets.  (2-D: y * nx + x.)
__host__ __device__ __forceinline__ int cell_key(const Grid &g, int cx, int cy, int cz)
{
    return (cy * g.nc[2] + cz) * g.nc[0] + cx;
}

// Cell of a position or -1 outside (neighbor_search.c:171-178).
__device__ __forceinline__ int cell_coords(const Grid &g, const dvec &p, int c[3])
{
    c[0] = c[1] = c[2] = 0;
#pragma unroll
    for (int d = 0; d < D_; d++) {
        c[d] = to_int_x86((p.c[d] - g.origin[d]) * g.inv);
        if (c[d] < 0 || c[d] >= g.nc[d]) return -1;
    }
    return cell_key(g, c[0], c[1], c[2]);
}

// A pointer that a kernel reads out of a table in device memory (the
// per-scene tables of the batched launches) is generic to the compiler, so
// every access through it becomes a flat instruction.  Flat accesses count
// against the LDS wait counter as well as the memory one, so each wait for
// an LDS read then also waits for every flat load or store in flight.  The
// tables only hold device buffers: gptr() reads the field as a global
// pointer, and the accesses through the result compile to global ones.
template <typename T>
__device__ __forceinline__ T *gptr(T *const &field)
{
    typedef __attribute__((address_space(1))) T GT;
    GT *g = *(GT *const *)&field;
    return (T *)g;
}

// ---------------------------------------------------------------------
// XCD-aware block order.  The dispatcher hands workgroups to the 8 XCDs
// round-robin (block b runs on XCD b % 8) and each XCD has its own L2.
// xcd_block() renumbers the blocks so that XCD k runs one contiguous range
// of logical blocks: rows that are neighbours in the cell-sorted order, and
// the vector entries they gather, then share one L2.  A bijection on
// [0, nb) for any nb.
// ---------------------------------------------------------------------
#define N_XCD 8

__device__ __forceinline__ int xcd_block(int b, int nb)
{
    const int q = nb / N_XCD, rem = nb % N_XCD;
    const int x = b % N_XCD, k = b / N_XCD;
    return x * q + (x < rem ? x : rem) + k;
}

// The scene of block b of a launch over several scenes' block ranges:
// the s with first[s] <= b < first[s + 1] (first[0] = 0, first[ns] = the
// launch's blocks, b below it).  Scenes without blocks are passed over.
__device__ __forceinline__ int flat_scene(const int *__restrict__ first, int ns, int b)
{
    int lo = 0, hi = ns;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------
// wavefront / block primitives (wave64)
// ---------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, WAVE);
    return v;  // valid in lane 0
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, WAVE);
    return v;
}

__device__ __forceinline__ long long wave_sum(long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, WAVE);
    return v;
}

// Deterministic block sum (fixed tree); result valid in thread 0.
template <typename T, int BLOCK>
__device__ __forceinline__ T block_sum(T v)
{
    __shared__ T part[BLOCK / WAVE];
    v = wave_sum(v);
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    if (lane == 0) part[w] = v;
    __syncthreads();
    T s = T(0);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < BLOCK / WAVE; k++) s += part[k];
    }
    __syncthreads();
    return s;
}

// Inclusive wave scan of ints.
__device__ __forceinline__ int wave_scan_incl(int v)
{
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        int t = __shfl_up(v, o, WAVE);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive block scan of ints; *total receives the block sum.
template <int BLOCK>
__device__ __forceinline__ int block_scan_excl(int v, int *total)
{
    __shared__ int wsum[BLOCK / WAVE + 1];
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    const int incl = wave_scan_incl(v);
    if (lane == WAVE - 1) wsum[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int k = 0; k < BLOCK / WAVE; k++) {
            int t = wsum[k];
            wsum[k] = run;
            run += t;
        }
        wsum[BLOCK / WAVE] = run;
    }
    __syncthreads();
    const int r = incl - v + wsum[w];
    *total = wsum[BLOCK / WAVE];
    __syncthreads();
    return r;
}
