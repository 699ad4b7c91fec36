// Device kernels of the preconditioned BiCGSTAB solve engine (gfx950).
//
// Every operation on the solve path is HBM-bandwidth bound (sparse
// matrix-vector products and level-1 vector work); nothing is a dense
// contraction, so there is no MFMA here.  Design points:
//  * CSR SpMV, CSR-Adaptive style: the host cuts the rows into blocks of
//    consecutive rows holding at most TILE nonzeros (or one long row).  A
//    256-thread workgroup streams its block's (value, column) pairs with fully
//    coalesced nontemporal loads, forms value*x[column] into LDS, then
//    reduces each row with a power-of-two group of lanes (1..64) chosen from
//    the block's row count, finishing with cross-lane xor shuffles inside the
//    wave.  A single row longer than TILE is reduced by the whole workgroup.
//    Every nonzero and every column index is read exactly once.  Workgroups
//    are remapped so each XCD walks one contiguous range of rows (x reuse in
//    its L2).
//  * Value / vector / accumulator types are template parameters: the FP64
//    path (double everywhere) and the mixed-precision inner solve (FP32 or
//    FP16-with-row-scale values, FP32 vectors, FP32 row sums) share the code.
//    Dot products and norms always accumulate in double.
//  * Epilogues are fused into the SpMV row write: D^{-1} scaling (factor
//    apply), b - A x (residuals), beta = 1 accumulation (Richardson sweeps),
//    and the dot products BiCGSTAB needs right after A-products.
//  * Reductions are deterministic: each workgroup writes its partial sums and
//    a one-workgroup finalize kernel adds them in a fixed order and updates
//    the solve state.  No float atomics, no cross-workgroup fences: the
//    kernel boundary publishes the partials (an agent-scope release fence per
//    workgroup writes back the XCD's L2 and costs more than the SpMV).
//  * The Krylov scalars and the exit decision live in device memory
//    (SolveState), so the host can enqueue many iterations without reading
//    anything back; every kernel returns at once when the state says done.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cm_layout.hpp"

namespace cm {

constexpr int MAX_RED_BLOCKS = 2048;  // workgroups of the reducing vector kernels

// Solve state kept in device memory.  variant: 0 AFIIM, 1 PDFIIM, 2 ACIFIM.
struct SolveState {
    double ro, rc, prc, m1, m2, alpha, omega, beta;
    double tol, eps;
    double res_outer;  // ||b - A x|| of the last refinement residual (mixed path)
    int iter, nmax, stag, msteps, maxmsteps, info, done, variant;
    int ntrace, trace_cap;
    double *trace;  // device array of trace_cap doubles or null
};

// ---------------------------------------------------------------- reductions

template <class T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Sum of K values over a workgroup of NT threads; result valid on thread 0.
template <int K, int NT = BS>
__device__ __forceinline__ void block_sum(double (&v)[K], double *lds /* K*NT/WAVE */)
{
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) lds[w * K + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; k++) {
            double s = 0.0;
            for (int q = 0; q < NT / WAVE; q++) s += lds[q * K + k];
            v[k] = s;
        }
}

// Workgroup index -> row block, XCD-aware.  Consecutive workgroups are
// dispatched round-robin over the 8 XCDs; remapping gives each XCD one
// contiguous range of row blocks (walked in order), so the x entries that
// neighbouring rows share stay in that XCD's L2.  A bijection of [0, nblk).
__device__ __forceinline__ int xcd_block(int i, int nblk)
{
    constexpr int NX = 8;
    const int q = nblk / NX, rem = nblk % NX;
    const int x = i % NX, k = i / NX;
    return x * q + min(x, rem) + k;
}

// Writes this workgroup's K sums to partials[slot*K + k] (plain stores; the
// finalize kernel that reads them runs after this kernel ends, so the kernel
// boundary makes them visible on every XCD).
template <int K>
__device__ __forceinline__ void block_partials(double (&v)[K], double *partials, double *lds, int slot)
{
    block_sum<K>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; k++) partials[(size_t)slot * K + k] = v[k];
}

// One workgroup of FBS threads: adds the nblk partials in a fixed order
// (deterministic) and hands the totals to fin.fin(state, totals).  The loads
// of a thread are independent, so the workgroup has all of them in flight.
constexpr int FBS = 1024;

template <int K, class Fin>
__global__ __launch_bounds__(FBS) void k_finalize(int nblk, const double *__restrict__ partials,
                                                  SolveState *st, Fin fin)
{
    if (st->done) return;
    __shared__ double red[K * FBS / WAVE];
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
    int b = threadIdx.x;
    for (; b + 3 * FBS < nblk; b += 4 * FBS) {
        double a0[K], a1[K], a2[K], a3[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            a0[k] = partials[(size_t)b * K + k];
            a1[k] = partials[(size_t)(b + FBS) * K + k];
            a2[k] = partials[(size_t)(b + 2 * FBS) * K + k];
            a3[k] = partials[(size_t)(b + 3 * FBS) * K + k];
        }
#pragma unroll
        for (int k = 0; k < K; k++) s[k] += ((a0[k] + a1[k]) + (a2[k] + a3[k]));
    }
    for (; b < nblk; b += FBS)
#pragma unroll
        for (int k = 0; k < K; k++) s[k] += partials[(size_t)b * K + k];
    block_sum<K, FBS>(s, red);
    if (threadIdx.x == 0) fin.fin(st, s);
}

// ------------------------------------------------------------ SpMV epilogues
//
// An epilogue sees each row's sum once (row, value) and may accumulate K
// per-row products into the workgroup's reduction; fin() runs in the
// finalize kernel with the grid totals.  XV counts the n-vectors an epilogue
// touches besides x and the y write (for the algorithmic byte count).

template <class Y>
struct EpiStore {  // y = A x
    static constexpr int K = 0;
    static constexpr int XV = 0;
    Y *y;
    template <class S>
    __device__ void row(int r, S s, double *) const { y[r] = (Y)s; }
    __device__ void fin(SolveState *, const double *) const {}
};

template <class Y, class D>
struct EpiScale {  // y = d .* (A x)   (factor apply: D^{-1} after H)
    static constexpr int K = 0;
    static constexpr int XV = 1;
    Y *y;
    const D *d;
    template <class S>
    __device__ void row(int r, S s, double *) const { y[r] = (Y)d[r] * (Y)s; }
    __device__ void fin(SolveState *, const double *) const {}
};

template <class Y>
struct EpiSubFrom {  // y = b - A x
    static constexpr int K = 0;
    static constexpr int XV = 1;
    Y *y;
    const Y *b;
    template <class S>
    __device__ void row(int r, S s, double *) const { y[r] = b[r] - (Y)s; }
    __device__ void fin(SolveState *, const double *) const {}
};

template <class Y>
struct EpiAccum {  // y = A x + y
    static constexpr int K = 0;
    static constexpr int XV = 1;
    Y *y;
    template <class S>
    __device__ void row(int r, S s, double *) const { y[r] = (Y)s + y[r]; }
    __device__ void fin(SolveState *, const double *) const {}
};

template <class Y, class D>
struct EpiAccumScale {  // z = (A x + y) .* d   (last Richardson sweep then D^{-1})
    static constexpr int K = 0;
    static constexpr int XV = 2;
    const Y *y;
    Y *z;
    const D *d;
    template <class S>
    __device__ void row(int r, S s, double *) const
    {
        const Y t = (Y)s + y[r];
        z[r] = t * (Y)d[r];
    }
    __device__ void fin(SolveState *, const double *) const {}
};

// Setup: r = b - A x; r0 = p = r; ro = ||r||, m1 = ro^2
// (AFIIM/src/bicgstab.cpp:149-165).
struct FinSetup {
    __device__ void fin(SolveState *st, const double *tot) const
    {
        st->ro = sqrt(tot[0]);
        st->m1 = st->ro * st->ro;
        st->rc = st->ro;
        st->iter = 0;
        st->stag = 0;
        st->msteps = 0;
        st->ntrace = 0;
        if (st->nmax <= 0) {
            st->done = 1;
            if (st->nmax == 0) st->info = 3;
        }
    }
};

template <class T>
struct EpiSetup {
    static constexpr int K = 1;
    static constexpr int XV = 3;
    T *r, *r0, *p;
    const T *b;
    __device__ void fin(SolveState *st, const double *tot) const { FinSetup{}.fin(st, tot); }
    template <class S>
    __device__ void row(int i, S s, double *acc) const
    {
        const T v = b[i] + (T)(-1.0) * (T)s;
        r[i] = v;
        r0[i] = v;
        p[i] = v;
        acc[0] += (double)v * (double)v;
    }
};

// v = A pc, alpha = m1 / (r0 . v)   (AFIIM/src/bicgstab.cpp:177-181;
// ACIFIM/src/bicgstab.cpp:205-213 adds the non-finite guard).
template <class T>
struct EpiAlpha {
    static constexpr int K = 1;
    static constexpr int XV = 1;
    T *v;
    const T *r0;
    template <class S>
    __device__ void row(int i, S s, double *acc) const
    {
        const T vi = (T)s;
        v[i] = vi;
        acc[0] += (double)r0[i] * (double)vi;
    }
    __device__ void fin(SolveState *st, const double *tot) const
    {
        const double a = st->m1 / tot[0];
        st->alpha = a;
        if (st->variant == 2 && !isfinite(a)) {
            st->iter = st->nmax;
            st->info = 3;
            st->done = 1;
        }
    }
};

// t = A sc, omega = (t . s) / (t . t)   (AFIIM/src/bicgstab.cpp:193-198).
template <class T>
struct EpiOmega {
    static constexpr int K = 2;
    static constexpr int XV = 1;
    T *t;
    const T *s;
    template <class S>
    __device__ void row(int i, S v, double *acc) const
    {
        const T ti = (T)v;
        t[i] = ti;
        acc[0] += (double)ti * (double)s[i];
        acc[1] += (double)ti * (double)ti;
    }
    __device__ void fin(SolveState *st, const double *tot) const { st->omega = tot[0] / tot[1]; }
};

// Refinement residual (FP64): r = b - A x in double, its FP32 copy for the
// inner solve, and ||r||^2.
struct EpiResid {
    static constexpr int K = 1;
    static constexpr int XV = 2;  // b read, FP32 copy written (half width)
    const double *b;
    float *r32;
    template <class S>
    __device__ void row(int i, S s, double *acc) const
    {
        const double v = b[i] - (double)s;
        r32[i] = (float)v;
        acc[0] += v * v;
    }
    __device__ void fin(SolveState *st, const double *tot) const { st->res_outer = sqrt(tot[0]); }
};

// ------------------------------------------------------------------- SpMV

template <class V>
__device__ __forceinline__ float ldv_nt(const V *p, int k) { return (float)__builtin_nontemporal_load(&p[k]); }
template <>
__device__ __forceinline__ float ldv_nt<__half>(const __half *p, int k)
{
    const unsigned short u = __builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p) + k);
    return __half2float(__ushort_as_half(u));
}
template <class V>
__device__ __forceinline__ float ldv(const V *p, int k) { return (float)p[k]; }
template <>
__device__ __forceinline__ float ldv<__half>(const __half *p, int k) { return __half2float(p[k]); }

// Value loads, plain and nontemporal.  Two functions rather than one with a
// flag: a select between a nontemporal and a plain load of the same address
// is folded into one load, and the nontemporal hint is lost.
template <class V, class Acc>
__device__ __forceinline__ Acc load_val(const V *p, int k)
{
    return (Acc)ldv(p, k);
}
template <>
__device__ __forceinline__ double load_val<double, double>(const double *p, int k)
{
    return p[k];
}
template <class V, class Acc>
__device__ __forceinline__ Acc load_val_nt(const V *p, int k)
{
    return (Acc)ldv_nt(p, k);
}
template <>
__device__ __forceinline__ double load_val_nt<double, double>(const double *p, int k)
{
    return __builtin_nontemporal_load(&p[k]);
}

// Matrix arrays are allocated with this many spare elements after nnz, so
// the 16-byte loads of spmv_vl may run past the last row block.
constexpr int CSR_PAD = 8;

typedef int v4i_t __attribute__((ext_vector_type(4)));

// 16 bytes of V (or int) from p, 16-byte aligned, nontemporal.
template <class V>
__device__ __forceinline__ v4i_t ld16_nt(const V *p)
{
    return __builtin_nontemporal_load(reinterpret_cast<const v4i_t *>(p));
}

// Row-block table of a matrix cut for one SpMV launch of nb = gridDim.x
// workgroups: blk[0..nb] are the blocks' first rows, blk[nb+1..2nb+1] the
// offsets of their first nonzeros (ptr at those rows).  Reading both halves
// at once saves the workgroup a dependent load of the row pointers before
// it can stream its nonzeros.
struct RowBlock {
    int r0, r1, rs, re;
};

__device__ __forceinline__ RowBlock row_block(const int *__restrict__ blk, int bi)
{
    const int nb1 = (int)gridDim.x + 1;
    return {blk[bi], blk[bi + 1], blk[nb1 + bi], blk[nb1 + bi + 1]};
}

// Rows of a staged block: a power-of-two group of tpr lanes per row (at most
// one wave), chosen from the row count.  The bounds of each thread's first
// row are loaded when the groups are formed, before the nonzeros are
// streamed, so the row reduction does not wait on them.
struct RowGroups {
    int tpr, groups, g, lane, a0, e0;
    __device__ __forceinline__ RowGroups(const int *__restrict__ ptr, const RowBlock &b)
    {
        const int nr = b.r1 - b.r0;
        tpr = 1;
        while (tpr < WAVE && tpr * 2 * nr <= BS) tpr *= 2;
        groups = BS / tpr;
        g = threadIdx.x / tpr;
        lane = threadIdx.x & (tpr - 1);
        a0 = e0 = 0;
        if (g < nr) {
            a0 = ptr[b.r0 + g] - b.rs;
            e0 = ptr[b.r0 + g + 1] - b.rs;
        }
    }
};

// Sums each row's staged products (prod[k] for the block's k-th nonzero)
// with its lane group and xor-shuffles, then hands the row sum (times the
// row scale, if any) to the epilogue on the group's first lane.
template <class Acc, class Epi>
__device__ __forceinline__ void reduce_rows(const RowGroups &G, const RowBlock &b, const Acc *prod,
                                            const int *__restrict__ ptr, const float *__restrict__ rowscale,
                                            const Epi &epi, double *acc)
{
    const int nr = b.r1 - b.r0;
    for (int base = 0; base < nr; base += G.groups) {
        const int rr = base + G.g;
        const bool live = rr < nr;
        Acc s = 0;
        if (live) {
            int a = G.a0, e = G.e0;
            if (base > 0) {
                a = ptr[b.r0 + rr] - b.rs;
                e = ptr[b.r0 + rr + 1] - b.rs;
            }
            for (int k = a + G.lane; k < e; k += G.tpr) s += prod[k];
        }
        for (int off = G.tpr / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, WAVE);
        if (live && G.lane == 0) {
            if (rowscale) s *= (Acc)rowscale[b.r0 + rr];
            epi.row(b.r0 + rr, s, acc);
        }
    }
}

// A block holding one row longer than the tile: the whole workgroup strides
// over it.  colf(k) gives the k-th nonzero's column.
template <class V, class X, class Acc, class Epi, class Col>
__device__ __forceinline__ void long_row(const RowBlock &b, const V *__restrict__ val, Col colf,
                                         const float *__restrict__ rowscale, const X *__restrict__ x,
                                         const Epi &epi, double *acc, double *red)
{
    double s = 0.0;
    for (int k = b.rs + (int)threadIdx.x; k < b.re; k += BS) s += (double)(load_val<V, Acc>(val, k) * (Acc)x[colf(k)]);
    double v1[1] = {s};
    block_sum<1>(v1, red);
    if (threadIdx.x == 0) {
        Acc rsum = (Acc)v1[0];
        if (rowscale) rsum *= (Acc)rowscale[b.r0];
        epi.row(b.r0, rsum, acc);
    }
}

// y = A x (+ epilogue).  A: CSR ptr/col/val (rowscale: optional per-row
// factor applied to the row sum, for FP16 values stored scaled); blk: the
// row-block table (row starts, then nonzero starts; 2 (gridDim.x + 1)
// entries), cut for TL staged nonzeros per block.  Products and row sums in
// Acc.  The solve-state test comes after the block header loads, so both are
// in flight together.
template <int TL, class V, class X, class Acc, class Epi>
__global__ __launch_bounds__(BS) void spmv_kernel(const int *__restrict__ ptr,
                                                  const int *__restrict__ col,
                                                  const V *__restrict__ val,
                                                  const float *__restrict__ rowscale,
                                                  const int *__restrict__ blk,
                                                  const X *__restrict__ x, Epi epi,
                                                  SolveState *st, double *partials)
{
    __shared__ Acc prod[TL];
    __shared__ double red[(Epi::K > 0 ? Epi::K : 1) * BS / WAVE];
    double acc[Epi::K > 0 ? Epi::K : 1];
#pragma unroll
    for (int k = 0; k < (Epi::K > 0 ? Epi::K : 1); k++) acc[k] = 0.0;

    const int tid = threadIdx.x;
    const int bi = xcd_block(blockIdx.x, gridDim.x);
    const RowBlock b = row_block(blk, bi);
    if (st && st->done) return;
    const int cnt = b.re - b.rs;

    if (cnt > TL) {
        long_row<V, X, Acc>(b, val, [&](int k) { return col[k]; }, rowscale, x, epi, acc, red);
    } else {
        const RowGroups G(ptr, b);
        int cidx[TL / BS];
        Acc cv[TL / BS];
#pragma unroll
        for (int u = 0; u < TL / BS; u++) {
            const int k = tid + u * BS;
            if (k < cnt) {  // streamed once: nontemporal, keeps x in the caches
                cidx[u] = __builtin_nontemporal_load(&col[b.rs + k]);
                cv[u] = load_val_nt<V, Acc>(val, b.rs + k);
            }
        }
#pragma unroll
        for (int u = 0; u < TL / BS; u++) {
            const int k = tid + u * BS;
            if (k < cnt) prod[k] = cv[u] * (Acc)x[cidx[u]];
        }
        __syncthreads();
        reduce_rows(G, b, prod, ptr, rowscale, epi, acc);
    }
    if constexpr (Epi::K > 0) block_partials<Epi::K>(acc, partials, red, bi);
}

// y = A x (+ epilogue), same contract and row blocks as spmv_kernel, with the
// block's column indices and values streamed by 16-byte vector loads into
// LDS first (from the 8-element-aligned start of the block; the arrays carry
// CSR_PAD spare elements) and read back one nonzero per lane for the x
// gather.  Against spmv_kernel this issues 4x fewer load instructions for the
// indices and 2x / 4x / 8x fewer for FP64 / FP32 / FP16 values, which is what
// bounds the narrow-value products; the gathers stay one nonzero per lane, so
// neighbouring lanes still read neighbouring x entries.
template <int TL, class V, class X, class Acc, class Epi>
__global__ __launch_bounds__(BS) void spmv_vl(const int *__restrict__ ptr, const int *__restrict__ col,
                                              const V *__restrict__ val, const float *__restrict__ rowscale,
                                              const int *__restrict__ blk, const X *__restrict__ x, Epi epi,
                                              SolveState *st, double *partials)
{
    constexpr int VB = (TL + CSR_PAD) * (int)sizeof(V) > TL * (int)sizeof(Acc) ? (TL + CSR_PAD) * (int)sizeof(V)
                                                                                : TL * (int)sizeof(Acc);
    __shared__ int scol[TL + CSR_PAD];
    __shared__ __attribute__((aligned(16))) unsigned char sbuf[VB];  // values, then products
    __shared__ double red[(Epi::K > 0 ? Epi::K : 1) * BS / WAVE];
    V *sval = reinterpret_cast<V *>(sbuf);
    Acc *prod = reinterpret_cast<Acc *>(sbuf);
    double acc[Epi::K > 0 ? Epi::K : 1];
#pragma unroll
    for (int k = 0; k < (Epi::K > 0 ? Epi::K : 1); k++) acc[k] = 0.0;

    const int tid = threadIdx.x;
    const int bi = xcd_block(blockIdx.x, gridDim.x);
    const RowBlock b = row_block(blk, bi);
    if (st && st->done) return;
    const int cnt = b.re - b.rs;

    if (cnt > TL) {
        long_row<V, X, Acc>(b, val, [&](int k) { return col[k]; }, rowscale, x, epi, acc, red);
    } else {
        const RowGroups G(ptr, b);
        const int base = b.rs & ~(CSR_PAD - 1), shift = b.rs - base;
        const int tot = shift + cnt;  // <= TL + CSR_PAD - 1
        constexpr int CW = 16 / (int)sizeof(int), VW = 16 / (int)sizeof(V);
        for (int q = tid; q < (tot + CW - 1) / CW; q += BS)
            *reinterpret_cast<v4i_t *>(&scol[q * CW]) = ld16_nt(col + base + q * CW);
        for (int q = tid; q < (tot + VW - 1) / VW; q += BS)
            *reinterpret_cast<v4i_t *>(&sbuf[q * 16]) = ld16_nt(val + base + q * VW);
        __syncthreads();
        Acc pr[TL / BS];
#pragma unroll
        for (int u = 0; u < TL / BS; u++) {
            const int k = tid + u * BS;
            if (k < cnt) {
                const V vv = sval[k + shift];
                pr[u] = load_val<V, Acc>(&vv, 0) * (Acc)x[scol[k + shift]];
            }
        }
        __syncthreads();  // values are dead: the products take their place
#pragma unroll
        for (int u = 0; u < TL / BS; u++) {
            const int k = tid + u * BS;
            if (k < cnt) prod[k] = pr[u];
        }
        __syncthreads();
        reduce_rows(G, b, prod, ptr, rowscale, epi, acc);
    }
    if constexpr (Epi::K > 0) block_partials<Epi::K>(acc, partials, red, bi);
}

// spmv_vl with 16-bit column indices: row block b's columns are stored as
// d16[k] = col[k] - cbase[b] (the block's smallest column), which fits
// whenever a block's columns span less than 65536 -- always for stencil and
// factor matrices in natural order.  The index stream halves again (2 bytes
// per nonzero); the gathers are unchanged.  Used by the FP32 / FP16 inner
// solves, whose products are bound per nonzero more than per byte, where it
// takes another 5-11% off.
template <int TL, class V, class X, class Acc, class Epi>
__global__ __launch_bounds__(BS) void spmv_vl16(const int *__restrict__ ptr, const uint16_t *__restrict__ d16,
                                                const int *__restrict__ cbase, const V *__restrict__ val,
                                                const float *__restrict__ rowscale, const int *__restrict__ blk,
                                                const X *__restrict__ x, Epi epi, SolveState *st, double *partials)
{
    constexpr int VB = (TL + CSR_PAD) * (int)sizeof(V) > TL * (int)sizeof(Acc) ? (TL + CSR_PAD) * (int)sizeof(V)
                                                                                : TL * (int)sizeof(Acc);
    __shared__ __attribute__((aligned(16))) uint16_t sd[TL + 2 * CSR_PAD];
    __shared__ __attribute__((aligned(16))) unsigned char sbuf[VB];  // values, then products
    __shared__ double red[(Epi::K > 0 ? Epi::K : 1) * BS / WAVE];
    V *sval = reinterpret_cast<V *>(sbuf);
    Acc *prod = reinterpret_cast<Acc *>(sbuf);
    double acc[Epi::K > 0 ? Epi::K : 1];
#pragma unroll
    for (int k = 0; k < (Epi::K > 0 ? Epi::K : 1); k++) acc[k] = 0.0;

    const int tid = threadIdx.x;
    const int bi = xcd_block(blockIdx.x, gridDim.x);
    const RowBlock b = row_block(blk, bi);
    const int cb = cbase[bi];
    if (st && st->done) return;
    const int cnt = b.re - b.rs;

    if (cnt > TL) {
        long_row<V, X, Acc>(b, val, [&](int k) { return cb + (int)d16[k]; }, rowscale, x, epi, acc, red);
    } else {
        const RowGroups G(ptr, b);
        const int base = b.rs & ~(CSR_PAD - 1), shift = b.rs - base;
        const int tot = shift + cnt;  // <= TL + CSR_PAD - 1
        constexpr int DW = 16 / (int)sizeof(uint16_t), VW = 16 / (int)sizeof(V);
        for (int q = tid; q < (tot + DW - 1) / DW; q += BS)
            *reinterpret_cast<v4i_t *>(&sd[q * DW]) = ld16_nt(d16 + base + q * DW);
        for (int q = tid; q < (tot + VW - 1) / VW; q += BS)
            *reinterpret_cast<v4i_t *>(&sbuf[q * 16]) = ld16_nt(val + base + q * VW);
        __syncthreads();
        Acc pr[TL / BS];
#pragma unroll
        for (int u = 0; u < TL / BS; u++) {
            const int k = tid + u * BS;
            if (k < cnt) {
                const V vv = sval[k + shift];
                pr[u] = load_val<V, Acc>(&vv, 0) * (Acc)x[cb + sd[k + shift]];
            }
        }
        __syncthreads();  // values are dead: the products take their place
#pragma unroll
        for (int u = 0; u < TL / BS; u++) {
            const int k = tid + u * BS;
            if (k < cnt) prod[k] = pr[u];
        }
        __syncthreads();
        reduce_rows(G, b, prod, ptr, rowscale, epi, acc);
    }
    if constexpr (Epi::K > 0) block_partials<Epi::K>(acc, partials, red, bi);
}

// Slot-major row blocks, one row per thread.  Rows are cut into blocks of
// SLOT_ROWS consecutive rows; inside a block the rows are ordered by length
// (longest first) and the nonzeros are stored slot by slot: first the first
// nonzero of every row, then the second of every row that has one, and so on.
// Thread i of the workgroup owns the block's i-th row in that order; its j-th
// nonzero sits at data start + soff[j] + i.  Consecutive lanes read
// consecutive column indices and values, and -- rows of a stencil-like matrix
// sharing their offset pattern -- gather x at consecutive addresses, which is
// what bounds the staged kernels on short, regular rows.  Each row is summed
// in its own nonzero order by one thread (no LDS staging, no shuffles), so
// the FP64 row sums equal a sequential row loop's bit for bit.
// Taken for matrices whose rows are at most SLOT_MAXLEN long.

// hdr: per block {first row, rows, data start, slot-table start, slots};
// lrow: per row position {row length << 16 | block-relative row};
// column indices as col (I16 = false) or as 16-bit offsets d16 from the
// block's smallest column cbase (I16 = true).
template <class V, class X, class Acc, class Epi, bool I16>
__global__ __launch_bounds__(BS) void spmv_slot(const int *__restrict__ hdr, const int *__restrict__ soff,
                                                const int *__restrict__ lrow, const int *__restrict__ col,
                                                const uint16_t *__restrict__ d16, const int *__restrict__ cbase,
                                                const V *__restrict__ val, const float *__restrict__ rowscale,
                                                const X *__restrict__ x, Epi epi, SolveState *st, double *partials)
{
    __shared__ double red[(Epi::K > 0 ? Epi::K : 1) * BS / WAVE];
    double acc[Epi::K > 0 ? Epi::K : 1];
#pragma unroll
    for (int k = 0; k < (Epi::K > 0 ? Epi::K : 1); k++) acc[k] = 0.0;

    const int bi = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = hdr[5 * bi], nr = hdr[5 * bi + 1], ds = hdr[5 * bi + 2], ss = hdr[5 * bi + 3];
    const int cb = I16 ? cbase[bi] : 0;
    if (st && st->done) return;
    const int i = threadIdx.x;
    if (i < nr) {
        // each product rounded before it is added, as in a plain row loop
        // (no fused multiply-add): the row sum is the sequential CSR sum
#pragma clang fp contract(off)
        const int lr = lrow[r0 + i];
        const int len = lr >> 16, row = r0 + (lr & 0xffff);
        Acc s = 0;
        for (int j = 0; j < len; j += 4) {
            int k[4];
            bool on[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                on[q] = j + q < len;
                k[q] = on[q] ? ds + soff[ss + j + q] + i : 0;
            }
            Acc v[4];
            int c[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (on[q]) {
                    v[q] = load_val_nt<V, Acc>(val, k[q]);
                    c[q] = I16 ? cb + (int)__builtin_nontemporal_load(&d16[k[q]]) : __builtin_nontemporal_load(&col[k[q]]);
                }
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (on[q]) s += v[q] * (Acc)x[c[q]];
        }
        if (rowscale) s *= (Acc)rowscale[row];
        epi.row(row, s, acc);
    }
    if constexpr (Epi::K > 0) block_partials<Epi::K>(acc, partials, red, bi);
}

// Diagonals plus a CSR remainder.  Matrices from structured grids keep most
// of their nonzeros on a few diagonals (offsets col - row shared by almost
// every row); those diagonals are stored as dense vectors dval[d * ldn + i] =
// A(i, i + off[d]) (zero where the entry is absent or the column falls
// outside the matrix; ldn = n rounded up to DIA_ALIGN, so every diagonal
// starts 16-byte aligned), and the entries on other diagonals as a short CSR
// remainder (rptr / rcol / rval; rptr null when empty).  The diagonal part
// reads no column indices, and consecutive lanes read consecutive values and
// consecutive x entries.  Each thread takes RPT consecutive rows, so one
// diagonal's values for them are one vector load (RPT * sizeof(V) bytes);
// diagonals are taken CHUNK at a time with all their loads in flight.  Each
// product is rounded before it is added (no fused multiply-add), in diagonal
// order, then the remainder in column order.

template <class V, int RPT>
struct DiaVec {  // RPT values of V loaded as one aligned vector
    typedef V type __attribute__((ext_vector_type(RPT)));
};
template <class V>
struct DiaVec<V, 1> {
    typedef V type;
};
template <int RPT>
struct DiaVec<__half, RPT> {
    typedef unsigned short type __attribute__((ext_vector_type(RPT)));
};
template <>
struct DiaVec<__half, 1> {
    typedef unsigned short type;
};

template <class V, class Acc>
__device__ __forceinline__ Acc dia_val(V v) { return (Acc)v; }
template <>
__device__ __forceinline__ float dia_val<__half, float>(__half v) { return __half2float(v); }

// The remainder entries of a workgroup's rows (consecutive rows) are one
// contiguous range [rs, re) of the remainder arrays.  With REM_LDS the
// workgroup walks it in chunks of DIA_REM_LDS entries: every thread loads a
// share of a chunk (columns and values, coalesced; the first chunk before the
// diagonals), gathers x for it and writes the products to LDS, and each row's
// thread adds the products of its own entries in the chunk, in column order,
// so a long row's entries are gathered by the whole workgroup instead of one
// after another by its thread.  Without REM_LDS each thread walks its rows'
// entries after the diagonals.  The host takes REM_LDS for a remainder of at
// least DIA_REM_LDS_MIN entries per row: on a sparser one most waves have no
// remainder row, and the workgroup-wide gather and barrier cost more than the
// few rows' own walks.
constexpr int DIA_REM_LDS = 1024;
constexpr double DIA_REM_LDS_MIN = 0.1;

// Adds rows i0 .. i0 + RPT - 1's diagonal products to s, diagonal by diagonal.
template <class V, class X, class Acc, int CHUNK, int RPT>
__device__ __forceinline__ void dia_rows(int n, int ldn, int nd, const int *__restrict__ off,
                                         const V *__restrict__ dval, const X *__restrict__ x, int i0, Acc (&s)[RPT])
{
#pragma clang fp contract(off)
    using VT = typename DiaVec<V, RPT>::type;
    for (int d0 = 0; d0 < nd; d0 += CHUNK) {
        VT v[CHUNK];
        Acc xv[CHUNK][RPT];
#pragma unroll
        for (int q = 0; q < CHUNK; q++) {
            const int d = d0 + q;
#pragma unroll
            for (int r = 0; r < RPT; r++) xv[q][r] = 0;
            if (d < nd) {
                const int c = i0 + off[d];
                v[q] = __builtin_nontemporal_load(reinterpret_cast<const VT *>(dval + (size_t)d * ldn + i0));
#pragma unroll
                for (int r = 0; r < RPT; r++)
                    if ((unsigned)(c + r) < (unsigned)n) xv[q][r] = (Acc)x[c + r];
            }
        }
#pragma unroll
        for (int q = 0; q < CHUNK; q++)
            if (d0 + q < nd) {
#pragma unroll
                for (int r = 0; r < RPT; r++) {
                    V e;
                    if constexpr (RPT == 1) {
                        __builtin_memcpy(&e, &v[q], sizeof(V));
                    } else {
                        const auto lane = v[q][r];
                        __builtin_memcpy(&e, &lane, sizeof(V));
                    }
                    s[r] += dia_val<V, Acc>(e) * xv[q][r];
                }
            }
    }
}

template <class V, class X, class Acc, class Epi, int CHUNK, int RPT, bool REM_LDS>
__global__ __launch_bounds__(BS) void spmv_dia(int n, int ldn, int nd, const int *__restrict__ off,
                                               const V *__restrict__ dval, const int *__restrict__ rptr,
                                               const int *__restrict__ rcol, const V *__restrict__ rval,
                                               const float *__restrict__ rowscale, const X *__restrict__ x, Epi epi,
                                               SolveState *st, double *partials)
{
    __shared__ double red[(Epi::K > 0 ? Epi::K : 1) * BS / WAVE];
    double acc[Epi::K > 0 ? Epi::K : 1];
#pragma unroll
    for (int k = 0; k < (Epi::K > 0 ? Epi::K : 1); k++) acc[k] = 0.0;

    const int bi = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = (bi * BS + (int)threadIdx.x) * RPT;
    if (st && st->done) return;
    if constexpr (!REM_LDS) {
        if (i0 < n) {
#pragma clang fp contract(off)
            Acc s[RPT];
#pragma unroll
            for (int r = 0; r < RPT; r++) s[r] = 0;
            // the rows' remainder bounds, loaded ahead of the diagonals so their
            // latency hides behind the diagonal loads
            int rb[RPT + 1];
#pragma unroll
            for (int r = 0; r <= RPT; r++) rb[r] = rptr ? rptr[min(i0 + r, n)] : 0;
            dia_rows<V, X, Acc, CHUNK, RPT>(n, ldn, nd, off, dval, x, i0, s);
#pragma unroll
            for (int r = 0; r < RPT; r++) {
                const int i = i0 + r;
                if (i >= n) break;
                Acc t = s[r];
                for (int k = rb[r]; k < rb[r + 1]; k++) t += load_val<V, Acc>(rval, k) * (Acc)x[rcol[k]];
                if (rowscale) t *= (Acc)rowscale[i];
                epi.row(i, t, acc);
            }
        }
    } else {
#pragma clang fp contract(off)
        __shared__ Acc prod[DIA_REM_LDS];
        constexpr int PER = DIA_REM_LDS / BS;
        // the workgroup's remainder range
        const int r0 = bi * BS * RPT;
        int rs = 0, re = 0;
        if (rptr) {
            rs = rptr[r0];
            re = rptr[min(r0 + BS * RPT, n)];
        }
        // this thread's share of the chunk [c0, c0 + DIA_REM_LDS)
        int pc[PER];
        Acc pv[PER];
        auto load_share = [&](int c0) {
            const int c1 = min(re, c0 + DIA_REM_LDS);
#pragma unroll
            for (int j = 0; j < PER; j++) {
                const int k = c0 + j * BS + (int)threadIdx.x;
                pc[j] = -1;
                pv[j] = 0;
                if (k < c1) {
                    pc[j] = rcol[k];
                    pv[j] = load_val<V, Acc>(rval, k);
                }
            }
        };
        load_share(rs);
        int rb[RPT + 1];
#pragma unroll
        for (int r = 0; r <= RPT; r++) rb[r] = rptr ? rptr[min(i0 + r, n)] : 0;
        Acc s[RPT];
#pragma unroll
        for (int r = 0; r < RPT; r++) s[r] = 0;
        if (i0 < n) dia_rows<V, X, Acc, CHUNK, RPT>(n, ldn, nd, off, dval, x, i0, s);
        for (int c0 = rs; c0 < re; c0 += DIA_REM_LDS) {  // uniform over the workgroup
            const int c1 = min(re, c0 + DIA_REM_LDS);
            if (c0 != rs) load_share(c0);
#pragma unroll
            for (int j = 0; j < PER; j++)
                if (pc[j] >= 0) prod[j * BS + (int)threadIdx.x] = pv[j] * (Acc)x[pc[j]];
            __syncthreads();
#pragma unroll
            for (int r = 0; r < RPT; r++)
                for (int k = max(rb[r], c0); k < min(rb[r + 1], c1); k++) s[r] += prod[k - c0];
            if (c1 < re) __syncthreads();  // the next chunk overwrites prod
        }
        if (i0 < n) {
#pragma unroll
            for (int r = 0; r < RPT; r++) {
                const int i = i0 + r;
                if (i >= n) break;
                Acc t = s[r];
                if (rowscale) t *= (Acc)rowscale[i];
                epi.row(i, t, acc);
            }
        }
    }
    if constexpr (Epi::K > 0) block_partials<Epi::K>(acc, partials, red, bi);
}

// ----------------------------------------------------------- vector kernels

// s = r - alpha v   (AFIIM/src/bicgstab.cpp:183-185)
template <class T>
__global__ __launch_bounds__(BS) void k_s_update(int n, const T *__restrict__ r, const T *__restrict__ v,
                                                 T *__restrict__ s, const SolveState *st)
{
    if (st->done) return;
    const T na = (T)(-st->alpha);
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) s[i] = r[i] + na * v[i];
}

// One pass over x, pc, sc, r, s, t, r0:
//   ||pc||^2, ||x_old||^2, ||r_old||^2;  x += alpha pc;  x += omega sc;
//   r = s - omega t;  r0 . r_new   (AFIIM/src/bicgstab.cpp:200-264).
// The partial sums go to FinPass.
template <class T>
__global__ __launch_bounds__(BS) void k_xr_update(int n, T *__restrict__ x, const T *__restrict__ pc,
                                                  const T *__restrict__ sc, T *__restrict__ r,
                                                  const T *__restrict__ s, const T *__restrict__ t,
                                                  const T *__restrict__ r0, SolveState *st,
                                                  double *partials)
{
    if (st->done) return;
    __shared__ double red[4 * BS / WAVE];
    const T a = (T)st->alpha, om = (T)st->omega, nom = (T)(-st->omega);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const T pci = pc[i], xi = x[i], ri = r[i];
        acc[0] += (double)pci * (double)pci;
        acc[1] += (double)xi * (double)xi;
        acc[2] += (double)ri * (double)ri;
        T xn = xi + a * pci;
        xn = xn + om * sc[i];
        x[i] = xn;
        const T rn = s[i] + nom * t[i];
        r[i] = rn;
        acc[3] += (double)r0[i] * (double)rn;
    }
    block_partials<4>(acc, partials, red, blockIdx.x);
}

// Exit tests of one pass on the grid totals of k_xr_update, in the
// reference's order (divergence, stagnation, convergence on the lagged
// ||r||); when the loop goes on, beta and the next pass.
struct FinPass {
    __device__ void fin(SolveState *st, const double *acc) const
    {
        const double pcn = sqrt(acc[0]), xnrm = sqrt(acc[1]);
        if (fabs(st->alpha) * pcn < st->eps * xnrm) st->stag++;
        else st->stag = 0;
        st->prc = st->rc;
        st->rc = sqrt(acc[2]);
        if (st->trace && st->ntrace < st->trace_cap) st->trace[st->ntrace] = st->rc;
        st->ntrace++;
        if (st->prc < st->rc) st->msteps++;
        else st->msteps = 0;
        if (st->msteps >= st->maxmsteps) {
            st->info = 1;
            st->done = 1;
        } else if (st->stag >= 3) {
            st->info = 2;
            st->done = 1;
        } else if (st->rc < st->tol * st->ro) {
            st->info = 0;
            st->done = 1;
        } else {
            st->m2 = st->m1;
            st->m1 = acc[3];
            st->beta = st->m1 * st->alpha / (st->omega * st->m2);
            st->iter++;
            if (st->iter >= st->nmax) {
                st->info = 3;
                st->done = 1;
            }
        }
    }
};

// p = (-beta omega) v + beta p;  p = p + r   (AFIIM/src/bicgstab.cpp:269-271)
template <class T>
__global__ __launch_bounds__(BS) void k_p_update(int n, T *__restrict__ p, const T *__restrict__ r,
                                                 const T *__restrict__ v, const SolveState *st)
{
    // runs after k_xr_update of the same pass; when that pass exited, p is dead
    if (st->done) return;
    const T be = (T)st->beta, c = (T)(-st->beta * st->omega);
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const T q = c * v[i] + be * p[i];
        p[i] = q + r[i];
    }
}

// Inner-solve setup from x = 0: r = r0 = p = b, ||b||^2 partials (FinSetup).
template <class T>
__global__ __launch_bounds__(BS) void k_setup_zero(int n, const T *__restrict__ b, T *__restrict__ x,
                                                   T *__restrict__ r, T *__restrict__ r0, T *__restrict__ p,
                                                   const SolveState *st, double *partials)
{
    __shared__ double red[BS / WAVE];
    double acc[1] = {0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        const T v = b[i];
        x[i] = (T)0;
        r[i] = v;
        r0[i] = v;
        p[i] = v;
        acc[0] += (double)v * (double)v;
    }
    block_partials<1>(acc, partials, red, blockIdx.x);
    (void)st;
}

// x64 += x32 (refinement update of the FP64 solution).
__global__ __launch_bounds__(BS) void k_x_accum(int n, double *__restrict__ x, const float *__restrict__ d)
{
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) x[i] += (double)d[i];
}

}  // namespace cm
