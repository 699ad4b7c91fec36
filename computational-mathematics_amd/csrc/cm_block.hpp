// Multi-right-hand-side (block) BiCGSTAB kernels.
//
// KB right-hand sides of one system are solved together.  Every column keeps
// its own BiCGSTAB recurrence, scalars, exit tests and counters -- the
// reference loop run once per column -- but the matrix products read A, G
// and H once for all columns (SpMM), which is where the bandwidth goes.
//
// Layout: block vectors are interleaved, element (row i, column j) at
// i*KB + j, so the KB values a nonzero multiplies sit in one contiguous
// segment (KB*8 bytes: one 128-byte line for KB = 16 in FP64).
//
// SpMM: rows are binned on the host once per matrix.  Short rows (at most
// SHORT_MAX nonzeros) get KB lanes each (64/KB rows per wave); long rows get
// a whole wave each, split into 64/KB nonzero groups times KB columns and
// reduced across the groups with xor shuffles.  Both kernels walk a row list,
// so a matrix with a few very long rows (the 27-point system with long-range
// couplings) does not stall the short ones.
//
// Per-column state: SolveState[KB]; a column whose done flag is set is
// frozen (its x, r, p are no longer written), and every kernel returns at
// once when all columns are done (BlockFlags::all_done).
#pragma once
#include "cm_kernels.hpp"

namespace cm {


struct BlockFlags {
    int all_done;
    int ncols;  // live columns (the rest of the KB slots are padding)
};

// ------------------------------------------------------------ block epilogues
// row(r, j, s, acc): row r, column j, row sum s; acc[K] per-column partials.

template <class Y>
struct BEpiStore {
    static constexpr int K = 0;
    static constexpr int XV = 0;
    Y *y;
    template <class S>
    __device__ void row(size_t e, int, S s, double *) const { y[e] = (Y)s; }
    __device__ void fin(SolveState *, const double *) const {}
};

template <class Y, class D, int KB>
struct BEpiScale {  // y = d[row] * (A x); element e of the block is row e / KB
    static_assert(KB > 0 && (KB & (KB - 1)) == 0, "block width is a power of two");
    static constexpr int K = 0;
    static constexpr int XV = 0;  // d is an n-vector, counted separately
    Y *y;
    const D *d;
    template <class S>
    __device__ void row(size_t e, int, S s, double *) const { y[e] = (Y)d[e / KB] * (Y)s; }
    __device__ void fin(SolveState *, const double *) const {}
};

template <class T>
struct BEpiSetup {  // r = b - A x; r0 = p = r; ||r||^2
    static constexpr int K = 1;
    static constexpr int XV = 3;
    T *r, *r0, *p;
    const T *b;
    template <class S>
    __device__ void row(size_t e, int, S s, double *acc) const
    {
        const T v = b[e] + (T)(-1.0) * (T)s;
        r[e] = v;
        r0[e] = v;
        p[e] = v;
        acc[0] += (double)v * (double)v;
    }
    __device__ void fin(SolveState *st, const double *tot) const { FinSetup{}.fin(st, tot); }
};

template <class T>
struct BEpiAlpha {
    static constexpr int K = 1;
    static constexpr int XV = 1;
    T *v;
    const T *r0;
    template <class S>
    __device__ void row(size_t e, int, S s, double *acc) const
    {
        const T vi = (T)s;
        v[e] = vi;
        acc[0] += (double)r0[e] * (double)vi;
    }
    __device__ void fin(SolveState *st, const double *tot) const { EpiAlpha<T>{nullptr, nullptr}.fin(st, tot); }
};

template <class T>
struct BEpiOmega {
    static constexpr int K = 2;
    static constexpr int XV = 1;
    T *t;
    const T *s;
    template <class S>
    __device__ void row(size_t e, int, S v, double *acc) const
    {
        const T ti = (T)v;
        t[e] = ti;
        acc[0] += (double)ti * (double)s[e];
        acc[1] += (double)ti * (double)ti;
    }
    __device__ void fin(SolveState *st, const double *tot) const { st->omega = tot[0] / tot[1]; }
};

struct BEpiResid {  // refinement residual: r = b - A x (FP64), FP32 copy, ||r||^2
    static constexpr int K = 1;
    static constexpr int XV = 2;
    const double *b;
    float *r32;
    template <class S>
    __device__ void row(size_t e, int, S s, double *acc) const
    {
        const double v = b[e] - (double)s;
        r32[e] = (float)v;
        acc[0] += v * v;
    }
    __device__ void fin(SolveState *st, const double *tot) const { st->res_outer = sqrt(tot[0]); }
};

// Sum over the workgroup of per-column partials.  Lanes j, j+KB, j+2KB, ...
// of a wave hold column j: xor-shuffle over the row bits, then LDS across
// waves; thread j (< KB) ends with column j's workgroup total.
template <int K, int KB>
__device__ __forceinline__ void block_col_partials(double (&acc)[K > 0 ? K : 1], double *lds /* (BS/WAVE)*K*KB */,
                                                   double *partials, int slot)
{
    if constexpr (K == 0) return;
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < K; k++)
        for (int off = WAVE / 2; off >= KB; off >>= 1) acc[k] += __shfl_xor(acc[k], off, WAVE);
    __syncthreads();
    if (lane < KB)
#pragma unroll
        for (int k = 0; k < K; k++) lds[(w * K + k) * KB + lane] = acc[k];
    __syncthreads();
    if (threadIdx.x < KB)
#pragma unroll
        for (int k = 0; k < K; k++) {
            double s = 0.0;
            for (int q = 0; q < BS / WAVE; q++) s += lds[(q * K + k) * KB + threadIdx.x];
            partials[((size_t)slot * K + k) * KB + threadIdx.x] = s;
        }
}

// Columns per lane in the SpMM kernels: a lane loads 16 bytes of an x row
// (4 FP32 or 2 FP64 columns) per nonzero, or the whole row when it is
// narrower; lanes per row = KB / CPL.
template <class T, int KB>
struct ColSplit {
    static constexpr int CPL = (int)(16 / sizeof(T)) < KB ? (int)(16 / sizeof(T)) : KB;
    static constexpr int LPRW = KB / CPL;
};

// CPL consecutive T from p (aligned to CPL * sizeof(T)).
template <class T, int CPL>
__device__ __forceinline__ void ld_cols(const T *p, T (&v)[CPL])
{
    typedef T vec_t __attribute__((ext_vector_type(CPL)));
    if constexpr (CPL == 1) {
        v[0] = *p;
    } else {
        const vec_t q = *reinterpret_cast<const vec_t *>(p);
#pragma unroll
        for (int i = 0; i < CPL; i++) v[i] = q[i];
    }
}

// Workgroup sums of per-lane partials acc[i][k] (column lane_group * CPL + i,
// value k): xor shuffles over the lanes that share a column group, LDS over
// waves, partials[(slot * K + k) * KB + column].  Fixed order.
template <int K, int KB, int CPL>
__device__ __forceinline__ void block_colgroup_partials(double (&acc)[CPL][K > 0 ? K : 1], double *lds,
                                                        double *partials, int slot)
{
    if constexpr (K == 0) return;
    constexpr int LPRW = KB / CPL;
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
#pragma unroll
    for (int i = 0; i < CPL; i++)
#pragma unroll
        for (int k = 0; k < K; k++)
            for (int off = WAVE / 2; off >= LPRW; off >>= 1) acc[i][k] += __shfl_xor(acc[i][k], off, WAVE);
    __syncthreads();
    if (lane < LPRW)
#pragma unroll
        for (int i = 0; i < CPL; i++)
#pragma unroll
            for (int k = 0; k < K; k++) lds[(w * K + k) * KB + lane * CPL + i] = acc[i][k];
    __syncthreads();
    if (threadIdx.x < KB)
#pragma unroll
        for (int k = 0; k < K; k++) {
            double s = 0.0;
            for (int q = 0; q < BS / WAVE; q++) s += lds[(q * K + k) * KB + threadIdx.x];
            partials[((size_t)slot * K + k) * KB + threadIdx.x] = s;
        }
}

// Short rows, rows from the list rows[0..cnt) (rows null: rows 0..cnt-1, no
// list to read): LPRW lanes per row, each with CPL columns; the row's
// nonzeros in order, four index / value loads issued ahead of their x loads.
template <class V, class T, class Acc, int KB, class Epi>
__global__ __launch_bounds__(BS) void spmm_short(int cnt, const int *__restrict__ rows, const int *__restrict__ ptr,
                                                 const int *__restrict__ col, const V *__restrict__ val,
                                                 const float *__restrict__ rowscale, const T *__restrict__ x,
                                                 Epi epi, const BlockFlags *fl, double *partials, int slot0)
{
    if (fl && fl->all_done) return;
    constexpr int CPL = ColSplit<T, KB>::CPL, LPRW = ColSplit<T, KB>::LPRW, KK = Epi::K > 0 ? Epi::K : 1;
    __shared__ double red[(BS / WAVE) * KK * KB];
    double acc[CPL][KK];
#pragma unroll
    for (int i = 0; i < CPL; i++)
#pragma unroll
        for (int k = 0; k < KK; k++) acc[i][k] = 0.0;
    const int c0 = (threadIdx.x % LPRW) * CPL;
    // a contiguous chunk of the row list per workgroup (XCD-remapped), walked
    // BS / LPRW rows at a time: the grid is capped, so are the partials
    constexpr int SR = BS / LPRW;
    const int chunk = ((cnt + (int)gridDim.x - 1) / (int)gridDim.x + SR - 1) / SR * SR;
    const int q_begin = xcd_block(blockIdx.x, gridDim.x) * chunk, q_end = min(cnt, q_begin + chunk);
    for (int q = q_begin + threadIdx.x / LPRW; q < q_end; q += SR) {
        const int r = rows ? rows[q] : q;
        Acc s[CPL];
#pragma unroll
        for (int i = 0; i < CPL; i++) s[i] = 0;
        int k = ptr[r];
        const int e = ptr[r + 1];
        for (; k + 4 <= e; k += 4) {
            int cc[4];
            Acc vv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                cc[u] = col[k + u];
                vv[u] = load_val<V, Acc>(val, k + u);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                T xv[CPL];
                ld_cols<T, CPL>(x + (size_t)cc[u] * KB + c0, xv);
#pragma unroll
                for (int i = 0; i < CPL; i++) s[i] += vv[u] * (Acc)xv[i];
            }
        }
        for (; k < e; k++) {
            T xv[CPL];
            ld_cols<T, CPL>(x + (size_t)col[k] * KB + c0, xv);
            const Acc vk = load_val<V, Acc>(val, k);
#pragma unroll
            for (int i = 0; i < CPL; i++) s[i] += vk * (Acc)xv[i];
        }
        const Acc sc = rowscale ? (Acc)rowscale[r] : (Acc)1;
#pragma unroll
        for (int i = 0; i < CPL; i++) epi.row((size_t)r * KB + c0 + i, c0 + i, rowscale ? s[i] * sc : s[i], acc[i]);
    }
    block_colgroup_partials<Epi::K, KB, CPL>(acc, red, partials, slot0 + blockIdx.x);
}

// Adds row r's diagonal products over this lane's CPL columns to s,
// CH diagonals in flight per step.
template <class V, class T, class Acc, int KB, int CPL, int CH>
__device__ __forceinline__ void spmm_dia_row(int n, int ldn, int nd, const int *__restrict__ off,
                                             const V *__restrict__ dval, const T *__restrict__ x, int r, int c0,
                                             Acc (&s)[CPL])
{
#pragma clang fp contract(off)
    for (int d0 = 0; d0 < nd; d0 += CH) {
        Acc v[CH];
        T xv[CH][CPL];
#pragma unroll
        for (int q = 0; q < CH; q++) {
            const int d = d0 + q;
            v[q] = 0;
#pragma unroll
            for (int i = 0; i < CPL; i++) xv[q][i] = 0;
            if (d < nd) {
                const int c = r + off[d];
                v[q] = load_val<V, Acc>(dval + (size_t)d * ldn, r);
                if ((unsigned)c < (unsigned)n) ld_cols<T, CPL>(x + (size_t)c * KB + c0, xv[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < CH; q++)
            if (d0 + q < nd) {
#pragma unroll
                for (int i = 0; i < CPL; i++) s[i] += v[q] * (Acc)xv[q][i];
            }
    }
}

// Diagonals plus a CSR remainder (the single-RHS spmv_dia layout: dval[d *
// ldn + i] = A(i, i + off[d]), remainder rptr / rcol / rval, rptr null when
// empty) over KB columns: LPRW lanes per row, each with CPL columns, rows
// walked in contiguous chunks per workgroup like spmm_short.  A diagonal's
// value is one load per row (shared by its lanes), its x segment the row's
// KB columns at row i + off[d], so consecutive rows read consecutive x
// segments and no column index is read for the diagonal part.  Row sums in
// diagonal order, then the remainder.  Rows with long remainders (lpos[row]
// >= 0; lpos null when there are none) take their remainder sums from racc
// (KB values per long row, written before this launch from the remainders'
// column-tiled segments) instead of walking the remainder themselves.
// With REM_LDS the workgroup takes its SR rows at a time, and when their
// remainder entries number at most 2 SR, each lane group loads one or two of
// them (column and value) before the diagonals, gathers their x rows after
// them and writes the KB-wide products to LDS, from which each row adds its
// own in column order; otherwise each row walks its own entries.  The host
// takes REM_LDS for a remainder of at least DIA_REM_LDS_MIN entries per row.
template <class V, class T, class Acc, int KB, class Epi, bool REM_LDS>
__global__ __launch_bounds__(BS) void spmm_dia(int n, int ldn, int nd, const int *__restrict__ off,
                                               const V *__restrict__ dval, const int *__restrict__ rptr,
                                               const int *__restrict__ rcol, const V *__restrict__ rval,
                                               const int *__restrict__ lpos, const T *__restrict__ racc,
                                               const float *__restrict__ rowscale, const T *__restrict__ x, Epi epi,
                                               const BlockFlags *fl, double *partials, int slot0)
{
    if (fl && fl->all_done) return;
    constexpr int CPL = ColSplit<T, KB>::CPL, LPRW = ColSplit<T, KB>::LPRW, KK = Epi::K > 0 ? Epi::K : 1;
    constexpr int CH = 4;  // diagonals in flight per step
    __shared__ double red[(BS / WAVE) * KK * KB];
    double acc[CPL][KK];
#pragma unroll
    for (int i = 0; i < CPL; i++)
#pragma unroll
        for (int k = 0; k < KK; k++) acc[i][k] = 0.0;
    const int c0 = (threadIdx.x % LPRW) * CPL;
    constexpr int SR = BS / LPRW;
    const int chunk = ((n + (int)gridDim.x - 1) / (int)gridDim.x + SR - 1) / SR * SR;
    const int r_begin = xcd_block(blockIdx.x, gridDim.x) * chunk, r_end = min(n, r_begin + chunk);
    if constexpr (!REM_LDS) {
        for (int r = r_begin + (int)threadIdx.x / LPRW; r < r_end; r += SR) {
#pragma clang fp contract(off)
            Acc s[CPL];
#pragma unroll
            for (int i = 0; i < CPL; i++) s[i] = 0;
            spmm_dia_row<V, T, Acc, KB, CPL, CH>(n, ldn, nd, off, dval, x, r, c0, s);
            const int lq = lpos ? lpos[r] : -1;
            if (lq >= 0) {
                T xr[CPL];
                ld_cols<T, CPL>(racc + (size_t)lq * KB + c0, xr);
#pragma unroll
                for (int i = 0; i < CPL; i++) s[i] += (Acc)xr[i];
            } else if (rptr) {
                for (int k = rptr[r]; k < rptr[r + 1]; k++) {
                    T xr[CPL];
                    ld_cols<T, CPL>(x + (size_t)rcol[k] * KB + c0, xr);
                    const Acc vk = load_val<V, Acc>(rval, k);
#pragma unroll
                    for (int i = 0; i < CPL; i++) s[i] += vk * (Acc)xr[i];
                }
            }
            const Acc sc = rowscale ? (Acc)rowscale[r] : (Acc)1;
#pragma unroll
            for (int i = 0; i < CPL; i++)
                epi.row((size_t)r * KB + c0 + i, c0 + i, rowscale ? s[i] * sc : s[i], acc[i]);
        }
    } else {
#pragma clang fp contract(off)
        constexpr int PER = 2;  // staged entries per lane group
        __shared__ Acc prod[PER * SR * KB];
        const int g = (int)threadIdx.x / LPRW;
        for (int p0 = r_begin; p0 < r_end; p0 += SR) {  // uniform over the workgroup
            const int r = p0 + g;
            const bool live = r < r_end;
            int q0 = 0, q1 = 0;
            if (rptr) {
                q0 = rptr[p0];
                q1 = rptr[min(p0 + SR, r_end)];
            }
            const bool stage = q1 - q0 <= PER * SR;
            int ec[PER];
            Acc ev[PER];
#pragma unroll
            for (int j = 0; j < PER; j++) {
                const int e = q0 + j * SR + g;
                ec[j] = -1;
                ev[j] = 0;
                if (stage && e < q1) {
                    ec[j] = rcol[e];
                    ev[j] = load_val<V, Acc>(rval, e);
                }
            }
            Acc s[CPL];
#pragma unroll
            for (int i = 0; i < CPL; i++) s[i] = 0;
            if (live) spmm_dia_row<V, T, Acc, KB, CPL, CH>(n, ldn, nd, off, dval, x, r, c0, s);
            const int lq = live && lpos ? lpos[r] : -1;
            if (live && lq >= 0) {
                T xr[CPL];
                ld_cols<T, CPL>(racc + (size_t)lq * KB + c0, xr);
#pragma unroll
                for (int i = 0; i < CPL; i++) s[i] += (Acc)xr[i];
            }
            if (stage) {
                if (q1 > q0) {
#pragma unroll
                    for (int j = 0; j < PER; j++)
                        if (ec[j] >= 0) {
                            T xr[CPL];
                            ld_cols<T, CPL>(x + (size_t)ec[j] * KB + c0, xr);
#pragma unroll
                            for (int i = 0; i < CPL; i++) prod[(j * SR + g) * KB + c0 + i] = ev[j] * (Acc)xr[i];
                        }
                    __syncthreads();
                    if (live && lq < 0)
                        for (int k = rptr[r]; k < rptr[r + 1]; k++)
#pragma unroll
                            for (int i = 0; i < CPL; i++) s[i] += prod[(k - q0) * KB + c0 + i];
                    __syncthreads();  // the next rows' products overwrite prod
                }
            } else if (live && lq < 0 && rptr) {
                for (int k = rptr[r]; k < rptr[r + 1]; k++) {
                    T xr[CPL];
                    ld_cols<T, CPL>(x + (size_t)rcol[k] * KB + c0, xr);
                    const Acc vk = load_val<V, Acc>(rval, k);
#pragma unroll
                    for (int i = 0; i < CPL; i++) s[i] += vk * (Acc)xr[i];
                }
            }
            if (live) {
                const Acc sc = rowscale ? (Acc)rowscale[r] : (Acc)1;
#pragma unroll
                for (int i = 0; i < CPL; i++)
                    epi.row((size_t)r * KB + c0 + i, c0 + i, rowscale ? s[i] * sc : s[i], acc[i]);
            }
        }
    }
    block_colgroup_partials<Epi::K, KB, CPL>(acc, red, partials, slot0 + blockIdx.x);
}

// Long rows: one wave per row; lane = column group (LPRW of them) + LPRW x
// nonzero group (64 / LPRW groups striding over the row), groups summed with
// xor shuffles.
template <class V, class T, class Acc, int KB, class Epi>
__global__ __launch_bounds__(BS) void spmm_long(int cnt, const int *__restrict__ rows, const int *__restrict__ ptr,
                                                const int *__restrict__ col, const V *__restrict__ val,
                                                const float *__restrict__ rowscale, const T *__restrict__ x, Epi epi,
                                                const BlockFlags *fl, double *partials, int slot0)
{
    if (fl && fl->all_done) return;
    constexpr int CPL = ColSplit<T, KB>::CPL, LPRW = ColSplit<T, KB>::LPRW, KK = Epi::K > 0 ? Epi::K : 1;
    constexpr int G = WAVE / LPRW;
    __shared__ double red[(BS / WAVE) * KK * KB];
    double acc[CPL][KK];
#pragma unroll
    for (int i = 0; i < CPL; i++)
#pragma unroll
        for (int k = 0; k < KK; k++) acc[i][k] = 0.0;
    const int lane = threadIdx.x & (WAVE - 1);
    const int c0 = (lane % LPRW) * CPL, g = lane / LPRW;
    // a contiguous chunk of the row list per workgroup, one row per wave at a
    // time (the loop bound is uniform over the wave: the shuffles below run
    // on every lane)
    constexpr int WPB = BS / WAVE;
    const int chunk = ((cnt + (int)gridDim.x - 1) / (int)gridDim.x + WPB - 1) / WPB * WPB;
    const int q_begin = blockIdx.x * chunk, q_end = min(cnt, q_begin + chunk);
    for (int q0 = q_begin; q0 < q_end; q0 += WPB) {
        const int q = q0 + threadIdx.x / WAVE;
        Acc s[CPL];
#pragma unroll
        for (int i = 0; i < CPL; i++) s[i] = 0;
        int r = -1;
        if (q < q_end) {
            r = rows[q];
            const int e = ptr[r + 1];
            int k = ptr[r] + g;
            for (; k + 3 * G < e; k += 4 * G) {
                int cc[4];
                Acc vv[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    cc[u] = col[k + u * G];
                    vv[u] = load_val<V, Acc>(val, k + u * G);
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    T xv[CPL];
                    ld_cols<T, CPL>(x + (size_t)cc[u] * KB + c0, xv);
#pragma unroll
                    for (int i = 0; i < CPL; i++) s[i] += vv[u] * (Acc)xv[i];
                }
            }
            for (; k < e; k += G) {
                T xv[CPL];
                ld_cols<T, CPL>(x + (size_t)col[k] * KB + c0, xv);
                const Acc vk = load_val<V, Acc>(val, k);
#pragma unroll
                for (int i = 0; i < CPL; i++) s[i] += vk * (Acc)xv[i];
            }
        }
#pragma unroll
        for (int i = 0; i < CPL; i++)
            for (int off = WAVE / 2; off >= LPRW; off >>= 1) s[i] += __shfl_xor(s[i], off, WAVE);
        if (q < q_end && g == 0) {
            const Acc sc = rowscale ? (Acc)rowscale[r] : (Acc)1;
#pragma unroll
            for (int i = 0; i < CPL; i++)
                epi.row((size_t)r * KB + c0 + i, c0 + i, rowscale ? s[i] * sc : s[i], acc[i]);
        }
    }
    block_colgroup_partials<Epi::K, KB, CPL>(acc, red, partials, slot0 + blockIdx.x);
}

// Long remainders cut into column tiles: ne = (long rows) x KB sums, each
// the total of its ntile segment sums (segment t at seg[t * ne + e]), added
// in tile order.
template <class T, class Acc>
__global__ __launch_bounds__(BS) void k_long_tiles(long long ne, int ntile, const T *__restrict__ seg,
                                                   T *__restrict__ racc, const BlockFlags *fl)
{
    if (fl && fl->all_done) return;
    for (long long e = (long long)blockIdx.x * BS + threadIdx.x; e < ne; e += (long long)gridDim.x * BS) {
        Acc s = (Acc)seg[e];
        for (int t = 1; t < ntile; t++) s += (Acc)seg[(size_t)t * ne + e];
        racc[e] = (T)s;
    }
}

// First stage of a two-stage column reduction: workgroup g sums partial
// slots [g * per, (g + 1) * per) of every column (same thread / column
// layout and fixed order as k_block_finalize) into slot g of out.
template <int K, int KB>
__global__ __launch_bounds__(FBS) void k_block_sum1(int nblk, int per, const double *__restrict__ partials,
                                                    double *__restrict__ out, const BlockFlags *fl)
{
    if (fl && fl->all_done) return;
    __shared__ double red[(FBS / WAVE) * K * KB];
    const int j = threadIdx.x % KB, lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    const int b0 = blockIdx.x * per, b1 = min(nblk, b0 + per);
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
    for (int b = b0 + threadIdx.x / KB; b < b1; b += FBS / KB)
#pragma unroll
        for (int k = 0; k < K; k++) s[k] += partials[((size_t)b * K + k) * KB + j];
#pragma unroll
    for (int k = 0; k < K; k++)
        for (int off = WAVE / 2; off >= KB; off >>= 1) s[k] += __shfl_xor(s[k], off, WAVE);
    if (lane < KB)
#pragma unroll
        for (int k = 0; k < K; k++) red[(w * K + k) * KB + lane] = s[k];
    __syncthreads();
    if (threadIdx.x < KB * K) {
        const int jj = threadIdx.x % KB, k = threadIdx.x / KB;
        double q = 0.0;
        for (int v = 0; v < FBS / WAVE; v++) q += red[(v * K + k) * KB + jj];
        out[((size_t)blockIdx.x * K + k) * KB + jj] = q;
    }
}

// One workgroup sums the nblk partials of every column at once: thread t
// takes column t % KB over the slots t / KB, t / KB + FBS / KB, ... (so a
// wave reads whole KB-column rows of the partials, coalesced), the lanes of a
// column are combined with xor shuffles, the waves in order through LDS, and
// thread j runs fin(state[j]) for live column j.  Fixed order throughout.
// Sets all_done when every live column is done.
template <int K, int KB, class Fin>
__global__ __launch_bounds__(FBS) void k_block_finalize(int nblk, const double *__restrict__ partials,
                                                        SolveState *st, BlockFlags *fl, Fin fin, int update_all)
{
    if (fl->all_done) return;
    __shared__ double red[(FBS / WAVE) * K * KB];
    const int j = threadIdx.x % KB, lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
    for (int b = threadIdx.x / KB; b < nblk; b += FBS / KB)
#pragma unroll
        for (int k = 0; k < K; k++) s[k] += partials[((size_t)b * K + k) * KB + j];
#pragma unroll
    for (int k = 0; k < K; k++)
        for (int off = WAVE / 2; off >= KB; off >>= 1) s[k] += __shfl_xor(s[k], off, WAVE);
    if (lane < KB)
#pragma unroll
        for (int k = 0; k < K; k++) red[(w * K + k) * KB + lane] = s[k];
    __syncthreads();
    const int ncols = fl->ncols;
    if (threadIdx.x < ncols) {
        double tot[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            double q = 0.0;
            for (int v = 0; v < FBS / WAVE; v++) q += red[(v * K + k) * KB + threadIdx.x];
            tot[k] = q;
        }
        if (!st[threadIdx.x].done) fin.fin(&st[threadIdx.x], tot);
    }
    __syncthreads();
    if (threadIdx.x == 0 && update_all) {
        int all = 1;
        for (int c = 0; c < ncols; c++) all &= (st[c].done != 0);
        fl->all_done = all;
    }
}

// ------------------------------------------------------ block vector kernels
// Column scalars are read per element from the column's state; done columns
// are left untouched.

template <class T, int KB>
__global__ __launch_bounds__(BS) void k_block_s(size_t ne, const T *__restrict__ r, const T *__restrict__ v,
                                                T *__restrict__ s, const SolveState *st, const BlockFlags *fl)
{
    if (fl->all_done) return;
    __shared__ T na[KB];
    __shared__ int dn[KB];
    if (threadIdx.x < KB) {
        na[threadIdx.x] = (T)(-st[threadIdx.x].alpha);
        dn[threadIdx.x] = st[threadIdx.x].done;
    }
    __syncthreads();
    for (size_t e = (size_t)blockIdx.x * BS + threadIdx.x; e < ne; e += (size_t)gridDim.x * BS) {
        const int j = (int)(e & (KB - 1));
        if (!dn[j]) s[e] = r[e] + na[j] * v[e];
    }
}

template <class T, int KB>
__global__ __launch_bounds__(BS) void k_block_xr(size_t ne, T *__restrict__ x, const T *__restrict__ pc,
                                                 const T *__restrict__ sc, T *__restrict__ r, const T *__restrict__ s,
                                                 const T *__restrict__ t, const T *__restrict__ r0,
                                                 const SolveState *st, const BlockFlags *fl, double *partials)
{
    if (fl->all_done) return;
    __shared__ T al[KB], om[KB];
    __shared__ int dn[KB];
    __shared__ double red[(BS / WAVE) * 4 * KB];
    if (threadIdx.x < KB) {
        al[threadIdx.x] = (T)st[threadIdx.x].alpha;
        om[threadIdx.x] = (T)st[threadIdx.x].omega;
        dn[threadIdx.x] = st[threadIdx.x].done;
    }
    __syncthreads();
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // the grid stride is a multiple of KB, so a thread always sees one column
    for (size_t e = (size_t)blockIdx.x * BS + threadIdx.x; e < ne; e += (size_t)gridDim.x * BS) {
        const int j = (int)(e & (KB - 1));
        if (dn[j]) continue;
        const T pci = pc[e], xi = x[e], ri = r[e];
        acc[0] += (double)pci * (double)pci;
        acc[1] += (double)xi * (double)xi;
        acc[2] += (double)ri * (double)ri;
        T xn = xi + al[j] * pci;
        xn = xn + om[j] * sc[e];
        x[e] = xn;
        const T rn = s[e] + (T)(-om[j]) * t[e];
        r[e] = rn;
        acc[3] += (double)r0[e] * (double)rn;
    }
    block_col_partials<4, KB>(acc, red, partials, blockIdx.x);
}

template <class T, int KB>
__global__ __launch_bounds__(BS) void k_block_p(size_t ne, T *__restrict__ p, const T *__restrict__ r,
                                                const T *__restrict__ v, const SolveState *st, const BlockFlags *fl)
{
    if (fl->all_done) return;
    __shared__ T be[KB], c[KB];
    __shared__ int dn[KB];
    if (threadIdx.x < KB) {
        be[threadIdx.x] = (T)st[threadIdx.x].beta;
        c[threadIdx.x] = (T)(-st[threadIdx.x].beta * st[threadIdx.x].omega);
        dn[threadIdx.x] = st[threadIdx.x].done;
    }
    __syncthreads();
    for (size_t e = (size_t)blockIdx.x * BS + threadIdx.x; e < ne; e += (size_t)gridDim.x * BS) {
        const int j = (int)(e & (KB - 1));
        if (dn[j]) continue;
        const T q = c[j] * v[e] + be[j] * p[e];
        p[e] = q + r[e];
    }
}

// Inner setup from x = 0: r = r0 = p = b; per-column ||b||^2 partials.
template <class T, int KB>
__global__ __launch_bounds__(BS) void k_block_setup_zero(size_t ne, const T *__restrict__ b, T *__restrict__ x,
                                                         T *__restrict__ r, T *__restrict__ r0, T *__restrict__ p,
                                                         double *partials)
{
    __shared__ double red[(BS / WAVE) * KB];
    double acc[1] = {0.0};
    for (size_t e = (size_t)blockIdx.x * BS + threadIdx.x; e < ne; e += (size_t)gridDim.x * BS) {
        const T v = b[e];
        x[e] = (T)0;
        r[e] = v;
        r0[e] = v;
        p[e] = v;
        acc[0] += (double)v * (double)v;
    }
    block_col_partials<1, KB>(acc, red, partials, blockIdx.x);
}

// Interleave / de-interleave between column-major host blocks (n x ncols)
// and the [n][KB] device layout; padding columns are zero.
template <class T, int KB>
__global__ __launch_bounds__(BS) void k_interleave(int n, int ncols, const double *__restrict__ cm, T *__restrict__ il)
{
    const size_t ne = (size_t)n * KB;
    for (size_t e = (size_t)blockIdx.x * BS + threadIdx.x; e < ne; e += (size_t)gridDim.x * BS) {
        const int j = (int)(e & (KB - 1));
        const size_t i = e / KB;
        il[e] = j < ncols ? (T)cm[(size_t)j * n + i] : (T)0;
    }
}

template <int KB>
__global__ __launch_bounds__(BS) void k_deinterleave(int n, int ncols, const double *__restrict__ il, double *__restrict__ cm)
{
    const size_t ne = (size_t)n * KB;
    for (size_t e = (size_t)blockIdx.x * BS + threadIdx.x; e < ne; e += (size_t)gridDim.x * BS) {
        const int j = (int)(e & (KB - 1));
        if (j < ncols) cm[(size_t)j * n + e / KB] = il[e];
    }
}

// x += d per live, not-yet-converged column (refinement update).
template <int KB>
__global__ __launch_bounds__(BS) void k_block_x_accum(size_t ne, double *__restrict__ x, const float *__restrict__ d,
                                                      const int *__restrict__ active)
{
    for (size_t e = (size_t)blockIdx.x * BS + threadIdx.x; e < ne; e += (size_t)gridDim.x * BS)
        if (active[e & (KB - 1)]) x[e] += (double)d[e];
}

}  // namespace cm
