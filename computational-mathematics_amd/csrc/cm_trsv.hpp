// Level-scheduled sparse triangular solves (ACIFIM apply modes 'I' and 'J',
// ACIFIM/src/bicgstab.cpp:405-406,421-427: mkl_sparse_d_trsv on the unit
// triangular L and U with their stored diagonal).
//
// The host groups the rows into dependency levels once per factor (a row's
// level is one more than the deepest row it reads); one launch per level then
// solves every row of that level, or one workgroup a run of consecutive small
// levels.  Rows of a level are independent, so no workgroup ever waits on
// another and no kernel can stall on a dependency: ordering comes from the
// stream, and inside a run from the workgroup's barriers.
#pragma once
#include "cm_kernels.hpp"

namespace cm {

// x[i] = (b[i] - sum_{j in triangle, j != i} T[i,j] x[j]) / T[i,i] for the
// rows of one level (T[i,i] = 1 when the row stores no diagonal).  The
// level's rows are rows[0..cnt) and their entries -- only the declared
// triangle's (lower: j <= i, upper: j >= i), as the TRIANGULAR descriptor
// makes mkl_sparse_d_trsv read (ACIFIM/src/bicgstab.cpp:170-176); a stray
// entry on the other side is never read while another level writes it --
// are tcol / tval[tptr[q] .. tptr[q + 1]), stored in level order so a
// level's entries are one contiguous run.  LPR lanes per row (a power of two
// dividing the wave): lane l takes the row's entries l, l + LPR, ..., so its
// gathers of x are issued together instead of one after another, and a level
// of a few thousand rows still spreads over many workgroups; the lanes' sums
// are added with xor shuffles in a fixed order.
// Row q of a level (in level order) by its LPR lanes; lane l holds the
// row's entries l, l + LPR, ...
template <class T, class V, int LPR>
__device__ __forceinline__ void trsv_row(int q, int l, const int *__restrict__ rows, const int *__restrict__ tptr,
                                         const int *__restrict__ tcol, const V *__restrict__ tval,
                                         const float *__restrict__ rowscale, const T *__restrict__ b, T *x)
{
    const int i = rows[q];
    const int e = tptr[q + 1];
    const double sc = rowscale ? (double)rowscale[i] : 1.0;
    double s = 0.0, d = 0.0, hasd = 0.0;
    for (int k = tptr[q] + l; k < e; k += LPR) {
        const int j = tcol[k];
        const double a = load_val<V, double>(tval, k) * sc;
        if (j == i) {
            d = a;
            hasd = 1.0;
        } else {
            s += a * (double)x[j];
        }
    }
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, LPR);
        d += __shfl_xor(d, o, LPR);
        hasd += __shfl_xor(hasd, o, LPR);
    }
    if (l == 0) x[i] = (T)(((double)b[i] - s) / (hasd > 0.0 ? d : 1.0));
}

template <class T, class V, int LPR>
__global__ __launch_bounds__(BS) void k_trsv_level(int cnt, const int *__restrict__ rows, const int *__restrict__ tptr,
                                                   const int *__restrict__ tcol, const V *__restrict__ tval,
                                                   const float *__restrict__ rowscale, const T *__restrict__ b,
                                                   T *__restrict__ x, const SolveState *st)
{
    static_assert(LPR >= 1 && LPR <= WAVE && (LPR & (LPR - 1)) == 0, "lanes per row: a power of two");
    if (st && st->done) return;
    const int q = (int)((blockIdx.x * BS + threadIdx.x) / LPR), l = (int)(threadIdx.x % LPR);
    if (q >= cnt) return;  // the row's LPR lanes leave together
    trsv_row<T, V, LPR>(q, l, rows, tptr, tcol, tval, rowscale, b, x);
}

// A run of consecutive small levels [l0, l1) in one workgroup of
// TRSV_RUN_THREADS threads: level after level, each row by LPR lanes as in
// k_trsv_level, with a workgroup barrier between levels, which makes the x
// entries one level wrote visible to the next level's rows.  Saves a launch
// per level where a level has too few rows to fill more than a workgroup
// (the first and last levels of a grid's wavefront).  A level's rows, row
// pointers, first TRSV_RUN_PREFETCH entries per lane and right-hand sides
// do not depend on x, so they are loaded while the level before is solved:
// after the barrier a level waits only for its x gathers.  loff: the levels'
// first positions in level order (nlev + 1 entries, on the device).
constexpr int TRSV_RUN_THREADS = 1024;
constexpr int TRSV_RUN_STEPS = 2;     // a level joins a run when its rows take at most this many steps
constexpr int TRSV_RUN_PREFETCH = 2;  // entries per lane loaded a level ahead

template <class V>
struct TrsvRow {  // one row of the next level, loaded ahead
    int i = -1, k0 = 0, k1 = 0;
    int col[TRSV_RUN_PREFETCH];
    double val[TRSV_RUN_PREFETCH];
    double b = 0.0, sc = 1.0;
};

template <class T, class V, int LPR>
__global__ __launch_bounds__(TRSV_RUN_THREADS) void k_trsv_run(int l0, int l1, const int *__restrict__ loff,
                                                               const int *__restrict__ rows,
                                                               const int *__restrict__ tptr,
                                                               const int *__restrict__ tcol, const V *__restrict__ tval,
                                                               const float *__restrict__ rowscale,
                                                               const T *__restrict__ b, T *x, const SolveState *st)
{
    static_assert(LPR >= 1 && LPR <= WAVE && (LPR & (LPR - 1)) == 0, "lanes per row: a power of two");
    if (st && st->done) return;
    constexpr int G = TRSV_RUN_THREADS / LPR;  // rows per step
    constexpr int P = TRSV_RUN_PREFETCH;
    const int g = (int)threadIdx.x / LPR, l = (int)(threadIdx.x % LPR);
    auto fetch = [&](int lev, TrsvRow<V> (&rw)[TRSV_RUN_STEPS]) {
        const int q0 = loff[lev], q1 = loff[lev + 1];
#pragma unroll
        for (int t = 0; t < TRSV_RUN_STEPS; t++) {
            const int q = q0 + g + t * G;
            rw[t].i = -1;
            if (q < q1) {
                const int i = rows[q];
                rw[t].i = i;
                rw[t].k0 = tptr[q];
                rw[t].k1 = tptr[q + 1];
                rw[t].b = (double)b[i];
                rw[t].sc = rowscale ? (double)rowscale[i] : 1.0;
#pragma unroll
                for (int e = 0; e < P; e++) {
                    const int k = rw[t].k0 + l + e * LPR;
                    rw[t].col[e] = -1;
                    rw[t].val[e] = 0.0;
                    if (k < rw[t].k1) {
                        rw[t].col[e] = tcol[k];
                        rw[t].val[e] = load_val<V, double>(tval, k);
                    }
                }
            }
        }
    };
    TrsvRow<V> cur[TRSV_RUN_STEPS], nxt[TRSV_RUN_STEPS];
    fetch(l0, cur);
    for (int lev = l0; lev < l1; lev++) {
        if (lev + 1 < l1) fetch(lev + 1, nxt);
#pragma unroll
        for (int t = 0; t < TRSV_RUN_STEPS; t++) {
            const TrsvRow<V> &rw = cur[t];
            if (rw.i < 0) continue;  // the row group's LPR lanes skip together
            double s = 0.0, d = 0.0, hasd = 0.0;
#pragma unroll
            for (int e = 0; e < P; e++) {
                const int j = rw.col[e];
                if (j < 0) continue;
                const double a = rw.val[e] * rw.sc;
                if (j == rw.i) {
                    d = a;
                    hasd = 1.0;
                } else {
                    s += a * (double)x[j];
                }
            }
            for (int k = rw.k0 + l + P * LPR; k < rw.k1; k += LPR) {
                const int j = tcol[k];
                const double a = load_val<V, double>(tval, k) * rw.sc;
                if (j == rw.i) {
                    d = a;
                    hasd = 1.0;
                } else {
                    s += a * (double)x[j];
                }
            }
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) {
                s += __shfl_xor(s, o, LPR);
                d += __shfl_xor(d, o, LPR);
                hasd += __shfl_xor(hasd, o, LPR);
            }
            if (l == 0) x[rw.i] = (T)((rw.b - s) / (hasd > 0.0 ? d : 1.0));
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < TRSV_RUN_STEPS; t++) cur[t] = nxt[t];
    }
}

}  // namespace cm
