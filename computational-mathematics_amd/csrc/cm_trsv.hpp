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
// (the first and last levels of a grid's wavefront).  loff: the levels'
// first positions in level order (nlev + 1 entries, on the device).
constexpr int TRSV_RUN_THREADS = 1024;
constexpr int TRSV_RUN_STEPS = 2;  // a level joins a run when its rows take at most this many steps

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
    const int g = (int)threadIdx.x / LPR, l = (int)(threadIdx.x % LPR);
    for (int lev = l0; lev < l1; lev++) {
        const int q1 = loff[lev + 1];
        for (int q = loff[lev] + g; q < q1; q += G) trsv_row<T, V, LPR>(q, l, rows, tptr, tcol, tval, rowscale, b, x);
        __syncthreads();
    }
}

}  // namespace cm
