// Level-scheduled sparse triangular solves (ACIFIM apply modes 'I' and 'J',
// ACIFIM/src/bicgstab.cpp:405-406,421-427: mkl_sparse_d_trsv on the unit
// triangular L and U with their stored diagonal).
//
// The host groups the rows into dependency levels once per factor (a row's
// level is one more than the deepest row it reads); one launch per level then
// solves every row of that level.  Rows of a level are independent, so no
// workgroup ever waits on another and no kernel can stall on a dependency:
// ordering comes from the stream.
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

}  // namespace cm
