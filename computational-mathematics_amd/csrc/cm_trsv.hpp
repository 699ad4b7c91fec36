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
// rows of one level (T[i,i] = 1 when the row stores no diagonal).  LPR lanes
// per row (a power of two dividing the wave): lane l takes the row's entries
// l, l + LPR, ..., so its gathers of x are issued together instead of one
// after another, and a level of a few thousand rows still spreads over many
// workgroups; the lanes' sums are added with xor shuffles in a fixed order.
// Only the declared triangle is read (lower: j <= i, upper: j >= i), as the
// TRIANGULAR descriptor makes mkl_sparse_d_trsv do
// (ACIFIM/src/bicgstab.cpp:170-176); a stray entry on the other side is
// ignored rather than read while another level writes it.
template <class T, class V, int LPR>
__global__ __launch_bounds__(BS) void k_trsv_level(int cnt, bool lower, const int *__restrict__ rows,
                                                   const int *__restrict__ ptr, const int *__restrict__ col,
                                                   const V *__restrict__ val, const float *__restrict__ rowscale,
                                                   const T *__restrict__ b, T *__restrict__ x,
                                                   const SolveState *st)
{
    static_assert(LPR >= 1 && LPR <= WAVE && (LPR & (LPR - 1)) == 0, "lanes per row: a power of two");
    if (st && st->done) return;
    const int q = (int)((blockIdx.x * BS + threadIdx.x) / LPR), l = (int)(threadIdx.x % LPR);
    if (q >= cnt) return;  // the row's LPR lanes leave together
    const int i = rows[q];
    const double sc = rowscale ? (double)rowscale[i] : 1.0;
    double s = 0.0, d = 0.0, hasd = 0.0;
    const int e = ptr[i + 1];
    for (int k = ptr[i] + l; k < e; k += LPR) {
        const int j = col[k];
        if (lower ? j > i : j < i) continue;
        const double a = load_val<V, double>(val, k) * sc;
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
