// Level-scheduled sparse triangular solves (ACIFIM apply modes 'I' and 'J',
// ACIFIM/src/bicgstab.cpp:405-406,421-427: mkl_sparse_d_trsv on the unit
// triangular L and U with their stored diagonal).
//
// The host groups the rows into dependency levels once per factor (a row's
// level is one more than the deepest row it reads); one launch per level then
// solves every row of that level with one thread per row.  Rows of a level
// are independent, so no workgroup ever waits on another and no kernel can
// stall on a dependency: ordering comes from the stream.
#pragma once
#include "cm_kernels.hpp"

namespace cm {

// x[i] = (b[i] - sum_{j in triangle, j != i} T[i,j] x[j]) / T[i,i] for the
// rows of one level.  Only the declared triangle is read (lower: j <= i,
// upper: j >= i), as the TRIANGULAR descriptor makes mkl_sparse_d_trsv do
// (ACIFIM/src/bicgstab.cpp:170-176); a stray entry on the other side is
// ignored rather than read while another level writes it.
template <class T, class V>
__global__ __launch_bounds__(BS) void k_trsv_level(int cnt, bool lower, const int *__restrict__ rows,
                                                   const int *__restrict__ ptr, const int *__restrict__ col,
                                                   const V *__restrict__ val, const float *__restrict__ rowscale,
                                                   const T *__restrict__ b, T *__restrict__ x,
                                                   const SolveState *st)
{
    if (st && st->done) return;
    const int q = blockIdx.x * BS + threadIdx.x;
    if (q >= cnt) return;
    const int i = rows[q];
    double s = (double)b[i], d = 1.0;
    const double sc = rowscale ? (double)rowscale[i] : 1.0;
    for (int k = ptr[i]; k < ptr[i + 1]; k++) {
        const int j = col[k];
        if (lower ? j > i : j < i) continue;
        const double a = load_val<V, double>(val, k) * sc;
        if (j == i) d = a;
        else s -= a * (double)x[j];
    }
    x[i] = (T)(s / d);
}

}  // namespace cm
