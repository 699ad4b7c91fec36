// Persistent batched solver: many small systems of uneven size in one launch.
//
// Small systems (a few thousand to a few tens of thousands of rows) are
// bound by launch and synchronisation latency on the stream path: every
// BiCGSTAB pass is ~16 kernels of a few microseconds of work each.  Here one
// workgroup of PBS threads runs a whole solve -- the reference loop
// (AFIIM/src/bicgstab.cpp:149-281, same constants, exit order and info
// codes; ACIFIM's non-finite alpha guard, ACIFIM/src/bicgstab.cpp:208-213) --
// with workgroup barriers in place of kernel boundaries, and a persistent
// grid pulls (system, right-hand side) items from a device work queue (one
// atomic counter) until it is empty.  Items are queued largest system first
// so the longest solves start early.  Every workgroup leaves the loop when
// the counter passes the item count, and every solve ends by NMAX, so the
// grid always drains.
//
// Per workgroup: SpMV rows go to groups of LPR lanes (xor-shuffle sums inside
// the group); vector updates stride over the rows; dot products and norms
// are summed in a fixed order (per-thread, wave shuffles, then the waves in
// index order), so a result does not depend on which workgroup ran it or
// what else was in the batch.  The nine work vectors of a solve live in a
// per-workgroup slice of an arena; a system's matrices are read from its
// resident FP64 copies.
//
// restart_true = 1 adds a true-residual check: when the loop converges on
// the recursive residual, r = b - A x is recomputed and, if it is not below
// tol * ||r_0||, BiCGSTAB restarts from the current x (at most MAX_CYCLES
// cycles, NMAX passes in total).  restart_true = 0 is the reference loop
// unchanged.
#pragma once
#include "cm_kernels.hpp"

namespace cm {

constexpr int PBS = 1024;       // threads per persistent workgroup
constexpr int LPR = 8;          // lanes per SpMV row
constexpr int MAX_CYCLES = 4;   // BiCGSTAB cycles with the true-residual check

struct BatchSys {
    int n, variant, maxmsteps, pad;
    const int *ap, *ac;
    const double *av;
    const int *gp, *gc;
    const double *gv;
    const int *hp, *hc;  // H, or G^T on the symmetric PDFIIM branch
    const double *hv;
    const double *idv;
};

struct BatchItem {
    int sys, nmax;
    double tol;
    const double *b;  // device, n values
    double *x;        // device, n values: x0 in, solution out
    double resval;    // out
    int iter, info;   // out
    int cycles, pad;  // out: BiCGSTAB cycles run
};

// Sum of K values over the workgroup, returned to every thread.  Fixed
// order: wave shuffles, then the PBS/WAVE wave sums in wave order.
template <int K>
__device__ __forceinline__ void wg_allsum(double (&v)[K], double *lds /* K * PBS / WAVE */)
{
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) lds[w * K + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        for (int q = 0; q < PBS / WAVE; q++) s += lds[q * K + k];
        v[k] = s;
    }
    __syncthreads();  // lds is reused by the next reduction
}

// y = M x over the workgroup; f(row, sum) is called once per row by the
// row group's first lane.
template <class F>
__device__ __forceinline__ void wg_spmv(int n, const int *__restrict__ ptr, const int *__restrict__ col,
                                        const double *__restrict__ val, const double *__restrict__ x, F f)
{
    const int lane = threadIdx.x & (LPR - 1), g = threadIdx.x / LPR;
    for (int r = g; r < n; r += PBS / LPR) {
        double s = 0.0;
        const int e = ptr[r + 1];
        for (int k = ptr[r] + lane; k < e; k += LPR) s += val[k] * x[col[k]];
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, WAVE);
        if (lane == 0) f(r, s);
    }
}

// out = G (D^-1 (H in)); tm is scratch.
__device__ __forceinline__ void wg_apply(const BatchSys &S, const double *in, double *tm, double *out)
{
    const double *idv = S.idv;
    wg_spmv(S.n, S.hp, S.hc, S.hv, in, [&](int r, double s) { tm[r] = idv[r] * s; });
    __syncthreads();
    wg_spmv(S.n, S.gp, S.gc, S.gv, tm, [&](int r, double s) { out[r] = s; });
    __syncthreads();
}

__device__ void wg_bicgstab(const BatchSys &S, BatchItem &it, double *w, double *lds, int restart_true)
{
    const int n = S.n;
    double *r0 = w, *r = w + n, *p = w + 2 * (size_t)n, *v = w + 3 * (size_t)n, *s = w + 4 * (size_t)n,
           *t = w + 5 * (size_t)n, *pc = w + 6 * (size_t)n, *sc = w + 7 * (size_t)n, *tm = w + 8 * (size_t)n;
    const double *b = it.b;
    double *x = it.x;
    const double eps = 1.11022302462516e-16;  // AFIIM/src/bicgstab.cpp:108
    int iter = 0, info = 3, cycles = 0;
    double ro0 = 0.0, resval = 0.0;
    const int ncyc = restart_true ? MAX_CYCLES : 1;
    for (int cyc = 0; cyc < ncyc; cyc++) {
        // r = b - A x; r0 = p = r; ro = ||r||   (AFIIM/src/bicgstab.cpp:149-165)
        double a1[1] = {0.0};
        wg_spmv(n, S.ap, S.ac, S.av, x, [&](int i, double q) {
            const double ri = b[i] + (-1.0) * q;
            r[i] = ri;
            r0[i] = ri;
            p[i] = ri;
            a1[0] += ri * ri;
        });
        wg_allsum<1>(a1, lds);
        const double ro = sqrt(a1[0]);
        cycles = cyc + 1;
        if (cyc == 0) {
            ro0 = ro;
            if (it.nmax <= 0) {  // NMAX < 0: the reference leaves info as the caller set it (-1 here)
                resval = ro / ro0;
                iter = 0;
                info = it.nmax == 0 ? 3 : -1;
                break;
            }
        } else if (ro < it.tol * ro0) {  // true residual below tol: done
            resval = ro / ro0;
            info = 0;
            break;
        }
        const double target = it.tol * ro0;
        double m1 = ro * ro, m2 = 0.0, rc = ro, prc = 0.0, alpha = 0.0, omega = 0.0, beta = 0.0;
        int stag = 0, msteps = 0;
        info = 3;
        bool conv = false;
        while (iter < it.nmax) {
            wg_apply(S, p, tm, pc);
            double a[1] = {0.0};
            wg_spmv(n, S.ap, S.ac, S.av, pc, [&](int i, double q) {
                v[i] = q;
                a[0] += r0[i] * q;
            });
            wg_allsum<1>(a, lds);
            alpha = m1 / a[0];
            if (S.variant == 2 && !isfinite(alpha)) {  // ACIFIM/src/bicgstab.cpp:208-213
                iter = it.nmax;
                info = 3;
                break;
            }
            for (int i = threadIdx.x; i < n; i += PBS) s[i] = r[i] + (-alpha) * v[i];
            __syncthreads();
            wg_apply(S, s, tm, sc);
            double o[2] = {0.0, 0.0};
            wg_spmv(n, S.ap, S.ac, S.av, sc, [&](int i, double q) {
                t[i] = q;
                o[0] += q * s[i];
                o[1] += q * q;
            });
            wg_allsum<2>(o, lds);
            omega = o[0] / o[1];
            double u[4] = {0.0, 0.0, 0.0, 0.0};
            for (int i = threadIdx.x; i < n; i += PBS) {
                const double pci = pc[i], xi = x[i], ri = r[i];
                u[0] += pci * pci;
                u[1] += xi * xi;
                u[2] += ri * ri;
                double xn = xi + alpha * pci;
                xn = xn + omega * sc[i];
                x[i] = xn;
                const double rn = s[i] + (-omega) * t[i];
                r[i] = rn;
                u[3] += r0[i] * rn;
            }
            wg_allsum<4>(u, lds);
            if (fabs(alpha) * sqrt(u[0]) < eps * sqrt(u[1])) stag++;
            else stag = 0;
            prc = rc;
            rc = sqrt(u[2]);
            if (prc < rc) msteps++;
            else msteps = 0;
            if (msteps >= S.maxmsteps) {
                info = 1;
                break;
            }
            if (stag >= 3) {
                info = 2;
                break;
            }
            if (rc < target) {
                info = 0;
                conv = true;
                break;
            }
            m2 = m1;
            m1 = u[3];
            beta = m1 * alpha / (omega * m2);
            const double c = -beta * omega;
            for (int i = threadIdx.x; i < n; i += PBS) {
                const double q = c * v[i] + beta * p[i];
                p[i] = q + r[i];
            }
            __syncthreads();
            iter++;
        }
        __syncthreads();
        resval = rc / ro0;
        if (!conv) break;  // divergence, stagnation, NMAX: as the reference reports it
    }
    if (threadIdx.x == 0) {
        it.resval = resval;
        it.iter = iter;
        it.info = info;
        it.cycles = cycles;
    }
}

__global__ __launch_bounds__(PBS) void k_batch_solve(int nitems, const BatchSys *__restrict__ sys,
                                                     BatchItem *items, double *arena, size_t slot_stride,
                                                     int *head, int restart_true)
{
    __shared__ double lds[4 * PBS / WAVE];
    __shared__ int item_sh;
    double *w = arena + (size_t)blockIdx.x * slot_stride;
    for (;;) {
        if (threadIdx.x == 0) item_sh = atomicAdd(head, 1);
        __syncthreads();
        const int item = item_sh;
        __syncthreads();
        if (item >= nitems) break;
        BatchItem &it = items[item];
        const BatchSys S = sys[it.sys];
        wg_bicgstab(S, it, w, lds, restart_true);
        __syncthreads();
    }
}

__global__ void k_zero_int(int *p) { p[threadIdx.x] = 0; }

}  // namespace cm
