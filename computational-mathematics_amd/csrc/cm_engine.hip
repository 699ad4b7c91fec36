// Host driver and C ABI of the BiCGSTAB solve engine (one GPU, gfx950).
//
// A solve is a device-resident state machine: the host uploads the system
// once, enqueues the setup kernel and then chunks of iterations on the
// system's own stream, and reads the solve state back once per chunk (at
// most two chunks outstanding), so the GPU never waits on the host inside a
// chunk.  Every kernel returns immediately once the state says done, so
// enqueuing past the exit is harmless and the iteration at which the loop
// stops is the reference's, decided on the device.
//
// Per BiCGSTAB pass (mode E / AFIIM / PDFIIM), 12 launches:
//   tm = D^{-1}(H p)   pc = G tm   v = A pc (+ r0.v)   [alpha]   s = r - alpha v
//   tm = D^{-1}(H s)   sc = G tm   t = A sc (+ t.s, t.t)   [omega]
//   x, r update (+ ||pc||, ||x||, ||r||, r0.r)   [exit tests, beta]   p update
// ([..] = one-workgroup finalize of the preceding reduction)
// (AFIIM/src/bicgstab.cpp:168-275, PDFIIM/src/bicgstab.cpp:186-293,
//  ACIFIM/src/bicgstab.cpp:197-304).
//
// Two solve paths share the Krylov core (Krylov<T, VA, VF>: vector, A-value
// and factor-value types):
//  * FP64 (double everywhere) -- the drop-in replacement of the reference;
//  * mixed precision -- FP64 iterative refinement (true residual b - A x,
//    solution and norm in FP64, with the FP64 A) around an inner BiCGSTAB on
//    FP32 A, FP32 or FP16 (per-row scaled) factors and FP32 vectors.  The
//    low-precision matrices are cast on the device from the resident FP64
//    ones and share their index arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "cm_kernels.hpp"
#include "cm_solve.h"
#include "cm_trsv.hpp"

namespace cm {

static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define CM_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(CM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

#define CM_TRY(expr)                  \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != CM_OK) return rc_; \
    } while (0)

// ------------------------------------------------------ device-side casts

// dst = (V) src
template <class V>
__global__ __launch_bounds__(BS) void k_cast(long long cnt, const double *__restrict__ src, V *__restrict__ dst)
{
    for (long long i = (long long)blockIdx.x * BS + threadIdx.x; i < cnt; i += (long long)gridDim.x * BS)
        dst[i] = (V)src[i];
}

// Row-scaled FP16 copy: scale[i] = max |row i| (1 where the row is zero),
// dst = src / scale[i] rounded to FP16 (so |dst| <= 1).
__global__ __launch_bounds__(BS) void k_half_rows(int n, const int *__restrict__ ptr, const double *__restrict__ src,
                                                  __half *__restrict__ dst, float *__restrict__ scale)
{
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    double mx = 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; k++) mx = fmax(mx, fabs(src[k]));
    const float sc = mx > 0.0 ? (float)mx : 1.0f;
    scale[i] = sc;
    for (int k = ptr[i]; k < ptr[i + 1]; k++) dst[k] = __float2half_rn((float)(src[k] / (double)sc));
}

// z = a .* d (mode I of the ACIFIM apply)
template <class T>
__global__ __launch_bounds__(BS) void k_vmul(int n, const T *__restrict__ a, const T *__restrict__ d,
                                             T *__restrict__ z, const SolveState *st)
{
    if (st && st->done) return;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) z[i] = a[i] * d[i];
}

// ------------------------------------------------------------ host helpers

// Row blocks for the adaptive SpMV: consecutive rows with at most TILE
// nonzeros and ROWCAP rows, or a single longer row.
static std::vector<int> make_row_blocks(int n, const int *ptr)
{
    std::vector<int> b;
    b.push_back(0);
    int rows = 0;
    long long nz = 0;
    for (int r = 0; r < n; r++) {
        const int len = ptr[r + 1] - ptr[r];
        if (rows > 0 && (nz + len > TILE || rows >= ROWCAP)) {
            b.push_back(r);
            rows = 0;
            nz = 0;
        }
        rows++;
        nz += len;
        if (len > TILE) {
            b.push_back(r + 1);
            rows = 0;
            nz = 0;
        }
    }
    if (rows > 0) b.push_back(n);
    return b;
}

template <class T>
static int dev_alloc(T **d, size_t count)
{
    *d = nullptr;
    CM_HIP(hipMalloc((void **)d, std::max<size_t>(count, 1) * sizeof(T)));
    return CM_OK;
}

template <class T>
static int dev_upload(T **d, const T *h, size_t count, hipStream_t s)
{
    CM_TRY(dev_alloc(d, count));
    if (h && count) CM_HIP(hipMemcpyAsync(*d, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    return CM_OK;
}

static inline int vec_grid(int n)
{
    int g = (n + BS - 1) / BS;
    return std::max(1, std::min(g, MAX_RED_BLOCKS));
}

static inline int long_grid(long long cnt)
{
    long long g = (cnt + BS - 1) / BS;
    return (int)std::max<long long>(1, std::min<long long>(g, 8192));
}

// A CSR matrix in HBM with its SpMV row blocks (and, for triangular factors
// used by the level-scheduled solves, its level schedule).  A low-precision
// copy shares ptr/col/blk/levels with the FP64 matrix it was cast from
// (owns_index = false).  FP16 values are stored divided by the row's largest
// magnitude, which rowscale holds.
template <class V>
struct DevCsr {
    int n = 0;
    long long nnz = 0;
    int nblk = 0;
    bool owns_index = true;
    int *ptr = nullptr, *col = nullptr, *blk = nullptr;
    V *val = nullptr;
    float *rowscale = nullptr;
    int nlev = 0;
    int *lrows = nullptr;
    std::vector<int> loff;
    bool present() const { return ptr != nullptr; }
    void release()
    {
        if (owns_index) {
            (void)hipFree(ptr);
            (void)hipFree(col);
            (void)hipFree(blk);
            (void)hipFree(lrows);
        }
        (void)hipFree(val);
        (void)hipFree(rowscale);
        ptr = col = blk = lrows = nullptr;
        val = nullptr;
        rowscale = nullptr;
    }
    ~DevCsr() { release(); }
};

static int check_csr(int n, const int *ptr, const int *col)
{
    // The kernels index x by the stored columns and stage rows by ptr: reject
    // anything that would read outside the arrays.
    if (!ptr || !col || ptr[0] != 0) return fail(CM_ERR_ARG, "CSR row pointer must start at 0");
    for (int i = 0; i < n; i++)
        if (ptr[i + 1] < ptr[i]) return fail(CM_ERR_ARG, "CSR row pointer not monotone");
    for (long long k = 0; k < ptr[n]; k++)
        if (col[k] < 0 || col[k] >= n) return fail(CM_ERR_ARG, "CSR column index out of range");
    return CM_OK;
}

static int upload_csr(DevCsr<double> &m, int n, const int *ptr, const int *col, const double *val, hipStream_t s)
{
    CM_TRY(check_csr(n, ptr, col));
    m.release();
    m.owns_index = true;
    m.n = n;
    m.nnz = ptr[n];
    std::vector<int> blk = make_row_blocks(n, ptr);
    m.nblk = (int)blk.size() - 1;
    CM_TRY(dev_upload(&m.ptr, ptr, (size_t)n + 1, s));
    CM_TRY(dev_upload(&m.col, col, (size_t)m.nnz, s));
    CM_TRY(dev_upload(&m.val, val, (size_t)m.nnz, s));
    CM_TRY(dev_upload(&m.blk, blk.data(), blk.size(), s));
    CM_HIP(hipStreamSynchronize(s));  // blk is a host temporary
    return CM_OK;
}

// Low-precision copy of a resident FP64 matrix, cast on the device.
template <class V>
static int cast_csr(DevCsr<V> &dst, const DevCsr<double> &src, hipStream_t s)
{
    dst.release();
    if (!src.present()) return CM_OK;
    dst.n = src.n;
    dst.nnz = src.nnz;
    dst.nblk = src.nblk;
    dst.owns_index = false;
    dst.ptr = src.ptr;
    dst.col = src.col;
    dst.blk = src.blk;
    dst.nlev = src.nlev;
    dst.lrows = src.lrows;
    dst.loff = src.loff;
    CM_TRY(dev_alloc(&dst.val, (size_t)dst.nnz));
    if constexpr (std::is_same<V, __half>::value) {
        CM_TRY(dev_alloc(&dst.rowscale, (size_t)dst.n));
        k_half_rows<<<(dst.n + BS - 1) / BS, BS, 0, s>>>(dst.n, src.ptr, src.val, dst.val, dst.rowscale);
    } else {
        k_cast<V><<<long_grid(dst.nnz), BS, 0, s>>>(dst.nnz, src.val, dst.val);
    }
    CM_HIP(hipGetLastError());
    return CM_OK;
}

// Dependency levels of a unit triangular factor: lower -> a row reads only
// smaller columns (rows in increasing order); upper -> larger columns.
static int build_levels(DevCsr<double> &m, const int *ptr, const int *col, bool lower, hipStream_t s)
{
    const int n = m.n;
    std::vector<int> lev(n, 0);
    int nlev = 0;
    for (int q = 0; q < n; q++) {
        const int i = lower ? q : n - 1 - q;
        int l = 0;
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int j = col[k];
            if ((lower && j < i) || (!lower && j > i)) l = std::max(l, lev[j] + 1);
        }
        lev[i] = l;
        nlev = std::max(nlev, l + 1);
    }
    m.nlev = nlev;
    m.loff.assign(nlev + 1, 0);
    for (int i = 0; i < n; i++) m.loff[lev[i] + 1]++;
    for (int l = 0; l < nlev; l++) m.loff[l + 1] += m.loff[l];
    std::vector<int> rows(n), fill(m.loff.begin(), m.loff.end() - 1);
    for (int i = 0; i < n; i++) rows[fill[lev[i]]++] = i;
    CM_TRY(dev_upload(&m.lrows, rows.data(), (size_t)n, s));
    CM_HIP(hipStreamSynchronize(s));
    return CM_OK;
}

// CSR of the transpose (host).
static void host_transpose(int n, const int *ptr, const int *col, const double *val, std::vector<int> &tp,
                           std::vector<int> &tc, std::vector<double> &tv)
{
    tp.assign(n + 1, 0);
    for (int k = 0; k < ptr[n]; k++) tp[col[k] + 1]++;
    for (int i = 0; i < n; i++) tp[i + 1] += tp[i];
    tc.resize(ptr[n]);
    tv.resize(ptr[n]);
    std::vector<int> fill(tp.begin(), tp.end() - 1);
    for (int i = 0; i < n; i++)
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            int q = fill[col[k]]++;
            tc[q] = i;
            tv[q] = val[k];
        }
}

// Full CSR of a symmetric matrix given its upper triangle (entries below the
// diagonal, if any, are ignored as the SYMMETRIC/UPPER descriptor does).
static void host_sym_expand(int n, const int *ptr, const int *col, const double *val, std::vector<int> &fp,
                            std::vector<int> &fc, std::vector<double> &fv)
{
    fp.assign(n + 1, 0);
    for (int i = 0; i < n; i++)
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int j = col[k];
            if (j < i) continue;
            fp[i + 1]++;
            if (j != i) fp[j + 1]++;
        }
    for (int i = 0; i < n; i++) fp[i + 1] += fp[i];
    fc.resize(fp[n]);
    fv.resize(fp[n]);
    std::vector<int> fill(fp.begin(), fp.end() - 1);
    // lower part first (column order ascending because rows are scanned in
    // order), then the stored upper part of the row
    for (int i = 0; i < n; i++)
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int j = col[k];
            if (j <= i) continue;
            int q = fill[j]++;
            fc[q] = i;
            fv[q] = val[k];
        }
    for (int i = 0; i < n; i++)
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            const int j = col[k];
            if (j < i) continue;
            int q = fill[i]++;
            fc[q] = j;
            fv[q] = val[k];
        }
}

// maxmsteps per package (AFIIM/src/bicgstab.cpp:114, PDFIIM/src/bicgstab.cpp:106,
// ACIFIM/src/bicgstab.cpp:132).
static int max_div_steps(int variant, int n)
{
    if (variant == 2) return std::min(std::max(n, 2000) / 50, 40);
    return std::min(std::max(n, 500) / 50, 20);
}

// Kernel categories of the per-kernel timing (cm_system_kernel_stats).
enum KernelCat { KC_A = 0, KC_F1 = 1, KC_F2 = 2, KC_LU = 3, KC_VEC = 4, KC_SETUP = 5, KC_FIN = 6, KC_N = 7 };

// Optional per-launch timing: an event pair around every launch of a chosen
// category, tagged with the BiCGSTAB pass it belongs to and its algorithmic
// bytes.  Launches the device skipped (after the exit) are dropped when the
// solve ends, so the averages cover executed kernels only.
struct Prof {
    bool on = false;
    unsigned mask = 0;  // categories whose launches are bracketed
    int pass = -1;      // pass being enqueued (-1: setup)
    bool p_update = false;
    struct Rec {
        int cat, pass;
        bool p_update;
        double bytes;
        hipEvent_t e0, e1;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    long long launches[KC_N] = {};
    double ms[KC_N] = {}, bytes[KC_N] = {};

    hipEvent_t get()
    {
        if (pool.empty()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    void reset_stats()
    {
        for (int c = 0; c < KC_N; c++) {
            launches[c] = 0;
            ms[c] = 0.0;
            bytes[c] = 0.0;
        }
    }
    // Fold the finished solve's records into the totals.  A pass executes
    // fully iff it computed its residual (pass < ntrace); its p update runs
    // only when the loop went on (pass < ntrace - 1).
    void collect(int ntrace)
    {
        for (Rec &r : recs) {
            bool ran = (r.pass < 0) || (r.pass < ntrace && (!r.p_update || r.pass < ntrace - 1));
            if (ran && r.e0 && r.e1) {
                float t = 0.f;
                if (hipEventElapsedTime(&t, r.e0, r.e1) == hipSuccess) {
                    launches[r.cat]++;
                    ms[r.cat] += t;
                    bytes[r.cat] += r.bytes;
                }
            }
            if (r.e0) pool.push_back(r.e0);
            if (r.e1) pool.push_back(r.e1);
        }
        recs.clear();
    }
    ~Prof()
    {
        for (Rec &r : recs) {
            if (r.e0) (void)hipEventDestroy(r.e0);
            if (r.e1) (void)hipEventDestroy(r.e1);
        }
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

// Host-side description of a system (CSR arrays owned by the caller).
struct HostSys {
    int variant, n, a_symmetric, nr;
    char mode;
    const int *Ai, *Aj;
    const double *Av;
    const int *Gi, *Gj;
    const double *Gv, *IDv;
    const int *Hi, *Hj;
    const double *Hv;
    const int *Li, *Lj;
    const double *Lv;
    const int *Ui, *Uj;
    const double *Uv;
};

// The BiCGSTAB core for vectors of type T, A values of type VA and factor
// values of type VF.
template <class T, class VA, class VF>
struct Krylov {
    int variant = 0, n = 0, nr = 1;
    char mode = 'E';
    bool sym = false;
    double eps = 1.11022302462516e-16;  // stagnation constant, AFIIM/src/bicgstab.cpp:108
    hipStream_t stream = nullptr;       // owned by the System
    DevCsr<VA> A;
    DevCsr<VF> G, H, GT, L, U;
    T *idv = nullptr;
    T *b = nullptr, *x = nullptr, *r0 = nullptr, *r = nullptr, *p = nullptr, *v = nullptr, *s = nullptr,
      *t = nullptr, *pc = nullptr, *sc = nullptr, *tm = nullptr, *t1 = nullptr, *t2 = nullptr, *t3 = nullptr;
    SolveState *st = nullptr;   // device
    SolveState *hst = nullptr;  // pinned host, 3 slots (init + 2 polling slots)
    double *partials = nullptr;
    double *trace = nullptr;
    int trace_cap_dev = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    Prof prof;

    ~Krylov()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        T *bufs[] = {idv, b, x, r0, r, p, v, s, t, pc, sc, tm, t1, t2, t3};
        for (T *q : bufs) (void)hipFree(q);
        (void)hipFree(partials);
        (void)hipFree(trace);
        (void)hipFree(st);
        if (hst) (void)hipHostFree(hst);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }

    bool uses_gh() const { return variant != 2 || mode == 'E' || mode == 'C'; }
    bool uses_lu() const { return variant == 2 && ((mode == 'C' && nr > 1) || mode == 'I' || mode == 'J'); }

    // vectors, state, partials, events
    int alloc_work(int nblk_a)
    {
        T **vecs[] = {&b, &x, &r0, &r, &p, &v, &s, &t, &pc, &sc, &tm, &t1, &t2, &t3};
        for (T **q : vecs) CM_TRY(dev_alloc(q, n));
        CM_HIP(hipMemsetAsync(x, 0, (size_t)n * sizeof(T), stream));
        CM_HIP(hipMemsetAsync(b, 0, (size_t)n * sizeof(T), stream));
        CM_HIP(hipMalloc((void **)&st, sizeof(SolveState)));
        CM_HIP(hipHostMalloc((void **)&hst, 3 * sizeof(SolveState), hipHostMallocDefault));
        const int maxblk = std::max(nblk_a, MAX_RED_BLOCKS);
        CM_HIP(hipMalloc((void **)&partials, (size_t)maxblk * 4 * sizeof(double)));
        for (auto &e : ev) CM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return CM_OK;
    }

    void copy_shape(int var, int nn, int nr_, char mode_, bool sym_, hipStream_t strm)
    {
        variant = var;
        n = nn;
        nr = nr_;
        mode = mode_;
        sym = sym_;
        stream = strm;
    }

    // Launch f on the stream, bracketed by timing events when profiling.
    template <class F>
    void timed(int cat, double bytes, F f)
    {
        if (!prof.on || !(prof.mask & (1u << cat))) {
            f();
            return;
        }
        Prof::Rec rec{cat, prof.pass, prof.p_update, bytes, prof.get(), prof.get()};
        if (rec.e0) (void)hipEventRecord(rec.e0, stream);
        f();
        if (rec.e1) (void)hipEventRecord(rec.e1, stream);
        prof.recs.push_back(rec);
    }

    // Compulsory bytes of one SpMV launch: values + column indices once, row
    // pointers, row blocks and row scales once, x once, y written once, plus
    // the epilogue's extra n-vectors.
    template <class V, class Epi>
    double spmv_bytes(const DevCsr<V> &m) const
    {
        return (double)m.nnz * (sizeof(V) + sizeof(int)) + 4.0 * (m.n + 1) + 4.0 * (m.nblk + 1) +
               (m.rowscale ? 4.0 * m.n : 0.0) + (double)sizeof(T) * m.n * (2 + Epi::XV);
    }

    template <class V, class Epi>
    void spmv(const DevCsr<V> &m, const T *xin, Epi epi, bool use_state, int cat, SolveState *fst = nullptr)
    {
        if (m.nblk == 0) return;
        using Acc = typename std::conditional<std::is_same<T, double>::value, double, float>::type;
        // fst: a separate state (refinement residual) instead of the solve's
        SolveState *kst = fst ? fst : st;
        timed(cat, spmv_bytes<V, Epi>(m), [&] {
            spmv_kernel<V, T, Acc, Epi><<<m.nblk, BS, 0, stream>>>(m.ptr, m.col, m.val, m.rowscale, m.blk, xin,
                                                                  epi, use_state ? kst : nullptr, partials);
        });
        if constexpr (Epi::K > 0)
            timed(KC_FIN, 8.0 * Epi::K * m.nblk,
                  [&] { k_finalize<Epi::K, Epi><<<1, FBS, 0, stream>>>(m.nblk, partials, kst, epi); });
    }

    // x = T^{-1} b for a unit triangular factor with its level schedule
    void trsv(const DevCsr<VF> &m, const T *bin, T *xout, bool use_state)
    {
        for (int l = 0; l < m.nlev; l++) {
            const int cnt = m.loff[l + 1] - m.loff[l];
            const int g = (cnt + BS - 1) / BS;
            timed(KC_LU, (double)cnt * 3.0 * sizeof(T), [&] {
                k_trsv_level<T, VF><<<g, BS, 0, stream>>>(cnt, m.lrows + m.loff[l], m.ptr, m.col, m.val,
                                                          m.rowscale, bin, xout, use_state ? st : nullptr);
            });
        }
    }

    // out = M in, per the reference's apply (ACIFIM/src/bicgstab.cpp:372-428);
    // for AFIIM / PDFIIM: out = G (IDv .* (H in)), H = G^T when symmetric.
    void enqueue_apply(const T *in, T *out, bool use_state)
    {
        if (variant != 2 || mode == 'E') {
            const DevCsr<VF> &Hm = sym ? GT : H;
            spmv(Hm, in, EpiScale<T, T>{tm, idv}, use_state, KC_F1);
            spmv(G, tm, EpiStore<T>{out}, use_state, KC_F2);
            return;
        }
        if (mode == 'I' || mode == 'J') {
            // I: t1 = L^{-1} b; t2 = t1 .* IDv; x = U^{-1} t2.   J: x = U^{-1} L^{-1} b.
            trsv(L, in, t1, use_state);
            if (mode == 'I') {
                const int g = vec_grid(n);
                timed(KC_VEC, 3.0 * sizeof(T) * n, [&] {
                    k_vmul<T><<<g, BS, 0, stream>>>(n, t1, idv, t2, use_state ? st : nullptr);
                });
                trsv(U, t2, out, use_state);
            } else {
                trsv(U, t1, out, use_state);
            }
            return;
        }
        // mode C: t1 = H b; repeat: t1 += H (b - L t1); t2 = t1 .* IDv;
        //         x = G t2; repeat: x += G (t2 - U x)
        if (nr <= 1) {
            spmv(H, in, EpiScale<T, T>{t2, idv}, use_state, KC_F1);
        } else {
            spmv(H, in, EpiStore<T>{t1}, use_state, KC_F1);
            for (int it = 1; it < nr; it++) {
                spmv(L, t1, EpiSubFrom<T>{t3, in}, use_state, KC_LU);
                if (it + 1 < nr) spmv(H, t3, EpiAccum<T>{t1}, use_state, KC_F1);
                else spmv(H, t3, EpiAccumScale<T, T>{t1, t2, idv}, use_state, KC_F1);
            }
        }
        spmv(G, t2, EpiStore<T>{out}, use_state, KC_F2);
        for (int it = 1; it < nr; it++) {
            spmv(U, out, EpiSubFrom<T>{t3, t2}, use_state, KC_LU);
            spmv(G, t3, EpiAccum<T>{out}, use_state, KC_F2);
        }
    }

    void enqueue_iteration()
    {
        const int g = vec_grid(n);
        const double nn = (double)n, w = (double)sizeof(T);
        prof.pass++;
        enqueue_apply(p, pc, true);
        spmv(A, pc, EpiAlpha<T>{v, r0}, true, KC_A);
        timed(KC_VEC, 3.0 * w * nn, [&] { k_s_update<T><<<g, BS, 0, stream>>>(n, r, v, s, st); });
        enqueue_apply(s, sc, true);
        spmv(A, sc, EpiOmega<T>{t, s}, true, KC_A);
        timed(KC_VEC, 9.0 * w * nn, [&] {
            k_xr_update<T><<<g, BS, 0, stream>>>(n, x, pc, sc, r, s, t, r0, st, partials);
        });
        timed(KC_FIN, 32.0 * g,
              [&] { k_finalize<4, FinPass><<<1, FBS, 0, stream>>>(g, partials, st, FinPass{}); });
        prof.p_update = true;
        timed(KC_VEC, 4.0 * w * nn, [&] { k_p_update<T><<<g, BS, 0, stream>>>(n, p, r, v, st); });
        prof.p_update = false;
    }

    // Writes the initial state (done = 0) of a solve of NMAX passes.
    int begin(double tol, int NMAX, int info_in, int trace_cap)
    {
        if (trace_cap > 0 && trace_cap > trace_cap_dev) {
            (void)hipFree(trace);
            trace = nullptr;
            CM_HIP(hipMalloc((void **)&trace, (size_t)trace_cap * sizeof(double)));
            trace_cap_dev = trace_cap;
        }
        SolveState init;
        std::memset(&init, 0, sizeof init);
        init.tol = tol;
        init.eps = eps;
        init.nmax = NMAX;
        init.maxmsteps = max_div_steps(variant, n);
        init.variant = variant;
        init.info = info_in;
        init.trace = trace_cap > 0 ? trace : nullptr;
        init.trace_cap = trace_cap > 0 ? trace_cap : 0;
        hst[0] = init;
        CM_HIP(hipMemcpyAsync(st, &hst[0], sizeof(SolveState), hipMemcpyHostToDevice, stream));
        prof.pass = -1;
        return CM_OK;
    }

    // r = b - A x, r0 = p = r (x as given)
    void setup_from_x() { spmv(A, x, EpiSetup<T>{r, r0, p, b}, true, KC_SETUP); }

    // x = 0, r = r0 = p = b
    void setup_zero()
    {
        const int g = vec_grid(n);
        timed(KC_SETUP, 5.0 * sizeof(T) * n,
              [&] { k_setup_zero<T><<<g, BS, 0, stream>>>(n, b, x, r, r0, p, st, partials); });
        timed(KC_FIN, 8.0 * g,
              [&] { k_finalize<1, FinSetup><<<1, FBS, 0, stream>>>(g, partials, st, FinSetup{}); });
    }

    // Runs passes until the device state says done.  Passes are enqueued in
    // chunks, at most two chunks outstanding, and the state is read back at
    // the end of each chunk.  The next chunk's size is the number of passes
    // the residual's recent contraction rate says are still needed (at least
    // 2, at most 32), so few no-op passes are enqueued past the exit.
    int run(int NMAX, SolveState &fin_state)
    {
        int qs[2];
        int nq = 0, slot = 0, last_chunk = 4;
        long long enq = 0;
        auto push = [&](int size) -> int {
            for (int i = 0; i < size; i++) enqueue_iteration();
            CM_HIP(hipMemcpyAsync(&hst[1 + slot], st, sizeof(SolveState), hipMemcpyDeviceToHost, stream));
            CM_HIP(hipEventRecord(ev[slot], stream));
            enq += size;
            qs[nq++] = slot;
            slot ^= 1;
            last_chunk = size;
            return CM_OK;
        };
        const long long nmax_passes = std::max(NMAX, 0);
        CM_TRY(push((int)std::min<long long>(4, std::max<long long>(nmax_passes, 1))));
        bool have_prev = false;
        int prev_iter = 0;
        double prev_rc = 0.0;
        for (;;) {
            if (nq < 2 && enq < nmax_passes) {
                long long want = std::min(2 * last_chunk, 32);
                if (have_prev) {
                    const SolveState &sn = fin_state;  // latest snapshot
                    if (sn.iter > prev_iter && sn.rc < prev_rc && sn.rc > 0.0 && prev_rc > 0.0) {
                        const double rate = std::pow(sn.rc / prev_rc, 1.0 / (sn.iter - prev_iter));
                        if (rate < 1.0) {
                            const double need = std::ceil(std::log(sn.tol * sn.ro / sn.rc) / std::log(rate)) + 1.0;
                            const long long ahead = enq - sn.iter;
                            want = (long long)std::max(2.0, std::min(32.0, need - (double)ahead));
                        }
                    }
                }
                want = std::min(want, nmax_passes - enq);
                if (want > 0) CM_TRY(push((int)want));
            }
            if (nq == 0) return fail(CM_ERR_HIP, "solve did not reach an exit");
            const int o = qs[0];
            qs[0] = qs[1];
            nq--;
            CM_HIP(hipEventSynchronize(ev[o]));
            const SolveState snap = hst[1 + o];
            if (have_prev || snap.iter > 0) {
                prev_iter = have_prev ? fin_state.iter : 0;
                prev_rc = have_prev ? fin_state.rc : snap.ro;
            }
            fin_state = snap;
            have_prev = true;
            if (snap.done) break;
        }
        CM_HIP(hipStreamSynchronize(stream));
        CM_HIP(hipGetLastError());
        if (prof.on) prof.collect(fin_state.ntrace);
        if (!fin_state.done) return fail(CM_ERR_HIP, "solve state did not reach done");
        return CM_OK;
    }
};

using K64 = Krylov<double, double, double>;
using K32 = Krylov<float, float, float>;
using K16 = Krylov<float, float, __half>;

static int create_k64(K64 &k, const HostSys &h, hipStream_t strm)
{
    const char mode = (h.variant == 2) ? h.mode : 'E';
    k.copy_shape(h.variant, h.n, h.nr, mode, h.variant == 1 && h.a_symmetric, strm);
    const int n = h.n;
    if (k.sym) {
        std::vector<int> fp, fc;
        std::vector<double> fv;
        host_sym_expand(n, h.Ai, h.Aj, h.Av, fp, fc, fv);
        CM_TRY(upload_csr(k.A, n, fp.data(), fc.data(), fv.data(), strm));
        std::vector<int> tp, tc;
        std::vector<double> tv;
        host_transpose(n, h.Gi, h.Gj, h.Gv, tp, tc, tv);
        CM_TRY(upload_csr(k.GT, n, tp.data(), tc.data(), tv.data(), strm));
    } else {
        CM_TRY(upload_csr(k.A, n, h.Ai, h.Aj, h.Av, strm));
        if (k.uses_gh()) CM_TRY(upload_csr(k.H, n, h.Hi, h.Hj, h.Hv, strm));
    }
    if (k.uses_gh()) CM_TRY(upload_csr(k.G, n, h.Gi, h.Gj, h.Gv, strm));
    if (k.uses_lu()) {
        CM_TRY(upload_csr(k.L, n, h.Li, h.Lj, h.Lv, strm));
        CM_TRY(upload_csr(k.U, n, h.Ui, h.Uj, h.Uv, strm));
        if (k.mode == 'I' || k.mode == 'J') {
            CM_TRY(build_levels(k.L, h.Li, h.Lj, true, strm));
            CM_TRY(build_levels(k.U, h.Ui, h.Uj, false, strm));
        }
    }
    CM_TRY(dev_upload(&k.idv, h.IDv, (size_t)n, strm));
    CM_TRY(k.alloc_work(k.A.nblk));
    CM_HIP(hipStreamSynchronize(strm));
    return CM_OK;
}

// Low-precision inner solver cast from the resident FP64 one.
template <class KI>
static int create_inner(KI &ki, const K64 &k)
{
    ki.copy_shape(k.variant, k.n, k.nr, k.mode, k.sym, k.stream);
    ki.eps = 5.9604644775390625e-08;  // 2^-24: the stagnation test at FP32 resolution
    CM_TRY(cast_csr(ki.A, k.A, k.stream));
    CM_TRY(cast_csr(ki.G, k.G, k.stream));
    CM_TRY(cast_csr(ki.H, k.H, k.stream));
    CM_TRY(cast_csr(ki.GT, k.GT, k.stream));
    CM_TRY(cast_csr(ki.L, k.L, k.stream));
    CM_TRY(cast_csr(ki.U, k.U, k.stream));
    CM_TRY(dev_alloc(&ki.idv, (size_t)k.n));
    k_cast<float><<<long_grid(k.n), BS, 0, k.stream>>>(k.n, k.idv, ki.idv);
    CM_TRY(ki.alloc_work(ki.A.nblk));
    CM_HIP(hipStreamSynchronize(k.stream));
    return CM_OK;
}

// Results of the last mixed-precision solve.
struct MixedStats {
    int outer = 0, inner_total = 0, info = 0;
    double resval = 0.0;
};

struct System {
    hipStream_t stream = nullptr;
    int n = 0, variant = 0;
    std::unique_ptr<K64> k64;
    std::unique_ptr<K32> k32;
    std::unique_ptr<K16> k16;
    SolveState *ost = nullptr;       // device state of the refinement residual
    SolveState *host_ost = nullptr;  // pinned
    MixedStats mixed;
    ~System()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        k16.reset();
        k32.reset();
        k64.reset();
        (void)hipFree(ost);
        if (host_ost) (void)hipHostFree(host_ost);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static int validate(const HostSys &h)
{
    if (h.n <= 0 || !h.Ai || !h.Aj || !h.Av || !h.IDv) return fail(CM_ERR_ARG, "null or empty system");
    if (h.variant < 0 || h.variant > 2) return fail(CM_ERR_ARG, "bad variant");
    const char mode = h.variant == 2 ? h.mode : 'E';
    if (mode != 'E' && mode != 'C' && mode != 'I' && mode != 'J') return fail(CM_ERR_ARG, "unknown apply mode");
    const bool sym = h.variant == 1 && h.a_symmetric;
    const bool need_gh = mode == 'E' || mode == 'C';
    if (!h.Gi || !h.Gj || !h.Gv) {
        if (need_gh) return fail(CM_ERR_ARG, "G factor missing");
    }
    if (need_gh && !sym && (!h.Hi || !h.Hj || !h.Hv)) return fail(CM_ERR_ARG, "H factor missing");
    const bool need_lu = h.variant == 2 && ((mode == 'C' && h.nr > 1) || mode == 'I' || mode == 'J');
    if (need_lu && (!h.Li || !h.Lj || !h.Lv || !h.Ui || !h.Uj || !h.Uv))
        return fail(CM_ERR_ARG, "L/U factors missing for this apply mode");
    return CM_OK;
}

static int system_create(System &S, const HostSys &h)
{
    CM_TRY(validate(h));
    S.n = h.n;
    S.variant = h.variant;
    CM_HIP(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
    S.k64 = std::make_unique<K64>();
    return create_k64(*S.k64, h, S.stream);
}

static void print_verbose(int variant, int iter, int info, double rel)
{
    if (info == 0) {
        std::cout << "Convergence in " << iter << " iterations with relative residual " << rel << "\n";
    } else if (info == 1) {
        std::cout << "BiCGSTAB diverged and iterations stopped with relative residual " << rel << " in iteration "
                  << iter << "\n";
        std::cout << "Try to increase the quality of preconditioning\n";
    } else if (info == 2) {
        if (variant == 1) {
            std::cout << "BiCGSTAB diverged and iterations stopped with relative residual " << rel
                      << " in iteration " << iter << "\n";
            std::cout << "Try to increase the quality of preconditioning\n";
        } else {
            std::cout << "BiCGSTAB stagnated with relative residual " << rel << " in iteration " << iter << "\n";
            std::cout << "Try increasing prescribed tolerance\n";
        }
    } else if (info == 3) {
        std::cout << "Convergence failed after " << iter << " iterations with relative residual " << rel << "\n";
    }
}

// FP64 solve on the resident b and x.
static int solve64(System &S, double tol, int NMAX, double *resval, int *iter, int *info, double *trace,
                   int trace_cap, int *ntrace)
{
    K64 &k = *S.k64;
    CM_TRY(k.begin(tol, NMAX, info ? *info : 0, trace ? trace_cap : 0));
    k.setup_from_x();
    SolveState fin;
    CM_TRY(k.run(NMAX, fin));
    *iter = fin.iter;
    if (info) *info = fin.info;
    *resval = fin.rc / fin.ro;
    if (ntrace) *ntrace = fin.ntrace;
    if (trace && trace_cap > 0) {
        const int cnt = std::min(fin.ntrace, trace_cap);
        if (cnt > 0) CM_HIP(hipMemcpy(trace, k.trace, (size_t)cnt * sizeof(double), hipMemcpyDeviceToHost));
    }
    return CM_OK;
}

// Mixed precision: FP64 refinement around the low-precision inner solve.
//   r = b - A x (FP64 A, FP64 vectors), rn = ||r||; stop when rn <= tol * rn_0;
//   solve A d = r with the inner BiCGSTAB to inner_tol (FP32 A, FP32/FP16
//   factors, FP32 vectors, from d = 0); x += d in FP64.
// info: 0 converged; 2 an outer step did not reduce ||r|| (refinement stalled);
// 1 a non-finite residual; 3 max_outer steps without convergence.
template <class KI>
static int solve_mixed(System &S, KI &ki, double tol, double inner_tol, int NMAX, int max_outer, MixedStats &out)
{
    K64 &k = *S.k64;
    const int n = S.n;
    const int g = vec_grid(n);
    out = MixedStats{};
    double res0 = -1.0, prev = -1.0;
    for (int outer = 0;; outer++) {
        std::memset(S.host_ost, 0, sizeof(SolveState));
        CM_HIP(hipMemcpyAsync(S.ost, S.host_ost, sizeof(SolveState), hipMemcpyHostToDevice, S.stream));
        k.spmv(k.A, k.x, EpiResid{k.b, ki.b}, true, KC_SETUP, S.ost);
        CM_HIP(hipMemcpyAsync(S.host_ost, S.ost, sizeof(SolveState), hipMemcpyDeviceToHost, S.stream));
        CM_HIP(hipStreamSynchronize(S.stream));
        const double rn = S.host_ost->res_outer;
        if (res0 < 0.0) res0 = rn;
        out.resval = res0 > 0.0 ? rn / res0 : 0.0;
        out.outer = outer;
        if (!std::isfinite(rn)) {
            out.info = 1;
            return CM_OK;
        }
        if (!(rn > tol * res0)) {
            out.info = 0;
            return CM_OK;
        }
        if (outer >= max_outer) {
            out.info = 3;
            return CM_OK;
        }
        if (prev >= 0.0 && !(rn < prev)) {
            out.info = 2;
            return CM_OK;
        }
        prev = rn;
        CM_TRY(ki.begin(inner_tol, NMAX, 0, 0));
        ki.setup_zero();
        SolveState fin;
        CM_TRY(ki.run(NMAX, fin));
        out.inner_total += fin.ntrace;
        k_x_accum<<<g, BS, 0, S.stream>>>(n, k.x, ki.x);
        CM_HIP(hipGetLastError());
    }
}

static int dropin_solve(System &S, const double *b, double *x, double tol, int NMAX, double *resval, int *iter,
                        int verbose, int *info)
{
    const int n = S.n;
    K64 &k = *S.k64;
    CM_HIP(hipMemcpyAsync(k.b, b, (size_t)n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    CM_HIP(hipMemcpyAsync(k.x, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    CM_TRY(solve64(S, tol, NMAX, resval, iter, info, nullptr, 0, nullptr));
    CM_HIP(hipMemcpy(x, k.x, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    // the reference prints only when verbose (AFIIM/src/bicgstab.cpp:227-281)
    if (verbose) print_verbose(S.variant, *iter, *info, *resval);
    return CM_OK;
}

}  // namespace cm

using namespace cm;

struct cm_system_s {
    System sys;
};

extern "C" {

const char *cm_last_error(void) { return g_last_error.c_str(); }

int cm_device_count(void)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int cm_system_create(cm_system *out, int variant, int n, int a_symmetric, const int *Ai, const int *Aj,
                     const double *Av, const int *Gi, const int *Gj, const double *Gv, const double *IDv,
                     const int *Hi, const int *Hj, const double *Hv, const int *Li, const int *Lj,
                     const double *Lv, const int *Ui, const int *Uj, const double *Uv, int nr, char mode)
{
    if (!out) return fail(CM_ERR_ARG, "null handle");
    *out = nullptr;
    if (cm_device_count() <= 0) return fail(CM_ERR_NODEV, "no GPU visible");
    auto h = std::make_unique<cm_system_s>();
    const HostSys hs{variant, n, a_symmetric, nr, mode, Ai, Aj, Av, Gi, Gj, Gv, IDv, Hi, Hj, Hv, Li, Lj, Lv, Ui, Uj, Uv};
    CM_TRY(system_create(h->sys, hs));
    *out = h.release();
    return CM_OK;
}

int cm_system_destroy(cm_system sys)
{
    delete sys;
    return CM_OK;
}

int cm_system_set_rhs(cm_system h, const double *b)
{
    if (!h || !b) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    CM_HIP(hipMemcpyAsync(S.k64->b, b, (size_t)S.n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    CM_HIP(hipStreamSynchronize(S.stream));
    return CM_OK;
}

int cm_system_set_x(cm_system h, const double *x)
{
    if (!h) return fail(CM_ERR_ARG, "null handle");
    System &S = h->sys;
    if (x) CM_HIP(hipMemcpyAsync(S.k64->x, x, (size_t)S.n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    else CM_HIP(hipMemsetAsync(S.k64->x, 0, (size_t)S.n * sizeof(double), S.stream));
    CM_HIP(hipStreamSynchronize(S.stream));
    return CM_OK;
}

int cm_system_get_x(cm_system h, double *x)
{
    if (!h || !x) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    CM_HIP(hipStreamSynchronize(S.stream));
    CM_HIP(hipMemcpy(x, S.k64->x, (size_t)S.n * sizeof(double), hipMemcpyDeviceToHost));
    return CM_OK;
}

int cm_system_solve(cm_system h, double tol, int NMAX, double *resval, int *iter, int *info, double *trace,
                    int trace_cap, int *ntrace)
{
    if (!h || !resval || !iter) return fail(CM_ERR_ARG, "null argument");
    return solve64(h->sys, tol, NMAX, resval, iter, info, trace, trace_cap, ntrace);
}

int cm_system_enable_mixed(cm_system h, int factor_bits)
{
    if (!h) return fail(CM_ERR_ARG, "null handle");
    System &S = h->sys;
    if (factor_bits != 32 && factor_bits != 16) return fail(CM_ERR_ARG, "factor_bits must be 32 or 16");
    if (!S.ost) {
        CM_HIP(hipMalloc((void **)&S.ost, sizeof(SolveState)));
        CM_HIP(hipHostMalloc((void **)&S.host_ost, sizeof(SolveState), hipHostMallocDefault));
    }
    if (factor_bits == 32 && !S.k32) {
        S.k32 = std::make_unique<K32>();
        CM_TRY(create_inner(*S.k32, *S.k64));
    }
    if (factor_bits == 16 && !S.k16) {
        S.k16 = std::make_unique<K16>();
        CM_TRY(create_inner(*S.k16, *S.k64));
    }
    return CM_OK;
}

int cm_system_solve_mixed(cm_system h, int factor_bits, double tol, double inner_tol, int NMAX, int max_outer,
                          double *resval, int *outer, int *inner_total, int *info)
{
    if (!h || !resval) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    MixedStats ms;
    if (factor_bits == 32) {
        if (!S.k32) return fail(CM_ERR_ARG, "FP32 inner solver not enabled (cm_system_enable_mixed)");
        CM_TRY(solve_mixed(S, *S.k32, tol, inner_tol, NMAX, max_outer, ms));
    } else if (factor_bits == 16) {
        if (!S.k16) return fail(CM_ERR_ARG, "FP16-factor inner solver not enabled (cm_system_enable_mixed)");
        CM_TRY(solve_mixed(S, *S.k16, tol, inner_tol, NMAX, max_outer, ms));
    } else {
        return fail(CM_ERR_ARG, "factor_bits must be 32 or 16");
    }
    S.mixed = ms;
    *resval = ms.resval;
    if (outer) *outer = ms.outer;
    if (inner_total) *inner_total = ms.inner_total;
    if (info) *info = ms.info;
    return CM_OK;
}

int cm_system_spmv(cm_system h, char which, const double *x, double *y)
{
    if (!h || !x || !y) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    K64 &k = *S.k64;
    const DevCsr<double> *m = nullptr;
    switch (which) {
    case 'A': m = &k.A; break;
    case 'G': m = &k.G; break;
    case 'H': m = k.sym ? &k.GT : &k.H; break;
    case 'L': m = &k.L; break;
    case 'U': m = &k.U; break;
    case 'T': m = &k.GT; break;
    default: break;
    }
    if (!m || !m->present()) return fail(CM_ERR_ARG, "matrix not present");
    CM_HIP(hipMemcpyAsync(k.t1, x, (size_t)S.n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    const bool prof_on = k.prof.on;  // kernel-level calls are not part of a solve's timing
    k.prof.on = false;
    k.spmv(*m, k.t1, EpiStore<double>{k.t2}, false, KC_A);
    k.prof.on = prof_on;
    CM_HIP(hipMemcpyAsync(y, k.t2, (size_t)S.n * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    CM_HIP(hipStreamSynchronize(S.stream));
    CM_HIP(hipGetLastError());
    return CM_OK;
}

int cm_system_apply(cm_system h, const double *v, double *out)
{
    if (!h || !v || !out) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    K64 &k = *S.k64;
    CM_HIP(hipMemcpyAsync(k.s, v, (size_t)S.n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    const bool prof_on = k.prof.on;
    k.prof.on = false;
    k.enqueue_apply(k.s, k.sc, false);
    k.prof.on = prof_on;
    CM_HIP(hipMemcpyAsync(out, k.sc, (size_t)S.n * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    CM_HIP(hipStreamSynchronize(S.stream));
    CM_HIP(hipGetLastError());
    return CM_OK;
}

long long cm_system_nnz(cm_system h, char which)
{
    if (!h) return -1;
    K64 &k = *h->sys.k64;
    const DevCsr<double> *m = nullptr;
    switch (which) {
    case 'A': m = &k.A; break;
    case 'G': m = &k.G; break;
    case 'H': m = k.sym ? &k.GT : &k.H; break;
    case 'L': m = &k.L; break;
    case 'U': m = &k.U; break;
    case 'T': m = &k.GT; break;
    default: break;
    }
    return (m && m->present()) ? m->nnz : -1;
}

int cm_system_profile(cm_system h, int mask)
{
    if (!h) return fail(CM_ERR_ARG, "null handle");
    System &S = h->sys;
    Prof *ps[] = {&S.k64->prof, S.k32 ? &S.k32->prof : nullptr, S.k16 ? &S.k16->prof : nullptr};
    for (Prof *p : ps)
        if (p) {
            p->on = (mask != 0);
            p->mask = (unsigned)mask;
            p->reset_stats();
        }
    return CM_OK;
}

int cm_system_kernel_stats(cm_system h, int category, long long *launches, double *total_ms, double *total_bytes)
{
    if (!h || category < 0 || category >= KC_N) return fail(CM_ERR_ARG, "bad category");
    System &S = h->sys;
    long long nl = 0;
    double ms = 0.0, by = 0.0;
    const Prof *ps[] = {&S.k64->prof, S.k32 ? &S.k32->prof : nullptr, S.k16 ? &S.k16->prof : nullptr};
    for (const Prof *p : ps)
        if (p) {
            nl += p->launches[category];
            ms += p->ms[category];
            by += p->bytes[category];
        }
    if (launches) *launches = nl;
    if (total_ms) *total_ms = ms;
    if (total_bytes) *total_bytes = by;
    return CM_OK;
}

static int dropin(int variant, int n, int a_symmetric, const int *Ai, const int *Aj, const double *Av,
                  const double *b, double tol, int NMAX, const int *Gi, const int *Gj, const double *Gv,
                  const double *IDv, const int *Hi, const int *Hj, const double *Hv, const int *Li, const int *Lj,
                  const double *Lv, const int *Ui, const int *Uj, const double *Uv, int nr, char mode, double *x,
                  double *resval, int *iter, int verbose, int *info)
{
    if (!b || !x || !resval || !iter || !info) return fail(CM_ERR_ARG, "null argument");
    if (cm_device_count() <= 0) return fail(CM_ERR_NODEV, "no GPU visible");
    System S;
    const HostSys hs{variant, n, a_symmetric, nr, mode, Ai, Aj, Av, Gi, Gj, Gv, IDv, Hi, Hj, Hv, Li, Lj, Lv, Ui, Uj, Uv};
    CM_TRY(system_create(S, hs));
    return dropin_solve(S, b, x, tol, NMAX, resval, iter, verbose, info);
}

int cm_afiim_epbicgstab(int n, double *Av, int *Aj, int *Ai, double *b, double tol, int NMAX, double *Gv, int *Gj,
                        int *Gi, double *IDv, double *Hv, int *Hj, int *Hi, double *x, double *resval, int *iter,
                        int verbose, int *info)
{
    return dropin(0, n, 0, Ai, Aj, Av, b, tol, NMAX, Gi, Gj, Gv, IDv, Hi, Hj, Hv, nullptr, nullptr, nullptr,
                  nullptr, nullptr, nullptr, 1, 'E', x, resval, iter, verbose, info);
}

int cm_pdfiim_epbicgstab(int n, int a_symmetric, const int *Ai, const int *Aj, const double *Av, const double *b,
                         double tol, int NMAX, const int *Gi, const int *Gj, const double *Gv, const double *IDv,
                         const int *Hi, const int *Hj, const double *Hv, double *x, double *resval, int *iter,
                         int verbose, int *info)
{
    return dropin(1, n, a_symmetric, Ai, Aj, Av, b, tol, NMAX, Gi, Gj, Gv, IDv, Hi, Hj, Hv, nullptr, nullptr,
                  nullptr, nullptr, nullptr, nullptr, 1, 'E', x, resval, iter, verbose, info);
}

int cm_acifim_pbicgstab(int n, double *Av, int *Aj, int *Ai, double *b, double tol, int NMAX, double *Gv, int *Gj,
                        int *Gi, double *IDv, double *Hv, int *Hj, int *Hi, double *Lv, int *Lj, int *Li,
                        double *Uv, int *Uj, int *Ui, int nr, char mode, double *x, double *resval, int *iter,
                        int verbose, int *info)
{
    return dropin(2, n, 0, Ai, Aj, Av, b, tol, NMAX, Gi, Gj, Gv, IDv, Hi, Hj, Hv, Li, Lj, Lv, Ui, Uj, Uv, nr, mode,
                  x, resval, iter, verbose, info);
}

}  // extern "C"
