// Host driver and C ABI of the BiCGSTAB solve engine (one GPU, gfx950).
//
// A solve is a device-resident state machine: the host uploads the system
// once, enqueues the setup kernel and then chunks of iterations on the
// system's own stream, and reads the 200-byte solve state back once per chunk
// (one chunk always in flight), so the GPU never waits on the host inside a
// chunk.  Every kernel returns immediately once the state says done, so
// enqueuing past the exit is harmless and the iteration at which the loop
// stops is the reference's, decided on the device.
//
// Per BiCGSTAB pass (mode E / AFIIM / PDFIIM), 9 kernels:
//   tm = D^{-1}(H p)   pc = G tm   v = A pc (+ r0.v -> alpha)   s = r - alpha v
//   tm = D^{-1}(H s)   sc = G tm   t = A sc (+ t.s, t.t -> omega)
//   x, r update (+ ||pc||, ||x||, ||r||, r0.r -> exit tests, beta)
//   p update
// (AFIIM/src/bicgstab.cpp:168-275, PDFIIM/src/bicgstab.cpp:186-293,
//  ACIFIM/src/bicgstab.cpp:197-304).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "cm_kernels.hpp"
#include "cm_solve.h"

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
static int dev_upload(T **d, const T *h, size_t count, hipStream_t s)
{
    *d = nullptr;
    if (count == 0) count = 1;
    CM_HIP(hipMalloc((void **)d, count * sizeof(T)));
    if (h) CM_HIP(hipMemcpyAsync(*d, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    return CM_OK;
}

// A CSR matrix in HBM with its SpMV row blocks.
struct DevCsr {
    int n = 0;
    long long nnz = 0;
    int nblk = 0;
    int *ptr = nullptr, *col = nullptr, *blk = nullptr;
    double *val = nullptr;
    bool present() const { return ptr != nullptr; }
    void release()
    {
        (void)hipFree(ptr);
        (void)hipFree(col);
        (void)hipFree(blk);
        (void)hipFree(val);
        ptr = col = blk = nullptr;
        val = nullptr;
    }
};

static int upload_csr(DevCsr &m, int n, const int *ptr, const int *col, const double *val,
                      hipStream_t s)
{
    // The kernels index x by the stored columns and stage rows by ptr: reject
    // anything that would read outside the arrays.
    if (!ptr || ptr[0] != 0) return fail(CM_ERR_ARG, "CSR row pointer must start at 0");
    for (int i = 0; i < n; i++)
        if (ptr[i + 1] < ptr[i]) return fail(CM_ERR_ARG, "CSR row pointer not monotone");
    for (long long k = 0; k < ptr[n]; k++)
        if (col[k] < 0 || col[k] >= n) return fail(CM_ERR_ARG, "CSR column index out of range");
    m.n = n;
    m.nnz = ptr[n];
    std::vector<int> blk = make_row_blocks(n, ptr);
    m.nblk = (int)blk.size() - 1;
    int rc;
    if ((rc = dev_upload(&m.ptr, ptr, (size_t)n + 1, s))) return rc;
    if ((rc = dev_upload(&m.col, col, (size_t)m.nnz, s))) return rc;
    if ((rc = dev_upload(&m.val, val, (size_t)m.nnz, s))) return rc;
    if ((rc = dev_upload(&m.blk, blk.data(), blk.size(), s))) return rc;
    // blk is a stack vector: finish the copy before it goes away
    CM_HIP(hipStreamSynchronize(s));
    return CM_OK;
}

// CSR of the transpose (host).
static void host_transpose(int n, const int *ptr, const int *col, const double *val,
                           std::vector<int> &tp, std::vector<int> &tc, std::vector<double> &tv)
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
static void host_sym_expand(int n, const int *ptr, const int *col, const double *val,
                            std::vector<int> &fp, std::vector<int> &fc, std::vector<double> &fv)
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

static inline int vec_grid(int n)
{
    int g = (n + BS - 1) / BS;
    return std::max(1, std::min(g, MAX_RED_BLOCKS));
}

// Kernel categories of the per-kernel timing (cm_system_kernel_stats).
enum KernelCat { KC_A = 0, KC_F1 = 1, KC_F2 = 2, KC_LU = 3, KC_VEC = 4, KC_SETUP = 5, KC_FIN = 6, KC_N = 7 };

// Optional per-launch timing: an event pair around every launch of a solve,
// tagged with its category, the BiCGSTAB pass it belongs to and its
// algorithmic bytes.  Launches the device skipped (after the exit) are
// dropped when the solve ends, so the averages cover executed kernels only.
struct Prof {
    bool on = false;
    unsigned mask = 0;    // categories whose launches are bracketed
    int pass = -1;        // pass being enqueued (-1: setup)
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

struct System {
    int variant = 0, n = 0, nr = 1;
    char mode = 'E';
    bool sym = false;
    hipStream_t stream = nullptr;
    DevCsr A, G, H, L, U, GT;
    double *idv = nullptr;
    // work vectors
    double *b = nullptr, *x = nullptr, *r0 = nullptr, *r = nullptr, *p = nullptr, *v = nullptr,
           *s = nullptr, *t = nullptr, *pc = nullptr, *sc = nullptr, *tm = nullptr, *t1 = nullptr,
           *t2 = nullptr, *t3 = nullptr;
    SolveState *st = nullptr;   // device
    SolveState *hst = nullptr;  // pinned host, 3 slots (init + 2 polling slots)
    double *partials = nullptr;
    double *trace = nullptr;
    int trace_cap_dev = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    Prof prof;

    ~System()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        A.release();
        G.release();
        H.release();
        L.release();
        U.release();
        GT.release();
        double *bufs[] = {idv, b, x, r0, r, p, v, s, t, pc, sc, tm, t1, t2, t3, partials, trace};
        for (double *q : bufs) (void)hipFree(q);
        (void)hipFree(st);
        if (hst) (void)hipHostFree(hst);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    const DevCsr *mat(char which) const
    {
        switch (which) {
        case 'A': return &A;
        case 'G': return &G;
        case 'H': return sym ? &GT : &H;
        case 'L': return &L;
        case 'U': return &U;
        case 'T': return &GT;
        default: return nullptr;
        }
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
    // pointers and row blocks once, x once, y written once, plus the
    // epilogue's extra n-vectors.
    template <class Epi>
    double spmv_bytes(const DevCsr &m) const
    {
        return (double)m.nnz * (sizeof(double) + sizeof(int)) + 4.0 * (m.n + 1) + 4.0 * (m.nblk + 1) +
               8.0 * m.n * (2 + Epi::XV);
    }

    template <class Epi>
    void spmv(const DevCsr &m, const double *xin, Epi epi, bool use_state, int cat = KC_A)
    {
        if (m.nblk == 0) return;
        timed(cat, spmv_bytes<Epi>(m), [&] {
            spmv_kernel<double, double, Epi><<<m.nblk, BS, 0, stream>>>(
                m.ptr, m.col, m.val, nullptr, m.blk, xin, epi, use_state ? st : nullptr, partials);
        });
        if constexpr (Epi::K > 0)
            timed(KC_FIN, 8.0 * Epi::K * m.nblk, [&] {
                k_finalize<Epi::K, Epi><<<1, FBS, 0, stream>>>(m.nblk, partials, st, epi);
            });
    }

    // out = M in, per the reference's apply (ACIFIM/src/bicgstab.cpp:372-428);
    // for AFIIM / PDFIIM: out = G (IDv .* (H in)), H = G^T when symmetric.
    void enqueue_apply(const double *in, double *out, bool use_state)
    {
        const DevCsr &Hm = sym ? GT : H;
        if (variant != 2 || mode == 'E') {
            spmv(Hm, in, EpiScale<double, double>{tm, idv}, use_state, KC_F1);
            spmv(G, tm, EpiStore<double>{out}, use_state, KC_F2);
            return;
        }
        // mode C: t1 = H b; repeat: t1 += H (b - L t1); t2 = t1 .* IDv;
        //         x = G t2; repeat: x += G (t2 - U x)
        if (nr <= 1) {
            spmv(H, in, EpiScale<double, double>{t2, idv}, use_state, KC_F1);
        } else {
            spmv(H, in, EpiStore<double>{t1}, use_state, KC_F1);
            for (int it = 1; it < nr; it++) {
                spmv(L, t1, EpiSubFrom<double>{t3, in}, use_state, KC_LU);
                if (it + 1 < nr) spmv(H, t3, EpiAccum<double>{t1}, use_state, KC_F1);
                else spmv(H, t3, EpiAccumScale<double, double>{t1, t2, idv}, use_state, KC_F1);
            }
        }
        spmv(G, t2, EpiStore<double>{out}, use_state, KC_F2);
        for (int it = 1; it < nr; it++) {
            spmv(U, out, EpiSubFrom<double>{t3, t2}, use_state, KC_LU);
            spmv(G, t3, EpiAccum<double>{out}, use_state, KC_F2);
        }
    }

    void enqueue_iteration()
    {
        const int g = vec_grid(n);
        const double nn = (double)n;
        prof.pass++;
        enqueue_apply(p, pc, true);
        spmv(A, pc, EpiAlpha{v, r0}, true, KC_A);
        timed(KC_VEC, 24.0 * nn, [&] { k_s_update<<<g, BS, 0, stream>>>(n, r, v, s, st); });
        enqueue_apply(s, sc, true);
        spmv(A, sc, EpiOmega{t, s}, true, KC_A);
        timed(KC_VEC, 72.0 * nn, [&] {
            k_xr_update<<<g, BS, 0, stream>>>(n, x, pc, sc, r, s, t, r0, st, partials);
        });
        timed(KC_FIN, 32.0 * g, [&] { k_finalize<4, FinPass><<<1, FBS, 0, stream>>>(g, partials, st, FinPass{}); });
        prof.p_update = true;
        timed(KC_VEC, 32.0 * nn, [&] { k_p_update<<<g, BS, 0, stream>>>(n, p, r, v, st); });
        prof.p_update = false;
    }
};

static int alloc_vec(double **d, int n)
{
    CM_HIP(hipMalloc((void **)d, (size_t)std::max(n, 1) * sizeof(double)));
    return CM_OK;
}

static int system_create(System *S, int variant, int n, int a_symmetric, const int *Ai,
                         const int *Aj, const double *Av, const int *Gi, const int *Gj,
                         const double *Gv, const double *IDv, const int *Hi, const int *Hj,
                         const double *Hv, const int *Li, const int *Lj, const double *Lv,
                         const int *Ui, const int *Uj, const double *Uv, int nr, char mode)
{
    if (n <= 0 || !Ai || !Aj || !Av || !Gi || !Gj || !Gv || !IDv) return fail(CM_ERR_ARG, "null or empty system");
    if (variant < 0 || variant > 2) return fail(CM_ERR_ARG, "bad variant");
    S->variant = variant;
    S->n = n;
    S->nr = nr;
    S->mode = (variant == 2) ? mode : 'E';
    S->sym = (variant == 1 && a_symmetric);
    if (variant == 2 && mode != 'E' && mode != 'C' && mode != 'I' && mode != 'J')
        return fail(CM_ERR_ARG, "unknown apply mode");
    if (variant == 2 && (mode == 'I' || mode == 'J'))
        return fail(CM_ERR_UNSUPPORTED, "apply modes I/J (triangular solves) are not in this build");
    if (!S->sym && (!Hi || !Hj || !Hv)) return fail(CM_ERR_ARG, "H factor missing");
    if (variant == 2 && mode == 'C' && nr > 1 && (!Li || !Lj || !Lv || !Ui || !Uj || !Uv))
        return fail(CM_ERR_ARG, "L/U factors missing for mode C");
    CM_HIP(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
    hipStream_t st = S->stream;
    int rc;
    if (S->sym) {
        std::vector<int> fp, fc;
        std::vector<double> fv;
        host_sym_expand(n, Ai, Aj, Av, fp, fc, fv);
        if ((rc = upload_csr(S->A, n, fp.data(), fc.data(), fv.data(), st))) return rc;
        std::vector<int> tp, tc;
        std::vector<double> tv;
        host_transpose(n, Gi, Gj, Gv, tp, tc, tv);
        if ((rc = upload_csr(S->GT, n, tp.data(), tc.data(), tv.data(), st))) return rc;
    } else {
        if ((rc = upload_csr(S->A, n, Ai, Aj, Av, st))) return rc;
        if ((rc = upload_csr(S->H, n, Hi, Hj, Hv, st))) return rc;
    }
    if ((rc = upload_csr(S->G, n, Gi, Gj, Gv, st))) return rc;
    if (variant == 2 && Li && Lj && Lv && Ui && Uj && Uv) {
        if ((rc = upload_csr(S->L, n, Li, Lj, Lv, st))) return rc;
        if ((rc = upload_csr(S->U, n, Ui, Uj, Uv, st))) return rc;
    }
    if ((rc = dev_upload(&S->idv, IDv, (size_t)n, st))) return rc;
    double **vecs[] = {&S->b, &S->x, &S->r0, &S->r, &S->p, &S->v, &S->s, &S->t,
                       &S->pc, &S->sc, &S->tm, &S->t1, &S->t2, &S->t3};
    for (double **q : vecs)
        if ((rc = alloc_vec(q, n))) return rc;
    CM_HIP(hipMemsetAsync(S->x, 0, (size_t)n * sizeof(double), st));
    CM_HIP(hipMemsetAsync(S->b, 0, (size_t)n * sizeof(double), st));
    CM_HIP(hipMalloc((void **)&S->st, sizeof(SolveState)));
    CM_HIP(hipHostMalloc((void **)&S->hst, 3 * sizeof(SolveState), hipHostMallocDefault));
    const int maxblk = std::max({S->A.nblk, S->G.nblk, MAX_RED_BLOCKS});
    CM_HIP(hipMalloc((void **)&S->partials, (size_t)maxblk * 4 * sizeof(double)));
    for (auto &e : S->ev) CM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    CM_HIP(hipStreamSynchronize(st));
    return CM_OK;
}

// maxmsteps per package (AFIIM/src/bicgstab.cpp:114, PDFIIM/src/bicgstab.cpp:106,
// ACIFIM/src/bicgstab.cpp:132).
static int max_div_steps(int variant, int n)
{
    if (variant == 2) return std::min(std::max(n, 2000) / 50, 40);
    return std::min(std::max(n, 500) / 50, 20);
}

static int system_solve(System *S, double tol, int NMAX, double *resval, int *iter, int *info,
                        double *trace, int trace_cap, int *ntrace)
{
    hipStream_t stream = S->stream;
    if (trace && trace_cap > 0 && trace_cap > S->trace_cap_dev) {
        (void)hipFree(S->trace);
        S->trace = nullptr;
        CM_HIP(hipMalloc((void **)&S->trace, (size_t)trace_cap * sizeof(double)));
        S->trace_cap_dev = trace_cap;
    }
    SolveState init;
    std::memset(&init, 0, sizeof init);
    init.tol = tol;
    init.eps = 1.11022302462516e-16;  // AFIIM/src/bicgstab.cpp:108
    init.nmax = NMAX;
    init.maxmsteps = max_div_steps(S->variant, S->n);
    init.variant = S->variant;
    init.info = info ? *info : 0;
    init.trace = (trace && trace_cap > 0) ? S->trace : nullptr;
    init.trace_cap = (trace && trace_cap > 0) ? trace_cap : 0;
    S->hst[0] = init;
    CM_HIP(hipMemcpyAsync(S->st, &S->hst[0], sizeof(SolveState), hipMemcpyHostToDevice, stream));
    S->prof.pass = -1;
    S->spmv(S->A, S->x, EpiSetup{S->r, S->r0, S->p, S->b}, true, KC_SETUP);

    // Passes are enqueued in chunks, at most two chunks outstanding, and the
    // state is read back at the end of each chunk.  The next chunk's size is
    // the number of passes the residual's recent contraction rate says are
    // still needed (at least 4, at most 32), so few no-op passes are
    // enqueued past the exit.
    struct Outstanding {
        int slot;
        long long upto;  // passes enqueued up to the end of this chunk
    };
    Outstanding q[2];
    int nq = 0, slot = 0, last_chunk = 4;
    long long enq = 0;
    auto push = [&](int size) -> int {
        for (int i = 0; i < size; i++) S->enqueue_iteration();
        CM_HIP(hipMemcpyAsync(&S->hst[1 + slot], S->st, sizeof(SolveState), hipMemcpyDeviceToHost, stream));
        CM_HIP(hipEventRecord(S->ev[slot], stream));
        enq += size;
        q[nq++] = Outstanding{slot, enq};
        slot ^= 1;
        last_chunk = size;
        return CM_OK;
    };
    int rc;
    const long long nmax_passes = std::max(NMAX, 0);
    if ((rc = push((int)std::min<long long>(4, std::max<long long>(nmax_passes, 1))))) return rc;
    SolveState fin_state;
    bool have_prev = false;
    int prev_iter = 0;
    double prev_rc = 0.0;
    for (;;) {
        if (nq < 2 && enq < nmax_passes) {
            long long want = std::min(2 * last_chunk, 32);
            if (have_prev) {
                // estimate from the last two snapshots
                const SolveState &sn = fin_state;
                if (sn.iter > prev_iter && sn.rc < prev_rc && sn.rc > 0.0 && prev_rc > 0.0) {
                    const double rate = std::pow(sn.rc / prev_rc, 1.0 / (sn.iter - prev_iter));
                    if (rate < 1.0) {
                        const double need = std::ceil(std::log(sn.tol * sn.ro / sn.rc) / std::log(rate)) + 1.0;
                        const long long ahead = enq - sn.iter;
                        want = (long long)std::max(4.0, std::min(32.0, need - (double)ahead));
                    }
                }
            }
            want = std::min(want, nmax_passes - enq);
            if (want > 0 && (rc = push((int)want))) return rc;
        }
        if (nq == 0) return fail(CM_ERR_HIP, "solve did not reach an exit");
        const Outstanding o = q[0];
        q[0] = q[1];
        nq--;
        CM_HIP(hipEventSynchronize(S->ev[o.slot]));
        const SolveState snap = S->hst[1 + o.slot];
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
    // the state is final once done is set (later kernels are no-ops)
    const SolveState &fin = fin_state;
    if (S->prof.on) S->prof.collect(fin.ntrace);
    if (!fin.done) return fail(CM_ERR_HIP, "solve state did not reach done");
    *iter = fin.iter;
    if (info) *info = fin.info;
    *resval = fin.rc / fin.ro;
    if (ntrace) *ntrace = fin.ntrace;
    if (trace && trace_cap > 0) {
        const int cnt = std::min(fin.ntrace, trace_cap);
        if (cnt > 0) CM_HIP(hipMemcpy(trace, S->trace, (size_t)cnt * sizeof(double), hipMemcpyDeviceToHost));
    }
    return CM_OK;
}

static void print_verbose(const System *S, int iter, int info, double rel, int NMAX)
{
    (void)NMAX;
    if (info == 0) {
        std::cout << "Convergence in " << iter << " iterations with relative residual " << rel << "\n";
    } else if (info == 1) {
        std::cout << "BiCGSTAB diverged and iterations stopped with relative residual " << rel
                  << " in iteration " << iter << "\n";
        std::cout << "Try to increase the quality of preconditioning\n";
    } else if (info == 2) {
        if (S->variant == 1) {
            std::cout << "BiCGSTAB diverged and iterations stopped with relative residual " << rel
                      << " in iteration " << iter << "\n";
            std::cout << "Try to increase the quality of preconditioning\n";
        } else {
            std::cout << "BiCGSTAB stagnated with relative residual " << rel << " in iteration " << iter
                      << "\n";
            std::cout << "Try increasing prescribed tolerance\n";
        }
    } else if (info == 3) {
        std::cout << "Convergence failed after " << iter << " iterations with relative residual " << rel
                  << "\n";
    }
}

static int dropin_solve(System &S, int n, const double *b, double *x, double tol, int NMAX,
                        double *resval, int *iter, int verbose, int *info)
{
    CM_HIP(hipMemcpyAsync(S.b, b, (size_t)n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    CM_HIP(hipMemcpyAsync(S.x, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    int rc = system_solve(&S, tol, NMAX, resval, iter, info, nullptr, 0, nullptr);
    if (rc) return rc;
    CM_HIP(hipMemcpy(x, S.x, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    // the reference prints only when verbose (AFIIM/src/bicgstab.cpp:227-281);
    // ACIFIM does not print the non-finite-alpha exit
    if (verbose && info) print_verbose(&S, *iter, *info, *resval, NMAX);
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

int cm_system_create(cm_system *out, int variant, int n, int a_symmetric, const int *Ai,
                     const int *Aj, const double *Av, const int *Gi, const int *Gj,
                     const double *Gv, const double *IDv, const int *Hi, const int *Hj,
                     const double *Hv, const int *Li, const int *Lj, const double *Lv,
                     const int *Ui, const int *Uj, const double *Uv, int nr, char mode)
{
    if (!out) return fail(CM_ERR_ARG, "null handle");
    *out = nullptr;
    if (cm_device_count() <= 0) return fail(CM_ERR_NODEV, "no GPU visible");
    auto h = std::make_unique<cm_system_s>();
    int rc = system_create(&h->sys, variant, n, a_symmetric, Ai, Aj, Av, Gi, Gj, Gv, IDv, Hi, Hj, Hv,
                           Li, Lj, Lv, Ui, Uj, Uv, nr, mode);
    if (rc) return rc;
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
    CM_HIP(hipMemcpyAsync(S.b, b, (size_t)S.n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    CM_HIP(hipStreamSynchronize(S.stream));
    return CM_OK;
}

int cm_system_set_x(cm_system h, const double *x)
{
    if (!h) return fail(CM_ERR_ARG, "null handle");
    System &S = h->sys;
    if (x) CM_HIP(hipMemcpyAsync(S.x, x, (size_t)S.n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    else CM_HIP(hipMemsetAsync(S.x, 0, (size_t)S.n * sizeof(double), S.stream));
    CM_HIP(hipStreamSynchronize(S.stream));
    return CM_OK;
}

int cm_system_get_x(cm_system h, double *x)
{
    if (!h || !x) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    CM_HIP(hipStreamSynchronize(S.stream));
    CM_HIP(hipMemcpy(x, S.x, (size_t)S.n * sizeof(double), hipMemcpyDeviceToHost));
    return CM_OK;
}

int cm_system_solve(cm_system h, double tol, int NMAX, double *resval, int *iter, int *info,
                    double *trace, int trace_cap, int *ntrace)
{
    if (!h || !resval || !iter) return fail(CM_ERR_ARG, "null argument");
    return system_solve(&h->sys, tol, NMAX, resval, iter, info, trace, trace_cap, ntrace);
}

int cm_system_spmv(cm_system h, char which, const double *x, double *y)
{
    if (!h || !x || !y) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    const DevCsr *m = S.mat(which);
    if (!m || !m->present()) return fail(CM_ERR_ARG, "matrix not present");
    CM_HIP(hipMemcpyAsync(S.t1, x, (size_t)S.n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    const bool prof_on = S.prof.on;  // kernel-level calls are not part of a solve's timing
    S.prof.on = false;
    S.spmv(*m, S.t1, EpiStore<double>{S.t2}, false);
    S.prof.on = prof_on;
    CM_HIP(hipMemcpyAsync(y, S.t2, (size_t)S.n * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    CM_HIP(hipStreamSynchronize(S.stream));
    CM_HIP(hipGetLastError());
    return CM_OK;
}

int cm_system_apply(cm_system h, const double *v, double *out)
{
    if (!h || !v || !out) return fail(CM_ERR_ARG, "null argument");
    System &S = h->sys;
    CM_HIP(hipMemcpyAsync(S.s, v, (size_t)S.n * sizeof(double), hipMemcpyHostToDevice, S.stream));
    const bool prof_on = S.prof.on;
    S.prof.on = false;
    S.enqueue_apply(S.s, S.sc, false);
    S.prof.on = prof_on;
    CM_HIP(hipMemcpyAsync(out, S.sc, (size_t)S.n * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    CM_HIP(hipStreamSynchronize(S.stream));
    CM_HIP(hipGetLastError());
    return CM_OK;
}

long long cm_system_nnz(cm_system h, char which)
{
    if (!h) return -1;
    const DevCsr *m = h->sys.mat(which);
    return (m && m->present()) ? m->nnz : -1;
}

int cm_system_profile(cm_system h, int mask)
{
    if (!h) return fail(CM_ERR_ARG, "null handle");
    h->sys.prof.on = (mask != 0);
    h->sys.prof.mask = (unsigned)mask;
    h->sys.prof.reset_stats();
    return CM_OK;
}

int cm_system_kernel_stats(cm_system h, int category, long long *launches, double *total_ms,
                           double *total_bytes)
{
    if (!h || category < 0 || category >= KC_N) return fail(CM_ERR_ARG, "bad category");
    const Prof &p = h->sys.prof;
    if (launches) *launches = p.launches[category];
    if (total_ms) *total_ms = p.ms[category];
    if (total_bytes) *total_bytes = p.bytes[category];
    return CM_OK;
}

int cm_afiim_epbicgstab(int n, double *Av, int *Aj, int *Ai, double *b, double tol, int NMAX,
                        double *Gv, int *Gj, int *Gi, double *IDv, double *Hv, int *Hj, int *Hi,
                        double *x, double *resval, int *iter, int verbose, int *info)
{
    if (!b || !x || !resval || !iter || !info) return fail(CM_ERR_ARG, "null argument");
    if (cm_device_count() <= 0) return fail(CM_ERR_NODEV, "no GPU visible");
    System S;
    int rc = system_create(&S, 0, n, 0, Ai, Aj, Av, Gi, Gj, Gv, IDv, Hi, Hj, Hv, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr, 1, 'E');
    if (rc) return rc;
    return dropin_solve(S, n, b, x, tol, NMAX, resval, iter, verbose, info);
}

int cm_pdfiim_epbicgstab(int n, int a_symmetric, const int *Ai, const int *Aj, const double *Av,
                         const double *b, double tol, int NMAX, const int *Gi, const int *Gj,
                         const double *Gv, const double *IDv, const int *Hi, const int *Hj,
                         const double *Hv, double *x, double *resval, int *iter, int verbose,
                         int *info)
{
    if (!b || !x || !resval || !iter || !info) return fail(CM_ERR_ARG, "null argument");
    if (cm_device_count() <= 0) return fail(CM_ERR_NODEV, "no GPU visible");
    System S;
    int rc = system_create(&S, 1, n, a_symmetric, Ai, Aj, Av, Gi, Gj, Gv, IDv, Hi, Hj, Hv, nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr, 1, 'E');
    if (rc) return rc;
    return dropin_solve(S, n, b, x, tol, NMAX, resval, iter, verbose, info);
}

int cm_acifim_pbicgstab(int n, double *Av, int *Aj, int *Ai, double *b, double tol, int NMAX,
                        double *Gv, int *Gj, int *Gi, double *IDv, double *Hv, int *Hj, int *Hi,
                        double *Lv, int *Lj, int *Li, double *Uv, int *Uj, int *Ui, int nr,
                        char mode, double *x, double *resval, int *iter, int verbose, int *info)
{
    if (!b || !x || !resval || !iter || !info) return fail(CM_ERR_ARG, "null argument");
    if (cm_device_count() <= 0) return fail(CM_ERR_NODEV, "no GPU visible");
    System S;
    int rc = system_create(&S, 2, n, 0, Ai, Aj, Av, Gi, Gj, Gv, IDv, Hi, Hj, Hv, Li, Lj, Lv, Ui, Uj,
                           Uv, nr, mode);
    if (rc) return rc;
    return dropin_solve(S, n, b, x, tol, NMAX, resval, iter, verbose, info);
}

}  // extern "C"
