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

    template <class Epi>
    void spmv(const DevCsr &m, const double *xin, Epi epi, bool use_state)
    {
        if (m.nblk == 0) return;
        spmv_kernel<double, double, Epi><<<m.nblk, BS, 0, stream>>>(
            m.ptr, m.col, m.val, nullptr, m.blk, xin, epi, use_state ? st : nullptr, partials);
    }

    // out = M in, per the reference's apply (ACIFIM/src/bicgstab.cpp:372-428);
    // for AFIIM / PDFIIM: out = G (IDv .* (H in)), H = G^T when symmetric.
    void enqueue_apply(const double *in, double *out, bool use_state)
    {
        const DevCsr &Hm = sym ? GT : H;
        if (variant != 2 || mode == 'E') {
            spmv(Hm, in, EpiScale<double, double>{tm, idv}, use_state);
            spmv(G, tm, EpiStore<double>{out}, use_state);
            return;
        }
        // mode C: t1 = H b; repeat: t1 += H (b - L t1); t2 = t1 .* IDv;
        //         x = G t2; repeat: x += G (t2 - U x)
        if (nr <= 1) {
            spmv(H, in, EpiScale<double, double>{t2, idv}, use_state);
        } else {
            spmv(H, in, EpiStore<double>{t1}, use_state);
            for (int it = 1; it < nr; it++) {
                spmv(L, t1, EpiSubFrom<double>{t3, in}, use_state);
                if (it + 1 < nr) spmv(H, t3, EpiAccum<double>{t1}, use_state);
                else spmv(H, t3, EpiAccumScale<double, double>{t1, t2, idv}, use_state);
            }
        }
        spmv(G, t2, EpiStore<double>{out}, use_state);
        for (int it = 1; it < nr; it++) {
            spmv(U, out, EpiSubFrom<double>{t3, t2}, use_state);
            spmv(G, t3, EpiAccum<double>{out}, use_state);
        }
    }

    void enqueue_iteration()
    {
        const int g = vec_grid(n);
        enqueue_apply(p, pc, true);
        spmv(A, pc, EpiAlpha{v, r0}, true);
        k_s_update<<<g, BS, 0, stream>>>(n, r, v, s, st);
        enqueue_apply(s, sc, true);
        spmv(A, sc, EpiOmega{t, s}, true);
        k_xr_update<<<g, BS, 0, stream>>>(n, x, pc, sc, r, s, t, r0, st, partials);
        k_p_update<<<g, BS, 0, stream>>>(n, p, r, v, st);
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
    S->spmv(S->A, S->x, EpiSetup{S->r, S->r0, S->p, S->b}, true);

    // chunks of iterations with one chunk in flight
    int chunk = 2, k = 0;
    auto enqueue_chunk = [&](int slot) -> int {
        for (int i = 0; i < chunk; i++) S->enqueue_iteration();
        CM_HIP(hipMemcpyAsync(&S->hst[1 + slot], S->st, sizeof(SolveState), hipMemcpyDeviceToHost, stream));
        CM_HIP(hipEventRecord(S->ev[slot], stream));
        chunk = std::min(chunk * 2, 32);
        return CM_OK;
    };
    int rc;
    if ((rc = enqueue_chunk(0))) return rc;
    int latest = 0;
    // every pass either exits or advances iter, so NMAX passes always end it
    const long long max_chunks = (long long)std::max(NMAX, 0) + 4;
    for (long long c = 0; c < max_chunks; c++) {
        const int cur = k & 1;
        if ((rc = enqueue_chunk(cur ^ 1))) return rc;
        latest = cur ^ 1;
        CM_HIP(hipEventSynchronize(S->ev[cur]));
        if (S->hst[1 + cur].done) break;
        k++;
    }
    CM_HIP(hipStreamSynchronize(stream));
    CM_HIP(hipGetLastError());
    // the latest copy is the final state (kernels after the exit are no-ops)
    const SolveState &fin = S->hst[1 + latest];
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
    S.spmv(*m, S.t1, EpiStore<double>{S.t2}, false);
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
    S.enqueue_apply(S.s, S.sc, false);
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
