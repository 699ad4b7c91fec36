/*
 * TEST INFRASTRUCTURE ONLY -- not part of the product.
 *
 * CPU restatement (plain C, FP64) of the reference's preconditioned BiCGSTAB
 * solve phase.  Only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark load this file's library, and only as the checker / the timed
 * CPU baseline.  The GPU product never links it.
 *
 * What it restates (reference files, relative to the reference root):
 *   or_csr_mv          mkl_sparse_d_mv(NON_TRANSPOSE, 1, A, GENERAL, x, 0, y)
 *                      AFIIM/src/bicgstab.cpp:151,172,175,178
 *   or_csr_mv_sym_upper  SYMMETRIC/UPPER descriptor (A stored as upper
 *                      triangle)  PDFIIM/src/bicgstab.cpp:146-153,169
 *   or_csr_mv_trans    SPARSE_OPERATION_TRANSPOSE on G
 *                      PDFIIM/src/bicgstab.cpp:190,207
 *   or_trsv_lower / or_trsv_upper   mkl_sparse_d_trsv, TRIANGULAR, NON_UNIT
 *                      ACIFIM/src/bicgstab.cpp:406,422,425-426
 *   or_apply           ACIFIM apply() modes E, C, I, J
 *                      ACIFIM/src/bicgstab.cpp:372-428
 *   or_bicgstab        EPBiCGSTAB / PBiCGSTAB control flow, constants and exit
 *                      semantics: AFIIM/src/bicgstab.cpp:87-297,
 *                      PDFIIM/src/bicgstab.cpp:80-310,
 *                      ACIFIM/src/bicgstab.cpp:97-327
 *
 * Parity pins (tests/test_oracle.py): SpMV against scipy.sparse in extended
 * precision; the apply modes against dense numpy products / closed-form
 * Richardson sums / scipy triangular solves; the Krylov recurrence against an
 * independent textbook right-preconditioned BiCGSTAB (van der Vorst 1992)
 * written with numpy; the exit codes against hand-derived forced cases.
 *
 * Reductions use a static contiguous split over `nthreads` OpenMP threads and
 * sum the per-thread partials in thread order, so a result is reproducible for
 * a fixed thread count (1 thread == strictly sequential order).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define OR_MAXT 256

static int g_threads = 1;

void or_set_threads(int t)
{
    if (t < 1) t = 1;
    if (t > OR_MAXT) t = OR_MAXT;
    g_threads = t;
}

int or_get_threads(void) { return g_threads; }

/* y = A x for a general CSR matrix (every stored entry is used). */
void or_csr_mv(int n, const int *Ai, const int *Aj, const double *Av,
               const double *x, double *y)
{
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int k = Ai[i]; k < Ai[i + 1]; k++) s += Av[k] * x[Aj[k]];
        y[i] = s;
    }
}

/* y = A x where only the upper triangle (including the diagonal) of a
 * symmetric A is stored.  Sequential scatter (the transpose half). */
void or_csr_mv_sym_upper(int n, const int *Ai, const int *Aj, const double *Av,
                         const double *x, double *y)
{
    for (int i = 0; i < n; i++) y[i] = 0.0;
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int k = Ai[i]; k < Ai[i + 1]; k++) {
            int j = Aj[k];
            if (j < i) continue; /* MKL reads only the declared triangle */
            s += Av[k] * x[j];
            if (j != i) y[j] += Av[k] * x[i];
        }
        y[i] += s;
    }
}

/* y = A^T x (scatter, sequential). */
void or_csr_mv_trans(int n, const int *Ai, const int *Aj, const double *Av,
                     const double *x, double *y)
{
    for (int i = 0; i < n; i++) y[i] = 0.0;
    for (int i = 0; i < n; i++)
        for (int k = Ai[i]; k < Ai[i + 1]; k++) y[Aj[k]] += Av[k] * x[i];
}

/* Solve L x = b, L lower triangular in CSR, stored (non-unit) diagonal. */
void or_trsv_lower(int n, const int *Li, const int *Lj, const double *Lv,
                   const double *b, double *x)
{
    for (int i = 0; i < n; i++) {
        double s = b[i], d = 0.0;
        for (int k = Li[i]; k < Li[i + 1]; k++) {
            int j = Lj[k];
            if (j < i) s -= Lv[k] * x[j];
            else if (j == i) d = Lv[k];
        }
        x[i] = s / d;
    }
}

/* Solve U x = b, U upper triangular in CSR, stored (non-unit) diagonal. */
void or_trsv_upper(int n, const int *Ui, const int *Uj, const double *Uv,
                   const double *b, double *x)
{
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i], d = 0.0;
        for (int k = Ui[i]; k < Ui[i + 1]; k++) {
            int j = Uj[k];
            if (j > i) s -= Uv[k] * x[j];
            else if (j == i) d = Uv[k];
        }
        x[i] = s / d;
    }
}

/* Deterministic parallel dot product: static split, partials in thread order. */
static double or_dot_impl(int n, const double *a, const double *b)
{
    double part[OR_MAXT];
    int T = g_threads;
    if (T > n) T = n > 0 ? n : 1;
#pragma omp parallel num_threads(T)
    {
        int t = omp_get_thread_num(), nt = omp_get_num_threads();
        long lo = (long)n * t / nt, hi = (long)n * (t + 1) / nt;
        double s = 0.0;
        for (long i = lo; i < hi; i++) s += a[i] * b[i];
        part[t] = s;
        if (t == 0) T = nt;
    }
    double s = 0.0;
    for (int t = 0; t < T; t++) s += part[t];
    return s;
}

double or_dot(int n, const double *a, const double *b) { return or_dot_impl(n, a, b); }
double or_nrm2(int n, const double *a) { return sqrt(or_dot_impl(n, a, a)); }

/* z = a .* b elementwise (vdMul). */
void or_vmul(int n, const double *a, const double *b, double *z)
{
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int i = 0; i < n; i++) z[i] = a[i] * b[i];
}

/* Preconditioner application, ACIFIM/src/bicgstab.cpp:372-428.
 * work holds 3n doubles (t1, t2, t3). */
void or_apply(int n, int nr, char mode,
              const int *Gi, const int *Gj, const double *Gv, const double *IDv,
              const int *Hi, const int *Hj, const double *Hv,
              const int *Li, const int *Lj, const double *Lv,
              const int *Ui, const int *Uj, const double *Uv,
              const double *b, double *x, double *work)
{
    double *t1 = work, *t2 = work + n, *t3 = work + 2 * (long)n;
    if (mode == 'J') {
        or_trsv_lower(n, Li, Lj, Lv, b, t1);
        or_trsv_upper(n, Ui, Uj, Uv, t1, x);
        return;
    }
    if (mode == 'E') {
        or_csr_mv(n, Hi, Hj, Hv, b, t1);
    } else if (mode == 'C') {
        or_csr_mv(n, Hi, Hj, Hv, b, t1);
        for (int it = 1; it < nr; it++) {
            or_csr_mv(n, Li, Lj, Lv, t1, t2);
#pragma omp parallel for schedule(static) num_threads(g_threads)
            for (int i = 0; i < n; i++) t3[i] = b[i] - t2[i];
            /* t1 = H t3 + 1.0 * t1 */
#pragma omp parallel for schedule(static) num_threads(g_threads)
            for (int i = 0; i < n; i++) {
                double s = 0.0;
                for (int k = Hi[i]; k < Hi[i + 1]; k++) s += Hv[k] * t3[Hj[k]];
                t1[i] = s + t1[i];
            }
        }
    } else {
        or_trsv_lower(n, Li, Lj, Lv, b, t1);
    }
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int i = 0; i < n; i++) t2[i] = t1[i] * IDv[i];
    if (mode == 'E') {
        or_csr_mv(n, Gi, Gj, Gv, t2, x);
    } else if (mode == 'C') {
        or_csr_mv(n, Gi, Gj, Gv, t2, x);
        for (int it = 1; it < nr; it++) {
            or_csr_mv(n, Ui, Uj, Uv, x, t1);
#pragma omp parallel for schedule(static) num_threads(g_threads)
            for (int i = 0; i < n; i++) t3[i] = t2[i] - t1[i];
#pragma omp parallel for schedule(static) num_threads(g_threads)
            for (int i = 0; i < n; i++) {
                double s = 0.0;
                for (int k = Gi[i]; k < Gi[i + 1]; k++) s += Gv[k] * t3[Gj[k]];
                x[i] = s + x[i];
            }
        }
    } else {
        or_trsv_upper(n, Ui, Uj, Uv, t2, x);
    }
}

/*
 * variant: 0 = AFIIM EPBiCGSTAB, 1 = PDFIIM EPBiCGSTAB, 2 = ACIFIM PBiCGSTAB.
 * issym (PDFIIM only): A holds the upper triangle of a symmetric matrix and
 *   M v = G (IDv .* (G^T v)); H is ignored.
 * trace (may be NULL): trace[k] = rc computed in pass k (the lagged ||r_k||),
 *   for k < trace_cap.  Returns the number of passes that computed rc.
 */
int or_bicgstab(int variant, int issym, int n,
                const int *Ai, const int *Aj, const double *Av,
                const double *b, double tol, int NMAX,
                const int *Gi, const int *Gj, const double *Gv, const double *IDv,
                const int *Hi, const int *Hj, const double *Hv,
                const int *Li, const int *Lj, const double *Lv,
                const int *Ui, const int *Uj, const double *Uv,
                int nr, char mode,
                double *x, double *resval, int *iter, int *info,
                double *trace, int trace_cap)
{
    const double eps = 1.11022302462516e-16;
    const int maxstagsteps = 3;
    int maxmsteps;
    if (variant == 2)
        maxmsteps = ((n > 2000) ? n : 2000) / 50 < 40 ? ((n > 2000) ? n : 2000) / 50 : 40;
    else
        maxmsteps = ((n > 500) ? n : 500) / 50 < 20 ? ((n > 500) ? n : 500) / 50 : 20;
    int stag = 0, msteps = 0, ntrace = 0;
    double prc, ro, a, om, m1, m2, rc, beta;
    size_t nb = (size_t)n * sizeof(double);
    double *p = malloc(nb), *r0 = malloc(nb), *r = malloc(nb), *v = malloc(nb),
           *s = malloc(nb), *t = malloc(nb), *pc = malloc(nb), *sc = malloc(nb),
           *tm = malloc(nb), *work = malloc(3 * nb);

#define OR_MV(out, in)                                                        \
    do {                                                                      \
        if (issym) or_csr_mv_sym_upper(n, Ai, Aj, Av, in, out);               \
        else or_csr_mv(n, Ai, Aj, Av, in, out);                               \
    } while (0)
#define OR_PREC(in, out)                                                      \
    do {                                                                      \
        if (variant == 2)                                                     \
            or_apply(n, nr, mode, Gi, Gj, Gv, IDv, Hi, Hj, Hv, Li, Lj, Lv,    \
                     Ui, Uj, Uv, in, out, work);                              \
        else {                                                                \
            if (issym) or_csr_mv_trans(n, Gi, Gj, Gv, in, out);               \
            else or_csr_mv(n, Hi, Hj, Hv, in, out);                           \
            or_vmul(n, IDv, out, tm);                                         \
            or_csr_mv(n, Gi, Gj, Gv, tm, out);                                \
        }                                                                     \
    } while (0)

    *iter = 0;
    OR_MV(r, x);
    #pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int i = 0; i < n; i++) p[i] = b[i] + (-1.0) * r[i];
    memcpy(r0, p, nb);
    memcpy(r, p, nb);
    ro = or_nrm2(n, r0);
    m1 = ro * ro;
    rc = ro;

    while (*iter < NMAX) {
        OR_PREC(p, pc);
        OR_MV(v, pc);
        a = m1 / or_dot(n, r0, v);
        if (variant == 2 && !isfinite(a)) {
            *iter = NMAX;
            *info = 3;
            break;
        }
        #pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int i = 0; i < n; i++) s[i] = r[i] + (-a) * v[i];
        OR_PREC(s, sc);
        OR_MV(t, sc);
        om = or_dot(n, t, s) / or_dot(n, t, t);
        if (fabs(a) * or_nrm2(n, pc) < eps * or_nrm2(n, x)) stag++;
        else stag = 0;
        #pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int i = 0; i < n; i++) x[i] = x[i] + a * pc[i];
        #pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int i = 0; i < n; i++) x[i] = x[i] + om * sc[i];
        prc = rc;
        rc = or_nrm2(n, r);
        if (trace && ntrace < trace_cap) trace[ntrace] = rc;
        ntrace++;
        if (prc < rc) msteps++;
        else msteps = 0;
        if (msteps >= maxmsteps) { *info = 1; break; }
        if (stag >= maxstagsteps) { *info = 2; break; }
        if (rc < tol * ro) { *info = 0; break; }
        #pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int i = 0; i < n; i++) r[i] = s[i] + (-om) * t[i];
        m2 = m1;
        m1 = or_dot(n, r0, r);
        beta = m1 * a / (om * m2);
        #pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int i = 0; i < n; i++) p[i] = (-beta * om) * v[i] + beta * p[i];
        #pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int i = 0; i < n; i++) p[i] = p[i] + r[i];
        (*iter)++;
    }
    if (*iter == NMAX) *info = 3;
    *resval = rc / ro;
#undef OR_MV
#undef OR_PREC
    free(p); free(r0); free(r); free(v); free(s); free(t);
    free(pc); free(sc); free(tm); free(work);
    return ntrace;
}
