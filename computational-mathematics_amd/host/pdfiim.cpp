// PDFIIM factor producer (host, OpenMP over rows, FP64).
//
// Column i of G and row i of H are computed independently for every i from a
// dynamic sparsity pattern S (a set of indices < i grown by level sets of A's
// graph and filtered after each solve):
//   g_S = -A(S,S)^{-1} A(S,i),   h_S = -A(S,S)^{-T} A(i,S)^T,
//   D(i) = 1 / ([h;1]^T A [g;1])
// (PDFIIM/src/parPrec.hpp:36-46, PDFIIM/src/parPrec.cpp:95-588).  The pattern
// bookkeeping follows the reference's order of insertion and traversal
// (PDFIIM/src/sparseAccumulator.hpp:268-482), because the order of S fixes the
// row/column order of the small dense systems and thus their rounding.  The
// dense solves use this file's own partial-pivoting LU (the reference calls
// LAPACKE dgetrf/dgetrs, PDFIIM/src/parPrec.cpp:288-290); pivots are chosen
// the LAPACK way (first index of the largest magnitude), the arithmetic
// differs from MKL's blocked kernels at rounding level.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "cm_host.h"
#include "factor_common.hpp"

namespace {

// Insertion-ordered index set over [0, n): link (-1 = absent), insertion rank,
// most-recent-first traversal.  PDFIIM/src/sparseAccumulator.hpp:268-482.
struct PatternSet {
    std::vector<int> link, rank;
    int head = -2, cursor = -2, count = 0;
    explicit PatternSet(int n) : link(n, -1), rank(n, -1) {}
    inline bool has(int k) const { return link[k] != -1; }
    inline void push(int k)
    {
        if (link[k] == -1) {
            link[k] = head;
            rank[k] = count;
            head = k;
            count++;
        }
    }
    inline void pop_latest()
    {
        int k = head;
        head = link[k];
        link[k] = -1;
        rank[k] = -1;
        count--;
    }
    inline void rewind() { cursor = head; }
    inline int next()
    {
        int k = cursor;
        cursor = link[k];
        return k;
    }
    void clear()
    {
        for (int j = 0; j < count; j++) {
            int k = head;
            head = link[k];
            link[k] = -1;
            rank[k] = -1;
        }
        count = 0;
    }
};

struct CsrView {
    const int *ptr, *idx;
    const double *val;
};

// Push every index < limit of row r (rows sorted ascending: stop at the first
// index >= limit).
inline void push_row_below(PatternSet &s, const CsrView &m, int r, int limit)
{
    for (int l = m.ptr[r]; l < m.ptr[r + 1]; l++) {
        if (m.idx[l] < limit) s.push(m.idx[l]);
        else break;
    }
}

// In-place LU with partial pivoting of a row-major len x len matrix.
void lu_factor(int len, double *B, int *piv)
{
    for (int j = 0; j < len; j++) {
        int p = j;
        double best = std::fabs(B[(size_t)j * len + j]);
        for (int r = j + 1; r < len; r++) {
            double a = std::fabs(B[(size_t)r * len + j]);
            if (a > best) {
                best = a;
                p = r;
            }
        }
        piv[j] = p;
        if (p != j)
            for (int c = 0; c < len; c++) std::swap(B[(size_t)j * len + c], B[(size_t)p * len + c]);
        const double d = B[(size_t)j * len + j];
        if (d == 0.0) continue;  // singular: the solves then propagate inf/NaN
        for (int r = j + 1; r < len; r++) {
            double *row = B + (size_t)r * len;
            const double l = row[j] / d;
            row[j] = l;
            if (l == 0.0) continue;
            const double *prow = B + (size_t)j * len;
            for (int c = j + 1; c < len; c++) row[c] -= l * prow[c];
        }
    }
}

// Solve B x = b given the LU factors (b overwritten by x).  Both
// substitutions are column-oriented: once an unknown is final it is
// subtracted from every later (forward) or earlier (backward) row, so each
// entry's operations run in a fixed order that a wave-parallel
// implementation (one lane per row) reproduces exactly.
void lu_solve_n(int len, const double *B, const int *piv, double *b)
{
    for (int j = 0; j < len; j++)
        if (piv[j] != j) std::swap(b[j], b[piv[j]]);
    for (int c = 0; c < len; c++)  // L y = b, unit diagonal
        for (int r = c + 1; r < len; r++) b[r] -= B[(size_t)r * len + c] * b[c];
    for (int c = len - 1; c >= 0; c--) {  // U x = y
        b[c] = b[c] / B[(size_t)c * len + c];
        for (int r = 0; r < c; r++) b[r] -= B[(size_t)r * len + c] * b[c];
    }
}

// Solve B^T x = b given the LU factors of B (b overwritten by x), in the
// same column-oriented form.
void lu_solve_t(int len, const double *B, const int *piv, double *b)
{
    for (int c = 0; c < len; c++) {  // U^T y = b
        b[c] = b[c] / B[(size_t)c * len + c];
        for (int r = c + 1; r < len; r++) b[r] -= B[(size_t)c * len + r] * b[c];
    }
    for (int c = len - 1; c >= 0; c--)  // L^T z = y, unit diagonal
        for (int r = 0; r < c; r++) b[r] -= B[(size_t)c * len + r] * b[c];
    for (int j = len - 1; j >= 0; j--)
        if (piv[j] != j) std::swap(b[j], b[piv[j]]);
}

double filter_norm(char filttype, const std::vector<double> &v, int len)
{
    double nrm = 1.0;
    if (filttype == '1') {
        for (int j = 0; j < len; j++) nrm += std::fabs(v[j]);
        nrm /= (len + 1);
    } else if (filttype == '2') {
        for (int j = 0; j < len; j++) nrm += v[j] * v[j];
        nrm = std::sqrt(nrm);
    } else {
        for (int j = 0; j < len; j++) nrm = (nrm < std::fabs(v[j])) ? std::fabs(v[j]) : nrm;
    }
    return nrm;
}

}  // namespace

extern "C" int cm_pdfiim(int n, const double *Av, const int *Aj, const int *Ai,
                         double **Gval, int **Gindj, int **Gindi, double **IDv,
                         double **Hval, int **Hindj, int **Hindi,
                         int prefill, int lfill, int postfill, double ftol, int warpsize,
                         char filttype, char patt, int nthreads)
{
    *Gval = *Hval = *IDv = nullptr;
    *Gindj = *Gindi = *Hindj = *Hindi = nullptr;
    if (n <= 0) return -1;
    if (patt != 'B' && patt != 'L' && patt != 'U') return -1;
    // The reference's 'L'/'U' level-set expansion walks a stale list cursor
    // (PDFIIM/src/parPrec.cpp:393-456: next() without rewind()), which is
    // undefined once a second pass uses the expanded set.  Only the single-pass
    // form, whose expansion never reaches the output, is defined.
    if (patt != 'B' && lfill > 1) return -1;
    // below 1 (other than -1) the reference's truncation pops an empty set
    if (warpsize != -1 && warpsize < 1) return -1;

    // A^T in CSR (PDFIIM/src/sparseLinearAlgebra.cpp:50-109).
    std::vector<int> tptr(n + 1, 0), tidx(Ai[n]);
    std::vector<double> tval(Ai[n]);
    for (int i = 0; i < n; i++)
        for (int k = Ai[i]; k < Ai[i + 1]; k++) tptr[Aj[k] + 1]++;
    for (int i = 0; i < n; i++) tptr[i + 1] += tptr[i];
    {
        std::vector<int> fill(tptr.begin(), tptr.end() - 1);
        for (int i = 0; i < n; i++)
            for (int k = Ai[i]; k < Ai[i + 1]; k++) {
                int q = fill[Aj[k]]++;
                tidx[q] = i;
                tval[q] = Av[k];
            }
    }
    const CsrView A{Ai, Aj, Av}, At{tptr.data(), tidx.data(), tval.data()};
    const bool use_at = (patt == 'U' || patt == 'B');
    const bool use_a = (patt == 'L' || patt == 'B');
    const double diag_floor = std::max(1e-4, ftol);

    std::vector<std::vector<double>> gcv(n), hrv(n);
    std::vector<std::vector<int>> gci(n), hri(n);
    gcv[0] = {1.0};
    gci[0] = {0};
    hrv[0] = {1.0};
    hri[0] = {0};
    std::vector<double> dinv(n);
    dinv[0] = 1. / Av[0];

#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        PatternSet col(n);
        std::vector<int> inds, piv;
        std::vector<double> B, rhs, rhs2;
        std::vector<std::pair<int, double>> tmp;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
        for (int i = 1; i < n; i++) {
            int c = 0, c2 = 0, len = 0, tlen = 0;
            double ng = 0.0, nh = 0.0;
            if (use_at) push_row_below(col, At, i, i);
            if (use_a) push_row_below(col, A, i, i);
            for (int k = 2; k <= prefill; k++) {
                col.rewind();
                len = col.count;
                for (int j = 0; j < len; j++) {
                    int r = col.next();
                    if (use_at) push_row_below(col, At, r, i);
                    if (use_a) push_row_below(col, A, r, i);
                }
            }
            for (int k = 1; k <= lfill; k++) {
                c = 0;
                c2 = 0;
                ng = 0.0;
                len = col.count;
                tlen = len;
                if (len == 0) break;
                if (warpsize != -1)
                    for (int j = len; j >= warpsize; j--) col.pop_latest();
                len = col.count;
                tlen = len;
                col.rewind();
                inds.resize(len);
                for (int j = 0; j < len; j++) {
                    int r = col.next();
                    inds[col.rank[r]] = r;
                }
                // dense A(S,S) in insertion order, floor on absent/zero diagonals
                B.assign((size_t)len * len, 0.0);
                rhs.assign(len, 0.0);
                rhs2.assign(len, 0.0);
                for (int j = 0; j < len; j++) {
                    const int r = inds[j];
                    double *row = B.data() + (size_t)j * len;
                    row[j] = diag_floor;
                    for (int kk = Ai[r]; kk < Ai[r + 1] && Aj[kk] < i; kk++)
                        if (col.has(Aj[kk])) row[col.rank[Aj[kk]]] = Av[kk];
                    if (row[j] == 0.0) row[j] = diag_floor;
                }
                for (int q = tptr[i]; q < tptr[i + 1]; q++)
                    if (col.has(tidx[q]) && tidx[q] < i) rhs[col.rank[tidx[q]]] = -tval[q];
                for (int q = Ai[i]; q < Ai[i + 1]; q++)
                    if (col.has(Aj[q]) && Aj[q] < i) rhs2[col.rank[Aj[q]]] = -Av[q];
                piv.resize(len);
                lu_factor(len, B.data(), piv.data());
                lu_solve_n(len, B.data(), piv.data(), rhs.data());
                lu_solve_t(len, B.data(), piv.data(), rhs2.data());
                col.clear();
                ng = filter_norm(filttype, rhs, len);
                nh = filter_norm(filttype, rhs2, len);
                c = 0;
                c2 = 0;
                tlen = len;
                for (int j = 0; j < len; j++)
                    if (std::fabs(rhs[j]) >= ftol * ng) {
                        col.push(inds[j]);
                        c++;
                    }
                for (int j = 0; j < len; j++)
                    if (std::fabs(rhs2[j]) >= ftol * nh) {
                        col.push(inds[j]);
                        c2++;
                    }
                if ((c == 0 && c2 == 0) || (c2 == warpsize)) break;
                // The next level set only feeds a further pass.
                if (k == lfill) break;
                len = col.count;
                for (int j = 0; j < len; j++) {
                    const int r = inds[j];  // first entries of the previous pass's order
                    if (use_at) push_row_below(col, At, r, i);
                    if (use_a) push_row_below(col, A, r, i);
                }
                for (int m = 2; m <= postfill; m++) {
                    col.rewind();
                    int len2 = col.count;
                    for (int j = 0; j < len2; j++) {
                        int r = col.next();
                        if (use_at) push_row_below(col, At, r, i);
                        if (use_a) push_row_below(col, A, r, i);
                    }
                }
                if (col.count == len) break;
            }
            // keep, sort by index, append the unit diagonal
            tmp.clear();
            for (int j = 0; j < tlen; j++)
                if (std::fabs(rhs[j]) >= ftol * ng) tmp.push_back({inds[j], rhs[j]});
            std::sort(tmp.begin(), tmp.end(), [](auto &a, auto &b) { return a.first < b.first; });
            gci[i].resize(tmp.size() + 1);
            gcv[i].resize(tmp.size() + 1);
            for (size_t q = 0; q < tmp.size(); q++) {
                gci[i][q] = tmp[q].first;
                gcv[i][q] = tmp[q].second;
            }
            gci[i][tmp.size()] = i;
            gcv[i][tmp.size()] = 1.0;
            tmp.clear();
            for (int j = 0; j < tlen; j++)
                if (std::fabs(rhs2[j]) >= ftol * nh) tmp.push_back({inds[j], rhs2[j]});
            std::sort(tmp.begin(), tmp.end(), [](auto &a, auto &b) { return a.first < b.first; });
            hri[i].resize(tmp.size() + 1);
            hrv[i].resize(tmp.size() + 1);
            for (size_t q = 0; q < tmp.size(); q++) {
                hri[i][q] = tmp[q].first;
                hrv[i][q] = tmp[q].second;
            }
            hri[i][tmp.size()] = i;
            hrv[i][tmp.size()] = 1.0;
            col.clear();
            // D(i) = 1 / (h^T A g) by sorted merges of A's rows with g
            double s = 0;
            const std::vector<int> &gj = gci[i];
            const std::vector<double> &gv = gcv[i];
            for (size_t j = 0; j < hri[i].size(); j++) {
                const int hj = hri[i][j];
                const double hv = hrv[i][j];
                int kk = Ai[hj];
                size_t cc = 0;
                while (kk < Ai[hj + 1] && cc < gj.size()) {
                    if (Aj[kk] == gj[cc]) {
                        s += hv * Av[kk] * gv[cc];
                        kk++;
                        cc++;
                    } else if (Aj[kk] < gj[cc]) {
                        kk++;
                    } else {
                        cc++;
                    }
                }
            }
            dinv[i] = 1. / s;
        }
    }

    // G by rows from its columns (rows come out with ascending columns, the
    // unit diagonal first); H rows are already sorted with the diagonal last.
    std::vector<int> gp(n + 1, 0);
    for (int i = 0; i < n; i++)
        for (int r : gci[i]) gp[r + 1]++;
    for (int i = 0; i < n; i++) gp[i + 1] += gp[i];
    std::vector<int> gj(gp[n]);
    std::vector<double> gvv(gp[n]);
    {
        std::vector<int> fill(gp.begin(), gp.end() - 1);
        for (int i = 0; i < n; i++)
            for (size_t q = 0; q < gci[i].size(); q++) {
                int p = fill[gci[i][q]]++;
                gj[p] = i;
                gvv[p] = gcv[i][q];
            }
    }
    std::vector<int> hp(n + 1, 0);
    for (int i = 0; i < n; i++) hp[i + 1] = hp[i] + (int)hri[i].size();
    std::vector<int> hj(hp[n]);
    std::vector<double> hvv(hp[n]);
    for (int i = 0; i < n; i++) {
        std::copy(hri[i].begin(), hri[i].end(), hj.begin() + hp[i]);
        std::copy(hrv[i].begin(), hrv[i].end(), hvv.begin() + hp[i]);
    }
    *Gindi = cmh::new_copy(gp);
    *Gindj = cmh::new_copy(gj);
    *Gval = cmh::new_copy(gvv);
    *Hindi = cmh::new_copy(hp);
    *Hindj = cmh::new_copy(hj);
    *Hval = cmh::new_copy(hvv);
    *IDv = cmh::new_copy(dinv);
    return 0;
}
