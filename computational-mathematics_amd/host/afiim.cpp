// AFIIM and ACIFIM factor producers (host, sequential, FP64).
//
// Both build M = G D^{-1} H ~ A^{-1} by the bordering recursion
//   A_{i+1} = [A_i b; c d]  ->  g = -G D^{-1} H b,  h = -c G D^{-1} H,
//   s = d - c G D^{-1} H b   (AFIIM/src/prec.hpp:24-36, ACIFIM/src/prec.hpp:21-35)
// with norm-based dropping.  Factor construction is not accelerated: it runs
// once per system and its output must be the factors the reference produces,
// bit for bit, because the solve's iteration counts depend on them.  The
// floating-point results depend on the order in which the sparse
// accumulators are filled and walked, so that order follows the reference
// (most-recent-first walks, "first touch stores" accumulation, the same
// operand order in every product).  Storage is this file's own: G is kept by
// columns and by rows in growable per-row lists, H by rows and by columns.
//
// Reference: AFIIM/src/prec.cpp:101-767 (afiim),
//            ACIFIM/src/prec.cpp:127-1052 (acifim).
#include <cmath>
#include <cstring>

#include "cm_host.h"
#include "factor_common.hpp"

namespace cmh {

void split_lower_upper(int n, const double *Av, const int *Aj, const int *Ai,
                       double zero_diag_value, Split &s)
{
    s.lptr.assign(n + 1, 0);
    s.uptr.assign(n + 1, 0);
    s.d.assign(n, 0.0);
    for (int i = 0; i < n; i++)
        for (int k = Ai[i]; k < Ai[i + 1]; k++) {
            int j = Aj[k];
            if (j > i) s.uptr[j + 1]++;
            else if (j < i) s.lptr[i + 1]++;
        }
    for (int i = 0; i < n; i++) {
        s.uptr[i + 1] += s.uptr[i];
        s.lptr[i + 1] += s.lptr[i];
    }
    s.lidx.resize(s.lptr[n]);
    s.lval.resize(s.lptr[n]);
    s.uidx.resize(s.uptr[n]);
    s.uval.resize(s.uptr[n]);
    std::vector<int> ufill(s.uptr.begin(), s.uptr.end() - 1);
    int lpos = 0;
    for (int i = 0; i < n; i++) {
        for (int k = Ai[i]; k < Ai[i + 1]; k++) {
            int j = Aj[k];
            if (j > i) {
                int q = ufill[j]++;
                s.uidx[q] = i;
                s.uval[q] = Av[k];
            } else if (j < i) {
                s.lidx[lpos] = j;
                s.lval[lpos] = Av[k];
                lpos++;
            } else {
                s.d[i] = Av[k];
            }
        }
        if (s.d[i] == 0.0) s.d[i] = zero_diag_value;
    }
}

namespace {

// Row-wise CSR from a list of rows held in order.
void rows_to_csr(int n, const std::vector<std::vector<Ent>> &rows,
                 std::vector<int> &ptr, std::vector<int> &idx, std::vector<double> &val)
{
    ptr.assign(n + 1, 0);
    for (int r = 0; r < n; r++) ptr[r + 1] = ptr[r] + (int)rows[r].size();
    idx.resize(ptr[n]);
    val.resize(ptr[n]);
    for (int r = 0; r < n; r++) {
        int q = ptr[r];
        for (const Ent &e : rows[r]) {
            idx[q] = e.i;
            val[q] = e.v;
            q++;
        }
    }
}

// Row-wise CSR from column lists (cols[c] holds (row, value) pairs): rows come
// out with ascending columns.
void cols_to_csr(int n, const std::vector<std::vector<Ent>> &cols,
                 std::vector<int> &ptr, std::vector<int> &idx, std::vector<double> &val)
{
    ptr.assign(n + 1, 0);
    for (int c = 0; c < n; c++)
        for (const Ent &e : cols[c]) ptr[e.i + 1]++;
    for (int r = 0; r < n; r++) ptr[r + 1] += ptr[r];
    idx.resize(ptr[n]);
    val.resize(ptr[n]);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int c = 0; c < n; c++)
        for (const Ent &e : cols[c]) {
            int q = fill[e.i]++;
            idx[q] = c;
            val[q] = e.v;
        }
}

// A compressed list of lists appended one list at a time (G by columns, H by
// rows, in formation order).
struct Packed {
    std::vector<int> ptr{0};
    std::vector<int> idx;
    std::vector<double> val;
    void close() { ptr.push_back((int)idx.size()); }
};

// Drop pass over a LIFO accumulator: zero every entry whose test is true.
// Returns how many entries test true (zeros already present count again, as
// in ACIFIM/src/prec.cpp:434-450).
template <class Pred>
int drop_pass(Spa &a, Pred pred)
{
    int cnt = 0;
    a.walk([&](int k) {
        if (pred(k)) {
            a.v[k] = 0.0;
            cnt++;
        }
    });
    return cnt;
}

// Moves the surviving (nonzero) entries of a LIFO accumulator, negated, into
// the packed list in walk order, appends the unit diagonal, and clears it.
void store_negated(Spa &a, int diag, Packed &out)
{
    a.walk([&](int k) {
        if (a.v[k] != 0.0) {
            out.idx.push_back(k);
            out.val.push_back(-a.v[k]);
        }
    });
    a.clear();
    out.idx.push_back(diag);
    out.val.push_back(1.0);
    out.close();
}

}  // namespace
}  // namespace cmh

using namespace cmh;

extern "C" void cm_afiim(int n, const double *Av, const int *Aj, const int *Ai,
                         double **Gval, int **Gindj, int **Gindi, double **IDv,
                         double **Hval, int **Hindj, int **Hindi, double dtol,
                         int elemPerRowCol, int growth, double eta, double shift)
{
    (void)elemPerRowCol;  // allocation hints of the reference; results do not depend on them
    (void)growth;
    Split A;
    split_lower_upper(n, Av, Aj, Ai, dtol, A);

    std::vector<double> gn(n, 1.0), hn(n, 1.0), un(n, 1.0), ln(n, 1.0);
    std::vector<double> alr(n, 1.0), auc(n, 1.0);
    // Scaled inf-norms of the columns of triu(A) and the rows of tril(A)
    // (AFIIM/src/prec.cpp:308-327).
    for (int c = 0; c < n; c++)
        for (int k = A.uptr[c]; k < A.uptr[c + 1]; k++) {
            int r = A.uidx[k];
            auc[c] = dmax(auc[c], std::fabs(A.uval[k]) /
                                      (std::sqrt(std::fabs(A.d[c])) * std::sqrt(std::fabs(A.d[r]))));
        }
    for (int r = 0; r < n; r++)
        for (int k = A.lptr[r]; k < A.lptr[r + 1]; k++) {
            int c = A.lidx[k];
            alr[r] = dmax(alr[r], std::fabs(A.lval[k]) /
                                      (std::sqrt(std::fabs(A.d[r])) * std::sqrt(std::fabs(A.d[c]))));
        }

    std::vector<double> idv(n);
    {
        double s0 = A.d[0] + shift * std::fabs(A.d[0]);
        idv[0] = (std::fabs(s0) < eta) ? 1.0 / (1e-4 + dtol) : 1.0 / s0;
    }

    Packed gcol, hrow;  // G by columns, H by rows, in formation order
    gcol.idx.push_back(0);
    gcol.val.push_back(1.0);
    gcol.close();
    hrow.idx.push_back(0);
    hrow.val.push_back(1.0);
    hrow.close();
    std::vector<std::vector<Ent>> grow(n), hcol(n);  // G by rows, H by columns
    grow[0].push_back({0, 1.0});
    hcol[0].push_back({0, 1.0});

    Spa tb, ttb, tc, ttc;
    tb.init(n);
    ttb.init(n);
    tc.init(n);
    ttc.init(n);

    for (int m = 1; m < n; m++) {
        // t_b = H b, b = A(0:m-1, m)
        for (int k = A.uptr[m]; k < A.uptr[m + 1]; k++) {
            const double u = A.uval[k];
            for (const Ent &e : hcol[A.uidx[k]]) tb.add(e.i, u * e.v);
        }
        // tt_b = G D^{-1} t_b
        tb.walk([&](int j) {
            const double t = tb.v[j];
            for (int q = gcol.ptr[j]; q < gcol.ptr[j + 1]; q++)
                ttb.add(gcol.idx[q], t * idv[j] * gcol.val[q]);
            un[m] = dmax(un[m], std::fabs(idv[j] * t));
        });
        // t_c = c G, c = A(m, 0:m-1)
        for (int k = A.lptr[m]; k < A.lptr[m + 1]; k++) {
            const double l = A.lval[k];
            for (const Ent &e : grow[A.lidx[k]]) tc.add(e.i, l * e.v);
        }
        // tt_c = t_c D^{-1} H
        tc.walk([&](int j) {
            const double t = tc.v[j];
            for (int q = hrow.ptr[j]; q < hrow.ptr[j + 1]; q++)
                ttc.add(hrow.idx[q], idv[j] * t * hrow.val[q]);
            ln[m] = dmax(ln[m], std::fabs(t * idv[j]));
        });
        // s = d + c G D^{-1} H b
        const double d = A.d[m];
        double sum = 0.0;
        tb.walk([&](int j) { sum += tb.v[j] * idv[j] * tc.v[j]; });
        tb.clear();
        tc.clear();
        double s = d + sum;
        double s1 = 0.0;
        for (int k = A.lptr[m]; k < A.lptr[m + 1]; k++) s1 -= A.lval[k] * ttb.v[A.lidx[k]];
        // filter and store g_m (column m of G)
        const double gden = dmin(un[m], alr[m]);
        drop_pass(ttb, [&](int j) {
            gn[j] = dmax(gn[j], std::fabs(ttb.v[j]));
            return std::fabs(ttb.v[j]) <= dtol * gn[j] / gden;
        });
        store_negated(ttb, m, gcol);
        double s2 = 0.0;
        for (int k = A.uptr[m]; k < A.uptr[m + 1]; k++) s2 -= A.uval[k] * ttc.v[A.uidx[k]];
        // filter and store h_m (row m of H)
        const double hden = dmin(ln[m], auc[m]);
        drop_pass(ttc, [&](int j) {
            hn[j] = dmax(hn[j], std::fabs(ttc.v[j]));
            return std::fabs(ttc.v[j]) <= dtol * hn[j] / hden;
        });
        store_negated(ttc, m, hrow);
        for (int q = gcol.ptr[m]; q < gcol.ptr[m + 1]; q++) grow[gcol.idx[q]].push_back({m, gcol.val[q]});
        for (int q = hrow.ptr[m]; q < hrow.ptr[m + 1]; q++) hcol[hrow.idx[q]].push_back({m, hrow.val[q]});
        s += (s1 + s2) + shift * std::fabs(d);
        if (std::fabs(s) <= eta) s = (1e-4 + dtol);
        idv[m] = 1. / s;
    }

    std::vector<int> ptr, idx;
    std::vector<double> val;
    rows_to_csr(n, grow, ptr, idx, val);
    *Gindi = new_copy(ptr);
    *Gindj = new_copy(idx);
    *Gval = new_copy(val);
    cols_to_csr(n, hcol, ptr, idx, val);
    *Hindi = new_copy(ptr);
    *Hindj = new_copy(idx);
    *Hval = new_copy(val);
    *IDv = new_copy(idv);
}

extern "C" int cm_acifim(int n, const double *Av, const int *Aj, const int *Ai,
                         double **Gval, int **Gindj, int **Gindi, double **IDv,
                         double **Hval, int **Hindj, int **Hindi,
                         double **Lval, int **Lindj, int **Lindi,
                         double **Uval, int **Uindj, int **Uindi,
                         double dtol1, double dtol2, int elemPerRowCol, int growth,
                         double eta, double shift, double coeff, int element_limit,
                         int filtit)
{
    (void)elemPerRowCol;
    (void)growth;
    *Gval = nullptr;
    *Gindj = *Gindi = nullptr;
    *Hval = nullptr;
    *Hindj = *Hindi = nullptr;
    *Lval = nullptr;
    *Lindj = *Lindi = nullptr;
    *Uval = nullptr;
    *Uindj = *Uindi = nullptr;
    *IDv = nullptr;

    Split A;
    split_lower_upper(n, Av, Aj, Ai, dtol2, A);
    std::vector<double> gn(n, 1.0), hn(n, 1.0), gnc(n, 1.0), hnr(n, 1.0);
    std::vector<double> idv(n);
    {
        double s0 = A.d[0] + shift * std::fabs(A.d[0]);
        idv[0] = (std::fabs(s0) < eta) ? 1.0 / (1e-4 + dtol2) : 1.0 / s0;
    }
    Packed gcol, hrow;
    gcol.idx.push_back(0);
    gcol.val.push_back(1.0);
    gcol.close();
    hrow.idx.push_back(0);
    hrow.val.push_back(1.0);
    hrow.close();
    std::vector<std::vector<Ent>> grow(n), hcol(n), urow(n), lcol(n);
    grow[0].push_back({0, 1.0});
    hcol[0].push_back({0, 1.0});
    for (int r = 0; r < n; r++) {
        urow[r].push_back({r, 1.0});  // unit diagonal of U, first in its row
        lcol[r].push_back({r, 1.0});  // unit diagonal of L, last in its row
    }

    Spa tb, ttb, tc, ttc;
    tb.init(n);
    ttb.init(n);
    tc.init(n);
    ttc.init(n);

    // Threshold filter followed by the element-limit passes with a growing
    // multiplier (ACIFIM/src/prec.cpp:417-450).
    auto limit_filter = [&](Spa &a, int cnt, auto pred) {
        int l = 1;
        double mult = 1.0;
        while ((a.cnt - cnt) > element_limit && l <= filtit) {
            cnt = drop_pass(a, [&](int j) { return pred(j, mult); });
            mult *= coeff;
            l++;
        }
        return cnt;
    };

    for (int m = 1; m < n; m++) {
        // t_b = H b
        for (int k = A.uptr[m]; k < A.uptr[m + 1]; k++) {
            const double u = A.uval[k];
            for (const Ent &e : hcol[A.uidx[k]]) tb.add(e.i, u * e.v);
        }
        // filter u_m and store it (scaled by D^{-1}) as column m of U
        {
            auto pu = [&](int j, double mult) {
                return std::fabs(idv[j] * tb.v[j]) <= mult * dtol1 / (gnc[j]);
            };
            int cnt = drop_pass(tb, [&](int j) { return std::fabs(idv[j] * tb.v[j]) <= dtol1 / (gnc[j]); });
            limit_filter(tb, cnt, pu);
        }
        tb.walk([&](int j) {
            const double t = tb.v[j];
            if (std::fabs(t) != 0.0) urow[j].push_back({m, idv[j] * t});
        });
        // tt_b = G D^{-1} t_b
        tb.walk([&](int j) {
            const double t = tb.v[j];
            if (std::fabs(t) != 0.0)
                for (int q = gcol.ptr[j]; q < gcol.ptr[j + 1]; q++)
                    ttb.add(gcol.idx[q], t * idv[j] * gcol.val[q]);
        });
        // t_c = c G
        for (int k = A.lptr[m]; k < A.lptr[m + 1]; k++) {
            const double l = A.lval[k];
            for (const Ent &e : grow[A.lidx[k]]) tc.add(e.i, l * e.v);
        }
        // filter l_m and store it (scaled by D^{-1}) as row m of L
        {
            auto pl = [&](int j, double mult) {
                return std::fabs(idv[j] * tc.v[j]) <= mult * dtol1 / (hnr[j]);
            };
            int cnt = drop_pass(tc, [&](int j) { return std::fabs(idv[j] * tc.v[j]) <= dtol1 / (hnr[j]); });
            limit_filter(tc, cnt, pl);
        }
        tc.walk([&](int j) {
            const double t = tc.v[j];
            if (std::fabs(t) != 0.0) lcol[j].push_back({m, idv[j] * t});
        });
        // tt_c = t_c D^{-1} H
        tc.walk([&](int j) {
            const double t = tc.v[j];
            if (std::fabs(t) != 0.0)
                for (int q = hrow.ptr[j]; q < hrow.ptr[j + 1]; q++)
                    ttc.add(hrow.idx[q], idv[j] * t * hrow.val[q]);
        });
        // Schur complement
        const double d = A.d[m];
        double sum = 0.0;
        tb.walk([&](int j) { sum += tb.v[j] * idv[j] * tc.v[j]; });
        tb.clear();
        tc.clear();
        double s = d + sum;
        double s1 = 0.0;
        for (int k = A.lptr[m]; k < A.lptr[m + 1]; k++) {
            const int j = A.lidx[k];
            s1 -= A.lval[k] * ttb.v[j];
            gnc[m] = dmax(gnc[m], std::fabs(ttb.v[j]));
        }
        // filter and store g_m
        {
            int cnt = drop_pass(ttb, [&](int j) {
                gn[j] = gn[j] + std::fabs(ttb.v[j]);
                return std::fabs(ttb.v[j]) <= dtol2 * gn[j];
            });
            limit_filter(ttb, cnt, [&](int j, double mult) {
                return std::fabs(ttb.v[j]) <= mult * dtol2 * gn[j];
            });
        }
        store_negated(ttb, m, gcol);
        double s2 = 0.0;
        for (int k = A.uptr[m]; k < A.uptr[m + 1]; k++) {
            const int j = A.uidx[k];
            s2 -= A.uval[k] * ttc.v[j];
            hnr[m] = dmax(hnr[m], std::fabs(ttc.v[j]));
        }
        // filter and store h_m
        {
            int cnt = drop_pass(ttc, [&](int j) {
                hn[j] = hn[j] + std::fabs(ttc.v[j]);
                return std::fabs(ttc.v[j]) <= dtol2 * hn[j];
            });
            limit_filter(ttc, cnt, [&](int j, double mult) {
                return std::fabs(ttc.v[j]) <= mult * dtol2 * hn[j];
            });
        }
        store_negated(ttc, m, hrow);
        for (int q = gcol.ptr[m]; q < gcol.ptr[m + 1]; q++) grow[gcol.idx[q]].push_back({m, gcol.val[q]});
        for (int q = hrow.ptr[m]; q < hrow.ptr[m + 1]; q++) hcol[hrow.idx[q]].push_back({m, hrow.val[q]});
        s += (s1 + s2) + shift * std::fabs(d);
        if (std::isnan(s)) return 1;
        if (std::fabs(s) <= eta) s = (1e-4 + dtol2);
        idv[m] = 1. / s;
    }

    std::vector<int> ptr, idx;
    std::vector<double> val;
    rows_to_csr(n, grow, ptr, idx, val);
    *Gindi = new_copy(ptr);
    *Gindj = new_copy(idx);
    *Gval = new_copy(val);
    rows_to_csr(n, urow, ptr, idx, val);
    *Uindi = new_copy(ptr);
    *Uindj = new_copy(idx);
    *Uval = new_copy(val);
    cols_to_csr(n, hcol, ptr, idx, val);
    *Hindi = new_copy(ptr);
    *Hindj = new_copy(idx);
    *Hval = new_copy(val);
    cols_to_csr(n, lcol, ptr, idx, val);
    *Lindi = new_copy(ptr);
    *Lindj = new_copy(idx);
    *Lval = new_copy(val);
    *IDv = new_copy(idv);
    return 0;
}

extern "C" void cm_free_d(double *p) { delete[] p; }
extern "C" void cm_free_i(int *p) { delete[] p; }
