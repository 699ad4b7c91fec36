// CPU harness of the engine's host-side index code (csrc/cm_layout.hpp),
// built with the address and undefined-behaviour sanitizers: random and
// edge-case CSR matrices (n = 1, empty rows, rows longer than every bin and
// tile limit, banded and stencil-like structure, long-range couplings over
// several column tiles) through every layout function, each result checked
// against what it must satisfy -- a partition of the nonzeros, consistent
// pointers, dependencies in earlier levels, the transpose of the transpose,
// a symmetric expansion.  Exit status 0 when every check holds; the
// sanitizers abort on any out-of-bounds access or undefined behaviour.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "cm_layout.hpp"

using namespace cm;

static int g_fail = 0;
#define CHECK(c, what)                                                              \
    do {                                                                            \
        if (!(c)) {                                                                 \
            std::fprintf(stderr, "FAIL %s: %s (line %d)\n", what, #c, __LINE__); \
            g_fail++;                                                               \
        }                                                                           \
    } while (0)

struct Csr {
    int n = 0;
    std::vector<int> ptr, col;
    std::vector<double> val;
};

// rows: column sets; columns sorted, no duplicates
static Csr from_rows(int n, const std::vector<std::set<int>> &rows)
{
    Csr A;
    A.n = n;
    A.ptr.assign(n + 1, 0);
    for (int i = 0; i < n; i++) {
        for (int c : rows[i]) {
            A.col.push_back(c);
            A.val.push_back(1.0 + 0.001 * (double)(A.col.size() % 997));
        }
        A.ptr[i + 1] = (int)A.col.size();
    }
    return A;
}

static Csr random_csr(std::mt19937_64 &g, int n, double density, int nlong, int longlen, double empty_frac)
{
    std::vector<std::set<int>> rows(n);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::uniform_int_distribution<int> cu(0, n - 1);
    for (int i = 0; i < n; i++) {
        if (u(g) < empty_frac) continue;
        const int k = std::max(1, (int)(density * n * u(g) * 2));
        for (int q = 0; q < k; q++) rows[i].insert(cu(g));
    }
    for (int q = 0; q < nlong; q++) {
        const int r = cu(g);
        while ((int)rows[r].size() < std::min(longlen, n)) rows[r].insert(cu(g));
    }
    return from_rows(n, rows);
}

// 3D 7-point stencil on an N^3 grid, plus `extra` random couplings on a
// share of the rows (the long-row structure of the 27-point workload)
static Csr stencil(std::mt19937_64 &g, int N, double frac, int extra)
{
    const int n = N * N * N;
    std::vector<std::set<int>> rows(n);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::uniform_int_distribution<int> cu(0, n - 1);
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                const int i = (z * N + y) * N + x;
                rows[i].insert(i);
                if (x > 0) rows[i].insert(i - 1);
                if (x < N - 1) rows[i].insert(i + 1);
                if (y > 0) rows[i].insert(i - N);
                if (y < N - 1) rows[i].insert(i + N);
                if (z > 0) rows[i].insert(i - N * N);
                if (z < N - 1) rows[i].insert(i + N * N);
                if (u(g) < frac)
                    while ((int)rows[i].size() < 7 + extra) rows[i].insert(cu(g));
            }
    return from_rows(n, rows);
}

static void check_validation(const Csr &A)
{
    CHECK(csr_problem(A.n, A.ptr.data(), A.col.data()) == nullptr, "valid CSR accepted");
    if (A.n < 2 || A.col.empty()) return;
    Csr B = A;
    B.col[B.col.size() / 2] = A.n;
    CHECK(csr_problem(B.n, B.ptr.data(), B.col.data()) != nullptr, "column n rejected");
    B.col[B.col.size() / 2] = -1;
    CHECK(csr_problem(B.n, B.ptr.data(), B.col.data()) != nullptr, "column -1 rejected");
    B = A;
    B.ptr[0] = 1;
    CHECK(csr_problem(B.n, B.ptr.data(), B.col.data()) != nullptr, "ptr[0] != 0 rejected");
    B = A;
    if (B.ptr[1] < B.ptr[2]) {
        B.ptr[1] = B.ptr[2] + 1;
        B.ptr[2] = B.ptr[1] - 2 >= 0 ? B.ptr[1] - 2 : 0;
        CHECK(csr_problem(B.n, B.ptr.data(), B.col.data()) != nullptr, "non-monotone ptr rejected");
    }
}

static void check_blocks(const Csr &A)
{
    for (int tile : {1024, 2048, 4096}) {
        int nblk = 0;
        std::vector<int> t = block_table(A.n, A.ptr.data(), &nblk, tile);
        CHECK((int)t.size() == 2 * (nblk + 1), "block table size");
        CHECK(t[0] == 0 && t[nblk] == A.n, "blocks cover the rows");
        for (int b = 0; b < nblk; b++) {
            const int r0 = t[b], r1 = t[b + 1];
            CHECK(r1 > r0, "blocks non-empty");
            const int nz = A.ptr[r1] - A.ptr[r0];
            CHECK(r1 - r0 == 1 || (nz <= tile && r1 - r0 <= ROWCAP), "block within tile and row cap");
            CHECK(t[nblk + 1 + b] == A.ptr[r0], "block first nonzero");
        }
    }
    std::vector<int> sh, lo;
    row_bins(A.n, A.ptr.data(), SHORT_MAX, sh, lo);
    CHECK((int)(sh.size() + lo.size()) == A.n, "row bins partition");
    for (int r : lo) CHECK(A.ptr[r + 1] - A.ptr[r] > SHORT_MAX, "long bin");
    for (int r : sh) CHECK(A.ptr[r + 1] - A.ptr[r] <= SHORT_MAX, "short bin");
}

static void check_transpose(const Csr &A)
{
    std::vector<int> tp, tc, pp, pc;
    std::vector<double> tv, pv;
    host_transpose(A.n, A.ptr.data(), A.col.data(), A.val.data(), tp, tc, tv);
    CHECK(csr_problem(A.n, tp.data(), tc.data()) == nullptr, "transpose is valid CSR");
    host_transpose(A.n, tp.data(), tc.data(), tv.data(), pp, pc, pv);
    CHECK(pp == A.ptr && pc == A.col && pv == A.val, "transpose twice is the identity");
}

static void check_sym(const Csr &A)
{
    if (A.n > 3000) return;
    std::vector<int> fp, fc;
    std::vector<double> fv;
    host_sym_expand(A.n, A.ptr.data(), A.col.data(), A.val.data(), fp, fc, fv);
    CHECK(csr_problem(A.n, fp.data(), fc.data()) == nullptr, "expansion is valid CSR");
    std::map<std::pair<int, int>, double> full, want;
    for (int i = 0; i < A.n; i++) {
        for (int k = fp[i]; k < fp[i + 1]; k++) {
            CHECK(k == fp[i] || fc[k] > fc[k - 1], "expanded row sorted");
            full[{i, fc[k]}] = fv[k];
        }
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) {
            const int j = A.col[k];
            if (j < i) continue;
            want[{i, j}] = A.val[k];
            want[{j, i}] = A.val[k];
        }
    }
    CHECK(full == want, "expansion = upper + its transpose");
}

static void check_levels(const Csr &A)
{
    for (bool lower : {true, false}) {
        LevelSchedule L = level_schedule(A.n, A.ptr.data(), A.col.data(), lower);
        CHECK((int)L.loff.size() == L.nlev + 1 && L.loff[0] == 0 && L.loff[L.nlev] == A.n, "level offsets");
        std::vector<int> lev(A.n, -1);
        for (int l = 0; l < L.nlev; l++)
            for (int q = L.loff[l]; q < L.loff[l + 1]; q++) {
                CHECK(lev[L.rows[q]] == -1, "row in one level");
                lev[L.rows[q]] = l;
            }
        for (int i = 0; i < A.n; i++)
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++) {
                const int j = A.col[k];
                if (lower ? j < i : j > i) CHECK(lev[j] < lev[i], "dependency in an earlier level");
            }
        // the triangle's entries in level order: each row's own, in its order, nothing else
        CHECK((int)L.tptr.size() == A.n + 1 && L.tptr[0] == 0 && L.tptr[A.n] == (int)L.tcol.size() &&
                  L.tsrc.size() == L.tcol.size(),
              "level-ordered triangle shape");
        for (int q = 0; q < A.n; q++) {
            const int i = L.rows[q];
            std::vector<int> want;
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; k++)
                if (lower ? A.col[k] <= i : A.col[k] >= i) want.push_back(k);
            std::vector<int> got(L.tsrc.begin() + L.tptr[q], L.tsrc.begin() + L.tptr[q + 1]);
            CHECK(got == want, "level-ordered row = the row's triangle entries");
            for (int p2 = L.tptr[q]; p2 < L.tptr[q + 1]; p2++) CHECK(L.tcol[p2] == A.col[L.tsrc[p2]], "column kept");
        }
    }
}

static void check_slots(const Csr &A)
{
    SlotLayout S;
    const bool ok = slot_layout(A.n, A.ptr.data(), A.col.data(), S);
    int longest = 0;
    for (int i = 0; i < A.n; i++) longest = std::max(longest, A.ptr[i + 1] - A.ptr[i]);
    CHECK(ok == (longest <= SLOT_MAXLEN), "slot format applies iff rows fit");
    if (!ok) return;
    std::vector<int> seen(A.col.size(), 0);
    for (size_t p = 0; p < S.src.size(); p++) {
        const int k = S.src[p];
        CHECK(k >= 0 && k < (int)A.col.size(), "slot source in range");
        if (k < 0 || k >= (int)A.col.size()) continue;
        seen[k]++;
        CHECK(S.col[p] == A.col[k], "slot column");
    }
    for (int c : seen) CHECK(c == 1, "every nonzero in one slot");
    for (int b = 0; b < S.nb; b++) {
        const int r0 = S.hdr[5 * b], nr = S.hdr[5 * b + 1], L = S.hdr[5 * b + 4];
        CHECK(r0 == b * SLOT_ROWS && nr == std::min(SLOT_ROWS, A.n - r0), "slot block rows");
        for (int i = 0; i < nr; i++) {
            const int len = S.lrow[r0 + i] >> 16, pos = S.lrow[r0 + i] & 0xffff;
            CHECK(len <= L && pos < nr, "slot row descriptor");
        }
        if (S.fits16)
            for (int p = S.hdr[5 * b + 2]; p < (b + 1 < S.nb ? S.hdr[5 * (b + 1) + 2] : (int)S.src.size()); p++)
                CHECK(S.d16[p] + S.cbase[b] == S.col[p], "16-bit offset");
    }
}

static void check_dia(const Csr &A)
{
    DiaLayout D;
    if (!dia_layout(A.n, A.ptr.data(), A.col.data(), D)) return;
    CHECK(D.nd >= 1 && D.nd <= DIA_MAX && D.ldn >= A.n && D.ldn % DIA_ALIGN == 0, "diagonal shape");
    for (int d = 1; d < D.nd; d++) CHECK(D.off[d] > D.off[d - 1], "offsets ascending");
    std::vector<int> seen(A.col.size(), 0);
    for (int d = 0; d < D.nd; d++)
        for (int i = 0; i < D.ldn; i++) {
            const int k = D.dsrc[(size_t)d * D.ldn + i];
            if (k < 0) continue;
            CHECK(i < A.n && k < (int)A.col.size(), "diagonal source in range");
            if (i >= A.n || k >= (int)A.col.size()) continue;
            CHECK(k >= A.ptr[i] && k < A.ptr[i + 1] && A.col[k] - i == D.off[d], "diagonal element placement");
            seen[k]++;
        }
    CHECK((int)D.rptr.size() == A.n + 1 && D.rptr[A.n] == (int)D.rcol.size(), "remainder pointers");
    for (int i = 0; i < A.n; i++)
        for (int q = D.rptr[i]; q < D.rptr[i + 1]; q++) {
            const int k = D.rsrc[q];
            CHECK(k >= A.ptr[i] && k < A.ptr[i + 1] && A.col[k] == D.rcol[q], "remainder entry");
            if (k >= 0 && k < (int)A.col.size()) seen[k]++;
        }
    for (int c : seen) CHECK(c == 1, "every nonzero on one diagonal or in the remainder");
    // long rows: their remainders split into tile segments
    CHECK(D.nseg == (long long)D.ntile * D.nlong && (long long)D.lptr.size() == D.nseg + 1, "segment count");
    CHECK(D.ntile >= 1 && D.ntile <= LONG_TILES_MAX && (long long)D.tw * D.ntile >= A.n, "tiles cover the columns");
    std::vector<int> lrow;
    for (int i = 0; i < A.n; i++) {
        const int len = D.rptr[i + 1] - D.rptr[i];
        CHECK((D.lpos[i] >= 0) == (len > SHORT_MAX), "long row flag");
        if (D.lpos[i] >= 0) {
            CHECK(D.lpos[i] == (int)lrow.size(), "long row numbering");
            lrow.push_back(i);
        }
    }
    CHECK((int)lrow.size() == D.nlong, "long row count");
    long long tot = 0;
    for (int q = 0; q < D.nlong; q++) {
        const int i = lrow[q];
        std::vector<int> cols;
        for (int t = 0; t < D.ntile; t++) {
            const long long g = (long long)t * D.nlong + q;
            for (int p = D.lptr[g]; p < D.lptr[g + 1]; p++) {
                CHECK(D.lcol[p] / D.tw == t, "segment column in its tile");
                CHECK(D.lsrc[p] >= A.ptr[i] && D.lsrc[p] < A.ptr[i + 1], "segment entry from its row");
                cols.push_back(D.lcol[p]);
            }
        }
        std::vector<int> want(D.rcol.begin() + D.rptr[i], D.rcol.begin() + D.rptr[i + 1]);
        CHECK(cols == want, "segments of a row = its remainder, in order");
        tot += (long long)cols.size();
    }
    CHECK(tot == (long long)D.lcol.size(), "segment entries");
    if (D.nlong && D.ntile > 1) CHECK(tot / ((long long)D.ntile * D.nlong) >= LONG_SEG_MIN, "segments not too short");
}

static void run(const char *name, const Csr &A)
{
    const int before = g_fail;
    check_validation(A);
    check_blocks(A);
    check_transpose(A);
    check_sym(A);
    check_levels(A);
    check_slots(A);
    check_dia(A);
    std::printf("%-28s n=%7d nnz=%9zu %s\n", name, A.n, A.col.size(), g_fail == before ? "ok" : "FAILED");
}

int main()
{
    std::mt19937_64 g(12345);
    run("single row", from_rows(1, {{0}}));
    run("single empty row", from_rows(1, {{}}));
    {
        std::vector<std::set<int>> rows(9);
        rows[4] = {0, 8};
        run("mostly empty rows", from_rows(9, rows));
    }
    run("random sparse", random_csr(g, 500, 0.01, 0, 0, 0.1));
    run("random with long rows", random_csr(g, 5000, 0.001, 12, 3000, 0.05));
    run("random dense small", random_csr(g, 40, 0.9, 0, 0, 0.0));
    for (int N : {5, 17})
        run("7-point stencil", stencil(g, N, 0.0, 0));
    run("stencil, 5% long rows", stencil(g, 20, 0.05, 150));
    run("stencil over several tiles", stencil(g, 44, 0.05, 120));  // n = 85184: 3 column tiles
    {
        // upper and lower triangular with random fill
        Csr R = random_csr(g, 800, 0.01, 0, 0, 0.0);
        std::vector<std::set<int>> up(R.n), lo(R.n);
        for (int i = 0; i < R.n; i++) {
            up[i].insert(i);
            lo[i].insert(i);
            for (int k = R.ptr[i]; k < R.ptr[i + 1]; k++) (R.col[k] > i ? up[i] : lo[i]).insert(R.col[k]);
        }
        run("upper triangular", from_rows(R.n, up));
        run("lower triangular", from_rows(R.n, lo));
    }
    if (g_fail) {
        std::fprintf(stderr, "%d check(s) failed\n", g_fail);
        return 1;
    }
    std::printf("all layout checks passed\n");
    return 0;
}
