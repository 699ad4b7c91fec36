// Seeded synthetic systems for the solve engine's workloads.
//
// The reference ships no generator (its only fixture is HB/sherman2), so the
// workload definitions live here.  Every random draw comes from a
// counter-based splitmix64 stream keyed by (seed, purpose, index), so a row's
// values do not depend on the order or the thread in which rows are built,
// and every path (CPU reference loop, FP64 GPU, mixed precision) sees the same
// bits.  Normals use Box-Muller (cos branch) on two 53-bit uniforms.
//
//   cm_gen_convdiff : d-dimensional (2 or 3) cell-centred convection-diffusion
//     on an N^d grid, homogeneous Dirichlet boundary eliminated.  Cell
//     diffusivity kappa = exp(z), z ~ N(0,1); face coefficient = harmonic mean
//     of the two cells (the cell's own kappa on a boundary face, kept on the
//     diagonal); first-order upwind for velocity (1,..,1) with per-axis
//     coefficient c = 2 Pe.  Row: lower neighbour on each axis -(k_f + c),
//     upper neighbour -k_f, diagonal sum(k_f) + d c.  A nonsymmetric
//     M-matrix.  Rows ordered x fastest, columns ascending.
//   cm_gen_stencil27 : 27-point stencil on N^3, off-diagonals
//     -(1 + 0.3 U[-1,1]); a fraction of rows gets U{lo..hi} extra distinct
//     columns (not on the stencil) valued -0.01 U[0,1]; diagonal = 1.05 x the
//     row's absolute off-diagonal sum.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "cm_host.h"

namespace {

inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Stream {
    uint64_t s;
    inline uint64_t next()
    {
        s += 0x9E3779B97F4A7C15ULL;
        return mix64(s);
    }
    inline double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
    inline double normal()
    {
        const double u1 = 1.0 - u01();  // (0, 1]
        const double u2 = u01();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

enum : uint64_t { KEY_KAPPA = 1, KEY_ROW27 = 2, KEY_UNIFORM = 3 };

inline Stream stream(uint64_t seed, uint64_t purpose, uint64_t index)
{
    return Stream{mix64(seed ^ mix64(purpose * 0x632BE59BD9B4E019ULL + mix64(index + 1)))};
}

}  // namespace

extern "C" int cm_gen_convdiff(int dim, int N, double peclet, unsigned long long seed,
                               int *n_out, long long *nnz_out, int *Ai, int *Aj, double *Av)
{
    if ((dim != 2 && dim != 3) || N < 1) return -1;
    const long long n = (dim == 2) ? (long long)N * N : (long long)N * N * N;
    const long long nnz = (2LL * dim + 1) * n - 2LL * dim * (n / N);
    if (n > 0x7fffffffLL || nnz > 0x7fffffffLL) return -2;
    *n_out = (int)n;
    *nnz_out = nnz;
    if (!Ai || !Aj || !Av) return 0;
    std::vector<double> kappa(n);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; i++) {
        Stream st = stream(seed, KEY_KAPPA, (uint64_t)i);
        kappa[i] = std::exp(st.normal());
    }
    const double c = 2.0 * peclet;
    const long long stride[3] = {1, N, (long long)N * N};
    // Row pointers: every row has 1 + (number of existing neighbours).
    Ai[0] = 0;
    for (long long i = 0; i < n; i++) {
        long long rem = i;
        int cnt = 1;
        for (int a = 0; a < dim; a++) {
            int x = (int)(rem % N);
            rem /= N;
            cnt += (x > 0) + (x < N - 1);
        }
        Ai[i + 1] = Ai[i] + cnt;
    }
    auto face = [&](long long i, long long j) {
        const long long lo = std::min(i, j), hi = std::max(i, j);
        const double a = kappa[lo], b = kappa[hi];
        return 2.0 * a * b / (a + b);
    };
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; i++) {
        int coord[3] = {0, 0, 0};
        long long rem = i;
        for (int a = 0; a < dim; a++) {
            coord[a] = (int)(rem % N);
            rem /= N;
        }
        double kf_lo[3], kf_hi[3];
        for (int a = 0; a < dim; a++) {
            kf_lo[a] = coord[a] > 0 ? face(i, i - stride[a]) : kappa[i];
            kf_hi[a] = coord[a] < N - 1 ? face(i, i + stride[a]) : kappa[i];
        }
        double diag = 0.0;
        for (int a = 0; a < dim; a++) diag += kf_lo[a] + kf_hi[a] + c;
        int q = Ai[i];
        for (int a = dim - 1; a >= 0; a--)
            if (coord[a] > 0) {
                Aj[q] = (int)(i - stride[a]);
                Av[q] = -(kf_lo[a] + c);
                q++;
            }
        Aj[q] = (int)i;
        Av[q] = diag;
        q++;
        for (int a = 0; a < dim; a++)
            if (coord[a] < N - 1) {
                Aj[q] = (int)(i + stride[a]);
                Av[q] = -kf_hi[a];
                q++;
            }
    }
    return 0;
}

namespace {

// Number of extra long-range entries of row i (0 for most rows); drawn first
// from the row's own stream so the count can be recomputed cheaply.
inline int extra_count(Stream &st, double frac, int lo, int hi)
{
    const double u = st.u01();
    const double w = st.u01();
    if (u >= frac) return 0;
    return lo + (int)std::floor(w * (hi - lo + 1));
}

}  // namespace

extern "C" int cm_gen_stencil27(int N, double extra_frac, int extra_lo, int extra_hi,
                                unsigned long long seed, int *n_out, long long *nnz_out,
                                int *Ai, int *Aj, double *Av)
{
    if (N < 1 || extra_lo < 0 || extra_hi < extra_lo) return -1;
    const long long n = (long long)N * N * N;
    if (n > 0x7fffffffLL) return -2;
    std::vector<int> rowlen(n);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; i++) {
        const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / ((long long)N * N));
        auto span = [&](int v) { return 1 + (v > 0) + (v < N - 1); };
        Stream st = stream(seed, KEY_ROW27, (uint64_t)i);
        int ex = extra_count(st, extra_frac, extra_lo, extra_hi);
        if (ex > n - 27) ex = (int)std::max(0LL, n - 27);
        rowlen[i] = span(x) * span(y) * span(z) + ex;
    }
    long long nnz = 0;
    for (long long i = 0; i < n; i++) nnz += rowlen[i];
    if (nnz > 0x7fffffffLL) return -2;
    *n_out = (int)n;
    *nnz_out = nnz;
    if (!Ai || !Aj || !Av) return 0;
    Ai[0] = 0;
    for (long long i = 0; i < n; i++) Ai[i + 1] = Ai[i] + rowlen[i];
#pragma omp parallel
    {
        std::vector<std::pair<int, double>> row;
        std::vector<int> cols;
#pragma omp for schedule(dynamic, 1024)
        for (long long i = 0; i < n; i++) {
            const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / ((long long)N * N));
            Stream st = stream(seed, KEY_ROW27, (uint64_t)i);
            int ex = extra_count(st, extra_frac, extra_lo, extra_hi);
            if (ex > n - 27) ex = (int)std::max(0LL, n - 27);
            row.clear();
            cols.clear();
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        if (x + dx < 0 || x + dx >= N || y + dy < 0 || y + dy >= N || z + dz < 0 ||
                            z + dz >= N)
                            continue;
                        const long long j = i + dx + (long long)N * (dy + (long long)N * dz);
                        cols.push_back((int)j);
                        if (j == i) continue;
                        row.push_back({(int)j, -(1.0 + 0.3 * (2.0 * st.u01() - 1.0))});
                    }
            std::sort(cols.begin(), cols.end());
            for (int e = 0; e < ex; e++) {
                int j;
                for (;;) {
                    j = (int)(st.next() % (uint64_t)n);
                    auto it = std::lower_bound(cols.begin(), cols.end(), j);
                    if (it == cols.end() || *it != j) {
                        cols.insert(it, j);
                        break;
                    }
                }
                row.push_back({j, -0.01 * st.u01()});
            }
            std::sort(row.begin(), row.end(), [](auto &a, auto &b) { return a.first < b.first; });
            double asum = 0.0;
            for (auto &e : row) asum += std::fabs(e.second);
            int q = Ai[i];
            bool placed = false;
            for (auto &e : row) {
                if (!placed && e.first > i) {
                    Aj[q] = (int)i;
                    Av[q] = 1.05 * asum;
                    q++;
                    placed = true;
                }
                Aj[q] = e.first;
                Av[q] = e.second;
                q++;
            }
            if (!placed) {
                Aj[q] = (int)i;
                Av[q] = 1.05 * asum;
            }
        }
    }
    return 0;
}

extern "C" void cm_gen_uniform(unsigned long long seed, long long count, double *x)
{
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < count; k++) {
        Stream st = stream(seed, KEY_UNIFORM, (uint64_t)k);
        x[k] = 2.0 * st.u01() - 1.0;
    }
}

extern "C" void cm_csr_mv_host(int n, const int *Ai, const int *Aj, const double *Av,
                               const double *x, double *y)
{
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int k = Ai[i]; k < Ai[i + 1]; k++) s += Av[k] * x[Aj[k]];
        y[i] = s;
    }
}
