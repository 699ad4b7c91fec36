// Matrix Market ingest for the solve engine's drivers and fixtures.
//
// Accepts what the reference's readers accept -- "coordinate real
// general|symmetric" matrices and "array real general" single-column vectors
// (AFIIM/src/rw_matrix.cpp:62-216, :246-309; PDFIIM/src/rw_matrix.cpp:56-179)
// -- and produces the same CSR: 0-based, rows sorted by column.  A symmetric
// file is either expanded to both triangles (AFIIM/ACIFIM readers) or kept as
// one triangle read with (I,J) swapped, so that a lower-triangle file becomes
// the upper triangle (PDFIIM reader).
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cm_host.h"
#include "factor_common.hpp"

namespace {

struct Banner {
    bool coordinate = false, array = false, real = false, general = false, symmetric = false;
};

bool read_banner(FILE *f, Banner &b)
{
    char line[1100];
    if (!fgets(line, sizeof line, f)) return false;
    std::string s(line);
    for (auto &ch : s) ch = (char)std::tolower((unsigned char)ch);
    if (s.rfind("%%matrixmarket", 0) != 0) return false;
    b.coordinate = s.find(" coordinate") != std::string::npos;
    b.array = s.find(" array") != std::string::npos;
    b.real = s.find(" real") != std::string::npos;
    b.general = s.find(" general") != std::string::npos;
    b.symmetric = s.find(" symmetric") != std::string::npos;
    return true;
}

// Skips comment lines; leaves the file positioned after the size line.
bool read_size_line(FILE *f, char *line, size_t cap)
{
    while (fgets(line, (int)cap, f)) {
        if (line[0] == '%') continue;
        const char *p = line;
        while (*p && std::isspace((unsigned char)*p)) p++;
        if (*p) return true;
    }
    return false;
}

}  // namespace

extern "C" int cm_read_mtx(const char *path, int symmetric_mode, int *n, int *nnz_out,
                           int **Ai, int **Aj, double **Av, int *is_symmetric)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    Banner b;
    if (!read_banner(f, b)) {
        fclose(f);
        return -2;
    }
    if (!b.coordinate || !b.real || !(b.general || b.symmetric)) {
        fclose(f);
        return -3;
    }
    char line[1100];
    int M = 0, N = 0, nz = 0;
    if (!read_size_line(f, line, sizeof line) || sscanf(line, "%d %d %d", &M, &N, &nz) != 3) {
        fclose(f);
        return -4;
    }
    if (M != N || M <= 0 || nz < 0) {
        fclose(f);
        return -5;
    }
    std::vector<int> I(nz), J(nz);
    std::vector<double> V(nz);
    for (int k = 0; k < nz; k++) {
        if (fscanf(f, "%d %d %lg", &I[k], &J[k], &V[k]) != 3) {
            fclose(f);
            return -6;
        }
        I[k]--;
        J[k]--;
        if (I[k] < 0 || I[k] >= M || J[k] < 0 || J[k] >= N) {
            fclose(f);
            return -7;
        }
    }
    fclose(f);
    if (b.symmetric) {
        if (symmetric_mode == 0) {
            for (int k = 0; k < nz; k++)
                if (I[k] != J[k]) {
                    I.push_back(J[k]);
                    J.push_back(I[k]);
                    V.push_back(V[k]);
                }
        } else {
            std::swap(I, J);
        }
    }
    const int total = (int)I.size();
    std::vector<int> ptr(M + 1, 0), idx(total);
    std::vector<double> val(total);
    for (int k = 0; k < total; k++) ptr[I[k] + 1]++;
    for (int i = 0; i < M; i++) ptr[i + 1] += ptr[i];
    {
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int k = 0; k < total; k++) {
            int q = fill[I[k]]++;
            idx[q] = J[k];
            val[q] = V[k];
        }
    }
    std::vector<std::pair<int, double>> row;
    for (int i = 0; i < M; i++) {
        row.clear();
        for (int q = ptr[i]; q < ptr[i + 1]; q++) row.push_back({idx[q], val[q]});
        std::stable_sort(row.begin(), row.end(), [](auto &a, auto &c) { return a.first < c.first; });
        for (int q = ptr[i]; q < ptr[i + 1]; q++) {
            idx[q] = row[q - ptr[i]].first;
            val[q] = row[q - ptr[i]].second;
        }
    }
    *n = M;
    *nnz_out = total;
    *Ai = cmh::new_copy(ptr);
    *Aj = cmh::new_copy(idx);
    *Av = cmh::new_copy(val);
    if (is_symmetric) *is_symmetric = b.symmetric ? 1 : 0;
    return 0;
}

extern "C" int cm_read_mtx_vector(const char *path, int *n, double **B)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    Banner b;
    if (!read_banner(f, b)) {
        fclose(f);
        return -2;
    }
    if (!b.array || !b.real) {
        fclose(f);
        return -3;
    }
    char line[1100];
    int M = 0, N = 0;
    if (!read_size_line(f, line, sizeof line) || sscanf(line, "%d %d", &M, &N) != 2) {
        fclose(f);
        return -4;
    }
    if (N != 1 || M <= 0) {
        fclose(f);
        return -5;
    }
    std::vector<double> v(M);
    for (int k = 0; k < M; k++)
        if (fscanf(f, "%lg", &v[k]) != 1) {
            fclose(f);
            return -6;
        }
    fclose(f);
    *n = M;
    *B = cmh::new_copy(v);
    return 0;
}
