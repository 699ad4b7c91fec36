// Drop-in demonstration: the reference's PDFIIM driver flow
// (PDFIIM/example/pdfiim.cpp:29-144 -- read A into a sparseMatrix, b from a
// file or b = A * 1 over the stored entries, build PDFIIM with filter 'I' and
// pattern 'B', solve with EPBiCGSTAB, write x) linked against the MI355X
// engine instead of MKL.  The matrix type is a local struct with the fields
// the reference's sparseMatrix<int,int,double> exposes
// (PDFIIM/src/sparseMatrix.hpp:68-75), so the template overloads of
// cm_dropin.hpp instantiate as they would with the reference's class.  A
// symmetric file is kept as its upper triangle with mtype SYMMETRIC
// (PDFIIM/src/rw_matrix.cpp:101-112), which selects the solve's symmetric
// branch (M = G D^{-1} G^T).
//
// usage: pdfiim_solve A.mtx b.mtx|none prefill lfill postfill ftol warpsize [x.out]
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "cm_dropin.hpp"

struct SpMat {  // CSR, 0-based (PDFIIM/src/constants.hpp:6-12)
    int stype = -1, mtype = -1;
    std::vector<double> v;
    std::vector<int> i, j;
    int r = 0, c = 0, nnz = 0;
};

int main(int argc, char **argv)
{
    if (argc < 8) {
        std::cout << "usage: " << argv[0] << " A.mtx b.mtx|none prefill lfill postfill ftol warpsize [x.out]\n";
        return 1;
    }
    SpMat A, G, H;
    std::vector<double> ID, B;
    {
        int n = 0, nnz = 0, sym = 0;
        int *Ai = nullptr, *Aj = nullptr;
        double *Av = nullptr;
        if (cm_read_mtx(argv[1], 1, &n, &nnz, &Ai, &Aj, &Av, &sym) != 0) {
            std::cout << "cannot read " << argv[1] << "\n";
            return 1;
        }
        A.stype = 0;
        A.mtype = sym ? 1 : 0;  // SYMMETRIC : GENERAL
        A.r = A.c = n;
        A.nnz = Ai[n];
        A.i.assign(Ai, Ai + n + 1);
        A.j.assign(Aj, Aj + Ai[n]);
        A.v.assign(Av, Av + Ai[n]);
        cm_free_i(Ai);
        cm_free_i(Aj);
        cm_free_d(Av);
    }
    const int n = A.r;
    if (std::strcmp(argv[2], "none") == 0) {
        // b = (stored entries of A) * 1, as the reference driver forms it
        // (PDFIIM/example/pdfiim.cpp:63-75)
        B.assign(n, 0.0);
        for (int r = 0; r < n; r++) {
            double s = 0.0;
            for (int k = A.i[r]; k < A.i[r + 1]; k++) s += A.v[k] * 1.0;
            B[r] = s;
        }
    } else {
        int m = 0;
        double *b = nullptr;
        if (cm_read_mtx_vector(argv[2], &m, &b) != 0 || m != n) {
            std::cout << "Coefficient matrix and right hand side have different dimensions";
            return 1;
        }
        B.assign(b, b + n);
        cm_free_d(b);
    }
    std::cout << "Coefficient Matrix A with n = " << n << " and nnz(A) = " << A.nnz << "\n\n";
    // argument clamps of PDFIIM/example/pdfiim.cpp:44-45
    const double ftol = std::max(std::atof(argv[6]), std::numeric_limits<double>::epsilon());
    const int prefill = std::max(std::atoi(argv[3]), 1), lfill = std::max(std::atoi(argv[4]), 1),
              postfill = std::max(std::atoi(argv[5]), 1), warpsize = std::atoi(argv[7]);
    const int NMAX = 1200;
    const double tol = 1e-8;
    std::vector<double> x(n, 0.0);
    auto t0 = std::chrono::steady_clock::now();
    pdfiim(A, G, ID, H, prefill, lfill, postfill, ftol, warpsize, 'I', 'B');
    auto t1 = std::chrono::steady_clock::now();
    std::cout << "Elapsed time for computation of AFIIM: " << std::chrono::duration<double>(t1 - t0).count()
              << " seconds\n";
    std::cout << "Nonzero elements in preconditioner (nnz(G))              : " << G.nnz << "\n";
    std::cout << "Nonzero elements in preconditioner (nnz(H))              : " << H.nnz << "\n";
    double resval = 0.0;
    int iter = 0, info = 0;
    t0 = std::chrono::steady_clock::now();
    EPBiCGSTAB(A, B, tol, NMAX, G, ID, H, x, &resval, &iter, 1, &info);
    t1 = std::chrono::steady_clock::now();
    std::cout << "Elapsed time for computation of PDFIIM-EPBiCSTAB: "
              << std::chrono::duration<double>(t1 - t0).count() << " seconds\n";
    std::cout << "iter " << iter << " info " << info << " resval " << resval << " symmetric " << (A.mtype == 1)
              << "\n";
    if (argc > 8) {
        FILE *f = std::fopen(argv[8], "w");
        if (f) {
            std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%d 1\n", n);
            for (int r = 0; r < n; r++) std::fprintf(f, "%20.19g\n", x[r]);
            std::fclose(f);
        }
    }
    return info == 0 ? 0 : 2;
}
