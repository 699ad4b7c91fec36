// Drop-in demonstration: the reference's AFIIM driver flow
// (AFIIM/example/example.cpp:10-140 -- read A and b, build AFIIM, solve with
// EPBiCGSTAB, write x) linked against the MI355X engine instead of MKL.  The
// solve call is the reference's own signature, provided by cm_dropin.hpp.
//
// usage: afiim_solve A.mtx b.mtx|none dtol [x.out]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "cm_dropin.hpp"

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::cout << "usage: " << argv[0] << " A.mtx b.mtx|none dtol [x.out]\n";
        return 1;
    }
    int n = 0, nnz = 0, sym = 0;
    int *Ai = nullptr, *Aj = nullptr;
    double *Av = nullptr, *B = nullptr;
    if (cm_read_mtx(argv[1], 0, &n, &nnz, &Ai, &Aj, &Av, &sym) != 0) {
        std::cout << "cannot read " << argv[1] << "\n";
        return 1;
    }
    if (std::strcmp(argv[2], "none") == 0) {
        // b = A * 1, as the reference's model right-hand side (AFIIM/src/misc.cpp:51-74)
        std::vector<double> ones(n, 1.0);
        B = new double[n];
        cm_csr_mv_host(n, Ai, Aj, Av, ones.data(), B);
    } else {
        int m = 0;
        if (cm_read_mtx_vector(argv[2], &m, &B) != 0 || m != n) {
            std::cout << "Coefficient matrix and right hand side have different dimensions\n";
            return 1;
        }
    }
    std::cout << "Coefficient Matrix A with n = " << n << " and nnz(A) = " << Ai[n] << "\n\n";
    double *Gv, *Hv, *IDv;
    int *Gi, *Gj, *Hi, *Hj;
    const double dtol = std::atof(argv[3]);
    const int NMAX = 500;
    const double tol = 1e-8;
    std::vector<double> x(n, 0.0);
    auto t0 = std::chrono::steady_clock::now();
    afiim(n, Av, Aj, Ai, &Gv, &Gj, &Gi, &IDv, &Hv, &Hj, &Hi, dtol, 10, 5, 0.0, 0.0);
    auto t1 = std::chrono::steady_clock::now();
    std::cout << "Elapsed time for computation of AFIIM: " << std::chrono::duration<double>(t1 - t0).count()
              << " seconds\n";
    std::cout << "Nonzero elements in preconditioner (nnz(G)+nnz(H)+nnz(D)): " << n + Gi[n] + Hi[n] << "\n\n";
    double resval = 0.0;
    int iter = 0, info = 0;
    t0 = std::chrono::steady_clock::now();
    EPBiCGSTAB(n, Av, Aj, Ai, B, tol, NMAX, Gv, Gj, Gi, IDv, Hv, Hj, Hi, x.data(), &resval, &iter, 1, &info);
    t1 = std::chrono::steady_clock::now();
    std::cout << "Elapsed time for computation of AFIIM-EPBiCSTAB: "
              << std::chrono::duration<double>(t1 - t0).count() << " seconds\n";
    std::cout << "iter " << iter << " info " << info << " resval " << resval << "\n";
    if (argc > 4) {
        FILE *f = std::fopen(argv[4], "w");
        if (f) {
            std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%d 1\n", n);
            for (int i = 0; i < n; i++) std::fprintf(f, "%20.19g\n", x[i]);
            std::fclose(f);
        }
    }
    cm_free_d(Av); cm_free_i(Ai); cm_free_i(Aj); cm_free_d(B);
    cm_free_d(Gv); cm_free_i(Gi); cm_free_i(Gj); cm_free_d(Hv); cm_free_i(Hi); cm_free_i(Hj); cm_free_d(IDv);
    return info == 0 ? 0 : 2;
}
