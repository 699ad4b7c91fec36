// Drop-in demonstration: the reference's ACIFIM driver flow
// (ACIFIM/example/example.cpp:17-214 -- read A and b, build ACIFIM(eps1,
// eps2) with the example's construction defaults, solve with PBiCGSTAB in
// apply mode E, C, I or J with nr relaxation steps, write x) linked against
// the MI355X engine instead of MKL.  The factor and solve calls are the
// reference's own signatures, provided by cm_dropin.hpp.
//
// usage: acifim_solve A.mtx b.mtx|none dtol1 dtol2 nr mode [x.out]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "cm_dropin.hpp"

int main(int argc, char **argv)
{
    if (argc < 7) {
        std::cout << "usage: " << argv[0] << " A.mtx b.mtx|none dtol1 dtol2 nr mode [x.out]\n";
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
        // b = A * 1 (the reference's form_model_rhs)
        std::vector<double> ones(n, 1.0);
        B = new double[n];
        cm_csr_mv_host(n, Ai, Aj, Av, ones.data(), B);
    } else {
        int m = 0;
        if (cm_read_mtx_vector(argv[2], &m, &B) != 0 || m != n) {
            std::cout << "Coefficient matrix and right hand side have different dimensions";
            return 1;
        }
    }
    std::cout << "Coefficient Matrix A with n = " << n << " and nnz(A) = " << Ai[n] << "\n\n";
    const double dtol1 = std::atof(argv[3]), dtol2 = std::atof(argv[4]);
    const int nr = std::atoi(argv[5]);
    const char mode = argv[6][0];
    // construction and solve settings of ACIFIM/example/example.cpp:45-52
    const int elemPerRowCol = 30, growth = 5, element_limit = 200, filtit = 30, NMAX = 2000, verbose = 0;
    const double eta = 0.0, shift = 0.0, coeff = 2.0, tol = 1e-8;
    double *Gv, *Hv, *IDv, *Lv, *Uv;
    int *Gi, *Gj, *Hi, *Hj, *Li, *Lj, *Ui, *Uj;
    auto t0 = std::chrono::steady_clock::now();
    const int cerr = acifim(n, Av, Aj, Ai, &Gv, &Gj, &Gi, &IDv, &Hv, &Hj, &Hi, &Lv, &Lj, &Li, &Uv, &Uj, &Ui, dtol1,
                            dtol2, elemPerRowCol, growth, eta, shift, coeff, element_limit, filtit);
    auto t1 = std::chrono::steady_clock::now();
    if (cerr != 0) {
        std::cout << "Formation of preconditioner failed try changing the choices of parameters\n";
        return 0;
    }
    std::cout << "Elapsed time for computation of AFIIM          : " << std::chrono::duration<double>(t1 - t0).count()
              << " seconds\n";
    std::cout << "Nonzero elements in LDU preconditioner         : " << -n + Li[n] + Ui[n] << "\n";
    std::cout << "Nonzero elements in GD^{-1}H preconditioner    : " << -n + Gi[n] + Hi[n] << "\n";
    std::vector<double> x(n, 0.0);
    double resval = 0.0;
    int iter = 0, info = 0;
    t0 = std::chrono::steady_clock::now();
    PBiCGSTAB(n, Av, Aj, Ai, B, tol, NMAX, Gv, Gj, Gi, IDv, Hv, Hj, Hi, Lv, Lj, Li, Uv, Uj, Ui, nr, mode, x.data(),
              &resval, &iter, verbose, &info);
    t1 = std::chrono::steady_clock::now();
    std::cout << "Elapsed time for computation of ACIFIM-PBiCSTAB: "
              << std::chrono::duration<double>(t1 - t0).count() << " seconds\n";
    std::cout << "iter " << iter << " info " << info << " resval " << resval << " mode " << mode << " nr " << nr
              << "\n";
    if (argc > 7) {
        FILE *f = std::fopen(argv[7], "w");
        if (f) {
            std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%d 1\n", n);
            for (int i = 0; i < n; i++) std::fprintf(f, "%20.19g\n", x[i]);
            std::fclose(f);
        }
    }
    cm_free_d(Av); cm_free_i(Ai); cm_free_i(Aj); cm_free_d(B);
    cm_free_d(Gv); cm_free_i(Gi); cm_free_i(Gj); cm_free_d(Hv); cm_free_i(Hi); cm_free_i(Hj); cm_free_d(IDv);
    cm_free_d(Lv); cm_free_i(Li); cm_free_i(Lj); cm_free_d(Uv); cm_free_i(Ui); cm_free_i(Uj);
    return info == 0 ? 0 : 2;
}
