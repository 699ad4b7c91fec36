// cm_dropin.hpp -- the reference's solve entry points, same names, arguments
// and meaning, running on the MI355X engine (libcmsolve.so).
//
// Include this header instead of the package's bicgstab.hpp and link
// libcmsolve.so (+ libcmhost.so for the factor producers) instead of MKL:
//   AFIIM : EPBiCGSTAB(n, Av, Aj, Ai, b, tol, NMAX, Gv, Gj, Gi, IDv, Hv, Hj, Hi,
//                      x, &resval, &iter, verbose, &info)
//           (AFIIM/src/bicgstab.hpp:82-100)
//   PDFIIM: EPBiCGSTAB(A, b, tol, NMAX, G, ID, H, x, &resval, &iter, verbose,
//                      &info)  with any sparseMatrix-like type exposing
//           r, mtype, i (row offsets), j (columns), v (values) as std::vector
//           (PDFIIM/src/bicgstab.hpp:75-86; mtype == 1 is SYMMETRIC)
//   ACIFIM: PBiCGSTAB(n, Av, Aj, Ai, b, tol, NMAX, Gv, Gj, Gi, IDv, Hv, Hj, Hi,
//                     Lv, Lj, Li, Uv, Uj, Ui, nr, mode, x, &resval, &iter,
//                     verbose, &info)  (ACIFIM/src/bicgstab.hpp:94-120)
// and the factor producers afiim() / acifim() / pdfiim() with the reference
// signatures (AFIIM/src/prec.hpp:99-114, ACIFIM/src/prec.hpp:126-151,
// PDFIIM/src/parPrec.hpp:86-96).
//
// The reference functions return void and never fail short of std::bad_alloc;
// an engine error (no GPU, HIP failure, bad CSR) is raised here as
// std::runtime_error carrying the engine's message.
#ifndef CM_DROPIN_HPP
#define CM_DROPIN_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "cm_host.h"
#include "cm_solve.h"

namespace cm_dropin_detail {
inline void check(int rc)
{
    if (rc != CM_OK) throw std::runtime_error(std::string("cm solve engine: ") + cm_last_error());
}
}  // namespace cm_dropin_detail

inline void EPBiCGSTAB(int n, double *Av, int *Aj, int *Ai, double *b, double tol, int NMAX, double *Gv, int *Gj,
                       int *Gi, double *IDv, double *Hv, int *Hj, int *Hi, double *x, double *resval, int *iter,
                       int verbose, int *info)
{
    cm_dropin_detail::check(cm_afiim_epbicgstab(n, Av, Aj, Ai, b, tol, NMAX, Gv, Gj, Gi, IDv, Hv, Hj, Hi, x, resval,
                                                iter, verbose, info));
}

template <class SparseMatrix>
inline void EPBiCGSTAB(SparseMatrix &A, std::vector<double> &b, double tol, int NMAX, SparseMatrix &G,
                       std::vector<double> &ID, SparseMatrix &H, std::vector<double> &x, double *resval, int *iter,
                       int verbose, int *info)
{
    const int sym = (A.mtype == 1);  // SYMMETRIC, PDFIIM/src/constants.hpp:10
    cm_dropin_detail::check(cm_pdfiim_epbicgstab(
        (int)A.r, sym, A.i.data(), A.j.data(), A.v.data(), b.data(), tol, NMAX, G.i.data(), G.j.data(),
        G.v.data(), ID.data(), sym ? nullptr : H.i.data(), sym ? nullptr : H.j.data(),
        sym ? nullptr : H.v.data(), x.data(), resval, iter, verbose, info));
}

inline void PBiCGSTAB(int n, double *Av, int *Aj, int *Ai, double *b, double tol, int NMAX, double *Gv, int *Gj,
                      int *Gi, double *IDv, double *Hv, int *Hj, int *Hi, double *Lv, int *Lj, int *Li, double *Uv,
                      int *Uj, int *Ui, int nr, char mode, double *x, double *resval, int *iter, int verbose,
                      int *info)
{
    cm_dropin_detail::check(cm_acifim_pbicgstab(n, Av, Aj, Ai, b, tol, NMAX, Gv, Gj, Gi, IDv, Hv, Hj, Hi, Lv, Lj, Li,
                                                Uv, Uj, Ui, nr, mode, x, resval, iter, verbose, info));
}

inline void afiim(int n, double *Av, int *Aj, int *Ai, double **Gval, int **Gindj, int **Gindi, double **IDv,
                  double **Hval, int **Hindj, int **Hindi, double dtol, int elemPerRowCol, int growth, double eta,
                  double shift)
{
    cm_afiim(n, Av, Aj, Ai, Gval, Gindj, Gindi, IDv, Hval, Hindj, Hindi, dtol, elemPerRowCol, growth, eta, shift);
}

inline int acifim(int n, double *Av, int *Aj, int *Ai, double **Gval, int **Gindj, int **Gindi, double **IDv,
                  double **Hval, int **Hindj, int **Hindi, double **Lval, int **Lindj, int **Lindi, double **Uval,
                  int **Uindj, int **Uindi, double dtol1, double dtol2, int elemPerRowCol, int growth, double eta,
                  double shift, double coeff, int element_limit, int filtit)
{
    return cm_acifim(n, Av, Aj, Ai, Gval, Gindj, Gindi, IDv, Hval, Hindj, Hindi, Lval, Lindj, Lindi, Uval, Uindj,
                     Uindi, dtol1, dtol2, elemPerRowCol, growth, eta, shift, coeff, element_limit, filtit);
}

// PDFIIM factors into sparseMatrix-like G (CSR, UPPERTRI) and H (CSR,
// LOWERTRI) and the inverse diagonal D, from A's stored CSR arrays, as the
// reference's pdfiim() fills them (PDFIIM/src/parPrec.cpp:536-587).  Every
// host thread is used (the reference's OpenMP default).
template <class SparseMatrix>
inline void pdfiim(SparseMatrix &A, SparseMatrix &G, std::vector<double> &D, SparseMatrix &H, int prefill,
                   int lfill, int postfill, double ftol, int warpsize, char filttype, char patt)
{
    const int n = (int)A.r;
    double *Gv = nullptr, *Hv = nullptr, *IDv = nullptr;
    int *Gi = nullptr, *Gj = nullptr, *Hi = nullptr, *Hj = nullptr;
    if (cm_pdfiim(n, A.v.data(), A.j.data(), A.i.data(), &Gv, &Gj, &Gi, &IDv, &Hv, &Hj, &Hi, prefill, lfill,
                  postfill, ftol, warpsize, filttype, patt, 0) != 0)
        throw std::runtime_error("cm pdfiim: unsupported arguments (patt 'L'/'U' need lfill == 1)");
    auto fill = [n](SparseMatrix &M, int mtype, double *v, int *j, int *i) {
        M.stype = 0;  // CSR (PDFIIM/src/constants.hpp:6)
        M.mtype = mtype;
        M.r = n;
        M.c = n;
        M.nnz = i[n];
        M.i.assign(i, i + n + 1);
        M.j.assign(j, j + i[n]);
        M.v.assign(v, v + i[n]);
        cm_free_d(v);
        cm_free_i(j);
        cm_free_i(i);
    };
    fill(G, 2, Gv, Gj, Gi);  // UPPERTRI
    fill(H, 3, Hv, Hj, Hi);  // LOWERTRI
    D.assign(IDv, IDv + n);
    cm_free_d(IDv);
}

#endif
