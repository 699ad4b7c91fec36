// TEST INFRASTRUCTURE ONLY -- not part of the product.
//
// extern "C" entry point around the reference's own ACIFIM construction,
// compiled from the unmodified reference sources by oracle/Makefile into
// oracle/_ref/.  Used by tests as the checker for the product's construction.
//   acifim   ACIFIM/src/prec.hpp:126-151, ACIFIM/src/prec.cpp:127-1052
#include "prec.hpp"

extern "C" {

int ref_acifim(int n, double *Av, int *Aj, int *Ai, double **Gv, int **Gj,
               int **Gi, double **IDv, double **Hv, int **Hj, int **Hi,
               double **Lv, int **Lj, int **Li, double **Uv, int **Uj, int **Ui,
               double dtol1, double dtol2, int elemPerRowCol, int growth,
               double eta, double shift, double coeff, int element_limit,
               int filtit)
{
    return acifim(n, Av, Aj, Ai, Gv, Gj, Gi, IDv, Hv, Hj, Hi, Lv, Lj, Li, Uv,
                  Uj, Ui, dtol1, dtol2, elemPerRowCol, growth, eta, shift, coeff,
                  element_limit, filtit);
}

void ref_free_d(double *p) { delete[] p; }
void ref_free_i(int *p) { delete[] p; }

}
