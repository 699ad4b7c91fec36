// TEST INFRASTRUCTURE ONLY -- not part of the product.
//
// extern "C" entry points around the reference's own AFIIM construction and
// Matrix Market readers, compiled from the unmodified reference sources by
// oracle/Makefile into oracle/_ref/.  Used by tests as the checker for the
// product's factor construction and to read the reference's fixtures.
//   afiim             AFIIM/src/prec.hpp:99-114, AFIIM/src/prec.cpp:101-767
//   read_coeff_matrix AFIIM/src/rw_matrix.cpp:62-216
//   read_rhs_vector   AFIIM/src/rw_matrix.cpp:246-309
#include "prec.hpp"
#include "rw_matrix.hpp"

extern "C" {

void ref_afiim(int n, double *Av, int *Aj, int *Ai, double **Gv, int **Gj,
               int **Gi, double **IDv, double **Hv, int **Hj, int **Hi,
               double dtol, int elemPerRowCol, int growth, double eta,
               double shift)
{
    afiim(n, Av, Aj, Ai, Gv, Gj, Gi, IDv, Hv, Hj, Hi, dtol, elemPerRowCol,
          growth, eta, shift);
}

void ref_read_coeff_matrix(const char *fn, int *n, double **Av, int **Aj, int **Ai)
{
    read_coeff_matrix(fn, n, Av, Aj, Ai);
}

void ref_read_rhs_vector(const char *fn, int *n, double **B)
{
    read_rhs_vector(fn, n, B);
}

void ref_free_d(double *p) { delete[] p; }
void ref_free_i(int *p) { delete[] p; }

}
