/*
 * cm_host.h -- host-side C ABI of the solve engine (libcmhost.so).
 *
 * Factor producers (run once per system on the host; the GPU solve consumes
 * their output), Matrix Market ingest, and the seeded workload generators.
 * All matrices are CSR, 0-based, int32 indices, columns ascending.
 * Arrays returned through double / int out-pointers are allocated with C++
 * new[]; release them with cm_free_d / cm_free_i (or delete[] from C++), as
 * the reference's callers do (AFIIM/example/example.cpp:127-137).
 */
#ifndef CM_HOST_H
#define CM_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* AFIIM factors M = G diag(IDv) H.  Same parameters and outputs as the
 * reference afiim() (AFIIM/src/prec.hpp:99-114, AFIIM/src/prec.cpp:101-767);
 * elemPerRowCol and growth are allocation hints and do not change results. */
void cm_afiim(int n, const double *Av, const int *Aj, const int *Ai,
              double **Gval, int **Gindj, int **Gindi, double **IDv,
              double **Hval, int **Hindj, int **Hindi, double dtol,
              int elemPerRowCol, int growth, double eta, double shift);

/* ACIFIM factors (G, IDv, H) plus the incomplete L and U.  Same parameters,
 * outputs and return value (0 ok, 1 NaN Schur complement) as the reference
 * acifim() (ACIFIM/src/prec.hpp:126-151, ACIFIM/src/prec.cpp:127-1052). */
int cm_acifim(int n, const double *Av, const int *Aj, const int *Ai,
              double **Gval, int **Gindj, int **Gindi, double **IDv,
              double **Hval, int **Hindj, int **Hindi,
              double **Lval, int **Lindj, int **Lindi,
              double **Uval, int **Uindj, int **Uindi,
              double dtol1, double dtol2, int elemPerRowCol, int growth,
              double eta, double shift, double coeff, int element_limit,
              int filtit);

/* PDFIIM factors (PDFIIM/src/parPrec.hpp:26-98, PDFIIM/src/parPrec.cpp:95-588)
 * for a general CSR matrix A.  filttype '1' mean, '2' euclidean, anything else
 * inf-norm; patt 'B' (both parts; 'L'/'U' are accepted only with lfill == 1,
 * see the implementation).  nthreads <= 0 uses every available thread.
 * Returns 0, or -1 for unsupported arguments. */
int cm_pdfiim(int n, const double *Av, const int *Aj, const int *Ai,
              double **Gval, int **Gindj, int **Gindi, double **IDv,
              double **Hval, int **Hindj, int **Hindi,
              int prefill, int lfill, int postfill, double ftol, int warpsize,
              char filttype, char patt, int nthreads);

void cm_free_d(double *p);
void cm_free_i(int *p);

/* Matrix Market ingest (AFIIM/src/rw_matrix.cpp:62-216, :246-309).
 * symmetric_mode 0: a symmetric file is expanded to both triangles (AFIIM /
 * ACIFIM reader); 1: only the stored triangle is kept, transposed to the upper
 * triangle, as PDFIIM's reader does (PDFIIM/src/rw_matrix.cpp:101-112).
 * *is_symmetric reports the banner.  Returns 0 or a negative error code. */
int cm_read_mtx(const char *path, int symmetric_mode, int *n, int *nnz,
                int **Ai, int **Aj, double **Av, int *is_symmetric);
int cm_read_mtx_vector(const char *path, int *n, double **B);

/* Seeded synthetic systems (the splitmix64 / Box-Muller generator described in
 * the engine's design notes).  Two-call protocol: with Ai/Aj/Av == NULL the
 * call only reports n and nnz; then call again with arrays of that size. */
int cm_gen_convdiff(int dim, int N, double peclet, unsigned long long seed,
                    int *n, long long *nnz, int *Ai, int *Aj, double *Av);
int cm_gen_stencil27(int N, double extra_frac, int extra_lo, int extra_hi,
                     unsigned long long seed, int *n, long long *nnz,
                     int *Ai, int *Aj, double *Av);
/* x[k] ~ U[-1, 1] from the seed, for k < n*ncols (column-major). */
void cm_gen_uniform(unsigned long long seed, long long count, double *x);
/* b = A x in a fixed sequential order (shared bitwise by every path). */
void cm_csr_mv_host(int n, const int *Ai, const int *Aj, const double *Av,
                    const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
