/*
 * cm_solve.h -- C ABI of the MI355X BiCGSTAB solve engine (libcmsolve.so).
 *
 * The three drop-in entry points take the same arguments, in the same order
 * and with the same meaning, as the reference solve functions they replace,
 * and add an int status (CM_OK or a negative CM_ERR_*).  They read A, b and
 * the factors, write only x (in/out), *resval, *iter and *info, and keep no
 * pointer past return.  info: 0 converged, 1 diverged, 2 stagnated, 3 NMAX
 * reached or non-finite alpha (ACIFIM).  The C++ overloads with the
 * reference's exact names live in cm_dropin.hpp.
 */
#ifndef CM_SOLVE_H
#define CM_SOLVE_H

#ifdef __cplusplus
extern "C" {
#endif

#define CM_OK 0
#define CM_ERR_HIP (-1)          /* HIP runtime error (message: cm_last_error) */
#define CM_ERR_ARG (-2)          /* invalid argument */
#define CM_ERR_UNSUPPORTED (-3)  /* option not available in this build */
#define CM_ERR_NODEV (-4)        /* no usable GPU */

/* Replaces AFIIM EPBiCGSTAB (AFIIM/src/bicgstab.hpp:82-100,
 * AFIIM/src/bicgstab.cpp:87-297): M = G diag(IDv) H, explicit apply. */
int cm_afiim_epbicgstab(int n, double *Av, int *Aj, int *Ai, double *b, double tol,
                        int NMAX, double *Gv, int *Gj, int *Gi, double *IDv,
                        double *Hv, int *Hj, int *Hi, double *x, double *resval,
                        int *iter, int verbose, int *info);

/* Replaces PDFIIM EPBiCGSTAB (PDFIIM/src/bicgstab.hpp:75-86,
 * PDFIIM/src/bicgstab.cpp:80-310), with sparseMatrix<int,int,double>
 * flattened to its CSR arrays.  a_symmetric != 0 means A holds only the upper
 * triangle of a symmetric matrix (mtype SYMMETRIC) and M = G D^{-1} G^T; H is
 * then not read and may be NULL. */
int cm_pdfiim_epbicgstab(int n, int a_symmetric, const int *Ai, const int *Aj,
                         const double *Av, const double *b, double tol, int NMAX,
                         const int *Gi, const int *Gj, const double *Gv,
                         const double *IDv, const int *Hi, const int *Hj,
                         const double *Hv, double *x, double *resval, int *iter,
                         int verbose, int *info);

/* Replaces ACIFIM PBiCGSTAB (ACIFIM/src/bicgstab.hpp:94-120,
 * ACIFIM/src/bicgstab.cpp:97-327) with apply() modes 'E' (explicit
 * G D^{-1} H), 'C' (nr Richardson sweeps with L and U), 'I' (L, U triangular
 * solves with D^{-1}) and 'J' (L, U solves) (ACIFIM/src/bicgstab.cpp:372-428). */
int cm_acifim_pbicgstab(int n, double *Av, int *Aj, int *Ai, double *b, double tol,
                        int NMAX, double *Gv, int *Gj, int *Gi, double *IDv,
                        double *Hv, int *Hj, int *Hi, double *Lv, int *Lj, int *Li,
                        double *Uv, int *Uj, int *Ui, int nr, char mode, double *x,
                        double *resval, int *iter, int verbose, int *info);

/* PDFIIM factors built on the GPU: same parameters, outputs (new[] arrays,
 * release with cm_free_d / cm_free_i or delete[]) and meaning as cm_pdfiim
 * in cm_host.h, i.e. the reference's pdfiim() (PDFIIM/src/parPrec.hpp:86-96,
 * PDFIIM/src/parPrec.cpp:95-588), and the same factors bit for bit as the
 * host producer.  One wavefront builds one row (pattern growth, the dense
 * LU solves and the filters), with the dense system in LDS for patterns of
 * up to 64 indices and in global memory up to 256 (CM_ERR_UNSUPPORTED
 * beyond, naming the row).  patt 'L'/'U' need
 * lfill == 1 and warpsize must be -1 or positive, as in cm_pdfiim.
 * device_ms (optional) receives the construction kernels' device time. */
int cm_pdfiim_device(int n, const double *Av, const int *Aj, const int *Ai,
                     double **Gval, int **Gindj, int **Gindi, double **IDv,
                     double **Hval, int **Hindj, int **Hindi, int prefill, int lfill,
                     int postfill, double ftol, int warpsize, char filttype, char patt,
                     double *device_ms);

/* ---- resident systems (A and factors kept in HBM across solves) ---- */

typedef struct cm_system_s *cm_system;

#define CM_VARIANT_AFIIM 0
#define CM_VARIANT_PDFIIM 1
#define CM_VARIANT_ACIFIM 2

/* Uploads A and the factors once.  variant selects the reference loop whose
 * constants and exit rules apply; L/U (ACIFIM modes 'C','I','J') may be NULL
 * otherwise; H may be NULL when a_symmetric (PDFIIM). */
int cm_system_create(cm_system *sys, int variant, int n, int a_symmetric,
                     const int *Ai, const int *Aj, const double *Av,
                     const int *Gi, const int *Gj, const double *Gv,
                     const double *IDv, const int *Hi, const int *Hj,
                     const double *Hv, const int *Li, const int *Lj,
                     const double *Lv, const int *Ui, const int *Uj,
                     const double *Uv, int nr, char mode);
int cm_system_destroy(cm_system sys);

/* Device-resident right-hand side and solution of the system (n doubles each):
 * set_rhs uploads b; set_x uploads x (or zeroes it when x == NULL); get_x
 * downloads x. */
int cm_system_set_rhs(cm_system sys, const double *b);
int cm_system_set_x(cm_system sys, const double *x);
int cm_system_get_x(cm_system sys, double *x);

/* FP64 solve on the resident b and x (x holds the initial guess).
 * trace (optional, host, trace_cap doubles) receives ||r_k|| of every pass
 * (the lagged residual the reference tests, AFIIM/src/bicgstab.cpp:213-216);
 * *ntrace = passes that computed it.  info is in/out as in the reference (it
 * is left unchanged when NMAX < 0). */
int cm_system_solve(cm_system sys, double tol, int NMAX, double *resval, int *iter,
                    int *info, double *trace, int trace_cap, int *ntrace);

/* Mixed precision on the resident b and x: FP64 iterative refinement (true
 * residual b - A x, the solution and its norm in FP64, with the FP64 A)
 * around an inner BiCGSTAB on an FP32 A, FP32 (factor_bits 32) or FP16
 * per-row-scaled (factor_bits 16) factors and FP32 vectors, solved to
 * inner_tol from zero at every refinement step.  enable_mixed casts the
 * low-precision copies on the device once (they share the FP64 index
 * arrays).  Stops when ||b - A x|| <= tol ||b - A x0||: *resval is that true
 * relative residual, *outer the refinement steps, *inner_total the inner
 * BiCGSTAB passes; info 0 converged, 1 non-finite residual, 2 a refinement
 * step did not reduce the residual, 3 max_outer steps reached. */
int cm_system_enable_mixed(cm_system sys, int factor_bits);
int cm_system_solve_mixed(cm_system sys, int factor_bits, double tol, double inner_tol, int NMAX,
                          int max_outer, double *resval, int *outer, int *inner_total, int *info);

/* Multiple right-hand sides of one system, solved together: every column
 * runs its own BiCGSTAB recurrence with its own scalars, exit tests and
 * counters -- the reference loop (AFIIM/src/bicgstab.cpp:149-281, called
 * once per right-hand side) -- while each product with A, G and H reads the
 * matrix once for all columns.  Apply mode E only (any ACIFIM system built
 * with another mode returns CM_ERR_UNSUPPORTED).
 *   block_set:   k columns (1..16), B and X0 column-major n x k (X0 may be
 *                null: zero initial guesses);
 *   block_solve: FP64, per-column resval[k], iter[k], info[k] with the
 *                meaning of cm_system_solve;
 *   block_solve_mixed: FP64 refinement of each column around the
 *                low-precision inner block solve (cm_system_enable_mixed
 *                first); per-column resval[k] and info[k] as in
 *                cm_system_solve_mixed; outer = refinement steps, inner_total
 *                = inner passes summed over steps (the longest column);
 *   block_get:   the k solutions, column-major n x k.
 * block_set with B = NULL keeps the resident block of k right-hand sides
 * (set before with the same k) and only resets x (to X0, or zero), so a
 * block can be re-solved without moving B over PCIe again. */
int cm_system_block_set(cm_system sys, int k, const double *B, const double *X0);
int cm_system_block_solve(cm_system sys, double tol, int NMAX, double *resval, int *iter, int *info);
int cm_system_block_solve_mixed(cm_system sys, int factor_bits, double tol, double inner_tol, int NMAX,
                                int max_outer, double *resval, int *outer, int *inner_total, int *info);
int cm_system_block_get(cm_system sys, double *X);
/* BiCGSTAB passes are replayed from hipGraphs captured on first use (on, the
 * default) or launched kernel by kernel (off); the results are identical.
 * cm_system_graphs covers the single-RHS solves (which launch kernel by
 * kernel anyway while cm_system_profile brackets launches), and
 * cm_system_block_graphs the block solves. */
int cm_system_graphs(cm_system sys, int on);
int cm_system_block_graphs(cm_system sys, int on);
/* Persistent batched solves: nitems right-hand sides of small resident
 * systems (item i: system systems[i], column B[i], initial guess and result
 * X[i], tol[i], NMAX[i]; systems may repeat and differ in size) in one
 * launch.  Each workgroup of a persistent grid pulls items from a device
 * queue and runs the whole FP64 reference loop (AFIIM/src/bicgstab.cpp:
 * 149-281) for one item with workgroup barriers; largest systems go first.
 * restart_true = 1 restarts an item that converged on the recursive residual
 * while its true residual is above tol (at most 4 cycles, NMAX passes in
 * total); 0 is the reference loop.  Apply mode E; per-item resval, iter,
 * info as in cm_system_solve.  Suited to systems whose solve is launch-bound
 * (up to some 10^4-10^5 rows); work vectors take 9 n doubles per resident
 * workgroup. */
int cm_batch_solve(int nitems, const cm_system *systems, const double *const *B, double *const *X,
                   const double *tol, const int *NMAX, int restart_true, double *resval, int *iter, int *info);

/* cm_batch_solve without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 * Same items, per-item results and loop as cm_batch_solve, but B[i] and X[i]
 * are DEVICE pointers (n doubles of item i's system; X[i] holds the initial
 * guess and receives the solution in place), and the work runs on `stream`
 * (a hipStream_t; NULL = the default stream; the call returns when it is
 * done).  The queue, the item and system tables and the work vectors live in
 * one fixed device array of the library (64 MiB of work vectors plus tables,
 * held on each GPU it runs on until the process ends): at most 4096 items
 * and 256 distinct systems per call, and 9 n_max doubles per workgroup must
 * fit the 64 MiB (n_max up to about 930,000 rows; fewer rows run more items
 * at once); larger calls return CM_ERR_ARG.  Calls are serialized inside
 * the library; concurrent calls from several processes on one GPU use
 * separate copies. */
int cm_batch_solve_noalloc(int nitems, const cm_system *systems, const double *const *B, double *const *X,
                           const double *tol, const int *NMAX, int restart_true, double *resval, int *iter,
                           int *info, void *stream);

/* Kernel timing of the current block's SpMM y = M X, M = 'A', 'G' or 'H'
 * (first factor; G^T on the symmetric PDFIIM branch), in FP64 (bits 64) or
 * the inner precision (32, 16; cm_system_enable_mixed first): reps
 * back-to-back products between HIP events on the system's stream; avg_us
 * per product and its compulsory bytes (matrix and row lists once, X and Y
 * once per column). */
int cm_system_block_time_spmm(cm_system sys, char which, int bits, int reps, double *avg_us, double *bytes);
/* Autotuner of the single-RHS SpMV: for each resident matrix (A, G, H,
 * G^T, L, U), times reps FP64 products with row blocks cut for 1024, 2048
 * and 4096 staged nonzeros per workgroup and keeps the fastest (the FP32 /
 * FP16 copies follow).  Then, for A, G, H and G^T (and L and U for ACIFIM
 * apply mode C with nr > 1, whose sweeps multiply by them) in every enabled
 * precision, times each available SpMV format (see cm_system_spmv_format)
 * and keeps the fastest.  tiles[6]: the tile choices, 0 for an absent
 * matrix.  Results of later solves are unchanged up to the summation order
 * of the row sums and of the row blocks' partial dot products. */
int cm_system_autotune(cm_system sys, int reps, int *tiles);
/* SpMV format of one resident matrix copy: bits 64, 32 or 16 (after
 * cm_system_enable_mixed), which 0..5 = A, G, H, G^T, L, U, format
 *   0  staged row blocks, 32-bit column indices;
 *   1  staged row blocks, 16-bit block-relative column offsets (FP32 / FP16
 *      copies whose blocks' column spans are below 65536);
 *   2  slot-major blocks of 256 rows, one row per thread (rows of at most
 *      64 nonzeros), 32-bit column indices;
 *   3  slot-major blocks with 16-bit column offsets;
 *   4  diagonals held by at least 60% of the rows stored as dense n-vectors
 *      (no column indices), the other entries as a CSR remainder, one row
 *      per thread (matrices of structured grids; FP16 two rows per thread);
 *   5  the same layout with twice the rows per thread and 9 diagonals per
 *      chunk (faster on matrices with many diagonals).
 * CM_ERR_ARG when the format is not available for that copy.  The
 * slot-major and diagonal indexes are built on first use (and shared by the
 * precisions). */
int cm_system_spmv_format(cm_system sys, int bits, int which, int format);
/* Current formats: formats[18] = {FP64, FP32, FP16} x {A, G, H, G^T, L, U},
 * -1 for an absent matrix or a precision not enabled. */
int cm_system_spmv_formats(cm_system sys, int *formats);
/* Diagonal index of one resident matrix (which 0..5 = A, G, H, G^T, L, U), built
 * if absent: *nd stored diagonals (0 when none is held by 60% of the rows),
 * their offsets (column - row, ascending) in offsets[0..*nd) (room for 32
 * needed), and *nrem entries left to the CSR remainder.  No counterpart in
 * the reference (a view of this engine's storage choice). */
int cm_system_dia_info(cm_system sys, int which, int *nd, int *offsets, long long *nrem);
/* Autotuner of the current block's SpMM row binning (bits 64, 32 or 16):
 * for A, G and H the short-row limit among 16, 32, 64, 128 nonzeros that
 * times fastest; short_max[3] receives the choices. */
int cm_system_block_autotune(cm_system sys, int bits, int reps, int *short_max);
/* Block products (SpMM) in the diagonal format (on, the default) for every
 * SpMV matrix whose diagonals held by at least 60% of the rows leave at most
 * 30% of its nonzeros to a CSR remainder; the other matrices, and all of
 * them when off, use the row-binned kernels.  The diagonal index and values
 * are built on the first block solve (or here) and shared with the
 * single-RHS diagonal format. */
int cm_system_block_dia(cm_system sys, int on);
/* Frees the block workspaces (one per block width used so far). */
int cm_system_block_release(cm_system sys);

/* Kernel-level entry points (host vectors in/out), for parity checks:
 * which = 'A', 'G', 'H', 'L', 'U' (and 'T' for the G^T copy of a symmetric
 * PDFIIM system): y = M x.  apply: out = M v with the system's preconditioner. */
int cm_system_spmv(cm_system sys, char which, const double *x, double *y);
int cm_system_apply(cm_system sys, const double *v, double *out);

/* Per-kernel timing of subsequent solves: every launch of a category whose
 * bit is set in mask (0 = off) is bracketed by HIP events on the system's
 * stream (each bracket costs a few microseconds of dispatch); launches the
 * device skipped after the exit are not counted.  Statistics accumulate until
 * the next cm_system_profile call.  Categories: 0 A products (with
 * their fused dots), 1 first factor product (H, or G^T, with D^{-1}), 2 G
 * products, 3 L/U products (mode C), 4 vector updates, 5 setup, 6 the
 * one-workgroup reduction finalizers.  total_bytes
 * is the algorithmic traffic of the counted launches: for the SpMV categories
 * the CSR figure nnz*(value bytes + 4) + 4*(n+1) (+ 4*n row scales for FP16
 * values) plus the vectors read and written, whatever storage format runs. */
int cm_system_profile(cm_system sys, int mask);
/* As cm_system_profile, but only the first `passes` passes of each solve
 * are bracketed (and launched kernel by kernel); the other passes replay
 * from the captured hipGraphs, so timing costs the solve little.  passes = 0
 * brackets every pass (cm_system_profile). */
int cm_system_profile_sampled(cm_system sys, int mask, int passes);
int cm_system_kernel_stats(cm_system sys, int category, long long *launches, double *total_ms,
                           double *total_bytes);
/* The bytes the same counted launches touch once in the storage format each
 * one ran in (row-block tables, slot tables, or the diagonals with their
 * zero fill and no column indices; equal to total_bytes' vector terms plus
 * the format's matrix arrays).  For the non-SpMV categories it equals
 * total_bytes. */
int cm_system_kernel_format_bytes(cm_system sys, int category, double *total_bytes);

/* Stored sizes (after symmetric expansion / transposition): which as above. */
long long cm_system_nnz(cm_system sys, char which);

const char *cm_last_error(void);
int cm_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
