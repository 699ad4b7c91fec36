/* cm_serve.h -- request loop over resident systems (continuous batching).
 *
 * The reference has no serving layer: each driver calls the solve once per
 * right-hand side and waits (AFIIM/example/example.cpp:100-118,
 * PDFIIM/example/pdfiim.cpp:127-138, ACIFIM/example/example.cpp:158-184).
 * A server keeps registered systems (cm_system handles, see cm_solve.h)
 * resident on the GPU and solves requests as they arrive:
 *   - `slots` worker threads with one HIP stream each; a free slot takes the
 *     oldest pending request whose system is idle together with every other
 *     pending request of that system (same tol and NMAX) up to 16 columns,
 *     and solves them as one block (cm_system_block_*); batches of different
 *     systems overlap on the device;
 *   - admission by HBM footprint against `hbm_budget` bytes (0: 90% of the
 *     device memory): resident systems plus block workspaces; idle
 *     workspaces are released least-recently-used first when a batch needs
 *     room, otherwise the batch waits for running ones to finish;
 *   - precision 64: FP64 BiCGSTAB per column (tol on the recursive residual,
 *     the reference's test); 32 / 16: FP64 refinement of each column to tol
 *     on the true residual around the FP32 inner solve with FP32 or FP16
 *     factors (inner tolerance 1e-4, at most 30 refinement steps).
 * A registered system must not be used through the cm_system_* calls while
 * the server runs, and must outlive the server. */
#ifndef CM_SERVE_H
#define CM_SERVE_H

#include "cm_solve.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cm_server_s *cm_server;

int cm_server_create(cm_server *srv, int slots, int precision, long long hbm_budget);
int cm_server_destroy(cm_server srv);
/* Batching and refinement settings (before the first submit): at most
 * max_cols (1..16) right-hand sides per block batch -- 1 disables merging; a
 * wider request still runs, alone -- the inner tolerance / refinement steps
 * of precision 32/16, and small_n: requests on systems of at most small_n
 * rows are solved together, every pending one of any small system in one
 * launch of the persistent batched kernel (cm_batch_solve; FP64, with the
 * true-residual restart under precision 32/16); 0 disables it. */
int cm_server_configure(cm_server srv, int max_cols, double inner_tol, int max_outer, int small_n);
/* Registers a system (apply mode E for ACIFIM); enables its mixed-precision
 * copies for precision 32/16.  sys_id: the id requests name. */
int cm_server_add_system(cm_server srv, cm_system sys, int *sys_id);
/* Queues a request of k (1..16) right-hand sides, B column-major n x k
 * (copied); returns at once with its id.  x0 = 0 as in the reference
 * drivers. */
int cm_server_submit(cm_server srv, int sys_id, int k, const double *B, double tol, int NMAX, long long *req_id);
/* Blocks until the request is done: X (n x k), per-column resval[k] and
 * info[k]; counts = {outer steps, inner passes, columns in its batch};
 * times_ms = {queue, service, latency} (arrival = submit). */
int cm_server_wait(cm_server srv, long long req_id, double *X, double *resval, int *info, int *counts,
                   double *times_ms);
/* stats[8] = {block batches, columns, most batches running at once,
 * workspace evictions, bytes accounted, budget, persistent batches, their
 * columns}. */
int cm_server_stats(cm_server srv, long long *stats);

#ifdef __cplusplus
}
#endif

#endif
