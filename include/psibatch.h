/*
 * psibatch.h -- C ABI of batched ensembles of independent streamfunction
 * solves (libpsisolve.so).
 *
 * Each member is one congrad_streamfunction solve (solve_streamfct.f90:
 * 153-327) on its own basin: its own grid and metrics, depth coefficients,
 * island perimeters (Dirichlet points, make_inv_sfc :381-422) and forcing.
 * The reference solves one basin per process and has no batched entry
 * point; this API exists for ensembles of uneven basins (many small
 * solves that one at a time cannot fill the GPU).  Members run the
 * mixed-precision path of psi_congrad (FP32 inner Jacobi-PCG, FP64 residual
 * replacement) to the PSI_CRIT_RELRES target, all of them advanced by one
 * kernel per launch; the launches are replayed from a captured hipGraph.
 *
 * Conventions as in psisolve.h: fields are caller-owned device arrays of
 * shape (ny+4, nx+4) per member, metrics are host arrays, every function
 * returns a psi_status.
 */
#ifndef PSIBATCH_H
#define PSIBATCH_H

#include <stdint.h>

#include "psisolve.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psi_batch psi_batch;

typedef struct psi_batch_member {
    psi_grid grid;                /* as for psi_ctx_create */
    const double *hur, *hvr;      /* device (ny+4, nx+4) */
    psi_islands islands;          /* perimeters held at Dirichlet 0; nisle = 0: none */
} psi_batch_member;

/* Work schedule of the batched kernel. */
typedef struct psi_batch_tune {
    int32_t tile_x, tile_y;       /* output tile per workgroup: 64x16, 64x8, 32x16, 32x8 */
    int32_t tiles_per_item;       /* tiles of one member per work item (1..64) */
    int32_t items_per_grab;       /* work items a workgroup takes per grab (1..1024) */
    int32_t groups_per_cu;        /* resident workgroups per CU (1..occupancy) */
    int32_t graph_launches;       /* launches per captured graph (even, 2..256) */
    double ns_per_point;          /* autotuner: ns per active point per launch */
} psi_batch_tune;

typedef struct psi_batch_stats {
    double elapsed_ms;            /* device time from start to the last member done */
    int32_t launches;             /* batched kernel launches issued */
    int32_t graph_replays;
    int32_t polls;                /* host reads of the per-member done flags */
    int32_t reserved;
    int64_t active_points;        /* active (Z != 0) points over all members */
} psi_batch_stats;

/* Builds every member's operator (make_coeff_streamfunction, :123-148, and
 * the Jacobi inverse with its Dirichlet set) into packed device arrays and
 * allocates all scratch.  The default schedule is 64x16 tiles, 4 tiles per
 * item, 2 items per grab, full occupancy, 16 launches per graph. */
int psi_batch_create(psi_batch **out, int32_t count, const psi_batch_member *members,
                     void *stream);
void psi_batch_destroy(psi_batch *batch);
const char *psi_batch_last_error(const psi_batch *batch);
int64_t psi_batch_device_bytes(const psi_batch *batch);
int64_t psi_batch_active_points(const psi_batch *batch);
/* Most resident workgroups per CU of the kernel for a tile shape (0: no such shape). */
int psi_batch_max_groups_per_cu(const psi_batch *batch, int32_t tile_x, int32_t tile_y);

int psi_batch_get_tune(const psi_batch *batch, psi_batch_tune *tune);
int psi_batch_set_tune(psi_batch *batch, const psi_batch_tune *tune);

/* Autotuner: times `launches` batched launches (after 4 untimed ones) in
 * which every member runs one full FP32 sweep, for every tile shape x
 * tiles_per_item in {2,4,8} x items_per_grab in {1,4} x {full, half}
 * occupancy on the members' own operators, and installs the fastest
 * schedule.  Caller data is untouched.  table (optional, table_cap rows of
 * 6 doubles: tile_x, tile_y, tiles_per_item, items_per_grab, groups_per_cu,
 * ns_per_point) receives every candidate; *table_n the number of
 * candidates. */
int psi_batch_autotune(psi_batch *batch, int32_t launches, psi_batch_tune *best,
                       double *table, int32_t table_cap, int32_t *table_n);

/* Solves every member: A sol[k] = forc[k] from the first guess in sol[k]
 * to ||forc - A sol||_2 <= relres_target ||forc||_2 over active points
 * (FP64 true residual).  opts: mode PSI_MODE_MIXED and criterion
 * PSI_CRIT_RELRES are required; max_iterations bounds each member's inner
 * iterations; inner_reduction sets the FP32 leg length (<= 0: 0.05).  forc, sol: host
 * arrays of count device pointers.  infos (optional, count entries):
 * iterations, converged, status, outer_iterations (FP64 residual checks),
 * relres_true, elapsed_ms = the member's completion time from the start of
 * the batch (device clock), sweep_launches = launches the member took part
 * in.  latency_ms (optional, count): the same completion times.  Returns
 * PSI_OK if every member converged, else the first failing member's status. */
int psi_batch_congrad(psi_batch *batch, const psi_opts *opts, const double *const *forc,
                      double *const *sol, psi_info *infos, double *latency_ms,
                      psi_batch_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* PSIBATCH_H */
