/*
 * lfrv_batch.h -- many novel-view solves at once on one GPU.
 *
 * The reference solves novel views strictly one after another: one process
 * per IBR_gradient view (gradientIBR.cpp:12-150), a loop of
 * sr_synthesize_view calls for IBR_color (IBR_video, main.cpp:359-403).  A
 * batch takes solves of any size and view count, each on its own context
 * (views, depths and cameras loaded), and runs them together:
 *
 *   - every solve's set-up (warps, splat structure, visibility, weights,
 *     starting image, target gradient) on its context's own stream;
 *   - the PCG iterations of every dsf = 1 IBR_gradient solve in one kernel
 *     launch per iteration (block -> (solve, tile) from a work list) plus one
 *     launch of the scalar steps of all solves, `chunk` iterations captured
 *     once into a hipGraph and replayed until every solve has converged (a
 *     converged solve's blocks return at once);
 *   - step s of every dsf >= 2 IBR_color solve with one primal-dual
 *     iteration per step in one launch (block -> (solve, tile) from a work
 *     list, each solve's buffers of step parity s & 1 and its own FISTA
 *     theta), on a stream of the batch concurrent with the graph; other
 *     IBR_color solves run their steps on their own streams.
 * The results equal the single-solve entry points' bit for bit
 * (lfrv_cg_solve, lfrv_sr_iterate: every block of a batched launch does what
 * the same block of the solve's own launch does, partial sums in the same
 * slots).
 */
#ifndef LFRV_BATCH_H
#define LFRV_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "lfrv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lfrv_batch lfrv_batch;

enum { LFRV_BATCH_CG = 0, LFRV_BATCH_PD = 1 };

typedef struct {
    int chunk;          /* PCG iterations per graph replay (even; 0: 8) */
    int use_graph;      /* 1: replay a captured hipGraph; 0: launch the same kernels directly */
} lfrv_batch_params;

typedef struct {
    int kind;                 /* LFRV_BATCH_CG or LFRV_BATCH_PD */
    int status;               /* lfrv_status of the solve */
    int iterations;           /* CG: largest per-channel iteration count; PD: FISTA steps */
    double true_rel_residual; /* CG: largest per-channel ||b - A x|| / ||b|| at exit */
    double done_ms;           /* completion seen by the host, from the start of lfrv_batch_run */
} lfrv_batch_result;

lfrv_status lfrv_batch_create(const lfrv_batch_params *p, lfrv_batch **out);
void lfrv_batch_destroy(lfrv_batch *b);
const char *lfrv_batch_last_error(const lfrv_batch *b);

/* Add a solve.  CG: an IBR_gradient context (dsf 1) solved to ||r|| <= tol ||b||
 * within maxit iterations per channel (`steps` unused).  PD: an IBR_color
 * context run for `steps` FISTA steps (tol, maxit unused).  The context must
 * hold its views and either depths + cameras or tau; it stays owned by the
 * caller and must outlive the batch's runs. */
lfrv_status lfrv_batch_add(lfrv_batch *b, lfrv_ctx *ctx, int kind, double tol, int maxit, int steps);

/* Run every solve from its context's inputs to its solution (ctx U).
 * results (nullable) receives one entry per added solve, in order. */
lfrv_status lfrv_batch_run(lfrv_batch *b, lfrv_batch_result *results);

#ifdef __cplusplus
}
#endif
#endif /* LFRV_BATCH_H */
