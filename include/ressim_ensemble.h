/* Batched FIM ensemble on one GPU: many two-phase (oil-water) models of uneven grid
 * sizes advanced together, every kernel launched once over a packed index space that
 * holds all members' cells.
 *
 * No counterpart in the reference: it steps one model per FimSolver / ReservoirSimulator
 * (rs/fim/timestep.rs:988-1890). A member is configured through the reference's own
 * setters on an rsf_sim handle (include/ressim_fim.h); rsf_ensemble_create copies each
 * member's model and stored state into the packed layout, after which the ensemble owns
 * the state it advances (the member handles are not changed and may be destroyed).
 *
 * Each member runs its own Newton iteration on its own substep schedule: per-member
 * scalars (phase, substep length, Newton and Krylov state) live on the device, and one
 * captured graph, a "tick", performs one Newton iteration of every member that is still
 * running (assembly, convergence test, FGMRES with a two-stage CPR preconditioner,
 * update, and the substep bookkeeping). Members at different substeps, or converged,
 * coexist in one tick.
 *
 * Members must be two-phase (three-phase mode off) with gravity off and no PVT table;
 * wells may be BHP-controlled or rate-controlled with a BHP limit (the reference's
 * Fischer-Burmeister control row).
 */
#ifndef RESSIM_ENSEMBLE_H
#define RESSIM_ENSEMBLE_H

#include "ressim_fim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rsf_ensemble rsf_ensemble;

typedef struct rsf_ensemble_options {
  int max_newton_iterations;       /* per substep attempt (default 20) */
  int min_newton_iteration_index;  /* first iteration that may be accepted (default 1) */
  double newton_residual_tol;      /* scaled family maxima and well control rows (default 1e-6) */
  double newton_mb_tol;            /* per-component material balance (default 1e-7) */
  double linear_rel_tol;           /* FGMRES relative residual (default 5e-3) */
  int krylov_dimension;            /* FGMRES iterations per solve, no restart (default 20, at most 200) */
  int pressure_max_iterations;     /* CPR pressure BiCGStab (default 50, the reference's CPR:
                                      fim/linear/gmres_block_jacobi.rs:25) */
  double pressure_rel_tol;         /* its relative residual (default 1e-6, :26) */
  int max_retries;                 /* halvings of one substep before a member is given up (default 12) */
  int autotune;                    /* 1: time the tile choices at creation (default 1) */
  int threads_per_block;           /* without autotune: 128, 256 or 512 (default 256) */
  int cells_per_thread[3];         /* without autotune: per size class, 1, 2 or 4 (default 1) */
  int use_graph;                   /* 1: replay the tick as a captured HIP graph; 0: launch its kernels
                                      directly, the same kernels in the same order (default 1) */
} rsf_ensemble_options;

void rsf_ensemble_default_options(rsf_ensemble_options* opt);

/* Copies the members' models and stored states; RSF_ERR_INVALID (message via
 * rsf_ensemble_last_error) for a member the ensemble does not support. */
int rsf_ensemble_create(rsf_sim* const* members, int count, const rsf_ensemble_options* opt, rsf_ensemble** out);
void rsf_ensemble_destroy(rsf_ensemble* e);

/* n_reports report steps of report_dt_days for every member, each with its own substep
 * control: a substep starts at min(2 x the last accepted one, what is left of the report
 * step) and is halved on failure, at most max_retries times in a row, after which the
 * member stops (status 2). Returns RSF_OK, or RSF_ERR_NOT_CONVERGED when a member stopped. */
int rsf_ensemble_run(rsf_ensemble* e, double report_dt_days, int n_reports);

/* Member m takes exactly the substeps dts[offsets[m] .. offsets[m+1]) (days); a substep
 * that does not converge stops that member (status 2). */
int rsf_ensemble_replay(rsf_ensemble* e, const double* dts, const int* offsets);

/* Cell pressures (bar) and water saturations of member m, natural order, n = its cells. */
int rsf_ensemble_member_state(rsf_ensemble* e, int m, double* pressure, double* sw, int n);

/* BHP per physical well (the member's well order) and the connection rate per completion
 * (the member's completion order, m3/d at reservoir conditions, > 0 production) at the
 * last assembled iterate. */
int rsf_ensemble_member_wells(rsf_ensemble* e, int m, double* bhp, int n_wells, double* q, int n_completions);

typedef struct rsf_ensemble_member_report {
  int status;             /* 0 running, 1 finished, 2 stopped (did not converge) */
  int substeps, retries, newton_iterations, linear_iterations, n_cells;
  double time_days;
} rsf_ensemble_member_report;
int rsf_ensemble_member_stats(rsf_ensemble* e, int m, rsf_ensemble_member_report* out);

typedef struct rsf_ensemble_tuning {
  int threads_per_block;
  int cells_per_thread[3];  /* size classes: cells < class_bounds[0], < class_bounds[1], larger */
  int class_bounds[2];
  int class_members[3];
  double probe_ms[3][3][3]; /* [threads 128/256/512][class][cells per thread 1/2/4]; -1 not timed */
  long ticks, graph_nodes;
  int n_members;
  long n_cells;
} rsf_ensemble_tuning;
int rsf_ensemble_tuning_info(rsf_ensemble* e, rsf_ensemble_tuning* out);

const char* rsf_ensemble_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
