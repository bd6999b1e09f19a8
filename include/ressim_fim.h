/*
 * ressim_fim — C ABI of the MI355X (gfx950) fully implicit (FIM) Newton
 * timestep for three-phase black-oil flow.
 *
 * The boundary mirrors the reference simulator's Rust API
 * (src/lib/ressim/src/frontend.rs, `ReservoirSimulator`): one opaque handle
 * per simulator, the same setter names, arguments, units (bar, m, day, mD,
 * cP) and defaults, and `step(dt_days)` as the unit of work. Rust `Result`
 * errors become integer status codes; a `Vec<f64>` getter becomes a
 * caller-owned buffer plus its length. Citations below name the reference
 * function each entry point replaces (paths relative to
 * src/lib/ressim/src/). Entry points marked EXTENSION have no reference
 * counterpart; they carry inputs the reference cannot express.
 *
 * Threading: a handle is single-threaded and non-reentrant, like the
 * reference's `&mut self`. All device work runs on a stream the handle owns.
 */
#ifndef RESSIM_FIM_H
#define RESSIM_FIM_H

#ifdef __cplusplus
extern "C" {
#endif

#define RSF_OK 0                /* success */
#define RSF_ERR_INVALID 1       /* invalid argument (reference: Err(String) from a setter) */
#define RSF_ERR_NOT_CONVERGED 2 /* substep controller gave up (reference: last_solver_warning) */
#define RSF_ERR_HIP 3           /* HIP runtime failure (message via rsf_last_warning) */

typedef struct rsf_sim rsf_sim;

/* Newton / linear options. Defaults (rsf_default_solver_options) are the
 * reference's: FimNewtonOptions::default (fim/newton.rs:900-919) with the
 * OpmAligned acceptance (fim/newton/convergence.rs:352-385) and
 * FimLinearSolveOptions::default (fim/linear/mod.rs:207-235). */
typedef struct rsf_solver_options {
  int max_newton_iterations;      /* 20 */
  int min_newton_iteration_index; /* 2 (OPM newton-min-iterations) */
  double newton_residual_tol;     /* 0 = OPM CNV/MB; > 0 = max scaled residual |R|/(max(pv/dt,1)/B) */
  double newton_mb_tol;           /* 1e-7, used with newton_residual_tol */
  double linear_rel_tol;          /* 5e-3 */
  double linear_abs_tol;          /* 1e-12 */
  int linear_max_iterations;      /* 20 (promoted to one restart cycle) */
  int linear_restart;             /* 30 */
  int precision;                  /* 0 FP64; 1 mixed: FP32 CPR stages and inner vectors, FP64 outer FGMRES */
  int ilu_ordering;               /* 0 natural order (reference); 1 red-black multicolour (EXTENSION) */
  int pressure_max_iterations;    /* 50 (fim/linear/gmres_block_jacobi.rs:25) */
  double pressure_rel_tol;        /* 1e-6 (fim/linear/gmres_block_jacobi.rs:26) */
} rsf_solver_options;

typedef struct rsf_step_report { /* FimStepReport subset, fim/newton.rs:921-944 */
  int converged;
  int newton_iterations;
  long linear_iterations;
  double final_scaled_residual;
} rsf_step_report;

typedef struct rsf_stats { /* FimStepStats subset, reporting.rs:191-243 */
  long newton_iterations, linear_iterations, substeps, retries, pressure_iterations, pressure_solves;
  double assembly_ms, linear_ms, total_ms;
} rsf_stats;

/* FimStepStats (reporting.rs:191-243): one record per rsf_step call, in call order
 * (fim_step_stats_history, lib.rs:175; stored by timestep.rs:742-800 on every exit of the
 * controller). Scalar fields only; the rungs are read with rsf_get_step_rungs. An absent
 * value (the reference's None) is NaN for a double and -1 for a code or a row. The rung
 * vectors and the replay and hotspot counters of the reference's other controller modes
 * are not kept: the OPM-aligned controller has neither. A dominant row
 * (FimRetryFailureDiagnostics::dominant_row, fim/newton.rs:486-487) is the row of the
 * dominant family's peak in the reference's unknown numbering: 3 per cell (water, oil,
 * gas), then one per well, then one per completion (fim/newton/convergence.rs:734-790). */
enum { RSF_LIMIT_MAX_GROWTH = 0, RSF_LIMIT_PID = 1, RSF_LIMIT_ITER = 2, RSF_LIMIT_RESTART_GROWTH = 3 };
/* "opm-max-growth", "opm-pid", "opm-iter", "opm-restart-growth" (timestep.rs:502-528) */
enum { RSF_RETRY_LINEAR_BAD = 0, RSF_RETRY_NONLINEAR_BAD = 1, RSF_RETRY_MIXED = 2 };
/* fim/newton.rs:168-182 FimRetryFailureClass */
enum { RSF_FAMILY_WATER = 0, RSF_FAMILY_OIL = 1, RSF_FAMILY_GAS = 2, RSF_FAMILY_WELL = 3, RSF_FAMILY_PERF = 4 };
/* fim/newton/convergence.rs:287-304 ResidualRowFamily */
typedef struct rsf_step_stats {
  double time_days, target_dt_days, advanced_dt_days;
  int accepted_substeps, linear_bad_retries, nonlinear_bad_retries, mixed_retries;
  double min_accepted_dt_days, max_accepted_dt_days, last_accepted_dt_days, last_growth_factor;
  int growth_limiter, last_retry_class, last_retry_dominant_family, n_rungs;
  double solver_ms, accepted_solver_ms, retry_solver_ms, assembly_ms, linear_solve_ms;
  long last_retry_dominant_row;
} rsf_step_stats;

/* FimAcceptedRungStats / FimRetryRungStats (reporting.rs:141-188): one Newton attempt of
 * the step. retried = 1 for a rejected attempt, which carries its class, dominant residual
 * family and dominant row (-1 for an accepted one); linear_iterations and
 * last_linear_converged describe the attempt's last linear solve (-1: no solve). */
typedef struct rsf_step_rung {
  int substep, retried, newton_iterations, linear_iterations, linear_solve_count;
  int retry_class, dominant_family, last_linear_converged;
  long linear_iteration_count;
  double dt_days, linear_solve_ms;
  long dominant_row;
} rsf_step_rung;

/* ReservoirSimulator::new(nx, ny, nz, porosity) — frontend.rs:36-123 */
int rsf_create(int nx, int ny, int nz, double porosity, rsf_sim** out);
void rsf_destroy(rsf_sim* sim);

/* frontend.rs:622 set_cell_dimensions */
int rsf_set_cell_dimensions(rsf_sim* sim, double dx, double dy, double dz);
/* frontend.rs:639 set_cell_dimensions_per_layer */
int rsf_set_cell_dimensions_per_layer(rsf_sim* sim, double dx, double dy, const double* dz, int nz);
/* frontend.rs:1291 set_permeability_field */
int rsf_set_permeability_field(rsf_sim* sim, const double* kx, const double* ky, const double* kz, int n);
/* frontend.rs:1243 set_permeability_per_layer: one value per layer for every cell of it */
int rsf_set_permeability_per_layer(rsf_sim* sim, const double* kx, const double* ky, const double* kz, int nz);
/* EXTENSION: permeability from ln k per cell (the adjoint's parameter): kx = ky = exp(ln k),
 * kz = kz_over_kh kx */
int rsf_set_perm_field_log(rsf_sim* sim, const double* ln_k, double kz_over_kh, int n);
/* EXTENSION: per-cell porosity (the reference has only the constructor's scalar) */
int rsf_set_porosity_field(rsf_sim* sim, const double* phi, int n);
/* EXTENSION: per-cell Leverett-J multiplier of the Brooks-Corey Pc_ow (capillary.rs:73-90) */
int rsf_set_capillary_scale_field(rsf_sim* sim, const double* scale, int n);
/* frontend.rs:613 set_initial_pressure */
int rsf_set_initial_pressure(rsf_sim* sim, double p);
/* frontend.rs:676 set_initial_saturation */
int rsf_set_initial_saturation(rsf_sim* sim, double sw);
/* frontend.rs:1031 set_initial_gas_saturation */
int rsf_set_initial_gas_saturation(rsf_sim* sim, double sg);
/* frontend.rs:1023 set_initial_rs */
int rsf_set_initial_rs(rsf_sim* sim, double rs);
/* EXTENSION: per-cell initial state (a loadState grid without the JSON, frontend.rs:931-1008) */
int rsf_set_initial_state_arrays(rsf_sim* sim, const double* p, const double* sw, const double* sg,
                                 const double* rs, int n);
/* frontend.rs:788 set_fluid_properties */
int rsf_set_fluid_properties(rsf_sim* sim, double mu_o, double mu_w);
/* frontend.rs:804 set_fluid_compressibilities */
int rsf_set_fluid_compressibilities(rsf_sim* sim, double c_o, double c_w);
/* frontend.rs:772 set_fluid_densities */
int rsf_set_fluid_densities(rsf_sim* sim, double rho_o, double rho_w);
/* frontend.rs:820 set_rock_properties */
int rsf_set_rock_properties(rsf_sim* sim, double c_r, double depth_reference_m, double b_o, double b_w);
/* frontend.rs:1183 set_gas_fluid_properties */
int rsf_set_gas_fluid_properties(rsf_sim* sim, double mu_g, double c_g, double rho_g);
/* frontend.rs:718 set_rel_perm_props */
int rsf_set_rel_perm_props(rsf_sim* sim, double s_wc, double s_or, double n_w, double n_o, double k_rw_max,
                           double k_ro_max);
/* frontend.rs:1095 set_three_phase_rel_perm_props */
int rsf_set_three_phase_rel_perm_props(rsf_sim* sim, double s_wc, double s_or, double s_gc, double s_gr,
                                       double s_org, double n_w, double n_o, double n_g, double k_rw_max,
                                       double k_ro_max, double k_rg_max);
/* frontend.rs:1145 set_three_phase_scal_tables (ThreePhaseScalTables, relperm.rs:59-80): n_swof rows
 * of (sw, krw, krow, pcow) and n_sgof rows of (sg, krg, krog, pcog), 4 doubles each; pcow/pcog may
 * be NaN (None) and are not read. The three-phase relperms then interpolate the tables
 * (relperm.rs:391-444), Stone II combining them with k_ro_max. Validated as relperm.rs:82-124;
 * set_three_phase_rel_perm_props afterwards drops them (frontend.rs:1139). */
int rsf_set_three_phase_scal_tables(rsf_sim* sim, const double* swof, int n_swof, const double* sgof, int n_sgof);
/* frontend.rs:854 set_capillary_params */
int rsf_set_capillary_params(rsf_sim* sim, double p_entry, double lambda);
/* frontend.rs:1161 set_gas_oil_capillary_params */
int rsf_set_gas_oil_capillary_params(rsf_sim* sim, double p_entry, double lambda);
/* frontend.rs:1090 set_three_phase_mode_enabled */
int rsf_set_three_phase_mode_enabled(rsf_sim* sim, int enabled);
/* frontend.rs:1208 set_gas_redissolution_enabled (default 1): 0 = DRSDT 0, a cell's Rs
 * may not rise above its stored value within a substep (fim/flash.rs:13-41,
 * fim/properties.rs:37-133, fim/flash_ad.rs:104-207, fim/state.rs:825-836) */
int rsf_set_gas_redissolution_enabled(rsf_sim* sim, int enabled);
/* frontend.rs:425 set_gravity_enabled */
int rsf_set_gravity_enabled(rsf_sim* sim, int enabled);
/* frontend.rs:375 set_fim_nested_well_solve (default 0): after each Newton update every
 * well converges its own rows with the cells frozen (fim/wells_inner.rs:523-598) in place
 * of the well relaxation (fim/state.rs:494-551), and acceptance also asks every well's own
 * convergence test (fim/newton.rs:1262-1285) */
int rsf_set_fim_nested_well_solve(rsf_sim* sim, int enabled);
/* frontend.rs:413 set_fim_corey_table_points */
int rsf_set_fim_corey_table_points(rsf_sim* sim, int points);
/* frontend.rs:1228 set_injected_fluid ("water" | "gas") */
int rsf_set_injected_fluid(rsf_sim* sim, const char* fluid);
/* frontend.rs:1011 set_pvt_table: n_rows rows of (p_bar, rs_m3m3, bo_m3m3, mu_o_cp, bg_m3m3, mu_g_cp) */
int rsf_set_pvt_table(rsf_sim* sim, const double* rows, int n_rows);
/* frontend.rs:438 set_well_control_modes ("rate" | "pressure") */
int rsf_set_well_control_modes(rsf_sim* sim, const char* injector_mode, const char* producer_mode);
/* frontend.rs:448 set_injector_enabled */
int rsf_set_injector_enabled(rsf_sim* sim, int enabled);
/* frontend.rs:453 set_target_well_rates */
int rsf_set_target_well_rates(rsf_sim* sim, double injector_rate_m3_day, double producer_rate_m3_day);
/* frontend.rs:474 set_target_well_surface_rates */
int rsf_set_target_well_surface_rates(rsf_sim* sim, double injector_rate_m3_day, double producer_rate_m3_day);
/* frontend.rs:503 set_well_bhp_limits */
int rsf_set_well_bhp_limits(rsf_sim* sim, double bhp_min, double bhp_max);
/* frontend.rs:190 add_well */
int rsf_add_well(rsf_sim* sim, int i, int j, int k, double bhp, double well_radius, double skin, int injector);
/* frontend.rs:204 add_well_with_id */
int rsf_add_well_with_id(rsf_sim* sim, int i, int j, int k, double bhp, double well_radius, double skin,
                         int injector, const char* physical_well_id);
/* frontend.rs:228 set_well_schedule (NaN or negative rates mean "unset", as in the reference) */
int rsf_set_well_schedule(rsf_sim* sim, const char* physical_well_id, const char* control_mode,
                          double target_rate_m3_day, double target_surface_rate_m3_day, double bhp_limit,
                          int enabled);
/* frontend.rs:300 set_well_datum */
int rsf_set_well_datum(rsf_sim* sim, const char* physical_well_id, double datum_depth_m,
                       double wellbore_density_kg_m3);
/* EXTENSION: per-well injected fluid (the reference has one global injected_fluid, lib.rs:208) */
int rsf_set_well_injected_fluid(rsf_sim* sim, const char* physical_well_id, const char* fluid);
/* Solver options (reference: FimNewtonOptions / FimLinearSolveOptions, set internally there) */
int rsf_default_solver_options(rsf_solver_options* out);
int rsf_set_solver_options(rsf_sim* sim, const rsf_solver_options* options);

/* frontend.rs:345 step(target_dt_days): substep controller fim/timestep.rs:988-1890 */
int rsf_step(rsf_sim* sim, double dt_days);
/* One Newton substep of fixed length from the stored state, committed on convergence:
 * run_fim_timestep (fim/newton.rs:970-976) plus the accept/write-back of fim/timestep.rs:1332-1353.
 * Used to replay an accepted-substep schedule. */
int rsf_run_timestep(rsf_sim* sim, double dt_days, rsf_step_report* out);
/* run_fim_timestep(sim, previous_state, initial_iterate, dt, options) (fim/newton.rs:970-976)
 * with an initial iterate of the caller's: the previous state is the stored one, the
 * Newton iterate starts from init_cell_state ([p, sw, hc] per cell, interleaved) with
 * init_regimes (0 saturated, 1 undersaturated), init_bhp per well and init_q per
 * completion. Any of them may be NULL for the stored state's own value (what
 * rsf_run_timestep uses). Committed on convergence. Not while the adjoint records. */
int rsf_run_timestep_from(rsf_sim* sim, double dt_days, const double* init_cell_state, const int* init_regimes,
                          const double* init_bhp, const double* init_q, rsf_step_report* out);

/* FimStepReport (fim/newton.rs:921-944) of one Newton attempt. The accepted_state is returned
 * through the output arrays of rsf_run_fim_timestep. final_update_inf_norm is NaN (not
 * tracked); the reference's property-evaluation, preconditioner-build and state-update
 * times are inside assembly_ms and linear_solve_time_ms here; failure_class and
 * failure_dominant_family are RSF_RETRY_* and RSF_FAMILY_* (failure_diagnostics), -1 when
 * converged; failure_dominant_row is its dominant row (numbered as in rsf_step_stats). */
typedef struct rsf_fim_step_report {
  int converged;
  int newton_iterations;
  double final_residual_inf_norm;
  double final_material_balance_inf_norm;
  double final_update_inf_norm;
  int has_last_linear_report; /* last_linear_report: Some */
  int last_linear_iterations;
  int last_linear_converged;
  long linear_solve_count;
  long linear_iteration_count;
  int applied_update_count;
  double retry_factor; /* fim/newton.rs:946-952: 0.25 linear-bad, 0.5 otherwise */
  int failure_class;
  int failure_dominant_family;
  double total_time_ms, assembly_ms, linear_solve_time_ms;
  long failure_dominant_row;
} rsf_fim_step_report;
/* run_fim_timestep(sim, previous_state, initial_iterate, dt, options) (fim/newton.rs:970-976):
 * one Newton attempt that commits nothing (the stored state, time and histories are
 * unchanged). previous_state is [p, sw, hc] per cell (interleaved) with its regimes
 * (0 saturated, 1 undersaturated), or both NULL for the stored state's FimState
 * (FimState::from_simulator, fim/state.rs:109-177); it sets the accumulation terms. The
 * initial iterate is init_* (each may be NULL: the previous state's cells and regimes, the
 * stored BHPs and connection rates). options NULL: the handle's. The attempt's final
 * iterate (accepted_state) goes to out_cell_state (3N), out_regimes (N), out_bhp (per well)
 * and out_q (per completion); any of them may be NULL. Returns RSF_OK when converged,
 * RSF_ERR_NOT_CONVERGED otherwise. Not while the adjoint records. */
int rsf_run_fim_timestep(rsf_sim* sim, double dt_days, const double* prev_cell_state, const int* prev_regimes,
                         const double* init_cell_state, const int* init_regimes, const double* init_bhp,
                         const double* init_q, const rsf_solver_options* options, rsf_fim_step_report* out,
                         double* out_cell_state, int* out_regimes, double* out_bhp, double* out_q);

/* frontend.rs:523 get_pressures, :549 get_sat_water, :554 get_sat_oil, :1080 get_sat_gas, :1085 get_rs */
int rsf_get_pressures(rsf_sim* sim, double* out, int n);
int rsf_get_sat_water(rsf_sim* sim, double* out, int n);
int rsf_get_sat_oil(rsf_sim* sim, double* out, int n);
int rsf_get_sat_gas(rsf_sim* sim, double* out, int n);
int rsf_get_rs(rsf_sim* sim, double* out, int n);
/* physical wells and completions (fim/wells.rs:544-594 topology) */
int rsf_num_wells(rsf_sim* sim);
int rsf_num_perforations(rsf_sim* sim);
/* consistent well state at the stored state: BHP per physical well, connection rate q per
 * completion in add order (fim/state.rs:109-177) */
int rsf_get_well_state(rsf_sim* sim, double* bhp, int n_wells, double* q, int n_perforations);
/* frontend.rs:518 get_time */
double rsf_get_time(rsf_sim* sim);
/* frontend.rs:593 get_last_fim_step_stats (cumulative counters) */
int rsf_get_stats(rsf_sim* sim, rsf_stats* out);
/* frontend.rs:597-600 getFimStepStatsHistory / :592-595 getLastFimStepStats: the number of
 * records, the record of one rsf_step call (index < count), and its rungs (returns the
 * rung count; at most cap are written) */
int rsf_step_stats_count(rsf_sim* sim);
int rsf_get_step_stats(rsf_sim* sim, int index, rsf_step_stats* out);
int rsf_get_step_rungs(rsf_sim* sim, int index, rsf_step_rung* out, int cap);
/* fim/newton.rs:439-481 classify_retry_failure_with_site as the controller applies it to a
 * rejected attempt's last linear solve (have_linear = 0: the attempt made none), and
 * fim/newton/convergence.rs:793-806 the dominant residual family of five scaled family
 * maxima (water, oil, gas, well, perforation; the last two only with wells). Host
 * functions, for tests of the rules. */
int rsf_debug_classify_retry(int have_linear, int converged, int direct, int coarse_applications,
                             double average_reduction, double last_reduction);
int rsf_debug_dominant_family(const double* family_maxima, int with_wells);
/* accepted substep lengths so far; returns the count */
int rsf_get_accepted_dts(rsf_sim* sim, double* out, int cap);
/* frontend.rs:588 get_last_solver_warning */
int rsf_last_warning(rsf_sim* sim, char* buf, int len);

/* Assembly at a given iterate against the stored state: assemble_fim_system_ad
 * (fim/assembly_ad.rs:782-983). residual: 3N + wells + perforations in the reference's
 * unknown order (cell-major [water, oil, gas], then well rows, then perforation rows);
 * blocks: 63N doubles, slot-major 7-point 3x3 blocks (slot s, entry e, cell i at
 * (s*9+e)*N+i). eliminate_wells=1 folds the well Schur complement into the blocks. */
int rsf_debug_assemble(rsf_sim* sim, double dt_days, const double* cell_state, const int* regimes,
                       const double* bhp, const double* q, int eliminate_wells, double* residual, double* blocks);
/* Newton direction: solve_linearized_system (fim/linear/mod.rs:277-397) of J dx = -R at an iterate */
int rsf_debug_newton_direction(rsf_sim* sim, double dt_days, const double* cell_state, const int* regimes,
                               const double* bhp, const double* q, double* dx, int* iterations);
/* The Newton loop's device record at an iterate (assembled as rsf_debug_assemble does, with
 * the wells eliminated): the acceptance inputs of run_fim_timestep (fim/newton.rs:1262-1300,
 * fim/newton/convergence.rs:478-612, fim/scaling.rs:107-187, fim/wells_inner.rs:600-640).
 * out[0..16]: scaled family maxima [water, oil, gas, well, perforation], CNV[3], MB[3], the
 * violating pore-volume fraction, its 3 %-PV rule (0/1), the OPM verdict with the strict
 * tiers and with the final-iteration tiers (0/1 each, without the minimum-iteration gate),
 * every well's own convergence test (0/1), |R|^2 over cells and tail, then out[17..21]: the
 * row of each family's peak (the first row holding the family maximum, numbered as in
 * rsf_step_stats; -1 for a family with no rows). out has 22 values. */
int rsf_debug_newton_record(rsf_sim* sim, double dt_days, const double* cell_state, const int* regimes,
                            const double* bhp, const double* q, double* out);
/* Nested per-well Newton at an iterate: solve_wells_locally (fim/wells_inner.rs:585-598)
 * with the cells frozen at cell_state/regimes (the stored state where null). bhp[n_wells]
 * and q[n_perforations] (completion order) are read and overwritten; per well: converged
 * (0/1), iterations and the final scaled residual peak (FimWellInnerSolveReport,
 * fim/wells_inner.rs:260-267). max_iterations < 0 keeps the default 50. */
int rsf_debug_solve_wells_locally(rsf_sim* sim, const double* cell_state, const int* regimes, double* bhp, double* q,
                                  int max_iterations, int* converged, int* iterations, double* peak);
/* converged Newton iterate of the last accepted substep, before write-back: primaries,
 * the substep's frozen regimes, BHP per well and q per completion (the `state` that
 * run_fim_timestep returns in FimStepReport, fim/newton.rs:921-944) */
int rsf_debug_last_iterate(rsf_sim* sim, double* cell_state, int* regimes, double* bhp, int n_wells, double* q,
                           int n_perforations);
/* substep initial iterate: regimes and primaries (fim/state.rs:109-177) */
int rsf_debug_initial_state(rsf_sim* sim, double* cell_state, int* regimes);
/* Kernel timing at the simulator's current Newton system (call after a step): each
 * kernel family is launched `reps` times on the handle's stream between HIP events.
 * avg_ms is the mean duration of one launch of the family's unit (one SpMV, one full
 * forward+backward ILU sweep, one assembly); bytes is its algorithmic HBM traffic. */
typedef struct rsf_kernel_timing {
  double spmv_ms, spmv_bytes;             /* 7-point 3x3 block SpMV, FP64 */
  double bilu_ms, bilu_bytes;             /* block-ILU0 apply, configured precision and ordering */
  int bilu_launches;                      /* kernel launches per apply (2 x levels) */
  double pilu_ms, pilu_bytes;             /* scalar pressure ILU0 apply */
  double pspmv_ms, pspmv_bytes;           /* scalar pressure SpMV */
  double assembly_ms, assembly_bytes;     /* cell + face assembly kernels */
  double copy_ms, copy_bytes;             /* device-to-device copy of the Jacobian (bandwidth probe) */
} rsf_kernel_timing;
int rsf_profile_kernels(rsf_sim* sim, int reps, rsf_kernel_timing* out);
/* EXTENSION: discrete adjoint of the FIM step (no reference counterpart; it
 * differentiates the step that run_fim_timestep, fim/newton.rs:970-976, solves).
 * rsf_adjoint_record starts recording the following accepted substeps (from rsf_step or
 * rsf_run_timestep), storing the state before every checkpoint_interval-th substep;
 * 0 stops recording. The objective is the surface oil produced by the producers
 * over the recorded substeps, J = sum_n dt_n sum_completions (fo/Bo)(x_n) q_n.
 * rsf_adjoint_gradient runs the backward pass (replaying each checkpointed segment,
 * recomputing the Jacobians and solving the transposed systems) and returns
 * dJ/d ln k per cell (a log-multiplier applied to kx, ky and kz together) and
 * dJ/d phi per cell. The simulator's state is the same afterwards. */
int rsf_adjoint_record(rsf_sim* sim, int checkpoint_interval);
int rsf_adjoint_objective(rsf_sim* sim, double* value, int* n_substeps);
typedef struct rsf_adjoint_report {
  int substeps;                  /* recorded substeps walked back */
  long linear_iterations;        /* FGMRES iterations of the transposed solves */
  long pressure_iterations;      /* BiCGStab iterations of the pass (replay included) */
  double replay_ms, solve_ms, total_ms; /* forward replay, transposed solves, whole pass */
} rsf_adjoint_report;
int rsf_adjoint_gradient(rsf_sim* sim, double* grad_lnk, double* grad_phi, int n, rsf_adjoint_report* report);
/* EXTENSION (measurement, no reference counterpart): live timing of the BiCGStab kernel
 * with the largest share of GPU time (k_rb_a in red-black ordering with nx even, the
 * pressure SpMV otherwise) inside the steps that follow. enable = 1: every launch of an
 * executed iteration carries HIP start/stop events that its dispatch timestamps
 * (hipExtLaunchKernel on the handle's stream); enable = k > 1: every k-th iteration's
 * launch only, which keeps the cost within a throughput measurement; 0: off. Enabling
 * resets the counters. */
typedef struct rsf_kernel_probe {
  long launches;
  double total_ms, avg_ms;
  double bytes_per_launch; /* algorithmic HBM bytes of one launch */
  char kernel[64];         /* the timed kernel */
} rsf_kernel_probe;
int rsf_set_kernel_probe(rsf_sim* sim, int enable);
int rsf_get_kernel_probe(rsf_sim* sim, rsf_kernel_probe* out);
/* EXTENSION: how the natural-order (ilu_ordering 0) ILU0 solves of the CPR preconditioner
 * run; the factors and solves are those of FimBlockIlu0Factors::apply
 * (fim/linear/gmres_block_jacobi.rs:190-237) and apply_pressure_ilu (:634-658) either way.
 * mode 1 (default): one launch per factorization (factorize_block_ilu0 :239-337,
 * factorize_pressure_ilu0 :746-808) and per forward or backward solve, tiles of
 * lanes_pressure / lanes_block grid lines (at most 256) per workgroup marching along x
 * (grids with nx + tile edges above 65535 fall back to mode 0); mode 0: one launch per
 * hyperplane level. Both give the same numbers. Takes effect at the next preconditioner
 * build. */
int rsf_set_ilu_sweep(rsf_sim* sim, int mode, int lanes_pressure, int lanes_block);
/* Step reporting: record_fim_step_report (reporting.rs:553-734) appends one
 * TimePointRates (reporting.rs:246-292) per accepted substep of rsf_step or
 * rsf_run_timestep: surface and reservoir rates [m3/day] at the accepted state,
 * cumulative material-balance errors [sc m3], PV-weighted average pressure [bar], mean
 * saturations, producing GOR and the fractions of rate-controlled wells held at their BHP
 * limit. The optional sweep metrics are not reported. rsf_rate_history_len gives the
 * count (-1 for a null handle); rsf_get_rate_history copies the first min(cap, count)
 * entries (getRateHistory, frontend.rs). */
typedef struct rsf_time_point {
  double time;
  double total_production_oil, total_production_liquid, total_production_liquid_reservoir;
  double total_injection, total_injection_reservoir;
  double material_balance_error_m3, material_balance_error_oil_m3;
  double avg_reservoir_pressure, avg_water_saturation;
  double total_production_gas, avg_gas_saturation, material_balance_error_gas_m3, producing_gor;
  double producer_bhp_limited_fraction, injector_bhp_limited_fraction;
  /* sweep efficiency, present when a sweep configuration is set (SweepMetrics,
   * reporting.rs:32-47); NaN where the reference has None */
  double sweep_e_a, sweep_e_v, sweep_e_vol, sweep_mobile_oil_recovered;
} rsf_time_point;
int rsf_rate_history_len(rsf_sim* sim);
int rsf_get_rate_history(rsf_sim* sim, rsf_time_point* out, int cap);
/* frontend.rs:568 get_rate_history_since: the entries from min(start_index, count) on;
 * copies the first min(cap, remaining) and returns the number remaining (-1 on invalid
 * arguments). start_index = count - 1 gives getLatestRatePoint (frontend.rs:579). */
int rsf_get_rate_history_since(rsf_sim* sim, int start_index, rsf_time_point* out, int cap);
/* frontend.rs:1216 set_sweep_config: sweep metrics computed at every report entry
 * (reporting.rs:52-123). geometry "areal", "vertical" or "both" (any other name acts as
 * "both"); a sat_water above swept_threshold marks a cell swept. NULL geometry turns the
 * metrics off. */
int rsf_set_sweep_config(rsf_sim* sim, const char* geometry, double swept_threshold, double initial_oil_saturation,
                         double residual_oil_saturation);
/* frontend.rs:931-1006 load_state: time, the stored cell state (sat_gas and rs may be NULL
 * for zeros), the well state as one BHP per completion (the reference's Vec<Well>; the
 * completions must already exist, in the order they were added) and the rate history,
 * whose last entry restores the cumulative reservoir injection and production. */
int rsf_load_state(rsf_sim* sim, double time_days, const double* pressure, const double* sat_water,
                   const double* sat_oil, const double* sat_gas, const double* rs, int n_cells,
                   const double* completion_bhp, int n_completions, const rsf_time_point* rate_history,
                   int n_history);
/* Linear-system capture, the fim-capture-v2 text format of fim/linear/capture.rs:149-640
 * (write_capture, parse_capture). The systems written are the ones the linear stack
 * solves: after the Schur elimination of the well tail, so their layout line is
 * "N 3 0 3N". mode 0: off; 1: every linear solve, to <dir>/fim_capture_<seq>.txt, as the
 * reference's FIM_CAPTURE_SEQUENCE_DIR (capture.rs:33); 2: only solves that do not
 * converge, as FIM_CAPTURE_DIR (capture.rs:26). */
int rsf_set_capture_dir(rsf_sim* sim, const char* dir, int mode);
typedef struct rsf_linear_report {
  int converged;
  int iterations;       /* FGMRES iterations */
  double residual_norm; /* of the reduced (Schur) system at exit */
  double rhs_norm;      /* of the full captured right-hand side */
} rsf_linear_report;
/* Replays a captured system J x = rhs (CSR with 0-based indices, the reference's
 * unknown order: 3 interleaved unknowns per cell, then the well tail) through the GPU
 * linear stack with the handle's solver options, as the reference's offline solver lab
 * replays captures through its CPU stack (fim/linear/solver_lab.rs). Cell couplings
 * off the 7-point stencil of the handle's grid are carried like the wells' cliques;
 * the tail (at most 4096 unknowns) is eliminated by its Schur complement and
 * recovered after the solve.
 * The model's wells are rebuilt at the next step. */
int rsf_solve_captured_system(rsf_sim* sim, int rows, const int* row_ptr, const int* col, const double* val,
                              const double* rhs, double* x, rsf_linear_report* report);
/* cnv_mb_from_parts (fim/newton/convergence.rs:478-612) on device arrays, without a
 * handle: the OPM CNV/MB acceptance measures of an iterate. residual holds the 3n cell
 * rows in the reference's order (index 3 i + c, c = water, oil, gas), pore_volumes the n
 * reference pore volumes, fvf the n [B_w, B_o, B_g] triples (index 3 i + c); relax_final
 * selects the final-iteration tolerances. out (device, 10 doubles) receives cnv[3],
 * mb[3], the violating pore-volume fraction, the 3 %-PV rule, would_accept and
 * would_accept_strict (the last three 0 or 1). Every pointer is device memory; the
 * kernels run on `stream` (a hipStream_t, NULL for the null stream). The reference's
 * peak-cell and binding-criterion diagnostics are not computed. Returns RSF_ERR_INVALID
 * for a host pointer or n outside [0, 2^28], with the message in rsf_cnv_mb_last_error.
 * It allocates its block records for the call and waits for the stream before freeing
 * them. */
int rsf_cnv_mb(const double* residual, const double* pore_volumes, const double* fvf, int n, int relax_final,
               double* out, void* stream);
/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * rsf_cnv_mb without any allocation and without waiting for the stream. Its block records
 * live in a fixed device array of the library (90,240 bytes per device, held until the
 * process ends), so any n in [0, RSF_CNV_MB_NOALLOC_MAX_CELLS] fits; a larger n is rejected
 * with RSF_ERR_INVALID. Calls that overlap on one device (several host threads, or
 * several streams in flight) share that array and race: serialise them. Results are
 * bit-identical to rsf_cnv_mb's. */
#define RSF_CNV_MB_NOALLOC_MAX_CELLS (1 << 28)
int rsf_cnv_mb_noalloc(const double* residual, const double* pore_volumes, const double* fvf, int n,
                       int relax_final, double* out, void* stream);
/* message of the calling thread's last failed rsf_cnv_mb* call */
int rsf_cnv_mb_last_error(char* buf, int len);
const char* rsf_version(void);

#ifdef __cplusplus
}
#endif
#endif
