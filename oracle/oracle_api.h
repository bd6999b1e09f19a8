/* Oracle C API: the CPU FP64 restatement of the reference FIM timestep.
 * TEST INFRASTRUCTURE ONLY: tests, the smoke check and the bench's
 * CPU-baseline leg load it; the product library never does. */
#ifndef ORACLE_API_H
#define ORACLE_API_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

orc_sim* orc_create(int nx, int ny, int nz, double phi);
void orc_destroy(orc_sim* sim);
int orc_set_cell_dimensions_per_layer(orc_sim* sim, double dx, double dy, const double* dz);
int orc_set_fields(orc_sim* sim, const double* phi, const double* kx, const double* ky, const double* kz);
int orc_set_pc_scale(orc_sim* sim, const double* scale);
int orc_set_initial_state(orc_sim* sim, const double* p, const double* sw, const double* sg, const double* rs);
int orc_set_fluid(orc_sim* sim, double mu_o, double mu_w, double c_o, double c_w, double rho_o, double rho_w);
int orc_set_reference_pressure(orc_sim* sim, double pressure);
int orc_set_rock_properties(orc_sim* sim, double c_r, double depth_ref, double b_o, double b_w);
int orc_set_gas_fluid_properties(orc_sim* sim, double mu_g, double c_g, double rho_g);
int orc_set_rel_perm_props(orc_sim* sim, double s_wc, double s_or, double n_w, double n_o, double krw, double kro);
int orc_set_three_phase_rel_perm_props(orc_sim* sim, double s_wc, double s_or, double s_gc, double s_gr,
                                       double s_org, double n_w, double n_o, double n_g, double krw, double kro,
                                       double krg);
int orc_set_capillary_params(orc_sim* sim, double p_entry, double lambda);
/* SWOF rows (sw, krw, krow) and SGOF rows (sg, krg, krog); 1 without three-phase props */
int orc_set_scal_tables(orc_sim* sim, const double* swof, int n_swof, const double* sgof, int n_sgof);
int orc_set_gas_oil_capillary_params(orc_sim* sim, double p_entry, double lambda);
int orc_set_flags(orc_sim* sim, int three_phase, int gravity, int corey_points, int injected_fluid);
int orc_set_pvt_table(orc_sim* sim, const double* rows, int n_rows);
int orc_set_well_controls(orc_sim* sim, int inj_rate, int prod_rate, int inj_enabled, double tgt_inj,
                          double tgt_prod, double tgt_inj_surf, double tgt_prod_surf, double bhp_min,
                          double bhp_max);
int orc_add_well(orc_sim* sim, int i, int j, int k, double bhp, double rw, double skin, int injector,
                 const char* id);
int orc_set_well_schedule(orc_sim* sim, const char* id, int rate_mode, double target, double surface,
                          double limit, int enabled);
int orc_set_well_datum(orc_sim* sim, const char* id, double datum, double density);
int orc_set_well_fluid(orc_sim* sim, const char* id, int fluid);
int orc_set_gas_redissolution(orc_sim* sim, int enabled);
int orc_set_newton_options(orc_sim* sim, int max_it, int min_it, double residual_tol, double lin_rel_tol,
                           int lin_max_it, int allow_direct);
int orc_step(orc_sim* sim, double dt);
int orc_run_substep(orc_sim* sim, double dt, int* newton_its, long* linear_its);
int orc_get_field(orc_sim* sim, int which, double* out);
int orc_num_wells(orc_sim* sim);
int orc_num_perfs(orc_sim* sim);
int orc_get_well_state(orc_sim* sim, double* bhp, double* q);
double orc_get_time(orc_sim* sim);
int orc_get_stats(orc_sim* sim, long* out);
int orc_get_accepted_dts(orc_sim* sim, double* out, int cap);
/* FimStepStats per step call (rs/reporting.rs:141-243); the layout of the library's
 * rsf_step_stats / rsf_step_rung, codes and rows as there (-1 / NaN: absent) */
typedef struct orc_step_stats {
  double time_days, target_dt_days, advanced_dt_days;
  int accepted_substeps, linear_bad_retries, nonlinear_bad_retries, mixed_retries;
  double min_accepted_dt_days, max_accepted_dt_days, last_accepted_dt_days, last_growth_factor;
  int growth_limiter, last_retry_class, last_retry_dominant_family, n_rungs;
  double solver_ms, accepted_solver_ms, retry_solver_ms, assembly_ms, linear_solve_ms;
  long last_retry_dominant_row;
} orc_step_stats;
typedef struct orc_step_rung {
  int substep, retried, newton_iterations, linear_iterations, linear_solve_count;
  int retry_class, dominant_family, last_linear_converged;
  long linear_iteration_count;
  double dt_days, linear_solve_ms;
  long dominant_row;
} orc_step_rung;
int orc_step_stats_count(orc_sim* sim);
int orc_classify_retry(int have_linear, int converged, int direct, int coarse_applications, double average_reduction,
                       double last_reduction);
int orc_dominant_family(const double* family_maxima, int with_wells);
/* residual_family_diagnostics (rs/fim/newton/convergence.rs:734-790) of a residual in the
 * reference's unknown numbering (3 per cell, then wells, then completions) with one scale
 * per row: each family's peak value and row (-1 with no rows), and the global family */
int orc_family_peaks(int n_cells, int n_wells, int n_perf, const double* residual, const double* scale,
                     double* peak_values, long* peak_rows);
int orc_get_step_stats(orc_sim* sim, int index, orc_step_stats* out);
int orc_get_step_rungs(orc_sim* sim, int index, orc_step_rung* out, int cap);
int orc_assemble(orc_sim* sim, double dt, const double* cell_state, const int* regimes, const double* bhp,
                 const double* q, int* n_unknowns, int* nnz);
int orc_get_assembly(double* R, int* ptr, int* col, double* val);
/* surface oil produced by the producers over the accepted substeps (adjoint objective) */
double orc_producer_oil(orc_sim* sim);
int orc_rate_history_len(orc_sim* sim);
int orc_get_rate_history(orc_sim* sim, double* out, int cap);
/* component inventories [water, oil, gas] (sc m3) summed over the stored state */
int orc_inventory_totals(orc_sim* sim, double* out);
int orc_residual_check(orc_sim* sim, double dt, const double* cell_state, const int* regimes, const double* bhp,
                       const double* q, double* out);
/* |R| / scale per row at an iterate (equation_scaling, rs/fim/scaling.rs:139-187), rows in
 * the reference's unknown numbering */
int orc_scaled_residual(orc_sim* sim, double dt, const double* cell_state, const int* regimes, const double* bhp,
                        const double* q, double* out);
int orc_get_initial_state(orc_sim* sim, double* cell_state, int* regimes);
int orc_get_initial_well_state(orc_sim* sim, double* bhp, double* q);
int orc_probe(orc_sim* sim, int what, double x, double y, double* out);
const char* orc_last_warning(orc_sim* sim);
int orc_chop(int n, const double* cells, const int* regimes, int n_wells, const double* bhp, const double* update,
             double relax, double* out);
int orc_cnv_mb(int n, const double* R, const double* pv, const double* fvf, int relax_final, double* out);
int orc_ad_selftest(double* out);
int orc_newton_direction(orc_sim* sim, double dt, const double* cell_state, const int* regimes, const double* bhp,
                         const double* q, double rel_tol, int allow_direct, double* dx, int* iterations);

/* sweep metrics of every report entry (rs/reporting.rs:52-123); NULL geometry: off */
int orc_set_sweep_config(orc_sim* sim, const char* geometry, double threshold, double initial_so, double sor);
/* nested per-well Newton (rs/fim/wells_inner.rs) */
int orc_set_nested_well_solve(orc_sim* sim, int enabled);
/* one well's local residual and row-major Jacobian; returns the local dimension */
int orc_well_local_system(orc_sim* sim, const double* cell_state, const int* regimes, const double* bhp,
                          const double* q, int well, double* r, double* J);
/* solve_wells_locally at an iterate; bhp and q are updated in place; max_iterations < 0 keeps 50 */
int orc_solve_wells_locally(orc_sim* sim, const double* cell_state, const int* regimes, double* bhp, double* q,
                            int max_iterations, int* converged, int* iterations, double* peak);
double orc_chop_bhp_update(double bhp, double raw_delta, double rel);
/* the substep controller and Newton relaxation pieces (rs/fim/timestep.rs:436-528,
 * rs/fim/newton/damping.rs:715-766, rs/fim/state.rs:494-551) */
int orc_relax_wells(orc_sim* sim, const double* cell_state, const int* regimes, double* bhp, double* q);
double orc_pid_dt(double dt, const double* err);
double orc_iteration_count_dt(double dt, int its);
double orc_growth_factor(double dt, int its, const double* err, int retries, int* limiter);
int orc_detect_oscillation(const double* cur, const double* p1, const double* p2);
double orc_next_relaxation_factor(double relax, int oscillating);
double orc_relative_change(int n, const double* a, const int* ra, const double* b, const int* rb);
/* threads of the oracle's parallel loops (n <= 0 keeps the OpenMP default); returns the count in
 * effect. Results do not depend on it: every sum keeps its sequential order. */
int orc_set_threads(int n);
/* linear-stack components on caller matrices (CSR, 0-based): scalar ILU0 (L U densely and the
 * apply), block ILU0 apply on 3x3 cell blocks, and the CPR pieces (weights, coarse matrix, a
 * pressure solve or the full apply); see oracle.cpp */
int orc_scalar_ilu0(int n, const int* ptr, const int* col, const double* val, const double* rhs, double* lu,
                    double* x);
int orc_block_ilu0_apply(int nb, const int* ptr, const int* col, const double* val, const double* rhs, double* x);
int orc_cpr_probe(int nb, const int* ptr, const int* col, const double* val, const double* rp, int mode, int max_it,
                  double tol, double* weights, double* pmat, double* xp, int* info);
/* discrete-adjoint terms (see oracle.cpp and oracle/adjoint.py) */
int orc_last_iterate(orc_sim* sim, double* cells, int* regimes, double* bhp, double* q);
int orc_assemble_prev(orc_sim* sim, double dt, const double* prev_cells, const int* prev_regimes, const double* cells,
                      const int* regimes, const double* bhp, const double* q, int* n_unknowns, int* nnz);
int orc_adjoint_state_terms(orc_sim* sim, double dt, const double* cells, const int* regimes, const double* bhp,
                            const double* q, double* j, double* djdx, double* dinv);
int orc_adjoint_param_terms(orc_sim* sim, double dt, const double* prev_cells, const int* prev_regimes,
                            const double* cells, const int* regimes, const double* bhp, const double* q,
                            const double* lambda, double* g_lnk, double* g_phi);
#ifdef __cplusplus
}
#endif
#endif
