// Oracle: CPU FP64 restatement of the reference FIM timestep (AD assembly,
// well Schur + CPR-FGMRES linear stack, OPM-aligned Newton, substep
// controller). TEST INFRASTRUCTURE ONLY: loaded by tests, the smoke check and
// the bench's CPU-baseline leg, never by the product library.
#include <chrono>
#include <limits>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "linear.hpp"
#include "model.hpp"
#include "oracle_api.h"

using namespace orc;

namespace {

struct CellState {
  double p, sw, hc;
  int regime;
};

struct State {  // rs/fim/state.rs:101-106
  std::vector<CellState> cells;
  std::vector<double> bhp;  // per physical well
  std::vector<double> q;    // per perforation (completion entry order)
};

struct NewtonOpts {  // rs/fim/newton.rs:878-919 (OpmAligned defaults)
  int max_iterations = 20;
  int min_iteration_index = 2;  // rs/fim/newton/convergence.rs:378
  double residual_tol = 0.0;    // >0: scaled-residual acceptance instead of OPM CNV
  double mb_tol = 1e-7;
  bool nested_well_solve = false;  // rs/fim/newton.rs:890-894 (default off: relaxation)
  LinOpts lin;
};

struct StepReport {
  bool converged = false;
  int newton_iterations = 0;
  long linear_iterations = 0;
  double final_scaled_residual = 0.0;
  int fail_class = -1, fail_family = -1;  // a rejected attempt's class and dominant family
  long fail_row = -1;                      // the row of that family's peak
  int lin_solves = 0, last_lin_its = -1, last_lin_conv = -1;
};

struct StepStatsRec {  // rs/reporting.rs:141-243 FimStepStats with its rungs
  orc_step_stats s;
  std::vector<orc_step_rung> rungs;
};

// rs/reporting.rs:24-35 SweepConfig; geometry 0 areal, 1 vertical, 2 both
struct SweepConfig {
  bool on = false;
  int geometry = 2;
  double threshold = 0.0, initial_so = 0.0, sor = 0.0;
};

// rs/reporting.rs:246-292 TimePointRates; the sweep metrics NaN where the reference has None
struct TimePoint {
  double time;
  double total_production_oil, total_production_liquid, total_production_liquid_reservoir;
  double total_injection, total_injection_reservoir;
  double material_balance_error_m3, material_balance_error_oil_m3;
  double avg_reservoir_pressure, avg_water_saturation;
  double total_production_gas, avg_gas_saturation, material_balance_error_gas_m3, producing_gor;
  double producer_bhp_limited_fraction, injector_bhp_limited_fraction;
  double sweep_e_a, sweep_e_v, sweep_e_vol, sweep_mobile_oil_recovered;
};

struct Sim {
  Model m;
  SweepConfig sweep;
  std::vector<TimePoint> rate_history;
  double cum_mb_water = 0.0, cum_mb_oil = 0.0, cum_mb_gas = 0.0, cum_injection_m3 = 0.0, cum_production_m3 = 0.0;
  NewtonOpts nopts;
  LinStats lstats;
  std::string warning;
  std::vector<double> accepted_dts;
  std::vector<StepStatsRec> step_stats;
  long total_newton = 0, total_linear = 0, total_substeps = 0, total_retries = 0;
  double last_assembly_ms = 0, last_linear_ms = 0;
  double assembly_ms = 0, linear_ms = 0;
  double producer_oil = 0.0;  // surface oil produced over the accepted substeps
  State last_iterate;         // converged iterate of the last accepted substep, before commit
};

int n_cells(const Model& m) { return m.n(); }

// ------------------------------------------------------------------ wells
template <class S>
struct WCell {
  S p, sw, hc;
  int regime;
  int id;  // cell index (for the Rs cap when gas redissolution is disabled)
};

template <class S>
void producer_fractions(const Model& m, const WCell<S>& c, S f[3]) {  // rs/fim/wells_ad.rs:52-80
  auto pr = m.cell_props(c.regime, c.p, c.sw, c.hc, m.rs_cap_of(c.id));
  auto mob = m.mobilities(c.sw, pr.sg, c.p, pr.rs);
  S lw = max_floor(mob.w, 0.0), lo = max_floor(mob.o, 0.0), lg = max_floor(mob.g, 0.0);
  S lt = max_floor(lw + lo + lg, EPS);
  f[0] = min_ceil(max_floor(lw / lt, 0.0), 1.0);
  f[1] = min_ceil(max_floor(lo / lt, 0.0), 1.0);
  f[2] = min_ceil(max_floor(lg / lt, 0.0), 1.0);
}

template <class S>
S connection_rate(const Model& m, double wi, double head, bool injector, const WCell<S>& c, const S& bhp) {
  // rs/fim/wells_ad.rs:146-190 (strict-sign clamp)
  auto pr = m.cell_props(c.regime, c.p, c.sw, c.hc, m.rs_cap_of(c.id));
  auto mob = m.mobilities(c.sw, pr.sg, c.p, pr.rs);
  S cm = max_floor(mob.w + mob.o + mob.g, 0.0);
  S raw = (cm * (c.p - bhp - head)) * wi;
  if (injector) return val(raw) < 0.0 ? raw : cst(raw, 0.0);
  return val(raw) > 0.0 ? raw : cst(raw, 0.0);
}

template <class S>
void rate_coefficients(const Model& m, bool injector, int fluid, const WCell<S>& c, S coef[3]) {
  // rs/fim/wells_ad.rs:193-238
  auto pr = m.cell_props(c.regime, c.p, c.sw, c.hc, m.rs_cap_of(c.id));
  if (injector) {
    if (fluid == 0) {
      coef[0] = m.inv_bw(c.p);
      coef[1] = cst(c.p, 0.0);
      coef[2] = cst(c.p, 0.0);
    } else {
      coef[0] = cst(c.p, 0.0);
      coef[1] = cst(c.p, 0.0);
      coef[2] = recip(max_floor(pr.bg, 1e-9));
    }
    return;
  }
  S f[3];
  producer_fractions(m, c, f);
  S bo = max_floor(pr.bo, 1e-9), bg = max_floor(pr.bg, 1e-9);
  S oc = f[1] / bo;
  coef[0] = f[0] * m.inv_bw(c.p);
  coef[1] = oc;
  coef[2] = f[2] / bg + oc * pr.rs;
}

template <class S>
S perforation_rate(const Model& m, bool injector, int fluid, bool surface, const WCell<S>& c, const S& q) {
  // rs/fim/wells_ad.rs:247-294 (surface) and :298-304 (reservoir)
  if (!surface) return injector ? max_floor(-q, 0.0) : max_floor(q, 0.0);
  auto pr = m.cell_props(c.regime, c.p, c.sw, c.hc, m.rs_cap_of(c.id));
  if (injector) {
    S cl = max_floor(-q, 0.0);
    return fluid == 0 ? cl * m.inv_bw(c.p) : cl / max_floor(pr.bg, 1e-9);
  }
  S f[3];
  producer_fractions(m, c, f);
  return max_floor(q, 0.0) * f[1] / max_floor(pr.bo, 1e-9);
}

WCell<double> wcell(const State& s, int id) {
  const auto& c = s.cells[id];
  return {c.p, c.sw, c.hc, c.regime, id};
}

// Total well rate for a trial BHP (rs/fim/wells.rs:389-408), f64.
double total_rate_for_bhp(const Model& m, const State& s, int w, double bhp) {
  Control c = m.control(w);
  bool inj = m.topo_wells[w].injector;
  int fluid = m.fluid_of(w);
  double tot = 0.0;
  for (int e : m.topo_wells[w].perfs) {
    double wi;
    if (!m.well_index(e, wi)) continue;
    const Well& wl = m.wells[e];
    auto cell = wcell(s, m.idx(wl.i, wl.j, wl.k));
    double q = connection_rate<double>(m, wi, wl.head_offset, inj, cell, bhp);
    if (!std::isfinite(q)) continue;
    tot += perforation_rate<double>(m, inj, fluid, c.uses_surface, cell, q);
  }
  return tot;
}

// rs/fim/wells.rs:431-504 solve_bhp_from_target
bool solve_bhp_from_target(const Model& m, const State& s, int w, double& bhp_out) {
  Control c = m.control(w);
  if (!c.enabled || !c.rate_controlled || !c.has_target) return false;
  const auto& perfs = m.topo_wells[w].perfs;
  if (perfs.empty()) return false;
  double mn = INFINITY, mx = -INFINITY;
  for (int e : perfs) {
    const Well& wl = m.wells[e];
    double z = s.cells[m.idx(wl.i, wl.j, wl.k)].p - wl.head_offset;
    mn = std::min(mn, z);
    mx = std::max(mx, z);
  }
  if (!std::isfinite(mn) || !std::isfinite(mx)) return false;
  double tgt = c.target;
  double lim = c.bhp_limit;
  double max_rate = total_rate_for_bhp(m, s, w, lim);
  if (tgt >= max_rate - 1e-9) {
    bhp_out = lim;
    return true;
  }
  bool inj = m.topo_wells[w].injector;
  double lo, hi;
  if (inj) { lo = std::min(mn, lim); hi = lim; }
  else { lo = lim; hi = std::max(mx, lim); }
  for (int it = 0; it < 64; ++it) {
    double mid = 0.5 * (lo + hi);
    double r = total_rate_for_bhp(m, s, w, mid);
    if (inj) { if (r < tgt) lo = mid; else hi = mid; }
    else { if (r > tgt) lo = mid; else hi = mid; }
  }
  bhp_out = 0.5 * (lo + hi);
  return true;
}

double connection_rate_state(const Model& m, const State& s, int e, double bhp) {  // rs/fim/wells.rs:1098-1127
  double wi;
  if (!m.well_index(e, wi)) return 0.0;
  const Well& wl = m.wells[e];
  auto cell = wcell(s, m.idx(wl.i, wl.j, wl.k));
  double q = connection_rate<double>(m, wi, wl.head_offset, wl.injector, cell, bhp);
  return std::isfinite(q) ? q : 0.0;
}

// rs/fim/state.rs:433-459
void enforce_control_bounds(const Model& m, State& s) {
  double pu = std::max(m.bhp_max, 1.0);
  for (auto& c : s.cells) pu = std::max(pu, c.p);
  pu += 500.0;
  for (size_t w = 0; w < s.bhp.size(); ++w) {
    Control c = m.control((int)w);
    double& b = s.bhp[w];
    if (m.topo_wells[w].injector) {
      double hi = std::max(pu, std::max(c.bhp_limit, m.bhp_max));
      b = std::min(std::max(b, 1e-6), hi);
    } else {
      double lo = std::max(std::min(c.bhp_limit, m.bhp_min), 1e-6);
      b = std::min(std::max(b, lo), pu);
    }
  }
}

// rs/fim/state.rs:494-551 (blend 0.9 / 25 bar; 0.75 / 10% / 250 m3/d)
void relax_wells(const Model& m, State& s) {
  for (size_t w = 0; w < s.bhp.size(); ++w) {
    Control c = m.control((int)w);
    double consistent;
    if (!c.enabled) consistent = c.bhp_target;
    else if (c.rate_controlled) {
      if (!solve_bhp_from_target(m, s, (int)w, consistent)) continue;
    } else consistent = c.bhp_target;
    double prop = s.bhp[w];
    double blended = prop + 0.9 * (consistent - prop);
    s.bhp[w] = std::max(consistent + std::min(std::max(blended - consistent, -25.0), 25.0), 1e-6);
    for (int e : m.topo_wells[w].perfs) {
      double cq = c.enabled ? connection_rate_state(m, s, e, s.bhp[w]) : 0.0;
      double pq = s.q[e];
      double bq = pq + 0.75 * (cq - pq);
      double tr = std::max(0.1 * std::fabs(cq), 250.0);
      double q = cq + std::min(std::max(bq - cq, -tr), tr);
      s.q[e] = !c.enabled ? 0.0 : (m.topo_wells[w].injector ? std::min(q, 0.0) : std::max(q, 0.0));
    }
  }
}

// ------------------------------------------------------------------ nested well solve
// rs/fim/wells_inner.rs:92-246 assemble_well_local_system: one well's control row and its
// perforations' rate-consistency rows in the well's own unknowns, the cells frozen. Local
// order: row/unknown 0 is the well (bhp), 1 + t the t-th perforation (q). The rows come from
// the residual functions of the global assembly, with AD slots [bhp, q].
struct WellLocal {
  int dim;
  std::vector<double> r, J;  // J row-major, dim x dim
};

WellLocal well_local_system(const Model& m, const State& s, int w) {
  using A2 = Ad<2>;
  const auto& perfs = m.topo_wells[w].perfs;
  const int np = (int)perfs.size(), dim = 1 + np;
  WellLocal L{dim, std::vector<double>(dim, 0.0), std::vector<double>((size_t)dim * dim, 0.0)};
  const Control c = m.control(w);
  const bool inj = m.topo_wells[w].injector;
  const int fluid = m.fluid_of(w);
  const double bhp = s.bhp[w];
  auto frozen = [&](int e) {
    const Well& wl = m.wells[e];
    const int id = m.idx(wl.i, wl.j, wl.k);
    const auto& cc = s.cells[id];
    return WCell<A2>{A2(cc.p), A2(cc.sw), A2(cc.hc), cc.regime, id};
  };
  if (!c.enabled || !c.rate_controlled) {
    L.r[0] = bhp - c.bhp_target;
    L.J[0] = 1.0;
  } else if (c.has_target) {
    // Fischer-Burmeister complementarity of the BHP and rate slacks (rs/fim/wells_ad.rs:538-572)
    double actual = 0.0;
    std::vector<double> drq(np);
    for (int t = 0; t < np; ++t) {
      const int e = perfs[t];
      A2 rate = perforation_rate<A2>(m, inj, fluid, c.uses_surface, frozen(e), A2::var(s.q[e], 1));
      actual += rate.v;
      drq[t] = rate.d[1];
    }
    const double sb = std::max(std::fabs(c.bhp_limit), 1.0), sr = std::max(std::fabs(c.target), 1.0);
    const double xa = (inj ? c.bhp_limit - bhp : bhp - c.bhp_limit) / sb;
    const double xb = (c.target - actual) / sr;
    const double nrm = std::sqrt((xa * xa + xb * xb) + 2.0 * FB_EPSILON * FB_EPSILON);
    const double dfa = xa / nrm - 1.0, dfb = xb / nrm - 1.0;
    L.r[0] = nrm - xa - xb;
    L.J[0] = dfa * (inj ? -1.0 : 1.0) / sb;
    for (int t = 0; t < np; ++t) L.J[1 + t] = dfb * (-1.0 / sr) * drq[t];
  }
  for (int t = 0; t < np; ++t) {
    const int e = perfs[t];
    double wi;
    if (!m.well_index(e, wi)) continue;
    A2 conn = connection_rate<A2>(m, wi, m.wells[e].head_offset, inj, frozen(e), A2::var(bhp, 0));
    L.r[1 + t] = s.q[e] - conn.v;
    L.J[(size_t)(1 + t) * dim + 1 + t] = 1.0;
    L.J[(size_t)(1 + t) * dim] = -conn.d[0];
  }
  return L;
}

// Dense LU with partial pivoting and its solve (the reference's nalgebra `lu().solve`):
// false when a pivot is exactly zero.
bool dense_lu_solve(std::vector<double> A, std::vector<double>& b, int n) {
  for (int k = 0; k < n; ++k) {
    int piv = k;
    double best = std::fabs(A[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[(size_t)i * n + k]) > best) {
        best = std::fabs(A[(size_t)i * n + k]);
        piv = i;
      }
    if (best == 0.0) return false;
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)piv * n + j]);
      std::swap(b[k], b[piv]);
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[(size_t)i * n + k] / A[(size_t)k * n + k];
      for (int j = k + 1; j < n; ++j) A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
      b[i] -= f * b[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double v = b[i];
    for (int j = i + 1; j < n; ++j) v -= A[(size_t)i * n + j] * b[j];
    b[i] = v / A[(size_t)i * n + i];
  }
  return true;
}

// rs/fim/wells_inner.rs:247-258 FimWellInnerSolveOptions::default (OPM MaxInnerIterWells,
// ToleranceWells, DbhpMaxRel)
struct WellInnerOpts {
  int max_iterations = 50;
  double tolerance = 1e-4;
  double dbhp_max_rel = 1.0;
};

struct WellInnerReport {  // rs/fim/wells_inner.rs:260-267
  bool converged = false;
  int iterations = 0;
  double peak = 0.0;
};

// rs/fim/wells_inner.rs:432-436 chop_bhp_update: |dbhp| <= rel |bhp|, then the 1 bar - 1 Pa floor
double chop_bhp_update(double bhp, double raw_delta, double rel) {
  const double cap = std::fabs(bhp) * rel;
  const double lim = raw_delta < -cap ? -cap : (raw_delta > cap ? cap : raw_delta);
  return std::fmax(bhp + lim, 1.0 - 1.0e-5);
}

// rs/fim/wells_inner.rs:474-521 well_convergence_status: scaled residual peak (row scales of
// rs/fim/scaling.rs:107-119) and, for a pressure-controlled well, the flow direction of every
// perforation (injectors q <= 0, producers q >= 0).
bool well_status(const Model& m, const State& s, int w, const WellLocal& L, const WellInnerOpts& o, double& peak) {
  const Control c = m.control(w);
  const bool inj = m.topo_wells[w].injector;
  const bool slacks = c.enabled && c.rate_controlled && c.has_target;
  const double wscale = slacks ? 1.0 : std::max(std::fabs(s.bhp[w]), 1.0);
  peak = std::fabs(L.r[0]) / wscale;
  bool direction = true;
  const auto& perfs = m.topo_wells[w].perfs;
  for (size_t t = 0; t < perfs.size(); ++t) {
    const double q = s.q[perfs[t]];
    peak = std::max(peak, std::fabs(L.r[1 + t]) / std::max(std::fabs(q), 1.0));
    direction = direction && (inj ? q <= 0.0 : q >= 0.0);
  }
  return peak <= o.tolerance && (c.rate_controlled || direction);
}

// rs/fim/wells_inner.rs:523-583 solve_well_locally: chopped Newton on the local rows; a
// singular local Jacobian or an exhausted budget keeps the last iterate, not converged.
WellInnerReport solve_well_locally(const Model& m, State& s, int w, const WellInnerOpts& o) {
  WellInnerReport rep;
  const auto& perfs = m.topo_wells[w].perfs;
  while (true) {
    WellLocal L = well_local_system(m, s, w);
    rep.converged = well_status(m, s, w, L, o, rep.peak);
    if (rep.converged || rep.iterations >= o.max_iterations) break;
    std::vector<double> d(L.dim);
    for (int k = 0; k < L.dim; ++k) d[k] = -L.r[k];
    if (!dense_lu_solve(L.J, d, L.dim)) break;
    s.bhp[w] = chop_bhp_update(s.bhp[w], d[0], o.dbhp_max_rel);
    for (size_t t = 0; t < perfs.size(); ++t) s.q[perfs[t]] += d[1 + t];
    rep.iterations++;
  }
  return rep;
}

// rs/fim/wells_inner.rs:585-598 solve_wells_locally, in well order
std::vector<WellInnerReport> solve_wells_locally(const Model& m, State& s, const WellInnerOpts& o) {
  std::vector<WellInnerReport> out;
  for (int w = 0; w < (int)s.bhp.size(); ++w) out.push_back(solve_well_locally(m, s, w, o));
  return out;
}

// rs/fim/wells_inner.rs:600-640 all_wells_converged: the same status at the state as it stands
bool all_wells_converged(const Model& m, const State& s, const WellInnerOpts& o) {
  for (int w = 0; w < (int)s.bhp.size(); ++w) {
    double peak;
    if (!well_status(m, s, w, well_local_system(m, s, w), o, peak)) return false;
  }
  return true;
}

// ------------------------------------------------------------------ state
// rs/fim/flash.rs:13-41 classify_cell_regime (rs_cap: the drsdt0 base, or infinity)
int classify(const Model& m, double p, double sg, double rs, double rs_cap) {
  if (!m.three_phase || !m.has_table) return SATURATED;
  if (sg > 1e-9) return SATURATED;
  double rs_sat = std::min(m.table.rs_sat_at(p), rs_cap);
  return rs < rs_sat - 1e-6 ? UNDERSATURATED : SATURATED;
}

// rs/fim/state.rs:109-177 FimState::from_simulator
State from_simulator(const Model& m) {
  State s;
  int n = m.n();
  s.cells.resize(n);
  for (int i = 0; i < n; ++i) {
    int r = classify(m, m.pressure[i], m.sg[i], m.rs[i], m.rs_cap_of(i));
    s.cells[i] = {m.pressure[i], m.sw[i], r == SATURATED ? m.sg[i] : m.rs[i], r};
  }
  int W = (int)m.topo_wells.size();
  s.bhp.resize(W);
  for (int w = 0; w < W; ++w) s.bhp[w] = m.control(w).bhp_target;
  s.q.assign(m.wells.size(), 0.0);
  for (int w = 0; w < W; ++w) {
    double b;
    if (solve_bhp_from_target(m, s, w, b)) s.bhp[w] = b;
  }
  for (size_t e = 0; e < m.wells.size(); ++e) {
    int w = m.perf_well[e];
    s.q[e] = m.control(w).enabled ? connection_rate_state(m, s, (int)e, s.bhp[w]) : 0.0;
  }
  return s;
}

// rs/fim/state.rs:899-914 write-back (flash per cell)
void write_back(Model& m, const State& s) {
  for (int i = 0; i < m.n(); ++i) {
    const auto& c = s.cells[i];
    auto pr = m.cell_props<double>(c.regime, c.p, c.sw, c.hc, m.rs_cap_of(i));
    double sg, so, rs;
    if (!m.three_phase) {
      sg = 0.0;
      so = std::max(1.0 - c.sw, 0.0);
      rs = 0.0;
    } else {
      sg = pr.sg;
      so = pr.so;
      rs = pr.rs;
    }
    m.pressure[i] = c.p;
    m.sw[i] = c.sw;
    m.sg[i] = sg;
    m.so[i] = so;
    m.rs[i] = rs;
  }
  for (size_t e = 0; e < m.wells.size(); ++e) m.wells[e].bhp = s.bhp[m.perf_well[e]];
}

// ------------------------------------------------------------------ assembly
struct Assembly {
  Vec R;
  Csr J;
};

// rs/fim/assembly_ad.rs:782-983 assemble_fim_system_ad
Assembly assemble(const Model& m, const State& prev, const State& s, double dt, bool residual_only) {
  int n = m.n(), W = (int)s.bhp.size(), P = (int)s.q.size();
  int nu = 3 * n + W + P;
  Assembly a;
  a.R.assign(nu, 0.0);
  // accumulation residual
  ORC_PARALLEL_FOR
  for (int i = 0; i < n; ++i) {
    const auto& c = s.cells[i];
    const auto& pc = prev.cells[i];
    double acc[3];
    m.accumulation<double>(i, c.regime, c.p, c.sw, c.hc, pc.regime, pc.p, pc.sw, pc.hc, acc);
    for (int e = 0; e < 3; ++e) a.R[3 * i + e] += acc[e];
  }
  // faces in natural order (cell, then its x, y, z faces); each face's terms are evaluated
  // independently (in parallel) and then summed into the rows in that order
  struct Face { int a, b; char dim; int ka, kb; };
  std::vector<Face> faces;
  faces.reserve(3 * (size_t)n);
  for (int k = 0; k < m.nz; ++k)
    for (int j = 0; j < m.ny; ++j)
      for (int i = 0; i < m.nx; ++i) {
        int id = m.idx(i, j, k);
        if (i + 1 < m.nx) faces.push_back({id, m.idx(i + 1, j, k), 'x', k, k});
        if (j + 1 < m.ny) faces.push_back({id, m.idx(i, j + 1, k), 'y', k, k});
        if (k + 1 < m.nz) faces.push_back({id, m.idx(i, j, k + 1), 'z', k, k + 1});
      }
  const int nf = (int)faces.size();
  std::vector<double> fgt(nf);
  ORC_PARALLEL_FOR
  for (int q = 0; q < nf; ++q) fgt[q] = DARCY_METRIC_FACTOR * m.geom_trans(faces[q].a, faces[q].b, faces[q].dim);
  {
    std::vector<std::array<double, 3>> ff(nf);
    ORC_PARALLEL_FOR
    for (int q = 0; q < nf; ++q) {
      const Face& fc = faces[q];
      if (fgt[q] <= 0.0) continue;
      const auto& ci = s.cells[fc.a];
      const auto& cj = s.cells[fc.b];
      m.face_flux<double>(fgt[q], fc.a, fc.b, ci.regime, cj.regime, m.depth_at_k(fc.ka), m.depth_at_k(fc.kb), ci.p,
                          ci.sw, ci.hc, cj.p, cj.sw, cj.hc, ff[q].data());
    }
    for (int q = 0; q < nf; ++q) {
      if (fgt[q] <= 0.0) continue;
      for (int e = 0; e < 3; ++e) {
        a.R[3 * faces[q].a + e] += ff[q][e] * dt;
        a.R[3 * faces[q].b + e] += -ff[q][e] * dt;
      }
    }
  }
  // wells: rs/fim/assembly_ad.rs:239-331
  std::vector<std::vector<double>> perf_rates(W);
  for (int e = 0; e < P; ++e) {
    int w = m.perf_well[e];
    bool inj = m.topo_wells[w].injector;
    int fluid = m.fluid_of(w);
    const Well& wl = m.wells[e];
    int id = m.idx(wl.i, wl.j, wl.k);
    auto cell = wcell(s, id);
    double q = s.q[e];
    double coef[3];
    rate_coefficients<double>(m, inj, fluid, cell, coef);
    for (int c = 0; c < 3; ++c) a.R[3 * id + c] += coef[c] * q * dt;
    double wi;
    if (m.well_index(e, wi)) {
      double conn = connection_rate<double>(m, wi, wl.head_offset, inj, cell, s.bhp[w]);
      a.R[3 * n + W + e] += q - conn;
    }
  }
  for (int w = 0; w < W; ++w) {
    Control c = m.control(w);
    bool inj = m.topo_wells[w].injector;
    int fluid = m.fluid_of(w);
    double bhp = s.bhp[w];
    if (!c.enabled || !c.rate_controlled) {
      a.R[3 * n + w] += bhp - c.bhp_target;
      continue;
    }
    if (!c.has_target) continue;
    double actual = 0.0;
    for (int e : m.topo_wells[w].perfs) {
      const Well& wl = m.wells[e];
      actual += perforation_rate<double>(m, inj, fluid, c.uses_surface, wcell(s, m.idx(wl.i, wl.j, wl.k)), s.q[e]);
    }
    double bs = inj ? c.bhp_limit - bhp : bhp - c.bhp_limit;
    double rsl = c.target - actual;
    double xa = bs / std::max(std::fabs(c.bhp_limit), 1.0), xb = rsl / std::max(std::fabs(c.target), 1.0);
    double nrm = std::sqrt((xa * xa + xb * xb) + 2.0 * FB_EPSILON * FB_EPSILON);
    a.R[3 * n + w] += nrm - xa - xb;
  }
  if (residual_only) return a;

  Triplets t;
  t.n = nu;
  auto scatter = [&](int rc, int cc, const double blk[3][3]) {  // :736-754
    for (int e = 0; e < 3; ++e)
      for (int v = 0; v < 3; ++v)
        if (std::fabs(blk[e][v]) > 1e-14) t.add(3 * rc + e, 3 * cc + v, blk[e][v]);
  };
  {
    std::vector<std::array<double, 9>> cb(n);
    ORC_PARALLEL_FOR
    for (int i = 0; i < n; ++i) {
      const auto& c = s.cells[i];
      const auto& pc = prev.cells[i];
      Ad<3> acc[3];
      m.accumulation<Ad<3>>(i, c.regime, Ad<3>::var(c.p, 0), Ad<3>::var(c.sw, 1), Ad<3>::var(c.hc, 2), pc.regime,
                            pc.p, pc.sw, pc.hc, acc);
      for (int e = 0; e < 3; ++e)
        for (int v = 0; v < 3; ++v) cb[i][3 * e + v] = acc[e].d[v];
    }
    for (int i = 0; i < n; ++i) {
      double blk[3][3];
      for (int e = 0; e < 3; ++e)
        for (int v = 0; v < 3; ++v) blk[e][v] = cb[i][3 * e + v];
      scatter(i, i, blk);
    }
  }
  {
    // per face: d flux / d (cell a unknowns) and d flux / d (cell b unknowns), times dt
    std::vector<std::array<double, 18>> fd(nf);
    ORC_PARALLEL_FOR
    for (int q = 0; q < nf; ++q) {
      const Face& fc = faces[q];
      if (fgt[q] <= 0.0) continue;
      const auto& ci = s.cells[fc.a];
      const auto& cj = s.cells[fc.b];
      Ad<6> f[3];
      m.face_flux<Ad<6>>(fgt[q], fc.a, fc.b, ci.regime, cj.regime, m.depth_at_k(fc.ka), m.depth_at_k(fc.kb),
                         Ad<6>::var(ci.p, 0), Ad<6>::var(ci.sw, 1), Ad<6>::var(ci.hc, 2), Ad<6>::var(cj.p, 3),
                         Ad<6>::var(cj.sw, 4), Ad<6>::var(cj.hc, 5), f);
      for (int e = 0; e < 3; ++e)
        for (int v = 0; v < 3; ++v) {
          fd[q][3 * e + v] = f[e].d[v] * dt;
          fd[q][9 + 3 * e + v] = f[e].d[3 + v] * dt;
        }
    }
    for (int q = 0; q < nf; ++q) {
      if (fgt[q] <= 0.0) continue;
      double bii[3][3], bij[3][3], bji[3][3], bjj[3][3];
      for (int e = 0; e < 3; ++e)
        for (int v = 0; v < 3; ++v) {
          bii[e][v] = fd[q][3 * e + v];
          bij[e][v] = fd[q][9 + 3 * e + v];
          bji[e][v] = -fd[q][3 * e + v];
          bjj[e][v] = -fd[q][9 + 3 * e + v];
        }
      scatter(faces[q].a, faces[q].a, bii);
      scatter(faces[q].a, faces[q].b, bij);
      scatter(faces[q].b, faces[q].a, bji);
      scatter(faces[q].b, faces[q].b, bjj);
    }
  }
  // wells: rs/fim/assembly_ad.rs:333-573, AD slots [p, sw, hc, bhp, q]
  std::vector<std::array<Ad<5>, 1>> rates(P);
  for (int e = 0; e < P; ++e) {
    int w = m.perf_well[e];
    bool inj = m.topo_wells[w].injector;
    int fluid = m.fluid_of(w);
    const Well& wl = m.wells[e];
    int id = m.idx(wl.i, wl.j, wl.k);
    const auto& c = s.cells[id];
    WCell<Ad<5>> ac{Ad<5>::var(c.p, 0), Ad<5>::var(c.sw, 1), Ad<5>::var(c.hc, 2), c.regime, id};
    Ad<5> bhp = Ad<5>::var(s.bhp[w], 3);
    Ad<5> q = Ad<5>::var(s.q[e], 4);
    Ad<5> coef[3];
    rate_coefficients<Ad<5>>(m, inj, fluid, ac, coef);
    int qcol = 3 * n + W + e;
    for (int r = 0; r < 3; ++r) {
      Ad<5> src = coef[r] * q * dt;
      for (int v = 0; v < 3; ++v) t.add_nz(3 * id + r, 3 * id + v, src.d[v]);
      t.add_nz(3 * id + r, qcol, src.d[4]);
    }
    double wi;
    if (m.well_index(e, wi)) {
      Ad<5> row = q - connection_rate<Ad<5>>(m, wi, wl.head_offset, inj, ac, bhp);
      int rr = 3 * n + W + e;
      for (int v = 0; v < 3; ++v) t.add_nz(rr, 3 * id + v, row.d[v]);
      t.add_nz(rr, 3 * n + w, row.d[3]);
      t.add_nz(rr, qcol, row.d[4]);
    }
    Control cc = m.control(w);
    rates[e][0] = perforation_rate<Ad<5>>(m, inj, fluid, cc.uses_surface, ac, q);
  }
  for (int w = 0; w < W; ++w) {
    Control c = m.control(w);
    bool inj = m.topo_wells[w].injector;
    int rr = 3 * n + w;
    if (!c.enabled || !c.rate_controlled) {
      t.add_nz(rr, 3 * n + w, 1.0);
      continue;
    }
    if (!c.has_target) continue;
    double actual = 0.0;
    for (int e : m.topo_wells[w].perfs) actual += rates[e][0].v;
    double bhp = s.bhp[w];
    double sb = std::max(std::fabs(c.bhp_limit), 1.0), sr = std::max(std::fabs(c.target), 1.0);
    double xa = (inj ? c.bhp_limit - bhp : bhp - c.bhp_limit) / sb;
    double xb = (c.target - actual) / sr;
    double nrm = std::sqrt((xa * xa + xb * xb) + 2.0 * FB_EPSILON * FB_EPSILON);
    double dfa = xa / nrm - 1.0, dfb = xb / nrm - 1.0;
    t.add_nz(rr, 3 * n + w, dfa * (inj ? -1.0 : 1.0) / sb);
    for (int e : m.topo_wells[w].perfs) {
      const Well& wl = m.wells[e];
      int id = m.idx(wl.i, wl.j, wl.k);
      const auto& d = rates[e][0].d;
      double f = dfb * (-1.0 / sr);
      for (int v = 0; v < 3; ++v) t.add_nz(rr, 3 * id + v, f * d[v]);
      t.add_nz(rr, 3 * n + W + e, f * d[4]);
    }
  }
  a.J = Csr::from(t);
  return a;
}

// ------------------------------------------------------------------ convergence
struct Scaling {
  Vec w, o, g, wc, pf;
};

// rs/fim/scaling.rs:139-187
Scaling equation_scaling(const Model& m, const State& s, double dt) {
  Scaling sc;
  int n = m.n();
  double dtd = std::max(dt, 1e-12);
  for (int i = 0; i < n; ++i) {
    double pvdt = std::max(std::fabs(m.pore_volume(i) / dtd), 1.0);
    double p = s.cells[i].p;
    double bo;
    if (m.has_table) {
      if (m.three_phase) {
        double b, mu;
        m.table.oil(p, m.rs[i], b, mu);
        bo = b;
      } else bo = m.table.saturated(p).bo;
    } else bo = m.base_bo(p);
    bo = std::max(bo, 1e-9);
    double bg = std::max(m.has_table ? m.table.saturated(p).bg : 1.0, 1e-9);
    double bw = m.water_fvf(p);
    sc.w.push_back(pvdt / bw);
    sc.o.push_back(pvdt / bo);
    sc.g.push_back(pvdt / bg);
  }
  for (size_t w = 0; w < s.bhp.size(); ++w) {
    Control c = m.control((int)w);
    bool slack = c.enabled && c.rate_controlled && c.has_target;
    sc.wc.push_back(slack ? 1.0 : std::max(std::fabs(s.bhp[w]), 1.0));
  }
  for (double q : s.q) sc.pf.push_back(std::max(std::fabs(q), 1.0));
  return sc;
}

struct Families {
  double f[5];
  long row[5];  // the row of each family's peak (-1: the family has no rows)
};

// rs/fim/newton/convergence.rs:325-345,734-790 residual family peaks (scaled) and their
// rows: each family's rows in order, a later row replacing the peak only when strictly
// larger, so the first row holding the maximum (the first row when none is positive). A
// NaN value is passed over here as std::max passes it over; the reference counts it as
// +inf, a case the Newton loop reports as a non-finite state before it gets here.
Families family_norms(const Vec& R, const Scaling& sc, int n) {
  Families fm;
  const long W = (long)sc.wc.size(), P = (long)sc.pf.size();
  for (int c = 0; c < 3; ++c) fm.f[c] = 0.0, fm.row[c] = c;
  fm.f[3] = fm.f[4] = 0.0;
  fm.row[3] = W ? 3L * n : -1;
  fm.row[4] = P ? 3L * n + W : -1;
  auto take = [&](int fam, double v, long row) {
    if (fm.f[fam] < v) fm.f[fam] = v, fm.row[fam] = row;
  };
  for (int i = 0; i < n; ++i) {
    take(0, std::fabs(R[3 * i]) / sc.w[i], 3L * i);
    take(1, std::fabs(R[3 * i + 1]) / sc.o[i], 3L * i + 1);
    take(2, std::fabs(R[3 * i + 2]) / sc.g[i], 3L * i + 2);
  }
  for (long k = 0; k < W; ++k) take(3, std::fabs(R[3 * n + k]) / sc.wc[k], 3L * n + k);
  for (long k = 0; k < P; ++k) take(4, std::fabs(R[3 * n + W + k]) / sc.pf[k], 3L * n + W + k);
  if (!W) fm.f[3] = INFINITY;
  if (!P) fm.f[4] = INFINITY;
  return fm;
}

struct CnvMb {
  double cnv[3], mb[3];
  double violating_fraction;
  bool pv_relax, accept;
};

// rs/fim/newton/convergence.rs:478-612 (residual is dt-integrated)
CnvMb cnv_mb(const Model& m, const State& s, const Vec& R, bool relax_final) {
  int n = m.n();
  double bavg[3] = {0, 0, 0};
  std::vector<std::array<double, 3>> fvf(n);
  for (int i = 0; i < n; ++i) {
    double p = s.cells[i].p;
    double bo;
    if (m.has_table) {
      if (m.three_phase) {
        double b, mu;
        m.table.oil(p, m.rs[i], b, mu);
        bo = b;
      } else bo = m.table.saturated(p).bo;
    } else bo = m.base_bo(p);
    fvf[i] = {m.water_fvf(p), std::max(bo, 1e-9), std::max(m.has_table ? m.table.saturated(p).bg : 1.0, 1e-9)};
    for (int c = 0; c < 3; ++c) bavg[c] += fvf[i][c];
  }
  for (double& b : bavg) b /= std::max(n, 1);
  double maxc[3] = {0, 0, 0}, rsum[3] = {0, 0, 0}, pvsum = 0.0, viol = 0.0;
  for (int i = 0; i < n; ++i) {
    double pv = std::max(m.pore_volume(i), 1e-9);
    pvsum += pv;
    double cm = 0.0;
    for (int c = 0; c < 3; ++c) {
      double r = R[3 * i + c];
      rsum[c] += r;
      maxc[c] = std::max(maxc[c], std::fabs(r) / pv);
      cm = std::max(cm, std::fabs(r) * bavg[c] / pv);
    }
    if (cm > 1e-2) viol += pv;
  }
  CnvMb o;
  for (int c = 0; c < 3; ++c) {
    o.cnv[c] = bavg[c] * maxc[c];
    o.mb[c] = std::fabs(bavg[c] * rsum[c]) / std::max(pvsum, 1e-9);
  }
  bool pv_relax = viol < 0.03 * pvsum;
  o.violating_fraction = viol / std::max(pvsum, 1e-9);
  o.pv_relax = pv_relax;
  double mbt = relax_final ? 1e-6 : 1e-7, cnvt = relax_final ? 1.0 : 1e-2;
  bool ok = true;
  for (int c = 0; c < 3; ++c) {
    bool cnv_ok = o.cnv[c] <= cnvt || (pv_relax && o.cnv[c] <= 1.0);
    ok = ok && cnv_ok && o.mb[c] <= mbt;
  }
  o.accept = ok;
  return o;
}

// rs/fim/newton/convergence.rs:648-711 opm_per_cell_chopped_update
Vec chop(const State& s, const Vec& upd, double relax, int n) {
  Vec c(upd.size());
  for (size_t i = 0; i < upd.size(); ++i) c[i] = upd[i] * relax;
  int W = (int)s.bhp.size();
  for (int w = 0; w < W; ++w) {
    int off = 3 * n + w;
    double b = s.bhp[w];
    double d = c[off];
    double cap = 1.0 * std::fabs(b);
    double lim = std::fabs(d) > cap ? (d > 0 ? cap : -cap) : d;
    c[off] = (b + lim < 1.0) ? 1.0 - b : lim;
  }
  for (int i = 0; i < n; ++i) {
    const auto& cell = s.cells[i];
    double dp = c[3 * i], dsw = c[3 * i + 1], dhc = c[3 * i + 2];
    double dsg, dso;
    if (cell.regime == SATURATED) { dsg = dhc; dso = -(dsw + dhc); }
    else { dsg = 0.0; dso = -dsw; }
    double mx = std::max(std::fabs(dsw), std::max(std::fabs(dso), std::fabs(dsg)));
    double a = mx > 0.2 ? 0.2 / mx : 1.0;
    c[3 * i + 1] = dsw * a;
    if (cell.regime == SATURATED) c[3 * i + 2] = dhc * a;
    else if (cell.hc + c[3 * i + 2] < 0.0) c[3 * i + 2] = -cell.hc;
    double cap = 0.3 * std::fabs(cell.p);
    if (std::fabs(dp) > cap) c[3 * i] = dp > 0 ? cap : -cap;
  }
  return c;
}

// rs/fim/newton/damping.rs:716-768
bool family_osc(double f0, double f1, double f2) {
  if (!(std::isfinite(f0) && std::isfinite(f1) && std::isfinite(f2)) || f0 <= 0.0) return false;
  double d1 = std::fabs(f0 - f2) / f0, d2 = std::fabs(f0 - f1) / f0;
  return d1 < 0.2 && 0.2 < d2;
}

// rs/fim/newton/damping.rs:724-757 detect_oscillation over the five families
// (water, oil, gas, well constraint, perforation flow)
int detect_oscillation(const double cur[5], const double p1[5], const double p2[5]) {
  int c = 0;
  for (int f = 0; f < 5; ++f) c += family_osc(cur[f], p1[f], p2[f]) ? 1 : 0;
  return c;
}

// rs/fim/newton/damping.rs:760-766 next_relaxation_factor (step 0.1, floor 0.5, one
// oscillating family suffices)
double next_relaxation_factor(double relax, int oscillating) {
  return oscillating >= 1 ? std::max(relax - 0.1, 0.5) : relax;
}

// rs/fim/newton.rs:970-3368 run_fim_timestep (OpmAligned branches)
// rs/fim/newton.rs:439-481 classify_retry_failure_with_site: 0 linear-bad, 1 nonlinear-bad,
// 2 mixed. No linear solve in the attempt: mixed; an unconverged solve: linear-bad; a
// converged direct solve: nonlinear-bad; a converged FGMRES-CPR solve: nonlinear-bad when
// its pressure stage was strong (:50-51 average reduction <= 0.25, last <= 0.5).
int retry_class(bool have, const LinReport& l) {
  if (!have) return 2;
  if (!l.converged) return 0;
  if (l.direct) return 1;
  return l.cpr_apps > 0 && l.cpr_avg <= 0.25 && l.cpr_last <= 0.5 ? 1 : 2;
}

// rs/fim/newton/convergence.rs:793-806 global peak: water, oil, gas, then the well and
// perforation families (present with wells); a later family wins only when strictly larger
int dominant_family(const double f[5], bool wells) {
  int g = 0;
  for (int k = 1; k < (wells ? 5 : 3); ++k)
    if (f[k] > f[g]) g = k;
  return g;
}

StepReport run_timestep(Sim& S, const State& prev, State& s, double dt) {
  const Model& m = S.m;
  int n = m.n();
  StepReport rep;
  LinReport last_lr;
  bool have_lr = false;
  auto fail = [&](const Families& fm) {
    rep.fail_class = retry_class(have_lr, last_lr);
    rep.fail_family = dominant_family(fm.f, !m.topo_wells.empty());
    rep.fail_row = fm.row[rep.fail_family];
  };
  Families h1, h2;
  for (double& x : h1.f) x = INFINITY;
  for (double& x : h2.f) x = INFINITY;
  double relax = 1.0;
  const auto& o = S.nopts;
  for (int it = 0; it < o.max_iterations; ++it) {
    auto t0 = std::chrono::steady_clock::now();
    Assembly a = assemble(m, prev, s, dt, false);
    auto t1 = std::chrono::steady_clock::now();
    S.assembly_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    Scaling sc = equation_scaling(m, s, dt);
    Families fm = family_norms(a.R, sc, n);
    relax = next_relaxation_factor(relax, detect_oscillation(fm.f, h1.f, h2.f));
    h2 = h1;
    h1 = fm;
    bool final_it = it + 1 == o.max_iterations;
    bool accept;
    double cell_norm = std::max(fm.f[0], std::max(fm.f[1], fm.f[2]));
    rep.final_scaled_residual = cell_norm;
    if (o.residual_tol > 0.0) {
      CnvMb cm = cnv_mb(m, s, a.R, false);
      bool mb_ok = cm.mb[0] <= o.mb_tol && cm.mb[1] <= o.mb_tol && cm.mb[2] <= o.mb_tol;
      accept = it >= o.min_iteration_index && cell_norm <= o.residual_tol && mb_ok;
    } else {
      accept = it >= o.min_iteration_index && cnv_mb(m, s, a.R, final_it).accept;
      // rs/fim/newton.rs:1262-1285: with the nested well solve on, every well must also pass
      // its own convergence test
      if (accept && o.nested_well_solve) accept = all_wells_converged(m, s, WellInnerOpts{});
    }
    if (accept) {
      rep.converged = true;
      rep.newton_iterations = it;
      return rep;
    }
    Vec rhs(a.R.size());
    for (size_t i = 0; i < rhs.size(); ++i) rhs[i] = -a.R[i];
    auto t2 = std::chrono::steady_clock::now();
    LinReport lr = solve_linearized(a.J, rhs, n, o.lin, &S.lstats);
    auto t3 = std::chrono::steady_clock::now();
    if (getenv("ORC_TRACE")) {
      fprintf(stderr, "newton it=%d fam=[%.3e %.3e %.3e %.3e %.3e] lin its=%d conv=%d red=%.3e relax=%.2f\n", it,
              fm.f[0], fm.f[1], fm.f[2], fm.f[3], fm.f[4], lr.iterations, (int)lr.converged, lr.reduction(), relax);
      int worst = 0;
      double wv = 0;
      for (int i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
          double v = std::fabs(a.R[3 * i + c]) / (c == 0 ? sc.w[i] : c == 1 ? sc.o[i] : sc.g[i]);
          if (v > wv) { wv = v; worst = 3 * i + c; }
        }
      int wc = worst / 3;
      fprintf(stderr, "  worst row %d (cell %d comp %d) scaled %.3e state p=%.3f sw=%.4f hc=%.5f reg=%d\n", worst, wc,
              worst % 3, wv, s.cells[wc].p, s.cells[wc].sw, s.cells[wc].hc, s.cells[wc].regime);
    }
    S.linear_ms += std::chrono::duration<double, std::milli>(t3 - t2).count();
    rep.linear_iterations += lr.iterations;
    rep.lin_solves++;
    rep.last_lin_its = lr.iterations;
    rep.last_lin_conv = lr.converged ? 1 : 0;
    last_lr = lr;
    have_lr = true;
    bool finite = true;
    for (double v : lr.x) finite = finite && std::isfinite(v);
    if (!lr.converged) {
      // rs/fim/newton/convergence.rs:385-389 relaxed linear acceptance
      if (!(finite && std::isfinite(lr.reduction()) && lr.reduction() < 0.01)) {
        rep.converged = false;
        rep.newton_iterations = it + 1;
        fail(fm);
        return rep;
      }
    }
    Vec up = chop(s, lr.x, relax, n);
    // rs/fim/state.rs:597,672-715 raw update, pressure floor, then well post-processing
    for (int i = 0; i < n; ++i) {
      s.cells[i].p += up[3 * i];
      s.cells[i].sw += up[3 * i + 1];
      s.cells[i].hc += up[3 * i + 2];
      s.cells[i].p = std::max(s.cells[i].p, 1e-6);
    }
    int W = (int)s.bhp.size();
    for (int w = 0; w < W; ++w) s.bhp[w] += up[3 * n + w];
    for (size_t e = 0; e < s.q.size(); ++e) s.q[e] += up[3 * n + W + e];
    // rs/fim/state.rs:717-740 finish_well_update: Relax, or NestedSolve with the flag on
    enforce_control_bounds(m, s);
    if (o.nested_well_solve) solve_wells_locally(m, s, WellInnerOpts{});
    else relax_wells(m, s);
    enforce_control_bounds(m, s);
    bool ok = true;
    for (auto& c : s.cells) ok = ok && std::isfinite(c.p) && std::isfinite(c.sw) && std::isfinite(c.hc);
    for (double b : s.bhp) ok = ok && std::isfinite(b);
    for (double q : s.q) ok = ok && std::isfinite(q);
    if (!ok) {
      rep.converged = false;
      rep.newton_iterations = it + 1;
      fail(fm);
      return rep;
    }
  }
  // rs/fim/newton.rs:3239-3300 post-loop residual-only check with relaxed tiers
  Assembly a = assemble(m, prev, s, dt, true);
  rep.newton_iterations = o.max_iterations;
  const Families fm_end = family_norms(a.R, equation_scaling(m, s, dt), n);
  if (o.residual_tol > 0.0) {
    const Families& fm = fm_end;
    double cell_norm = std::max(fm.f[0], std::max(fm.f[1], fm.f[2]));
    rep.final_scaled_residual = cell_norm;
    CnvMb cm = cnv_mb(m, s, a.R, false);
    rep.converged = cell_norm <= o.residual_tol && cm.mb[0] <= o.mb_tol && cm.mb[1] <= o.mb_tol && cm.mb[2] <= o.mb_tol;
  } else {
    rep.converged = cnv_mb(m, s, a.R, true).accept;
  }
  if (!rep.converged) fail(fm_end);
  return rep;
}

// Run one substep of length dt from the stored simulator state; commit on success.
// Surface oil produced by the producers in one accepted substep at its converged
// iterate: dt sum_e (fo/Bo)(x) q_e, the oil source term of the producer rows
// (rs/fim/assembly_ad.rs:239-331). The objective of the discrete-adjoint extension.
double producer_oil(const Model& m, const State& s, double dt) {
  double tot = 0.0;
  for (size_t e = 0; e < s.q.size(); ++e) {
    int w = m.perf_well[e];
    if (m.topo_wells[w].injector) continue;
    if (!m.control(w).enabled) continue;
    const Well& wl = m.wells[e];
    double coef[3];
    rate_coefficients<double>(m, false, m.fluid_of(w), wcell(s, m.idx(wl.i, wl.j, wl.k)), coef);
    tot += coef[1] * s.q[e] * dt;
  }
  return tot;
}

// Report sums of the stored state: [0..2] water, oil, gas inventory at surface conditions
// on the geometric pore volume (rs/fim/timestep.rs:1849-1888, with get_b_o_cell of
// rs/pvt.rs:769-779 and get_b_g of :835-841), [3] pv p and [4] pv over cells with a
// positive finite pore volume, [5] Sw, [6] Sg (rs/reporting.rs:304-327, 646-666).
void report_sums(const Model& m, double out[7]) {
  for (int k = 0; k < 7; ++k) out[k] = 0.0;
  for (int i = 0; i < m.n(); ++i) {
    const double pv = m.pore_volume(i), p = m.pressure[i], pvf = std::max(pv, 1e-9);
    double bo, mu;
    if (m.has_table) {
      if (m.three_phase) m.table.oil(p, m.rs[i], bo, mu);
      else bo = m.table.saturated(p).bo;
    } else {
      bo = m.base_bo(p);
    }
    bo = std::max(bo, 1e-9);
    out[0] += m.sw[i] * pv * m.inv_bw(p);
    out[1] += m.so[i] * pvf / bo;
    if (m.three_phase) {
      const double bg = std::max(m.has_table ? m.table.saturated(p).bg : 1.0, 1e-9);
      out[2] += m.sg[i] * pvf / bg + (m.has_table ? m.so[i] * pvf * m.rs[i] / bo : 0.0);
    }
    if (pv > 0.0 && std::isfinite(pv)) {
      out[3] += p * pv;
      out[4] += pv;
    }
    out[5] += m.sw[i];
    out[6] += m.sg[i];
  }
}

// rs/reporting.rs:553-734 record_fim_step_report at the accepted state s (already written
// back). Perforation rates: rs/fim/wells.rs:834-856 producer phase split at the perforated
// cell, :1634-1666 component rates at surface conditions; injectors inject the well's
// fluid (rs/reporting.rs:125-131).
// rs/reporting.rs:52-123 compute_sweep_metrics on the written-back fields, in the
// reference's loop order (columns j, i; layers k inside)
void sweep_metrics(const Sim& S, TimePoint& tp) {
  const double nan = std::numeric_limits<double>::quiet_NaN();
  tp.sweep_e_a = tp.sweep_e_v = tp.sweep_e_vol = tp.sweep_mobile_oil_recovered = nan;
  if (!S.sweep.on) return;
  const Model& m = S.m;
  const int nx = m.nx, ny = m.ny, nz = m.nz;
  double swept_cells = 0.0, swept_columns = 0.0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double column = 0.0;
      for (int k = 0; k < nz; ++k) {
        const double w = m.sw[(size_t)k * nx * ny + j * nx + i] > S.sweep.threshold ? 1.0 : 0.0;
        swept_cells += w;
        column = std::max(column, w);
      }
      swept_columns += column;
    }
  const double total = double(nx) * ny * nz;
  tp.sweep_e_vol = total > 0.0 ? swept_cells / total : 0.0;
  const double e_a = nx * ny > 0 ? swept_columns / (nx * ny) : 0.0;
  if (S.sweep.geometry == 0) {
    tp.sweep_e_a = e_a;
    tp.sweep_e_v = 1.0;
  } else if (S.sweep.geometry == 1) {
    tp.sweep_e_a = 1.0;
    tp.sweep_e_v = tp.sweep_e_vol;
  } else {
    const double initial = total * std::max(S.sweep.initial_so - S.sweep.sor, 0.0);
    double remaining = 0.0;
    for (size_t c = 0; c < (size_t)total; ++c) remaining += std::max(m.so[c] - S.sweep.sor, 0.0);
    tp.sweep_mobile_oil_recovered = initial > 1e-12 ? std::min(std::max(1.0 - remaining / initial, 0.0), 1.0) : 0.0;
  }
}

void record_report(Sim& S, const State& s, double dt, double d_water, double oil_removed, double d_gas,
                   const double sums[7]) {
  const Model& m = S.m;
  int prod_rc = 0, inj_rc = 0, prod_lim = 0, inj_lim = 0;
  for (size_t w = 0; w < m.topo_wells.size(); ++w) {
    const Control c = m.control((int)w);
    if (!(c.enabled && c.rate_controlled)) continue;
    const bool inj = m.topo_wells[w].injector;
    const bool lim = inj ? s.bhp[w] >= c.bhp_limit - 1e-6 : s.bhp[w] <= c.bhp_limit + 1e-6;
    (inj ? inj_rc : prod_rc)++;
    if (lim) (inj ? inj_lim : prod_lim)++;
  }
  double prod_oil = 0, prod_liq = 0, prod_liq_res = 0, inj_tot = 0, inj_res = 0, inj_w_res = 0, prod_w_res = 0;
  double inj_w_sc = 0, prod_w_sc = 0, prod_gas = 0, inj_g_sc = 0;
  for (size_t e = 0; e < m.wells.size(); ++e) {
    const int w = m.perf_well[e];
    const Well& wl = m.wells[e];
    const int id = m.idx(wl.i, wl.j, wl.k);
    const double q = s.q[e], p = m.pressure[id], sw = m.sw[id], sg = m.sg[id], rs = m.rs[id];
    const bool table3 = m.three_phase && m.has_table;
    double bo, mu;
    if (table3) m.table.oil(p, rs, bo, mu);
    else bo = m.base_bo(p);
    const double bg = table3 ? m.table.saturated(p).bg : 1.0;
    const double ibw = m.inv_bw(p);
    if (m.topo_wells[w].injector) {
      inj_res += std::max(-q, 0.0);
      if (m.fluid_of(w) == 0) {
        const double cw = q * ibw;
        inj_tot += std::max(-cw, 0.0);
        inj_w_res += std::max(-q, 0.0);
        inj_w_sc += std::max(-cw, 0.0);
      } else {
        const double cg = q / std::max(bg, 1e-9);
        inj_tot += std::max(-cg, 0.0);
        inj_g_sc += std::max(-cg, 0.0);
      }
      continue;
    }
    const auto mob = m.mobilities<double>(sw, sg, p, rs);
    const double lw = std::max(mob.w, 0.0), lo = std::max(mob.o, 0.0), lg = std::max(mob.g, 0.0);
    const double lt = std::max(lw + lo + lg, 2.220446049250313e-16);
    auto c01 = [](double x) { return std::min(std::max(x, 0.0), 1.0); };
    const double fw = c01(lw / lt), fo = c01(lo / lt), fg = c01(lg / lt);
    const double cw = q * fw * ibw, co = q * fo / std::max(bo, 1e-9);
    const double cg = q * fg / std::max(bg, 1e-9) + co * std::max(rs, 0.0);
    prod_liq_res += std::max(q, 0.0);
    prod_w_res += std::max(q, 0.0) * fw;
    prod_w_sc += std::max(cw, 0.0);
    prod_oil += std::max(co, 0.0);
    prod_liq += std::max(cw, 0.0) + std::max(co, 0.0);
    prod_gas += std::max(cg, 0.0);
  }
  S.cum_injection_m3 += inj_w_res * dt;
  S.cum_production_m3 += prod_w_res * dt;
  S.cum_mb_water += (inj_w_sc - prod_w_sc) * dt - d_water;
  S.cum_mb_oil += prod_oil * dt - oil_removed;
  if (m.three_phase) S.cum_mb_gas += (inj_g_sc - prod_gas) * dt - d_gas;
  TimePoint tp{};
  tp.time = m.time + dt;
  tp.total_production_oil = prod_oil;
  tp.total_production_liquid = prod_liq;
  tp.total_production_liquid_reservoir = prod_liq_res;
  tp.total_injection = inj_tot;
  tp.total_injection_reservoir = inj_res;
  tp.material_balance_error_m3 = std::fabs(S.cum_mb_water);
  tp.material_balance_error_oil_m3 = std::fabs(S.cum_mb_oil);
  tp.avg_reservoir_pressure = sums[4] > 0.0 ? sums[3] / sums[4] : 0.0;
  tp.avg_water_saturation = m.n() > 0 ? sums[5] / m.n() : 0.0;
  tp.total_production_gas = prod_gas;
  tp.avg_gas_saturation = m.n() > 0 ? sums[6] / m.n() : 0.0;
  tp.material_balance_error_gas_m3 = std::fabs(S.cum_mb_gas);
  tp.producing_gor = prod_oil > 1e-9 ? prod_gas / prod_oil : 0.0;  // rs/reporting.rs:21
  tp.producer_bhp_limited_fraction = prod_rc > 0 ? double(prod_lim) / prod_rc : 0.0;
  tp.injector_bhp_limited_fraction = inj_rc > 0 ? double(inj_lim) / inj_rc : 0.0;
  sweep_metrics(S, tp);
  S.rate_history.push_back(tp);
}

// accept (rs/fim/timestep.rs:1329-1353): inventories around the write-back, then the report
void commit(Sim& S, const State& s, double dt) {
  double before[7], after[7];
  report_sums(S.m, before);
  write_back(S.m, s);
  report_sums(S.m, after);
  record_report(S, s, dt, after[0] - before[0], before[1] - after[1], after[2] - before[2], after);
}

bool substep(Sim& S, double dt, StepReport& rep) {
  Model& m = S.m;
  State prev = from_simulator(m);
  State s = prev;
  rep = run_timestep(S, prev, s, dt);
  S.total_newton += rep.newton_iterations;
  S.total_linear += rep.linear_iterations;
  if (!rep.converged) return false;
  S.producer_oil += producer_oil(m, s, dt);
  S.last_iterate = s;
  commit(S, s, dt);
  m.time += dt;
  S.accepted_dts.push_back(dt);
  S.total_substeps++;
  return true;
}

double opm_relative_change(const State& a, const State& b) {  // rs/fim/timestep.rs:436-466
  double delta = 0.0, denom = 0.0;
  for (size_t i = 0; i < a.cells.size(); ++i) {
    const auto& o = a.cells[i];
    const auto& c = b.cells[i];
    double dp = c.p - o.p;
    delta += dp * dp;
    denom += c.p * c.p;
    auto sgof = [](const CellState& x) { return x.regime == SATURATED ? std::max(x.hc, 0.0) : 0.0; };
    double ns = sgof(c), os = sgof(o);
    double nso = 1.0 - c.sw - ns, oso = 1.0 - o.sw - os;
    double pairs[3][2] = {{c.sw, o.sw}, {ns, os}, {nso, oso}};
    for (auto& pr : pairs) {
      double ds = pr[0] - pr[1];
      delta += ds * ds;
      denom += pr[0] * pr[0];
    }
  }
  return denom > 0.0 ? delta / denom : 0.0;
}

// rs/fim/timestep.rs:469-482 opm_pid_dt: errors are the last three relative changes, oldest first
double opm_pid_dt(double dt, const double err[3]) {
  const double tol = 0.1;  // OPM_TIME_STEP_CONTROL_TOLERANCE
  if (err[2] > tol) return dt * tol / err[2];
  if (err[0] == 0.0 || err[1] == 0.0 || err[2] == 0.0) return std::numeric_limits<double>::max();
  return dt * std::pow(err[1] / err[2], 0.075) * std::pow(tol / err[2], 0.175) *
         std::pow(err[1] * err[1] / (err[0] * err[2]), 0.01);
}

// rs/fim/timestep.rs:484-494 opm_iteration_count_dt (target 8, decay 1.0, growth 3.2)
double opm_iteration_count_dt(double dt, int newton_iterations) {
  const double its = newton_iterations, tgt = 8.0;
  return its > tgt ? dt / (1.0 + (its - tgt) / tgt * 1.0) : dt * (1.0 + (tgt - its) / tgt * 3.2);
}

// rs/fim/timestep.rs:502-528 opm_accepted_step_growth_decision: the tightest of the 3x
// ceiling, the PID and the iteration-count proposals, then 2x after a retry. Returns the
// factor; limiter 0 max growth, 1 PID, 2 iteration count, 3 restart growth.
double opm_growth_factor(double dt, int newton_iterations, const double err[3], int retries, int* limiter) {
  double f = 3.0;
  int lim = 0;
  const double cand[2] = {opm_pid_dt(dt, err) / dt, opm_iteration_count_dt(dt, newton_iterations) / dt};
  for (int k = 0; k < 2; ++k)
    if (cand[k] < f) {
      f = cand[k];
      lim = 1 + k;
    }
  if (retries > 0 && 2.0 < f) {
    f = 2.0;
    lim = 3;
  }
  if (limiter) *limiter = lim;
  return f;
}

// rs/fim/timestep.rs:988-1890 substep controller (OpmAligned growth, retry x0.33), with
// its FimStepStats record (timestep.rs:742-800)
int step(Sim& S, double target) {
  Model& m = S.m;
  m.refresh_head_offsets();
  double done = 0.0;
  double last_dt = target, growth = 3.0;
  double err[3] = {0.1, 0.1, 0.1};
  int substeps = 0;
  StepStatsRec sr{};
  orc_step_stats& st = sr.s;
  st.target_dt_days = target;
  st.min_accepted_dt_days = st.max_accepted_dt_days = st.last_accepted_dt_days = NAN;
  st.last_growth_factor = NAN;
  st.growth_limiter = st.last_retry_class = st.last_retry_dominant_family = -1;
  st.last_retry_dominant_row = -1;
  auto finish = [&](int rc) {
    st.time_days = m.time;
    st.advanced_dt_days = done;
    st.accepted_substeps = substeps;
    st.n_rungs = (int)sr.rungs.size();
    S.step_stats.push_back(sr);
    return rc;
  };
  const int max_substeps = 100000;  // rs/fim/timestep.rs:991 MAX_SUBSTEPS
  while (done < target - 1e-12 && substeps < max_substeps) {
    double remaining = target - done;
    double trial;
    if (substeps == 0) {
      if (m.time <= 1e-12) trial = std::min(remaining, remaining <= 1.0 + 1e-12 ? 0.25 : 1.0);
      else trial = remaining;
    } else {
      trial = std::min(remaining, last_dt * growth);
    }
    int retries = 0;
    for (;;) {
      State prev = from_simulator(m);
      State s = prev;
      StepReport rep = run_timestep(S, prev, s, trial);
      S.total_newton += rep.newton_iterations;
      S.total_linear += rep.linear_iterations;
      orc_step_rung rung{};
      rung.substep = substeps;
      rung.retried = rep.converged ? 0 : 1;
      rung.newton_iterations = rep.newton_iterations;
      rung.linear_iterations = rep.last_lin_its;
      rung.last_linear_converged = rep.last_lin_conv;
      rung.linear_solve_count = rep.lin_solves;
      rung.linear_iteration_count = rep.linear_iterations;
      rung.dt_days = trial;
      rung.retry_class = rep.converged ? -1 : rep.fail_class;
      rung.dominant_family = rep.converged ? -1 : rep.fail_family;
      rung.dominant_row = rep.converged ? -1 : rep.fail_row;
      sr.rungs.push_back(rung);
      if (rep.converged) {
        double rc = opm_relative_change(prev, s);
        err[0] = err[1];
        err[1] = err[2];
        err[2] = rc;
        int lim = 0;
        growth = opm_growth_factor(trial, rep.newton_iterations, err, retries, &lim);
        st.last_growth_factor = growth;
        st.growth_limiter = lim;
        st.min_accepted_dt_days = std::isnan(st.min_accepted_dt_days) ? trial : std::min(st.min_accepted_dt_days, trial);
        st.max_accepted_dt_days = std::isnan(st.max_accepted_dt_days) ? trial : std::max(st.max_accepted_dt_days, trial);
        st.last_accepted_dt_days = trial;
        last_dt = trial;
        S.producer_oil += producer_oil(m, s, trial);
        commit(S, s, trial);
        m.time += trial;
        done += trial;
        S.accepted_dts.push_back(trial);
        S.total_substeps++;
        substeps++;
        break;
      }
      st.last_retry_class = rep.fail_class;
      st.last_retry_dominant_family = rep.fail_family;
      st.last_retry_dominant_row = rep.fail_row;
      if (rep.fail_class == 0) st.linear_bad_retries++;
      else if (rep.fail_class == 1) st.nonlinear_bad_retries++;
      else st.mixed_retries++;
      retries++;
      S.total_retries++;
      // rs/fim/timestep.rs:1589-1767: next trial x0.33, then the collapse exit and the
      // retry budget of 16 per substep (:1000); t is the reference's time_days + time_stepped
      const double next = trial * 0.33;
      char msg[160];
      if (!std::isfinite(next) || next <= 1e-12) {
        snprintf(msg, sizeof msg, "FIM Newton step collapsed timestep at t=%.6f days after %d iterations",
                 m.time + done, rep.newton_iterations);
        S.warning = msg;
        return finish(2);
      }
      if (retries >= 16) {
        snprintf(msg, sizeof msg, "FIM Newton step exceeded retry budget at t=%.6f days after %d retries", m.time + done,
                 retries);
        S.warning = msg;
        return finish(2);
      }
      trial = next;
    }
  }
  if (done < target - 1e-12) {  // rs/fim/timestep.rs:1772-1778
    char msg[160];
    snprintf(msg, sizeof msg,
             "FIM adaptive timestep hit MAX_SUBSTEPS before completing requested dt (advanced %.6f of %.6f days)", done,
             target);
    S.warning = msg;
    return finish(2);
  }
  return finish(0);
}

Sim* as(orc_sim* p) { return reinterpret_cast<Sim*>(p); }

void resize_fields(Model& m) {
  int n = m.n();
  m.phi.assign(n, 0.2);
  m.kx.assign(n, 100.0);
  m.ky.assign(n, 100.0);
  m.kz.assign(n, 10.0);
  m.pc_scale.assign(n, 1.0);
  m.pressure.assign(n, 300.0);
  m.sw.assign(n, 0.3);
  m.so.assign(n, 0.7);
  m.sg.assign(n, 0.0);
  m.rs.assign(n, 0.0);
}

}  // namespace

// ======================================================================= C API
extern "C" {

orc_sim* orc_create(int nx, int ny, int nz, double phi) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return nullptr;
  Sim* s = new Sim();
  s->m.nx = nx;
  s->m.ny = ny;
  s->m.nz = nz;
  s->m.dz.assign(nz, 1.0);
  resize_fields(s->m);
  std::fill(s->m.phi.begin(), s->m.phi.end(), phi);
  return reinterpret_cast<orc_sim*>(s);
}

void orc_destroy(orc_sim* p) { delete as(p); }

int orc_set_cell_dimensions_per_layer(orc_sim* p, double dx, double dy, const double* dz) {
  Model& m = as(p)->m;
  if (!(dx > 0) || !(dy > 0)) return 1;
  m.dx = dx;
  m.dy = dy;
  for (int k = 0; k < m.nz; ++k) {
    if (!(dz[k] > 0)) return 1;
    m.dz[k] = dz[k];
  }
  return 0;
}

int orc_set_fields(orc_sim* p, const double* phi, const double* kx, const double* ky, const double* kz) {
  Model& m = as(p)->m;
  int n = m.n();
  for (int i = 0; i < n; ++i)
    if (!(kx[i] > 0) || !(ky[i] > 0) || !(kz[i] > 0) || !(phi[i] > 0)) return 1;
  m.phi.assign(phi, phi + n);
  m.kx.assign(kx, kx + n);
  m.ky.assign(ky, ky + n);
  m.kz.assign(kz, kz + n);
  return 0;
}

int orc_set_pc_scale(orc_sim* p, const double* scale) {
  Model& m = as(p)->m;
  m.pc_scale.assign(scale, scale + m.n());
  return 0;
}

int orc_set_initial_state(orc_sim* p, const double* pr, const double* sw, const double* sg, const double* rs) {
  Model& m = as(p)->m;
  int n = m.n();
  for (int i = 0; i < n; ++i) {
    m.pressure[i] = pr[i];
    m.sw[i] = sw[i];
    m.sg[i] = sg[i];
    m.rs[i] = rs[i];
    m.so[i] = std::max(1.0 - sw[i] - sg[i], 0.0);
  }
  return 0;
}

int orc_set_fluid(orc_sim* p, double mu_o, double mu_w, double c_o, double c_w, double rho_o, double rho_w) {
  Model& m = as(p)->m;
  m.mu_o = mu_o;
  m.mu_w = mu_w;
  m.c_o = c_o;
  m.c_w = c_w;
  m.rho_o = rho_o;
  m.rho_w = rho_w;
  return 0;
}

// rs/frontend.rs:611-618: set_initial_pressure also references the water PVT and the
// rock compressibility to the initial pressure
int orc_set_reference_pressure(orc_sim* p, double pressure) {
  Model& m = as(p)->m;
  m.p_ref_w = pressure;
  m.p_ref_rock = pressure;
  return 0;
}

int orc_set_rock_properties(orc_sim* p, double c_r, double depth_ref, double b_o, double b_w) {
  Model& m = as(p)->m;
  m.c_r = c_r;
  m.depth_ref = depth_ref;
  m.b_o = b_o;
  m.b_w = b_w;
  return 0;
}

int orc_set_gas_fluid_properties(orc_sim* p, double mu_g, double c_g, double rho_g) {
  Model& m = as(p)->m;
  m.mu_g = mu_g;
  m.c_g = c_g;
  m.rho_g = rho_g;
  return 0;
}

int orc_set_rel_perm_props(orc_sim* p, double s_wc, double s_or, double n_w, double n_o, double krw, double kro) {
  Model& m = as(p)->m;
  m.scal = {s_wc, s_or, n_w, n_o, krw, kro};
  return 0;
}

int orc_set_three_phase_rel_perm_props(orc_sim* p, double s_wc, double s_or, double s_gc, double s_gr, double s_org,
                                       double n_w, double n_o, double n_g, double krw, double kro, double krg) {
  Model& m = as(p)->m;
  m.has_scal3 = true;
  m.scal3 = {s_wc, s_or, s_gc, s_gr, s_org, n_w, n_o, n_g, krw, kro, krg};
  return 0;
}

// rs/frontend.rs:1145-1158 set_three_phase_scal_tables: rows (s, kr, kro) of SWOF and SGOF;
// the three-phase props must exist (1 otherwise); validation is the caller's
int orc_set_scal_tables(orc_sim* p, const double* swof, int n_swof, const double* sgof, int n_sgof) {
  Model& m = as(p)->m;
  if (!m.has_scal3) return 1;
  m.scal3.swof.assign(n_swof, {});
  m.scal3.sgof.assign(n_sgof, {});
  for (int r = 0; r < n_swof; ++r) m.scal3.swof[r] = {swof[3 * r], swof[3 * r + 1], swof[3 * r + 2]};
  for (int r = 0; r < n_sgof; ++r) m.scal3.sgof[r] = {sgof[3 * r], sgof[3 * r + 1], sgof[3 * r + 2]};
  return 0;
}

int orc_set_capillary_params(orc_sim* p, double p_entry, double lambda) {
  Model& m = as(p)->m;
  m.pc.p_entry = p_entry;
  m.pc.lambda = lambda;
  return 0;
}

int orc_set_gas_oil_capillary_params(orc_sim* p, double p_entry, double lambda) {
  Model& m = as(p)->m;
  m.has_pc_og = true;
  m.pc_og.p_entry = p_entry;
  m.pc_og.lambda = lambda;
  return 0;
}

int orc_set_flags(orc_sim* p, int three_phase, int gravity, int corey_points, int injected_fluid) {
  Model& m = as(p)->m;
  m.three_phase = three_phase != 0;
  m.gravity = gravity != 0;
  m.corey_points = corey_points;
  m.injected_fluid = injected_fluid;
  return 0;
}

int orc_set_pvt_table(orc_sim* p, const double* rows, int n_rows) {
  Model& m = as(p)->m;
  std::vector<PvtRow> r(n_rows);
  for (int i = 0; i < n_rows; ++i)
    r[i] = {rows[6 * i], rows[6 * i + 1], rows[6 * i + 2], rows[6 * i + 3], rows[6 * i + 4], rows[6 * i + 5]};
  m.table.build(r, m.c_o);  // rs/frontend.rs:1011-1021 uses the current c_o
  m.has_table = true;
  for (int i = 0; i < m.n(); ++i) m.rs[i] = m.table.rs_sat_at(m.pressure[i]);
  return 0;
}

int orc_set_well_controls(orc_sim* p, int inj_rate, int prod_rate, int inj_enabled, double tgt_inj, double tgt_prod,
                          double tgt_inj_surf, double tgt_prod_surf, double bhp_min, double bhp_max) {
  Model& m = as(p)->m;
  m.inj_rate = inj_rate != 0;
  m.prod_rate = prod_rate != 0;
  m.inj_enabled = inj_enabled != 0;
  m.tgt_inj = tgt_inj;
  m.tgt_prod = tgt_prod;
  m.has_tgt_inj_surf = tgt_inj_surf > 0.0;
  m.tgt_inj_surf = tgt_inj_surf;
  m.has_tgt_prod_surf = tgt_prod_surf > 0.0;
  m.tgt_prod_surf = tgt_prod_surf;
  m.bhp_min = bhp_min;
  m.bhp_max = bhp_max;
  return 0;
}

int orc_add_well(orc_sim* p, int i, int j, int k, double bhp, double rw, double skin, int injector, const char* id) {
  Model& m = as(p)->m;
  if (i < 0 || j < 0 || k < 0 || i >= m.nx || j >= m.ny || k >= m.nz) return 1;
  if (!(rw > 0)) return 1;
  for (auto& w : m.wells)
    if (w.i == i && w.j == j && w.k == k) return 1;  // rs/frontend.rs:155-164
  Well w{};
  w.i = i;
  w.j = j;
  w.k = k;
  w.bhp = bhp;
  w.rw = rw;
  w.skin = skin;
  w.injector = injector != 0;
  w.id = id ? id : "";
  m.wells.push_back(w);
  m.build_topology();
  return 0;
}

int orc_set_well_schedule(orc_sim* p, const char* id, int rate_mode, double target, double surface, double limit,
                          int enabled) {
  Model& m = as(p)->m;
  bool any = false;
  for (auto& w : m.wells)
    if (w.id == id) {
      w.schedule.has = true;
      w.schedule.rate = rate_mode != 0;
      w.schedule.has_target = std::isfinite(target) && target >= 0.0;
      w.schedule.target = target;
      w.schedule.has_surface = std::isfinite(surface) && surface >= 0.0;
      w.schedule.surface = surface;
      w.schedule.has_limit = std::isfinite(limit);
      w.schedule.limit = limit;
      w.schedule.enabled = enabled != 0;
      any = true;
    }
  return any ? 0 : 1;
}

int orc_set_well_datum(orc_sim* p, const char* id, double datum, double density) {
  Model& m = as(p)->m;
  bool any = false;
  for (auto& w : m.wells)
    if (w.id == id) {
      w.has_datum = std::isfinite(datum);
      w.datum = datum;
      w.has_density = std::isfinite(density) && density >= 0.0;
      w.density = density;
      any = true;
    }
  return any ? 0 : 1;
}

int orc_set_well_fluid(orc_sim* p, const char* id, int fluid) {
  Model& m = as(p)->m;
  bool any = false;
  for (auto& w : m.wells)
    if (w.id == id) {
      w.fluid = fluid;
      any = true;
    }
  return any ? 0 : 1;
}

int orc_set_gas_redissolution(orc_sim* p, int enabled) {
  as(p)->m.gas_redissolution = enabled != 0;
  return 0;
}

int orc_set_newton_options(orc_sim* p, int max_it, int min_it, double residual_tol, double lin_rel_tol,
                           int lin_max_it, int allow_direct) {
  Sim* s = as(p);
  s->nopts.max_iterations = max_it;
  s->nopts.min_iteration_index = min_it;
  s->nopts.residual_tol = residual_tol;
  s->nopts.lin.rel_tol = lin_rel_tol;
  s->nopts.lin.max_iterations = lin_max_it;
  s->nopts.lin.restart = std::max(30, lin_max_it);
  s->nopts.lin.allow_direct = allow_direct != 0;
  return 0;
}

int orc_step(orc_sim* p, double dt) { return step(*as(p), dt); }

int orc_run_substep(orc_sim* p, double dt, int* newton_its, long* linear_its) {
  Sim* s = as(p);
  s->m.refresh_head_offsets();
  StepReport rep;
  bool ok = substep(*s, dt, rep);
  if (newton_its) *newton_its = rep.newton_iterations;
  if (linear_its) *linear_its = rep.linear_iterations;
  return ok ? 0 : 2;
}

int orc_get_field(orc_sim* p, int which, double* out) {
  Model& m = as(p)->m;
  const std::vector<double>* f = nullptr;
  switch (which) {
    case 0: f = &m.pressure; break;
    case 1: f = &m.sw; break;
    case 2: f = &m.so; break;
    case 3: f = &m.sg; break;
    case 4: f = &m.rs; break;
    default: return 1;
  }
  std::copy(f->begin(), f->end(), out);
  return 0;
}

int orc_num_wells(orc_sim* p) { return (int)as(p)->m.topo_wells.size(); }
int orc_num_perfs(orc_sim* p) { return (int)as(p)->m.wells.size(); }

// Perforation q and physical-well BHP of the consistent state at the stored simulator state.
int orc_get_well_state(orc_sim* p, double* bhp, double* q) {
  Model& m = as(p)->m;
  m.refresh_head_offsets();
  State s = from_simulator(m);
  for (size_t w = 0; w < s.bhp.size(); ++w) bhp[w] = s.bhp[w];
  for (size_t e = 0; e < s.q.size(); ++e) q[e] = s.q[e];
  return 0;
}

double orc_get_time(orc_sim* p) { return as(p)->m.time; }

int orc_get_stats(orc_sim* p, long* out) {
  Sim* s = as(p);
  out[0] = s->total_newton;
  out[1] = s->total_linear;
  out[2] = s->total_substeps;
  out[3] = s->total_retries;
  out[4] = s->lstats.bicgstab_iterations;
  return 0;
}

int orc_classify_retry(int have_linear, int converged, int direct, int coarse_applications, double average_reduction,
                       double last_reduction) {
  LinReport l;
  l.converged = converged != 0;
  l.direct = direct != 0;
  l.cpr_apps = coarse_applications;
  l.cpr_avg = average_reduction;
  l.cpr_last = last_reduction;
  return retry_class(have_linear != 0, l);
}

int orc_dominant_family(const double* family_maxima, int with_wells) {
  return dominant_family(family_maxima, with_wells != 0);
}

int orc_family_peaks(int n_cells, int n_wells, int n_perf, const double* residual, const double* scale,
                     double* peak_values, long* peak_rows) {
  if (n_cells < 1 || n_wells < 0 || n_perf < 0 || !residual || !scale || !peak_values || !peak_rows) return -1;
  const size_t m = 3 * (size_t)n_cells + n_wells + n_perf;
  Vec R(residual, residual + m);
  Scaling sc;
  for (int i = 0; i < n_cells; ++i) {
    sc.w.push_back(scale[3 * i]);
    sc.o.push_back(scale[3 * i + 1]);
    sc.g.push_back(scale[3 * i + 2]);
  }
  for (int k = 0; k < n_wells; ++k) sc.wc.push_back(scale[3 * n_cells + k]);
  for (int k = 0; k < n_perf; ++k) sc.pf.push_back(scale[3 * n_cells + n_wells + k]);
  const Families fm = family_norms(R, sc, n_cells);
  for (int k = 0; k < 5; ++k) peak_values[k] = fm.f[k], peak_rows[k] = fm.row[k];
  return dominant_family(fm.f, n_wells > 0);
}

int orc_step_stats_count(orc_sim* p) { return (int)as(p)->step_stats.size(); }

int orc_get_step_stats(orc_sim* p, int index, orc_step_stats* out) {
  Sim* s = as(p);
  if (!out || index < 0 || index >= (int)s->step_stats.size()) return -1;
  *out = s->step_stats[index].s;
  return 0;
}

int orc_get_step_rungs(orc_sim* p, int index, orc_step_rung* out, int cap) {
  Sim* s = as(p);
  if (index < 0 || index >= (int)s->step_stats.size() || cap < 0 || (cap > 0 && !out)) return -1;
  const auto& r = s->step_stats[index].rungs;
  for (int i = 0; i < (int)r.size() && i < cap; ++i) out[i] = r[i];
  return (int)r.size();
}

int orc_get_accepted_dts(orc_sim* p, double* out, int cap) {
  Sim* s = as(p);
  int n = (int)s->accepted_dts.size();
  for (int i = 0; i < n && i < cap; ++i) out[i] = s->accepted_dts[i];
  return n;
}

// Assemble R and J at the stored simulator state (previous = current) with a trial
// state offset; used by parity tests of the assembly kernels. Returns CSR sizes.
struct OrcAsm {
  Vec R;
  Csr J;
};
static OrcAsm g_last_asm;

int orc_assemble(orc_sim* p, double dt, const double* cell_state, const int* regimes, const double* bhp,
                 const double* q, int* n_unknowns, int* nnz) {
  Sim* S = as(p);
  Model& m = S->m;
  m.refresh_head_offsets();
  State prev = from_simulator(m);
  State s = prev;
  int n = m.n();
  if (cell_state) {
    for (int i = 0; i < n; ++i) {
      s.cells[i].p = cell_state[3 * i];
      s.cells[i].sw = cell_state[3 * i + 1];
      s.cells[i].hc = cell_state[3 * i + 2];
      if (regimes) s.cells[i].regime = regimes[i];
    }
  }
  if (bhp)
    for (size_t w = 0; w < s.bhp.size(); ++w) s.bhp[w] = bhp[w];
  if (q)
    for (size_t e = 0; e < s.q.size(); ++e) s.q[e] = q[e];
  Assembly a = assemble(m, prev, s, dt, false);
  g_last_asm.R = a.R;
  g_last_asm.J = a.J;
  *n_unknowns = a.J.n;
  *nnz = (int)a.J.col.size();
  return 0;
}

// The reference's acceptance measures at an iterate against the stored state:
// out[0..4] scaled family peaks (water, oil, gas, well, perforation;
// rs/fim/newton/convergence.rs:734-790), out[5..7] CNV and out[8..10] MB
// (rs/fim/newton/convergence.rs:478-612).
// The iterate (cell primaries interleaved [p, sw, hc], regimes, BHPs, completion rates)
// assembled against the stored state as the previous one.
struct IterateAssembly {
  State s;
  Assembly a;
  Scaling sc;
};
static IterateAssembly assemble_iterate(Sim* S, double dt, const double* cell_state, const int* regimes,
                                        const double* bhp, const double* q) {
  Model& m = S->m;
  m.refresh_head_offsets();
  State prev = from_simulator(m);
  State s = prev;
  int n = m.n();
  for (int i = 0; i < n; ++i) {
    s.cells[i].p = cell_state[3 * i];
    s.cells[i].sw = cell_state[3 * i + 1];
    s.cells[i].hc = cell_state[3 * i + 2];
    s.cells[i].regime = regimes[i];
  }
  for (size_t w = 0; w < s.bhp.size(); ++w) s.bhp[w] = bhp[w];
  for (size_t e = 0; e < s.q.size(); ++e) s.q[e] = q[e];
  Assembly a = assemble(m, prev, s, dt, false);
  Scaling sc = equation_scaling(m, s, dt);
  return {std::move(s), std::move(a), std::move(sc)};
}

int orc_residual_check(orc_sim* p, double dt, const double* cell_state, const int* regimes, const double* bhp,
                       const double* q, double* out) {
  Sim* S = as(p);
  const Model& m = S->m;
  const int n = m.n();
  IterateAssembly it = assemble_iterate(S, dt, cell_state, regimes, bhp, q);
  const State& s = it.s;
  Families fm = family_norms(it.a.R, it.sc, n);
  CnvMb cm = cnv_mb(m, s, it.a.R, false);
  CnvMb cf = cnv_mb(m, s, it.a.R, true);
  for (int k = 0; k < 5; ++k) out[k] = fm.f[k];
  for (int k = 0; k < 3; ++k) {
    out[5 + k] = cm.cnv[k];
    out[8 + k] = cm.mb[k];
  }
  // the acceptance inputs of rs/fim/newton.rs:1262-1300: the violating pore-volume
  // fraction and its 3 % rule, the verdict with the strict and with the final-iteration
  // tiers, and every well's own convergence test (rs/fim/wells_inner.rs:600-640)
  out[11] = cm.violating_fraction;
  out[12] = cm.pv_relax ? 1.0 : 0.0;
  out[13] = cm.accept ? 1.0 : 0.0;
  out[14] = cf.accept ? 1.0 : 0.0;
  out[15] = all_wells_converged(m, s, WellInnerOpts{}) ? 1.0 : 0.0;
  for (int k = 0; k < 5; ++k) out[16 + k] = (double)fm.row[k];
  return 0;
}

int orc_scaled_residual(orc_sim* p, double dt, const double* cell_state, const int* regimes, const double* bhp,
                        const double* q, double* out) {
  Sim* S = as(p);
  const int n = S->m.n();
  IterateAssembly it = assemble_iterate(S, dt, cell_state, regimes, bhp, q);
  const Scaling& sc = it.sc;
  const Vec& R = it.a.R;
  for (int i = 0; i < n; ++i) {
    out[3 * i] = std::fabs(R[3 * i]) / sc.w[i];
    out[3 * i + 1] = std::fabs(R[3 * i + 1]) / sc.o[i];
    out[3 * i + 2] = std::fabs(R[3 * i + 2]) / sc.g[i];
  }
  const size_t W = sc.wc.size();
  for (size_t k = 0; k < W; ++k) out[3 * n + k] = std::fabs(R[3 * n + k]) / sc.wc[k];
  for (size_t k = 0; k < sc.pf.size(); ++k) out[3 * n + W + k] = std::fabs(R[3 * n + W + k]) / sc.pf[k];
  return 0;
}

double orc_producer_oil(orc_sim* p) { return as(p)->producer_oil; }

int orc_rate_history_len(orc_sim* p) { return (int)as(p)->rate_history.size(); }

// the first min(cap, count) TimePointRates, 20 doubles each in declaration order
int orc_get_rate_history(orc_sim* p, double* out, int cap) {
  const auto& h = as(p)->rate_history;
  const int k = std::min(cap, (int)h.size());
  static_assert(sizeof(TimePoint) == 20 * sizeof(double), "20 doubles per entry");
  if (k > 0) std::memcpy(out, h.data(), (size_t)k * sizeof(TimePoint));
  return k;
}

// component inventories [water, oil, gas] in sc m3 summed over the cells of the stored state
// (rs/fim/properties.rs:146-166)
int orc_inventory_totals(orc_sim* p, double* out) {
  Model& m = as(p)->m;
  State s = from_simulator(m);
  out[0] = out[1] = out[2] = 0.0;
  for (int i = 0; i < m.n(); ++i) {
    const auto& c = s.cells[i];
    auto pr = m.cell_props<double>(c.regime, c.p, c.sw, c.hc, m.rs_cap_of(i));
    double inv[3];
    m.inventory<double>(m.pore_volume_g<double>(i, c.p), c.p, c.sw, pr, inv);
    for (int k = 0; k < 3; ++k) out[k] += inv[k];
  }
  return 0;
}

int orc_get_assembly(double* R, int* ptr, int* col, double* val) {
  std::copy(g_last_asm.R.begin(), g_last_asm.R.end(), R);
  std::copy(g_last_asm.J.ptr.begin(), g_last_asm.J.ptr.end(), ptr);
  std::copy(g_last_asm.J.col.begin(), g_last_asm.J.col.end(), col);
  std::copy(g_last_asm.J.val.begin(), g_last_asm.J.val.end(), val);
  return 0;
}

// Initial (from_simulator) state: regimes and hydrocarbon variable per cell.
int orc_get_initial_state(orc_sim* p, double* cell_state, int* regimes) {
  Model& m = as(p)->m;
  m.refresh_head_offsets();
  State s = from_simulator(m);
  for (int i = 0; i < m.n(); ++i) {
    cell_state[3 * i] = s.cells[i].p;
    cell_state[3 * i + 1] = s.cells[i].sw;
    cell_state[3 * i + 2] = s.cells[i].hc;
    regimes[i] = s.cells[i].regime;
  }
  return 0;
}

// initial well iterate of FimState::from_simulator (rs/fim/state.rs:109-177)
int orc_get_initial_well_state(orc_sim* p, double* bhp, double* q) {
  Model& m = as(p)->m;
  m.refresh_head_offsets();
  State s = from_simulator(m);
  std::copy(s.bhp.begin(), s.bhp.end(), bhp);
  std::copy(s.q.begin(), s.q.end(), q);
  return 0;
}

// Property probes for known-answer tests (value and d/dx by AD).
int orc_probe(orc_sim* p, int what, double x, double y, double* out) {
  Model& m = as(p)->m;
  switch (what) {
    case 0: {  // saturated PVT at p=x: rs, bo, bg, mu_o, mu_g and d/dp
      auto s = m.table.saturated(Ad<1>::var(x, 0));
      Ad<1> v[5] = {s.rs, s.bo, s.bg, s.mu_o, s.mu_g};
      for (int i = 0; i < 5; ++i) { out[2 * i] = v[i].v; out[2 * i + 1] = v[i].d[0]; }
      return 0;
    }
    case 1: {  // oil (bo, mu) at p=x, rs=y with d/dp, d/drs
      Ad<2> bo, mu;
      m.table.oil(Ad<2>::var(x, 0), Ad<2>::var(y, 1), bo, mu);
      out[0] = bo.v; out[1] = bo.d[0]; out[2] = bo.d[1];
      out[3] = mu.v; out[4] = mu.d[0]; out[5] = mu.d[1];
      return 0;
    }
    case 2: {  // two-phase table relperm at sw=x: krw, dkrw, kro, dkro
      Ad<1> krw, kro;
      m.scal.table(Ad<1>::var(x, 0), m.corey_points, krw, kro);
      out[0] = krw.v; out[1] = krw.d[0]; out[2] = kro.v; out[3] = kro.d[0];
      return 0;
    }
    case 3: {  // Stone II kro at sw=x, sg=y
      Ad<2> k = m.scal3.kro_stone2(Ad<2>::var(x, 0), Ad<2>::var(y, 1));
      out[0] = k.v; out[1] = k.d[0]; out[2] = k.d[1];
      return 0;
    }
    case 4: {  // Pc_ow at sw=x (cell scale 1)
      Ad<1> pc = m.pc.pc(Ad<1>::var(x, 0), m.scal, 1.0);
      out[0] = pc.v; out[1] = pc.d[0];
      return 0;
    }
    case 5: {  // Peaceman WI of completion index (int)x
      double wi = 0.0;
      bool ok = m.well_index((int)x, wi);
      out[0] = ok ? wi : NAN;
      return 0;
    }
    case 6: {  // geometric transmissibility between cells (int)x, (int)y along dim given by out[0]
      char dim = out[0] == 0.0 ? 'x' : (out[0] == 1.0 ? 'y' : 'z');
      out[0] = m.geom_trans((int)x, (int)y, dim);
      return 0;
    }
    case 8: {  // three-phase relperms at sw=x, sg=y: krw, krg, krow, krog, each with its derivative
      Ad<1> sw = Ad<1>::var(x, 0), sg = Ad<1>::var(y, 0);
      Ad<1> v[4] = {m.scal3.krw(sw), m.scal3.krg(sg), m.scal3.kro_w(sw), m.scal3.kro_g(sg)};
      for (int i = 0; i < 4; ++i) { out[2 * i] = v[i].v; out[2 * i + 1] = v[i].d[0]; }
      return 0;
    }
    default:
      return 1;
  }
}

const char* orc_last_warning(orc_sim* p) { return as(p)->warning.c_str(); }

// Per-cell OPM chop of a raw update (known-answer tests of rs/fim/newton/tests.rs:159-292).
int orc_chop(int n, const double* cells, const int* regimes, int n_wells, const double* bhp, const double* update,
             double relax, double* out) {
  State s;
  s.cells.resize(n);
  for (int i = 0; i < n; ++i) s.cells[i] = {cells[3 * i], cells[3 * i + 1], cells[3 * i + 2], regimes[i]};
  s.bhp.assign(bhp, bhp + n_wells);
  Vec up(update, update + 3 * n + n_wells);
  Vec c = chop(s, up, relax, n);
  std::copy(c.begin(), c.end(), out);
  return 0;
}

// CNV/MB from raw parts (rs/fim/newton/convergence.rs:478-612); out = cnv[3], mb[3],
// violating pv fraction, pv-rule relaxes, accept.
int orc_cnv_mb(int n, const double* R, const double* pv, const double* fvf, int relax_final, double* out) {
  double bavg[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) bavg[c] += fvf[3 * i + c];
  for (double& b : bavg) b /= std::max(n, 1);
  double maxc[3] = {0, 0, 0}, rsum[3] = {0, 0, 0}, pvsum = 0, viol = 0;
  for (int i = 0; i < n; ++i) {
    double v = std::max(pv[i], 1e-9);
    pvsum += v;
    double cm = 0;
    for (int c = 0; c < 3; ++c) {
      double r = R[3 * i + c];
      rsum[c] += r;
      maxc[c] = std::max(maxc[c], std::fabs(r) / v);
      cm = std::max(cm, std::fabs(r) * bavg[c] / v);
    }
    if (cm > 1e-2) viol += v;
  }
  bool relax = viol < 0.03 * pvsum;
  double mbt = relax_final ? 1e-6 : 1e-7, cnvt = relax_final ? 1.0 : 1e-2;
  bool ok = true;
  for (int c = 0; c < 3; ++c) {
    out[c] = bavg[c] * maxc[c];
    out[3 + c] = std::fabs(bavg[c] * rsum[c]) / std::max(pvsum, 1e-9);
    ok = ok && (out[c] <= cnvt || (relax && out[c] <= 1.0)) && out[3 + c] <= mbt;
  }
  out[6] = viol / std::max(pvsum, 1e-9);
  out[7] = relax ? 1.0 : 0.0;
  out[8] = ok ? 1.0 : 0.0;
  return 0;
}

// Closed-form AD examples of rs/fim/ad.rs:429-499, evaluated with this oracle's Ad.
int orc_ad_selftest(double* out) {
  Ad<2> x = Ad<2>::var(2.0, 0), y = Ad<2>::var(3.0, 1);
  Ad<2> f = x * y / (x + y);
  out[0] = f.v; out[1] = f.d[0]; out[2] = f.d[1];
  Ad<1> e = exp_(sqrt_(Ad<1>::var(4.0, 0)));
  out[3] = e.v; out[4] = e.d[0];
  Ad<1> pw = powf_(Ad<1>::var(4.0, 0), 1.5);
  out[5] = pw.v; out[6] = pw.d[0];
  Ad<1> r = recip(Ad<1>::var(2.0, 0));
  out[7] = r.v; out[8] = r.d[0];
  Ad<1> z = Ad<1>::var(3.0, 0);
  Ad<1> g = 10.0 - 2.0 * z + z / 2.0;
  out[9] = g.v; out[10] = g.d[0];
  Ad<1> m = max_of(Ad<1>::var(1.0, 0), Ad<1>(2.0));
  out[11] = m.v; out[12] = m.d[0];
  Ad<1> m2 = max_floor(Ad<1>::var(1.0, 0), 0.5);
  out[13] = m2.v; out[14] = m2.d[0];
  return 0;
}

// Newton direction of J dx = -R through this oracle's linear stack (Schur + CPR-FGMRES).
int orc_set_sweep_config(orc_sim* p, const char* geometry, double threshold, double initial_so, double sor) {
  Sim* S = as(p);
  if (!geometry) {
    S->sweep = SweepConfig{};
    return 0;
  }
  const std::string g = geometry;
  S->sweep.on = true;
  S->sweep.geometry = g == "areal" ? 0 : (g == "vertical" ? 1 : 2);
  S->sweep.threshold = threshold;
  S->sweep.initial_so = initial_so;
  S->sweep.sor = sor;
  return 0;
}

int orc_set_nested_well_solve(orc_sim* p, int enabled) {
  as(p)->nopts.nested_well_solve = enabled != 0;
  return 0;
}

// State at a given iterate for the per-well functions (the stored state where null)
static State iterate_state(Model& m, const double* cell_state, const int* regimes, const double* bhp,
                           const double* q) {
  m.refresh_head_offsets();
  State s = from_simulator(m);
  if (cell_state)
    for (int i = 0; i < m.n(); ++i) {
      s.cells[i].p = cell_state[3 * i];
      s.cells[i].sw = cell_state[3 * i + 1];
      s.cells[i].hc = cell_state[3 * i + 2];
      if (regimes) s.cells[i].regime = regimes[i];
    }
  if (bhp)
    for (size_t w = 0; w < s.bhp.size(); ++w) s.bhp[w] = bhp[w];
  if (q)
    for (size_t e = 0; e < s.q.size(); ++e) s.q[e] = q[e];
  return s;
}

int orc_well_local_system(orc_sim* p, const double* cell_state, const int* regimes, const double* bhp,
                          const double* q, int well, double* r, double* J) {
  Model& m = as(p)->m;
  if (well < 0 || well >= (int)m.topo_wells.size()) return 1;
  State s = iterate_state(m, cell_state, regimes, bhp, q);
  WellLocal L = well_local_system(m, s, well);
  std::copy(L.r.begin(), L.r.end(), r);
  std::copy(L.J.begin(), L.J.end(), J);
  return L.dim;
}

int orc_solve_wells_locally(orc_sim* p, const double* cell_state, const int* regimes, double* bhp, double* q,
                            int max_iterations, int* converged, int* iterations, double* peak) {
  Model& m = as(p)->m;
  State s = iterate_state(m, cell_state, regimes, bhp, q);
  WellInnerOpts o;
  if (max_iterations >= 0) o.max_iterations = max_iterations;
  auto reps = solve_wells_locally(m, s, o);
  for (size_t w = 0; w < reps.size(); ++w) {
    bhp[w] = s.bhp[w];
    if (converged) converged[w] = reps[w].converged ? 1 : 0;
    if (iterations) iterations[w] = reps[w].iterations;
    if (peak) peak[w] = reps[w].peak;
  }
  for (size_t e = 0; e < s.q.size(); ++e) q[e] = s.q[e];
  return 0;
}

double orc_chop_bhp_update(double bhp, double raw_delta, double rel) { return chop_bhp_update(bhp, raw_delta, rel); }

// relax_wells at an iterate (rs/fim/state.rs:494-551); bhp and q are updated in place
int orc_relax_wells(orc_sim* p, const double* cell_state, const int* regimes, double* bhp, double* q) {
  Model& m = as(p)->m;
  State s = iterate_state(m, cell_state, regimes, bhp, q);
  relax_wells(m, s);
  std::copy(s.bhp.begin(), s.bhp.end(), bhp);
  std::copy(s.q.begin(), s.q.end(), q);
  return 0;
}

double orc_pid_dt(double dt, const double* err) { return opm_pid_dt(dt, err); }
double orc_iteration_count_dt(double dt, int its) { return opm_iteration_count_dt(dt, its); }
double orc_growth_factor(double dt, int its, const double* err, int retries, int* limiter) {
  return opm_growth_factor(dt, its, err, retries, limiter);
}
int orc_detect_oscillation(const double* cur, const double* p1, const double* p2) {
  return detect_oscillation(cur, p1, p2);
}
double orc_next_relaxation_factor(double relax, int oscillating) { return next_relaxation_factor(relax, oscillating); }

// opm_relative_change between two cell states ([p, sw, hc] per cell and regimes)
double orc_relative_change(int n, const double* a, const int* ra, const double* b, const int* rb) {
  State sa, sb;
  sa.cells.resize(n);
  sb.cells.resize(n);
  for (int i = 0; i < n; ++i) {
    sa.cells[i] = {a[3 * i], a[3 * i + 1], a[3 * i + 2], ra[i]};
    sb.cells[i] = {b[3 * i], b[3 * i + 1], b[3 * i + 2], rb[i]};
  }
  return opm_relative_change(sa, sb);
}

int orc_newton_direction(orc_sim* p, double dt, const double* cell_state, const int* regimes, const double* bhp,
                         const double* q, double rel_tol, int allow_direct, double* dx, int* iterations) {
  int nu, nnz;
  orc_assemble(p, dt, cell_state, regimes, bhp, q, &nu, &nnz);
  Vec rhs(g_last_asm.R.size());
  for (size_t i = 0; i < rhs.size(); ++i) rhs[i] = -g_last_asm.R[i];
  LinOpts o;
  o.rel_tol = rel_tol;
  o.max_iterations = 300;
  o.restart = 300;
  o.allow_direct = allow_direct != 0;
  LinStats st;
  LinReport r = solve_linearized(g_last_asm.J, rhs, as(p)->m.n(), o, &st);
  std::copy(r.x.begin(), r.x.end(), dx);
  if (iterations) *iterations = r.iterations;
  return r.converged ? 0 : 2;
}

}  // extern "C"

// ---------------------------------------------------------------- threads
#ifdef _OPENMP
#include <omp.h>
#endif
extern "C" {
// threads of the parallel loops (n <= 0: the OpenMP default); returns the count in effect
int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}
}  // extern "C"

// ---------------------------------------------------------------- linear-stack probes
// The CPR components of linear.hpp on caller-supplied matrices, so that each can be pinned
// on its own (the reference's unit tests of rs/fim/linear/gmres_block_jacobi.rs:2860-3360).
extern "C" {

static Csr probe_csr(int n, const int* ptr, const int* col, const double* val) {
  Csr m;
  m.n = n;
  m.ptr.assign(ptr, ptr + n + 1);
  m.col.assign(col, col + ptr[n]);
  m.val.assign(val, val + ptr[n]);
  return m;
}

// factorize_pressure_ilu0 (:746-808) of a scalar CSR matrix: the product L U written densely
// (n x n, row-major; L unit lower, U with its diagonal) and the apply (:634-658) to rhs.
int orc_scalar_ilu0(int n, const int* ptr, const int* col, const double* val, const double* rhs, double* lu,
                    double* x) {
  Csr A = probe_csr(n, ptr, col, val);
  std::vector<std::vector<std::pair<int, double>>> rows(n);
  for (int i = 0; i < n; ++i)
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) rows[i].push_back({A.col[k], A.val[k]});
  ScalarIlu0 f;
  f.factor(rows);
  if (lu) {
    Dense L(n), U(n);
    for (int i = 0; i < n; ++i) {
      L(i, i) = 1.0;
      for (auto& e : f.Lr[i]) L(i, e.first) = e.second;
      U(i, i) = f.ud[i];
      for (auto& e : f.Ur[i]) U(i, e.first) = e.second;
    }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += L(i, k) * U(k, j);
        lu[(size_t)i * n + j] = s;
      }
  }
  if (rhs && x) {
    Vec r(rhs, rhs + n);
    Vec y = f.apply(r);
    std::copy(y.begin(), y.end(), x);
  }
  return 0;
}

// factorize_block_ilu0 (:239-354) of a cell system of nb 3x3 blocks and its apply (:190-237)
int orc_block_ilu0_apply(int nb, const int* ptr, const int* col, const double* val, const double* rhs, double* x) {
  Csr A = probe_csr(3 * nb, ptr, col, val);
  BlockIlu0 f;
  f.factor(A, nb);
  Vec r(rhs, rhs + 3 * nb);
  Vec y = f.apply(r);
  std::copy(y.begin(), y.end(), x);
  return 0;
}

// The CPR of a cell system of nb 3x3 blocks (:1022-1363 build): quasi-IMPES weights (nb x 3),
// the coarse pressure matrix (dense nb x nb) and a pressure solve of rp. mode 0: the
// pressure stage as the preconditioner applies it (the dense inverse up to 299 rows, else
// BiCGStab with ILU0, 50 iterations, 1e-6 relative); mode 1: BiCGStab with max_it and the
// absolute tolerance tol (solve_pressure_with_bicgstab, :660-730). info: [dense, BiCGStab
// iterations]. z (optional, 3 nb): the full two-stage apply (:471-501) of rp read as a cell
// vector when rp has 3 nb entries (mode 2).
int orc_cpr_probe(int nb, const int* ptr, const int* col, const double* val, const double* rp, int mode, int max_it,
                  double tol, double* weights, double* pmat, double* xp, int* info) {
  Csr A = probe_csr(3 * nb, ptr, col, val);
  LinStats st;
  Cpr M;
  M.stats = &st;
  M.build(A, nb);
  if (weights)
    for (int b = 0; b < nb; ++b)
      for (int l = 0; l < 3; ++l) weights[3 * b + l] = M.w[b][l];
  if (pmat) {
    std::fill(pmat, pmat + (size_t)nb * nb, 0.0);
    for (int i = 0; i < nb; ++i)
      for (auto& e : M.prow[i]) pmat[(size_t)i * nb + e.first] = e.second;
  }
  Vec out;
  if (mode == 2) {
    Vec v(rp, rp + 3 * nb);
    out = M.apply(A, v);
  } else if (mode == 1) {
    Vec b(rp, rp + nb);
    out = M.bicgstab(b, max_it, tol);
  } else {  // the coarse solve of apply (:585-610), on rp as the restricted residual
    Vec b(rp, rp + nb);
    if (M.dense) {
      out.assign(nb, 0.0);
      for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) out[i] += M.pinv(i, j) * b[j];
    } else {
      out = M.bicgstab(b, 50, std::max(norm2(b), EPS) * 1e-6);
    }
  }
  if (xp) std::copy(out.begin(), out.end(), xp);
  if (info) {
    info[0] = M.dense ? 1 : 0;
    info[1] = (int)st.bicgstab_iterations;
  }
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- discrete adjoint terms
// The pieces of the adjoint of the FIM step that the test driver (oracle/adjoint.py)
// combines with this file's AD Jacobian: the converged iterate of a substep, its objective
// term with the gradient, the inventory Jacobian that couples a substep to the next, and
// lambda^T dR/dtheta for theta = a per-cell log-multiplier of kx, ky, kz and the porosity.
extern "C" {

static State probe_state(const Model& m, const double* cells, const int* regimes, const double* bhp, const double* q) {
  State s = from_simulator(m);
  const int n = m.n();
  for (int i = 0; i < n; ++i) s.cells[i] = {cells[3 * i], cells[3 * i + 1], cells[3 * i + 2], regimes[i]};
  if (bhp) std::copy(bhp, bhp + s.bhp.size(), s.bhp.begin());
  if (q) std::copy(q, q + s.q.size(), s.q.begin());
  return s;
}

// converged iterate of the last accepted substep: cells [p, sw, hc], regimes, BHP per well,
// q per completion
int orc_last_iterate(orc_sim* p, double* cells, int* regimes, double* bhp, double* q) {
  const State& s = as(p)->last_iterate;
  if (s.cells.empty()) return 1;
  for (size_t i = 0; i < s.cells.size(); ++i) {
    cells[3 * i] = s.cells[i].p;
    cells[3 * i + 1] = s.cells[i].sw;
    cells[3 * i + 2] = s.cells[i].hc;
    regimes[i] = s.cells[i].regime;
  }
  std::copy(s.bhp.begin(), s.bhp.end(), bhp);
  std::copy(s.q.begin(), s.q.end(), q);
  return 0;
}

// assemble_fim_system_ad at `cells` against an explicit previous state (prev_cells,
// prev_regimes) instead of the stored one; read with orc_get_assembly
int orc_assemble_prev(orc_sim* p, double dt, const double* prev_cells, const int* prev_regimes, const double* cells,
                      const int* regimes, const double* bhp, const double* q, int* n_unknowns, int* nnz) {
  Model& m = as(p)->m;
  State prev = probe_state(m, prev_cells, prev_regimes, nullptr, nullptr);
  State s = probe_state(m, cells, regimes, bhp, q);
  Assembly a = assemble(m, prev, s, dt, false);
  g_last_asm.R = a.R;
  g_last_asm.J = a.J;
  *n_unknowns = a.J.n;
  *nnz = (int)a.J.col.size();
  return 0;
}

// objective term dt sum_producers (fo/Bo) q at an iterate (producer_oil) and its gradient
// over the unknowns (3n cells, then wells, then completions); dinv: the inventory Jacobian
// d inv_i / d x_i, 9 per cell (row = component, column = unknown)
int orc_adjoint_state_terms(orc_sim* p, double dt, const double* cells, const int* regimes, const double* bhp,
                            const double* q, double* j, double* djdx, double* dinv) {
  Model& m = as(p)->m;
  State s = probe_state(m, cells, regimes, bhp, q);
  const int n = m.n(), W = (int)s.bhp.size(), P = (int)s.q.size();
  std::fill(djdx, djdx + 3 * n + W + P, 0.0);
  double tot = 0.0;
  for (int e = 0; e < P; ++e) {
    int w = m.perf_well[e];
    if (m.topo_wells[w].injector || !m.control(w).enabled) continue;
    const Well& wl = m.wells[e];
    const int id = m.idx(wl.i, wl.j, wl.k);
    const auto& c = s.cells[id];
    WCell<Ad<3>> ac{Ad<3>::var(c.p, 0), Ad<3>::var(c.sw, 1), Ad<3>::var(c.hc, 2), c.regime, id};
    Ad<3> coef[3];
    rate_coefficients<Ad<3>>(m, false, m.fluid_of(w), ac, coef);
    tot += coef[1].v * s.q[e] * dt;
    for (int v = 0; v < 3; ++v) djdx[3 * id + v] += coef[1].d[v] * s.q[e] * dt;
    djdx[3 * n + W + e] += coef[1].v * dt;
  }
  *j = tot;
  for (int i = 0; i < n; ++i) {
    const auto& c = s.cells[i];
    Ad<3> acc[3];
    m.accumulation<Ad<3>>(i, c.regime, Ad<3>::var(c.p, 0), Ad<3>::var(c.sw, 1), Ad<3>::var(c.hc, 2), c.regime, c.p,
                          c.sw, c.hc, acc);
    for (int r = 0; r < 3; ++r)
      for (int v = 0; v < 3; ++v) dinv[9 * i + 3 * r + v] = acc[r].d[v];
  }
  return 0;
}

// g_lnk += lambda^T dR/d(ln k multiplier), g_phi += lambda^T dR/dphi for the residual of one
// substep at (prev, x): the accumulation is linear in phi, each face term linear in its
// transmissibility (d ln T / d ln k_a = k_b/(k_a+k_b) on x and y faces, kz_b dz_a /
// (kz_b dz_a + kz_a dz_b) on z faces), each connection rate linear in its Peaceman index
// (d ln WI / d ln k = 1)
int orc_adjoint_param_terms(orc_sim* p, double dt, const double* prev_cells, const int* prev_regimes,
                            const double* cells, const int* regimes, const double* bhp, const double* q,
                            const double* lambda, double* g_lnk, double* g_phi) {
  Model& m = as(p)->m;
  State prev = probe_state(m, prev_cells, prev_regimes, nullptr, nullptr);
  State s = probe_state(m, cells, regimes, bhp, q);
  const int n = m.n(), W = (int)s.bhp.size();
  for (int i = 0; i < n; ++i) {
    const auto& c = s.cells[i];
    const auto& pc = prev.cells[i];
    double acc[3];
    m.accumulation<double>(i, c.regime, c.p, c.sw, c.hc, pc.regime, pc.p, pc.sw, pc.hc, acc);
    double t = 0.0;
    for (int r = 0; r < 3; ++r) t += lambda[3 * i + r] * acc[r];
    g_phi[i] += t / m.phi[i];
  }
  auto face = [&](int a, int b, char dim, int ka, int kb) {
    const double gt = DARCY_METRIC_FACTOR * m.geom_trans(a, b, dim);
    if (gt <= 0.0) return;
    const auto& ci = s.cells[a];
    const auto& cj = s.cells[b];
    double f[3];
    m.face_flux<double>(gt, a, b, ci.regime, cj.regime, m.depth_at_k(ka), m.depth_at_k(kb), ci.p, ci.sw, ci.hc, cj.p,
                        cj.sw, cj.hc, f);
    double t = 0.0;  // lambda^T dR/d ln T
    for (int r = 0; r < 3; ++r) t += (lambda[3 * a + r] - lambda[3 * b + r]) * f[r] * dt;
    double wa, wb;
    if (dim == 'z') {
      const double da = m.dz[ka], db = m.dz[kb];
      const double den = m.kz[b] * da + m.kz[a] * db;
      wa = m.kz[b] * da / den;
      wb = m.kz[a] * db / den;
    } else {
      const double ka_ = dim == 'x' ? m.kx[a] : m.ky[a], kb_ = dim == 'x' ? m.kx[b] : m.ky[b];
      wa = kb_ / (ka_ + kb_);
      wb = ka_ / (ka_ + kb_);
    }
    g_lnk[a] += t * wa;
    g_lnk[b] += t * wb;
  };
  for (int k = 0; k < m.nz; ++k)
    for (int jj = 0; jj < m.ny; ++jj)
      for (int i = 0; i < m.nx; ++i) {
        const int id = m.idx(i, jj, k);
        if (i + 1 < m.nx) face(id, m.idx(i + 1, jj, k), 'x', k, k);
        if (jj + 1 < m.ny) face(id, m.idx(i, jj + 1, k), 'y', k, k);
        if (k + 1 < m.nz) face(id, m.idx(i, jj, k + 1), 'z', k, k + 1);
      }
  for (size_t e = 0; e < s.q.size(); ++e) {
    const int w = m.perf_well[e];
    const Well& wl = m.wells[e];
    const int id = m.idx(wl.i, wl.j, wl.k);
    double wi;
    if (!m.well_index((int)e, wi)) continue;
    const double conn = connection_rate<double>(m, wi, wl.head_offset, m.topo_wells[w].injector, wcell(s, id), s.bhp[w]);
    g_lnk[id] += lambda[3 * n + W + e] * (-conn);
  }
  return 0;
}

}  // extern "C"
