// Oracle: simulator model and the generic (value-or-dual) physics of the FIM
// residual. TEST INFRASTRUCTURE ONLY (see ad.hpp).
//
// Each function restates one reference generic, cited at its definition.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "ad.hpp"

namespace orc {

constexpr double DARCY_METRIC_FACTOR = 8.5269888e-3;  // rs/fim/assembly.rs:29
constexpr double GRAVITY = 9.80665;                   // rs/fim/flux.rs:130
constexpr double FB_EPSILON = 1e-6;                   // rs/fim/wells.rs:1582
constexpr double EPS = std::numeric_limits<double>::epsilon();

enum Regime : int { SATURATED = 0, UNDERSATURATED = 1 };

// ---------------------------------------------------------------- PVT table
// rs/pvt.rs:40-135 (rows, oil branches, saturated rows) and :227-425.
struct PvtRow {
  double p, rs, bo, mu_o, bg, mu_g;
};
struct PvtBranch {
  double rs;
  std::vector<PvtRow> rows;
};
constexpr double PVTO_RS_TOL = 1e-6;  // rs/pvt.rs:17

struct PvtTable {
  std::vector<PvtRow> rows, sat;
  std::vector<PvtBranch> branches;
  double c_o = 0.0;

  // rs/pvt.rs:86-135
  void build(const std::vector<PvtRow>& in, double c_o_) {
    rows = in;
    c_o = c_o_;
    branches.clear();
    for (const auto& r : in) {
      if (!branches.empty()) {
        auto& b = branches.back();
        bool same_rs = std::fabs(r.rs - b.rs) <= PVTO_RS_TOL;
        bool higher = b.rows.empty() ? true : (r.p >= b.rows.back().p - 1e-9);
        if (same_rs && higher) {
          b.rows.push_back(r);
          continue;
        }
      }
      branches.push_back({r.rs, {r}});
    }
    for (auto& b : branches)
      std::stable_sort(b.rows.begin(), b.rows.end(), [](const PvtRow& a, const PvtRow& c) { return a.p < c.p; });
    std::stable_sort(branches.begin(), branches.end(), [](const PvtBranch& a, const PvtBranch& c) { return a.rs < c.rs; });
    sat.clear();
    for (auto& b : branches) sat.push_back(b.rows.front());
    std::stable_sort(sat.begin(), sat.end(), [](const PvtRow& a, const PvtRow& c) { return a.p < c.p; });
  }

  template <class S>
  struct Sat {
    S rs, bo, bg, mu_o, mu_g;
  };

  // rs/pvt.rs:259-342 interpolate_saturated_generic
  template <class S>
  Sat<S> saturated(const S& p) const {
    if (sat.empty()) return {cst(p, 0.0), cst(p, 1.0), cst(p, 1.0), cst(p, 1.0), cst(p, 0.02)};
    double pv = val(p);
    if (pv <= sat[0].p) {
      const auto& r = sat[0];
      return {cst(p, r.rs), cst(p, r.bo), cst(p, r.bg), cst(p, r.mu_o), cst(p, r.mu_g)};
    }
    size_t last = sat.size() - 1;
    if (pv >= sat[last].p) {
      const auto& r = sat[last];
      S bo = exp_((p * (-c_o)) + (c_o * r.p)) * r.bo;
      S bg = cst(p, r.bg * r.p) / p;
      return {cst(p, r.rs), bo, bg, cst(p, r.mu_o), cst(p, r.mu_g)};
    }
    for (size_t i = 0; i < last; ++i) {
      const auto& r0 = sat[i];
      const auto& r1 = sat[i + 1];
      if (pv >= r0.p && pv <= r1.p) {
        double dp = r1.p - r0.p;
        if (dp < 1e-6) return {cst(p, r0.rs), cst(p, r0.bo), cst(p, r0.bg), cst(p, r0.mu_o), cst(p, r0.mu_g)};
        S t = (p - r0.p) / dp;
        auto lerp = [&](double a, double b) { return t * (b - a) + a; };
        S inv_bg = lerp(1.0 / r0.bg, 1.0 / r1.bg);
        S inv_bg_mu = lerp(1.0 / (r0.bg * r0.mu_g), 1.0 / (r1.bg * r1.mu_g));
        S inv_bo = lerp(1.0 / r0.bo, 1.0 / r1.bo);
        S inv_bo_mu = lerp(1.0 / (r0.bo * r0.mu_o), 1.0 / (r1.bo * r1.mu_o));
        return {lerp(r0.rs, r1.rs), recip(inv_bo), recip(inv_bg), inv_bo / inv_bo_mu, inv_bg / inv_bg_mu};
      }
    }
    const auto& r = sat[last];
    return {cst(p, r.rs), cst(p, r.bo), cst(p, r.bg), cst(p, r.mu_o), cst(p, r.mu_g)};
  }

  // rs/pvt.rs:344-382 branch_props_generic
  template <class S>
  void branch_props(const PvtBranch& b, const S& p, S& bo, S& mu) const {
    const auto& rows_ = b.rows;
    if (rows_.empty()) { bo = cst(p, 1.0); mu = cst(p, 1.0); return; }
    if (rows_.size() == 1) {
      const auto& r = rows_[0];
      S ex = p - r.p;
      bo = exp_(ex * (-c_o)) * r.bo;
      mu = exp_(ex * c_o) * r.mu_o;
      return;
    }
    double pv = val(p);
    if (pv <= rows_[0].p) { bo = cst(p, rows_[0].bo); mu = cst(p, rows_[0].mu_o); return; }
    for (size_t i = 0; i + 1 < rows_.size(); ++i) {
      const auto& r0 = rows_[i];
      const auto& r1 = rows_[i + 1];
      if (pv >= r0.p && pv <= r1.p) {
        double dp = r1.p - r0.p;
        if (std::fabs(dp) < 1e-9) { bo = cst(p, r0.bo); mu = cst(p, r0.mu_o); return; }
        S t = (p - r0.p) / dp;
        S inv_bo = t * (1.0 / r1.bo - 1.0 / r0.bo) + 1.0 / r0.bo;
        S inv_bo_mu = t * (1.0 / (r1.bo * r1.mu_o) - 1.0 / (r0.bo * r0.mu_o)) + 1.0 / (r0.bo * r0.mu_o);
        bo = recip(inv_bo);
        mu = inv_bo / inv_bo_mu;
        return;
      }
    }
    const auto& l = rows_.back();
    S ex = p - l.p;
    bo = max_floor(exp_(ex * (-c_o)) * l.bo, 1e-9);
    mu = exp_(ex * c_o) * l.mu_o;
  }

  // rs/pvt.rs:227-246 branch_bounds
  bool bounds(double rs, int& lo, int& hi) const {
    if (branches.empty()) return false;
    int n = (int)branches.size();
    if (rs <= branches[0].rs + PVTO_RS_TOL) { lo = hi = 0; return true; }
    if (rs >= branches[n - 1].rs - PVTO_RS_TOL) { lo = hi = n - 1; return true; }
    for (int i = 0; i + 1 < n; ++i) {
      if (rs >= branches[i].rs - PVTO_RS_TOL && rs <= branches[i + 1].rs + PVTO_RS_TOL) {
        lo = i; hi = i + 1; return true;
      }
    }
    lo = hi = n - 1;
    return true;
  }

  // rs/pvt.rs:384-425 interpolate_oil_generic -> (bo, mu_o)
  template <class S>
  void oil(const S& p, const S& rs, S& bo, S& mu) const {
    if (sat.empty()) { bo = cst(p, 1.0); mu = cst(p, 1.0); return; }
    auto s = saturated(p);
    double rs_sat = val(s.rs);
    if (val(rs) >= rs_sat - 1e-6) { bo = s.bo; mu = s.mu_o; return; }
    int lo, hi;
    if (bounds(val(rs), lo, hi)) {
      const auto& L = branches[lo];
      const auto& H = branches[hi];
      S bo_l, mu_l;
      branch_props(L, p, bo_l, mu_l);
      if (std::fabs(H.rs - L.rs) <= PVTO_RS_TOL) { bo = bo_l; mu = mu_l; return; }
      S t = (rs - L.rs) / (H.rs - L.rs);
      double shift = H.rows[0].p - L.rows[0].p;
      S p_lo = p - t * shift;
      S p_hi = p + (1.0 - t) * shift;
      S bo_lo, mu_lo, bo_hi, mu_hi;
      branch_props(L, p_lo, bo_lo, mu_lo);
      branch_props(H, p_hi, bo_hi, mu_hi);
      S inv_bo = t * (recip(bo_hi) - recip(bo_lo)) + recip(bo_lo);
      S inv_bo_mu = t * (recip(bo_hi * mu_hi) - recip(bo_lo * mu_lo)) + recip(bo_lo * mu_lo);
      bo = recip(inv_bo);
      mu = inv_bo / inv_bo_mu;
      return;
    }
    bo = s.bo;
    mu = s.mu_o;
  }

  // f64 row interpolation used by getters/classification: rs/pvt.rs:250-252
  double rs_sat_at(double p) const { return sat.empty() ? 0.0 : val(saturated(p).rs); }
};

// ---------------------------------------------------------------- SCAL
struct Scal2 {  // rs/relperm.rs:531-541, two-phase RockFluidProps
  double s_wc = 0.1, s_or = 0.1, n_w = 2.0, n_o = 2.0, k_rw_max = 1.0, k_ro_max = 1.0;
  double krw(double sw) const {  // f64 k_rw
    double denom = 1.0 - s_wc - s_or;
    double se = std::min(std::max((sw - s_wc) / denom, 0.0), 1.0);
    return std::pow(se, n_w) * k_rw_max;
  }
  double kro(double sw) const {
    double denom = 1.0 - s_wc - s_or;
    double se = std::min(std::max(((1.0 - s_or) - sw) / denom, 0.0), 1.0);
    return std::pow(se, n_o) * k_ro_max;
  }
  template <class S> S krw_g(const S& sw) const {
    double denom = 1.0 - s_wc - s_or;
    return powf_(min_ceil(max_floor((sw - s_wc) / denom, 0.0), 1.0), n_w) * k_rw_max;
  }
  template <class S> S kro_g(const S& sw) const {
    double denom = 1.0 - s_wc - s_or;
    return powf_(min_ceil(max_floor((cst(sw, 1.0 - s_or) - sw) / denom, 0.0), 1.0), n_o) * k_ro_max;
  }
  // rs/relperm.rs:602-635 corey_table_generic (points knots across [s_wc, 1-s_or])
  template <class S> void table(const S& sw, int points, S& krw_o, S& kro_o) const {
    points = std::max(points, 2);
    double lo = s_wc, hi = 1.0 - s_or;
    double span = std::max(hi - lo, 1e-12);
    double step = span / double(points - 1);
    auto knot = [&](int i, double& x, double& w, double& o) {
      x = std::min(lo + double(i) * step, hi);
      w = krw(x);
      o = kro(x);
    };
    double v = val(sw);
    double x, w, o;
    if (v <= lo) { knot(0, x, w, o); krw_o = cst(sw, w); kro_o = cst(sw, o); return; }
    if (v >= hi) { knot(points - 1, x, w, o); krw_o = cst(sw, w); kro_o = cst(sw, o); return; }
    int seg = std::min((int)std::floor((v - lo) / step), points - 2);
    double x0, w0, o0, x1, w1, o1;
    knot(seg, x0, w0, o0);
    knot(seg + 1, x1, w1, o1);
    S t = min_ceil(max_floor((sw - x0) / (x1 - x0), 0.0), 1.0);
    krw_o = t * (w1 - w0) + w0;
    kro_o = t * (o1 - o0) + o0;
  }
};

// rs/relperm.rs:8-39 interpolate_piecewise_generic over rows of (x, y1, y2): the segment is
// chosen on the value (the left one at a knot), constant beyond either end, and a segment
// of zero width gives its right value
template <class S>
S table_kr(const std::vector<std::array<double, 3>>& rows, const S& x, int col) {
  if (rows.empty()) return cst(x, 0.0);
  const double xv = val(x);
  if (xv <= rows[0][0]) return cst(x, rows[0][col]);
  for (size_t i = 0; i + 1 < rows.size(); ++i) {
    const double x0 = rows[i][0], x1 = rows[i + 1][0];
    if (xv <= x1) {
      if (std::fabs(x1 - x0) <= EPS) return cst(x, rows[i + 1][col]);
      S f = min_ceil(max_floor((x - x0) / (x1 - x0), 0.0), 1.0);
      return f * (rows[i + 1][col] - rows[i][col]) + rows[i][col];
    }
  }
  return cst(x, rows.back()[col]);
}

struct Scal3 {  // rs/relperm.rs:391-472 three-phase generics, with the optional SWOF/SGOF tables
  double s_wc, s_or, s_gc, s_gr, s_org, n_w, n_o, n_g, k_rw_max, k_ro_max, k_rg_max;
  // ThreePhaseScalTables (rs/relperm.rs:59-80): rows (sw, krw, krow) and (sg, krg, krog);
  // the capillary columns of the tables are not read by the reference's relperm
  std::vector<std::array<double, 3>> swof, sgof;
  bool tables() const { return !swof.empty(); }
  template <class S> S krw(const S& sw) const {
    if (tables()) return table_kr(swof, sw, 1);
    double denom = 1.0 - s_wc - s_or;
    return powf_(min_ceil(max_floor((sw - s_wc) / denom, 0.0), 1.0), n_w) * k_rw_max;
  }
  template <class S> S krg(const S& sg) const {
    if (tables()) return table_kr(sgof, sg, 1);
    double denom = 1.0 - s_wc - s_gc - s_gr;
    return powf_(min_ceil(max_floor((sg - s_gc) / denom, 0.0), 1.0), n_g) * k_rg_max;
  }
  template <class S> S kro_w(const S& sw) const {
    if (tables()) return table_kr(swof, sw, 2);
    double denom = 1.0 - s_wc - s_or;
    return powf_(min_ceil(max_floor((cst(sw, 1.0 - s_or) - sw) / denom, 0.0), 1.0), n_o) * k_ro_max;
  }
  template <class S> S kro_g(const S& sg) const {
    if (tables()) return table_kr(sgof, sg, 2);
    double denom = 1.0 - s_wc - s_org;
    return powf_(min_ceil(max_floor((cst(sg, 1.0 - s_wc - s_org) - sg) / denom, 0.0), 1.0), n_o) * k_ro_max;
  }
  // rs/relperm.rs:446-471 (zero derivative at or beyond either clamp bound)
  template <class S> S kro_stone2(const S& sw, const S& sg) const {
    double m = k_ro_max;
    S kw = kro_w(sw), kg = kro_g(sg), rw = krw(sw), rg = krg(sg);
    S v = ((kw / m + rw) * (kg / m + rg) - rw - rg) * m;
    if (val(v) <= 0.0 || val(v) >= m) return cst(v, std::min(std::max(val(v), 0.0), m));
    return v;
  }
};

struct CapBC {  // rs/capillary.rs:73-90 Brooks-Corey oil-water, cap 20*p_entry
  double p_entry = 5.0, lambda = 2.0;
  template <class S> S pc(const S& sw, const Scal2& rock, double scale) const {
    double denom = 1.0 - rock.s_wc - rock.s_or;
    S se = min_ceil(max_floor((sw - rock.s_wc) / denom, 0.0), 1.0);
    if (val(se) >= 1.0) return cst(sw, 0.0);
    double pe = p_entry * scale;  // Leverett-J per-cell scale (extension; scale=1 is the reference)
    double pcmax = pe * 20.0;
    if (val(se) <= 0.0) return cst(sw, pcmax);
    S v = powf_(se, -1.0 / lambda) * pe;
    return min_ceil(max_floor(v, 0.0), pcmax);
  }
};

struct CapOG {  // rs/capillary.rs:151-171
  double p_entry = 0.0, lambda = 2.0;
  template <class S> S pc(const S& sg, const Scal3& rock) const {
    double denom = 1.0 - rock.s_wc - rock.s_org;
    if (denom <= 0.0) return cst(sg, p_entry * 20.0);
    S so = cst(sg, 1.0 - rock.s_wc) - sg;
    S se = min_ceil(max_floor((so - rock.s_org) / denom, 0.0), 1.0);
    double pcmax = p_entry * 20.0;
    if (val(se) <= 0.0) return cst(sg, pcmax);
    S v = powf_(se, -1.0 / lambda) * p_entry;
    return min_ceil(max_floor(v, 0.0), pcmax);
  }
};

// ---------------------------------------------------------------- wells
struct WellSchedule {
  bool has = false;
  bool rate = false;
  bool has_target = false, has_surface = false, has_limit = false;
  double target = 0.0, surface = 0.0, limit = 0.0;
  bool enabled = true;
};

struct Well {  // rs/well.rs:17-29
  int i, j, k;
  double bhp, rw, skin;
  bool injector;
  std::string id;  // empty = legacy fingerprint grouping
  WellSchedule schedule;
  bool has_datum = false, has_density = false;
  double datum = 0.0, density = 0.0;
  double head_offset = 0.0;
  int fluid = -1;  // per-well injected fluid (extension): -1 inherit, 0 water, 1 gas
};

struct PhysWell {
  int rep;  // representative completion index
  bool injector;
  std::vector<int> perfs;
};

struct Control {  // rs/fim/wells.rs:626-634
  bool enabled, rate_controlled, uses_surface;
  bool has_target;
  double target, bhp_limit, bhp_target;
};

// ---------------------------------------------------------------- model
struct Model {
  int nx, ny, nz;
  double dx = 10.0, dy = 10.0;
  std::vector<double> dz;
  double depth_ref = 0.0;
  std::vector<double> phi, kx, ky, kz, pc_scale;
  // FluidProperties::default_pvt, rs/lib.rs:67-78
  double mu_o = 1.0, mu_w = 0.5, c_o = 1e-5, c_w = 3e-6, rho_o = 800.0, rho_w = 1000.0;
  double b_o = 1.0, b_w = 1.0, p_ref_w = 300.0, c_r = 0.0, p_ref_rock = 300.0;
  double mu_g = 0.02, c_g = 1e-4, rho_g = 10.0;
  Scal2 scal;
  bool has_scal3 = false;
  Scal3 scal3{};
  CapBC pc;
  bool has_pc_og = false;
  CapOG pc_og;
  bool gravity = false, three_phase = false, has_table = false;
  PvtTable table;
  int corey_points = 21;  // rs/lib.rs:90
  int injected_fluid = 1;  // 0 water, 1 gas (rs/frontend.rs:114 default Gas)
  bool inj_rate = false, prod_rate = false, inj_enabled = true;
  double tgt_inj = 0.0, tgt_prod = 0.0;
  bool has_tgt_inj_surf = false, has_tgt_prod_surf = false;
  double tgt_inj_surf = 0.0, tgt_prod_surf = 0.0;
  double bhp_min = -100.0, bhp_max = 2000.0;
  std::vector<Well> wells;
  // dynamic state (rs/lib.rs:92-217)
  std::vector<double> pressure, sw, so, sg, rs;
  // rs/frontend.rs:1208 set_gas_redissolution_enabled (default true, rs/frontend.rs:119).
  // Disabled (DRSDT 0): a cell's Rs may not rise above its stored value of the substep.
  bool gas_redissolution = true;
  double rs_cap_of(int id) const { return gas_redissolution ? INFINITY : rs[id]; }
  double time = 0.0;

  int n() const { return nx * ny * nz; }
  int idx(int i, int j, int k) const { return k * nx * ny + j * nx + i; }  // rs/grid.rs:252-254
  double dz_at(int id) const { return dz[id / (nx * ny)]; }
  double pore_volume(int id) const { return dx * dy * dz_at(id) * phi[id]; }  // rs/grid.rs:17-19
  double depth_at_k(int k) const {  // rs/grid.rs:256-263
    double d = depth_ref;
    for (int l = 0; l < k; ++l) d += dz[l];
    return d + dz[k] * 0.5;
  }
  // rs/grid.rs:278-311
  double geom_trans(int a, int b, char dim) const {
    int ka = a / (nx * ny), kb = b / (nx * ny);
    double dz1 = dz[ka], dz2 = dz[kb];
    auto hm = [](double x, double y) { return (x + y == 0.0) ? 0.0 : 2.0 * x * y / (x + y); };
    if (dim == 'x') return hm(kx[a], kx[b]) * (dy * dz1) / dx;
    if (dim == 'y') return hm(ky[a], ky[b]) * (dx * dz1) / dy;
    double denom = kz[b] * dz1 + kz[a] * dz2;
    if (denom <= 0.0) return 0.0;
    return 2.0 * kz[a] * kz[b] * (dx * dy) / denom;
  }

  // rs/pvt.rs:603-606
  template <class S> S inv_bw(const S& p) const {
    S x = (p - p_ref_w) * c_w;
    return (1.0 + x * (1.0 + x * 0.5)) / std::max(b_w, 1e-9);
  }
  double water_fvf(double p) const { return 1.0 / std::max(inv_bw(p), 1e-9); }
  template <class S> S base_bo(const S& p) const {  // rs/fim/properties.rs:135-138
    return max_floor(exp_(p * (-c_o)) * b_o, 1e-9);
  }
  template <class S> S rho_w_g(const S& p) const { return inv_bw(p) * rho_w; }
  template <class S> S rho_o_g(const S& p, const S& rs) const {  // rs/pvt.rs:988-996
    if (has_table) {
      S bo, mu;
      table.oil(p, rs, bo, mu);
      return (rs * rho_g + rho_o) / max_floor(bo, 1e-9);
    }
    return cst(p, rho_o) / base_bo(p);
  }
  template <class S> S rho_g_g(const S& p) const {  // rs/pvt.rs:1000-1008
    if (has_table) return cst(p, rho_g) / max_floor(table.saturated(p).bg, 1e-9);
    return cst(p, rho_g);
  }
  template <class S> S mu_o_g(const S& p, const S& rs) const {  // rs/pvt.rs:674-684
    if (has_table) {
      if (three_phase) {
        S bo, mu;
        table.oil(p, rs, bo, mu);
        return mu;
      }
      return table.saturated(p).mu_o;
    }
    return cst(p, mu_o);
  }
  template <class S> S mu_g_g(const S& p) const {
    if (has_table) return table.saturated(p).mu_g;
    return cst(p, mu_g);
  }

  template <class S> struct Mob { S w, o, g; };
  // rs/mobility.rs:196-249 phase_mobilities_for_state_generic
  template <class S> Mob<S> mobilities(const S& sw_, const S& sg_, const S& p, const S& rs_) const {
    S muw = cst(p, mu_w);
    if (three_phase) {
      if (!has_scal3) {
        return {scal.krw_g(sw_) / muw, scal.kro_g(sw_) / mu_o_g(p, rs_), cst(p, 0.0)};
      }
      return {scal3.krw(sw_) / muw, scal3.kro_stone2(sw_, sg_) / mu_o_g(p, rs_), scal3.krg(sg_) / mu_g_g(p)};
    }
    S krw, kro;
    if (corey_points > 0) {
      scal.table(sw_, corey_points, krw, kro);
    } else {
      krw = scal.krw_g(sw_);
      kro = scal.kro_g(sw_);
    }
    return {krw / muw, kro / mu_o_g(p, rs_), cst(p, 0.0)};
  }

  template <class S> struct Props { S so, sg, rs, bo, bg; };

  // rs/fim/flash_ad.rs:19-59 (implicit-function derivative through the bisection)
  template <class S>
  S solve_rs_for_dissolved(const S& p, const S& sw_, const S& sg_, const S& pv, const S& dissolved, double rs_upper) const {
    double rs_star;
    {  // rs/step.rs:66-104 bisection
      double target = std::max(val(dissolved), 0.0);
      double so_ = std::max(1.0 - val(sw_) - val(sg_), 0.0);
      if (target <= 0.0 || rs_upper <= 0.0 || so_ <= 1e-12 || val(pv) <= 0.0) {
        rs_star = 0.0;
      } else {
        double lo = 0.0, hi = std::max(rs_upper, 0.0);
        for (int it = 0; it < 64; ++it) {
          double mid = 0.5 * (lo + hi);
          double bo, mu;
          table.oil(val(p), mid, bo, mu);
          double dm = (so_ * val(pv) / std::max(bo, 1e-9)) * mid;
          if (dm < target) lo = mid; else hi = mid;
        }
        rs_star = 0.5 * (lo + hi);
      }
    }
    S so_g = max_floor(1.0 - sw_ - sg_, 0.0);
    S target = max_floor(dissolved, 0.0);
    double dgdrs;
    {
      Ad<1> rl = Ad<1>::var(rs_star, 0);
      Ad<1> bo, mu;
      table.oil(Ad<1>(val(p)), rl, bo, mu);
      Ad<1> g = (Ad<1>(val(so_g)) * val(pv) / max_floor(bo, 1e-9)) * rl;
      dgdrs = g.d[0];
    }
    if (std::fabs(dgdrs) < 1e-14) return cst(p, rs_star);
    S bo, mu;
    table.oil(p, cst(p, rs_star), bo, mu);
    S g_theta = (so_g * pv / max_floor(bo, 1e-9)) * rs_star - target;
    return cst(p, rs_star) - g_theta / dgdrs;
  }

  // rs/fim/flash_ad.rs:104-207. rs_cap finite: gas redissolution disabled, Rs held at or
  // below min(rs_cap, Rs_sat(p)) and the free gas at or above its transported amount.
  template <class S>
  void split_gas(const S& p, const S& pv, const S& sw_, const S& free_sc, const S& dissolved_in, S& sg_o, S& so_o, S& rs_o,
                 double rs_cap = INFINITY) const {
    const bool capped = std::isfinite(rs_cap);
    S total_hc = max_floor(1.0 - sw_, 0.0);
    S bg = max_floor(table.saturated(p).bg, 1e-9);
    S free_t = max_floor(free_sc, 0.0);
    S sg_t = min_of(max_floor((free_t * bg) / max_floor(pv, 1e-9), 0.0), total_hc);
    S so_t = max_floor(total_hc - sg_t, 0.0);
    S dissolved = max_floor(dissolved_in, 0.0);
    S rs_max = max_floor(table.saturated(p).rs, 0.0);
    S rs_sat = rs_max, bo_sat, mu_sat;
    if (capped) {
      rs_sat = min_of(cst(p, std::max(rs_cap, 0.0)), rs_max);
      table.oil(p, rs_sat, bo_sat, mu_sat);
      bo_sat = max_floor(bo_sat, 1e-9);
      S max_t = (so_t * pv / bo_sat) * rs_sat;
      if (val(dissolved) <= val(max_t) + 1e-9) {
        rs_o = solve_rs_for_dissolved(p, sw_, sg_t, pv, dissolved, val(rs_sat));
        sg_o = sg_t;
        so_o = so_t;
        return;
      }
    } else {
      table.oil(p, rs_max, bo_sat, mu_sat);
      bo_sat = max_floor(bo_sat, 1e-9);
    }
    S total_gas = free_t + dissolved;
    S max_all = (total_hc * pv / bo_sat) * rs_sat;
    if (!capped && val(total_gas) <= val(max_all) + 1e-9) {
      rs_o = solve_rs_for_dissolved(p, sw_, cst(p, 0.0), pv, total_gas, val(rs_sat));
      sg_o = cst(p, 0.0);
      so_o = total_hc;
      return;
    }
    S denom = recip(bg) - (rs_sat / bo_sat);
    S sg_s = std::fabs(val(denom)) > 1e-12 ? ((total_gas / pv) - (total_hc * rs_sat / bo_sat)) / denom : sg_t;
    S sg = min_of(max_of(sg_s, capped ? sg_t : cst(p, 0.0)), total_hc);
    sg_o = sg;
    so_o = max_floor(total_hc - sg, 0.0);
    rs_o = rs_sat;
  }

  // rs/fim/properties.rs:37-133 cell_props_generic; rs_cap finite = the cell's stored Rs
  // when gas redissolution is disabled (the drsdt0 base of the reference)
  template <class S>
  Props<S> cell_props(int regime, const S& p, const S& sw_, const S& hc, double rs_cap = INFINITY) const {
    S raw_hc = 1.0 - sw_;
    if (!three_phase || !has_table) {
      S total = max_floor(raw_hc, 0.0);
      S sg_ = min_of(max_floor(hc, 0.0), total);
      S so_ = max_floor(1.0 - sw_ - sg_, 0.0);
      return {so_, sg_, cst(p, 0.0), base_bo(p), cst(p, 1.0)};
    }
    const bool capped = std::isfinite(rs_cap);
    if (regime == SATURATED) {
      S sg_ = hc;
      auto s = table.saturated(p);
      S rs_ = s.rs;
      if (capped) rs_ = min_ceil(rs_, rs_cap);
      S so_ = raw_hc - sg_;
      S bo, mu;
      table.oil(p, rs_, bo, mu);
      return {so_, sg_, rs_, bo, s.bg};
    }
    auto s = table.saturated(p);
    S rs_cap_v = max_floor(s.rs, 0.0);
    if (capped) rs_cap_v = min_of(rs_cap_v, cst(p, std::max(rs_cap, 0.0)));
    S rs_trial = hc;
    S sg_, so_, rs_;
    if (val(rs_trial) <= val(rs_cap_v) + 1e-6) {
      sg_ = cst(p, 0.0);
      so_ = raw_hc;
      rs_ = rs_trial;
    } else {
      S bo_t, mu_t;
      table.oil(p, rs_trial, bo_t, mu_t);
      bo_t = max_floor(bo_t, 1e-9);
      S dissolved = raw_hc * rs_trial / bo_t;
      split_gas(p, cst(p, 1.0), sw_, cst(p, 0.0), dissolved, sg_, so_, rs_, rs_cap);
    }
    S bo, mu;
    table.oil(p, rs_, bo, mu);
    return {so_, sg_, rs_, bo, table.saturated(p).bg};
  }

  // rs/fim/properties.rs:146-151
  template <class S> S pore_volume_g(int id, const S& p) const {
    return exp_((p - p_ref_rock) * c_r) * pore_volume(id);
  }

  // rs/fim/properties.rs:155-166
  template <class S>
  void inventory(const S& pv, const S& p, const S& sw_, const Props<S>& pr, S out[3]) const {
    out[0] = pv * sw_ * inv_bw(p);
    out[1] = pv * pr.so / max_floor(pr.bo, 1e-9);
    out[2] = pv * pr.sg / max_floor(pr.bg, 1e-9) + out[1] * pr.rs;
  }

  // rs/fim/properties.rs:171-212
  template <class S>
  void accumulation(int id, int regime, const S& p, const S& sw_, const S& hc, int pregime, double pp, double psw,
                    double phc, S out[3]) const {
    auto pr = cell_props(regime, p, sw_, hc, rs_cap_of(id));
    S pv = pore_volume_g(id, p);
    S cur[3];
    inventory(pv, p, sw_, pr, cur);
    auto pp_ = cell_props<double>(pregime, pp, psw, phc, rs_cap_of(id));
    double ppv = pore_volume_g<double>(id, pp);
    double prev[3];
    inventory<double>(ppv, pp, psw, pp_, prev);
    for (int c = 0; c < 3; ++c) out[c] = cur[c] - prev[c];
  }

  template <class S> S pc_og_g(const S& sg_) const {  // rs/fim/flux.rs:134-139
    if (has_pc_og && has_scal3) return pc_og.pc(sg_, scal3);
    return cst(sg_, 0.0);
  }

  // rs/fim/flux.rs:42-118 face_flux_terms_generic (sc units per day)
  template <class S>
  void face_flux(double geom_t, int id_i, int id_j, int reg_i, int reg_j, double depth_i, double depth_j, const S& pi,
                 const S& swi, const S& hci, const S& pj, const S& swj, const S& hcj, S flux[3]) const {
    auto pri = cell_props(reg_i, pi, swi, hci, rs_cap_of(id_i));
    auto prj = cell_props(reg_j, pj, swj, hcj, rs_cap_of(id_j));
    S rwi = rho_w_g(pi), rwj = rho_w_g(pj);
    S roi = rho_o_g(pi, pri.rs), roj = rho_o_g(pj, prj.rs);
    S rgi = rho_g_g(pi), rgj = rho_g_g(pj);
    S pcw_i = pc.pc(swi, scal, pc_scale[id_i]);
    S pcw_j = pc.pc(swj, scal, pc_scale[id_j]);
    S pcg_i = pc_og_g(pri.sg), pcg_j = pc_og_g(prj.sg);
    auto grav = [&](const S& a, const S& b) {
      if (!gravity) return cst(a, 0.0);
      return ((a + b) * 0.5) * (GRAVITY * (depth_i - depth_j) * 1e-5);
    };
    S gw = grav(rwi, rwj), go = grav(roi, roj), gg = grav(rgi, rgj);
    S dphi_w = (pi - pj) - (pcw_i - pcw_j) - gw;
    S dphi_o = (pi - pj) - go;
    S dphi_g = (pi - pj) + (pcg_i - pcg_j) - gg;
    auto mi = mobilities(swi, pri.sg, pi, pri.rs);
    auto mj = mobilities(swj, prj.sg, pj, prj.rs);
    bool uw = val(dphi_w) >= 0.0, uo = val(dphi_o) >= 0.0, ug = val(dphi_g) >= 0.0;
    S lw = uw ? mi.w : mj.w;
    S p_w = uw ? pi : pj;
    S lo = uo ? mi.o : mj.o;
    S bo_o = uo ? pri.bo : prj.bo;
    S rs_o = uo ? pri.rs : prj.rs;
    S lg = ug ? mi.g : mj.g;
    S bg_g = ug ? pri.bg : prj.bg;
    S qw = lw * dphi_w * geom_t * inv_bw(p_w);
    S qo_res = lo * dphi_o * geom_t;
    S qo = qo_res / max_floor(bo_o, 1e-9);
    S qg_free = lg * dphi_g * geom_t / max_floor(bg_g, 1e-9);
    S qg = qg_free + qo * rs_o;
    flux[0] = qw;
    flux[1] = qo;
    flux[2] = qg;
  }

  // ---- wells -------------------------------------------------------------
  std::vector<PhysWell> topo_wells;
  std::vector<int> perf_well;  // perforation -> physical well

  // rs/fim/wells.rs:544-594
  void build_topology() {
    topo_wells.clear();
    perf_well.assign(wells.size(), -1);
    std::vector<std::string> keys;
    for (size_t e = 0; e < wells.size(); ++e) {
      const Well& w = wells[e];
      std::string key;
      if (!w.id.empty()) {
        key = "id:" + w.id;
      } else {
        char buf[256];
        snprintf(buf, sizeof buf, "fp:%d:%d:%d:%a:%a:%a", int(w.injector), w.i, w.j, w.bhp, w.rw, w.skin);
        key = buf;
      }
      int found = -1;
      for (size_t g = 0; g < keys.size(); ++g)
        if (keys[g] == key) { found = (int)g; break; }
      if (found < 0) {
        found = (int)topo_wells.size();
        keys.push_back(key);
        topo_wells.push_back({(int)e, w.injector, {}});
      }
      perf_well[e] = found;
      topo_wells[found].perfs.push_back((int)e);
    }
  }

  int fluid_of(int well_idx) const {  // rs/fim/wells.rs:618-624 (+ per-well extension)
    if (!three_phase) return 0;
    const Well& rep = wells[topo_wells[well_idx].rep];
    return rep.fluid >= 0 ? rep.fluid : injected_fluid;
  }

  // rs/fim/wells.rs:691-798
  Control control(int well_idx) const {
    const Well& w = wells[topo_wells[well_idx].rep];
    bool inj = topo_wells[well_idx].injector;
    bool expl = w.schedule.has;
    Control c{};
    c.enabled = expl ? w.schedule.enabled : (inj ? inj_enabled : true);
    bool fam_rate = inj ? (inj_rate && inj_enabled) : prod_rate;
    c.rate_controlled = !c.enabled ? false : (expl ? w.schedule.rate : fam_rate);
    c.uses_surface = expl ? w.schedule.has_surface : (inj ? has_tgt_inj_surf : has_tgt_prod_surf);
    c.has_target = false;
    c.target = 0.0;
    if (c.rate_controlled) {
      if (expl) {
        double t = w.schedule.has_surface ? w.schedule.surface : (w.schedule.has_target ? w.schedule.target : 0.0);
        c.has_target = true;
        c.target = std::max(t, 0.0);
      } else if (fam_rate) {
        c.has_target = true;
        c.target = inj ? std::max(has_tgt_inj_surf ? tgt_inj_surf : tgt_inj, 0.0)
                       : std::max(has_tgt_prod_surf ? tgt_prod_surf : tgt_prod, 0.0);
      }
    }
    double fam_limit = inj ? bhp_max : bhp_min;
    c.bhp_limit = expl ? (w.schedule.has_limit ? w.schedule.limit : fam_limit) : fam_limit;
    c.bhp_target = w.bhp;
    return c;
  }

  // rs/fim/wells.rs:800-826 Peaceman WI (returns false when undefined)
  bool well_index(int e, double& wi) const {
    const Well& w = wells[e];
    int id = idx(w.i, w.j, w.k);
    double a = kx[id], b = ky[id];
    if (!std::isfinite(a) || !std::isfinite(b) || a <= 0.0 || b <= 0.0) return false;
    double r_eq = 0.28 * std::sqrt(std::sqrt(a / b) * dx * dx + std::sqrt(b / a) * dy * dy) /
                  (std::pow(a / b, 0.25) + std::pow(b / a, 0.25));
    if (!std::isfinite(r_eq) || r_eq <= w.rw) return false;
    double kavg = std::sqrt(a * b);
    double denom = std::log(r_eq / w.rw) + w.skin;
    if (!std::isfinite(kavg) || kavg <= 0.0 || !std::isfinite(denom) || std::fabs(denom) <= EPS) return false;
    wi = DARCY_METRIC_FACTOR * 2.0 * M_PI * kavg * dz_at(id) / denom;
    return true;
  }

  // rs/well_control.rs:192-280 head offsets, refreshed once per report step
  void refresh_head_offsets() {
    if (!gravity) {
      for (auto& w : wells) w.head_offset = 0.0;
      return;
    }
    std::vector<double> off(wells.size(), 0.0);
    for (size_t e = 0; e < wells.size(); ++e) {
      int g = perf_well[e];
      const auto& grp = topo_wells[g].perfs;
      double datum = std::numeric_limits<double>::infinity();
      bool have_datum = false;
      for (int m : grp)
        if (wells[m].has_datum && std::isfinite(wells[m].datum)) { datum = wells[m].datum; have_datum = true; break; }
      if (!have_datum)
        for (int m : grp) datum = std::min(datum, depth_at_k(wells[m].k));
      double dens = -1.0;
      for (int m : grp)
        if (wells[m].has_density && std::isfinite(wells[m].density) && wells[m].density >= 0.0) { dens = wells[m].density; break; }
      if (dens < 0.0) {
        double tot = 0.0, cnt = 0.0;
        for (int m : grp) {
          const Well& w = wells[m];
          int id = idx(w.i, w.j, w.k);
          double p = pressure[id];
          double d;
          if (w.injector) {
            int fl = three_phase ? (w.fluid >= 0 ? w.fluid : injected_fluid) : 0;
            d = fl == 1 ? rho_g_g(p) : rho_w_g(p);
          } else {
            double fw, fo, fg;
            producer_fracs_state(id, fw, fo, fg);
            d = fw * rho_w_g(p) + fo * rho_o_g(p, std::max(rs[id], 0.0)) + fg * rho_g_g(p);
          }
          if (std::isfinite(d) && d > 0.0) { tot += d; cnt += 1.0; }
        }
        dens = cnt > 0.0 ? tot / cnt : 0.0;
      }
      double o = dens * GRAVITY * (depth_at_k(wells[e].k) - datum) * 1e-5;
      off[e] = std::isfinite(o) ? o : 0.0;
    }
    for (size_t e = 0; e < wells.size(); ++e) wells[e].head_offset = off[e];
  }

  // Producer phase fractions from the stored simulator state of one cell (used only
  // for the wellbore density of a producer head offset).
  void producer_fracs_state(int id, double& fw, double& fo, double& fg) const {
    double p = pressure[id];
    double sg_ = three_phase ? sg[id] : 0.0;
    double rs_ = three_phase ? rs[id] : 0.0;
    auto m = mobilities<double>(sw[id], sg_, p, rs_);
    double lw = std::max(m.w, 0.0), lo = std::max(m.o, 0.0), lg = std::max(m.g, 0.0);
    double lt = std::max(lw + lo + lg, EPS);
    fw = std::min(std::max(lw / lt, 0.0), 1.0);
    fo = std::min(std::max(lo / lt, 0.0), 1.0);
    fg = std::min(std::max(lg / lt, 0.0), 1.0);
  }
};

}  // namespace orc
