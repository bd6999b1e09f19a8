// Black-oil property model of the FIM residual, evaluated per cell on the GPU
// (and on the host for the handful of perforated cells the well logic needs).
//
// Every function is generic over T = double or Dual<N> and restates one
// reference routine, cited at its definition. Tables live in a small device
// (or host) buffer addressed by the pointers in PhysParams.
#pragma once
#include "rsf_dual.hpp"

namespace rsf {

constexpr double kDarcy = 8.5269888e-3;  // rs/fim/assembly.rs:29 mD*m2/(m*cP) -> m3/d/bar
constexpr double kGravity = 9.80665;     // rs/fim/flux.rs:130
constexpr double kEps = 2.220446049250313e-16;
constexpr double kPvtoRsTol = 1e-6;      // rs/pvt.rs:17

enum { kSaturated = 0, kUndersaturated = 1 };

struct Scal2P {  // two-phase Corey (rs/relperm.rs:531-541)
  double s_wc, s_or, n_w, n_o, k_rw_max, k_ro_max;
};
struct Scal3P {  // three-phase Corey + Stone II (rs/relperm.rs:391-472)
  double s_wc, s_or, s_gc, s_gr, s_org, n_w, n_o, n_g, k_rw_max, k_ro_max, k_rg_max;
};

struct PhysParams {
  int three_phase, has_table, has_scal3, has_pc_og, gravity, corey_points;
  double mu_o, mu_w, c_o, c_w, rho_o, rho_w, b_o, b_w, p_ref_w, c_r, p_ref_rock, mu_g, rho_g;
  Scal2P scal;
  Scal3P scal3;
  double pc_entry, pc_lambda, pcog_entry, pcog_lambda;
  // PVT table: saturated rows (p, rs, bo, mu_o, bg, mu_g) sorted by p; oil
  // branches sorted by rs, each a run of (p, bo, mu_o) rows sorted by p.
  int n_sat, n_br;
  double tbl_c_o;
  const double* sat;
  const double* br_rs;
  const int* br_off;
  const int* br_len;
  const double* br_rows;
  // gas redissolution disabled (rs/frontend.rs:1208): the stored Rs per cell (device),
  // the cap of each cell's Rs within a substep; null when redissolution is enabled
  const double* rs_base;
  // SWOF / SGOF tables of the three-phase props (rs/relperm.rs:59-80): rows (sw, krw, krow)
  // and (sg, krg, krog); n_swof = 0 without tables
  int n_swof, n_sgof;
  const double* swof;
  const double* sgof;
};

// the Rs cap of cell i (infinite: redissolution enabled); device code only
RSF_HD double rs_cap_at(const PhysParams& P, int i) { return P.rs_base ? P.rs_base[i] : INFINITY; }

template <class T>
struct SatProps {
  T rs, bo, bg, mu_o, mu_g;
};

// rs/pvt.rs:259-342 saturated-curve interpolation (segment chosen on the value)
template <class T>
RSF_HD SatProps<T> pvt_saturated(const PhysParams& P, const T& p) {
  SatProps<T> o;
  if (P.n_sat == 0) {
    o.rs = lift(p, 0.0); o.bo = lift(p, 1.0); o.bg = lift(p, 1.0); o.mu_o = lift(p, 1.0); o.mu_g = lift(p, 0.02);
    return o;
  }
  const double* s = P.sat;
  double pv = value(p);
  auto row = [&](int r) {
    o.rs = lift(p, s[6 * r + 1]); o.bo = lift(p, s[6 * r + 2]); o.mu_o = lift(p, s[6 * r + 3]);
    o.bg = lift(p, s[6 * r + 4]); o.mu_g = lift(p, s[6 * r + 5]);
  };
  if (pv <= s[0]) { row(0); return o; }
  int last = P.n_sat - 1;
  if (pv >= s[6 * last]) {
    const double* r = s + 6 * last;
    o.rs = lift(p, r[1]);
    o.bo = ex(p * (-P.tbl_c_o) + P.tbl_c_o * r[0]) * r[2];
    o.bg = lift(p, r[4] * r[0]) / p;
    o.mu_o = lift(p, r[3]);
    o.mu_g = lift(p, r[5]);
    return o;
  }
  for (int i = 0; i < last; ++i) {
    const double* a = s + 6 * i;
    const double* b = a + 6;
    if (pv >= a[0] && pv <= b[0]) {
      double dp = b[0] - a[0];
      if (dp < 1e-6) { row(i); return o; }
      T t = (p - a[0]) / dp;
      T ibg = t * (1.0 / b[4] - 1.0 / a[4]) + 1.0 / a[4];
      T ibgm = t * (1.0 / (b[4] * b[5]) - 1.0 / (a[4] * a[5])) + 1.0 / (a[4] * a[5]);
      T ibo = t * (1.0 / b[2] - 1.0 / a[2]) + 1.0 / a[2];
      T ibom = t * (1.0 / (b[2] * b[3]) - 1.0 / (a[2] * a[3])) + 1.0 / (a[2] * a[3]);
      o.rs = t * (b[1] - a[1]) + a[1];
      o.bo = inv(ibo);
      o.bg = inv(ibg);
      o.mu_o = ibo / ibom;
      o.mu_g = ibg / ibgm;
      return o;
    }
  }
  row(last);
  return o;
}

// rs/pvt.rs:344-382 one undersaturated oil branch at pressure p
template <class T>
RSF_HD void pvt_branch(const PhysParams& P, int b, const T& p, T& bo, T& mu) {
  int len = P.br_len[b];
  const double* r = P.br_rows + 3 * P.br_off[b];
  if (len == 0) { bo = lift(p, 1.0); mu = lift(p, 1.0); return; }
  if (len == 1) {
    T e = p - r[0];
    bo = ex(e * (-P.tbl_c_o)) * r[1];
    mu = ex(e * P.tbl_c_o) * r[2];
    return;
  }
  double pv = value(p);
  if (pv <= r[0]) { bo = lift(p, r[1]); mu = lift(p, r[2]); return; }
  for (int i = 0; i + 1 < len; ++i) {
    const double* a = r + 3 * i;
    const double* c = a + 3;
    if (pv >= a[0] && pv <= c[0]) {
      double dp = c[0] - a[0];
      if (fabs(dp) < 1e-9) { bo = lift(p, a[1]); mu = lift(p, a[2]); return; }
      T t = (p - a[0]) / dp;
      T ibo = t * (1.0 / c[1] - 1.0 / a[1]) + 1.0 / a[1];
      T ibom = t * (1.0 / (c[1] * c[2]) - 1.0 / (a[1] * a[2])) + 1.0 / (a[1] * a[2]);
      bo = inv(ibo);
      mu = ibo / ibom;
      return;
    }
  }
  const double* l = r + 3 * (len - 1);
  T e = p - l[0];
  bo = floor_at(ex(e * (-P.tbl_c_o)) * l[1], 1e-9);
  mu = ex(e * P.tbl_c_o) * l[2];
}

// rs/pvt.rs:384-425 oil FVF and viscosity at (p, Rs)
template <class T>
RSF_HD void pvt_oil(const PhysParams& P, const T& p, const T& rs, T& bo, T& mu) {
  if (P.n_sat == 0) { bo = lift(p, 1.0); mu = lift(p, 1.0); return; }
  SatProps<T> s = pvt_saturated(P, p);
  double rsv = value(rs);
  if (rsv >= value(s.rs) - 1e-6) { bo = s.bo; mu = s.mu_o; return; }
  int nb = P.n_br;
  if (nb == 0) { bo = s.bo; mu = s.mu_o; return; }
  int lo = nb - 1, hi = nb - 1;
  if (rsv <= P.br_rs[0] + kPvtoRsTol) {
    lo = hi = 0;
  } else if (rsv >= P.br_rs[nb - 1] - kPvtoRsTol) {
    lo = hi = nb - 1;
  } else {
    for (int i = 0; i + 1 < nb; ++i)
      if (rsv >= P.br_rs[i] - kPvtoRsTol && rsv <= P.br_rs[i + 1] + kPvtoRsTol) { lo = i; hi = i + 1; break; }
  }
  T bl, ml;
  pvt_branch(P, lo, p, bl, ml);
  double rlo = P.br_rs[lo], rhi = P.br_rs[hi];
  if (fabs(rhi - rlo) <= kPvtoRsTol) { bo = bl; mu = ml; return; }
  T t = (rs - rlo) / (rhi - rlo);
  double shift = P.br_rows[3 * P.br_off[hi]] - P.br_rows[3 * P.br_off[lo]];
  T plo = p - t * shift;
  T phi = p + (1.0 - t) * shift;
  T b0, m0, b1, m1;
  pvt_branch(P, lo, plo, b0, m0);
  pvt_branch(P, hi, phi, b1, m1);
  T ib = t * (inv(b1) - inv(b0)) + inv(b0);
  T ibm = t * (inv(b1 * m1) - inv(b0 * m0)) + inv(b0 * m0);
  bo = inv(ib);
  mu = ib / ibm;
}

// rs/pvt.rs:603-606
template <class T>
RSF_HD T inv_bw(const PhysParams& P, const T& p) {
  T x = (p - P.p_ref_w) * P.c_w;
  return (1.0 + x * (1.0 + x * 0.5)) / (P.b_w > 1e-9 ? P.b_w : 1e-9);
}

template <class T>
RSF_HD T base_bo(const PhysParams& P, const T& p) {  // rs/fim/properties.rs:135-138
  return floor_at(ex(p * (-P.c_o)) * P.b_o, 1e-9);
}

// ---------------------------------------------------------------- relperm
template <class T>
RSF_HD T clamp01(const T& x) { return ceil_at(floor_at(x, 0.0), 1.0); }

RSF_HD double corey_krw(const Scal2P& s, double sw) {
  double se = (sw - s.s_wc) / (1.0 - s.s_wc - s.s_or);
  se = se < 0.0 ? 0.0 : (se > 1.0 ? 1.0 : se);
  return pow(se, s.n_w) * s.k_rw_max;
}
RSF_HD double corey_kro(const Scal2P& s, double sw) {
  double se = ((1.0 - s.s_or) - sw) / (1.0 - s.s_wc - s.s_or);
  se = se < 0.0 ? 0.0 : (se > 1.0 ? 1.0 : se);
  return pow(se, s.n_o) * s.k_ro_max;
}

// rs/relperm.rs:602-635: Corey sampled at `points` equal knots, linear between
template <class T>
RSF_HD void corey_table(const Scal2P& s, int points, const T& sw, T& krw, T& kro) {
  points = points < 2 ? 2 : points;
  double lo = s.s_wc, hi = 1.0 - s.s_or;
  double span = hi - lo > 1e-12 ? hi - lo : 1e-12;
  double step = span / double(points - 1);
  double v = value(sw);
  if (v <= lo) { krw = lift(sw, corey_krw(s, lo)); kro = lift(sw, corey_kro(s, lo)); return; }
  if (v >= hi) {
    double x = fmin(lo + double(points - 1) * step, hi);
    krw = lift(sw, corey_krw(s, x));
    kro = lift(sw, corey_kro(s, x));
    return;
  }
  int seg = (int)floor((v - lo) / step);
  if (seg > points - 2) seg = points - 2;
  double x0 = fmin(lo + double(seg) * step, hi), x1 = fmin(lo + double(seg + 1) * step, hi);
  double w0 = corey_krw(s, x0), w1 = corey_krw(s, x1), o0 = corey_kro(s, x0), o1 = corey_kro(s, x1);
  T t = clamp01((sw - x0) / (x1 - x0));
  krw = t * (w1 - w0) + w0;
  kro = t * (o1 - o0) + o0;
}

// rs/relperm.rs:8-39 interpolate_piecewise_generic over rows of 3 (x, y1, y2): the segment
// is chosen on the value (the left one at a knot), constant beyond either end, a segment of
// zero width gives its right value
template <class T>
RSF_HD T table_kr(const double* rows, int n, const T& x, int col) {
  if (n <= 0) return lift(x, 0.0);
  const double xv = value(x);
  if (xv <= rows[0]) return lift(x, rows[col]);
  for (int i = 0; i + 1 < n; ++i) {
    const double x0 = rows[3 * i], x1 = rows[3 * i + 3];
    if (xv <= x1) {
      if (fabs(x1 - x0) <= kEps) return lift(x, rows[3 * i + 3 + col]);
      const double y0 = rows[3 * i + col], y1 = rows[3 * i + 3 + col];
      return clamp01((x - x0) / (x1 - x0)) * (y1 - y0) + y0;
    }
  }
  return lift(x, rows[3 * (n - 1) + col]);
}

template <class T>
RSF_HD T s3_krw(const Scal3P& s, const T& sw) {
  return pw(clamp01((sw - s.s_wc) / (1.0 - s.s_wc - s.s_or)), s.n_w) * s.k_rw_max;
}
template <class T>
RSF_HD T s3_krg(const Scal3P& s, const T& sg) {
  return pw(clamp01((sg - s.s_gc) / (1.0 - s.s_wc - s.s_gc - s.s_gr)), s.n_g) * s.k_rg_max;
}
template <class T>
RSF_HD T s3_krow(const Scal3P& s, const T& sw) {
  return pw(clamp01(((1.0 - s.s_or) - sw) / (1.0 - s.s_wc - s.s_or)), s.n_o) * s.k_ro_max;
}
template <class T>
RSF_HD T s3_krog(const Scal3P& s, const T& sg) {
  return pw(clamp01(((1.0 - s.s_wc - s.s_org) - sg) / (1.0 - s.s_wc - s.s_org)), s.n_o) * s.k_ro_max;
}
// the three-phase relperms with the SWOF/SGOF tables when set (rs/relperm.rs:391-444)
template <class T>
RSF_HD T p3_krw(const PhysParams& P, const T& sw) {
  return P.n_swof ? table_kr(P.swof, P.n_swof, sw, 1) : s3_krw(P.scal3, sw);
}
template <class T>
RSF_HD T p3_krg(const PhysParams& P, const T& sg) {
  return P.n_swof ? table_kr(P.sgof, P.n_sgof, sg, 1) : s3_krg(P.scal3, sg);
}
template <class T>
RSF_HD T p3_krow(const PhysParams& P, const T& sw) {
  return P.n_swof ? table_kr(P.swof, P.n_swof, sw, 2) : s3_krow(P.scal3, sw);
}
template <class T>
RSF_HD T p3_krog(const PhysParams& P, const T& sg) {
  return P.n_swof ? table_kr(P.sgof, P.n_sgof, sg, 2) : s3_krog(P.scal3, sg);
}

// rs/relperm.rs:446-471 Stone II, derivative frozen at or beyond [0, kro_max]
template <class T>
RSF_HD T stone2(const PhysParams& P, const T& sw, const T& sg) {
  double m = P.scal3.k_ro_max;
  T rw = p3_krw(P, sw), rg = p3_krg(P, sg);
  T v = ((p3_krow(P, sw) / m + rw) * (p3_krog(P, sg) / m + rg) - rw - rg) * m;
  double vv = value(v);
  if (vv <= 0.0 || vv >= m) return lift(v, vv < 0.0 ? 0.0 : (vv > m ? m : vv));
  return v;
}

// rs/capillary.rs:73-90 Brooks-Corey Pc_ow (per-cell Leverett-J scale; 1 = reference)
template <class T>
RSF_HD T pc_ow(const PhysParams& P, const T& sw, double scale) {
  T se = clamp01((sw - P.scal.s_wc) / (1.0 - P.scal.s_wc - P.scal.s_or));
  if (value(se) >= 1.0) return lift(sw, 0.0);
  double pe = P.pc_entry * scale;
  double pmax = pe * 20.0;
  if (value(se) <= 0.0) return lift(sw, pmax);
  return ceil_at(floor_at(pw(se, -1.0 / P.pc_lambda) * pe, 0.0), pmax);
}

// rs/capillary.rs:151-171 Pc_og (zero unless configured with three-phase SCAL)
template <class T>
RSF_HD T pc_og(const PhysParams& P, const T& sg) {
  if (!(P.has_pc_og && P.has_scal3)) return lift(sg, 0.0);
  double denom = 1.0 - P.scal3.s_wc - P.scal3.s_org;
  double pmax = P.pcog_entry * 20.0;
  if (denom <= 0.0) return lift(sg, pmax);
  T se = clamp01(((1.0 - P.scal3.s_wc) - sg - P.scal3.s_org) / denom);
  if (value(se) <= 0.0) return lift(sg, pmax);
  return ceil_at(floor_at(pw(se, -1.0 / P.pcog_lambda) * P.pcog_entry, 0.0), pmax);
}

template <class T>
RSF_HD T visc_oil(const PhysParams& P, const T& p, const T& rs) {  // rs/pvt.rs:674-684
  if (!P.has_table) return lift(p, P.mu_o);
  if (P.three_phase) {
    T bo, mu;
    pvt_oil(P, p, rs, bo, mu);
    return mu;
  }
  return pvt_saturated(P, p).mu_o;
}

template <class T>
struct Mobility {
  T w, o, g;
};

// rs/mobility.rs:196-249
template <class T>
RSF_HD Mobility<T> mobilities(const PhysParams& P, const T& sw, const T& sg, const T& p, const T& rs) {
  Mobility<T> m;
  if (P.three_phase) {
    if (!P.has_scal3) {
      Scal3P s{P.scal.s_wc, P.scal.s_or, 0, 0, 0, P.scal.n_w, P.scal.n_o, 1, P.scal.k_rw_max, P.scal.k_ro_max, 0};
      m.w = s3_krw(s, sw) / P.mu_w;
      m.o = s3_krow(s, sw) / visc_oil(P, p, rs);
      m.g = lift(p, 0.0);
      return m;
    }
    m.w = p3_krw(P, sw) / P.mu_w;
    m.o = stone2(P, sw, sg) / visc_oil(P, p, rs);
    m.g = p3_krg(P, sg) / (P.has_table ? pvt_saturated(P, p).mu_g : lift(p, P.mu_g));
    return m;
  }
  T krw, kro;
  if (P.corey_points > 0) {
    corey_table(P.scal, P.corey_points, sw, krw, kro);
  } else {
    Scal3P s{P.scal.s_wc, P.scal.s_or, 0, 0, 0, P.scal.n_w, P.scal.n_o, 1, P.scal.k_rw_max, P.scal.k_ro_max, 0};
    krw = s3_krw(s, sw);
    kro = s3_krow(s, sw);
  }
  m.w = krw / P.mu_w;
  m.o = kro / visc_oil(P, p, rs);
  m.g = lift(p, 0.0);
  return m;
}

template <class T>
struct Fluids {
  T so, sg, rs, bo, bg;
};

// rs/step.rs:66-104 bisection for Rs carrying `target` dissolved gas, with the
// implicit-function derivative of rs/fim/flash_ad.rs:19-59.
template <class T>
RSF_HD T rs_for_dissolved(const PhysParams& P, const T& p, const T& sw, const T& dissolved, double rs_upper) {
  double target = value(dissolved) > 0.0 ? value(dissolved) : 0.0;
  double so = 1.0 - value(sw);
  so = so > 0.0 ? so : 0.0;
  double star = 0.0;
  double pv = value(p);
  if (!(target <= 0.0 || rs_upper <= 0.0 || so <= 1e-12)) {
    double lo = 0.0, hi = rs_upper > 0.0 ? rs_upper : 0.0;
    for (int it = 0; it < 64; ++it) {
      double mid = 0.5 * (lo + hi);
      double b, m;
      pvt_oil<double>(P, pv, mid, b, m);
      double g = (so / (b > 1e-9 ? b : 1e-9)) * mid;
      if (g < target) lo = mid; else hi = mid;
    }
    star = 0.5 * (lo + hi);
  }
  // dg/drs at rs* with p, so frozen
  Dual<1> r1 = dvar<1>(star, 0), b1, m1;
  pvt_oil(P, dconst<1>(pv), r1, b1, m1);
  Dual<1> g1 = (dconst<1>(so) / floor_at(b1, 1e-9)) * r1;
  double dg = g1.d[0];
  if (fabs(dg) < 1e-14) return lift(p, star);
  T so_t = floor_at(1.0 - sw, 0.0);
  T b, m;
  pvt_oil(P, p, lift(p, star), b, m);
  T gth = (so_t / floor_at(b, 1e-9)) * star - floor_at(dissolved, 0.0);
  return lift(p, star) - gth / dg;
}

// rs/fim/properties.rs:37-133 cell fluids (frozen regime); the undersaturated overflow
// follows rs/fim/flash_ad.rs:104-207 with unit pore volume. A finite rs_cap (gas
// redissolution disabled: the cell's stored Rs) caps the saturated Rs at
// min(Rs_sat(p), rs_cap) in both regimes and in the overflow split.
template <class T>
RSF_HD Fluids<T> cell_fluids(const PhysParams& P, int regime, const T& p, const T& sw, const T& hc,
                             double rs_cap = INFINITY) {
  Fluids<T> f;
  T raw = 1.0 - sw;
  if (!P.three_phase || !P.has_table) {
    T tot = floor_at(raw, 0.0);
    f.sg = pick_min(floor_at(hc, 0.0), tot);
    f.so = floor_at(1.0 - sw - f.sg, 0.0);
    f.rs = lift(p, 0.0);
    f.bo = base_bo(P, p);
    f.bg = lift(p, 1.0);
    return f;
  }
  const bool capped = rs_cap < INFINITY;
  SatProps<T> s = pvt_saturated(P, p);
  if (regime == kSaturated) {
    f.sg = hc;
    f.rs = capped ? ceil_at(s.rs, rs_cap) : s.rs;
    f.so = raw - hc;
    T mu;
    pvt_oil(P, p, f.rs, f.bo, mu);
    f.bg = s.bg;
    return f;
  }
  T cap = floor_at(s.rs, 0.0);
  if (capped) cap = pick_min(lift(p, rs_cap > 0.0 ? rs_cap : 0.0), cap);
  if (value(hc) <= value(cap) + 1e-6) {
    f.sg = lift(p, 0.0);
    f.so = raw;
    f.rs = hc;
  } else {
    T bt, mt;
    pvt_oil(P, p, hc, bt, mt);
    T dissolved = raw * hc / floor_at(bt, 1e-9);
    T tot = floor_at(raw, 0.0);
    T bg = floor_at(s.bg, 1e-9);
    T bsat, msat;
    pvt_oil(P, p, cap, bsat, msat);
    bsat = floor_at(bsat, 1e-9);
    T gas = floor_at(dissolved, 0.0);
    T maxd = (tot / bsat) * cap;
    if (value(gas) <= value(maxd) + 1e-9) {
      f.rs = rs_for_dissolved(P, p, sw, gas, value(cap));
      f.sg = lift(p, 0.0);
      f.so = tot;
    } else {
      T den = inv(bg) - cap / bsat;
      T sgs = fabs(value(den)) > 1e-12 ? (gas - tot * cap / bsat) / den : lift(p, 0.0);
      T sg = pick_min(pick_max(sgs, lift(p, 0.0)), tot);
      f.sg = sg;
      f.so = floor_at(tot - sg, 0.0);
      f.rs = cap;
    }
  }
  T mu;
  pvt_oil(P, p, f.rs, f.bo, mu);
  f.bg = s.bg;
  return f;
}

// rs/fim/properties.rs:146-166 component inventory [water, oil, gas] in sc m3
template <class T>
RSF_HD void inventory(const PhysParams& P, double pv_ref, int regime, const T& p, const T& sw, const T& hc,
                      T out[3], double rs_cap = INFINITY) {
  Fluids<T> f = cell_fluids(P, regime, p, sw, hc, rs_cap);
  T pv = ex((p - P.p_ref_rock) * P.c_r) * pv_ref;
  out[0] = pv * sw * inv_bw(P, p);
  out[1] = pv * f.so / floor_at(f.bo, 1e-9);
  out[2] = pv * f.sg / floor_at(f.bg, 1e-9) + out[1] * f.rs;
}

template <class T>
RSF_HD T rho_oil(const PhysParams& P, const T& p, const T& rs) {  // rs/pvt.rs:988-996
  if (P.has_table) {
    T bo, mu;
    pvt_oil(P, p, rs, bo, mu);
    return (rs * P.rho_g + P.rho_o) / floor_at(bo, 1e-9);
  }
  return lift(p, P.rho_o) / base_bo(P, p);
}
template <class T>
RSF_HD T rho_gas(const PhysParams& P, const T& p) {  // rs/pvt.rs:1000-1008
  if (P.has_table) return lift(p, P.rho_g) / floor_at(pvt_saturated(P, p).bg, 1e-9);
  return lift(p, P.rho_g);
}

// Everything a TPFA face needs from one cell (rs/fim/flux.rs:42-118), with
// partials against the cell's [p, sw, hc]: upwinded mobility/FVF products
// m_w = lambda_w/B_w, m_o = lambda_o/B_o, m_g = lambda_g/B_g, m_s = m_o*Rs,
// phase densities and capillary pressures.
template <class T>
struct FaceCell {
  T mw, mo, mg, ms, rw, ro, rg, pcw, pcg;
};

template <class T>
RSF_HD FaceCell<T> face_cell(const PhysParams& P, int regime, const T& p, const T& sw, const T& hc, double pc_scale,
                             double rs_cap = INFINITY) {
  FaceCell<T> c;
  Fluids<T> f = cell_fluids(P, regime, p, sw, hc, rs_cap);
  Mobility<T> m = mobilities(P, sw, f.sg, p, f.rs);
  c.mw = m.w * inv_bw(P, p);
  c.mo = m.o / floor_at(f.bo, 1e-9);
  c.mg = m.g / floor_at(f.bg, 1e-9);
  c.ms = c.mo * f.rs;
  c.rw = inv_bw(P, p) * P.rho_w;
  c.ro = rho_oil(P, p, f.rs);
  c.rg = rho_gas(P, p);
  c.pcw = pc_ow(P, sw, pc_scale);
  c.pcg = pc_og(P, f.sg);
  return c;
}

}  // namespace rsf
