// Oracle: the FIM linear stack (well Schur elimination, CPR-FGMRES with a
// quasi-IMPES pressure stage and block-ILU0). TEST INFRASTRUCTURE ONLY.
#pragma once
#include <algorithm>
#include <cmath>
#include <map>
#include <limits>
#include <vector>

#include "model.hpp"

// Loops whose iterations are independent run on OpenMP threads; every sum keeps its
// sequential order, so results do not depend on the thread count.
#ifdef _OPENMP
#define ORC_PARALLEL_FOR _Pragma("omp parallel for schedule(static)")
#else
#define ORC_PARALLEL_FOR
#endif

namespace orc {

using Vec = std::vector<double>;

struct Triplets {
  int n = 0;
  std::vector<int> r, c;
  std::vector<double> v;
  void add(int row, int col, double val) {
    r.push_back(row);
    c.push_back(col);
    v.push_back(val);
  }
  // rs/fim/assembly_ad.rs:126-130 add_if_nonzero
  void add_nz(int row, int col, double val) {
    if (std::fabs(val) > 1e-14) add(row, col, val);
  }
};

// CSR with sorted, duplicate-summed columns (sprs TriMatI::to_csr semantics).
struct Csr {
  int n = 0;
  std::vector<int> ptr, col;
  Vec val;
  static Csr from(const Triplets& t) {
    Csr m;
    m.n = t.n;
    std::vector<std::vector<std::pair<int, double>>> rows(t.n);
    for (size_t k = 0; k < t.r.size(); ++k) rows[t.r[k]].push_back({t.c[k], t.v[k]});
    m.ptr.assign(t.n + 1, 0);
    // each row sorted and its duplicates summed in insertion order, rows independent
    ORC_PARALLEL_FOR
    for (int i = 0; i < t.n; ++i) {
      auto& rw = rows[i];
      std::stable_sort(rw.begin(), rw.end(), [](auto& a, auto& b) { return a.first < b.first; });
      size_t w = 0;
      for (size_t k = 0; k < rw.size(); ++k) {
        if (w > 0 && rw[w - 1].first == rw[k].first)
          rw[w - 1].second += rw[k].second;
        else
          rw[w++] = rw[k];
      }
      rw.resize(w);
    }
    for (int i = 0; i < t.n; ++i) m.ptr[i + 1] = m.ptr[i] + (int)rows[i].size();
    m.col.resize(m.ptr[t.n]);
    m.val.resize(m.ptr[t.n]);
    ORC_PARALLEL_FOR
    for (int i = 0; i < t.n; ++i)
      for (size_t k = 0; k < rows[i].size(); ++k) {
        m.col[m.ptr[i] + k] = rows[i][k].first;
        m.val[m.ptr[i] + k] = rows[i][k].second;
      }
    return m;
  }
  double at(int i, int j) const {
    auto b = col.begin() + ptr[i], e = col.begin() + ptr[i + 1];
    auto it = std::lower_bound(b, e, j);
    if (it != e && *it == j) return val[it - col.begin()];
    return 0.0;
  }
  // rs/fim/linear/gmres_block_jacobi.rs:1365-1375 (row-sequential sums)
  Vec mul(const Vec& x) const {
    Vec y(n, 0.0);
    ORC_PARALLEL_FOR
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[col[k]];
      y[i] = s;
    }
    return y;
  }
};

inline double norm2(const Vec& a) {
  double s = 0.0;
  for (double x : a) s += x * x;
  return std::sqrt(s);
}
inline double dot(const Vec& a, const Vec& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// Dense matrix helpers (row-major).
struct Dense {
  int n = 0;
  Vec a;
  Dense() = default;
  explicit Dense(int n_) : n(n_), a((size_t)n_ * n_, 0.0) {}
  double& operator()(int i, int j) { return a[(size_t)i * n + j]; }
  double operator()(int i, int j) const { return a[(size_t)i * n + j]; }
};

// Gauss-Jordan inverse with partial pivoting; false if singular.
inline bool dense_inverse(const Dense& m, Dense& inv) {
  int n = m.n;
  Dense w = m;
  inv = Dense(n);
  for (int i = 0; i < n; ++i) inv(i, i) = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    double best = std::fabs(w(c, c));
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(w(r, c)) > best) { best = std::fabs(w(r, c)); piv = r; }
    if (!(best > 0.0) || !std::isfinite(best)) return false;
    if (piv != c)
      for (int k = 0; k < n; ++k) { std::swap(w(c, k), w(piv, k)); std::swap(inv(c, k), inv(piv, k)); }
    double d = 1.0 / w(c, c);
    for (int k = 0; k < n; ++k) { w(c, k) *= d; inv(c, k) *= d; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double f = w(r, c);
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) { w(r, k) -= f * w(c, k); inv(r, k) -= f * inv(c, k); }
    }
  }
  for (double v : inv.a)
    if (!std::isfinite(v)) return false;
  return true;
}

// rs/fim/linear/gmres_block_jacobi.rs:888-901 invert_tail_block (diag-reciprocal fallback)
inline Dense invert_or_diag(const Dense& m) {
  Dense inv;
  if (dense_inverse(m, inv)) return inv;
  Dense f(m.n);
  for (int i = 0; i < m.n; ++i) {
    double d = m(i, i);
    f(i, i) = std::fabs(d) > EPS ? 1.0 / d : 1.0;
  }
  return f;
}

struct LinOpts {  // rs/fim/linear/mod.rs:207-235 defaults
  int restart = 30, max_iterations = 20;
  double rel_tol = 5e-3, abs_tol = 1e-12;
  bool eliminate_wells = true;
  bool allow_direct = true;  // rows <= 512 -> direct (rs/fim/linear/mod.rs:40,313-335)
};

struct LinReport {
  Vec x;
  bool converged = false;
  int iterations = 0;
  double rhs_norm = 0.0, final_residual = 0.0;
  double reduction() const { return final_residual / std::max(rhs_norm, 1e-30); }
  // backend and CPR diagnostics (rs/fim/linear/mod.rs:238-250): a direct solve, or the
  // pressure-stage applications with their average and last residual reduction
  bool direct = false;
  int cpr_apps = 0;
  double cpr_avg = 1.0, cpr_last = 1.0;
};

// Direct dense LU solve for tiny systems (the reference uses sparse LU natively).
inline bool direct_solve(const Csr& A, const Vec& b, Vec& x) {
  int n = A.n;
  Dense m(n);
  for (int i = 0; i < n; ++i)
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) m(i, A.col[k]) = A.val[k];
  std::vector<int> perm(n);
  Vec rhs = b;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    double best = std::fabs(m(c, c));
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(m(r, c)) > best) { best = std::fabs(m(r, c)); piv = r; }
    if (!(best > 1e-300)) return false;
    if (piv != c) {
      for (int k = 0; k < n; ++k) std::swap(m(c, k), m(piv, k));
      std::swap(rhs[c], rhs[piv]);
    }
    for (int r = c + 1; r < n; ++r) {
      double f = m(r, c) / m(c, c);
      if (f == 0.0) continue;
      for (int k = c; k < n; ++k) m(r, k) -= f * m(c, k);
      rhs[r] -= f * rhs[c];
    }
  }
  x.assign(n, 0.0);
  for (int r = n - 1; r >= 0; --r) {
    double s = rhs[r];
    for (int k = r + 1; k < n; ++k) s -= m(r, k) * x[k];
    x[r] = s / m(r, r);
  }
  for (double v : x)
    if (!std::isfinite(v)) return false;
  return true;
}

// ---------------------------------------------------------------- CPR
// rs/fim/linear/gmres_block_jacobi.rs:239-354 factorize_block_ilu0 and :190-237 apply.
struct BlockIlu0 {
  int nb = 0;
  struct Blk { int j; double m[9]; };
  std::vector<std::vector<Blk>> L, U;
  std::vector<std::array<double, 9>> Dinv;

  static void mm(const double* a, const double* b, double* c) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += a[i * 3 + k] * b[k * 3 + j];
        c[i * 3 + j] = s;
      }
  }
  static std::array<double, 9> inv3(const double* b) {
    Dense m(3);
    for (int i = 0; i < 9; ++i) m.a[i] = b[i];
    Dense r = invert_or_diag(m);
    std::array<double, 9> o;
    for (int i = 0; i < 9; ++i) o[i] = r.a[i];
    return o;
  }

  void factor(const Csr& A, int n_blocks) {
    nb = n_blocks;
    std::vector<std::map<int, std::array<double, 9>>> rows(nb);
    for (int rb = 0; rb < nb; ++rb) {
      std::vector<int> cols{rb};
      for (int l = 0; l < 3; ++l) {
        int r = rb * 3 + l;
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
          int c = A.col[k];
          if (c >= 3 * nb || std::fabs(A.val[k]) <= EPS) continue;
          cols.push_back(c / 3);
        }
      }
      for (int cb : cols) {
        if (rows[rb].count(cb)) continue;
        std::array<double, 9> blk;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) blk[i * 3 + j] = A.at(rb * 3 + i, cb * 3 + j);
        rows[rb][cb] = blk;
      }
    }
    L.assign(nb, {});
    U.assign(nb, {});
    Dinv.assign(nb, {});
    for (int i = 0; i < nb; ++i) {
      std::vector<int> ks;
      for (auto& kv : rows[i])
        if (kv.first < i) ks.push_back(kv.first);
      for (int k : ks) {
        double lik[9];
        mm(rows[i][k].data(), Dinv[k].data(), lik);
        for (auto& u : U[k]) {
          auto it = rows[i].find(u.j);
          if (it != rows[i].end()) {
            double t[9];
            mm(lik, u.m, t);
            for (int q = 0; q < 9; ++q) it->second[q] -= t[q];
          }
        }
        Blk b;
        b.j = k;
        std::copy(lik, lik + 9, b.m);
        L[i].push_back(b);
        rows[i].erase(k);
      }
      std::array<double, 9> diag{};
      auto it = rows[i].find(i);
      if (it != rows[i].end()) diag = it->second;
      Dinv[i] = inv3(diag.data());
      for (auto& kv : rows[i])
        if (kv.first > i) {
          Blk b;
          b.j = kv.first;
          std::copy(kv.second.begin(), kv.second.end(), b.m);
          U[i].push_back(b);
        }
    }
    flatten(L, lp, lb);
    flatten(U, up_, ub);
  }
  // the factors' block rows in contiguous arrays for the sweeps (same blocks, same order)
  std::vector<int> lp, up_;
  std::vector<Blk> lb, ub;
  static void flatten(const std::vector<std::vector<Blk>>& rows, std::vector<int>& p, std::vector<Blk>& b) {
    p.assign(rows.size() + 1, 0);
    for (size_t i = 0; i < rows.size(); ++i) p[i + 1] = p[i] + (int)rows[i].size();
    b.clear();
    b.reserve(p.back());
    for (auto& r : rows) b.insert(b.end(), r.begin(), r.end());
  }

  Vec apply(const Vec& r) const {
    std::vector<std::array<double, 3>> y(nb), x(nb);
    for (int i = 0; i < nb; ++i) {
      double v[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
      for (int q = lp[i]; q < lp[i + 1]; ++q) {
        const Blk& l = lb[q];
        const auto& yk = y[l.j];
        for (int a = 0; a < 3; ++a) v[a] -= l.m[a * 3] * yk[0] + l.m[a * 3 + 1] * yk[1] + l.m[a * 3 + 2] * yk[2];
      }
      y[i] = {v[0], v[1], v[2]};
    }
    for (int i = nb - 1; i >= 0; --i) {
      double v[3] = {y[i][0], y[i][1], y[i][2]};
      for (int q = up_[i]; q < up_[i + 1]; ++q) {
        const Blk& u = ub[q];
        const auto& xj = x[u.j];
        for (int a = 0; a < 3; ++a) v[a] -= u.m[a * 3] * xj[0] + u.m[a * 3 + 1] * xj[1] + u.m[a * 3 + 2] * xj[2];
      }
      const auto& D = Dinv[i];
      for (int a = 0; a < 3; ++a) x[i][a] = D[a * 3] * v[0] + D[a * 3 + 1] * v[1] + D[a * 3 + 2] * v[2];
    }
    Vec out(r.size(), 0.0);
    for (int i = 0; i < nb; ++i)
      for (int a = 0; a < 3; ++a) out[3 * i + a] = x[i][a];
    return out;
  }
};

// rs/fim/linear/gmres_block_jacobi.rs:746-808 factorize_pressure_ilu0, :634-658 apply
struct ScalarIlu0 {
  int n = 0;
  std::vector<std::vector<std::pair<int, double>>> Lr, Ur;
  Vec ud;
  static double row_entry(const std::vector<std::pair<int, double>>& e, int c) {
    auto it = std::lower_bound(e.begin(), e.end(), c, [](const std::pair<int, double>& a, int b) { return a.first < b; });
    if (it != e.end() && it->first == c) return it->second;
    return 0.0;
  }
  void factor(const std::vector<std::vector<std::pair<int, double>>>& rows) {
    n = (int)rows.size();
    Lr.assign(n, {});
    Ur.assign(n, {});
    ud.assign(n, 1.0);
    for (int i = 0; i < n; ++i) {
      std::vector<std::pair<int, double>> lo, up;
      double diag = 0.0;
      for (auto& e : rows[i]) {
        if (e.first < i) lo.push_back(e);
        else if (e.first == i) diag = e.second;
        else up.push_back(e);
      }
      for (auto& e : lo) {
        double s = e.second;
        for (auto& l : Lr[i]) {
          if (l.first >= e.first) break;
          s -= l.second * row_entry(Ur[l.first], e.first);
        }
        double d = ud[e.first];
        Lr[i].push_back({e.first, std::fabs(d) > EPS ? s / d : 0.0});
      }
      double u = diag;
      for (auto& l : Lr[i]) u -= l.second * row_entry(Ur[l.first], i);
      if (std::fabs(u) <= EPS) u = std::fabs(diag) > EPS ? diag : 1.0;
      ud[i] = u;
      for (auto& e : up) {
        double s = e.second;
        for (auto& l : Lr[i]) s -= l.second * row_entry(Ur[l.first], e.first);
        Ur[i].push_back({e.first, s});
      }
    }
    flatten(Lr, lp, lc, lv);
    flatten(Ur, up_, uc, uv);
  }
  // the factors' rows in contiguous arrays for the sweeps (same entries, same order)
  std::vector<int> lp, lc, up_, uc;
  Vec lv, uv;
  static void flatten(const std::vector<std::vector<std::pair<int, double>>>& rows, std::vector<int>& p,
                      std::vector<int>& c, Vec& v) {
    p.assign(rows.size() + 1, 0);
    for (size_t i = 0; i < rows.size(); ++i) p[i + 1] = p[i] + (int)rows[i].size();
    c.resize(p.back());
    v.resize(p.back());
    for (size_t i = 0; i < rows.size(); ++i)
      for (size_t k = 0; k < rows[i].size(); ++k) {
        c[p[i] + k] = rows[i][k].first;
        v[p[i] + k] = rows[i][k].second;
      }
  }
  Vec apply(const Vec& r) const {
    Vec y(n, 0.0), x(n, 0.0);
    for (int i = 0; i < n; ++i) {
      double s = r[i];
      for (int k = lp[i]; k < lp[i + 1]; ++k) s -= lv[k] * y[lc[k]];
      y[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int k = up_[i]; k < up_[i + 1]; ++k) s -= uv[k] * x[uc[k]];
      x[i] = std::fabs(ud[i]) > EPS ? s / ud[i] : s;
    }
    return x;
  }
};

struct LinStats {
  long fgmres_iterations = 0, bicgstab_iterations = 0, pressure_solves = 0;
};

// Reduced (well-free) CPR preconditioner: rs/fim/linear/gmres_block_jacobi.rs:1022-1363 build,
// :471-501 apply, :503-610 pressure stage, :660-730 BiCGStab.
struct Cpr {
  int nb = 0;
  std::vector<std::array<double, 3>> w;  // quasi-IMPES restriction weights
  std::vector<std::vector<std::pair<int, double>>> prow;
  bool dense = false;
  Dense pinv;
  ScalarIlu0 pilu;
  BlockIlu0 bilu;
  LinStats* stats = nullptr;
  // rs/fim/linear/gmres_block_jacobi.rs:380-420 PressureCorrectionAccumulator
  mutable int apps = 0;
  mutable double ratio_sum = 0.0, ratio_last = 1.0;

  void build(const Csr& A, int n_blocks) {
    nb = n_blocks;
    w.assign(nb, {0, 0, 0});
    ORC_PARALLEL_FOR
    for (int b = 0; b < nb; ++b) {
      Dense blk(3);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) blk(i, j) = A.at(3 * b + i, 3 * b + j);
      Dense inv = invert_or_diag(blk);  // :961-975 QuasiImpes: row 0 of the inverse
      double r[3] = {inv(0, 0), inv(0, 1), inv(0, 2)};
      double mx = 0.0;
      for (double x : r) mx = std::max(mx, std::fabs(x));
      if (mx > EPS)
        for (double& x : r) x /= mx;  // :123-133 normalize_weights
      w[b] = {r[0], r[1], r[2]};
    }
    // :1160-1220 coarse rows A_p[i][j] = sum_l w_i[l] * A[3i+l][3j] (prolongation e0)
    prow.assign(nb, {});
    ORC_PARALLEL_FOR
    for (int b = 0; b < nb; ++b) {
      std::map<int, double> coef;
      for (int l = 0; l < 3; ++l) {
        double rw = w[b][l];
        if (std::fabs(rw) <= EPS) continue;
        int r = 3 * b + l;
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
          int c = A.col[k];
          if (c % 3 != 0) continue;  // prolongation is e0: only pressure columns
          coef[c / 3] += rw * A.val[k];
        }
      }
      for (auto& kv : coef)
        if (std::fabs(kv.second) > 1e-14) prow[b].push_back({kv.first, kv.second});
    }
    dense = nb > 0 && nb <= 299;  // :24 PRESSURE_DIRECT_SOLVE_ROW_THRESHOLD
    if (dense) {
      Dense m(nb);
      for (int i = 0; i < nb; ++i)
        for (auto& e : prow[i]) m(i, e.first) = e.second;
      dense = dense_inverse(m, pinv);
    }
    pilu.factor(prow);
    bilu.factor(A, nb);
  }

  Vec pmul(const Vec& x) const {
    Vec y(nb, 0.0);
    ORC_PARALLEL_FOR
    for (int i = 0; i < nb; ++i) {
      double s = 0.0;
      for (auto& e : prow[i]) s += e.second * x[e.first];
      y[i] = s;
    }
    return y;
  }

  Vec bicgstab(const Vec& b, int max_it, double tol) const {
    Vec x(nb, 0.0);
    Vec r = b;  // x0 = 0
    if (norm2(r) <= tol) return x;
    Vec rh = r, v(nb, 0.0), p(nb, 0.0);
    double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
    for (int it = 0; it < max_it; ++it) {
      if (stats) stats->bicgstab_iterations++;
      double rho = dot(rh, r);
      if (!std::isfinite(rho) || std::fabs(rho) < EPS) break;
      if (it == 0) {
        p = r;
      } else {
        double beta = (rho / rho_prev) * (alpha / omega);
        ORC_PARALLEL_FOR
        for (int i = 0; i < nb; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
      }
      Vec ph = pilu.apply(p);
      v = pmul(ph);
      double vd = dot(rh, v);
      if (!std::isfinite(vd) || std::fabs(vd) < EPS) break;
      alpha = rho / vd;
      Vec s(nb);
      ORC_PARALLEL_FOR
      for (int i = 0; i < nb; ++i) s[i] = r[i] - alpha * v[i];
      if (norm2(s) <= tol) {
        ORC_PARALLEL_FOR
        for (int i = 0; i < nb; ++i) x[i] += alpha * ph[i];
        break;
      }
      Vec sh = pilu.apply(s);
      Vec t = pmul(sh);
      double tt = dot(t, t);
      if (!std::isfinite(tt) || tt <= EPS) break;
      omega = dot(t, s) / tt;
      if (!std::isfinite(omega) || std::fabs(omega) < EPS) break;
      ORC_PARALLEL_FOR
      for (int i = 0; i < nb; ++i) {
        x[i] += alpha * ph[i] + omega * sh[i];
        r[i] = s[i] - omega * t[i];
      }
      if (norm2(r) <= tol) break;
      rho_prev = rho;
    }
    return x;
  }

  Vec apply(const Csr& A, const Vec& v) const {
    Vec rp(nb);
    for (int b = 0; b < nb; ++b) rp[b] = w[b][0] * v[3 * b] + w[b][1] * v[3 * b + 1] + w[b][2] * v[3 * b + 2];
    Vec xp;
    if (stats) stats->pressure_solves++;
    if (dense) {
      xp.assign(nb, 0.0);
      for (int i = 0; i < nb; ++i) {
        double s = 0.0;
        for (int j = 0; j < nb; ++j) s += pinv(i, j) * rp[j];
        xp[i] = s;
      }
    } else {
      double tol = std::max(norm2(rp), EPS) * 1e-6;  // :26 PRESSURE_DEFECT_CORRECTION_REL_TOL
      xp = bicgstab(rp, 50, tol);                    // :25 max 50 iterations
    }
    {  // :584-609 the coarse solve's residual reduction |rp - P xp| / |rp| (0 for |rp| <= eps)
      const double bn = norm2(rp);
      Vec px = pmul(xp);
      for (int i = 0; i < nb; ++i) px[i] = rp[i] - px[i];
      const double ratio = bn > std::numeric_limits<double>::epsilon() ? norm2(px) / bn : 0.0;
      apps++;
      ratio_sum += ratio;
      ratio_last = ratio;
    }
    Vec res(v.size(), 0.0);
    for (int b = 0; b < nb; ++b) res[3 * b] += xp[b];
    Vec Ar = A.mul(res);
    Vec corr(v.size());
    for (size_t i = 0; i < v.size(); ++i) corr[i] = v[i] - Ar[i];
    Vec sm = bilu.apply(corr);
    for (size_t i = 0; i < v.size(); ++i) res[i] += sm[i];
    return res;
  }
};

// rs/fim/linear/gmres_block_jacobi.rs:1470-1665 run_true_flexible_gmres
inline LinReport fgmres(const Csr& A, const Vec& b, const LinOpts& o, const Cpr& M, LinStats* st) {
  LinReport rep;
  int n = A.n;
  double bn = norm2(b);
  rep.rhs_norm = bn;
  double tol = o.abs_tol + o.rel_tol * std::max(bn, EPS);
  if (n == 0) { rep.converged = true; return rep; }
  int restart = std::max(o.restart, 2);
  int max_it = std::max(o.max_iterations, restart);
  Vec x(n, 0.0);
  int iters = 0;
  auto residual = [&](const Vec& s) {
    Vec r = A.mul(s);
    ORC_PARALLEL_FOR
    for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
    return r;
  };
  while (iters < max_it) {
    Vec r = residual(x);
    double beta = norm2(r);
    if (beta <= tol) {
      rep.x = x; rep.converged = true; rep.iterations = iters; rep.final_residual = beta;
      return rep;
    }
    int m = std::min(restart, max_it - iters);
    std::vector<Vec> V, Z;
    Vec v0(n);
    ORC_PARALLEL_FOR
    for (int i = 0; i < n; ++i) v0[i] = r[i] / beta;
    V.push_back(v0);
    std::vector<Vec> H(m + 1, Vec(m, 0.0));
    Vec cs(m, 0.0), sn(m, 0.0), g(m + 1, 0.0);
    g[0] = beta;
    Vec origin = x, cand = x;
    for (int j = 0; j < m; ++j) {
      Vec z = M.apply(A, V[j]);
      Vec wv = A.mul(z);
      Z.push_back(z);
      for (int p = 0; p <= j; ++p) {
        double h = dot(V[p], wv);
        H[p][j] = h;
        ORC_PARALLEL_FOR
        for (int i = 0; i < n; ++i) wv[i] -= V[p][i] * h;
      }
      double nn = norm2(wv);
      H[j + 1][j] = nn;
      if (nn > EPS) {
        ORC_PARALLEL_FOR
        for (int i = 0; i < n; ++i) wv[i] /= nn;
        V.push_back(wv);
      }
      for (int p = 0; p < j; ++p) {
        double a = H[p][j], c = H[p + 1][j];
        H[p][j] = cs[p] * a + sn[p] * c;
        H[p + 1][j] = -sn[p] * a + cs[p] * c;
      }
      double a = H[j][j], c = H[j + 1][j];
      double rad = std::hypot(a, c);
      double co = 1.0, si = 0.0;
      if (rad > EPS) { co = a / rad; si = c / rad; }
      cs[j] = co;
      sn[j] = si;
      H[j][j] = co * a + si * c;
      H[j + 1][j] = -si * a + co * c;
      double g0 = g[j], g1 = g[j + 1];
      g[j] = co * g0 + si * g1;
      g[j + 1] = -si * g0 + co * g1;
      iters++;
      if (st) st->fgmres_iterations++;
      int cols = j + 1;
      Vec y(cols, 0.0);
      for (int row = cols - 1; row >= 0; --row) {
        double s = g[row];
        for (int cc = row + 1; cc < cols; ++cc) s -= H[row][cc] * y[cc];
        double d = H[row][row];
        y[row] = std::fabs(d) > EPS ? s / d : 0.0;
      }
      cand = origin;
      Vec upd(n, 0.0);
      ORC_PARALLEL_FOR
      for (int i = 0; i < n; ++i)
        for (int cc = 0; cc < cols; ++cc) upd[i] += Z[cc][i] * y[cc];
      ORC_PARALLEL_FOR
      for (int i = 0; i < n; ++i) cand[i] += upd[i];
      Vec cr = residual(cand);
      double tr = norm2(cr);
      if (tr <= tol) {
        rep.x = cand; rep.converged = true; rep.iterations = iters; rep.final_residual = tr;
        return rep;
      }
      if (nn <= EPS || iters >= max_it) {
        rep.x = cand; rep.converged = false; rep.iterations = iters; rep.final_residual = tr;
        return rep;
      }
    }
    x = cand;
  }
  rep.x = x;
  rep.converged = false;
  rep.iterations = iters;
  rep.final_residual = norm2(residual(x));
  return rep;
}

inline LinReport solve_cells(const Csr& A, const Vec& b, int n_blocks, const LinOpts& o, LinStats* st) {
  if (o.allow_direct && A.n <= 512) {  // rs/fim/linear/mod.rs:313-335
    LinReport r;
    r.rhs_norm = norm2(b);
    if (direct_solve(A, b, r.x)) {
      Vec ax = A.mul(r.x);
      for (int i = 0; i < A.n; ++i) ax[i] = b[i] - ax[i];
      r.final_residual = norm2(ax);
      r.converged = true;
      r.iterations = 1;
      r.direct = true;
      return r;
    }
  }
  Cpr M;
  M.stats = st;
  M.build(A, n_blocks);
  LinReport r = fgmres(A, b, o, M, st);
  r.cpr_apps = M.apps;
  r.cpr_avg = M.apps ? M.ratio_sum / M.apps : 1.0;
  r.cpr_last = M.apps ? M.ratio_last : 1.0;
  return r;
}

// rs/fim/linear/well_schur.rs:70-231 + :382-471. The tail starts at 3*n_cells.
inline LinReport solve_linearized(const Csr& J, const Vec& rhs, int n_cells, const LinOpts& o, LinStats* st) {
  int ns = 3 * n_cells;
  int m = J.n - ns;
  if (!o.eliminate_wells || m == 0) {
    if (o.allow_direct && J.n <= 512) {
      LinReport r;
      r.rhs_norm = norm2(rhs);
      if (direct_solve(J, rhs, r.x)) {
        Vec ax = J.mul(r.x);
        for (int i = 0; i < J.n; ++i) ax[i] = rhs[i] - ax[i];
        r.final_residual = norm2(ax);
        r.converged = true;
        r.iterations = 1;
        r.direct = true;
        return r;
      }
    }
    return solve_cells(J, rhs, n_cells, o, st);
  }
  if (o.allow_direct && J.n <= 512) {
    LinReport r;
    r.rhs_norm = norm2(rhs);
    if (direct_solve(J, rhs, r.x)) {
      Vec ax = J.mul(r.x);
      for (int i = 0; i < J.n; ++i) ax[i] = rhs[i] - ax[i];
      r.final_residual = norm2(ax);
      r.converged = true;
      r.iterations = 1;
      r.direct = true;
      return r;
    }
  }
  Dense T(m);
  for (int a = 0; a < m; ++a)
    for (int k = J.ptr[ns + a]; k < J.ptr[ns + a + 1]; ++k)
      if (J.col[k] >= ns) T(a, J.col[k] - ns) = J.val[k];
  Dense Ti = invert_or_diag(T);
  std::vector<std::vector<std::pair<int, double>>> jwr(m);
  for (int a = 0; a < m; ++a)
    for (int k = J.ptr[ns + a]; k < J.ptr[ns + a + 1]; ++k)
      if (J.col[k] < ns && std::fabs(J.val[k]) > EPS) jwr[a].push_back({J.col[k], J.val[k]});
  std::vector<std::pair<int, std::vector<std::pair<int, double>>>> jrw;
  for (int r = 0; r < ns; ++r) {
    std::vector<std::pair<int, double>> e;
    for (int k = J.ptr[r]; k < J.ptr[r + 1]; ++k)
      if (J.col[k] >= ns && std::fabs(J.val[k]) > EPS) e.push_back({J.col[k] - ns, J.val[k]});
    if (!e.empty()) jrw.push_back({r, e});
  }
  Vec rt(m), tir(m, 0.0);
  for (int a = 0; a < m; ++a) rt[a] = rhs[ns + a];
  for (int a = 0; a < m; ++a)
    for (int c = 0; c < m; ++c) tir[a] += Ti(a, c) * rt[c];
  Vec rr(rhs.begin(), rhs.begin() + ns);
  for (auto& re : jrw) {
    double corr = 0.0;
    for (auto& e : re.second) corr += e.second * tir[e.first];
    rr[re.first] -= corr;
  }
  Triplets t;
  t.n = ns;
  for (int r = 0; r < ns; ++r)
    for (int k = J.ptr[r]; k < J.ptr[r + 1]; ++k)
      if (J.col[k] < ns) t.add(r, J.col[k], J.val[k]);
  for (auto& re : jrw) {
    Vec wgt(m, 0.0);
    for (auto& e : re.second)
      for (int c = 0; c < m; ++c) wgt[c] += e.second * Ti(e.first, c);
    for (int c = 0; c < m; ++c) {
      if (std::fabs(wgt[c]) <= EPS) continue;
      for (auto& jw : jwr[c]) t.add(re.first, jw.first, -wgt[c] * jw.second);
    }
  }
  Csr R = Csr::from(t);
  LinOpts ro = o;
  LinReport red = solve_cells(R, rr, n_cells, ro, st);
  // recover dx_W = T^-1 (r_W - J_WR dx_R)
  Vec x(J.n, 0.0);
  for (int i = 0; i < ns; ++i) x[i] = red.x.empty() ? 0.0 : red.x[i];
  Vec tr(m);
  for (int a = 0; a < m; ++a) {
    double s = rt[a];
    for (auto& e : jwr[a]) s -= e.second * x[e.first];
    tr[a] = s;
  }
  for (int a = 0; a < m; ++a) {
    double s = 0.0;
    for (int c = 0; c < m; ++c) s += Ti(a, c) * tr[c];
    x[ns + a] = s;
  }
  LinReport rep;
  rep.x = x;
  Vec ax = J.mul(x);
  for (int i = 0; i < J.n; ++i) ax[i] = rhs[i] - ax[i];
  rep.final_residual = norm2(ax);
  rep.rhs_norm = norm2(rhs);
  double tol = o.abs_tol + o.rel_tol * std::max(rep.rhs_norm, EPS);
  rep.converged = red.converged && rep.final_residual <= tol;
  rep.iterations = red.iterations;
  rep.direct = red.direct;
  rep.cpr_apps = red.cpr_apps;
  rep.cpr_avg = red.cpr_avg;
  rep.cpr_last = red.cpr_last;
  return rep;
}

}  // namespace orc
