// Host-side simulator object of the HIP FIM path: model description, device
// buffers, the linear stack (well Schur -> CPR-FGMRES) and the Newton /
// substep drivers. All device work is issued on the simulator's own stream.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/ressim_fim.h"

#include <algorithm>
#include <array>
#include <charconv>
#include <filesystem>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "rsf_kernels.hpp"
#include "rsf_sweep.hpp"

namespace rsf {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define RSF_CHECK(expr)                                                                            \
  do {                                                                                             \
    hipError_t _e = (expr);                                                                        \
    if (_e != hipSuccess)                                                                          \
      throw ::rsf::HipError(std::string(#expr) + " failed: " + hipGetErrorString(_e));             \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    RSF_CHECK(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  // The host source may be a temporary: wait for the copy before returning.
  void upload(const T* h, size_t count, hipStream_t s) {
    alloc(count);
    if (count) {
      RSF_CHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
      RSF_CHECK(hipStreamSynchronize(s));
    }
  }
  void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
  void download(T* h, size_t count, hipStream_t s) const {
    if (count) RSF_CHECK(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

inline int nblk(long n) { return (int)((n + kBlock - 1) / kBlock); }

// ------------------------------------------------------------------ model
struct WellSchedule {
  bool has = false, rate = false, has_target = false, has_surface = false, has_limit = false, enabled = true;
  double target = 0, surface = 0, limit = 0;
};

struct Completion {  // rs/well.rs:17-29 (one completion per cell)
  int i, j, k;
  double bhp, rw, skin;
  bool injector;
  std::string id;
  WellSchedule schedule;
  bool has_datum = false, has_density = false;
  double datum = 0, density = 0, head = 0;
  int fluid = -1;  // -1 inherit the global injected fluid, 0 water, 1 gas
};

struct PhysWell {
  int rep;
  bool injector;
  std::vector<int> perfs;
};

// rs/reporting.rs:246-292 TimePointRates (the optional sweep metrics are not kept)
struct TimePoint {
  double time;
  double total_production_oil, total_production_liquid, total_production_liquid_reservoir;
  double total_injection, total_injection_reservoir;
  double material_balance_error_m3, material_balance_error_oil_m3;
  double avg_reservoir_pressure, avg_water_saturation;
  double total_production_gas, avg_gas_saturation, material_balance_error_gas_m3, producing_gor;
  double producer_bhp_limited_fraction, injector_bhp_limited_fraction;
  // sweep efficiency (rs/reporting.rs:32-47 SweepMetrics): NaN where the reference has None
  double sweep_e_a, sweep_e_v, sweep_e_vol, sweep_mobile_oil_recovered;
};

// rs/reporting.rs:24-35 SweepConfig; geometry 0 areal, 1 vertical, 2 both (or any other name)
struct SweepConfig {
  bool on = false;
  int geometry = 2;
  double swept_threshold = 0.0, initial_oil_saturation = 0.0, residual_oil_saturation = 0.0;
};

struct ControlState {  // rs/fim/wells.rs:626-634
  bool enabled, rate, surface, has_target;
  double target, limit, bhp_target;
};

struct Options {
  int max_newton = 20;
  int min_newton = 2;          // OPM newton-min-iterations (rs/fim/newton/convergence.rs:378)
  double residual_tol = 0.0;   // > 0: scaled-residual acceptance with MB <= mb_tol
  double mb_tol = 1e-7;
  double lin_rel_tol = 5e-3;   // rs/fim/linear/mod.rs:207-235
  double lin_abs_tol = 1e-12;
  int lin_max_it = 20;
  int restart = 30;
  int precision = 0;           // 0 FP64, 1 mixed (FP32 preconditioner, FP32 inner vectors)
  int ordering = 0;            // 0 natural hyperplane ILU (reference order), 1 red-black
  int pressure_max_it = 50;
  double pressure_rel_tol = 1e-6;
};

struct Stats {
  long newton = 0, linear = 0, substeps = 0, retries = 0, bicgstab = 0, pressure_solves = 0;
  double assembly_ms = 0, linear_ms = 0, total_ms = 0;
};

struct Sim {
  // ---- model (host) ----
  int nx, ny, nz, n;
  double dx = 10, dy = 10, depth_ref = 0;
  std::vector<double> dz, phi, kx, ky, kz, pc_scale;
  double mu_o = 1.0, mu_w = 0.5, c_o = 1e-5, c_w = 3e-6, rho_o = 800, rho_w = 1000;
  bool gas_redissolution = true;  // rs/frontend.rs:1208 (default true, rs/frontend.rs:119)
  double b_o = 1, b_w = 1, p_ref_w = 300, c_r = 0, p_ref_rock = 300, mu_g = 0.02, c_g = 1e-4, rho_g = 10;
  Scal2P scal{0.1, 0.1, 2.0, 2.0, 1.0, 1.0};
  bool has_scal3 = false;
  Scal3P scal3{};
  std::vector<double> swof_h, sgof_h;  // SWOF / SGOF rows of 3 (empty: Corey)
  double pc_entry = 5.0, pc_lambda = 2.0, pcog_entry = 0.0, pcog_lambda = 2.0;
  bool has_pc_og = false, gravity = false, three_phase = false;
  int corey_points = 21;
  int injected_fluid = 1;
  bool inj_rate = false, prod_rate = false, inj_enabled = true;
  double tgt_inj = 0, tgt_prod = 0, tgt_inj_surf = 0, tgt_prod_surf = 0;
  bool has_tgt_inj_surf = false, has_tgt_prod_surf = false;
  double bhp_min = -100, bhp_max = 2000;
  std::vector<Completion> wells;
  std::vector<PhysWell> topo;
  std::vector<int> perf_well;
  // PVT (host mirror of the device table)
  bool has_table = false;
  std::vector<double> sat_h, br_rs_h, br_rows_h;
  std::vector<int> br_off_h, br_len_h;
  double tbl_c_o = 0;
  // stored simulator state (host copy kept in sync with device on demand)
  std::vector<double> h_p, h_sw, h_so, h_sg, h_rs;
  double time = 0;
  Options opt;
  Stats stats;
  std::string warning;
  long invalid_gen = 0;  // bumped by every invalid-argument message
  std::vector<double> accepted_dts;
  // rs/reporting.rs:141-243 FimStepStats with its rungs, one record per step() call
  struct StepRecord {
    rsf_step_stats s;
    std::vector<rsf_step_rung> rungs;
  };
  std::vector<StepRecord> step_stats;
  // rs/fim/linear/gmres_block_jacobi.rs:380-420 PressureCorrectionAccumulator: the residual
  // reductions of the pressure solves of the current linear solve
  int cpr_apps = 0;
  double cpr_ratio_sum = 0.0, cpr_ratio_last = 1.0;
  // per CPR application of the current linear solve: |rp - P xp|^2 and |rp|^2
  DevBuf<double> cpr_red;
  int cpr_red_cap = 0;
  template <class PT>
  void pressure_reduction(const PT* rp, const PT* xp);
  // dense routes for small systems (rsf_dense.hpp): the direct solve of the reduced system
  // and the exact inverse of a small CPR coarse matrix ([P | P^-1] after elimination)
  DevBuf<double> dense_m;
  DevBuf<int> dense_flag;
  bool coarse_dense = false;
  bool solve_direct(double& final_res);
  bool solve_reduced(double rhs_norm_full, int& iters, double& final_res, bool& direct);

  // ---- device ----
  hipStream_t stream = nullptr;
  bool dev_ready = false, state_on_dev = false, state_dirty_host = false;
  Grid grid{};
  DevBuf<double> d_Tx, d_Ty, d_Tz, d_depth, d_pv, d_pcs;
  DevBuf<double> d_sat, d_br_rs, d_br_rows, d_swof, d_sgof;
  DevBuf<int> d_br_off, d_br_len;
  DevBuf<double> s_p, s_sw, s_so, s_sg, s_rs;  // stored state
  DevBuf<double> x_p, x_sw, x_hc, x0_p, x0_sw, x0_hc, acc0;
  DevBuf<int> x_reg, x0_reg;
  DevBuf<double> R, A, pack;
  // wells
  DevBuf<int> w_off, w_cell, w_flags, w_cdest, w_coff, w_pref;
  DevBuf<double> w_wi, w_head, w_target, w_limit, w_bhpt, w_bhp, w_q, w_tailR, w_wk, w_well, w_extra;
  int n_extra = 0;
  DevBuf<int> x_rows, x_ptr, x_col, x_erow;  // extras CSR over rows, and the row of each extra
  DevBuf<int> x_wave;                         // first extras row of each 64-row wavefront
  DevBuf<int> x_wave2;                        // the same per red-black pressure wavefront
  DevBuf<int> x_rows_c, x_col_c;              // extras rows / columns, colour-compacted
  int n_wave2 = 0;
  int n_extra_rows = 0;
  std::vector<int> perf_cells_h;
  DevBuf<int> d_perf_cells;
  DevBuf<double> d_gather;
  // solver
  DevBuf<double> b, dxv, tmp, cand, tdx;
  std::vector<DevBuf<double>*> V;
  std::vector<DevBuf<double>*> Zd;
  std::vector<DevBuf<float>*> Zf;
  DevBuf<double*> zptr_d;
  DevBuf<float*> zptr_f;
  DevBuf<double> dscal, part, bsc, bpart;  // bpart: BiCGStab block partials
  // FGMRES least-squares state on the device: Hessenberg, rotations, g, y, and two status
  // slots (the last inner step's new-vector norm and squared true residual)
  DevBuf<double> gm;
  double *gm_H = nullptr, *gm_cs = nullptr, *gm_sn = nullptr, *gm_g = nullptr, *gm_y = nullptr, *gm_st = nullptr;
  // Live timing of the pressure SpMV (k_pspmv_dot) inside the BiCGStab: HIP events
  // on the solve's own stream around every launch of an executed iteration.
  bool probe_on = false;
  int probe_every = 1;  // time every probe_every-th iteration's launch
  std::vector<hipEvent_t> probe_ev;  // start/stop pairs, reused per chunk
  long probe_launches = 0;
  double probe_ms = 0.0;
  DevBuf<double> cw;  // quasi-IMPES weights FP64
  // FP64 preconditioner
  DevBuf<double> Dinv_d, Pm_d, pdinv_d, wpt_d, Pe_d, A_d;  // A_d: colour-compacted copy of A
  DevBuf<double> bc_d[11];
  // FP32 preconditioner
  DevBuf<float> A_f, Dinv_f, Pm_f, pdinv_f, wpt_f, Pe_f, Ef, corr_f;
  DevBuf<float> bc_f[11];
  DevBuf<double> corr_d;
  // level schedules
  std::vector<int> lvl_ptr[2];
  DevBuf<int> lvl_rows[2];
  // Natural-order ILU solves (ordering 0): one launch per sweep (rsf_sweep.hpp) when
  // sweep_mode is 1, one launch per hyperplane level when 0. Tiles of sweep_lanes_p lines
  // for the pressure system and sweep_lanes_b for the block system.
  int sweep_mode = 1;
  int sweep_lanes_p = 256, sweep_lanes_b = 64;
  bool sweep_on = false;  // the current preconditioner's natural-order solves use the sweeps
  SweepTiles sweep_tp{}, sweep_tb{};
  DevBuf<float> sl_pf, sl_bf;   // sweep-layout coefficients, FP32 preconditioner
  DevBuf<double> sl_pd, sl_bd;  // FP64 preconditioner
  DevBuf<unsigned long long> sweep_halo_p, sweep_halo_b;  // tagged halo entries per system
  DevBuf<unsigned long long> sweep_halo_f;  // tagged D^-1 of the block factorization's march
  DevBuf<double> sweep_vin, sweep_vout;  // right-hand side and result in march order
  DevBuf<int> sweep_err;
  unsigned sweep_epoch = 0;

  Sim(int nx_, int ny_, int nz_, double phi0) : nx(nx_), ny(ny_), nz(nz_), n(nx_ * ny_ * nz_) {
    dz.assign(nz, 1.0);
    phi.assign(n, phi0);
    kx.assign(n, 100.0);
    ky.assign(n, 100.0);
    kz.assign(n, 10.0);
    pc_scale.assign(n, 1.0);
    h_p.assign(n, 300.0);
    h_sw.assign(n, 0.3);
    h_so.assign(n, 0.7);
    h_sg.assign(n, 0.0);
    h_rs.assign(n, 0.0);
  }
  ~Sim() {
    for (auto e : probe_ev) (void)hipEventDestroy(e);
    for (auto* v : V) delete v;
    for (auto* v : Zd) delete v;
    for (auto* v : Zf) delete v;
    if (stream) (void)hipStreamDestroy(stream);
  }

  int idx(int i, int j, int k) const { return k * nx * ny + j * nx + i; }
  double depth_at_k(int k) const {  // rs/grid.rs:256-263
    double d = depth_ref;
    for (int l = 0; l < k; ++l) d += dz[l];
    return d + dz[k] * 0.5;
  }
  double pore_volume(int id) const { return dx * dy * dz[id / (nx * ny)] * phi[id]; }

  // ---- PVT table (rs/pvt.rs:86-135): branches of equal Rs, saturated rows ----
  void set_table(const double* rows, int nr) {
    struct Row { double p, rs, bo, mu, bg, mg; };
    struct Br { double rs; std::vector<Row> rows; };
    std::vector<Br> brs;
    for (int i = 0; i < nr; ++i) {
      Row r{rows[6 * i], rows[6 * i + 1], rows[6 * i + 2], rows[6 * i + 3], rows[6 * i + 4], rows[6 * i + 5]};
      if (!brs.empty() && fabs(r.rs - brs.back().rs) <= kPvtoRsTol &&
          (brs.back().rows.empty() || r.p >= brs.back().rows.back().p - 1e-9)) {
        brs.back().rows.push_back(r);
      } else {
        brs.push_back({r.rs, {r}});
      }
    }
    for (auto& b : brs) std::stable_sort(b.rows.begin(), b.rows.end(), [](const Row& a, const Row& c) { return a.p < c.p; });
    std::stable_sort(brs.begin(), brs.end(), [](const Br& a, const Br& c) { return a.rs < c.rs; });
    std::vector<Row> sat;
    for (auto& b : brs) sat.push_back(b.rows.front());
    std::stable_sort(sat.begin(), sat.end(), [](const Row& a, const Row& c) { return a.p < c.p; });
    sat_h.clear();
    for (auto& r : sat) sat_h.insert(sat_h.end(), {r.p, r.rs, r.bo, r.mu, r.bg, r.mg});
    br_rs_h.clear();
    br_off_h.clear();
    br_len_h.clear();
    br_rows_h.clear();
    for (auto& b : brs) {
      br_rs_h.push_back(b.rs);
      br_off_h.push_back((int)(br_rows_h.size() / 3));
      br_len_h.push_back((int)b.rows.size());
      for (auto& r : b.rows) br_rows_h.insert(br_rows_h.end(), {r.p, r.bo, r.mu});
    }
    tbl_c_o = c_o;  // rs/frontend.rs:1011-1021 captures the current c_o
    has_table = true;
    for (int i = 0; i < n; ++i) h_rs[i] = pvt_saturated<double>(host_params(), h_p[i]).rs;
    dev_ready = false;
  }

  PhysParams params(bool device) const {
    PhysParams P{};
    P.three_phase = three_phase;
    P.has_table = has_table;
    P.has_scal3 = has_scal3;
    P.has_pc_og = has_pc_og;
    P.gravity = gravity;
    P.corey_points = corey_points;
    P.mu_o = mu_o; P.mu_w = mu_w; P.c_o = c_o; P.c_w = c_w; P.rho_o = rho_o; P.rho_w = rho_w;
    P.b_o = b_o; P.b_w = b_w; P.p_ref_w = p_ref_w; P.c_r = c_r; P.p_ref_rock = p_ref_rock;
    P.mu_g = mu_g; P.rho_g = rho_g;
    P.scal = scal;
    P.scal3 = scal3;
    P.pc_entry = pc_entry; P.pc_lambda = pc_lambda; P.pcog_entry = pcog_entry; P.pcog_lambda = pcog_lambda;
    P.n_sat = has_table ? (int)(sat_h.size() / 6) : 0;
    P.n_br = has_table ? (int)br_rs_h.size() : 0;
    P.tbl_c_o = tbl_c_o;
    P.n_swof = (int)(swof_h.size() / 3);
    P.n_sgof = (int)(sgof_h.size() / 3);
    if (device) {
      P.sat = d_sat.p; P.br_rs = d_br_rs.p; P.br_off = d_br_off.p; P.br_len = d_br_len.p; P.br_rows = d_br_rows.p;
      P.rs_base = gas_redissolution ? nullptr : s_rs.p;
      P.swof = d_swof.p; P.sgof = d_sgof.p;
    } else {
      P.sat = sat_h.data(); P.br_rs = br_rs_h.data(); P.br_off = br_off_h.data(); P.br_len = br_len_h.data();
      P.br_rows = br_rows_h.data();
      P.swof = swof_h.data(); P.sgof = sgof_h.data();
    }
    return P;
  }
  PhysParams host_params() const { return params(false); }

  // ---- wells ----
  void build_topology() {  // rs/fim/wells.rs:544-594
    topo.clear();
    perf_well.assign(wells.size(), -1);
    std::vector<std::string> keys;
    for (size_t e = 0; e < wells.size(); ++e) {
      const auto& w = wells[e];
      std::string key;
      if (!w.id.empty()) {
        key = "id:" + w.id;
      } else {
        char buf[256];
        snprintf(buf, sizeof buf, "fp:%d:%d:%d:%a:%a:%a", int(w.injector), w.i, w.j, w.bhp, w.rw, w.skin);
        key = buf;
      }
      int g = -1;
      for (size_t k = 0; k < keys.size(); ++k)
        if (keys[k] == key) g = (int)k;
      if (g < 0) {
        g = (int)topo.size();
        keys.push_back(key);
        topo.push_back({(int)e, w.injector, {}});
      }
      perf_well[e] = g;
      topo[g].perfs.push_back((int)e);
    }
    dev_ready = false;
  }

  int fluid_of(int w) const {
    if (!three_phase) return 0;
    const auto& rep = wells[topo[w].rep];
    return rep.fluid >= 0 ? rep.fluid : injected_fluid;
  }

  ControlState control(int w) const {  // rs/fim/wells.rs:691-798
    const auto& cw_ = wells[topo[w].rep];
    bool inj = topo[w].injector, expl = cw_.schedule.has;
    ControlState c{};
    c.enabled = expl ? cw_.schedule.enabled : (inj ? inj_enabled : true);
    bool fam = inj ? (inj_rate && inj_enabled) : prod_rate;
    c.rate = !c.enabled ? false : (expl ? cw_.schedule.rate : fam);
    c.surface = expl ? cw_.schedule.has_surface : (inj ? has_tgt_inj_surf : has_tgt_prod_surf);
    c.has_target = false;
    c.target = 0;
    if (c.rate) {
      if (expl) {
        double t = cw_.schedule.has_surface ? cw_.schedule.surface : (cw_.schedule.has_target ? cw_.schedule.target : 0.0);
        c.has_target = true;
        c.target = std::max(t, 0.0);
      } else if (fam) {
        c.has_target = true;
        c.target = inj ? std::max(has_tgt_inj_surf ? tgt_inj_surf : tgt_inj, 0.0)
                       : std::max(has_tgt_prod_surf ? tgt_prod_surf : tgt_prod, 0.0);
      }
    }
    double fl = inj ? bhp_max : bhp_min;
    c.limit = expl ? (cw_.schedule.has_limit ? cw_.schedule.limit : fl) : fl;
    c.bhp_target = cw_.bhp;
    return c;
  }

  bool well_index(int e, double& wi) const {  // rs/fim/wells.rs:800-826 Peaceman
    const auto& w = wells[e];
    int id = idx(w.i, w.j, w.k);
    double a = kx[id], b = ky[id];
    if (!std::isfinite(a) || !std::isfinite(b) || a <= 0 || b <= 0) return false;
    double req = 0.28 * std::sqrt(std::sqrt(a / b) * dx * dx + std::sqrt(b / a) * dy * dy) /
                 (std::pow(a / b, 0.25) + std::pow(b / a, 0.25));
    if (!std::isfinite(req) || req <= w.rw) return false;
    double ka = std::sqrt(a * b), den = std::log(req / w.rw) + w.skin;
    if (!std::isfinite(ka) || ka <= 0 || !std::isfinite(den) || std::fabs(den) <= kEps) return false;
    wi = kDarcy * 2.0 * M_PI * ka * dz[w.k] / den;
    return true;
  }

  // host evaluation helpers for one cell iterate
  struct CellIt {
    double p, sw, hc;
    int reg;
    double rs_cap;  // stored Rs when gas redissolution is disabled, else infinity
  };
  double conn_rate(int e, const CellIt& c, double bhp) const {  // rs/fim/wells.rs:1098-1127
    double wi;
    if (!well_index(e, wi)) return 0.0;
    PhysParams P = host_params();
    Fluids<double> f = cell_fluids<double>(P, c.reg, c.p, c.sw, c.hc, c.rs_cap);
    Mobility<double> m = mobilities<double>(P, c.sw, f.sg, c.p, f.rs);
    double cm = std::max(m.w + m.o + m.g, 0.0);
    double raw = wi * cm * (c.p - (bhp + wells[e].head));
    if (!std::isfinite(raw)) return 0.0;
    return wells[e].injector ? std::min(raw, 0.0) : std::max(raw, 0.0);
  }
  double surface_rate(int w, const CellIt& c, double q) const {  // rs/fim/wells.rs:1147-1169
    PhysParams P = host_params();
    Fluids<double> f = cell_fluids<double>(P, c.reg, c.p, c.sw, c.hc, c.rs_cap);
    if (topo[w].injector) {
      if (fluid_of(w) == 0) return std::max(-q, 0.0) * inv_bw<double>(P, c.p);
      return std::max(-q, 0.0) / std::max(f.bg, 1e-9);
    }
    Mobility<double> m = mobilities<double>(P, c.sw, f.sg, c.p, f.rs);
    double lw = std::max(m.w, 0.0), lo = std::max(m.o, 0.0), lg = std::max(m.g, 0.0);
    double lt = std::max(lw + lo + lg, kEps);
    double fo = std::min(std::max(lo / lt, 0.0), 1.0);
    return std::max(q, 0.0) * fo / std::max(f.bo, 1e-9);
  }
  double total_rate_for_bhp(int w, const std::vector<CellIt>& cells, double bhp) const {  // :389-408
    ControlState c = control(w);
    bool inj = topo[w].injector;
    double tot = 0;
    for (int e : topo[w].perfs) {
      double wi;
      if (!well_index(e, wi)) continue;
      double q = conn_rate(e, cells[e], bhp);
      if (c.surface) tot += surface_rate(w, cells[e], q);
      else tot += inj ? std::max(-q, 0.0) : std::max(q, 0.0);
    }
    return tot;
  }
  bool bhp_from_target(int w, const std::vector<CellIt>& cells, double& out) const {  // :431-504
    ControlState c = control(w);
    if (!c.enabled || !c.rate || !c.has_target || topo[w].perfs.empty()) return false;
    double mn = INFINITY, mx = -INFINITY;
    for (int e : topo[w].perfs) {
      double z = cells[e].p - wells[e].head;
      mn = std::min(mn, z);
      mx = std::max(mx, z);
    }
    if (!std::isfinite(mn) || !std::isfinite(mx)) return false;
    double maxr = total_rate_for_bhp(w, cells, c.limit);
    if (c.target >= maxr - 1e-9) {
      out = c.limit;
      return true;
    }
    bool inj = topo[w].injector;
    double lo = inj ? std::min(mn, c.limit) : c.limit, hi = inj ? c.limit : std::max(mx, c.limit);
    for (int it = 0; it < 64; ++it) {
      double mid = 0.5 * (lo + hi);
      double r = total_rate_for_bhp(w, cells, mid);
      if (inj) { if (r < c.target) lo = mid; else hi = mid; }
      else { if (r > c.target) lo = mid; else hi = mid; }
    }
    out = 0.5 * (lo + hi);
    return true;
  }

  std::vector<int> dperf_order;  // device perforation slot -> completion index
  std::vector<double> last_bhp, last_q;

  // ---- discrete adjoint: recording of the forward substeps, backward pass ----
  struct StoredState {  // everything a substep starts from
    std::vector<double> p, sw, so, sg, rs, bhp, head;
    double time = 0.0;
  };
  struct Iterate {  // converged Newton iterate of a substep
    std::vector<double> p, sw, hc, bhp, q;  // bhp per well, q in completion order
    std::vector<int> reg;
  };
  struct AdjointRecord {
    bool on = false;
    int interval = 10;
    std::vector<double> dts, jn;         // accepted substeps and their objective terms
    std::map<int, StoredState> ckpt;     // stored state before substep k, every `interval`
    std::map<int, Iterate> iter;         // interval 1: the converged iterate of every substep too
  } adj;
  void save_iterate(Iterate& it, const std::vector<double>& bhp, const std::vector<double>& q);
  DevBuf<double> d_phi, d_kx, d_ky, d_kz, d_dz;  // fields behind the parameter derivatives
  DevBuf<double> A2, E2, adj_lam, adj_g_lnk, adj_g_phi, adj_djdx, adj_jv, adj_djdq, adj_d3;
  DevBuf<int> x_tperm;  // extras entry (r, c) -> entry (c, r)
  bool cpr_transposed = false;  // A holds J^T: precondition with the transposed forward CPR
  void save_state(StoredState& st);
  void restore_state(const StoredState& st);
  void upload_bhp_targets();
  double objective_substep(double dt);
  // recording hooks around every substep attempt / accepted substep
  void adjoint_before_substep() {
    if (!adj.on) return;
    int k = (int)adj.dts.size();
    if (k % adj.interval == 0 && !adj.ckpt.count(k)) save_state(adj.ckpt[k]);
  }
  void adjoint_after_substep(double dt, const std::vector<double>& bhp, const std::vector<double>& q) {
    if (!adj.on) return;
    // with a checkpoint before every substep the backward pass needs no forward replay
    // once the converged iterates are kept as well
    if (adj.interval == 1) save_iterate(adj.iter[(int)adj.dts.size()], bhp, q);
    adj.dts.push_back(dt);
    adj.jn.push_back(objective_substep(dt));
  }
  struct AdjointTimes {
    long linear_its = 0, pressure_its = 0;
    double replay_ms = 0.0, solve_ms = 0.0, total_ms = 0.0;
  };
  void adjoint_gradient(std::vector<double>& g_lnk, std::vector<double>& g_phi, AdjointTimes& tm);
  std::vector<CellIt> gather_wells();
  WellDev well_dev();
  void upload_wells(const std::vector<double>& bhp, const std::vector<double>& q);
  void refresh_head_offsets();
  // bring the device up to date with the model, keeping a newer device state
  void ensure_device() {
    if (state_dirty_host && s_p.p) download_state();
    if (!dev_ready) setup_device();
  }
  void setup_device();
  void setup_wells_device();
  void upload_state();
  void download_state();
  void init_substep(std::vector<double>& bhp, std::vector<double>& q);
  void assemble(double dt, bool apply_schur);
  template <class PT> bool linear_solve(double rhs_norm_full, int& iters, double& final_res);
  template <class PT> void cpr_build();
  template <class PT> void cpr_apply(const double* v, PT* z);
  template <class PT> void bicgstab(PT* x, const PT* bvec, int& its);
  template <class PT> void bicgstab_rb(PT* x, const PT* bvec, int& its);
  int pressure_its_hint = 0;  // iterations of the last pressure solve (sizes the first chunk)
  void set_extras_pattern(const std::vector<std::pair<int, int>>& rc);
  // step reporting (rs/reporting.rs:553-734): one entry per accepted substep, and the
  // cumulative balances behind it; off while the adjoint replays recorded substeps
  std::vector<TimePoint> rate_history;
  double cum_mb_water = 0.0, cum_mb_oil = 0.0, cum_mb_gas = 0.0, cum_injection_m3 = 0.0, cum_production_m3 = 0.0;
  bool report_on = true;
  SweepConfig sweep;
  void report_sums(double out[7]);
  void sweep_metrics(TimePoint& tp);
  void commit_substep(double dt, const std::vector<double>& bhp, const std::vector<double>& q);
  void record_report(double dt, const std::vector<double>& bhp, const std::vector<double>& q, double d_water,
                     double oil_removed, double d_gas, const double sums[7]);
  // nested per-well Newton (rs/fim/wells_inner.rs) in place of the well relaxation, with
  // every well's own convergence test in the acceptance (setFimNestedWellSolve)
  bool nested_wells = false;
  DevBuf<int> w_inner_i;
  DevBuf<double> w_inner_peak;
  std::vector<int> inner_i;        // per well: converged, iterations of the last inner solve
  std::vector<double> inner_peak;  // per well: scaled residual peak of the last inner solve
  void solve_wells_locally(std::vector<double>& bhp, std::vector<double>& q, int max_it);
  DevBuf<double> rec;  // Newton record (kRec* layout)
  void newton_record(double dt, double* rec_h);
  bool opm_verdict(const double* rec_h, bool relax_final) const;
  // linear-system capture: 0 off, 1 every solve, 2 solves that do not converge
  std::string capture_dir;
  int capture_mode = 0;
  long capture_seq = 0;
  std::vector<std::pair<int, int>> extras_rc;  // host copy of the extras pattern
  void write_capture(int newton_iteration, const char* reason);
  void write_capture_file(int newton_iteration, const char* reason);
  void solve_captured(int rows, const int* ptr, const int* col, const double* val, const double* rhs, double* x,
                      int& iters, double& res, double& rhs_norm, bool& conv);
  template <class PT> void pilu_apply(const PT* r, PT* out, const double* skip = nullptr);
  template <class PT> void bilu_apply(const PT* r, PT* out, const PT* Am);
  template <int NC, class PT> void factor_sweep(const double* Am, const PT* Pm, PT* dinv);
  template <int NC, class PT> void sweep_pack(const PT* M, const PT* dinv, PT* sl);
  template <int NC, class PT> void sweep_solve(const PT* sl, const PT* r, PT* out, const double* skip);
  void sweep_check();
  double dev_dot(const double* a, const double* b2, int len);
  template <class X, class Y> double dev_dot_t(const X* a, const Y* b2, int len);
  struct SubstepResult {
    bool converged;
    int newton;
    long linear;
    double scaled;
    // a rejected attempt's RSF_RETRY_* class, RSF_FAMILY_* dominant family and the row of
    // its peak in the reference's unknown numbering
    int fail_class = -1, fail_family = -1;
    long fail_row = -1;
    int lin_solves = 0, last_lin_its = -1, last_lin_conv = -1;
    double lin_ms = 0.0;
    double mb_norm = NAN;  // largest MB measure at the last iterate checked
    int updates = 0;       // Newton corrections applied to the iterate
    double asm_ms = 0.0;
  };
  SubstepResult run_timestep(double dt, std::vector<double>& bhp, std::vector<double>& q);
  int step(double target);
  // One substep from the stored state; an initial iterate other than the stored state's
  // own may be given (cell primaries interleaved [p, sw, hc], regimes, BHP per well, q per
  // completion; each may be null)
  int run_substep(double dt, SubstepResult& out, const double* init_cells = nullptr, const int* init_regimes = nullptr,
                  const double* init_bhp = nullptr, const double* init_q = nullptr);
  void write_back(const std::vector<double>& bhp);
  double rel_change();
  void sync() { RSF_CHECK(hipStreamSynchronize(stream)); }
};

}  // namespace rsf

// the C ABI's simulator handle (include/ressim_fim.h)
struct rsf_sim {
  rsf::Sim* s;
};
