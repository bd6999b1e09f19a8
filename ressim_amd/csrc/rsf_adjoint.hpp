// Discrete adjoint of the FIM timestep: host side (included into rsf_lib.hip, inside
// namespace rsf). The forward pass records every accepted substep (dt and its
// objective term) and stores the state before every `interval`-th substep. The
// backward pass replays each segment from its checkpoint, keeps the states and
// converged iterates of the segment, and walks it backwards:
//   1. reassemble the Jacobian at (x_m; state before m) with the wells eliminated,
//   2. b = -dj/dx + (d inv(x_m)/dx)^T lambda_{m+1} on the cells, -dj/d(bhp, q) on the wells,
//   3. solve J^T lambda_m = b: with J = [[A, B], [C, D]] and S = A - B D^-1 C,
//        S^T lambda_c = b_c - C^T D^-T b_w,   lambda_w = D^-T (b_w - B^T lambda_c),
//      where D (one well: [[alpha, beta^T], [gamma, I]]) has a closed-form inverse
//      and S^T is the transposed stencil store plus transposed clique extras,
//   4. dJ/dtheta += lambda_m^T dR_m/dtheta (porosity through the inventories,
//      log-permeability through the transmissibilities and the Peaceman indices).
// The wells' lagged datum heads are held fixed (they vanish without gravity) and
// a well's committed BHP carried into the next substep has no effect on the
// gradient (a BHP-controlled well's multiplier only meets its own BHP column).

void Sim::save_state(StoredState& st) {
  download_state();
  st.p = h_p;
  st.sw = h_sw;
  st.so = h_so;
  st.sg = h_sg;
  st.rs = h_rs;
  st.bhp.resize(wells.size());
  st.head.resize(wells.size());
  for (size_t e = 0; e < wells.size(); ++e) {
    st.bhp[e] = wells[e].bhp;
    st.head[e] = wells[e].head;
  }
  st.time = time;
}

void Sim::save_iterate(Iterate& it, const std::vector<double>& bhp, const std::vector<double>& q) {
  it.p.resize(n);
  it.sw.resize(n);
  it.hc.resize(n);
  it.reg.resize(n);
  x_p.download(it.p.data(), n, stream);
  x_sw.download(it.sw.data(), n, stream);
  x_hc.download(it.hc.data(), n, stream);
  x_reg.download(it.reg.data(), n, stream);
  sync();
  it.bhp = bhp;
  it.q = q;
}

void Sim::restore_state(const StoredState& st) {
  h_p = st.p;
  h_sw = st.sw;
  h_so = st.so;
  h_sg = st.sg;
  h_rs = st.rs;
  upload_state();
  for (size_t e = 0; e < wells.size(); ++e) {
    wells[e].bhp = st.bhp[e];
    wells[e].head = st.head[e];
  }
  if (!dperf_order.empty()) {
    std::vector<double> head;
    for (int e : dperf_order) head.push_back(wells[e].head);
    w_head.upload(head, stream);
  }
  upload_bhp_targets();
  time = st.time;
}

// objective term of the substep just converged: surface oil of the producers
double Sim::objective_substep(double dt) {
  int Pn = (int)dperf_order.size();
  if (!Pn) return 0.0;
  adj_jv.alloc(Pn);
  adj_djdq.alloc(Pn);
  adj_djdx.alloc(3 * (size_t)n);
  hipLaunchKernelGGL((k_fill<double>), dim3(nblk(3L * n)), dim3(kBlock), 0, stream, 3 * n, adj_djdx.p, 0.0);
  PhysParams P = params(true);
  hipLaunchKernelGGL(k_adj_objective, dim3(nblk(Pn)), dim3(kBlock), 0, stream, P, n, dt, well_dev(), x_p.p, x_sw.p,
                     x_hc.p, x_reg.p, adj_jv.p, adj_djdq.p, adj_djdx.p);
  std::vector<double> jv(Pn);
  adj_jv.download(jv.data(), Pn, stream);
  sync();
  double j = 0.0;
  for (double v : jv) j += v;
  return j;
}

namespace {
// y = D^-T r for one well's tail (rows: control, perforations), D = [[alpha, beta^T], [gamma, I]]:
// y_b = (r_b - gamma.r_q) / s, y_q = r_q - beta y_b, s = alpha - beta.gamma
// (singular s: the control row alone, as the forward recovery does)
void tail_solve_transposed(double alpha, double s, int np, const double* beta, const double* gamma, double r_b,
                           const double* r_q, double& y_b, double* y_q) {
  bool singular = !(std::fabs(s) > 0.0) || !std::isfinite(s);
  double gr = 0.0;
  for (int t = 0; t < np; ++t) gr += gamma[t] * r_q[t];
  y_b = singular ? (std::fabs(alpha) > kEps ? r_b / alpha : r_b) : (r_b - gr) / s;
  for (int t = 0; t < np; ++t) y_q[t] = r_q[t] - (singular ? 0.0 : beta[t] * y_b);
}
}  // namespace

void Sim::adjoint_gradient(std::vector<double>& g_lnk_h, std::vector<double>& g_phi_h, AdjointTimes& tm) {
  using clk = std::chrono::steady_clock;
  auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  const auto t_start = clk::now();
  tm = AdjointTimes{};
  const long bicg0 = stats.bicgstab;
  const int N = (int)adj.dts.size();
  const int W = (int)topo.size(), Pn = (int)dperf_order.size();
  g_lnk_h.assign(n, 0.0);
  g_phi_h.assign(n, 0.0);
  if (N == 0) return;
  StoredState end_state;
  save_state(end_state);
  const bool was_on = adj.on;
  adj.on = false;
  const Options saved = opt;
  // transposed solves to a tight relative tolerance, 1e-10 or the configured linear
  // tolerance when that is tighter (FP64 outer FGMRES, the configured preconditioner
  // precision); the forward replay keeps the forward options
  Options adj_opt = opt;
  adj_opt.lin_rel_tol = std::min(1e-10, opt.lin_rel_tol);
  adj_opt.lin_abs_tol = 0.0;
  adj_opt.lin_max_it = 400;
  adj_opt.restart = 60;
  adj_lam.alloc(3 * (size_t)n);
  adj_g_lnk.alloc(n);
  adj_g_phi.alloc(n);
  adj_d3.alloc(3 * std::max((size_t)Pn, (size_t)1));
  A2.alloc(63 * (size_t)n);
  E2.alloc(9 * std::max((size_t)n_extra, (size_t)1));
  hipLaunchKernelGGL((k_fill<double>), dim3(nblk(n)), dim3(kBlock), 0, stream, n, adj_g_lnk.p, 0.0);
  hipLaunchKernelGGL((k_fill<double>), dim3(nblk(n)), dim3(kBlock), 0, stream, n, adj_g_phi.p, 0.0);
  PhysParams P = params(true);
  std::vector<double> g_wells(n, 0.0);  // Peaceman-index part, summed on the host
  bool have_next = false;               // adj_lam holds lambda_{m+1} (cells)
  // checkpoints in decreasing order
  std::vector<int> starts;
  for (auto& kv : adj.ckpt) starts.push_back(kv.first);
  std::sort(starts.begin(), starts.end());
  int seg_end = N;
  for (int si = (int)starts.size() - 1; si >= 0; --si) {
    const int c0 = starts[si];
    if (c0 >= seg_end) continue;
    // forward replay of the segment: states before each substep and converged iterates
    std::vector<StoredState> pre(seg_end - c0);
    std::vector<Iterate> its(seg_end - c0);
    const auto t_rep = clk::now();
    if (seg_end - c0 == 1 && adj.iter.count(c0)) {  // kept by the forward pass: no replay
      pre[0] = adj.ckpt.at(c0);
      its[0] = adj.iter.at(c0);
    } else {
    restore_state(adj.ckpt.at(c0));
    // replays of recorded substeps are not new steps: no report entries, also on a throw
    struct ReportOff {
      bool& f;
      explicit ReportOff(bool& x) : f(x) { f = false; }
      ~ReportOff() { f = true; }
    } report_off(report_on);
    for (int m = c0; m < seg_end; ++m) {
      save_state(pre[m - c0]);
      SubstepResult r;
      if (run_substep(adj.dts[m], r) != 0)
        throw std::runtime_error("adjoint: replay of a recorded substep did not converge");
      Iterate& it = its[m - c0];
      it.p.resize(n);
      it.sw.resize(n);
      it.hc.resize(n);
      it.reg.resize(n);
      x_p.download(it.p.data(), n, stream);
      x_sw.download(it.sw.data(), n, stream);
      x_hc.download(it.hc.data(), n, stream);
      x_reg.download(it.reg.data(), n, stream);
      sync();
      it.bhp = last_bhp;
      it.q = last_q;
    }
    }
    tm.replay_ms += ms(t_rep, clk::now());
    // backward through the segment
    for (int m = seg_end - 1; m >= c0; --m) {
      const double dt = adj.dts[m];
      const Iterate& it = its[m - c0];
      restore_state(pre[m - c0]);
      std::vector<double> bhp0, q0;
      init_substep(bhp0, q0);  // acc0 and the frozen regimes of substep m
      x_p.upload(it.p, stream);
      x_sw.upload(it.sw, stream);
      x_hc.upload(it.hc, stream);
      x_reg.upload(it.reg, stream);
      upload_wells(it.bhp, it.q);
      assemble(dt, true);
      objective_substep(dt);  // dj/dx (adj_djdx), dj/dq (adj_djdq)
      // cell right-hand side
      hipLaunchKernelGGL(k_adj_rhs, dim3(nblk(n)), dim3(kBlock), 0, stream, P, n, d_pv.p, x_p.p, x_sw.p, x_hc.p,
                         x_reg.p, have_next ? (const double*)adj_lam.p : nullptr, (const double*)adj_djdx.p, b.p);
      // well tail data of this Jacobian
      std::vector<double> wk(16 * (size_t)std::max(Pn, 1)), ww(4 * (size_t)std::max(W, 1)), djdq(std::max(Pn, 1));
      std::vector<int> off(W + 1);
      if (Pn) {
        w_wk.download(wk.data(), 16 * (size_t)Pn, stream);
        w_well.download(ww.data(), 4 * (size_t)W, stream);
        adj_djdq.download(djdq.data(), Pn, stream);
        w_off.download(off.data(), W + 1, stream);
      }
      sync();
      // b_c -= C^T D^-T b_w  (b_w: 0 on the control rows, -dj/dq on the perforation rows)
      std::vector<double> yb(W, 0.0), yq(Pn, 0.0), d3(3 * (size_t)std::max(Pn, 1), 0.0);
      for (int w = 0; w < W; ++w) {
        int p0 = off[w], np = off[w + 1] - p0;
        std::vector<double> beta(np), gamma(np), rq(np);
        for (int t = 0; t < np; ++t) {
          beta[t] = wk[16 * (p0 + t) + 11];
          gamma[t] = wk[16 * (p0 + t) + 9];
          rq[t] = -djdq[p0 + t];
        }
        tail_solve_transposed(ww[4 * w], ww[4 * w + 1], np, beta.data(), gamma.data(), 0.0, rq.data(), yb[w],
                              yq.data() + p0);
        for (int t = 0; t < np; ++t) {
          const double* e = &wk[16 * (p0 + t)];
          for (int v = 0; v < 3; ++v) d3[3 * (p0 + t) + v] = -(e[3 + v] * yq[p0 + t] + e[6 + v] * yb[w]);
        }
      }
      if (Pn) {
        adj_d3.upload(d3, stream);
        hipLaunchKernelGGL(k_adj_scatter3, dim3(nblk(Pn)), dim3(kBlock), 0, stream, Pn, n, w_cell.p, adj_d3.p, b.p);
      }
      // S^T: transposed stencil store and clique extras in place of S
      hipLaunchKernelGGL(k_transpose_blocks, dim3(nblk(n)), dim3(kBlock), 0, stream, grid, A.p, A2.p);
      if (n_extra > 0)
        hipLaunchKernelGGL(k_transpose_extras, dim3(nblk(n_extra)), dim3(kBlock), 0, stream, n_extra, x_tperm.p,
                           w_extra.p, E2.p);
      std::swap(A.p, A2.p);
      std::swap(A.n, A2.n);
      std::swap(w_extra.p, E2.p);
      std::swap(w_extra.n, E2.n);
      const auto t_sol = clk::now();
      double bn = std::sqrt(dev_dot(b.p, b.p, 3 * n));
      int lits = 0;
      double fr = 0.0;
      cpr_transposed = true;
      opt = adj_opt;
      if (opt.precision == 1) {
        cpr_build<float>();
        linear_solve<float>(bn, lits, fr);
      } else {
        cpr_build<double>();
        linear_solve<double>(bn, lits, fr);
      }
      cpr_transposed = false;
      opt = saved;
      sync();
      tm.solve_ms += ms(t_sol, clk::now());
      std::swap(A.p, A2.p);
      std::swap(A.n, A2.n);
      std::swap(w_extra.p, E2.p);
      std::swap(w_extra.n, E2.n);
      tm.linear_its += lits;
      RSF_CHECK(hipMemcpyAsync(adj_lam.p, dxv.p, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, stream));
      have_next = true;
      // parameter terms of lambda_m^T dR_m/dtheta
      hipLaunchKernelGGL(k_adj_grad_phi, dim3(nblk(n)), dim3(kBlock), 0, stream, P, n, d_pv.p, d_phi.p, x_p.p,
                         x_sw.p, x_hc.p, x_reg.p, acc0.p, adj_lam.p, adj_g_phi.p);
      hipLaunchKernelGGL(k_adj_grad_lnk, dim3(nblk(n)), dim3(kBlock), 0, stream, grid, (int)gravity, dt, d_Tx.p,
                         d_Ty.p, d_Tz.p, d_depth.p, x_p.p, pack.p, d_kx.p, d_ky.p, d_kz.p, d_dz.p, adj_lam.p,
                         adj_g_lnk.p);
      // wells: lambda_w = D^-T (b_w - B^T lambda_c); the perforation rows q - WI lambda_t dp
      // scale with the Peaceman index, d/d ln k = -connection rate
      if (Pn) {
        std::vector<double> lc(5 * (size_t)Pn);
        hipLaunchKernelGGL(k_gather4, dim3(nblk(Pn)), dim3(kBlock), 0, stream, Pn, w_cell.p, adj_lam.p,
                           adj_lam.p + n, adj_lam.p + 2 * (size_t)n, adj_lam.p, d_gather.p);
        d_gather.download(lc.data(), 4 * (size_t)Pn, stream);  // k_gather4 writes 4 per completion
        sync();
        for (int w = 0; w < W; ++w) {
          int p0 = off[w], np = off[w + 1] - p0;
          std::vector<double> beta(np), gamma(np), rq(np), lq(np);
          for (int t = 0; t < np; ++t) {
            const double* e = &wk[16 * (p0 + t)];
            beta[t] = e[11];
            gamma[t] = e[9];
            double bt = e[0] * lc[4 * (p0 + t)] + e[1] * lc[4 * (p0 + t) + 1] + e[2] * lc[4 * (p0 + t) + 2];
            rq[t] = -djdq[p0 + t] - bt;
          }
          double lb;
          tail_solve_transposed(ww[4 * w], ww[4 * w + 1], np, beta.data(), gamma.data(), 0.0, rq.data(), lb,
                                lq.data());
          for (int t = 0; t < np; ++t) {
            int cell = perf_cells_h[p0 + t];
            g_wells[cell] += lq[t] * (-wk[16 * (p0 + t) + 12]);
          }
        }
      }
    }
    seg_end = c0;
  }
  adj_g_lnk.download(g_lnk_h.data(), n, stream);
  adj_g_phi.download(g_phi_h.data(), n, stream);
  sync();
  for (int i = 0; i < n; ++i) g_lnk_h[i] += g_wells[i];
  opt = saved;
  restore_state(end_state);
  adj.on = was_on;
  tm.pressure_its = stats.bicgstab - bicg0;
  tm.total_ms = ms(t_start, clk::now());
}
