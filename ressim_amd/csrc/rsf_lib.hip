// HIP FIM timestep library: simulator driver and the C ABI declared in
// include/ressim_fim.h.
#include "../../include/ressim_fim.h"
#include "rsf_sim.hpp"
#include "rsf_adjoint_kernels.hpp"
#include "rsf_dense.hpp"

#include <hip/hip_ext.h>

#include <cfloat>

namespace rsf {

// ------------------------------------------------------------------ device setup
void Sim::setup_device() {
  if (!stream) RSF_CHECK(hipStreamCreate(&stream));
  grid = Grid{nx, ny, nz, n};
  std::vector<double> Tx(n, 0.0), Ty(n, 0.0), Tz(n, 0.0), pv(n), depth(nz);
  auto hm = [](double a, double b) { return (a + b == 0.0) ? 0.0 : 2.0 * a * b / (a + b); };
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        int id = idx(i, j, k);
        // rs/grid.rs:278-311 geometric transmissibility, times the Darcy factor
        if (i + 1 < nx) Tx[id] = kDarcy * hm(kx[id], kx[id + 1]) * (dy * dz[k]) / dx;
        if (j + 1 < ny) Ty[id] = kDarcy * hm(ky[id], ky[id + nx]) * (dx * dz[k]) / dy;
        if (k + 1 < nz) {
          int up = id + nx * ny;
          double den = kz[up] * dz[k] + kz[id] * dz[k + 1];
          Tz[id] = den <= 0.0 ? 0.0 : kDarcy * (2.0 * kz[id] * kz[up] * (dx * dy) / den);
        }
        pv[id] = pore_volume(id);
      }
  for (int k = 0; k < nz; ++k) depth[k] = depth_at_k(k);
  d_Tx.upload(Tx, stream);
  d_Ty.upload(Ty, stream);
  d_Tz.upload(Tz, stream);
  d_depth.upload(depth, stream);
  d_pv.upload(pv, stream);
  d_pcs.upload(pc_scale, stream);
  d_phi.upload(phi, stream);
  d_kx.upload(kx, stream);
  d_ky.upload(ky, stream);
  d_kz.upload(kz, stream);
  d_dz.upload(dz, stream);
  std::vector<double> one{0.0};
  std::vector<int> ione{0};
  d_sat.upload(has_table ? sat_h : one, stream);
  d_br_rs.upload(has_table ? br_rs_h : one, stream);
  d_br_rows.upload(has_table ? br_rows_h : one, stream);
  d_br_off.upload(has_table ? br_off_h : ione, stream);
  d_br_len.upload(has_table ? br_len_h : ione, stream);
  d_swof.upload(swof_h.empty() ? one : swof_h, stream);
  d_sgof.upload(sgof_h.empty() ? one : sgof_h, stream);
  for (auto* b_ : {&s_p, &s_sw, &s_so, &s_sg, &s_rs, &x_p, &x_sw, &x_hc, &x0_p, &x0_sw, &x0_hc}) b_->alloc(n);
  x_reg.alloc(n);
  x0_reg.alloc(n);
  acc0.alloc(3 * (size_t)n);
  R.alloc(3 * (size_t)n);
  A.alloc(63 * (size_t)n);
  pack.alloc((size_t)kPack * n);
  b.alloc(3 * (size_t)n);
  dxv.alloc(3 * (size_t)n);
  tmp.alloc(3 * (size_t)n);
  cand.alloc(3 * (size_t)n);
  cw.alloc(3 * (size_t)n);
  dscal.alloc(64);
  rec.alloc(kRecCount);
  part.alloc(1024 * 16);
  bpart.alloc(10 * (size_t)kPartStride);
  bsc.alloc(kBsCount);
  // level schedules: natural hyperplanes x+y+z and red-black parity
  for (int mode = 0; mode < 2; ++mode) {
    int L = mode == 0 ? nx + ny + nz - 2 : 2;
    std::vector<int> cnt(L + 1, 0), rows(n);
    auto lv = [&](int id) {
      int z = id / (nx * ny), r = id % (nx * ny), y = r / nx, x = r % nx;
      return mode == 0 ? x + y + z : ((x + y + z) & 1);
    };
    for (int id = 0; id < n; ++id) cnt[lv(id) + 1]++;
    for (int l = 0; l < L; ++l) cnt[l + 1] += cnt[l];
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int id = 0; id < n; ++id) rows[pos[lv(id)]++] = id;
    lvl_ptr[mode] = cnt;
    lvl_rows[mode].upload(rows, stream);
  }
  setup_wells_device();
  upload_state();
  dev_ready = true;
}

// Off-stencil cell couplings ("extras") of the matrix the linear stack solves: (row,
// col) cell pairs sorted by row then column, a symmetric pattern. Builds the CSR over
// the rows that have extras, the per-wavefront row maps of the pressure kernels, the
// colour-compacted copies, and the transpose map, and sizes the value store (9 doubles
// per entry).
void Sim::set_extras_pattern(const std::vector<std::pair<int, int>>& rc) {
  extras_rc = rc;
  n_extra = (int)rc.size();
  std::vector<int> zero{0};
  std::vector<int> xrows, xptr{0}, xcol;
  for (int k = 0; k < n_extra; ++k) {
    if (xrows.empty() || xrows.back() != rc[k].first) {
      if (!xrows.empty()) xptr.push_back(k);
      xrows.push_back(rc[k].first);
    }
    xcol.push_back(rc[k].second);
  }
  if (!xrows.empty()) xptr.push_back(n_extra);
  n_extra_rows = (int)xrows.size();
  w_extra.alloc(9 * std::max((size_t)n_extra, (size_t)1));
  x_rows.upload(xrows.empty() ? zero : xrows, stream);
  x_ptr.upload(xptr.size() < 2 ? std::vector<int>{0, 0} : xptr, stream);
  x_col.upload(xcol.empty() ? zero : xcol, stream);
  // per wavefront of the pressure SpMV grid: first extras row at or after its first cell
  int nwave = (n + kRBlock * kRUnroll + 63) / 64;
  std::vector<int> xwave(nwave + 1);
  for (int wv = 0; wv <= nwave; ++wv)
    xwave[wv] = int(std::lower_bound(xrows.begin(), xrows.end(), wv * 64) - xrows.begin());
  x_wave.upload(xwave, stream);
  // per wavefront of the red-black pressure kernels (up to 64 cell pairs of one grid
  // row, nx even): first extras row at or after the wavefront's first cell
  if (nx % 2 == 0) {
    const int nxh = nx / 2, wpr = (nxh + 63) / 64;
    n_wave2 = wpr * ny * nz;
    std::vector<int> xwave2(n_wave2 + 1);
    for (int wv = 0; wv <= n_wave2; ++wv) {
      const int row = wv / wpr, seg = wv % wpr;
      const long first_cell = 2L * ((long)row * nxh + std::min(seg * 64, nxh));
      xwave2[wv] = int(std::lower_bound(xrows.begin(), xrows.end(), first_cell) - xrows.begin());
    }
    x_wave2.upload(xwave2, stream);
  } else {
    n_wave2 = 0;
    x_wave2.upload(std::vector<int>{0}, stream);
  }
  // extras rows and columns in the colour-compacted order of the red-black pressure
  // solve (red cell of pair q at q, black at n/2 + q; defined for nx even)
  std::vector<int> xrowc(xrows.size()), xcolc(xcol.size());
  auto cpos = [&](int i) {
    int z = i / (nx * ny), y = (i / nx) % ny, x = i % nx;
    return (i >> 1) + (((x + y + z) & 1) ? n / 2 : 0);
  };
  for (size_t k = 0; k < xrows.size(); ++k) xrowc[k] = cpos(xrows[k]);
  for (size_t k = 0; k < xcol.size(); ++k) xcolc[k] = cpos(xcol[k]);
  x_rows_c.upload(xrowc.empty() ? zero : xrowc, stream);
  x_col_c.upload(xcolc.empty() ? zero : xcolc, stream);
  std::vector<int> erow;
  for (int k = 0; k < n_extra; ++k) erow.push_back(rc[k].first);
  x_erow.upload(erow.empty() ? zero : erow, stream);
  // the pattern is symmetric: entry (r, c) has a partner (c, r)
  std::vector<int> tperm(n_extra, 0);
  {
    std::map<std::pair<int, int>, int> pos;
    for (int k = 0; k < n_extra; ++k) pos[rc[k]] = k;
    for (int k = 0; k < n_extra; ++k) {
      auto it = pos.find({rc[k].second, rc[k].first});
      if (it == pos.end()) throw std::runtime_error("extras pattern is not symmetric");
      tperm[k] = it->second;
    }
  }
  x_tperm.upload(tperm.empty() ? zero : tperm, stream);
}

void Sim::setup_wells_device() {
  int W = (int)topo.size();
  std::vector<int> off(W + 1, 0), cells, flags(W), cdest, coff(W + 1, 0);
  std::vector<double> wi, head, target(W), limit(W), bhpt(W);
  perf_cells_h.clear();
  std::vector<int> order;  // device perforation -> completion index
  for (int w = 0; w < W; ++w) {
    ControlState c = control(w);
    bool fb = c.enabled && c.rate && c.has_target;
    flags[w] = (topo[w].injector ? 1 : 0) | (fb ? 2 : 0) | (c.surface ? 4 : 0) | (fluid_of(w) == 1 ? 8 : 0) |
               (c.enabled ? 16 : 0);
    target[w] = c.target;
    limit[w] = c.limit;
    bhpt[w] = c.bhp_target;
    for (int e : topo[w].perfs) {
      order.push_back(e);
      const auto& cp = wells[e];
      cells.push_back(idx(cp.i, cp.j, cp.k));
      double v = 0.0;
      well_index(e, v);
      wi.push_back(v);
      head.push_back(cp.head);
    }
    off[w + 1] = (int)order.size();
  }
  // clique destinations for rate-controlled (Fischer-Burmeister) wells
  struct Ex { int row, col, w, t, u; };
  std::vector<Ex> ex;
  for (int w = 0; w < W; ++w) {
    int np = off[w + 1] - off[w];
    coff[w + 1] = coff[w] + np * np;
    bool fb = flags[w] & 2;
    for (int t = 0; t < np; ++t)
      for (int u = 0; u < np; ++u) {
        int ci = cells[off[w] + t], cj = cells[off[w] + u];
        int dd = -1;
        if (t != u && fb) {
          int zi = ci / (nx * ny), ri = ci % (nx * ny), yi = ri / nx, xi = ri % nx;
          int d = cj - ci;
          if (d == -1 && xi > 0) dd = 1;
          else if (d == 1 && xi + 1 < nx) dd = 2;
          else if (d == -nx && yi > 0) dd = 3;
          else if (d == nx && yi + 1 < ny) dd = 4;
          else if (d == -nx * ny && zi > 0) dd = 5;
          else if (d == nx * ny && zi + 1 < nz) dd = 6;
          else ex.push_back({ci, cj, w, t, u});
        }
        cdest.push_back(dd);
      }
  }
  std::sort(ex.begin(), ex.end(), [](const Ex& a, const Ex& c) { return a.row != c.row ? a.row < c.row : a.col < c.col; });
  for (size_t k = 0; k < ex.size(); ++k)
    cdest[coff[ex[k].w] + ex[k].t * (off[ex[k].w + 1] - off[ex[k].w]) + ex[k].u] = 7 + (int)k;
  std::vector<std::pair<int, int>> rc;
  for (const auto& e : ex) rc.push_back({e.row, e.col});
  set_extras_pattern(rc);
  perf_cells_h = cells;
  std::vector<int> zero{0};
  std::vector<double> dzero{0.0};
  w_off.upload(off, stream);
  w_cell.upload(cells.empty() ? zero : cells, stream);
  w_flags.upload(flags.empty() ? zero : flags, stream);
  w_pref.upload(order.empty() ? zero : order, stream);
  w_cdest.upload(cdest.empty() ? zero : cdest, stream);
  w_coff.upload(coff, stream);
  w_wi.upload(wi.empty() ? dzero : wi, stream);
  w_head.upload(head.empty() ? dzero : head, stream);
  w_target.upload(target.empty() ? dzero : target, stream);
  w_limit.upload(limit.empty() ? dzero : limit, stream);
  w_bhpt.upload(bhpt.empty() ? dzero : bhpt, stream);
  w_bhp.alloc(std::max(W, 1));
  w_q.alloc(std::max((int)order.size(), 1));
  w_tailR.alloc(std::max(W + (int)order.size(), 1));
  w_wk.alloc(16 * std::max((size_t)order.size(), (size_t)1));
  w_well.alloc(4 * std::max(W, 1));
  w_inner_i.alloc(2 * std::max(W, 1));
  w_inner_peak.alloc(std::max(W, 1));
  d_perf_cells.upload(cells.empty() ? zero : cells, stream);
  d_gather.alloc(5 * std::max((size_t)cells.size(), (size_t)1));
  tdx.alloc(std::max(W + (int)order.size(), 1));
  dperf_order = order;
}

void Sim::upload_state() {
  s_p.upload(h_p, stream);
  s_sw.upload(h_sw, stream);
  s_so.upload(h_so, stream);
  s_sg.upload(h_sg, stream);
  s_rs.upload(h_rs, stream);
  state_on_dev = true;
  state_dirty_host = false;
}

void Sim::download_state() {
  if (!state_dirty_host) return;
  s_p.download(h_p.data(), n, stream);
  s_sw.download(h_sw.data(), n, stream);
  s_so.download(h_so.data(), n, stream);
  s_sg.download(h_sg.data(), n, stream);
  s_rs.download(h_rs.data(), n, stream);
  sync();
  state_dirty_host = false;
}

// rs/well_control.rs:192-280: datum head per completion, constant over a report step
void Sim::refresh_head_offsets() {
  if (!gravity) {
    for (auto& w : wells) w.head = 0.0;
  } else {
    // stored state of completion cells
    std::vector<double> g(4 * wells.size());
    if (!wells.empty()) {
      std::vector<int> cells;
      for (auto& w : wells) cells.push_back(idx(w.i, w.j, w.k));
      DevBuf<int> dc;
      dc.upload(cells, stream);
      DevBuf<double> out;
      out.alloc(g.size());
      hipLaunchKernelGGL(k_gather4, dim3(nblk(cells.size())), dim3(kBlock), 0, stream, (int)cells.size(), dc.p, s_p.p,
                         s_sw.p, s_sg.p, s_rs.p, out.p);
      out.download(g.data(), g.size(), stream);
      sync();
    }
    PhysParams P = host_params();
    std::vector<double> offs(wells.size(), 0.0);
    for (size_t e = 0; e < wells.size(); ++e) {
      int gw = perf_well[e];
      const auto& grp = topo[gw].perfs;
      double datum = INFINITY;
      bool have = false;
      for (int m : grp)
        if (wells[m].has_datum && std::isfinite(wells[m].datum)) { datum = wells[m].datum; have = true; break; }
      if (!have)
        for (int m : grp) datum = std::min(datum, depth_at_k(wells[m].k));
      double dens = -1.0;
      for (int m : grp)
        if (wells[m].has_density && std::isfinite(wells[m].density) && wells[m].density >= 0.0) { dens = wells[m].density; break; }
      if (dens < 0.0) {
        double tot = 0.0, cnt = 0.0;
        for (int m : grp) {
          double p = g[4 * m], sw = g[4 * m + 1], sg = g[4 * m + 2], rs = g[4 * m + 3];
          double d;
          if (wells[m].injector) {
            int fl = three_phase ? (wells[m].fluid >= 0 ? wells[m].fluid : injected_fluid) : 0;
            d = fl == 1 ? rho_gas<double>(P, p) : inv_bw<double>(P, p) * rho_w;
          } else {
            double sg_ = three_phase ? sg : 0.0, rs_ = three_phase ? rs : 0.0;
            Mobility<double> mb = mobilities<double>(P, sw, sg_, p, rs_);
            double lw = std::max(mb.w, 0.0), lo = std::max(mb.o, 0.0), lg = std::max(mb.g, 0.0);
            double lt = std::max(lw + lo + lg, kEps);
            double fw = std::min(std::max(lw / lt, 0.0), 1.0), fo = std::min(std::max(lo / lt, 0.0), 1.0),
                   fg = std::min(std::max(lg / lt, 0.0), 1.0);
            d = fw * inv_bw<double>(P, p) * rho_w + fo * rho_oil<double>(P, p, std::max(rs, 0.0)) +
                fg * rho_gas<double>(P, p);
          }
          if (std::isfinite(d) && d > 0.0) { tot += d; cnt += 1.0; }
        }
        dens = cnt > 0.0 ? tot / cnt : 0.0;
      }
      double o = dens * kGravity * (depth_at_k(wells[e].k) - datum) * 1e-5;
      offs[e] = std::isfinite(o) ? o : 0.0;
    }
    for (size_t e = 0; e < wells.size(); ++e) wells[e].head = offs[e];
  }
  if (dev_ready && !dperf_order.empty()) {
    std::vector<double> head;
    for (int e : dperf_order) head.push_back(wells[e].head);
    w_head.upload(head, stream);
  }
}

// ------------------------------------------------------------------ substep init
void Sim::init_substep(std::vector<double>& bhp, std::vector<double>& q) {
  PhysParams P = params(true);
  hipLaunchKernelGGL(k_classify, dim3(nblk(n)), dim3(kBlock), 0, stream, P, n, s_p.p, s_sw.p, s_sg.p, s_rs.p, x_p.p,
                     x_sw.p, x_hc.p, x_reg.p);
  RSF_CHECK(hipMemcpyAsync(x0_p.p, x_p.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  RSF_CHECK(hipMemcpyAsync(x0_sw.p, x_sw.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  RSF_CHECK(hipMemcpyAsync(x0_hc.p, x_hc.p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  RSF_CHECK(hipMemcpyAsync(x0_reg.p, x_reg.p, n * sizeof(int), hipMemcpyDeviceToDevice, stream));
  hipLaunchKernelGGL(k_inventory, dim3(nblk(n)), dim3(kBlock), 0, stream, P, n, d_pv.p, x0_p.p, x0_sw.p, x0_hc.p,
                     x0_reg.p, acc0.p);
  int W = (int)topo.size();
  bhp.assign(W, 0.0);
  q.assign(wells.size(), 0.0);
  for (int w = 0; w < W; ++w) bhp[w] = control(w).bhp_target;
  if (wells.empty()) return;
  std::vector<CellIt> cells = gather_wells();
  for (int w = 0; w < W; ++w) {
    double v;
    if (bhp_from_target(w, cells, v)) bhp[w] = v;
  }
  for (size_t e = 0; e < wells.size(); ++e) {
    int w = perf_well[e];
    q[e] = control(w).enabled ? conn_rate((int)e, cells[e], bhp[w]) : 0.0;
  }
}

std::vector<Sim::CellIt> Sim::gather_wells() {
  int m = (int)perf_cells_h.size();
  std::vector<double> g(5 * (size_t)m);
  hipLaunchKernelGGL(k_gather_cells, dim3(nblk(m)), dim3(kBlock), 0, stream, m, d_perf_cells.p, x_p.p, x_sw.p,
                     x_hc.p, x_reg.p, s_rs.p, d_gather.p);
  d_gather.download(g.data(), g.size(), stream);
  sync();
  std::vector<CellIt> cells(wells.size());
  for (int t = 0; t < m; ++t)
    cells[dperf_order[t]] = {g[5 * t], g[5 * t + 1], g[5 * t + 2], (int)g[5 * t + 3],
                             gas_redissolution ? INFINITY : g[5 * t + 4]};
  return cells;
}

// ------------------------------------------------------------------ assembly
void Sim::assemble(double dt, bool apply_schur) {
  PhysParams P = params(true);
  hipLaunchKernelGGL(k_cells, dim3(nblk(n)), dim3(kBlock), 0, stream, P, n, d_pv.p, d_pcs.p, x_p.p, x_sw.p, x_hc.p,
                     x_reg.p, acc0.p, R.p, A.p, pack.p);
  hipLaunchKernelGGL(k_face_rows, dim3(nblk(n)), dim3(kBlock), 0, stream, grid, (int)gravity, dt, d_Tx.p, d_Ty.p,
                     d_Tz.p, d_depth.p, x_p.p, pack.p, R.p, A.p);
  int W = (int)topo.size();
  if (W > 0) {
    WellDev wd = well_dev();
    int maxnp = 1;
    for (auto& t : topo) maxnp = std::max(maxnp, (int)t.perfs.size());
    if (12 * (size_t)maxnp * sizeof(double) > 65536)
      throw HipError("a well has more than 682 completions (LDS limit of the well kernels)");
    hipLaunchKernelGGL(k_wells_perf, dim3(W, (maxnp + 63) / 64), dim3(64), 0, stream, P, n, dt, wd, x_p.p, x_sw.p,
                       x_hc.p, x_reg.p, R.p, A.p);
    hipLaunchKernelGGL(k_wells, dim3(W), dim3(kWellBlock), maxnp * sizeof(double), stream, wd);
    if (apply_schur)
      hipLaunchKernelGGL(k_wells_schur, dim3(W, (maxnp * maxnp + kWellBlock - 1) / kWellBlock), dim3(kWellBlock), 0,
                         stream, n, wd, A.p);
  }
  RSF_CHECK(hipGetLastError());
}

WellDev Sim::well_dev() {
  WellDev wd{};
  wd.n_wells = (int)topo.size();
  wd.n_perf = (int)dperf_order.size();
  wd.well_off = w_off.p;
  wd.perf_cell = w_cell.p;
  wd.perf_wi = w_wi.p;
  wd.perf_head = w_head.p;
  wd.perf_ref = w_pref.p;
  wd.well_flags = w_flags.p;
  wd.well_target = w_target.p;
  wd.well_limit = w_limit.p;
  wd.well_bhp_target = w_bhpt.p;
  wd.clique_dest = w_cdest.p;
  wd.clique_off = w_coff.p;
  wd.extra_val = w_extra.p;
  wd.n_extra = std::max(n_extra, 1);
  wd.bhp = w_bhp.p;
  wd.q = w_q.p;
  wd.tail_R = w_tailR.p;
  wd.wk = w_wk.p;
  wd.wwell = w_well.p;
  return wd;
}

void Sim::upload_wells(const std::vector<double>& bhp, const std::vector<double>& q) {
  if (topo.empty()) return;
  std::vector<double> qd;
  for (int e : dperf_order) qd.push_back(q[e]);
  w_bhp.upload(bhp, stream);
  w_q.upload(qd, stream);
}

// ------------------------------------------------------------------ reductions
template <class X, class Y>
double Sim::dev_dot_t(const X* a, const Y* b2, int len) {
  int nb = std::min(nblk(len), 1024);
  hipLaunchKernelGGL((k_dot_partial<X, Y>), dim3(nb), dim3(kBlock), 0, stream, len, a, b2, part.p);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, stream, nb, part.p, dscal.p + 63);
  double h;
  RSF_CHECK(hipMemcpyAsync(&h, dscal.p + 63, sizeof(double), hipMemcpyDeviceToHost, stream));
  sync();
  return h;
}
double Sim::dev_dot(const double* a, const double* b2, int len) { return dev_dot_t<double, double>(a, b2, len); }

// dot into a device slot without synchronising
template <class X, class Y>
static void dot_to(Sim& S, const X* a, const Y* b2, int len, double* slot) {
  int nb = std::min(nblk(len), 1024);
  hipLaunchKernelGGL((k_dot_partial<X, Y>), dim3(nb), dim3(kBlock), 0, S.stream, len, a, b2, S.part.p);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, S.stream, nb, S.part.p, slot);
}

// ------------------------------------------------------------------ preconditioner
template <class PT> struct PBuf;
template <> struct PBuf<double> {
  static DevBuf<double>& Dinv(Sim& s) { return s.Dinv_d; }
  static DevBuf<double>& Pm(Sim& s) { return s.Pm_d; }
  static DevBuf<double>& pd(Sim& s) { return s.pdinv_d; }
  static DevBuf<double>& w(Sim& s) { return s.wpt_d; }
  static DevBuf<double>& Pe(Sim& s) { return s.Pe_d; }
  static DevBuf<double>* bc(Sim& s) { return s.bc_d; }
  static DevBuf<double>& corr(Sim& s) { return s.corr_d; }
  static const double* Amat(Sim& s) { return s.grid.pcmp ? s.A_d.p : s.A.p; }
  static const double* Emat(Sim& s) { return s.w_extra.p; }
  static double* Z(Sim& s, int j) { return s.Zd[j]->p; }
  static double** Zptr(Sim& s) { return s.zptr_d.p; }
};
template <> struct PBuf<float> {
  static DevBuf<float>& Dinv(Sim& s) { return s.Dinv_f; }
  static DevBuf<float>& Pm(Sim& s) { return s.Pm_f; }
  static DevBuf<float>& pd(Sim& s) { return s.pdinv_f; }
  static DevBuf<float>& w(Sim& s) { return s.wpt_f; }
  static DevBuf<float>& Pe(Sim& s) { return s.Pe_f; }
  static DevBuf<float>* bc(Sim& s) { return s.bc_f; }
  static DevBuf<float>& corr(Sim& s) { return s.corr_f; }
  static const float* Amat(Sim& s) { return s.A_f.p; }
  static const float* Emat(Sim& s) { return s.Ef.p; }
  static float* Z(Sim& s, int j) { return s.Zf[j]->p; }
  static float** Zptr(Sim& s) { return s.zptr_f.p; }
};

template <class PT>
void Sim::cpr_build() {
  using B = PBuf<PT>;
  B::Dinv(*this).alloc(9 * (size_t)n);
  B::Pm(*this).alloc(7 * (size_t)n);
  B::pd(*this).alloc(n);
  B::w(*this).alloc(3 * (size_t)n);
  B::Pe(*this).alloc(std::max(n_extra, 1));
  B::corr(*this).alloc(3 * (size_t)n);
  for (int k = 0; k < 11; ++k) B::bc(*this)[k].alloc(n);
  // red-black with an even nx: colour-compacted preconditioner planes (every kernel
  // that reads them indexes them through pslot)
  grid.pcmp = (opt.ordering == 1 && nx % 2 == 0) ? 1 : 0;
  if (std::is_same<PT, float>::value) {
    A_f.alloc(63 * (size_t)n);
    Ef.alloc(9 * (size_t)std::max(n_extra, 1));
    if (grid.pcmp)
      hipLaunchKernelGGL((k_copy_pslot<float>), dim3(nblk(n), 63), dim3(kBlock), 0, stream, grid, A.p, A_f.p);
    else
      hipLaunchKernelGGL((k_copy<double, float>), dim3(nblk(63L * n)), dim3(kBlock), 0, stream, 63 * n, A.p, A_f.p);
    if (n_extra > 0)
      hipLaunchKernelGGL((k_copy<double, float>), dim3(nblk(9L * n_extra)), dim3(kBlock), 0, stream, 9 * n_extra,
                         w_extra.p, Ef.p);
  } else if (grid.pcmp) {
    A_d.alloc(63 * (size_t)n);
    hipLaunchKernelGGL((k_copy_pslot<double>), dim3(nblk(n), 63), dim3(kBlock), 0, stream, grid, A.p, A_d.p);
  }
  if (cpr_transposed) {  // A holds J^T: the transpose of the forward CPR
    hipLaunchKernelGGL((k_cpr_weights_t<PT>), dim3(nblk(n)), dim3(kBlock), 0, stream, n, A.p, cw.p, B::w(*this).p);
    hipLaunchKernelGGL((k_cpr_coarse_t<PT>), dim3(nblk(n)), dim3(kBlock), 0, stream, grid, A.p, cw.p,
                       B::Pm(*this).p);
    if (n_extra > 0)
      hipLaunchKernelGGL((k_cpr_extra_t<PT>), dim3(nblk(n_extra)), dim3(kBlock), 0, stream, n, n_extra, x_col.p,
                         cw.p, w_extra.p, B::Pe(*this).p);
  } else {
    hipLaunchKernelGGL((k_cpr_setup<PT>), dim3(nblk(n)), dim3(kBlock), 0, stream, grid, A.p, cw.p, B::Pm(*this).p,
                       B::w(*this).p);
    if (n_extra > 0)
      hipLaunchKernelGGL((k_cpr_extra<PT>), dim3(nblk(n_extra)), dim3(kBlock), 0, stream, n, n_extra, x_erow.p,
                         cw.p, w_extra.p, B::Pe(*this).p);
  }
  // rs/fim/linear/gmres_block_jacobi.rs:24,903-917: a coarse matrix of at most 299 rows is
  // inverted exactly (Gauss-Jordan, partial pivoting); a singular one keeps BiCGStab
  coarse_dense = false;
  if (!cpr_transposed && n <= kCoarseDenseMaxRows) {
    dense_m.alloc(2 * (size_t)n * n);
    dense_flag.alloc(1);
    hipLaunchKernelGGL((k_pdense_fill<PT>), dim3(nblk(n)), dim3(kBlock), 0, stream, grid, (const PT*)B::Pm(*this).p,
                       n_extra_rows, (const int*)x_rows.p, (const int*)x_ptr.p, (const int*)x_col.p,
                       (const PT*)B::Pe(*this).p, dense_m.p);
    hipLaunchKernelGGL(k_gauss_jordan, dim3(1), dim3(kDenseBlock), 0, stream, n, 2 * n, dense_m.p, 0.0, dense_flag.p);
    int singular = 1;
    dense_flag.download(&singular, 1, stream);
    sync();
    coarse_dense = singular == 0;
  }
  int mode = opt.ordering;
  sweep_on = false;
  if (mode == 0 && sweep_mode == 1) {
    // natural order in one launch per factorization and per sweep (tiles of grid lines)
    const SweepTiles tp = sweep_tiles(nx, ny, nz, sweep_lanes_p), tb = sweep_tiles(nx, ny, nz, sweep_lanes_b);
    sweep_on = tp.steps <= kSweepMaxSteps && tb.steps <= kSweepMaxSteps;
    if (sweep_on) {
      // tagged words: one per float, two per double, per value
      auto halo = [](const SweepTiles& t, int words) {
        return (size_t)words * t.nty * t.ntz * t.steps * (t.ty + t.tz);
      };
      const int tw = sizeof(PT) / 4;
      auto fresh = [&](DevBuf<unsigned long long>& h, size_t count, const SweepTiles& now, const SweepTiles& was) {
        const bool same = h.n == count && now.ty == was.ty && now.tz == was.tz && now.steps == was.steps;
        if (same) return;
        h.alloc(count);  // a new shape starts from tag 0, which no launch uses
        hipLaunchKernelGGL((k_fill<unsigned long long>), dim3(nblk((long)count)), dim3(kBlock), 0, stream,
                           (int)count, h.p, 0ull);
      };
      // every launch on a halo buffer rewrites all of its entries (one system, one tile
      // shape per buffer), so no entry outlives the next launch's tags
      fresh(sweep_halo_p, halo(tp, tw), tp, sweep_tp);
      fresh(sweep_halo_b, halo(tb, 3 * tw), tb, sweep_tb);
      fresh(sweep_halo_f, halo(tb, 9 * tw), tb, sweep_tb);
      sweep_tp = tp;
      sweep_tb = tb;
      if (!sweep_err.p) {
        sweep_err.alloc(1);
        hipLaunchKernelGGL((k_fill<int>), dim3(1), dim3(kBlock), 0, stream, 1, sweep_err.p, 0);
      }
    }
  }
  if (sweep_on) {
    factor_sweep<3, PT>(A.p, nullptr, B::Dinv(*this).p);
    factor_sweep<1, PT>(nullptr, B::Pm(*this).p, B::pd(*this).p);
    auto cells = [](const SweepTiles& t) { return (size_t)t.nty * t.ntz * t.steps * t.ty * t.tz; };
    DevBuf<PT>& slp = std::is_same<PT, float>::value ? (DevBuf<PT>&)sl_pf : (DevBuf<PT>&)sl_pd;
    DevBuf<PT>& slb = std::is_same<PT, float>::value ? (DevBuf<PT>&)sl_bf : (DevBuf<PT>&)sl_bd;
    slp.alloc(cells(sweep_tp) * sweep_rec(1));
    slb.alloc(cells(sweep_tb) * sweep_rec(3));
    const size_t vbytes = std::max(cells(sweep_tp), 3 * cells(sweep_tb)) * sizeof(PT);
    sweep_vin.alloc((vbytes + 7) / 8);
    sweep_vout.alloc((vbytes + 7) / 8);
    sweep_pack<1, PT>(B::Pm(*this).p, B::pd(*this).p, slp.p);
    sweep_pack<3, PT>(B::Amat(*this), B::Dinv(*this).p, slb.p);
  } else {
    const auto& lp = lvl_ptr[mode];
    for (size_t l = 0; l + 1 < lp.size(); ++l) {
      int cnt = lp[l + 1] - lp[l];
      if (!cnt) continue;
      hipLaunchKernelGGL((k_bilu_factor_level<PT>), dim3(nblk(cnt)), dim3(kBlock), 0, stream, grid, mode,
                         lvl_rows[mode].p + lp[l], cnt, A.p, B::Dinv(*this).p);
      hipLaunchKernelGGL((k_pilu_factor_level<PT>), dim3(nblk(cnt)), dim3(kBlock), 0, stream, grid, mode,
                         lvl_rows[mode].p + lp[l], cnt, B::Pm(*this).p, B::pd(*this).p);
    }
  }
  RSF_CHECK(hipGetLastError());
}

// The natural-order DILU factorization in one launch: NC = 3 the block factor of the
// Jacobian (block tiles), NC = 1 the pressure ILU0 (pressure tiles). Operands: 63 FP64
// values per cell for the block factor, one step of them in flight; 7 for the pressure.
template <int NC, class PT>
void Sim::factor_sweep(const double* Am, const PT* Pm, PT* dinv) {
  const SweepTiles& st = NC == 1 ? sweep_tp : sweep_tb;
  FactorSweepArgs<PT> a;
  a.g = grid;
  a.st = st;
  a.A = Am;
  a.P = Pm;
  a.dinv = dinv;
  DevBuf<unsigned long long>& halo = NC == 1 ? sweep_halo_p : sweep_halo_f;
  const size_t nT = (size_t)st.nty * st.ntz;
  const int words = NC * NC * (int)(sizeof(PT) / 4);
  a.halo_y = halo.p;
  a.halo_z = halo.p + (size_t)words * nT * st.steps * st.tz;
  a.err = sweep_err.p;
  a.tag0 = (++sweep_epoch) << 16;
  constexpr int D = NC == 1 ? 4 : 1;
  hipLaunchKernelGGL((k_ilu_factor_sweep<NC, PT, D, D>), dim3((unsigned)nT), dim3(st.ty * st.tz), 0, stream, a);
}

template <int NC, class PT>
void Sim::sweep_pack(const PT* M, const PT* dinv, PT* sl) {
  const SweepTiles& st = NC == 1 ? sweep_tp : sweep_tb;
  const long total = (long)st.nty * st.ntz * st.steps * st.ty * st.tz;
  hipLaunchKernelGGL((k_sweep_pack<NC, PT>), dim3(nblk(total)), dim3(kBlock), 0, stream, grid, st, M, dinv, sl);
}

// One natural-order solve (forward then backward) of the DILU factor packed in sl:
// the right-hand side gathered into march order, two sweep launches, the solution
// scattered back. Loads in flight per thread stay within the 63 the vector-memory counter
// tracks: the pressure sweeps run 8 steps of coefficients and 4 of halo entries ahead,
// the block sweeps 3 (FP32) or 1 (FP64).
template <int NC, class PT>
void Sim::sweep_solve(const PT* sl, const PT* r, PT* out, const double* skip) {
  const SweepTiles& st = NC == 1 ? sweep_tp : sweep_tb;
  DevBuf<unsigned long long>& halo = NC == 1 ? sweep_halo_p : sweep_halo_b;
  SweepArgs<PT> a;
  a.g = grid;
  a.st = st;
  a.sl = sl;
  a.r = (const PT*)sweep_vin.p;
  a.out = (PT*)sweep_vout.p;
  const size_t nT = (size_t)st.nty * st.ntz;
  a.cells = nT * st.steps * st.ty * st.tz;
  hipLaunchKernelGGL((k_sweep_gather<NC, PT>), dim3(nblk((long)a.cells)), dim3(kBlock), 0, stream, grid, st, r,
                     (PT*)sweep_vin.p, skip);
  const int words = NC * (int)(sizeof(PT) / 4);
  a.halo_y = halo.p;
  a.halo_z = halo.p + (size_t)words * nT * st.steps * st.tz;
  a.err = sweep_err.p;
  a.skip = skip;
  const int nl = st.ty * st.tz;
  auto launch = [&](auto fwd, auto bwd) {
    a.tag0 = (++sweep_epoch) << 16;
    hipLaunchKernelGGL(fwd, dim3((unsigned)nT), dim3(nl), 0, stream, a);
    a.tag0 = (++sweep_epoch) << 16;
    hipLaunchKernelGGL(bwd, dim3((unsigned)nT), dim3(nl), 0, stream, a);
  };
  constexpr int D = NC == 1 ? 8 : (sizeof(PT) == 4 ? 3 : 1);
  constexpr int HD = NC == 1 ? 4 : D;
  launch(k_ilu_sweep<NC, PT, false, D, HD>, k_ilu_sweep<NC, PT, true, D, HD>);
  hipLaunchKernelGGL((k_sweep_scatter<NC, PT>), dim3(nblk((long)a.cells)), dim3(kBlock), 0, stream, grid, st,
                     (const PT*)sweep_vout.p, out, skip);
}

// A sweep whose wait on a neighbouring tile gave up leaves a device flag; the host reads it
// once per linear solve and fails the solve.
void Sim::sweep_check() {
  if (!sweep_err.p) return;
  int e = 0;
  sweep_err.download(&e, 1, stream);
  sync();
  if (e) {
    hipLaunchKernelGGL((k_fill<int>), dim3(1), dim3(kBlock), 0, stream, 1, sweep_err.p, 0);
    throw HipError("natural-order ILU sweep: a wait on a neighbouring tile gave up");
  }
}

template <class PT>
void Sim::bilu_apply(const PT* r, PT* out, const PT* Am) {
  using B = PBuf<PT>;
  int mode = opt.ordering;
  const auto& lp = lvl_ptr[mode];
  const int* rows = lvl_rows[mode].p;
  if (mode == 1) {
    int nred = lp[1], nblack = lp[2] - lp[1];
    if (nblack)
      hipLaunchKernelGGL((k_bilu_rb_black<PT, PT>), dim3(nblk(nblack)), dim3(kBlock), 0, stream, grid, rows + nred,
                         nblack, Am, B::Dinv(*this).p, r, out);
    if (nred)
      hipLaunchKernelGGL((k_bilu_rb_red<PT, PT>), dim3(nblk(nred)), dim3(kBlock), 0, stream, grid, rows, nred, Am,
                         B::Dinv(*this).p, r, out);
    return;
  }
  if (sweep_on) {
    sweep_solve<3, PT>(std::is_same<PT, float>::value ? (const PT*)sl_bf.p : (const PT*)sl_bd.p, r, out, nullptr);
    return;
  }
  int L = (int)lp.size() - 1;
  for (int l = 0; l < L; ++l) {
    int cnt = lp[l + 1] - lp[l];
    if (cnt)
      hipLaunchKernelGGL((k_bilu_fwd_level<PT, PT>), dim3(nblk(cnt)), dim3(kBlock), 0, stream, grid, mode,
                         rows + lp[l], cnt, Am, B::Dinv(*this).p, r, out);
  }
  for (int l = L - 1; l >= 0; --l) {
    int cnt = lp[l + 1] - lp[l];
    if (cnt)
      hipLaunchKernelGGL((k_bilu_bwd_level<PT, PT>), dim3(nblk(cnt)), dim3(kBlock), 0, stream, grid, mode,
                         rows + lp[l], cnt, Am, B::Dinv(*this).p, out);
  }
}

template <class PT>
void Sim::pilu_apply(const PT* r, PT* out, const double* skip) {
  using B = PBuf<PT>;
  int mode = opt.ordering;
  const auto& lp = lvl_ptr[mode];
  const int* rows = lvl_rows[mode].p;
  if (mode == 1) {
    int nred = lp[1], nblack = lp[2] - lp[1];
    if (nblack)
      hipLaunchKernelGGL((k_pilu_rb_black<PT>), dim3(nblk(nblack)), dim3(kBlock), 0, stream, grid, rows + nred,
                         nblack, B::Pm(*this).p, B::pd(*this).p, r, out, skip);
    if (nred)
      hipLaunchKernelGGL((k_pilu_rb_red<PT>), dim3(nblk(nred)), dim3(kBlock), 0, stream, grid, rows, nred,
                         B::Pm(*this).p, B::pd(*this).p, r, out, skip);
    return;
  }
  if (sweep_on) {
    sweep_solve<1, PT>(std::is_same<PT, float>::value ? (const PT*)sl_pf.p : (const PT*)sl_pd.p, r, out, skip);
    return;
  }
  int L = (int)lp.size() - 1;
  for (int l = 0; l < L; ++l) {
    int cnt = lp[l + 1] - lp[l];
    if (cnt)
      hipLaunchKernelGGL((k_pilu_fwd_level<PT>), dim3(nblk(cnt)), dim3(kBlock), 0, stream, grid, mode, rows + lp[l],
                         cnt, B::Pm(*this).p, B::pd(*this).p, r, out, skip);
  }
  for (int l = L - 1; l >= 0; --l) {
    int cnt = lp[l + 1] - lp[l];
    if (cnt)
      hipLaunchKernelGGL((k_pilu_bwd_level<PT>), dim3(nblk(cnt)), dim3(kBlock), 0, stream, grid, mode, rows + lp[l],
                         cnt, B::Pm(*this).p, B::pd(*this).p, out, skip);
  }
}

// Launch shape of the general pressure kernels: up to kRBlock threads per block and at
// most kRGrid blocks; a small system takes smaller blocks (down to one wavefront) so that
// its tiles spread over every CU.
static int rthreads(int n) {
  int t = kRBlock;
  while (t > 64 && (n + t * kRUnroll - 1) / (t * kRUnroll) < kRGrid) t /= 2;
  return t;
}
static int rgrid(int n) {
  const int t = rthreads(n);
  return std::max(1, std::min(kRGrid, (n + t * kRUnroll - 1) / (t * kRUnroll)));
}

// y = P x (+ clique extras); with part set, also the dot partials of the BiCGStab
// With ev0/ev1 set the launch carries them as start/stop events: the dispatch itself
// timestamps the kernel's execution (no extra packets on the stream).
template <class PT>
static void pmatvec(Sim& S, const PT* x, PT* y, const PT* u = nullptr, const PT* w = nullptr,
                    const double* bs = nullptr, double* part = nullptr, hipEvent_t ev0 = nullptr,
                    hipEvent_t ev1 = nullptr) {
  using B = PBuf<PT>;
  const int* xw = S.n_extra_rows > 0 ? (const int*)S.x_wave.p : nullptr;
  if (ev0) {
    hipExtLaunchKernelGGL((k_pspmv_dot<PT>), dim3(rgrid(S.n)), dim3(rthreads(S.n)), 0, S.stream, ev0, ev1, 0, S.grid,
                          (const PT*)B::Pm(S).p, x, y, xw, (const int*)S.x_rows.p, (const int*)S.x_ptr.p,
                          (const int*)S.x_col.p, (const PT*)B::Pe(S).p, u, w, bs, part);
    return;
  }
  hipLaunchKernelGGL((k_pspmv_dot<PT>), dim3(rgrid(S.n)), dim3(rthreads(S.n)), 0, S.stream, S.grid, B::Pm(S).p, x, y, xw,
                     S.x_rows.p, S.x_ptr.p, S.x_col.p, B::Pe(S).p, u, w, bs, part);
}

// rs/fim/linear/gmres_block_jacobi.rs:660-730 right-preconditioned BiCGStab, x0 = 0,
// scalars on the device (see the kernels). An iteration is 9 launches in red-black
// ordering (5 + 2 x 2 x levels in natural ordering). The host enqueues iterations
// in chunks of 4, 8, 16, ... and reads the done flag between chunks.
template <class PT>
void Sim::bicgstab(PT* x, const PT* bv, int& its) {
  if (grid.pcmp) {
    bicgstab_rb<PT>(x, bv, its);
    return;
  }
  using B = PBuf<PT>;
  DevBuf<PT>* w = B::bc(*this);
  PT *r = w[0].p, *rh = w[1].p, *p = w[2].p, *v = w[3].p, *s = w[4].p, *t = w[5].p, *ph = w[6].p, *sh = w[7].p;
  double* bs = bsc.p;
  const double* skip = bs + kBsDone;
  const int m = rgrid(n);
  dim3 gr(m), blk(rthreads(n));
  double* part0 = bpart.p;
  double* part1 = part0 + kPartStride;
  double* part2 = part1 + kPartStride;
  double* part3 = part2 + kPartStride;
  const double rtol = opt.pressure_rel_tol;
  const int max_it = opt.pressure_max_it;
  hipLaunchKernelGGL((k_bicg_init<PT>), gr, blk, 0, stream, n, bv, x, r, rh, p, v, bs, part0);
  int k = 0, chunk = std::max(4, std::min(max_it, pressure_its_hint));
  double h[kBsRr + 1] = {};
  std::vector<int> probe_it;  // iteration of each timed SpMV launch in this chunk
  auto spmv = [&](const PT* xin, PT* yout, const PT* u, const PT* w, double* part, int it) {
    if (probe_on && it % probe_every == 0) {
      size_t e = 2 * probe_it.size();
      while (probe_ev.size() < e + 2) {
        hipEvent_t ev;
        RSF_CHECK(hipEventCreate(&ev));
        probe_ev.push_back(ev);
      }
      pmatvec<PT>(*this, xin, yout, u, w, bs, part, probe_ev[e], probe_ev[e + 1]);
      probe_it.push_back(it);
    } else {
      pmatvec<PT>(*this, xin, yout, u, w, bs, part);
    }
  };
  while (k < max_it) {
    int last = std::min(k + chunk, max_it);
    probe_it.clear();
    for (++k; k <= last; ++k) {
      hipLaunchKernelGGL((k_bicg_p<PT>), gr, blk, 0, stream, n, k, bs, (const double*)part0, m, r, v, p, rtol);
      pilu_apply<PT>(p, ph, skip);
      spmv(ph, v, rh, nullptr, part1, k);
      hipLaunchKernelGGL((k_bicg_s<PT>), gr, blk, 0, stream, n, k, bs, (const double*)part1, m, r, v, s, part2);
      pilu_apply<PT>(s, sh, skip);
      spmv(sh, t, nullptr, s, part3, k);
      hipLaunchKernelGGL((k_bicg_update<PT>), gr, blk, 0, stream, n, bs, (const double*)part2, (const double*)part3,
                         m, ph, sh, s, t, rh, x, r, part0);
    }
    k = last;
    chunk *= 2;
    static_assert(kBsIts < kBsRr && kBsDone < kBsRr && kBsConvS < kBsRr, "scalars and flags are read together");
    RSF_CHECK(hipMemcpyAsync(h, bs, sizeof(h), hipMemcpyDeviceToHost, stream));
    sync();
    // launches of iterations the solve did not reach returned at once: not timed
    for (size_t q = 0; q < probe_it.size(); ++q) {
      if (probe_it[q] > (int)h[kBsIts]) continue;
      float ms = 0.0f;
      RSF_CHECK(hipEventElapsedTime(&ms, probe_ev[2 * q], probe_ev[2 * q + 1]));
      probe_ms += ms;
      probe_launches++;
    }
    if (h[kBsDone] != 0.0 || h[kBsConvS] != 0.0) break;
  }
  its = (int)h[kBsIts];
  pressure_its_hint = its;
}

// The same solve on the colour-compacted red-black system (five launches per
// iteration; the kernels describe the fusion). Every vector of
// the solve is kept in the compacted order: b is gathered into it by the first kernel
// and x scattered back by the last. The iteration number lives on the device and the
// solve ends there; the host enqueues iterations in chunks of 4, 8, 16, ... and reads
// the flags between chunks. With the kernel probe on, the A kernel carries start/stop
// events; launches of full iterations (2 .. its) are timed.
template <class PT>
void Sim::bicgstab_rb(PT* x, const PT* bv, int& its) {
  using B = PBuf<PT>;
  DevBuf<PT>* w = B::bc(*this);
  PT *r = w[0].p, *rh = w[1].p, *v = w[3].p, *s = w[4].p, *t = w[5].p, *ph = w[6].p, *sh = w[7].p;
  PT *p0 = w[2].p, *p1 = w[8].p;
  PT *chat = w[9].p, *xc = w[10].p;
  const PT* Pm = B::Pm(*this).p;
  const PT* d = B::pd(*this).p;
  const PT* Pe = B::Pe(*this).p;
  double* bs = bsc.p;
  // kRBlock / 64 wavefronts per block, at most kRGrid blocks; a small grid takes smaller
  // blocks (down to 2 wavefronts) so that its rows spread over every CU
  const int waves = rb_waves(grid);
  int wpb = kRBlock / 64;
  while (wpb > 2 && (waves + wpb - 1) / wpb < kRGrid) wpb /= 2;
  const int m = std::max(1, std::min(kRGrid, (waves + wpb - 1) / wpb));
  dim3 gr(m), blk(64 * wpb);
  // partial regions of up to four values per block
  double* part0 = bpart.p;
  double* part1a = part0 + 2 * kPartStride;
  double* part1b = part1a + 2 * kPartStride;
  double* part2 = part1b + 2 * kPartStride;
  double* part3 = part2 + 2 * kPartStride;
  const double rtol = opt.pressure_rel_tol;
  const int max_it = std::max(0, opt.pressure_max_it);
  const int ext = n_extra_rows > 0 ? 1 : 0;
  hipLaunchKernelGGL((k_rb_init<PT>), gr, blk, 0, stream, grid, Pm, bv, xc, r, rh, p0, v, chat, bs, part0);
  if (ext)
    hipLaunchKernelGGL((k_rb_extras_t<PT>), dim3(n_extra_rows), dim3(64), 0, stream, n_extra_rows, x_rows_c.p,
                       x_ptr.p, x_col_c.p, x_tperm.p, Pe, (const PT*)rh, chat);
  auto iteration = [&](hipEvent_t e0, hipEvent_t e1) {
    if (e0)
      hipExtLaunchKernelGGL((k_rb_a<PT>), gr, blk, 0, stream, e0, e1, 0, grid, max_it, bs, (const double*)part0,
                            (const double*)part2, (const double*)part3, m, Pm, d, (const PT*)chat, xc, r,
                            (const PT*)s, (const PT*)t, ph, (const PT*)sh, (const PT*)v, p0, p1, rtol, part1a);
    else
      hipLaunchKernelGGL((k_rb_a<PT>), gr, blk, 0, stream, grid, max_it, bs, (const double*)part0,
                         (const double*)part2, (const double*)part3, m, Pm, d, (const PT*)chat, xc, r, (const PT*)s,
                         (const PT*)t, ph, (const PT*)sh, (const PT*)v, p0, p1, rtol, part1a);
    hipLaunchKernelGGL((k_rb_red<PT>), gr, blk, 0, stream, grid, bs, Pm, d, (const PT*)p0, (const PT*)p1, ph, v,
                       (const PT*)chat, part1b, ext ? (const int*)x_wave2.p : nullptr, x_rows.p, x_ptr.p, x_col_c.p,
                       Pe);
    hipLaunchKernelGGL((k_rb_c<PT>), gr, blk, 0, stream, grid, bs, (const double*)part1a, (const double*)part1b, m, Pm,
                       d, (const PT*)r, (const PT*)rh, (const PT*)ph, v, s, sh, ext, part2);
    hipLaunchKernelGGL((k_rb_red<PT>), gr, blk, 0, stream, grid, bs, Pm, d, (const PT*)s, (const PT*)nullptr, sh, t,
                       (const PT*)nullptr, (double*)nullptr, (const int*)nullptr, (const int*)nullptr,
                       (const int*)nullptr, (const int*)nullptr, (const PT*)nullptr);
    hipLaunchKernelGGL((k_rb_e<PT>), gr, blk, 0, stream, grid, bs, Pm, (const PT*)sh, (const PT*)s, (const PT*)rh, t,
                       ext ? (const int*)x_wave2.p : nullptr, x_rows.p, x_ptr.p, x_col_c.p, Pe, part3);
  };
  // iterations 1 .. max_it, then one more A that applies the last update. The host reads
  // the flags after chunks of iterations that double from the previous solve's count
  // (at least 4): a solve that runs to the cap then costs one round trip, and launches
  // past an early convergence return at once.
  int launched = 0, chunk = std::max(4, std::min(max_it + 1, pressure_its_hint + 1));
  double h[kBsRr + 1] = {};
  std::vector<int> probe_it;  // iteration of each timed launch in this chunk
  while (launched < max_it + 1) {
    const int c = std::min(chunk, max_it + 1 - launched);
    probe_it.clear();
    for (int i = 0; i < c; ++i) {
      if (!probe_on || (launched + i + 1) % probe_every != 0) {
        iteration(nullptr, nullptr);
        continue;
      }
      size_t e = 2 * probe_it.size();
      while (probe_ev.size() < e + 2) {
        hipEvent_t ev;
        RSF_CHECK(hipEventCreate(&ev));
        probe_ev.push_back(ev);
      }
      iteration(probe_ev[e], probe_ev[e + 1]);
      probe_it.push_back(launched + i + 1);
    }
    launched += c;
    chunk *= 2;
    RSF_CHECK(hipMemcpyAsync(h, bs, sizeof(h), hipMemcpyDeviceToHost, stream));
    sync();
    for (size_t q = 0; q < probe_it.size(); ++q) {
      // iteration 1 and the closing update are lighter launches: not timed
      if (probe_it[q] < 2 || probe_it[q] > (int)h[kBsIts]) continue;
      float ms = 0.0f;
      RSF_CHECK(hipEventElapsedTime(&ms, probe_ev[2 * q], probe_ev[2 * q + 1]));
      probe_ms += ms;
      probe_launches++;
    }
    if (h[kBsDone] != 0.0 || h[kBsConvS] != 0.0) break;
  }
  hipLaunchKernelGGL((k_rb_scatter<PT>), gr, blk, 0, stream, grid, (const PT*)xc, x);
  its = (int)h[kBsIts];
  pressure_its_hint = its;
}

// rs/fim/linear/gmres_block_jacobi.rs:584-609: the residual reduction of a coarse solve,
// |rp - P xp| / |rp| (0 for |rp| <= DBL_EPSILON), from the true residual. The two squared
// norms go to the device slots of this application; linear_solve reads them once at its end.
template <class PT>
void Sim::pressure_reduction(const PT* rp, const PT* xp) {
  using B = PBuf<PT>;
  if (cpr_apps < cpr_red_cap) {
    PT* y = B::bc(*this)[3].p;  // free once the coarse solve has returned
    pmatvec<PT>(*this, xp, y);
    const int nb = std::min(nblk(n), 1024);
    hipLaunchKernelGGL((k_resid_norms_partial<PT>), dim3(nb), dim3(kBlock), 0, stream, n, rp, (const PT*)y, part.p,
                       1024);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, stream, nb, part.p, cpr_red.p + 2 * cpr_apps);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, stream, nb, part.p + 1024, cpr_red.p + 2 * cpr_apps + 1);
  }
  cpr_apps++;
}

// rs/fim/linear/gmres_block_jacobi.rs:471-501 two-stage CPR application
template <class PT>
void Sim::cpr_apply(const double* v, PT* z) {
  using B = PBuf<PT>;
  DevBuf<PT>* w = B::bc(*this);
  // the BiCGStab work vectors occupy bc[0..7]; rp and xp live in corr's spare planes
  PT* corr = B::corr(*this).p;
  PT* rp = corr;          // plane 0 used as restriction rhs before corr is formed
  PT* xp = corr + n;      // plane 1 holds the pressure correction
  if (cpr_transposed) {
    // z = y + w (x) C^-T e0^T (v - J^T y),  y = ILU(J^T)^-1 v
    hipLaunchKernelGGL((k_copy<double, PT>), dim3(nblk(3L * n)), dim3(kBlock), 0, stream, 3 * n, v, corr);
    bilu_apply<PT>(corr, z, B::Amat(*this));
    hipLaunchKernelGGL((k_cpr_restrict_t<PT, PT, PT>), dim3(nblk(n)), dim3(kBlock), 0, stream, grid,
                       B::Amat(*this), (const PT*)z, v, w[0].p);
    if (n_extra_rows > 0)
      hipLaunchKernelGGL((k_cpr_restrict_t_extra<PT, PT, PT>), dim3(n_extra_rows), dim3(64), 0, stream, n,
                         n_extra_rows, x_rows.p, x_ptr.p, x_col.p, B::Emat(*this), std::max(n_extra, 1),
                         (const PT*)z, w[0].p);
    hipLaunchKernelGGL((k_copy<PT, PT>), dim3(nblk(n)), dim3(kBlock), 0, stream, n, w[0].p, rp);
    int its = 0;
    bicgstab<PT>(xp, rp, its);
    stats.bicgstab += its;
    stats.pressure_solves++;
    hipLaunchKernelGGL((k_prolong_w<PT, PT>), dim3(nblk(n)), dim3(kBlock), 0, stream, n, B::w(*this).p,
                       (const PT*)xp, z);
    return;
  }
  hipLaunchKernelGGL((k_restrict<PT, double>), dim3(nblk(n)), dim3(kBlock), 0, stream, n, B::w(*this).p, v, rp);
  int its = 0;
  // coarse solution into z plane 0 (temporary), then moved to xp: the exact inverse of a
  // coarse matrix of at most 299 rows (:903-917), else BiCGStab
  if (coarse_dense)
    hipLaunchKernelGGL((k_pdense_apply<PT>), dim3((64 * n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, n,
                       (const double*)dense_m.p, (const PT*)rp, z);
  else
    bicgstab<PT>(z, rp, its);
  pressure_reduction<PT>(rp, z);
  stats.bicgstab += its;
  stats.pressure_solves++;
  hipLaunchKernelGGL((k_copy<PT, PT>), dim3(nblk(n)), dim3(kBlock), 0, stream, n, z, xp);
  // corr = v - A(:, p) xp, needs xp intact: compute into z (3 planes) then swap into corr
  hipLaunchKernelGGL((k_cpr_corr<PT, PT, double, PT>), dim3(nblk(n)), dim3(kBlock), 0, stream, grid, B::Amat(*this),
                     xp, v, z);
  if (n_extra_rows > 0)
    hipLaunchKernelGGL((k_cpr_corr_extra<PT, PT, PT>), dim3(n_extra_rows), dim3(64), 0, stream, n,
                       n_extra_rows, x_rows.p, x_ptr.p, x_col.p, B::Emat(*this), std::max(n_extra, 1), xp, z);
  // keep xp in the BiCGStab scratch (bc[0]) while the block ILU overwrites corr
  hipLaunchKernelGGL((k_copy<PT, PT>), dim3(nblk(n)), dim3(kBlock), 0, stream, n, xp, w[0].p);
  hipLaunchKernelGGL((k_copy<PT, PT>), dim3(nblk(3L * n)), dim3(kBlock), 0, stream, 3 * n, z, corr);
  bilu_apply<PT>(corr, z, B::Amat(*this));
  hipLaunchKernelGGL((k_prolong_add<PT, PT>), dim3(nblk(n)), dim3(kBlock), 0, stream, n, w[0].p, z);
}

// y = A x (+ extras), x of any type, y FP64
template <class X>
static void matvec(Sim& S, const X* x, double* y) {
  hipLaunchKernelGGL((k_spmv<double, X, double>), dim3(nblk(S.n)), dim3(kBlock), 0, S.stream, S.grid, S.A.p, x, y);
  if (S.n_extra_rows > 0)
    hipLaunchKernelGGL((k_spmv_extra<double, X, double>), dim3(S.n_extra_rows), dim3(64), 0, S.stream, S.n,
                       S.n_extra_rows, S.x_rows.p, S.x_ptr.p, S.x_col.p, S.w_extra.p, std::max(S.n_extra, 1), x, y, 1.0);
}
static void residual(Sim& S, const double* x, const double* bv, double* y) {
  hipLaunchKernelGGL((k_residual<double, double>), dim3(nblk(S.n)), dim3(kBlock), 0, S.stream, S.grid, S.A.p, x, bv, y);
  if (S.n_extra_rows > 0)
    hipLaunchKernelGGL((k_spmv_extra<double, double, double>), dim3(S.n_extra_rows), dim3(64), 0, S.stream,
                       S.n, S.n_extra_rows, S.x_rows.p, S.x_ptr.p, S.x_col.p, S.w_extra.p, std::max(S.n_extra, 1), x, y,
                       -1.0);
}

// rs/fim/linear/gmres_block_jacobi.rs:1470-1665 right-preconditioned FGMRES with the
// true residual checked after every inner step; solves A dx = b on the reduced system.
template <class PT>
bool Sim::linear_solve(double rhs_norm_full, int& iters, double& final_res) {
  using B = PBuf<PT>;
  cpr_apps = 0;
  cpr_ratio_sum = 0.0;
  cpr_ratio_last = 1.0;
  int N3 = 3 * n;
  int restart = std::max(opt.restart, 2);
  int max_it = std::max(opt.lin_max_it, restart);
  cpr_red_cap = max_it + 1;
  cpr_red.alloc(2 * (size_t)cpr_red_cap);
  while ((int)V.size() < restart + 1) {
    V.push_back(new DevBuf<double>());
    V.back()->alloc(N3);
  }
  std::vector<PT*> zp;
  if (std::is_same<PT, float>::value) {
    while ((int)Zf.size() < restart) {
      Zf.push_back(new DevBuf<float>());
      Zf.back()->alloc(N3);
    }
    for (int j = 0; j < restart; ++j) zp.push_back((PT*)Zf[j]->p);
    zptr_f.upload((float* const*)zp.data(), zp.size(), stream);
  } else {
    while ((int)Zd.size() < restart) {
      Zd.push_back(new DevBuf<double>());
      Zd.back()->alloc(N3);
    }
    for (int j = 0; j < restart; ++j) zp.push_back((PT*)Zd[j]->p);
    zptr_d.upload((double* const*)zp.data(), zp.size(), stream);
  }
  gm.alloc((size_t)(restart + 1) * restart + 2 * restart + (restart + 1) + restart + 2);
  gm_H = gm.p;
  gm_cs = gm_H + (size_t)(restart + 1) * restart;
  gm_sn = gm_cs + restart;
  gm_g = gm_sn + restart;
  gm_y = gm_g + restart + 1;
  gm_st = gm_y + restart;
  double bn = std::sqrt(dev_dot(b.p, b.p, N3));
  double tol_red = opt.lin_abs_tol + opt.lin_rel_tol * std::max(bn, kEps);
  double tol_full = opt.lin_abs_tol + opt.lin_rel_tol * std::max(rhs_norm_full, kEps);
  hipLaunchKernelGGL((k_fill<double>), dim3(nblk(N3)), dim3(kBlock), 0, stream, N3, dxv.p, 0.0);
  iters = 0;
  final_res = bn;
  bool conv = false;
  while (iters < max_it) {
    // the first cycle starts from dx = 0, where r = b and |r| = |b| (already formed)
    const double* r0 = b.p;
    double beta = bn;
    if (iters > 0) {
      residual(*this, dxv.p, b.p, tmp.p);
      beta = std::sqrt(dev_dot(tmp.p, tmp.p, N3));
      r0 = tmp.p;
    }
    final_res = beta;
    if (beta <= tol_red) {
      conv = true;
      break;
    }
    int m = std::min(restart, max_it - iters);
    hipLaunchKernelGGL((k_fill<double>), dim3(nblk(N3)), dim3(kBlock), 0, stream, N3, V[0]->p, 0.0);
    hipLaunchKernelGGL((k_axpy<double, double>), dim3(nblk(N3)), dim3(kBlock), 0, stream, N3, 1.0 / beta, r0, V[0]->p);
    // the least-squares problem lives on the device (k_fgmres_column); per inner step the
    // host reads two numbers: the new vector's norm (breakdown test) and the true residual
    hipLaunchKernelGGL(k_fgmres_begin, dim3(1), dim3(kBlock), 0, stream, m, restart, beta, gm_H, gm_cs, gm_sn, gm_g);
    bool done = false;
    int steps = 0;  // inner steps of this cycle: each writes cand = dx + Z y
    for (int j = 0; j < m; ++j) {
      PT* zj = zp[j];
      cpr_apply<PT>(V[j]->p, zj);
      double* wv = V[j + 1]->p;
      matvec<PT>(*this, zj, wv);
      // modified Gram-Schmidt: h_q = v_q . w, w -= h_q v_q for q = 0 .. j, then |w|^2; each
      // update shares its pass over w with the next dot (k_mgs_axpy_dot)
      dot_to<double, double>(*this, V[0]->p, wv, N3, dscal.p);
      const int nbd = std::min(nblk(N3), 1024);
      for (int q = 0; q <= j; ++q) {
        hipLaunchKernelGGL(k_mgs_axpy_dot, dim3(nbd), dim3(kBlock), 0, stream, N3, -1.0, (const double*)dscal.p + q,
                           (const double*)V[q]->p, wv, q < j ? (const double*)V[q + 1]->p : (const double*)nullptr,
                           part.p);
        hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, stream, nbd, (const double*)part.p,
                           dscal.p + q + 1);
      }
      hipLaunchKernelGGL(k_fgmres_column, dim3(1), dim3(64), 0, stream, j, restart, kEps, (const double*)dscal.p, gm_H,
                         gm_cs, gm_sn, gm_g, gm_y, gm_st);
      hipLaunchKernelGGL((k_scale_inv_sqrt_if<double>), dim3(nblk(N3)), dim3(kBlock), 0, stream, N3,
                         (const double*)dscal.p + j + 1, kEps, (const double*)wv, wv);
      iters++;
      steps++;
      const int cols = j + 1;
      hipLaunchKernelGGL((k_combine<PT>), dim3(nblk(N3)), dim3(kBlock), 0, stream, N3, cols, dxv.p,
                         (const PT* const*)B::Zptr(*this), (const double*)gm_y, cand.p);
      residual(*this, cand.p, b.p, tmp.p);
      dot_to<double, double>(*this, tmp.p, tmp.p, N3, gm_st + 1);
      double st[2];
      RSF_CHECK(hipMemcpyAsync(st, gm_st, sizeof(st), hipMemcpyDeviceToHost, stream));
      sync();
      const double nn = st[0], tr = std::sqrt(st[1]);
      final_res = tr;
      if (tr <= tol_red) {
        conv = true;
        done = true;
        break;
      }
      if (nn <= kEps || iters >= max_it) {
        done = true;
        break;
      }
    }
    if (steps > 0) std::swap(dxv.p, cand.p);  // the cycle's last iterate becomes dx
    if (done) break;
  }
  stats.linear += iters;
  if (cpr_apps > 0) {  // the coarse solves' residual reductions (:380-420, :584-609)
    const int k = std::min(cpr_apps, cpr_red_cap);
    std::vector<double> rr(2 * (size_t)k);
    cpr_red.download(rr.data(), rr.size(), stream);
    sync();
    for (int a = 0; a < k; ++a) {
      const double bn = std::sqrt(rr[2 * a + 1]);
      const double ratio = bn > DBL_EPSILON ? std::sqrt(rr[2 * a]) / bn : 0.0;
      cpr_ratio_sum += ratio;
      cpr_ratio_last = ratio;
    }
    cpr_apps = k;
  }
  if (sweep_on) sweep_check();
  // rs/fim/linear/well_schur.rs:410-471: full-system residual equals the reduced one
  return conv && final_res <= tol_full;
}

// rs/fim/linear/mod.rs:40,313-335: a system of at most 512 rows (cells and well tail) is
// solved directly; a singular one falls back to the iterative solver. The wells' tail is
// already eliminated in closed form, so the reduced cell system is factored densely here and
// the tail recovered by the Newton update, the same solution to rounding.
bool Sim::solve_direct(double& final_res) {
  const int m = 3 * n;
  dense_m.alloc((size_t)m * (m + 1));
  dense_flag.alloc(1);
  hipLaunchKernelGGL(k_dense_system, dim3(nblk(m)), dim3(kBlock), 0, stream, grid, (const double*)A.p, n_extra_rows,
                     (const int*)x_rows.p, (const int*)x_ptr.p, (const int*)x_col.p, (const double*)w_extra.p,
                     std::max(n_extra, 1), (const double*)b.p, dense_m.p);
  hipLaunchKernelGGL(k_gauss_jordan, dim3(1), dim3(kDenseBlock), 0, stream, m, m + 1, dense_m.p, 1e-300, dense_flag.p);
  int singular = 1;
  dense_flag.download(&singular, 1, stream);
  sync();
  if (singular) return false;
  hipLaunchKernelGGL(k_dense_solution, dim3(nblk(m)), dim3(kBlock), 0, stream, n, (const double*)dense_m.p, dxv.p);
  residual(*this, dxv.p, b.p, tmp.p);
  final_res = std::sqrt(dev_dot(tmp.p, tmp.p, m));
  stats.linear += 1;
  return true;
}

// rs/fim/linear/mod.rs:287-397 solve_linearized_system on the reduced system: the direct
// route for small systems, else CPR-FGMRES in the configured precision
bool Sim::solve_reduced(double rhs_norm_full, int& iters, double& final_res, bool& direct) {
  direct = false;
  const int full_rows = 3 * n + (int)topo.size() + (int)wells.size();
  if (!cpr_transposed && full_rows <= kDenseMaxRows && solve_direct(final_res)) {
    iters = 1;
    direct = true;
    cpr_apps = 0;
    return true;
  }
  if (opt.precision == 1) {
    cpr_build<float>();
    return linear_solve<float>(rhs_norm_full, iters, final_res);
  }
  cpr_build<double>();
  return linear_solve<double>(rhs_norm_full, iters, final_res);
}

// ------------------------------------------------------------------ Newton
// One Newton record (kRec* layout) at the current iterate: the acceptance measures, the
// tail families and sum of squares, every well's own convergence test, |R_cells|^2, and
// the finiteness flag of the last update. One host read.
void Sim::newton_record(double dt, double* rec_h) {
  PhysParams P = params(true);
  const int nb = std::min(nblk(n), 1024);
  hipLaunchKernelGGL(k_conv_partial, dim3(nb), dim3(kBlock), 0, stream, P, n, dt, d_pv.p, x_p.p, s_rs.p, R.p, part.p);
  hipLaunchKernelGGL(k_conv_reduce, dim3(1), dim3(kRecBlock), 0, stream, nb, n, (const double*)part.p, dscal.p, rec.p);
  hipLaunchKernelGGL(k_conv_violation, dim3(nb), dim3(kBlock), 0, stream, n, d_pv.p, R.p, dscal.p, part.p);
  hipLaunchKernelGGL(k_conv_finish, dim3(1), dim3(kRecBlock), 0, stream, nb, (const double*)part.p, (const double*)dscal.p,
                     well_dev(), rec.p);
  dot_to<double, double>(*this, R.p, R.p, 3 * n, rec.p + kRecCells2);
  rec.download(rec_h, kRecCount, stream);
  sync();
}

// rs/fim/newton.rs:970-3368 run_fim_timestep. The iterate lives on the device, wells
// included (W.bhp, W.q); the host reads one Newton record per iteration besides the
// linear solver's own reads, and the well state once at the end.
// The last linear solve of a Newton attempt, as the retry classification reads it
// (rs/fim/newton.rs:921-944 last_linear_report with its CPR diagnostics).
struct LinearOutcome {
  bool have = false, converged = false, direct = false;
  int apps = 0;
  double avg = 1.0, last = 1.0;
};

// rs/fim/newton.rs:439-481 classify_retry_failure_with_site: an attempt without a linear
// solve is mixed, an unconverged solve linear-bad, a converged direct solve nonlinear-bad,
// and a converged FGMRES-CPR solve nonlinear-bad when its pressure stage was strong
// (average reduction <= 0.25 and last <= 0.5, :50-51), mixed otherwise.
static int retry_class(const LinearOutcome& l) {
  if (!l.have) return RSF_RETRY_MIXED;
  if (!l.converged) return RSF_RETRY_LINEAR_BAD;
  if (l.direct) return RSF_RETRY_NONLINEAR_BAD;
  const bool strong = l.apps > 0 && l.avg <= 0.25 && l.last <= 0.5;
  return strong ? RSF_RETRY_NONLINEAR_BAD : RSF_RETRY_MIXED;
}

// rs/fim/newton/convergence.rs:793-806: the family of the largest scaled residual in the
// order water, oil, gas, well, perforation; a later family wins only when strictly larger
// (NaN never does). The well families exist only with wells.
static int dominant_family(const double f[5], bool wells) {
  int g = 0;
  for (int k = 1; k < (wells ? 5 : 3); ++k)
    if (f[k] > f[g]) g = k;
  return g;
}

// rs/fim/newton/convergence.rs:734-790 the row of a family's peak in the reference's
// unknown numbering (3 per cell, then one per well, then one per completion) from the
// peak's item (cell, well or completion index).
static long family_row(int family, long item, long n, long n_wells) {
  return family < 3 ? 3 * item + family : family == 3 ? 3 * n + item : 3 * n + n_wells + item;
}

// rs/fim/newton/convergence.rs:478-612 the OPM verdict of a Newton record: every
// component's CNV within 1e-2 (or within 1.0 when the cells over 1e-2 hold less than 3 % of
// the pore volume) and MB within 1e-7; with the final-iteration tiers CNV 1.0 and MB 1e-6.
bool Sim::opm_verdict(const double* rc, bool relax_final) const {
  const double* cnv = rc + kRecCnv;
  const double* mb = rc + kRecMb;
  const bool relaxpv = rc[kRecViolFrac] < 0.03;
  const double mbt = relax_final ? 1e-6 : 1e-7, cnvt = relax_final ? 1.0 : 1e-2;
  bool ok = true;
  for (int c = 0; c < 3; ++c) ok = ok && (cnv[c] <= cnvt || (relaxpv && cnv[c] <= 1.0)) && mb[c] <= mbt;
  return ok;
}

Sim::SubstepResult Sim::run_timestep(double dt, std::vector<double>& bhp, std::vector<double>& q) {
  SubstepResult res{false, 0, 0, 0.0};
  LinearOutcome lo;
  const int W = (int)topo.size(), Pn = (int)wells.size();
  // f: the five family maxima, item: the item of each family's peak
  auto fail = [&](const double f[5], const double item[5]) {
    res.fail_class = retry_class(lo);
    res.fail_family = dominant_family(f, !topo.empty());
    res.fail_row = family_row(res.fail_family, (long)item[res.fail_family], n, W);
  };
  double h1[5], h2[5], item1[5];
  for (int k = 0; k < 5; ++k) h1[k] = h2[k] = INFINITY, item1[k] = 0.0;
  double relax = 1.0;
  double rc[kRecCount];
  int maxnp = 1;
  for (auto& t : topo) maxnp = std::max(maxnp, (int)t.perfs.size());
  const size_t lds_relax = (size_t)kWellRelaxArrays * maxnp * sizeof(double);
  const size_t lds_inner = (size_t)kWellInnerArrays * maxnp * sizeof(double);
  if (W && (lds_relax > 65536 || lds_inner > 65536))
    throw HipError("a well has more than 910 completions (LDS limit of the well kernels)");
  hipLaunchKernelGGL((k_fill<double>), dim3(1), dim3(64), 0, stream, (int)kRecCount, rec.p, 0.0);
  upload_wells(bhp, q);
  auto finish = [&]() {
    if (W) {
      std::vector<double> qd(std::max(Pn, 1));
      w_bhp.download(bhp.data(), W, stream);
      if (Pn) w_q.download(qd.data(), Pn, stream);
      sync();
      for (int t = 0; t < Pn; ++t) q[dperf_order[t]] = qd[t];
    }
    return res;
  };
  for (int it = 0; it < opt.max_newton; ++it) {
    auto t0 = std::chrono::steady_clock::now();
    assemble(dt, true);
    newton_record(dt, rc);
    auto t1 = std::chrono::steady_clock::now();
    stats.assembly_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    res.asm_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (rc[kRecBad] != 0.0) {  // the previous update left a non-finite state
      res.converged = false;
      res.newton = it;
      fail(h1, item1);  // the residual of the iterate that took the update
      return finish();
    }
    const double* fam = rc + kRecFam;
    const double* mb = rc + kRecMb;
    double fm[5] = {fam[0], fam[1], fam[2], rc[kRecTail], rc[kRecTail + 1]};
    // rs/fim/newton/damping.rs:716-768 oscillation relaxation
    int osc = 0;
    for (int k = 0; k < 5; ++k) {
      double f0 = fm[k], f1 = h1[k], f2 = h2[k];
      if (std::isfinite(f0) && std::isfinite(f1) && std::isfinite(f2) && f0 > 0.0) {
        double d1 = std::fabs(f0 - f2) / f0, d2 = std::fabs(f0 - f1) / f0;
        if (d1 < 0.2 && 0.2 < d2) osc++;
      }
    }
    if (osc >= 1) relax = std::max(relax - 0.1, 0.5);
    for (int k = 0; k < 5; ++k) {
      h2[k] = h1[k];
      h1[k] = fm[k];
      item1[k] = rc[kRecItem + k];
    }
    bool final_it = it + 1 == opt.max_newton;
    double cell_norm = std::max(fam[0], std::max(fam[1], fam[2]));
    res.scaled = cell_norm;
    res.mb_norm = std::max(mb[0], std::max(mb[1], mb[2]));
    bool accept;
    if (opt.residual_tol > 0.0) {
      accept = it >= opt.min_newton && cell_norm <= opt.residual_tol && mb[0] <= opt.mb_tol && mb[1] <= opt.mb_tol &&
               mb[2] <= opt.mb_tol;
    } else {
      accept = it >= opt.min_newton && opm_verdict(rc, final_it);
      // rs/fim/newton.rs:1262-1285: with the nested well solve on, every well must also pass
      // its own convergence test at this iterate
      if (accept && nested_wells && W) accept = rc[kRecWellsOk] != 0.0;
    }
    if (accept) {
      res.converged = true;
      res.newton = it;
      return finish();
    }
    // rhs = -R, reduced by the well Schur complement
    auto t2 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL((k_fill<double>), dim3(nblk(3L * n)), dim3(kBlock), 0, stream, 3 * n, b.p, 0.0);
    hipLaunchKernelGGL((k_axpy<double, double>), dim3(nblk(3L * n)), dim3(kBlock), 0, stream, 3 * n, -1.0, R.p, b.p);
    double rhs_full = std::sqrt(rc[kRecCells2] + rc[kRecTail2]);
    if (W) hipLaunchKernelGGL(k_wells_rhs, dim3(W), dim3(64), 0, stream, n, well_dev(), R.p, b.p);
    int its = 0;
    double fr = 0.0;
    bool direct = false;
    const bool lconv = solve_reduced(rhs_full, its, fr, direct);
    res.linear += its;
    res.lin_solves++;
    res.last_lin_its = its;
    res.last_lin_conv = lconv ? 1 : 0;
    lo.have = true;
    lo.converged = lconv;
    lo.direct = direct;
    lo.apps = cpr_apps;
    lo.avg = cpr_apps ? cpr_ratio_sum / cpr_apps : 1.0;
    lo.last = cpr_apps ? cpr_ratio_last : 1.0;
    // rs/fim/linear/capture.rs: FIM_CAPTURE_SEQUENCE_DIR (every solve) / FIM_CAPTURE_DIR (failed solves)
    if (capture_mode == 1 || (capture_mode == 2 && !lconv)) write_capture(it, lconv ? "sequence" : "linear-not-converged");
    auto t3 = std::chrono::steady_clock::now();
    const double lin_ms = std::chrono::duration<double, std::milli>(t3 - t2).count();
    stats.linear_ms += lin_ms;
    res.lin_ms += lin_ms;
    double red = fr / std::max(rhs_full, 1e-30);
    if (!lconv) {
      // rs/fim/newton/convergence.rs:385-389 relaxed linear acceptance
      if (!(std::isfinite(red) && red < 0.01)) {
        res.converged = false;
        res.newton = it + 1;
        fail(fm, rc + kRecItem);
        return finish();
      }
    }
    // cells: per-cell chop + update
    res.updates++;
    hipLaunchKernelGGL(k_chop_update, dim3(nblk(n)), dim3(kBlock), 0, stream, n, relax, x_reg.p, dxv.p, x_p.p, x_sw.p,
                       x_hc.p);
    // wells: tail recovery, BHP chop (dbhp <= |bhp|, floor 1 bar) and raw q update, control
    // bounds, then (rs/fim/state.rs:717-740 finish_well_update) the relaxation or, with the
    // flag on, the nested per-well Newton, and the bounds again
    const WellDev wd = well_dev();
    if (W) {
      hipLaunchKernelGGL(k_wells_recover, dim3(W), dim3(kWellBlock), 12 * (size_t)maxnp * sizeof(double), stream, n,
                         wd, dxv.p, tdx.p);
      hipLaunchKernelGGL(k_well_update, dim3(nblk((long)wd.n_wells + wd.n_perf)), dim3(kBlock), 0, stream, wd,
                         (const double*)tdx.p, relax);
    }
    const int nb = std::min(nblk(n), 1024);
    hipLaunchKernelGGL(k_max_finite, dim3(nb), dim3(kBlock), 0, stream, n, x_p.p, x_sw.p, x_hc.p, part.p);
    hipLaunchKernelGGL(k_well_bounds, dim3(1), dim3(kRecBlock), 0, stream, wd, nb, (const double*)part.p, bhp_min, bhp_max, 0,
                       rec.p);
    if (W) {
      if (nested_wells)
        hipLaunchKernelGGL(k_wells_inner, dim3(W), dim3(kWellBlock), lds_inner, stream, params(true), wd, 50, 1e-4, 1.0,
                           x_p.p, x_sw.p, x_hc.p, x_reg.p, w_inner_i.p, w_inner_peak.p);
      else
        hipLaunchKernelGGL(k_wells_relax, dim3(W), dim3(kWellBlock), lds_relax, stream, params(true), wd, x_p.p,
                           x_sw.p, x_hc.p, x_reg.p);
    }
    hipLaunchKernelGGL(k_well_bounds, dim3(1), dim3(kRecBlock), 0, stream, wd, nb, (const double*)part.p, bhp_min, bhp_max, 1,
                       rec.p);
    RSF_CHECK(hipGetLastError());
  }
  // rs/fim/newton.rs:3239-3300 post-loop check with relaxed tiers
  assemble(dt, true);
  newton_record(dt, rc);
  res.newton = opt.max_newton;
  if (rc[kRecBad] != 0.0) {
    res.converged = false;
    fail(h1, item1);
    return finish();
  }
  const double* fam = rc + kRecFam;
  const double* mb = rc + kRecMb;
  double cell_norm = std::max(fam[0], std::max(fam[1], fam[2]));
  res.scaled = cell_norm;
  res.mb_norm = std::max(mb[0], std::max(mb[1], mb[2]));
  if (opt.residual_tol > 0.0) {
    res.converged = cell_norm <= opt.residual_tol && mb[0] <= opt.mb_tol && mb[1] <= opt.mb_tol && mb[2] <= opt.mb_tol;
  } else {
    res.converged = opm_verdict(rc, true);
  }
  if (!res.converged) {
    const double f[5] = {fam[0], fam[1], fam[2], rc[kRecTail], rc[kRecTail + 1]};
    fail(f, rc + kRecItem);
  }
  return finish();
}

// rs/fim/wells_inner.rs:585-598 solve_wells_locally, batched: one workgroup per well over
// the current device iterate of the cells, with bhp and q in and out through the host and
// the per-well reports kept (the Newton loop launches k_wells_inner on its own device
// iterate directly).
void Sim::solve_wells_locally(std::vector<double>& bhp, std::vector<double>& q, int max_it) {
  const int W = (int)topo.size(), Pn = (int)dperf_order.size();
  if (!W) return;
  int maxnp = 1;
  for (auto& t : topo) maxnp = std::max(maxnp, (int)t.perfs.size());
  const size_t lds = (size_t)kWellInnerArrays * maxnp * sizeof(double);
  if (lds > 65536) throw HipError("a well has more than 910 completions (LDS limit of the inner well kernel)");
  upload_wells(bhp, q);
  hipLaunchKernelGGL(k_wells_inner, dim3(W), dim3(kWellBlock), lds, stream, params(true), well_dev(), max_it, 1e-4,
                     1.0, x_p.p, x_sw.p, x_hc.p, x_reg.p, w_inner_i.p, w_inner_peak.p);
  RSF_CHECK(hipGetLastError());
  std::vector<double> qd(std::max(Pn, 1));
  inner_i.resize(2 * (size_t)W);
  inner_peak.resize(W);
  w_bhp.download(bhp.data(), W, stream);
  if (Pn) w_q.download(qd.data(), Pn, stream);
  w_inner_i.download(inner_i.data(), 2 * (size_t)W, stream);
  w_inner_peak.download(inner_peak.data(), W, stream);
  sync();
  for (int t = 0; t < Pn; ++t) q[dperf_order[t]] = qd[t];
}

void Sim::write_back(const std::vector<double>& bhp) {
  PhysParams P = params(true);
  hipLaunchKernelGGL(k_write_back, dim3(nblk(n)), dim3(kBlock), 0, stream, P, n, x_p.p, x_sw.p, x_hc.p, x_reg.p,
                     s_p.p, s_sw.p, s_so.p, s_sg.p, s_rs.p);
  state_dirty_host = true;
  for (size_t e = 0; e < wells.size(); ++e) wells[e].bhp = bhp[perf_well[e]];
  // a committed BHP becomes the next substep's BHP target (rs/fim/state.rs:909-913)
  upload_bhp_targets();
}

void Sim::report_sums(double out[7]) {
  PhysParams P = params(true);
  hipLaunchKernelGGL(k_report_sums, dim3(kReportBlocks), dim3(kBlock), 0, stream, P, n, d_pv.p, s_p.p, s_sw.p, s_so.p,
                     s_sg.p, s_rs.p, part.p);
  std::vector<double> h(7 * (size_t)kReportBlocks);
  part.download(h.data(), h.size(), stream);
  sync();
  for (int k = 0; k < 7; ++k) out[k] = 0.0;
  for (int b = 0; b < kReportBlocks; ++b)
    for (int k = 0; k < 7; ++k) out[k] += h[7 * (size_t)b + k];
}

// Accept of a substep (rs/fim/timestep.rs:1329-1353): the inventories before and after
// the write-back, then the step report.
void Sim::commit_substep(double dt, const std::vector<double>& bhp, const std::vector<double>& q) {
  double before[7] = {0, 0, 0, 0, 0, 0, 0};
  if (report_on) report_sums(before);
  write_back(bhp);
  if (!report_on) return;
  double after[7];
  report_sums(after);
  record_report(dt, bhp, q, after[0] - before[0], before[1] - after[1], after[2] - before[2], after);
}

// rs/reporting.rs:553-734 record_fim_step_report at the accepted state (now written
// back): per-perforation rates from rs/fim/wells.rs:834-856 (producer phase split at the
// perforated cell) and :1634-1666 (component rates at surface conditions), the
// rate-controlled wells held at their BHP limit, cumulative balances and averages.
void Sim::record_report(double dt, const std::vector<double>& bhp, const std::vector<double>& q, double d_water,
                        double oil_removed, double d_gas, const double sums[7]) {
  const int W = (int)topo.size(), np = (int)dperf_order.size();
  std::vector<double> g(4 * (size_t)std::max(np, 1));
  if (np) {
    hipLaunchKernelGGL(k_gather4, dim3(nblk(np)), dim3(kBlock), 0, stream, np, d_perf_cells.p, s_p.p, s_sw.p, s_sg.p,
                       s_rs.p, d_gather.p);
    d_gather.download(g.data(), 4 * (size_t)np, stream);
    sync();
  }
  const PhysParams P = params(false);
  int prod_rc = 0, inj_rc = 0, prod_lim = 0, inj_lim = 0;
  for (int w = 0; w < W; ++w) {
    const ControlState c = control(w);
    if (!(c.enabled && c.rate)) continue;
    const bool inj = topo[w].injector;
    const bool lim = inj ? bhp[w] >= c.limit - 1e-6 : bhp[w] <= c.limit + 1e-6;
    (inj ? inj_rc : prod_rc)++;
    if (lim) (inj ? inj_lim : prod_lim)++;
  }
  double prod_oil = 0, prod_liq = 0, prod_liq_res = 0, inj_tot = 0, inj_res = 0, inj_w_res = 0, prod_w_res = 0;
  double inj_w_sc = 0, prod_w_sc = 0, prod_gas = 0, inj_g_sc = 0;
  for (int t = 0; t < np; ++t) {
    const int e = dperf_order[t], w = perf_well[e];
    const double qe = q[e], p = g[4 * t], sw = g[4 * t + 1], sg = g[4 * t + 2], rs = g[4 * t + 3];
    const bool table3 = three_phase && has_table;
    double bo, mu;
    if (table3) pvt_oil<double>(P, p, rs, bo, mu);
    else bo = base_bo<double>(P, p);
    const double bg = table3 ? pvt_saturated<double>(P, p).bg : 1.0;
    const double ibw = inv_bw<double>(P, p);
    double comp[3];
    if (topo[w].injector) {
      if (fluid_of(w) == 0) {
        comp[0] = qe * ibw;
        comp[1] = comp[2] = 0.0;
      } else {
        comp[0] = comp[1] = 0.0;
        comp[2] = qe / std::max(bg, 1e-9);
      }
      inj_res += std::max(-qe, 0.0);
      if (fluid_of(w) == 0) {
        inj_tot += std::max(-comp[0], 0.0);
        inj_w_res += std::max(-qe, 0.0);
        inj_w_sc += std::max(-comp[0], 0.0);
      } else {
        inj_tot += std::max(-comp[2], 0.0);
        inj_g_sc += std::max(-comp[2], 0.0);
      }
      continue;
    }
    const Mobility<double> mob = mobilities<double>(P, sw, sg, p, rs);
    const double lw = std::max(mob.w, 0.0), lo = std::max(mob.o, 0.0), lg = std::max(mob.g, 0.0);
    const double lt = std::max(lw + lo + lg, 2.220446049250313e-16);
    auto c01 = [](double x) { return std::min(std::max(x, 0.0), 1.0); };
    const double fw = c01(lw / lt), fo = c01(lo / lt), fg = c01(lg / lt);
    comp[0] = qe * fw * ibw;
    comp[1] = qe * fo / std::max(std::max(bo, 1e-9), 1e-9);
    comp[2] = qe * fg / std::max(std::max(bg, 1e-9), 1e-9) + comp[1] * std::max(rs, 0.0);
    prod_liq_res += std::max(qe, 0.0);
    prod_w_res += std::max(qe, 0.0) * fw;
    prod_w_sc += std::max(comp[0], 0.0);
    prod_oil += std::max(comp[1], 0.0);
    prod_liq += std::max(comp[0], 0.0) + std::max(comp[1], 0.0);
    prod_gas += std::max(comp[2], 0.0);
  }
  cum_injection_m3 += inj_w_res * dt;
  cum_production_m3 += prod_w_res * dt;
  cum_mb_water += (inj_w_sc - prod_w_sc) * dt - d_water;
  cum_mb_oil += prod_oil * dt - oil_removed;
  if (three_phase) cum_mb_gas += (inj_g_sc - prod_gas) * dt - d_gas;
  TimePoint tp{};
  tp.time = time + dt;
  tp.total_production_oil = prod_oil;
  tp.total_production_liquid = prod_liq;
  tp.total_production_liquid_reservoir = prod_liq_res;
  tp.total_injection = inj_tot;
  tp.total_injection_reservoir = inj_res;
  tp.material_balance_error_m3 = std::fabs(cum_mb_water);
  tp.material_balance_error_oil_m3 = std::fabs(cum_mb_oil);
  tp.avg_reservoir_pressure = sums[4] > 0.0 ? sums[3] / sums[4] : 0.0;
  tp.avg_water_saturation = n > 0 ? sums[5] / n : 0.0;
  tp.total_production_gas = prod_gas;
  tp.avg_gas_saturation = n > 0 ? sums[6] / n : 0.0;
  tp.material_balance_error_gas_m3 = std::fabs(cum_mb_gas);
  tp.producing_gor = prod_oil > 1e-9 ? prod_gas / prod_oil : 0.0;  // MIN_GOR_OIL_RATE_SC_DAY, reporting.rs:21
  tp.producer_bhp_limited_fraction = prod_rc > 0 ? double(prod_lim) / prod_rc : 0.0;
  tp.injector_bhp_limited_fraction = inj_rc > 0 ? double(inj_lim) / inj_rc : 0.0;
  sweep_metrics(tp);
  rate_history.push_back(tp);
}

// rs/reporting.rs:52-123 compute_sweep_metrics over the written-back state, when a sweep
// configuration is set (else every sweep field is NaN, the reference's None)
void Sim::sweep_metrics(TimePoint& tp) {
  const double nan = std::numeric_limits<double>::quiet_NaN();
  tp.sweep_e_a = tp.sweep_e_v = tp.sweep_e_vol = tp.sweep_mobile_oil_recovered = nan;
  if (!sweep.on) return;
  const int nxy = nx * ny;
  const int nb = std::max(1, std::min(kReportBlocks, (nxy + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_sweep_sums, dim3(nb), dim3(kBlock), 0, stream, nx, ny, nz, sweep.swept_threshold,
                     sweep.residual_oil_saturation, s_sw.p, s_so.p, part.p);
  std::vector<double> h(3 * (size_t)nb);
  part.download(h.data(), h.size(), stream);
  sync();
  double swept = 0.0, cols = 0.0, remaining = 0.0;
  for (int b = 0; b < nb; ++b) {
    swept += h[3 * (size_t)b];
    cols += h[3 * (size_t)b + 1];
    remaining += h[3 * (size_t)b + 2];
  }
  const double total = double(n);
  tp.sweep_e_vol = total > 0.0 ? swept / total : 0.0;
  const double e_a = nxy > 0 ? cols / nxy : 0.0;
  if (sweep.geometry == 0) {
    tp.sweep_e_a = e_a;
    tp.sweep_e_v = 1.0;
  } else if (sweep.geometry == 1) {
    tp.sweep_e_a = 1.0;
    tp.sweep_e_v = tp.sweep_e_vol;
  } else {
    const double per_cell = std::max(sweep.initial_oil_saturation - sweep.residual_oil_saturation, 0.0);
    const double initial = total * per_cell;
    tp.sweep_mobile_oil_recovered =
        initial > 1e-12 ? std::min(std::max(1.0 - remaining / initial, 0.0), 1.0) : 0.0;
  }
}

void Sim::upload_bhp_targets() {
  if (topo.empty()) return;
  std::vector<double> bt(topo.size());
  for (size_t w = 0; w < topo.size(); ++w) bt[w] = control((int)w).bhp_target;
  w_bhpt.upload(bt, stream);
}

double Sim::rel_change() {  // rs/fim/timestep.rs:436-466
  int nb = std::min(nblk(n), 1024);
  hipLaunchKernelGGL(k_rel_change, dim3(nb), dim3(kBlock), 0, stream, n, x0_p.p, x0_sw.p, x0_hc.p, x0_reg.p, x_p.p,
                     x_sw.p, x_hc.p, x_reg.p, part.p);
  std::vector<double> h(2 * (size_t)nb);
  part.download(h.data(), h.size(), stream);
  sync();
  double a = 0, c = 0;
  for (int k = 0; k < nb; ++k) {
    a += h[2 * k];
    c += h[2 * k + 1];
  }
  return c > 0.0 ? a / c : 0.0;
}

int Sim::run_substep(double dt, SubstepResult& out, const double* init_cells, const int* init_regimes,
                     const double* init_bhp, const double* init_q) {
  ensure_device();
  adjoint_before_substep();
  std::vector<double> bhp, q;
  init_substep(bhp, q);
  // rs/fim/newton.rs:970-976: the previous state stays the stored one (accumulation terms,
  // regime freeze of x0); only the Newton iterate starts elsewhere
  if (init_cells) {
    std::vector<double> p(n), sw(n), hc(n);
    for (int i = 0; i < n; ++i) {
      p[i] = init_cells[3 * i];
      sw[i] = init_cells[3 * i + 1];
      hc[i] = init_cells[3 * i + 2];
    }
    x_p.upload(p, stream);
    x_sw.upload(sw, stream);
    x_hc.upload(hc, stream);
  }
  if (init_regimes) x_reg.upload(init_regimes, n, stream);
  if (init_bhp) std::copy(init_bhp, init_bhp + bhp.size(), bhp.begin());
  if (init_q) std::copy(init_q, init_q + q.size(), q.begin());
  out = run_timestep(dt, bhp, q);
  stats.newton += out.newton;
  if (!out.converged) return 2;
  adjoint_after_substep(dt, bhp, q);
  commit_substep(dt, bhp, q);
  last_bhp = bhp;
  last_q = q;
  time += dt;
  accepted_dts.push_back(dt);
  stats.substeps++;
  return 0;
}

// rs/fim/timestep.rs:988-1890 substep controller (OPM PID + iteration growth, retry x0.33),
// with its FimStepStats record (timestep.rs:742-800; reporting.rs:141-243)
int Sim::step(double target) {
  warning.clear();  // rs/fim/timestep.rs:1022: every step starts without a solver warning
  ensure_device();
  auto t0 = std::chrono::steady_clock::now();
  refresh_head_offsets();
  double done = 0.0, last_dt = target, growth = 3.0;
  double err[3] = {0.1, 0.1, 0.1};
  int substeps = 0;
  int rc = 0;
  StepRecord sr{};
  rsf_step_stats& st = sr.s;
  st.target_dt_days = target;
  st.min_accepted_dt_days = st.max_accepted_dt_days = st.last_accepted_dt_days = NAN;
  st.last_growth_factor = NAN;
  st.growth_limiter = st.last_retry_class = st.last_retry_dominant_family = -1;
  st.last_retry_dominant_row = -1;
  const double asm0 = stats.assembly_ms, lin0 = stats.linear_ms;
  const int max_substeps = 100000;  // rs/fim/timestep.rs:991 MAX_SUBSTEPS
  while (done < target - 1e-12 && substeps < max_substeps) {
    double remaining = target - done;
    double trial;
    if (substeps == 0) trial = time <= 1e-12 ? std::min(remaining, remaining <= 1.0 + 1e-12 ? 0.25 : 1.0) : remaining;
    else trial = std::min(remaining, last_dt * growth);
    int retries = 0;
    for (;;) {
      adjoint_before_substep();
      std::vector<double> bhp, q;
      init_substep(bhp, q);
      auto ta = std::chrono::steady_clock::now();
      SubstepResult r = run_timestep(trial, bhp, q);
      const double attempt_ms =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count();
      stats.newton += r.newton;
      rsf_step_rung rung{};
      rung.substep = substeps;
      rung.retried = r.converged ? 0 : 1;
      rung.newton_iterations = r.newton;
      rung.linear_iterations = r.last_lin_its;
      rung.last_linear_converged = r.last_lin_conv;
      rung.linear_solve_count = r.lin_solves;
      rung.linear_iteration_count = r.linear;
      rung.dt_days = trial;
      rung.linear_solve_ms = r.lin_ms;
      rung.retry_class = r.converged ? -1 : r.fail_class;
      rung.dominant_family = r.converged ? -1 : r.fail_family;
      rung.dominant_row = r.converged ? -1 : r.fail_row;
      sr.rungs.push_back(rung);
      st.solver_ms += attempt_ms;
      if (r.converged) {
        st.accepted_solver_ms += attempt_ms;
        adjoint_after_substep(trial, bhp, q);
        double rc_ = rel_change();
        err[0] = err[1];
        err[1] = err[2];
        err[2] = rc_;
        double tol = 0.1, dt_pid;
        if (err[2] > tol) dt_pid = trial * tol / err[2];
        else if (err[0] == 0.0 || err[1] == 0.0 || err[2] == 0.0) dt_pid = 1e300;
        else
          dt_pid = trial * std::pow(err[1] / err[2], 0.075) * std::pow(tol / err[2], 0.175) *
                   std::pow(err[1] * err[1] / (err[0] * err[2]), 0.01);
        double its = r.newton, tg = 8.0;
        double dt_it = its > tg ? trial / (1.0 + (its - tg) / tg * 1.0) : trial * (1.0 + (tg - its) / tg * 3.2);
        // rs/fim/timestep.rs:502-528: the tightest of the 3x ceiling, the PID and the
        // iteration-count proposals, then 2x after a retry
        double f = 3.0;
        int lim = RSF_LIMIT_MAX_GROWTH;
        const double cand[2] = {dt_pid / trial, dt_it / trial};
        for (int k = 0; k < 2; ++k)
          if (cand[k] < f) {
            f = cand[k];
            lim = k == 0 ? RSF_LIMIT_PID : RSF_LIMIT_ITER;
          }
        if (retries > 0 && 2.0 < f) {
          f = 2.0;
          lim = RSF_LIMIT_RESTART_GROWTH;
        }
        growth = f;
        st.last_growth_factor = f;
        st.growth_limiter = lim;
        st.min_accepted_dt_days = std::isnan(st.min_accepted_dt_days) ? trial : std::min(st.min_accepted_dt_days, trial);
        st.max_accepted_dt_days = std::isnan(st.max_accepted_dt_days) ? trial : std::max(st.max_accepted_dt_days, trial);
        st.last_accepted_dt_days = trial;
        last_dt = trial;
        commit_substep(trial, bhp, q);
        last_bhp = bhp;
        last_q = q;
        time += trial;
        done += trial;
        accepted_dts.push_back(trial);
        stats.substeps++;
        substeps++;
        break;
      }
      st.retry_solver_ms += attempt_ms;
      st.last_retry_class = r.fail_class;
      st.last_retry_dominant_family = r.fail_family;
      st.last_retry_dominant_row = r.fail_row;
      if (r.fail_class == RSF_RETRY_LINEAR_BAD) st.linear_bad_retries++;
      else if (r.fail_class == RSF_RETRY_NONLINEAR_BAD) st.nonlinear_bad_retries++;
      else st.mixed_retries++;
      retries++;
      stats.retries++;
      // rs/fim/timestep.rs:1589-1767: the next trial, then the collapse and retry-budget
      // exits (16 retries per substep, :1000); the reported time is the reference's
      // time_days + time_stepped
      const double next = trial * 0.33;
      char msg[160];
      if (!std::isfinite(next) || next <= 1e-12) {
        snprintf(msg, sizeof msg, "FIM Newton step collapsed timestep at t=%.6f days after %d iterations", time + done,
                 r.newton);
        warning = msg;
        rc = 2;
        break;
      }
      if (retries >= 16) {
        snprintf(msg, sizeof msg, "FIM Newton step exceeded retry budget at t=%.6f days after %d retries", time + done,
                 retries);
        warning = msg;
        rc = 2;
        break;
      }
      trial = next;
    }
    if (rc) break;
  }
  if (!rc && substeps == max_substeps && done < target - 1e-12) {  // rs/fim/timestep.rs:1772-1778
    char msg[160];
    snprintf(msg, sizeof msg,
             "FIM adaptive timestep hit MAX_SUBSTEPS before completing requested dt (advanced %.6f of %.6f days)", done,
             target);
    warning = msg;
    rc = 2;
  }
  sync();
  stats.total_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  st.time_days = time;
  st.advanced_dt_days = done;
  st.accepted_substeps = substeps;
  st.n_rungs = (int)sr.rungs.size();
  st.assembly_ms = stats.assembly_ms - asm0;
  st.linear_solve_ms = stats.linear_ms - lin0;
  step_stats.push_back(std::move(sr));
  return rc;
}

#include "rsf_adjoint.hpp"
#include "rsf_capture.hpp"

}  // namespace rsf

// ======================================================================= C ABI
using rsf::Sim;

namespace {
thread_local std::string g_err;

template <class F>
int guard(rsf_sim* h, F&& f) {
  if (!h || !h->s) return RSF_ERR_INVALID;
  try {
    return f(*h->s);
  } catch (const rsf::HipError& e) {
    h->s->warning = e.what();
    g_err = e.what();
    return RSF_ERR_HIP;
  } catch (const std::exception& e) {
    h->s->warning = e.what();
    h->s->invalid_gen++;
    g_err = e.what();
    return RSF_ERR_INVALID;
  }
}

// guard with the entry point's name: an invalid argument without a message of its own
// reports "<entry point>: invalid argument" in rsf_last_warning
template <class F>
int guard(rsf_sim* h, const char* fn, F&& f) {
  if (!h || !h->s) return RSF_ERR_INVALID;
  const long gen = h->s->invalid_gen;
  const int rc = guard(h, std::forward<F>(f));
  if (rc == RSF_ERR_INVALID && h->s->invalid_gen == gen) h->s->warning = std::string(fn) + ": invalid argument";
  return rc;
}

// A setter's Err(String): the message goes to rsf_last_warning
int invalid(Sim& s, const std::string& msg) {
  s.warning = msg;
  s.invalid_gen++;
  return RSF_ERR_INVALID;
}

std::string num(double v) {
  char b[40];
  std::snprintf(b, sizeof(b), "%g", v);
  return b;
}

std::string lower(const char* s) {
  std::string o = s ? s : "";
  for (auto& c : o) c = (char)tolower(c);
  while (!o.empty() && isspace((unsigned char)o.back())) o.pop_back();
  while (!o.empty() && isspace((unsigned char)o.front())) o.erase(o.begin());
  return o;
}
}  // namespace

extern "C" {

int rsf_create(int nx, int ny, int nz, double porosity, rsf_sim** out) {
  if (!out || nx <= 0 || ny <= 0 || nz <= 0 || !(porosity > 0.0)) return RSF_ERR_INVALID;
  try {
    auto* h = new rsf_sim{new Sim(nx, ny, nz, porosity)};
    *out = h;
    return RSF_OK;
  } catch (...) {
    return RSF_ERR_INVALID;
  }
}

void rsf_destroy(rsf_sim* h) {
  if (!h) return;
  delete h->s;
  delete h;
}

int rsf_set_cell_dimensions(rsf_sim* h, double dx, double dy, double dz) {
  return guard(h, __func__, [&](Sim& s) {
    // rs/frontend.rs:622-636
    if (!std::isfinite(dx) || !std::isfinite(dy) || !std::isfinite(dz))
      return invalid(s, "Cell dimensions must be finite numbers");
    if (!(dx > 0) || !(dy > 0) || !(dz > 0))
      return invalid(s, "Cell dimensions must be positive, got dx=" + num(dx) + ", dy=" + num(dy) + ", dz=" + num(dz));
    s.dx = dx; s.dy = dy;
    s.dz.assign(s.nz, dz);
    s.dev_ready = false;
    return RSF_OK;
  });
}

int rsf_set_cell_dimensions_per_layer(rsf_sim* h, double dx, double dy, const double* dz, int nz) {
  return guard(h, __func__, [&](Sim& s) {
    if (!(dx > 0) || !(dy > 0) || nz != s.nz || !dz) return RSF_ERR_INVALID;
    for (int k = 0; k < nz; ++k)
      if (!(dz[k] > 0)) return RSF_ERR_INVALID;
    s.dx = dx; s.dy = dy;
    s.dz.assign(dz, dz + nz);
    s.dev_ready = false;
    return RSF_OK;
  });
}

int rsf_set_permeability_field(rsf_sim* h, const double* kx, const double* ky, const double* kz, int n) {
  return guard(h, __func__, [&](Sim& s) {
    if (n != s.n || !kx || !ky || !kz) return RSF_ERR_INVALID;
    for (int i = 0; i < n; ++i)
      if (!(kx[i] > 0) || !(ky[i] > 0) || !(kz[i] > 0) || !std::isfinite(kx[i] + ky[i] + kz[i])) return RSF_ERR_INVALID;
    s.kx.assign(kx, kx + n); s.ky.assign(ky, ky + n); s.kz.assign(kz, kz + n);
    s.dev_ready = false;
    return RSF_OK;
  });
}

int rsf_set_permeability_per_layer(rsf_sim* h, const double* kx, const double* ky, const double* kz, int nz) {
  return guard(h, __func__, [&](Sim& s) {
    if (nz != s.nz || !kx || !ky || !kz)
      return invalid(s, "Permeability vectors must have length equal to nz (" + std::to_string(s.nz) + ")");
    for (int k = 0; k < nz; ++k) {
      if (!std::isfinite(kx[k]) || !std::isfinite(ky[k]) || !std::isfinite(kz[k]))
        return invalid(s, "Permeability for layer " + std::to_string(k) + " must be finite");
      if (kx[k] <= 0.0 || ky[k] <= 0.0 || kz[k] <= 0.0)
        return invalid(s, "Permeability for layer " + std::to_string(k) + " must be positive, got px=" + num(kx[k]) +
                              ", py=" + num(ky[k]) + ", pz=" + num(kz[k]));
    }
    const int nxy = s.nx * s.ny;
    std::vector<double> x(s.n), y(s.n), z(s.n);
    for (int i = 0; i < s.n; ++i) {
      x[i] = kx[i / nxy];
      y[i] = ky[i / nxy];
      z[i] = kz[i / nxy];
    }
    return rsf_set_permeability_field(h, x.data(), y.data(), z.data(), s.n);
  });
}

int rsf_set_perm_field_log(rsf_sim* h, const double* ln_k, double kz_over_kh, int n) {
  return guard(h, __func__, [&](Sim& s) {
    if (n != s.n || !ln_k || !(kz_over_kh > 0) || !std::isfinite(kz_over_kh)) return RSF_ERR_INVALID;
    std::vector<double> k(n);
    for (int i = 0; i < n; ++i) {
      k[i] = std::exp(ln_k[i]);
      if (!(k[i] > 0) || !std::isfinite(k[i] * kz_over_kh)) return RSF_ERR_INVALID;
    }
    s.kx = k;
    s.ky = k;
    s.kz.resize(n);
    for (int i = 0; i < n; ++i) s.kz[i] = kz_over_kh * k[i];
    s.dev_ready = false;
    return RSF_OK;
  });
}

int rsf_set_porosity_field(rsf_sim* h, const double* phi, int n) {
  return guard(h, __func__, [&](Sim& s) {
    if (n != s.n || !phi) return RSF_ERR_INVALID;
    for (int i = 0; i < n; ++i)
      if (!(phi[i] > 0) || !(phi[i] <= 1.0)) return RSF_ERR_INVALID;
    s.phi.assign(phi, phi + n);
    s.dev_ready = false;
    return RSF_OK;
  });
}

int rsf_set_capillary_scale_field(rsf_sim* h, const double* scale, int n) {
  return guard(h, __func__, [&](Sim& s) {
    if (n != s.n || !scale) return RSF_ERR_INVALID;
    for (int i = 0; i < n; ++i)
      if (!(scale[i] >= 0) || !std::isfinite(scale[i])) return RSF_ERR_INVALID;
    s.pc_scale.assign(scale, scale + n);
    s.dev_ready = false;
    return RSF_OK;
  });
}

static int set_state(Sim& s) {
  if (s.dev_ready) s.upload_state();
  return RSF_OK;
}

int rsf_set_gas_redissolution_enabled(rsf_sim* h, int enabled) {
  return guard(h, __func__, [&](Sim& s) {
    s.gas_redissolution = enabled != 0;
    return RSF_OK;
  });
}

int rsf_set_initial_pressure(rsf_sim* h, double p) {
  return guard(h, __func__, [&](Sim& s) {
    s.download_state();
    std::fill(s.h_p.begin(), s.h_p.end(), p);
    // frontend.rs:611-618: the water PVT and the rock compressibility are referenced to it
    s.p_ref_w = p;
    s.p_ref_rock = p;
    s.dev_ready = false;
    return set_state(s);
  });
}

int rsf_set_initial_saturation(rsf_sim* h, double sw) {
  return guard(h, __func__, [&](Sim& s) {
    s.download_state();
    for (int i = 0; i < s.n; ++i) {
      s.h_sw[i] = sw;
      s.h_so[i] = std::max(1.0 - sw - s.h_sg[i], 0.0);
    }
    return set_state(s);
  });
}

int rsf_set_initial_gas_saturation(rsf_sim* h, double sg) {
  return guard(h, __func__, [&](Sim& s) {
    s.download_state();
    double g = std::min(std::max(sg, 0.0), 1.0);
    for (int i = 0; i < s.n; ++i) {
      double c = std::min(g, 1.0 - s.h_sw[i]);
      s.h_sg[i] = c;
      s.h_so[i] = std::max(1.0 - s.h_sw[i] - c, 0.0);
    }
    return set_state(s);
  });
}

int rsf_set_initial_rs(rsf_sim* h, double rs) {
  return guard(h, __func__, [&](Sim& s) {
    s.download_state();
    std::fill(s.h_rs.begin(), s.h_rs.end(), rs);
    return set_state(s);
  });
}

int rsf_set_initial_state_arrays(rsf_sim* h, const double* p, const double* sw, const double* sg, const double* rs,
                                 int n) {
  return guard(h, __func__, [&](Sim& s) {
    if (n != s.n || !p || !sw || !sg || !rs) return RSF_ERR_INVALID;
    s.download_state();
    for (int i = 0; i < n; ++i) {
      s.h_p[i] = p[i]; s.h_sw[i] = sw[i]; s.h_sg[i] = sg[i]; s.h_rs[i] = rs[i];
      s.h_so[i] = std::max(1.0 - sw[i] - sg[i], 0.0);
    }
    return set_state(s);
  });
}

int rsf_set_fluid_properties(rsf_sim* h, double mu_o, double mu_w) {
  return guard(h, __func__, [&](Sim& s) {
    // rs/frontend.rs:786-800
    if (!std::isfinite(mu_o) || !std::isfinite(mu_w)) return invalid(s, "Fluid viscosities must be finite numbers");
    if (!(mu_o > 0) || !(mu_w > 0))
      return invalid(s, "Fluid viscosities must be positive, got mu_o=" + num(mu_o) + ", mu_w=" + num(mu_w));
    s.mu_o = mu_o; s.mu_w = mu_w;
    return RSF_OK;
  });
}

int rsf_set_fluid_compressibilities(rsf_sim* h, double c_o, double c_w) {
  return guard(h, __func__, [&](Sim& s) {
    // rs/frontend.rs:802-816
    if (!std::isfinite(c_o) || !std::isfinite(c_w)) return invalid(s, "Fluid compressibilities must be finite numbers");
    if (!(c_o >= 0) || !(c_w >= 0))
      return invalid(s, "Fluid compressibilities must be non-negative, got c_o=" + num(c_o) + ", c_w=" + num(c_w));
    s.c_o = c_o; s.c_w = c_w;
    return RSF_OK;
  });
}

int rsf_set_fluid_densities(rsf_sim* h, double rho_o, double rho_w) {
  return guard(h, __func__, [&](Sim& s) {
    // rs/frontend.rs:771-784
    if (!std::isfinite(rho_o) || !std::isfinite(rho_w)) return invalid(s, "Fluid densities must be finite numbers");
    if (!(rho_o > 0) || !(rho_w > 0))
      return invalid(s, "Fluid densities must be positive, got rho_o=" + num(rho_o) + ", rho_w=" + num(rho_w));
    s.rho_o = rho_o; s.rho_w = rho_w;
    return RSF_OK;
  });
}

int rsf_set_rock_properties(rsf_sim* h, double c_r, double depth_ref, double b_o, double b_w) {
  return guard(h, __func__, [&](Sim& s) {
    // rs/frontend.rs:818-847
    if (!std::isfinite(c_r) || !std::isfinite(depth_ref) || !std::isfinite(b_o) || !std::isfinite(b_w))
      return invalid(s, "Rock properties must be finite numbers");
    if (!(c_r >= 0)) return invalid(s, "Rock compressibility must be non-negative, got " + num(c_r));
    if (!(b_o > 0) || !(b_w > 0))
      return invalid(s, "Volume expansion factors must be positive, got b_o=" + num(b_o) + ", b_w=" + num(b_w));
    s.c_r = c_r; s.depth_ref = depth_ref; s.b_o = b_o; s.b_w = b_w;
    s.dev_ready = false;
    return RSF_OK;
  });
}

int rsf_set_gas_fluid_properties(rsf_sim* h, double mu_g, double c_g, double rho_g) {
  return guard(h, __func__, [&](Sim& s) {
    if (!(mu_g > 0) || !(c_g >= 0) || !(rho_g > 0)) return RSF_ERR_INVALID;
    s.mu_g = mu_g; s.c_g = c_g; s.rho_g = rho_g;
    return RSF_OK;
  });
}

int rsf_set_rel_perm_props(rsf_sim* h, double s_wc, double s_or, double n_w, double n_o, double k_rw_max,
                           double k_ro_max) {
  return guard(h, __func__, [&](Sim& s) {
    // rs/frontend.rs:716-769
    for (double v : {s_wc, s_or, n_w, n_o, k_rw_max, k_ro_max})
      if (!std::isfinite(v)) return invalid(s, "Relative permeability parameters must be finite numbers");
    if (!(s_wc >= 0 && s_wc < 1)) return invalid(s, "S_wc must be in [0, 1), got " + num(s_wc));
    if (!(s_or >= 0 && s_or < 1)) return invalid(s, "S_or must be in [0, 1), got " + num(s_or));
    if (!(s_wc + s_or < 1))
      return invalid(s, "Invalid saturation endpoints: S_wc + S_or must be < 1.0, got " + num(s_wc + s_or));
    if (!(n_w > 0) || !(n_o > 0))
      return invalid(s, "Corey exponents must be positive, got n_w=" + num(n_w) + ", n_o=" + num(n_o));
    if (!(k_rw_max >= 0 && k_rw_max <= 1)) return invalid(s, "k_rw_max must be in [0, 1], got " + num(k_rw_max));
    if (!(k_ro_max > 0 && k_ro_max <= 1)) return invalid(s, "k_ro_max must be in (0, 1], got " + num(k_ro_max));
    s.scal = rsf::Scal2P{s_wc, s_or, n_w, n_o, k_rw_max, k_ro_max};
    return RSF_OK;
  });
}

int rsf_set_three_phase_rel_perm_props(rsf_sim* h, double s_wc, double s_or, double s_gc, double s_gr, double s_org,
                                       double n_w, double n_o, double n_g, double k_rw_max, double k_ro_max,
                                       double k_rg_max) {
  return guard(h, __func__, [&](Sim& s) {
    if (!(s_wc + s_or + s_gc + s_gr < 1.0) || !(s_wc + s_org < 1.0) || !(n_w > 0) || !(n_o > 0) || !(n_g > 0))
      return RSF_ERR_INVALID;
    s.scal3 = rsf::Scal3P{s_wc, s_or, s_gc, s_gr, s_org, n_w, n_o, n_g, k_rw_max, k_ro_max, k_rg_max};
    s.has_scal3 = true;
    // rs/frontend.rs:1139: new props drop any SWOF/SGOF tables
    if (!s.swof_h.empty()) s.dev_ready = false;
    s.swof_h.clear();
    s.sgof_h.clear();
    return RSF_OK;
  });
}

// rs/relperm.rs:82-124 validate_swof_rows / validate_sgof_rows: rows of 4 (s, kr, kro, pc);
// the capillary column may be NaN (None) and is not read
static std::string validate_scal_rows(const char* name, const char* sat, const double* rows, int n) {
  if (n < 2 || !rows) return std::string(name) + " table must contain at least two rows";
  double prev = -INFINITY;
  for (int i = 0; i < n; ++i) {
    const double x = rows[4 * i], a = rows[4 * i + 1], b = rows[4 * i + 2];
    if (!std::isfinite(x) || !std::isfinite(a) || !std::isfinite(b))
      return std::string(name) + " row " + std::to_string(i) + " must contain finite values";
    if (!(x >= 0.0 && x <= 1.0) || !(a >= 0.0 && a <= 1.0) || !(b >= 0.0 && b <= 1.0))
      return std::string(name) + " row " + std::to_string(i) + " must stay within [0, 1]";
    if (x <= prev) return std::string(name) + " " + sat + " must be strictly increasing at row " + std::to_string(i);
    prev = x;
  }
  return "";
}

int rsf_set_three_phase_scal_tables(rsf_sim* h, const double* swof, int n_swof, const double* sgof, int n_sgof) {
  return guard(h, __func__, [&](Sim& s) {
    std::string e = validate_scal_rows("SWOF", "saturation", swof, n_swof);
    if (e.empty()) e = validate_scal_rows("SGOF", "saturation", sgof, n_sgof);
    if (!e.empty()) return invalid(s, e);
    if (!s.has_scal3) return invalid(s, "Three-phase relperm props must be configured before SWOF/SGOF tables");
    s.swof_h.clear();
    s.sgof_h.clear();
    for (int i = 0; i < n_swof; ++i) s.swof_h.insert(s.swof_h.end(), {swof[4 * i], swof[4 * i + 1], swof[4 * i + 2]});
    for (int i = 0; i < n_sgof; ++i) s.sgof_h.insert(s.sgof_h.end(), {sgof[4 * i], sgof[4 * i + 1], sgof[4 * i + 2]});
    s.dev_ready = false;
    return RSF_OK;
  });
}

int rsf_set_capillary_params(rsf_sim* h, double p_entry, double lambda) {
  return guard(h, __func__, [&](Sim& s) {
    // rs/frontend.rs:852-867
    if (!std::isfinite(p_entry) || !std::isfinite(lambda)) return invalid(s, "Capillary parameters must be finite numbers");
    if (!(p_entry >= 0)) return invalid(s, "Capillary entry pressure must be non-negative, got " + num(p_entry));
    if (!(lambda > 0)) return invalid(s, "Capillary lambda must be positive, got " + num(lambda));
    s.pc_entry = p_entry; s.pc_lambda = lambda;
    return RSF_OK;
  });
}

int rsf_set_gas_oil_capillary_params(rsf_sim* h, double p_entry, double lambda) {
  return guard(h, __func__, [&](Sim& s) {
    if (!(p_entry >= 0) || !(lambda > 0)) return RSF_ERR_INVALID;
    s.pcog_entry = p_entry; s.pcog_lambda = lambda; s.has_pc_og = true;
    return RSF_OK;
  });
}

int rsf_set_three_phase_mode_enabled(rsf_sim* h, int enabled) {
  return guard(h, __func__, [&](Sim& s) { s.three_phase = enabled != 0; s.dev_ready = false; return RSF_OK; });
}

int rsf_set_gravity_enabled(rsf_sim* h, int enabled) {
  return guard(h, __func__, [&](Sim& s) { s.gravity = enabled != 0; return RSF_OK; });
}

int rsf_set_fim_nested_well_solve(rsf_sim* h, int enabled) {
  return guard(h, __func__, [&](Sim& s) { s.nested_wells = enabled != 0; return RSF_OK; });
}

int rsf_set_fim_corey_table_points(rsf_sim* h, int points) {
  return guard(h, __func__, [&](Sim& s) { s.corey_points = points; return RSF_OK; });
}

int rsf_set_injected_fluid(rsf_sim* h, const char* fluid) {
  return guard(h, __func__, [&](Sim& s) {
    std::string f = lower(fluid);
    if (f == "water") s.injected_fluid = 0;
    else if (f == "gas") s.injected_fluid = 1;
    else return RSF_ERR_INVALID;
    s.dev_ready = false;
    return RSF_OK;
  });
}

int rsf_set_pvt_table(rsf_sim* h, const double* rows, int n_rows) {
  return guard(h, __func__, [&](Sim& s) {
    if (!rows || n_rows <= 0) return RSF_ERR_INVALID;
    for (int i = 0; i < 6 * n_rows; ++i)
      if (!std::isfinite(rows[i])) return RSF_ERR_INVALID;
    s.download_state();
    s.set_table(rows, n_rows);
    if (s.state_on_dev) s.upload_state();
    return RSF_OK;
  });
}

int rsf_set_well_control_modes(rsf_sim* h, const char* injector_mode, const char* producer_mode) {
  return guard(h, __func__, [&](Sim& s) {
    s.inj_rate = lower(injector_mode) == "rate";
    s.prod_rate = lower(producer_mode) == "rate";
    s.dev_ready = false;
    return RSF_OK;
  });
}

int rsf_set_injector_enabled(rsf_sim* h, int enabled) {
  return guard(h, __func__, [&](Sim& s) { s.inj_enabled = enabled != 0; s.dev_ready = false; return RSF_OK; });
}

int rsf_set_target_well_rates(rsf_sim* h, double injector_rate, double producer_rate) {
  return guard(h, __func__, [&](Sim& s) {
    if (!(injector_rate >= 0) || !(producer_rate >= 0)) return RSF_ERR_INVALID;
    s.tgt_inj = injector_rate; s.tgt_prod = producer_rate;
    s.dev_ready = false;
    return RSF_OK;
  });
}

int rsf_set_target_well_surface_rates(rsf_sim* h, double injector_rate, double producer_rate) {
  return guard(h, __func__, [&](Sim& s) {
    if (!(injector_rate >= 0) || !(producer_rate >= 0)) return RSF_ERR_INVALID;
    s.has_tgt_inj_surf = injector_rate > 0; s.tgt_inj_surf = injector_rate;
    s.has_tgt_prod_surf = producer_rate > 0; s.tgt_prod_surf = producer_rate;
    s.dev_ready = false;
    return RSF_OK;
  });
}

int rsf_set_well_bhp_limits(rsf_sim* h, double bhp_min, double bhp_max) {
  return guard(h, __func__, [&](Sim& s) {
    if (!std::isfinite(bhp_min) || !std::isfinite(bhp_max) || bhp_min > bhp_max) return RSF_ERR_INVALID;
    s.bhp_min = bhp_min; s.bhp_max = bhp_max;
    s.dev_ready = false;
    return RSF_OK;
  });
}

int rsf_add_well_with_id(rsf_sim* h, int i, int j, int k, double bhp, double rw, double skin, int injector,
                         const char* id) {
  return guard(h, __func__, [&](Sim& s) {
    // rs/frontend.rs:124-166
    if (i < 0 || j < 0 || k < 0 || i >= s.nx || j >= s.ny || k >= s.nz)
      return invalid(s, "Well indices out of bounds: (i=" + std::to_string(i) + ", j=" + std::to_string(j) +
                            ", k=" + std::to_string(k) + ") for grid (" + std::to_string(s.nx) + ", " +
                            std::to_string(s.ny) + ", " + std::to_string(s.nz) + ")");
    if (!std::isfinite(bhp)) return invalid(s, "BHP must be finite, got: " + num(bhp));
    if (!(rw > 0) || !std::isfinite(rw)) return invalid(s, "Well radius must be positive and finite, got: " + num(rw));
    if (!std::isfinite(skin)) return invalid(s, "Skin factor must be finite, got: " + num(skin));
    for (auto& w : s.wells)
      if (w.i == i && w.j == j && w.k == k)
        return invalid(s, "Multiple well completions in the same cell are not supported: (" + std::to_string(i) + ", " +
                              std::to_string(j) + ", " + std::to_string(k) + ")");
    rsf::Completion c{};
    c.i = i; c.j = j; c.k = k; c.bhp = bhp; c.rw = rw; c.skin = skin; c.injector = injector != 0;
    c.id = id ? id : "";
    s.wells.push_back(c);
    s.build_topology();
    return RSF_OK;
  });
}

int rsf_add_well(rsf_sim* h, int i, int j, int k, double bhp, double rw, double skin, int injector) {
  return rsf_add_well_with_id(h, i, j, k, bhp, rw, skin, injector, nullptr);
}

int rsf_set_well_schedule(rsf_sim* h, const char* id, const char* mode, double target_rate, double target_surface_rate,
                          double bhp_limit, int enabled) {
  return guard(h, __func__, [&](Sim& s) {
    std::string m = lower(mode);
    if (!(m == "rate" || m == "pressure" || m.empty()) || !id || !*id) return RSF_ERR_INVALID;
    bool any = false;
    for (auto& w : s.wells)
      if (w.id == id) {
        w.schedule.has = true;
        w.schedule.rate = m == "rate";
        w.schedule.has_target = std::isfinite(target_rate) && target_rate >= 0;
        w.schedule.target = target_rate;
        w.schedule.has_surface = std::isfinite(target_surface_rate) && target_surface_rate >= 0;
        w.schedule.surface = target_surface_rate;
        w.schedule.has_limit = std::isfinite(bhp_limit);
        w.schedule.limit = bhp_limit;
        w.schedule.enabled = enabled != 0;
        any = true;
      }
    s.dev_ready = false;
    return any ? RSF_OK : RSF_ERR_INVALID;
  });
}

int rsf_set_well_datum(rsf_sim* h, const char* id, double datum_depth_m, double wellbore_density) {
  return guard(h, __func__, [&](Sim& s) {
    bool any = false;
    for (auto& w : s.wells)
      if (id && w.id == id) {
        w.has_datum = std::isfinite(datum_depth_m);
        w.datum = datum_depth_m;
        w.has_density = std::isfinite(wellbore_density) && wellbore_density >= 0;
        w.density = wellbore_density;
        any = true;
      }
    return any ? RSF_OK : RSF_ERR_INVALID;
  });
}

int rsf_set_well_injected_fluid(rsf_sim* h, const char* id, const char* fluid) {
  return guard(h, __func__, [&](Sim& s) {
    std::string f = lower(fluid);
    int v = f == "water" ? 0 : (f == "gas" ? 1 : -2);
    if (v < -1) return RSF_ERR_INVALID;
    bool any = false;
    for (auto& w : s.wells)
      if (id && w.id == id) { w.fluid = v; any = true; }
    s.dev_ready = false;
    return any ? RSF_OK : RSF_ERR_INVALID;
  });
}

static bool options_valid(const rsf_solver_options* o) {
  return o && o->max_newton_iterations > 0 && o->linear_restart >= 2 && o->linear_restart <= 60 &&
         o->precision >= 0 && o->precision <= 1 && o->ilu_ordering >= 0 && o->ilu_ordering <= 1;
}
static rsf::Options options_from(const rsf_solver_options* o) {
  rsf::Options d;
  d.max_newton = o->max_newton_iterations;
  d.min_newton = o->min_newton_iteration_index;
  d.residual_tol = o->newton_residual_tol;
  d.mb_tol = o->newton_mb_tol;
  d.lin_rel_tol = o->linear_rel_tol;
  d.lin_abs_tol = o->linear_abs_tol;
  d.lin_max_it = o->linear_max_iterations;
  d.restart = o->linear_restart;
  d.precision = o->precision;
  d.ordering = o->ilu_ordering;
  d.pressure_max_it = o->pressure_max_iterations;
  d.pressure_rel_tol = o->pressure_rel_tol;
  return d;
}

int rsf_set_solver_options(rsf_sim* h, const rsf_solver_options* o) {
  return guard(h, __func__, [&](Sim& s) {
    if (!options_valid(o)) return RSF_ERR_INVALID;
    s.opt = options_from(o);
    return RSF_OK;
  });
}

int rsf_default_solver_options(rsf_solver_options* o) {
  if (!o) return RSF_ERR_INVALID;
  rsf::Options d;
  o->max_newton_iterations = d.max_newton;
  o->min_newton_iteration_index = d.min_newton;
  o->newton_residual_tol = d.residual_tol;
  o->newton_mb_tol = d.mb_tol;
  o->linear_rel_tol = d.lin_rel_tol;
  o->linear_abs_tol = d.lin_abs_tol;
  o->linear_max_iterations = d.lin_max_it;
  o->linear_restart = d.restart;
  o->precision = d.precision;
  o->ilu_ordering = d.ordering;
  o->pressure_max_iterations = d.pressure_max_it;
  o->pressure_rel_tol = d.pressure_rel_tol;
  return RSF_OK;
}

int rsf_step(rsf_sim* h, double dt_days) {
  return guard(h, __func__, [&](Sim& s) {
    if (!(dt_days > 0) || !std::isfinite(dt_days)) return RSF_ERR_INVALID;
    int rc = s.step(dt_days);
    return rc == 0 ? RSF_OK : RSF_ERR_NOT_CONVERGED;
  });
}

int rsf_run_timestep(rsf_sim* h, double dt_days, rsf_step_report* out) {
  return guard(h, __func__, [&](Sim& s) {
    if (!(dt_days > 0) || !std::isfinite(dt_days)) return RSF_ERR_INVALID;
    s.ensure_device();
    s.refresh_head_offsets();
    Sim::SubstepResult r{};
    int rc = s.run_substep(dt_days, r);
    if (out) {
      out->converged = r.converged ? 1 : 0;
      out->newton_iterations = r.newton;
      out->linear_iterations = r.linear;
      out->final_scaled_residual = r.scaled;
    }
    s.sync();
    return rc == 0 ? RSF_OK : RSF_ERR_NOT_CONVERGED;
  });
}

int rsf_run_timestep_from(rsf_sim* h, double dt_days, const double* init_cell_state, const int* init_regimes,
                          const double* init_bhp, const double* init_q, rsf_step_report* out) {
  return guard(h, __func__, [&](Sim& s) {
    if (!(dt_days > 0) || !std::isfinite(dt_days)) return RSF_ERR_INVALID;
    if (s.adj.on) {
      return invalid(s, "an initial iterate cannot be given while the adjoint records substeps");
    }
    s.ensure_device();
    s.refresh_head_offsets();
    Sim::SubstepResult r{};
    int rc = s.run_substep(dt_days, r, init_cell_state, init_regimes, init_bhp, init_q);
    if (out) {
      out->converged = r.converged ? 1 : 0;
      out->newton_iterations = r.newton;
      out->linear_iterations = r.linear;
      out->final_scaled_residual = r.scaled;
    }
    s.sync();
    return rc == 0 ? RSF_OK : RSF_ERR_NOT_CONVERGED;
  });
}

int rsf_run_fim_timestep(rsf_sim* h, double dt_days, const double* prev_cell_state, const int* prev_regimes,
                         const double* init_cell_state, const int* init_regimes, const double* init_bhp,
                         const double* init_q, const rsf_solver_options* options, rsf_fim_step_report* out,
                         double* out_cell_state, int* out_regimes, double* out_bhp, double* out_q) {
  return guard(h, __func__, [&](Sim& s) {
    if (!(dt_days > 0) || !std::isfinite(dt_days)) return RSF_ERR_INVALID;
    if ((prev_cell_state == nullptr) != (prev_regimes == nullptr))
      return invalid(s, "a previous state needs both its cell primaries and its regimes");
    if (options && !options_valid(options)) return RSF_ERR_INVALID;
    if (s.adj.on) return invalid(s, "run_fim_timestep cannot run while the adjoint records substeps");
    s.ensure_device();
    s.refresh_head_offsets();
    const rsf::Options saved = s.opt;
    if (options) s.opt = options_from(options);
    const int n = s.n;
    std::vector<double> bhp, q;
    s.init_substep(bhp, q);  // the stored state's FimState, BHPs and connection rates
    auto upload_cells = [&](const double* cells, const int* reg, rsf::DevBuf<double>& p, rsf::DevBuf<double>& sw,
                            rsf::DevBuf<double>& hc, rsf::DevBuf<int>& rg) {
      std::vector<double> a(n), b(n), c(n);
      for (int i = 0; i < n; ++i) {
        a[i] = cells[3 * i];
        b[i] = cells[3 * i + 1];
        c[i] = cells[3 * i + 2];
      }
      p.upload(a, s.stream);
      sw.upload(b, s.stream);
      hc.upload(c, s.stream);
      rg.upload(reg, n, s.stream);
    };
    if (prev_cell_state) {
      // fim/newton.rs:970-976: the accumulation terms come from previous_state; the initial
      // iterate defaults to it
      upload_cells(prev_cell_state, prev_regimes, s.x0_p, s.x0_sw, s.x0_hc, s.x0_reg);
      upload_cells(prev_cell_state, prev_regimes, s.x_p, s.x_sw, s.x_hc, s.x_reg);
      hipLaunchKernelGGL(rsf::k_inventory, dim3(rsf::nblk(n)), dim3(rsf::kBlock), 0, s.stream, s.params(true), n,
                         s.d_pv.p, s.x0_p.p, s.x0_sw.p, s.x0_hc.p, s.x0_reg.p, s.acc0.p);
    }
    if (init_cell_state) {
      std::vector<double> a(n), b(n), c(n);
      for (int i = 0; i < n; ++i) {
        a[i] = init_cell_state[3 * i];
        b[i] = init_cell_state[3 * i + 1];
        c[i] = init_cell_state[3 * i + 2];
      }
      s.x_p.upload(a, s.stream);
      s.x_sw.upload(b, s.stream);
      s.x_hc.upload(c, s.stream);
    }
    if (init_regimes) s.x_reg.upload(init_regimes, n, s.stream);
    if (init_bhp) std::copy(init_bhp, init_bhp + bhp.size(), bhp.begin());
    if (init_q) std::copy(init_q, init_q + q.size(), q.begin());
    const auto t0 = std::chrono::steady_clock::now();
    Sim::SubstepResult r{};
    try {
      r = s.run_timestep(dt_days, bhp, q);
    } catch (...) {
      s.opt = saved;
      throw;
    }
    s.opt = saved;
    s.stats.newton += r.newton;
    const double total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (out_cell_state || out_regimes) {
      std::vector<double> a(n), b(n), c(n);
      std::vector<int> rg(n);
      s.x_p.download(a.data(), n, s.stream);
      s.x_sw.download(b.data(), n, s.stream);
      s.x_hc.download(c.data(), n, s.stream);
      s.x_reg.download(rg.data(), n, s.stream);
      s.sync();
      if (out_cell_state)
        for (int i = 0; i < n; ++i) {
          out_cell_state[3 * i] = a[i];
          out_cell_state[3 * i + 1] = b[i];
          out_cell_state[3 * i + 2] = c[i];
        }
      if (out_regimes) std::copy(rg.begin(), rg.end(), out_regimes);
    }
    if (out_bhp) std::copy(bhp.begin(), bhp.end(), out_bhp);
    if (out_q) std::copy(q.begin(), q.end(), out_q);
    if (out) {
      out->converged = r.converged ? 1 : 0;
      out->newton_iterations = r.newton;
      out->final_residual_inf_norm = r.scaled;
      out->final_material_balance_inf_norm = r.mb_norm;
      out->final_update_inf_norm = NAN;
      out->has_last_linear_report = r.lin_solves > 0 ? 1 : 0;
      out->last_linear_iterations = r.last_lin_its;
      out->last_linear_converged = r.last_lin_conv;
      out->linear_solve_count = r.lin_solves;
      out->linear_iteration_count = r.linear;
      out->applied_update_count = r.updates;
      // fim/newton.rs:946-952 retry_factor_for_failure
      out->retry_factor = r.converged ? 0.5 : (r.fail_class == RSF_RETRY_LINEAR_BAD ? 0.25 : 0.5);
      out->failure_class = r.converged ? -1 : r.fail_class;
      out->failure_dominant_family = r.converged ? -1 : r.fail_family;
      out->failure_dominant_row = r.converged ? -1 : r.fail_row;
      out->total_time_ms = total;
      out->assembly_ms = r.asm_ms;
      out->linear_solve_time_ms = r.lin_ms;
    }
    s.sync();
    return r.converged ? RSF_OK : RSF_ERR_NOT_CONVERGED;
  });
}

static int get_field(rsf_sim* h, int which, double* out, int n) {
  return guard(h, __func__, [&](Sim& s) {
    if (!out || n != s.n) return RSF_ERR_INVALID;
    s.download_state();
    const std::vector<double>* f[] = {&s.h_p, &s.h_sw, &s.h_so, &s.h_sg, &s.h_rs};
    std::copy(f[which]->begin(), f[which]->end(), out);
    return RSF_OK;
  });
}
int rsf_get_pressures(rsf_sim* h, double* out, int n) { return get_field(h, 0, out, n); }
int rsf_get_sat_water(rsf_sim* h, double* out, int n) { return get_field(h, 1, out, n); }
int rsf_get_sat_oil(rsf_sim* h, double* out, int n) { return get_field(h, 2, out, n); }
int rsf_get_sat_gas(rsf_sim* h, double* out, int n) { return get_field(h, 3, out, n); }
int rsf_get_rs(rsf_sim* h, double* out, int n) { return get_field(h, 4, out, n); }

int rsf_num_wells(rsf_sim* h) { return (h && h->s) ? (int)h->s->topo.size() : -1; }
int rsf_num_perforations(rsf_sim* h) { return (h && h->s) ? (int)h->s->wells.size() : -1; }

int rsf_get_well_state(rsf_sim* h, double* bhp, int n_wells, double* q, int n_perf) {
  return guard(h, __func__, [&](Sim& s) {
    if (n_wells != (int)s.topo.size() || n_perf != (int)s.wells.size()) return RSF_ERR_INVALID;
    s.ensure_device();
    s.refresh_head_offsets();
    std::vector<double> b, q_;
    s.init_substep(b, q_);
    s.sync();
    for (int w = 0; w < n_wells; ++w) bhp[w] = b[w];
    for (int e = 0; e < n_perf; ++e) q[e] = q_[e];
    return RSF_OK;
  });
}

double rsf_get_time(rsf_sim* h) { return (h && h->s) ? h->s->time : NAN; }

int rsf_get_stats(rsf_sim* h, rsf_stats* out) {
  return guard(h, __func__, [&](Sim& s) {
    if (!out) return RSF_ERR_INVALID;
    out->newton_iterations = s.stats.newton;
    out->linear_iterations = s.stats.linear;
    out->substeps = s.stats.substeps;
    out->retries = s.stats.retries;
    out->pressure_iterations = s.stats.bicgstab;
    out->pressure_solves = s.stats.pressure_solves;
    out->assembly_ms = s.stats.assembly_ms;
    out->linear_ms = s.stats.linear_ms;
    out->total_ms = s.stats.total_ms;
    return RSF_OK;
  });
}

int rsf_step_stats_count(rsf_sim* h) {
  if (!h || !h->s) return RSF_ERR_INVALID;
  return (int)h->s->step_stats.size();
}

int rsf_get_step_stats(rsf_sim* h, int index, rsf_step_stats* out) {
  return guard(h, __func__, [&](Sim& s) {
    if (!out || index < 0 || index >= (int)s.step_stats.size()) return RSF_ERR_INVALID;
    *out = s.step_stats[index].s;
    return RSF_OK;
  });
}

int rsf_get_step_rungs(rsf_sim* h, int index, rsf_step_rung* out, int cap) {
  return guard(h, __func__, [&](Sim& s) {
    if (index < 0 || index >= (int)s.step_stats.size() || cap < 0 || (cap > 0 && !out)) return (int)RSF_ERR_INVALID;
    const auto& r = s.step_stats[index].rungs;
    std::copy(r.begin(), r.begin() + std::min<size_t>(r.size(), (size_t)cap), out);
    return (int)r.size();
  });
}

int rsf_debug_classify_retry(int have_linear, int converged, int direct, int coarse_applications,
                             double average_reduction, double last_reduction) {
  rsf::LinearOutcome l;
  l.have = have_linear != 0;
  l.converged = converged != 0;
  l.direct = direct != 0;
  l.apps = coarse_applications;
  l.avg = average_reduction;
  l.last = last_reduction;
  return rsf::retry_class(l);
}

int rsf_debug_dominant_family(const double* family_maxima, int with_wells) {
  if (!family_maxima) return RSF_ERR_INVALID;
  return rsf::dominant_family(family_maxima, with_wells != 0);
}

int rsf_get_accepted_dts(rsf_sim* h, double* out, int cap) {
  if (!h || !h->s) return RSF_ERR_INVALID;
  int n = (int)h->s->accepted_dts.size();
  for (int i = 0; i < n && i < cap && out; ++i) out[i] = h->s->accepted_dts[i];
  return n;
}

int rsf_last_warning(rsf_sim* h, char* buf, int len) {
  if (!h || !h->s || !buf || len <= 0) return RSF_ERR_INVALID;
  snprintf(buf, len, "%s", h->s->warning.c_str());
  return RSF_OK;
}

int rsf_debug_assemble(rsf_sim* h, double dt_days, const double* cell_state, const int* regimes, const double* bhp,
                       const double* q, int eliminate_wells, double* residual, double* blocks) {
  return guard(h, __func__, [&](Sim& s) {
    s.ensure_device();
    s.refresh_head_offsets();
    std::vector<double> b, q_;
    s.init_substep(b, q_);
    int n = s.n;
    if (cell_state) {
      std::vector<double> p(n), sw(n), hc(n);
      for (int i = 0; i < n; ++i) { p[i] = cell_state[3 * i]; sw[i] = cell_state[3 * i + 1]; hc[i] = cell_state[3 * i + 2]; }
      s.x_p.upload(p, s.stream);
      s.x_sw.upload(sw, s.stream);
      s.x_hc.upload(hc, s.stream);
    }
    if (regimes) s.x_reg.upload(regimes, n, s.stream);
    if (bhp) for (size_t w = 0; w < b.size(); ++w) b[w] = bhp[w];
    if (q) for (size_t e = 0; e < q_.size(); ++e) q_[e] = q[e];
    s.upload_wells(b, q_);
    s.assemble(dt_days, eliminate_wells != 0);
    int W = (int)s.topo.size(), P = (int)s.wells.size();
    std::vector<double> R(3 * (size_t)n), tail(W + P);
    s.R.download(R.data(), R.size(), s.stream);
    if (W) s.w_tailR.download(tail.data(), W + P, s.stream);
    if (blocks) s.A.download(blocks, 63 * (size_t)n, s.stream);
    s.sync();
    if (residual) {
      for (int i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) residual[3 * i + c] = R[(size_t)c * n + i];
      for (int w = 0; w < W; ++w) residual[3 * n + w] = tail[w];
      for (int t = 0; t < P; ++t) residual[3 * n + W + s.dperf_order[t]] = tail[W + t];
    }
    return RSF_OK;
  });
}

int rsf_debug_newton_direction(rsf_sim* h, double dt_days, const double* cell_state, const int* regimes,
                               const double* bhp, const double* q, double* dx, int* iterations) {
  return guard(h, __func__, [&](Sim& s) {
    int rc = rsf_debug_assemble(h, dt_days, cell_state, regimes, bhp, q, 1, nullptr, nullptr);
    if (rc) return rc;
    int n = s.n, W = (int)s.topo.size(), P = (int)s.wells.size();
    std::vector<double> tailR(W + P);
    if (W) s.w_tailR.download(tailR.data(), W + P, s.stream);
    s.sync();
    hipLaunchKernelGGL((rsf::k_fill<double>), dim3(rsf::nblk(3L * n)), dim3(rsf::kBlock), 0, s.stream, 3 * n, s.b.p, 0.0);
    hipLaunchKernelGGL((rsf::k_axpy<double, double>), dim3(rsf::nblk(3L * n)), dim3(rsf::kBlock), 0, s.stream, 3 * n,
                       -1.0, s.R.p, s.b.p);
    double t2 = 0;
    for (double v : tailR) t2 += v * v;
    double full = std::sqrt(s.dev_dot(s.b.p, s.b.p, 3 * n) + t2);
    if (W) hipLaunchKernelGGL(rsf::k_wells_rhs, dim3(W), dim3(64), 0, s.stream, n, s.well_dev(), s.R.p, s.b.p);
    int its = 0;
    double fr = 0;
    bool direct = false;
    s.solve_reduced(full, its, fr, direct);
    std::vector<double> d(3 * (size_t)n), td(W + P);
    s.dxv.download(d.data(), d.size(), s.stream);
    if (W) {
      int maxnp = 1;
      for (auto& t : s.topo) maxnp = std::max(maxnp, (int)t.perfs.size());
      hipLaunchKernelGGL(rsf::k_wells_recover, dim3(W), dim3(rsf::kWellBlock), 12 * (size_t)maxnp * sizeof(double),
                         s.stream, n, s.well_dev(), s.dxv.p, s.tdx.p);
      s.tdx.download(td.data(), W + P, s.stream);
    }
    s.sync();
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) dx[3 * i + c] = d[(size_t)c * n + i];
    for (int w = 0; w < W; ++w) dx[3 * n + w] = td[w];
    for (int t = 0; t < P; ++t) dx[3 * n + W + s.dperf_order[t]] = td[W + t];
    if (iterations) *iterations = its;
    return RSF_OK;
  });
}

int rsf_debug_newton_record(rsf_sim* h, double dt_days, const double* cell_state, const int* regimes,
                            const double* bhp, const double* q, double* out) {
  return guard(h, __func__, [&](Sim& s) {
    if (!out) return RSF_ERR_INVALID;
    int rc = rsf_debug_assemble(h, dt_days, cell_state, regimes, bhp, q, 1, nullptr, nullptr);
    if (rc) return rc;
    double r[rsf::kRecCount];
    hipLaunchKernelGGL((rsf::k_fill<double>), dim3(1), dim3(64), 0, s.stream, (int)rsf::kRecCount, s.rec.p, 0.0);
    s.newton_record(dt_days, r);
    for (int k = 0; k < 3; ++k) {
      out[k] = r[rsf::kRecFam + k];
      out[5 + k] = r[rsf::kRecCnv + k];
      out[8 + k] = r[rsf::kRecMb + k];
    }
    out[3] = r[rsf::kRecTail];
    out[4] = r[rsf::kRecTail + 1];
    out[11] = r[rsf::kRecViolFrac];
    out[12] = r[rsf::kRecViolFrac] < 0.03 ? 1.0 : 0.0;
    out[13] = s.opm_verdict(r, false) ? 1.0 : 0.0;
    out[14] = s.opm_verdict(r, true) ? 1.0 : 0.0;
    out[15] = r[rsf::kRecWellsOk];
    out[16] = r[rsf::kRecCells2] + r[rsf::kRecTail2];
    const long W = (long)s.topo.size(), P = (long)s.dperf_order.size();
    for (int k = 0; k < 5; ++k) {
      const bool none = (k == 3 && !W) || (k == 4 && !P);
      out[17 + k] = none ? -1.0 : (double)rsf::family_row(k, (long)r[rsf::kRecItem + k], s.n, W);
    }
    return RSF_OK;
  });
}

int rsf_debug_solve_wells_locally(rsf_sim* h, const double* cell_state, const int* regimes, double* bhp,
                                  double* q, int max_iterations, int* converged, int* iterations, double* peak) {
  return guard(h, __func__, [&](Sim& s) {
    if (!bhp || !q) return RSF_ERR_INVALID;
    int rc = rsf_debug_assemble(h, 1.0, cell_state, regimes, bhp, q, 0, nullptr, nullptr);
    if (rc) return rc;
    const int W = (int)s.topo.size(), P = (int)s.wells.size();
    std::vector<double> b(bhp, bhp + W), qq(q, q + P);
    s.solve_wells_locally(b, qq, max_iterations < 0 ? 50 : max_iterations);
    for (int w = 0; w < W; ++w) {
      bhp[w] = b[w];
      if (converged) converged[w] = s.inner_i[2 * w];
      if (iterations) iterations[w] = s.inner_i[2 * w + 1];
      if (peak) peak[w] = s.inner_peak[w];
    }
    for (int e = 0; e < P; ++e) q[e] = qq[e];
    return RSF_OK;
  });
}

int rsf_debug_last_iterate(rsf_sim* h, double* cell_state, int* regimes, double* bhp, int n_wells, double* q,
                           int n_perf) {
  return guard(h, __func__, [&](Sim& s) {
    if (!s.dev_ready || s.accepted_dts.empty() || !cell_state || !regimes) return RSF_ERR_INVALID;
    if (n_wells != (int)s.last_bhp.size() || n_perf != (int)s.last_q.size()) return RSF_ERR_INVALID;
    int n = s.n;
    std::vector<double> p(n), sw(n), hc(n);
    std::vector<int> rg(n);
    s.x_p.download(p.data(), n, s.stream);
    s.x_sw.download(sw.data(), n, s.stream);
    s.x_hc.download(hc.data(), n, s.stream);
    s.x_reg.download(rg.data(), n, s.stream);
    s.sync();
    for (int i = 0; i < n; ++i) {
      cell_state[3 * i] = p[i]; cell_state[3 * i + 1] = sw[i]; cell_state[3 * i + 2] = hc[i];
      regimes[i] = rg[i];
    }
    for (int w = 0; w < n_wells; ++w) bhp[w] = s.last_bhp[w];
    for (int e = 0; e < n_perf; ++e) q[e] = s.last_q[e];
    return RSF_OK;
  });
}

int rsf_debug_initial_state(rsf_sim* h, double* cell_state, int* regimes) {
  return guard(h, __func__, [&](Sim& s) {
    s.ensure_device();
    s.refresh_head_offsets();
    std::vector<double> b, q_;
    s.init_substep(b, q_);
    int n = s.n;
    std::vector<double> p(n), sw(n), hc(n);
    std::vector<int> rg(n);
    s.x_p.download(p.data(), n, s.stream);
    s.x_sw.download(sw.data(), n, s.stream);
    s.x_hc.download(hc.data(), n, s.stream);
    s.x_reg.download(rg.data(), n, s.stream);
    s.sync();
    for (int i = 0; i < n; ++i) {
      cell_state[3 * i] = p[i]; cell_state[3 * i + 1] = sw[i]; cell_state[3 * i + 2] = hc[i];
      regimes[i] = rg[i];
    }
    return RSF_OK;
  });
}

int rsf_profile_kernels(rsf_sim* h, int reps, rsf_kernel_timing* out) {
  return guard(h, __func__, [&](Sim& s) {
    if (!out || reps <= 0 || !s.dev_ready) return RSF_ERR_INVALID;
    using namespace rsf;
    int n = s.n;
    long F = (long)(s.nx - 1) * s.ny * s.nz + (long)s.nx * (s.ny - 1) * s.nz + (long)s.nx * s.ny * (s.nz - 1);
    double nnzb = n + 2.0 * F;
    bool f32 = s.opt.precision == 1;
    double P = f32 ? 4.0 : 8.0;
    hipEvent_t e0, e1;
    RSF_CHECK(hipEventCreate(&e0));
    RSF_CHECK(hipEventCreate(&e1));
    auto timeit = [&](auto&& fn) {
      fn();  // warm
      RSF_CHECK(hipEventRecord(e0, s.stream));
      for (int r = 0; r < reps; ++r) fn();
      RSF_CHECK(hipEventRecord(e1, s.stream));
      RSF_CHECK(hipEventSynchronize(e1));
      float ms = 0;
      RSF_CHECK(hipEventElapsedTime(&ms, e0, e1));
      return double(ms) / reps;
    };
    // make sure the preconditioner of the current system exists
    if (f32) s.cpr_build<float>(); else s.cpr_build<double>();
    out->spmv_ms = timeit([&] {
      hipLaunchKernelGGL((k_spmv<double, double, double>), dim3(nblk(n)), dim3(kBlock), 0, s.stream, s.grid, s.A.p,
                         s.dxv.p, s.tmp.p);
    });
    out->spmv_bytes = nnzb * 72.0 + 2.0 * 3.0 * n * 8.0;
    int L = (int)s.lvl_ptr[s.opt.ordering].size() - 1;
    out->bilu_launches = s.opt.ordering == 1 ? 2 : 2 * L;
    if (f32) {
      out->bilu_ms = timeit([&] { s.bilu_apply<float>(s.corr_f.p, s.Zf.empty() ? s.corr_f.p : s.Zf[0]->p, s.A_f.p); });
      out->pilu_ms = timeit([&] { s.pilu_apply<float>(s.bc_f[0].p, s.bc_f[1].p); });
      out->pspmv_ms = timeit([&] { pmatvec<float>(s, s.bc_f[0].p, s.bc_f[1].p); });
    } else {
      out->bilu_ms = timeit([&] {
        s.bilu_apply<double>(s.corr_d.p, s.Zd.empty() ? s.tmp.p : s.Zd[0]->p, PBuf<double>::Amat(s));
      });
      out->pilu_ms = timeit([&] { s.pilu_apply<double>(s.bc_d[0].p, s.bc_d[1].p); });
      out->pspmv_ms = timeit([&] { pmatvec<double>(s, s.bc_d[0].p, s.bc_d[1].p); });
    }
    // forward reads the lower blocks, backward the upper ones. Level-scheduled: D^-1 in
    // both sweeps, the rhs read, out written, then read and written again. Red-black: each
    // cell's D^-1 once (a neighbour's comes from cache), the rhs read and out written once.
    out->bilu_bytes = s.opt.ordering == 1 ? 2.0 * F * 9.0 * P + n * 9.0 * P + 2.0 * 3.0 * n * P
                                          : 2.0 * F * 9.0 * P + 2.0 * n * 9.0 * P + 4.0 * 3.0 * n * P;
    out->pilu_bytes = 2.0 * F * P + 2.0 * n * P + 4.0 * n * P;
    out->pspmv_bytes = (n + 2.0 * F) * P + 2.0 * n * P;
    PhysParams Pp = s.params(true);
    double dt = s.accepted_dts.empty() ? 1.0 : s.accepted_dts.back();
    out->assembly_ms = timeit([&] {
      hipLaunchKernelGGL(k_cells, dim3(nblk(n)), dim3(kBlock), 0, s.stream, Pp, n, s.d_pv.p, s.d_pcs.p, s.x_p.p,
                         s.x_sw.p, s.x_hc.p, s.x_reg.p, s.acc0.p, s.R.p, s.A.p, s.pack.p);
      hipLaunchKernelGGL(k_face_rows, dim3(nblk(n)), dim3(kBlock), 0, s.stream, s.grid, (int)s.gravity, dt, s.d_Tx.p,
                         s.d_Ty.p, s.d_Tz.p, s.d_depth.p, s.x_p.p, s.pack.p, s.R.p, s.A.p);
    });
    // state (current and previous) in, transmissibility, regimes, residual and blocks out
    out->assembly_bytes = 2.0 * 3 * n * 8 + F * 8.0 + 2.0 * n * 4 + 3.0 * n * 8 + nnzb * 72.0;
    DevBuf<double> cp;
    cp.alloc(63 * (size_t)n);
    out->copy_ms = timeit([&] {
      RSF_CHECK(hipMemcpyAsync(cp.p, s.A.p, 63 * (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
    });
    out->copy_bytes = 2.0 * 63 * n * 8;
    s.sync();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return RSF_OK;
  });
}

int rsf_adjoint_record(rsf_sim* h, int checkpoint_interval) {
  return guard(h, __func__, [&](Sim& s) {
    if (checkpoint_interval < 0) return RSF_ERR_INVALID;
    s.adj = Sim::AdjointRecord{};
    s.adj.on = checkpoint_interval > 0;
    s.adj.interval = std::max(checkpoint_interval, 1);
    return RSF_OK;
  });
}

int rsf_adjoint_objective(rsf_sim* h, double* value, int* n_substeps) {
  return guard(h, __func__, [&](Sim& s) {
    if (!value) return RSF_ERR_INVALID;
    double j = 0.0;
    for (double v : s.adj.jn) j += v;
    *value = j;
    if (n_substeps) *n_substeps = (int)s.adj.dts.size();
    return RSF_OK;
  });
}

int rsf_adjoint_gradient(rsf_sim* h, double* grad_lnk, double* grad_phi, int n, rsf_adjoint_report* report) {
  return guard(h, __func__, [&](Sim& s) {
    if (n != s.n || !grad_lnk || !grad_phi) return RSF_ERR_INVALID;
    if (s.adj.dts.empty() || !s.adj.ckpt.count(0)) return RSF_ERR_INVALID;
    s.ensure_device();
    std::vector<double> gl, gp;
    Sim::AdjointTimes tm;
    s.adjoint_gradient(gl, gp, tm);
    std::copy(gl.begin(), gl.end(), grad_lnk);
    std::copy(gp.begin(), gp.end(), grad_phi);
    if (report) {
      report->substeps = (int)s.adj.dts.size();
      report->linear_iterations = tm.linear_its;
      report->pressure_iterations = tm.pressure_its;
      report->replay_ms = tm.replay_ms;
      report->solve_ms = tm.solve_ms;
      report->total_ms = tm.total_ms;
    }
    return RSF_OK;
  });
}

int rsf_set_kernel_probe(rsf_sim* h, int enable) {
  return guard(h, __func__, [&](Sim& s) {
    if (enable < 0) return RSF_ERR_INVALID;
    s.probe_on = enable != 0;
    s.probe_every = std::max(enable, 1);
    s.probe_launches = 0;
    s.probe_ms = 0.0;
    return RSF_OK;
  });
}

int rsf_set_ilu_sweep(rsf_sim* h, int mode, int lanes_pressure, int lanes_block) {
  return guard(h, __func__, [&](Sim& s) {
    if (mode < 0 || mode > 1 || lanes_pressure < 1 || lanes_pressure > rsf::kSweepMaxLanes || lanes_block < 1 ||
        lanes_block > rsf::kSweepMaxLanes)
      return RSF_ERR_INVALID;
    s.sweep_mode = mode;
    s.sweep_lanes_p = lanes_pressure;
    s.sweep_lanes_b = lanes_block;
    return RSF_OK;
  });
}

int rsf_get_kernel_probe(rsf_sim* h, rsf_kernel_probe* out) {
  return guard(h, __func__, [&](Sim& s) {
    if (!out) return RSF_ERR_INVALID;
    long F = (long)(s.nx - 1) * s.ny * s.nz + (long)s.nx * (s.ny - 1) * s.nz + (long)s.nx * s.ny * (s.nz - 1);
    double P = s.opt.precision == 1 ? 4.0 : 8.0;
    out->launches = s.probe_launches;
    out->total_ms = s.probe_ms;
    out->avg_ms = s.probe_launches ? s.probe_ms / s.probe_launches : 0.0;
    if (s.grid.pcmp) {
      // k_rb_a, a full iteration: the black rows' off-diagonal entries that exist (one
      // endpoint of every face); D^-1 of every cell (own black, neighbouring red); x, ph,
      // sh, s, t, the previous p and v read; chat on the black half; x, r and p written,
      // ph on the black half
      out->bytes_per_launch = F * P + s.n * P + 7.0 * s.n * P + 0.5 * s.n * P + 3.5 * s.n * P;
      std::snprintf(out->kernel, sizeof(out->kernel), "%s", P == 4.0 ? "k_rb_a<float>" : "k_rb_a<double>");
    } else {
      // k_pspmv_dot: stencil entries that exist, x, y and the dot operand once each,
      // plus the clique extras (value, column) and their input values
      out->bytes_per_launch = (s.n + 2.0 * F) * P + 3.0 * s.n * P + s.n_extra * (P + 4.0 + P);
      std::snprintf(out->kernel, sizeof(out->kernel), "%s",
                    P == 4.0 ? "k_pspmv_dot<float>" : "k_pspmv_dot<double>");
    }
    return RSF_OK;
  });
}

static_assert(sizeof(rsf::TimePoint) == sizeof(rsf_time_point), "report record layout");

int rsf_rate_history_len(rsf_sim* h) {
  if (!h || !h->s) return -1;
  return (int)h->s->rate_history.size();
}

static_assert(sizeof(rsf_time_point) == sizeof(rsf::TimePoint) && sizeof(rsf_time_point) == 20 * sizeof(double),
              "the report entry is copied as 20 doubles in declaration order");
int rsf_get_rate_history(rsf_sim* h, rsf_time_point* out, int cap) {
  return guard(h, __func__, [&](Sim& s) {
    if (!out || cap < 0) return RSF_ERR_INVALID;
    const int k = std::min(cap, (int)s.rate_history.size());
    std::memcpy(out, s.rate_history.data(), (size_t)k * sizeof(rsf_time_point));
    return RSF_OK;
  });
}

int rsf_get_rate_history_since(rsf_sim* h, int start_index, rsf_time_point* out, int cap) {
  if (!h || !h->s || start_index < 0 || cap < 0 || (cap > 0 && !out)) return -1;
  const Sim& s = *h->s;
  const int len = (int)s.rate_history.size();
  const int start = std::min(start_index, len);  // frontend.rs: start_index.min(len)
  const int k = std::min(cap, len - start);
  if (k > 0) std::memcpy(out, s.rate_history.data() + start, (size_t)k * sizeof(rsf_time_point));
  return len - start;
}

int rsf_set_sweep_config(rsf_sim* h, const char* geometry, double swept_threshold, double initial_oil_saturation,
                         double residual_oil_saturation) {
  return guard(h, __func__, [&](Sim& s) {
    if (!geometry) {
      s.sweep = rsf::SweepConfig{};
      return RSF_OK;
    }
    const std::string g = geometry;
    s.sweep.on = true;
    s.sweep.geometry = g == "areal" ? 0 : (g == "vertical" ? 1 : 2);
    s.sweep.swept_threshold = swept_threshold;
    s.sweep.initial_oil_saturation = initial_oil_saturation;
    s.sweep.residual_oil_saturation = residual_oil_saturation;
    return RSF_OK;
  });
}

int rsf_load_state(rsf_sim* h, double time_days, const double* pressure, const double* sat_water,
                   const double* sat_oil, const double* sat_gas, const double* rs, int n_cells,
                   const double* completion_bhp, int n_completions, const rsf_time_point* rate_history,
                   int n_history) {
  return guard(h, __func__, [&](Sim& s) {
    if (n_cells != s.n || !pressure || !sat_water || !sat_oil) {
      return invalid(s, "Mismatch grid size. Expected " + std::to_string(s.n) + ", got " + std::to_string(n_cells));
    }
    if (n_completions != (int)s.wells.size() || (n_completions > 0 && !completion_bhp) || n_history < 0 ||
        (n_history > 0 && !rate_history))
      return RSF_ERR_INVALID;
    s.download_state();
    for (int i = 0; i < s.n; ++i) {
      s.h_p[i] = pressure[i];
      s.h_sw[i] = sat_water[i];
      s.h_so[i] = sat_oil[i];
      s.h_sg[i] = sat_gas ? sat_gas[i] : 0.0;
      s.h_rs[i] = rs ? rs[i] : 0.0;
    }
    s.time = time_days;
    for (int e = 0; e < n_completions; ++e) s.wells[e].bhp = completion_bhp[e];
    s.rate_history.assign(reinterpret_cast<const rsf::TimePoint*>(rate_history),
                          reinterpret_cast<const rsf::TimePoint*>(rate_history) + n_history);
    s.warning.clear();
    // rs/frontend.rs:998-1000: the step statistics start over
    s.step_stats.clear();
    if (!s.rate_history.empty()) {
      s.cum_injection_m3 = s.rate_history.back().total_injection_reservoir;
      s.cum_production_m3 = s.rate_history.back().total_production_liquid_reservoir;
    }
    s.dev_ready = false;
    return set_state(s);
  });
}

int rsf_set_capture_dir(rsf_sim* h, const char* dir, int mode) {
  return guard(h, __func__, [&](Sim& s) {
    if (mode < 0 || mode > 2 || (mode != 0 && (!dir || !*dir))) return RSF_ERR_INVALID;
    s.capture_mode = mode;
    s.capture_dir = mode ? dir : "";
    return RSF_OK;
  });
}

int rsf_solve_captured_system(rsf_sim* h, int rows, const int* row_ptr, const int* col, const double* val,
                              const double* rhs, double* x, rsf_linear_report* report) {
  return guard(h, __func__, [&](Sim& s) {
    if (rows <= 0 || !row_ptr || !col || !val || !rhs || !x) return RSF_ERR_INVALID;
    if (row_ptr[0] != 0) return RSF_ERR_INVALID;
    for (int r = 0; r < rows; ++r)
      if (row_ptr[r + 1] < row_ptr[r]) return RSF_ERR_INVALID;
    int its = 0;
    double res = 0.0, bn = 0.0;
    bool conv = false;
    s.solve_captured(rows, row_ptr, col, val, rhs, x, its, res, bn, conv);
    if (report) {
      report->converged = conv ? 1 : 0;
      report->iterations = its;
      report->residual_norm = res;
      report->rhs_norm = bn;
    }
    return RSF_OK;
  });
}

const char* rsf_version(void) { return "ressim_fim-hip 0.1 gfx950"; }

}  // extern "C"

// batched ensemble of two-phase members (include/ressim_ensemble.h)
#include "rsf_ensemble.hpp"
