// Linear-system capture (rs/fim/linear/capture.rs:149-640): the fim-capture-v2 text
// format, written for the systems the GPU linear stack solves, and the replay of a
// captured system through that stack. Included inside namespace rsf by rsf_lib.hip.
//
// A system is written in the reference's interleaved unknown order (cell i's water,
// oil and gas rows are 3i, 3i+1, 3i+2). What the GPU solves is the system after the
// Schur elimination of the well tail (rs/fim/linear/well_schur.rs:70-209), so its
// captures carry the layout (N, 3, 0, 3N): no well unknowns.

// f64 as Rust's `{:e}` prints it: the shortest digits that round-trip, an exponent
// without '+' or leading zeros ("1.5e-8", "4e0", "1e12", "0e0", "inf", "NaN").
inline std::string rust_exp(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
  std::string s(buf, r.ptr);
  const size_t e = s.find('e');
  if (e == std::string::npos) return s;
  std::string mant = s.substr(0, e), ex = s.substr(e + 1);
  bool neg = false;
  size_t k = 0;
  if (k < ex.size() && (ex[k] == '+' || ex[k] == '-')) neg = ex[k++] == '-';
  while (k + 1 < ex.size() && ex[k] == '0') ++k;
  return mant + "e" + (neg ? "-" : "") + ex.substr(k);
}

// stencil slot of cell j in cell i's row, -1 when j is not i or a 7-point neighbour
inline int stencil_slot(int nx, int ny, int nz, int i, int j) {
  if (i == j) return 0;
  const int nxy = nx * ny, z = i / nxy, y = (i % nxy) / nx, x = i % nx, d = j - i;
  if (d == -1 && x > 0) return 1;
  if (d == 1 && x + 1 < nx) return 2;
  if (d == -nx && y > 0) return 3;
  if (d == nx && y + 1 < ny) return 4;
  if (d == -nxy && z > 0) return 5;
  if (d == nxy && z + 1 < nz) return 6;
  return -1;
}

// Writes the system on the device (A and the extras, right-hand side b) to
// <capture_dir>/fim_capture_<seq>.txt, creating the directory. Like the reference
// (capture.rs:149-170) a failed write is reported on stderr and the solve goes on.
void Sim::write_capture(int newton_iteration, const char* reason) {
  try {
    std::filesystem::create_directories(capture_dir);
    write_capture_file(newton_iteration, reason);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "FIM capture: failed to write system %ld: %s\n", capture_seq, e.what());
  }
}

void Sim::write_capture_file(int newton_iteration, const char* reason) {
  const int N3 = 3 * n;
  std::vector<double> Ah(63 * (size_t)n), bh(N3), Eh(9 * (size_t)std::max(n_extra, 1));
  A.download(Ah.data(), Ah.size(), stream);
  b.download(bh.data(), bh.size(), stream);
  if (n_extra) w_extra.download(Eh.data(), 9 * (size_t)n_extra, stream);
  sync();
  // the extras of each cell row, by column cell
  std::vector<std::vector<std::pair<int, int>>> xr(n);
  for (int k = 0; k < n_extra; ++k) xr[extras_rc[k].first].push_back({extras_rc[k].second, k});
  const int nxy = nx * ny;
  std::string dir = capture_dir;
  char name[64];
  std::snprintf(name, sizeof(name), "/fim_capture_%05ld.txt", capture_seq++);
  std::FILE* f = std::fopen((dir + name).c_str(), "w");
  if (!f) throw std::runtime_error("cannot write capture file in " + dir);
  std::string body;
  long nnz = 0;
  std::vector<std::pair<int, double>> row;
  for (int i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      row.clear();
      const int cand[7] = {i - nxy, i - nx, i - 1, i, i + 1, i + nx, i + nxy};
      for (int j : cand) {
        if (j < 0 || j >= n) continue;
        const int s = stencil_slot(nx, ny, nz, i, j);
        if (s < 0) continue;
        for (int bb = 0; bb < 3; ++bb) row.push_back({3 * j + bb, Ah[((size_t)s * 9 + a * 3 + bb) * n + i]});
      }
      for (const auto& jk : xr[i])
        for (int bb = 0; bb < 3; ++bb)
          row.push_back({3 * jk.first + bb, Eh[(size_t)(a * 3 + bb) * n_extra + jk.second]});
      std::sort(row.begin(), row.end(), [](const auto& p, const auto& q) { return p.first < q.first; });
      for (const auto& cv : row) {
        body += std::to_string(3 * i + a) + " " + std::to_string(cv.first) + " " + rust_exp(cv.second) + "\n";
        ++nnz;
      }
    }
  }
  std::string head = "fim-capture-v2\n";
  head += "newton_iteration " + std::to_string(newton_iteration) + "\n";
  head += std::string("failure_reason ") + reason + "\n";
  head += "dominant_family none\ndominant_item_index 0\n";
  head += "layout " + std::to_string(n) + " 3 0 " + std::to_string(N3) + "\n";
  head += "equation_scaling 0\n";
  head += "rows " + std::to_string(N3) + "\ncols " + std::to_string(N3) + "\nnnz " + std::to_string(nnz) + "\n";
  head += "rhs\n";
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) head += rust_exp(bh[(size_t)c * n + i]) + "\n";
  head += "triplets\n";
  bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size() &&
            std::fwrite(body.data(), 1, body.size(), f) == body.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) throw std::runtime_error("short write of a capture file in " + dir);
}

// Dense LU with partial pivoting of an m x m row-major matrix, in place; false when
// a pivot vanishes.
inline bool lu_factor(std::vector<double>& M, int m, std::vector<int>& piv) {
  piv.resize(m);
  for (int k = 0; k < m; ++k) {
    int p = k;
    for (int r = k + 1; r < m; ++r)
      if (std::fabs(M[(size_t)r * m + k]) > std::fabs(M[(size_t)p * m + k])) p = r;
    piv[k] = p;
    if (M[(size_t)p * m + k] == 0.0 || !std::isfinite(M[(size_t)p * m + k])) return false;
    if (p != k)
      for (int c = 0; c < m; ++c) std::swap(M[(size_t)k * m + c], M[(size_t)p * m + c]);
    for (int r = k + 1; r < m; ++r) {
      const double l = M[(size_t)r * m + k] / M[(size_t)k * m + k];
      M[(size_t)r * m + k] = l;
      if (l != 0.0)
        for (int c = k + 1; c < m; ++c) M[(size_t)r * m + c] -= l * M[(size_t)k * m + c];
    }
  }
  return true;
}
// The factorization swaps whole rows (multipliers included), so the solve applies every
// swap first, then the unit-lower and the upper triangle.
inline void lu_solve(const std::vector<double>& M, int m, const std::vector<int>& piv, double* x) {
  for (int k = 0; k < m; ++k) std::swap(x[k], x[piv[k]]);
  for (int k = 0; k < m; ++k)
    for (int r = k + 1; r < m; ++r) x[r] -= M[(size_t)r * m + k] * x[k];
  for (int k = m - 1; k >= 0; --k) {
    for (int c = k + 1; c < m; ++c) x[k] -= M[(size_t)k * m + c] * x[c];
    x[k] /= M[(size_t)k * m + k];
  }
}

// Solves a captured system J x = rhs (CSR, 0-based, interleaved cell unknowns first,
// then the well tail) through the GPU linear stack. Cell-cell couplings on the 7-point
// stencil of the handle's grid go into the stencil blocks, any others into the extras
// (as the wells' cliques do). The tail is eliminated on the host by its Schur complement
// (a dense LU of the tail block), whose fill lands the same way. The reduced system is
// then solved exactly as a Newton step's (CPR-FGMRES in the handle's precision and
// ordering), and the tail recovered as x_w = J_ww^-1 (b_w - J_wr x_r).
void Sim::solve_captured(int rows, const int* ptr, const int* col, const double* val, const double* rhs, double* x,
                         int& iters, double& res, double& rhs_norm, bool& conv) {
  ensure_device();
  // from here on the device holds the captured system: the model's own wells (and
  // their extras) are rebuilt at the next step, also when this call fails
  dev_ready = false;
  const int N3 = 3 * n;
  if (rows < N3) throw std::invalid_argument("captured system has fewer rows than 3 x cells");
  const int T = rows - N3;
  if (T > 4096) throw std::invalid_argument("captured well tail larger than 4096 unknowns");
  std::vector<double> Ah(63 * (size_t)n, 0.0), bh(N3), D((size_t)T * T, 0.0), bw(T);
  struct Ent {
    int r, c;
    double v;
  };
  std::vector<Ent> jrw, jwr;
  std::map<std::pair<int, int>, std::array<double, 9>> extras;
  auto extra_block = [&](int i, int j) -> std::array<double, 9>& {
    auto it = extras.find({i, j});
    if (it == extras.end()) it = extras.emplace(std::make_pair(i, j), std::array<double, 9>{}).first;
    return it->second;
  };
  for (int r = 0; r < rows; ++r) {
    for (int e = ptr[r]; e < ptr[r + 1]; ++e) {
      const int c = col[e];
      const double v = val[e];
      if (c < 0 || c >= rows) throw std::invalid_argument("captured column index out of range");
      if (r < N3 && c < N3) {
        const int i = r / 3, a = r % 3, j = c / 3, bb = c % 3;
        const int s = stencil_slot(nx, ny, nz, i, j);
        if (s >= 0) Ah[((size_t)s * 9 + a * 3 + bb) * n + i] += v;
        else extra_block(i, j)[a * 3 + bb] += v;
      } else if (r < N3) {
        jrw.push_back({r, c - N3, v});
      } else if (c < N3) {
        jwr.push_back({r - N3, c, v});
      } else {
        D[(size_t)(r - N3) * T + (c - N3)] += v;
      }
    }
  }
  double r2 = 0.0;
  for (int r = 0; r < rows; ++r) r2 += rhs[r] * rhs[r];
  rhs_norm = std::sqrt(r2);
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) bh[(size_t)c * n + i] = rhs[3 * i + c];
  for (int k = 0; k < T; ++k) bw[k] = rhs[N3 + k];
  std::vector<int> piv;
  if (T > 0) {
    if (!lu_factor(D, T, piv)) throw std::invalid_argument("singular well tail block");
    // Y = J_ww^-1 J_wr over the cell unknowns J_wr touches; yb = J_ww^-1 b_w
    std::vector<int> ycols;
    std::map<int, int> ypos;
    for (const auto& e : jwr)
      if (!ypos.count(e.c)) {
        ypos[e.c] = (int)ycols.size();
        ycols.push_back(e.c);
      }
    const int nc = (int)ycols.size();
    std::vector<double> Y((size_t)nc * T, 0.0);  // column-major: column q at Y[q*T]
    for (const auto& e : jwr) Y[(size_t)ypos[e.c] * T + e.r] += e.v;
    for (int q = 0; q < nc; ++q) lu_solve(D, T, piv, &Y[(size_t)q * T]);
    std::vector<double> yb = bw;
    lu_solve(D, T, piv, yb.data());
    // reduced system: J_rr - J_rw Y and b_r - J_rw yb, row by row of J_rw
    std::map<int, std::vector<std::pair<int, double>>> rw;
    for (const auto& e : jrw) rw[e.r].push_back({e.c, e.v});
    for (const auto& kv : rw) {
      const int r = kv.first, i = r / 3, a = r % 3;
      double db = 0.0;
      for (const auto& kc : kv.second) db += kc.second * yb[kc.first];
      bh[(size_t)a * n + i] -= db;
      for (int q = 0; q < nc; ++q) {
        double f = 0.0;
        for (const auto& kc : kv.second) f += kc.second * Y[(size_t)q * T + kc.first];
        const int c = ycols[q], j = c / 3, bb = c % 3;
        const int s = stencil_slot(nx, ny, nz, i, j);
        if (s >= 0) Ah[((size_t)s * 9 + a * 3 + bb) * n + i] -= f;
        else extra_block(i, j)[a * 3 + bb] -= f;
      }
    }
  }
  // the extras pattern must be symmetric: absent partners are zero blocks
  std::vector<std::pair<int, int>> keys;
  for (const auto& kv : extras) keys.push_back(kv.first);
  for (const auto& k : keys)
    if (!extras.count({k.second, k.first})) extras.emplace(std::make_pair(k.second, k.first), std::array<double, 9>{});
  std::vector<std::pair<int, int>> rc;
  for (const auto& kv : extras) rc.push_back(kv.first);  // std::map order: row, then column
  set_extras_pattern(rc);
  if (n_extra) {
    std::vector<double> Eh(9 * (size_t)n_extra);
    int k = 0;
    for (const auto& kv : extras) {
      for (int e = 0; e < 9; ++e) Eh[(size_t)e * n_extra + k] = kv.second[e];
      ++k;
    }
    w_extra.upload(Eh.data(), Eh.size(), stream);
  }
  A.upload(Ah.data(), Ah.size(), stream);
  b.upload(bh.data(), bh.size(), stream);
  if (opt.precision == 1) {
    cpr_build<float>();
    conv = linear_solve<float>(rhs_norm, iters, res);
  } else {
    cpr_build<double>();
    conv = linear_solve<double>(rhs_norm, iters, res);
  }
  std::vector<double> xr(N3);
  dxv.download(xr.data(), xr.size(), stream);
  sync();
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) x[3 * i + c] = xr[(size_t)c * n + i];
  if (T > 0) {
    std::vector<double> t = bw;
    for (const auto& e : jwr) t[e.r] -= e.v * x[e.c];
    lu_solve(D, T, piv, t.data());
    for (int k = 0; k < T; ++k) x[N3 + k] = t[k];
  }
}
