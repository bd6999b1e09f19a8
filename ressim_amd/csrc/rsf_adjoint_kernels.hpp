// Kernels of the discrete adjoint of the FIM step (backward pass over recorded
// substeps). The forward residual of substep n is
//   R_n = inv(x_n) - inv(x_{n-1}) + dt sum_faces F(T_f, x_n) - dt sources(x_n, q_n),
// with well rows (perforation rates, control) in the tail. The objective is the
// surface oil produced by the producers: J = sum_n dt_n sum_e coef_o(x_n) q_e,n.
// Parameters: a per-cell log-permeability multiplier applied to kx, ky and kz
// together, and per-cell porosity.
#pragma once
#include "rsf_kernels.hpp"

namespace rsf {

// Objective of one substep, per perforation: j_e = dt coef_o(x) q_e for enabled
// producers (0 otherwise), dj/dq_e, and dj/dx of the completed cell (djdx is
// cleared by the caller; completions sit in distinct cells).
__global__ void k_adj_objective(PhysParams P, int n, double dt, WellDev W, const double* __restrict__ p,
                                const double* __restrict__ sw, const double* __restrict__ hc,
                                const int* __restrict__ reg, double* __restrict__ jv, double* __restrict__ djdq,
                                double* __restrict__ djdx) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= W.n_perf) return;
  int w = 0;
  while (w + 1 < W.n_wells && W.well_off[w + 1] <= e) ++w;
  int flags = W.well_flags[w];
  bool inj = flags & 1, enabled = (flags & 16) != 0;
  jv[e] = 0.0;
  djdq[e] = 0.0;
  if (inj || !enabled) return;
  int c = W.perf_cell[e];
  D3 dp = dvar<3>(p[c], 0), ds = dvar<3>(sw[c], 1), dh = dvar<3>(hc[c], 2);
  D3 coef[3], cm;
  perforation_terms(P, false, false, dp, ds, dh, reg[c], coef, cm, rs_cap_at(P, c));
  double q = W.q[e];
  jv[e] = coef[1].v * q * dt;
  djdq[e] = coef[1].v * dt;
#pragma unroll
  for (int v = 0; v < 3; ++v) djdx[v * n + c] = coef[1].d[v] * q * dt;
}

// Cell part of the adjoint right-hand side of substep m:
//   b = -dj/dx + (d inv(x_m)/dx)^T lambda_{m+1}   (lam_next null: last substep)
__global__ __launch_bounds__(kBlock) void k_adj_rhs(PhysParams P, int n, const double* __restrict__ pv_ref,
                                                    const double* __restrict__ p, const double* __restrict__ sw,
                                                    const double* __restrict__ hc, const int* __restrict__ reg,
                                                    const double* __restrict__ lam_next,
                                                    const double* __restrict__ djdx, double* __restrict__ b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double out[3] = {-djdx[i], -djdx[n + i], -djdx[2 * n + i]};
  if (lam_next) {
    D3 dp = dvar<3>(p[i], 0), ds = dvar<3>(sw[i], 1), dh = dvar<3>(hc[i], 2);
    D3 inv3[3];
    inventory(P, pv_ref[i], reg[i], dp, ds, dh, inv3, rs_cap_at(P, i));
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double l = lam_next[r * n + i];
#pragma unroll
      for (int v = 0; v < 3; ++v) out[v] += l * inv3[r].d[v];
    }
  }
#pragma unroll
  for (int v = 0; v < 3; ++v) b[v * n + i] = out[v];
}

// Porosity part of lambda_m^T dR_m/dtheta: the inventories of both states scale
// with the pore volume, so dR_i/dphi_i = (inv(x_m) - inv(x_{m-1}))_i / phi_i.
__global__ __launch_bounds__(kBlock) void k_adj_grad_phi(PhysParams P, int n, const double* __restrict__ pv_ref,
                                                         const double* __restrict__ phi, const double* __restrict__ p,
                                                         const double* __restrict__ sw, const double* __restrict__ hc,
                                                         const int* __restrict__ reg, const double* __restrict__ acc0,
                                                         const double* __restrict__ lam, double* __restrict__ g_phi) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double inv3[3];
  inventory<double>(P, pv_ref[i], reg[i], p[i], sw[i], hc[i], inv3, rs_cap_at(P, i));
  double acc = 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r) acc += lam[r * n + i] * (inv3[r] - acc0[r * n + i]);
  g_phi[i] += acc / phi[i];
}

// Permeability part through the TPFA transmissibilities: a face flux is linear in
// its T, so lambda^T dR/dT_f = dt (lambda_lo - lambda_hi).F_f / T_f, and
//   x, y faces: T ~ 2 k_a k_b / (k_a + k_b)      -> d ln T / d ln k_a = k_b / (k_a + k_b),
//   z faces:    T ~ k_a k_b / (k_b dz_a + k_a dz_b) -> d ln T / d ln k_a = k_b dz_a / (k_b dz_a + k_a dz_b)
// (rs/grid.rs:278-311). Each cell sums its own six faces: no atomics.
__global__ __launch_bounds__(kBlock, RSF_FACE_MIN_WAVES) void k_adj_grad_lnk(
    Grid g, int gravity, double dt, const double* __restrict__ Tx, const double* __restrict__ Ty,
    const double* __restrict__ Tz, const double* __restrict__ depth_k, const double* __restrict__ p,
    const double* __restrict__ pack, const double* __restrict__ kx, const double* __restrict__ ky,
    const double* __restrict__ kz, const double* __restrict__ dz, const double* __restrict__ lam,
    double* __restrict__ g_lnk) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n = g.n;
  if (i >= n) return;
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  const PackCell me{pack, n, i};
  double acc = 0.0;
#pragma unroll 1
  for (int s = 1; s <= 6; ++s) {
    int j = slot_nbr(g, i, s, x, y, z);
    if (j < 0) continue;
    bool minus = (s & 1) != 0;
    int lo = minus ? j : i, hi = minus ? i : j;
    const double* Tarr = s <= 2 ? Tx : (s <= 4 ? Ty : Tz);
    double T = Tarr[lo];
    if (!(T > 0.0)) continue;
    const PackCell nb{pack, n, j};
    int zl = minus ? (s == 5 ? z - 1 : z) : z;
    int zu = minus ? z : (s == 6 ? z + 1 : z);
    double G = gravity ? kGravity * (depth_k[zl] - depth_k[zu]) * 1e-5 : 0.0;
    double F[3], Ga[3][3], Gb[3][3];
    if (minus) face_flux(nb, me, p[j], p[i], T, G, F, Ga, Gb);
    else face_flux(me, nb, p[i], p[j], T, G, F, Ga, Gb);
    double lf = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) lf += (lam[c * n + lo] - lam[c * n + hi]) * F[c];
    double dlnT;
    if (s <= 4) {
      const double* k = s <= 2 ? kx : ky;
      dlnT = k[j] / (k[i] + k[j]);
    } else {
      double dzi = dz[z], dzj = dz[s == 5 ? z - 1 : z + 1];
      dlnT = kz[j] * dzi / (kz[j] * dzi + kz[i] * dzj);
    }
    acc += dt * lf * dlnT;
  }
  g_lnk[i] += acc;
}

// S^T of the stencil BSR store: block (i, slot s) = (block (j, opposite slot))^T
// for the neighbour j through s; the diagonal block transposes in place.
__global__ void k_transpose_blocks(Grid g, const double* __restrict__ A, double* __restrict__ AT) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n = g.n;
  if (i >= n) return;
  int x, y, z;
  cell_ijk(g, i, x, y, z);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int v = 0; v < 3; ++v) AT[(size_t)(c * 3 + v) * n + i] = A[(size_t)(v * 3 + c) * n + i];
  for (int s = 1; s <= 6; ++s) {
    int j = slot_nbr(g, i, s, x, y, z);
    int o = opposite(s);
    for (int c = 0; c < 3; ++c)
      for (int v = 0; v < 3; ++v)
        AT[(size_t)(s * 9 + c * 3 + v) * n + i] = j < 0 ? 0.0 : A[(size_t)(o * 9 + v * 3 + c) * n + j];
  }
}

// transposed clique extras: entry k = (row r, col c) takes (entry tperm[k] = (c, r))^T
__global__ void k_transpose_extras(int n_extra, const int* __restrict__ tperm, const double* __restrict__ E,
                                   double* __restrict__ ET) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_extra) return;
  int t = tperm[k];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int v = 0; v < 3; ++v) ET[(size_t)(c * 3 + v) * n_extra + k] = E[(size_t)(v * 3 + c) * n_extra + t];
}

// ------------------------------------------------ transposed CPR (adjoint solves)
// With A holding J^T, the preconditioner is the transpose of the forward two-stage
// CPR: weights w from the forward diagonal (column 0 of inv(diag J^T)), coarse
// matrix C^T = e0^T J^T W (pressure-unknown row of J^T, columns combined with the
// neighbours' weights), restriction by the pressure row, prolongation by w, and
// the block-ILU0 stage first:  z = y + w (x) C^-T e0^T (r - J^T y),  y = ILU(J^T)^-1 r.

// forward quasi-IMPES weights from the transposed diagonal
template <class PT>
__global__ void k_cpr_weights_t(int n, const double* __restrict__ A, double* __restrict__ w, PT* __restrict__ w_pt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m[9], o[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) m[k] = A[(size_t)k * n + i];
  inv3x3(m, o);
  double r[3] = {o[0], o[3], o[6]};
  double mx = fmax(fabs(r[0]), fmax(fabs(r[1]), fabs(r[2])));
  if (mx > kEps)
    for (int l = 0; l < 3; ++l) r[l] /= mx;
  for (int l = 0; l < 3; ++l) {
    w[l * n + i] = r[l];
    w_pt[l * n + i] = PT(r[l]);
  }
}

// C^T in the stencil store: slot s of cell i = sum_l (J^T block (i, s))[0][l] w_l(neighbour)
template <class PT>
__global__ void k_cpr_coarse_t(Grid g, const double* __restrict__ A, const double* __restrict__ w,
                               PT* __restrict__ Pm) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n = g.n;
  if (i >= n) return;
  int cx, cy, cz;
  cell_ijk(g, i, cx, cy, cz);
  const int pi = pslot(g, i, cx, cy, cz);
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    int j = s == 0 ? i : slot_nbr(g, i, s, cx, cy, cz);
    double acc = 0.0;
    if (j >= 0)
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        double wl = w[l * n + j];
        if (fabs(wl) > kEps) acc += A[(size_t)(s * 9 + l) * n + i] * wl;
      }
    Pm[(size_t)s * n + pi] = PT(acc);
  }
}

template <class PT>
__global__ void k_cpr_extra_t(int n, int n_extra, const int* __restrict__ ecol, const double* __restrict__ w,
                              const double* __restrict__ E, PT* __restrict__ Pe) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_extra) return;
  int c = ecol[k];
  double acc = 0.0;
  for (int l = 0; l < 3; ++l) {
    double wl = w[l * n + c];
    if (fabs(wl) > kEps) acc += E[(size_t)l * n_extra + k] * wl;
  }
  Pe[k] = PT(acc);
}

// rp = e0^T (r - J^T y): the pressure-unknown row of the residual after the ILU stage
template <class M, class Y, class PT>
__global__ void k_cpr_restrict_t(Grid g, const M* __restrict__ A, const Y* __restrict__ y, const double* __restrict__ r,
                                 PT* __restrict__ rp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n = g.n;
  if (i >= n) return;
  int cx, cy, cz;
  cell_ijk(g, i, cx, cy, cz);
  const int ai = pslot(g, i, cx, cy, cz);  // A is the preconditioner's copy
  double acc = r[i];
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    int j = s == 0 ? i : slot_nbr(g, i, s, cx, cy, cz);
    if (j < 0) continue;
#pragma unroll
    for (int v = 0; v < 3; ++v) acc -= double(A[(size_t)(s * 9 + v) * n + ai]) * double(y[v * n + j]);
  }
  rp[i] = PT(acc);
}

// extras rows of the same: rp[row] -= sum_k (E_k)[0][:] . y[col_k]   (one wavefront per row)
template <class M, class Y, class PT>
__global__ void k_cpr_restrict_t_extra(int n, int n_rows, const int* __restrict__ rows, const int* __restrict__ ptr,
                                       const int* __restrict__ col, const M* __restrict__ E, int n_extra,
                                       const Y* __restrict__ y, PT* __restrict__ rp) {
  int t = blockIdx.x;
  if (t >= n_rows) return;
  int r = rows[t];
  double acc = 0.0;
  for (int k = ptr[t] + threadIdx.x; k < ptr[t + 1]; k += 64) {
    int j = col[k];
    for (int v = 0; v < 3; ++v) acc += double(E[(size_t)v * n_extra + k]) * double(y[v * n + j]);
  }
  acc = wave_sum(acc);
  if (threadIdx.x == 0) rp[r] = PT(double(rp[r]) - acc);
}

// z += w (x) xp   (prolongation by the weights)
template <class PT, class Z>
__global__ void k_prolong_w(int n, const PT* __restrict__ w, const PT* __restrict__ xp, Z* __restrict__ z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = double(xp[i]);
#pragma unroll
  for (int c = 0; c < 3; ++c) z[c * n + i] = Z(double(z[c * n + i]) + double(w[c * n + i]) * x);
}

// b[:, cell_e] += d_e for a list of completions (distinct cells)
__global__ void k_adj_scatter3(int m, int n, const int* __restrict__ cells, const double* __restrict__ d,
                               double* __restrict__ b) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  int c = cells[e];
#pragma unroll
  for (int v = 0; v < 3; ++v) b[v * n + c] += d[3 * e + v];
}
}  // namespace rsf
