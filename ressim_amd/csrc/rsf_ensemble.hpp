// Batched FIM ensemble (include/ressim_ensemble.h): many two-phase oil-water models of
// uneven grid sizes advanced together on one GPU.
//
// Packed index space. Member m owns cells [off_m, off_m + n_m) of every cell array; its
// cells are in its own natural order (rs/grid.rs:252-254), so a cell's stencil
// neighbours are found from the member's (nx, ny, nz) exactly as for a single model.
// Kernels run over "tiles": a workgroup takes up to threads x cells-per-thread cells of
// ONE member, so every per-member scalar a kernel needs (substep length, phase, Krylov
// coefficients) is uniform over the workgroup. The tile maps (natural cells, red cells,
// black cells) are built once, with the cells per thread of each member's size class
// picked by timing (the autotuner below).
//
// Per member the FIM residual is the reference's (rs/fim/assembly_ad.rs:782-983) for two
// components (water, oil) in two unknowns (p, sw): accumulation through `inventory`,
// TPFA faces through `face_cell`, both shared with the single-model path
// (rsf_physics.hpp). Wells: the connection rate of a completion is eliminated in closed
// form, q = WI lambda_t (p - bhp - head) with the reference's strict-sign clamp
// (rs/fim/wells.rs:1098-1127), so a BHP-controlled well adds only cell terms and a
// rate-controlled well keeps one unknown, its BHP, with the Fischer-Burmeister control
// row of rs/fim/wells_ad.rs:538-572. At convergence this is the reference's system: its
// perforation rows q - conn(p, bhp) = 0 hold exactly. The BHP unknown is eliminated by
// its scalar Schur complement: (J - B H^T / s) dx = -R + B r_w / s, applied exactly in
// the operator (a rank-one term per well) and on the stencil pattern in the
// preconditioner.
//
// Linear solver per member: FGMRES (one cycle of up to krylov_dimension iterations,
// classical Gram-Schmidt applied twice, Givens rotations on the device) right-
// preconditioned by the reference's two-stage CPR (rs/fim/linear/gmres_block_jacobi.rs:
// quasi-IMPES weights, coarse pressure matrix, pressure BiCGStab with ILU(0), block
// ILU(0) on the corrected residual) in FP32, both ILU(0)s in red-black order (the
// single-model path's mixed red-black mode): red D^-1 = A_rr^-1, black
// D = A_bb - sum A_br D_r^-1 A_rb. The pressure BiCGStab of every member runs in one
// launch, one workgroup per member, with its reductions inside the workgroup.
// Acceptance (residual_tol mode of the single-model Newton loop): the scaled family
// maxima and |well control rows| <= tol and each component's material balance <= mb_tol,
// with the reference's relaxed linear acceptance (reduction < 0.01,
// rs/fim/newton/convergence.rs:385-389) and per-cell chop (rs/fim/newton/convergence.rs:648-711).
//
// One Newton iteration of every running member ("tick") is a fixed kernel sequence over
// device-resident control state, captured once as a HIP graph and relaunched until every
// member has finished.
//
// Compiled in the library's translation unit (included at the end of rsf_lib.hip).
#pragma once
#include "rsf_sim.hpp"
#include "../../include/ressim_ensemble.h"

#include <mutex>

namespace rsf {
namespace ens {

constexpr int kMaxThreads = 512;
constexpr int kMaxKrylov = 200;
constexpr int kPackQ = 3;  // face pack: m_w, m_o, Pc_ow, each (value, d/dp, d/dsw)
using D2 = Dual<2>;

enum Phase { kStart = 0, kIter = 1, kOk = 2, kFail = 3, kFinished = 4, kDead = 5 };
enum Action { kCommit = 1, kBegin = 2 };

struct Tile {
  int m, first, count;
};

struct Member {
  int nx, ny, nz, n, off, nred, w0, nw, t0, nt;
  double bhp_min;
  PhysParams P;
};

struct Ctl {
  int phase, action, it, solve, kact, k, upd;
  int substeps, retries, streak, newton, linear, report, replay_k;
  double time, t_rem, dt, dt_next, beta, red;
  double fam[2], mb[2], wres;
};

struct RunDev {
  int mode;  // 0 report steps with substep control, 1 replay
  int n_reports, max_retries, max_newton, min_newton, p_max_it;
  double report_dt, tol, mb_tol, lin_tol, p_tol;
  const double* replay_dt;
  const int* replay_off;
};

struct Dev {
  int Nt, K, nM, nW, pstride;
  const Member* mem;
  Ctl* ctl;
  const RunDev* run;
  int* alive;
  // cells
  double *p, *sw, *p0, *sw0, *acc0;
  const double *pv, *pcs, *Tx, *Ty, *Tz;
  double *R, *J, *pack;
  // wells (packed over members; completions in each member's device order)
  const int *w_flags, *w_pf;
  const double *w_target, *w_limit, *w_bhpt;
  double *w_bhp, *w_bhp0, *w_s, *w_rw;
  const int* pf_cell;
  const double *pf_wi, *pf_head;
  double *pf_b, *pf_h, *pf_q;
  // preconditioner: FP32, colour-compacted per member
  float *Mo, *Dg, *Di;
  // CPR pressure stage (FP32, natural): weights (2 planes), pressure matrix (7 planes),
  // its ILU(0) diagonal inverse, restricted residual, pressure solution, BiCGStab work
  float *Wq, *Pp, *pdi, *rp, *xp, *bw;
  double* cv;  // corrected residual of the second stage (FP64, 2 planes)
  // per cell: bit s-1 set when the slot-s neighbour exists (s = 1..6), bit 6 the colour
  // (x + y + z odd); read by the pressure BiCGStab instead of recomputing (x, y, z)
  const unsigned char* nmask;
  // Krylov
  double *V, *Wv, *b, *dx;
  float* Z;
  double *H, *g, *cs, *sn, *y, *hc;
  double* part;
};

__device__ __forceinline__ Grid mgrid(const Member& M) { return Grid{M.nx, M.ny, M.nz, M.n, 0}; }
__device__ __forceinline__ int colour_of(const Member& M, int i) {
  int x, y, z;
  cell_ijk(mgrid(M), i, x, y, z);
  return (x + y + z) & 1;
}
// colour-compacted slot of local cell i (every aligned pair 2t, 2t+1 holds one cell of
// each colour, for any grid shape, so the t-th cell of a colour is i >> 1)
__device__ __forceinline__ int cslot(const Member& M, int i) { return (i >> 1) + (colour_of(M, i) ? M.nred : 0); }
// local cell of the t-th cell of colour c
__device__ __forceinline__ int cell_of_slot(const Member& M, int c, int t) {
  return colour_of(M, 2 * t) == c ? 2 * t : 2 * t + 1;
}

__device__ __forceinline__ bool running(const Dev& d, int m) { return d.ctl[m].phase == kIter; }
__device__ __forceinline__ bool kactive(const Dev& d, int m) { return d.ctl[m].kact == 1; }
__device__ __forceinline__ bool solving(const Dev& d, int m) {
  const Ctl& c = d.ctl[m];
  return c.phase == kIter && c.solve;
}

// 2x2 inverse; a singular or non-finite block becomes the identity
__device__ __forceinline__ void inv2(float a, float b, float c, float e, float out[4]) {
  const float det = a * e - b * c;
  if (!(fabsf(det) > 0.0f) || !isfinite(det)) {
    out[0] = 1.0f; out[1] = 0.0f; out[2] = 0.0f; out[3] = 1.0f;
    return;
  }
  const float r = 1.0f / det;
  out[0] = e * r; out[1] = -b * r; out[2] = -c * r; out[3] = a * r;
}

// ------------------------------------------------------------------ control
__device__ void control_member(Dev& d, int m, int* alive) {
  const RunDev& r = *d.run;
  const Member& M = d.mem[m];
  Ctl c = d.ctl[m];
  c.action = 0;
  if (c.phase == kOk) {
    c.substeps++;
    c.time += c.dt;
    c.t_rem -= c.dt;
    c.streak = 0;
    c.action |= kCommit;
    for (int w = M.w0; w < M.w0 + M.nw; ++w) d.w_bhp0[w] = d.w_bhp[w];
    if (r.mode == 1) {
      c.replay_k++;
      const int cnt = r.replay_off[m + 1] - r.replay_off[m];
      if (c.replay_k >= cnt) {
        c.phase = kFinished;
      } else {
        c.dt = r.replay_dt[r.replay_off[m] + c.replay_k];
        c.phase = kStart;
      }
    } else {
      c.dt_next = fmin(2.0 * c.dt, r.report_dt);
      if (c.t_rem <= 1e-9 * r.report_dt) {
        c.report++;
        c.t_rem = r.report_dt;
        c.phase = c.report >= r.n_reports ? kFinished : kStart;
      } else {
        c.phase = kStart;
      }
      c.dt = fmin(c.dt_next, c.t_rem);
    }
  } else if (c.phase == kFail) {
    c.retries++;
    c.streak++;
    for (int w = M.w0; w < M.w0 + M.nw; ++w) d.w_bhp[w] = d.w_bhp0[w];
    if (r.mode == 1 || c.streak > r.max_retries) {
      c.phase = kDead;
    } else {
      c.dt *= 0.5;
      c.phase = kStart;
    }
  }
  if (c.phase == kStart) {
    c.action |= kBegin;
    c.phase = kIter;
    c.it = 0;
    for (int w = M.w0; w < M.w0 + M.nw; ++w) d.w_bhp[w] = d.w_bhp0[w];
  }
  c.solve = 0;
  c.kact = 0;
  c.upd = 0;
  d.ctl[m] = c;
  if (c.phase == kIter) atomicAdd(alive, 1);
}

// One workgroup, one thread per member: the substep bookkeeping of the previous tick's
// outcome, and the count of members still running.
__global__ void k_control(Dev d) {
  __shared__ int s_alive;
  if (threadIdx.x == 0) s_alive = 0;
  __syncthreads();
  for (int m = threadIdx.x; m < d.nM; m += blockDim.x) control_member(d, m, &s_alive);
  __syncthreads();
  if (threadIdx.x == 0) *d.alive = s_alive;
}


// commit (start := iterate) and begin (iterate := start, inventory of the start)
__global__ void k_cellstate(Dev d, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  const int a = d.ctl[t.m].action;
  if (!a) return;
  const Member& M = d.mem[t.m];
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    if (a & kCommit) {
      d.p0[g] = d.p[g];
      d.sw0[g] = d.sw[g];
    }
    if (a & kBegin) {
      d.p[g] = d.p0[g];
      d.sw[g] = d.sw0[g];
      double inv3[3];
      inventory<double>(M.P, d.pv[g], 0, d.p0[g], d.sw0[g], 0.0, inv3);
      d.acc0[g] = inv3[0];
      d.acc0[d.Nt + g] = inv3[1];
    }
  }
}

// ------------------------------------------------------------------ assembly
// accumulation residual, its diagonal block and the face pack (rs/fim/properties.rs:171-255)
__global__ __launch_bounds__(kMaxThreads) void k_cells(Dev d, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  if (!running(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const int N = d.Nt;
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    const D2 dp = dvar<2>(d.p[g], 0), ds = dvar<2>(d.sw[g], 1), dh = dconst<2>(0.0);
    D2 inv3[3];
    inventory(M.P, d.pv[g], 0, dp, ds, dh, inv3);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      d.R[r * N + g] = inv3[r].v - d.acc0[r * N + g];
#pragma unroll
      for (int v = 0; v < 2; ++v) d.J[(size_t)(r * 2 + v) * N + g] = inv3[r].d[v];
    }
    const FaceCell<D2> fc = face_cell(M.P, 0, dp, ds, dh, d.pcs[g]);
    const D2 q[kPackQ] = {fc.mw, fc.mo, fc.pcw};
#pragma unroll
    for (int k = 0; k < kPackQ; ++k) {
      d.pack[(size_t)(3 * k) * N + g] = q[k].v;
      d.pack[(size_t)(3 * k + 1) * N + g] = q[k].d[0];
      d.pack[(size_t)(3 * k + 2) * N + g] = q[k].d[1];
    }
  }
}

// TPFA flux of face a -> b (a the lower index) for water and oil without gravity, with
// partials against a's and b's (p, sw) (rs/fim/flux.rs:42-118; upwind on the sign of
// each phase potential difference, >= 0 picks a)
__device__ __forceinline__ void face_flux2(const double* __restrict__ pk, int N, int a, int b, double pa, double pb,
                                           double T, double F[2], double Ga[2][2], double Gb[2][2]) {
  auto q = [&](int c, int k, int e) { return pk[(size_t)(3 * k + e) * N + c]; };
  const double dw = (pa - pb) - (q(a, 2, 0) - q(b, 2, 0));
  const double dox = pa - pb;
  const double ddw_a[2] = {1.0 - q(a, 2, 1), -q(a, 2, 2)};
  const double ddw_b[2] = {-1.0 + q(b, 2, 1), q(b, 2, 2)};
  const double ddo_a[2] = {1.0, 0.0}, ddo_b[2] = {-1.0, 0.0};
  const bool uw = dw >= 0.0, uo = dox >= 0.0;
  const int cw = uw ? a : b, co = uo ? a : b;
  const double mw = q(cw, 0, 0), mo = q(co, 1, 0);
  F[0] = mw * dw * T;
  F[1] = mo * dox * T;
#pragma unroll
  for (int v = 0; v < 2; ++v) {
    const double mwa = uw ? q(a, 0, 1 + v) : 0.0, mwb = uw ? 0.0 : q(b, 0, 1 + v);
    const double moa = uo ? q(a, 1, 1 + v) : 0.0, mob = uo ? 0.0 : q(b, 1, 1 + v);
    Ga[0][v] = (mw * ddw_a[v] + mwa * dw) * T;
    Gb[0][v] = (mw * ddw_b[v] + mwb * dw) * T;
    Ga[1][v] = (mo * ddo_a[v] + moa * dox) * T;
    Gb[1][v] = (mo * ddo_b[v] + mob * dox) * T;
  }
}

// six faces per cell into its row (each face evaluated by both cells in the same
// lower -> upper orientation, so the rows receive exactly opposite fluxes)
__global__ __launch_bounds__(kMaxThreads) void k_faces(Dev d, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  if (!running(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const Grid G = mgrid(M);
  const int N = d.Nt;
  const double dt = d.ctl[t.m].dt;
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    int x, y, z;
    cell_ijk(G, i, x, y, z);
    double r[2], dg[2][2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      r[c] = d.R[c * N + g];
#pragma unroll
      for (int v = 0; v < 2; ++v) dg[c][v] = d.J[(size_t)(c * 2 + v) * N + g];
    }
    const double pi = d.p[g];
#pragma unroll 1
    for (int s = 1; s <= 6; ++s) {
      const int j = slot_nbr(G, i, s, x, y, z);
      double off[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
      if (j >= 0) {
        const bool minus = (s & 1) != 0;
        const int lo = M.off + (minus ? j : i);
        const double T = (s <= 2 ? d.Tx : (s <= 4 ? d.Ty : d.Tz))[lo];
        if (T > 0.0) {
          const int gj = M.off + j;
          double F[2], Ga[2][2], Gb[2][2];
          if (minus) {
            face_flux2(d.pack, N, gj, g, d.p[gj], pi, T, F, Ga, Gb);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
              r[c] -= F[c] * dt;
#pragma unroll
              for (int v = 0; v < 2; ++v) {
                dg[c][v] -= Gb[c][v] * dt;
                off[c][v] = -Ga[c][v] * dt;
              }
            }
          } else {
            face_flux2(d.pack, N, g, gj, pi, d.p[gj], T, F, Ga, Gb);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
              r[c] += F[c] * dt;
#pragma unroll
              for (int v = 0; v < 2; ++v) {
                dg[c][v] += Ga[c][v] * dt;
                off[c][v] = Gb[c][v] * dt;
              }
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int v = 0; v < 2; ++v) d.J[(size_t)(s * 4 + c * 2 + v) * N + g] = off[c][v];
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      d.R[c * N + g] = r[c];
#pragma unroll
      for (int v = 0; v < 2; ++v) d.J[(size_t)(c * 2 + v) * N + g] = dg[c][v];
    }
  }
}

// One thread per member, its wells in order (deterministic, no atomics): completion
// source terms into the cell rows, and for a rate-controlled well the control row with
// B (d cell rows / d bhp), H (d control / d cells) and s (d control / d bhp). Launched with
// 256 threads per block; the bound gives the loop over wells and completions the register
// budget of one wave per SIMD (with the default bound it spilled to scratch).
__global__ __launch_bounds__(256) void k_wells(Dev d) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= d.nM || !running(d, m)) return;
  const Member& M = d.mem[m];
  const int N = d.Nt;
  const double dt = d.ctl[m].dt;
  for (int w = M.w0; w < M.w0 + M.nw; ++w) {
    const int flags = d.w_flags[w];
    const bool inj = flags & 1, fb = (flags & 2) != 0, surf = (flags & 4) != 0, enabled = (flags & 16) != 0;
    const double bhp = fb ? d.w_bhp[w] : d.w_bhpt[w];
    double actual = 0.0, drb_sum = 0.0;
    for (int e = d.w_pf[w]; e < d.w_pf[w + 1]; ++e) {
      d.pf_b[2 * e] = d.pf_b[2 * e + 1] = 0.0;
      d.pf_h[2 * e] = d.pf_h[2 * e + 1] = 0.0;
      d.pf_q[e] = 0.0;
      if (!enabled) continue;
      const int g = M.off + d.pf_cell[e];
      const D2 dp = dvar<2>(d.p[g], 0), ds = dvar<2>(d.sw[g], 1), dh = dconst<2>(0.0);
      D2 coef[3], cm;
      perforation_terms(M.P, inj, false, dp, ds, dh, 0, coef, cm, INFINITY);
      const D2 raw = (cm * (dp - (bhp + d.pf_head[e]))) * d.pf_wi[e];
      const bool live = inj ? raw.v < 0.0 : raw.v > 0.0;
      const D2 q = live ? raw : dconst<2>(0.0);
      const double dq_dbhp = live ? -cm.v * d.pf_wi[e] : 0.0;
      d.pf_q[e] = q.v;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        d.R[r * N + g] += coef[r].v * q.v * dt;
#pragma unroll
        for (int v = 0; v < 2; ++v) d.J[(size_t)(r * 2 + v) * N + g] += (coef[r].d[v] * q.v + coef[r].v * q.d[v]) * dt;
        d.pf_b[2 * e + r] = coef[r].v * dq_dbhp * dt;
      }
      if (fb) {
        // 1/Bw or fo/Bo, selected value by value (a reference picked at run time put coef
        // in scratch memory)
        const double fv = inj ? coef[0].v : coef[1].v;
        const double fd0 = inj ? coef[0].d[0] : coef[1].d[0], fd1 = inj ? coef[0].d[1] : coef[1].d[1];
        double drq;
        const double rate = control_rate(inj, surf, q.v, fv, drq);
        const double cl = inj ? (-q.v >= 0.0 ? -q.v : 0.0) : (q.v >= 0.0 ? q.v : 0.0);
        actual += rate;
        d.pf_h[2 * e] = drq * q.d[0] + (surf ? cl * fd0 : 0.0);
        d.pf_h[2 * e + 1] = drq * q.d[1] + (surf ? cl * fd1 : 0.0);
        drb_sum += drq * dq_dbhp;
      }
    }
    if (enabled && fb) {
      double rw, alpha, dfb, sr;
      fb_control_row(inj, bhp, d.w_limit[w], d.w_target[w], actual, rw, alpha, dfb, sr);
      const double f = dfb * (-1.0 / sr);
      for (int e = d.w_pf[w]; e < d.w_pf[w + 1]; ++e) {
        d.pf_h[2 * e] *= f;
        d.pf_h[2 * e + 1] *= f;
      }
      d.w_rw[w] = rw;
      d.w_s[w] = alpha + f * drb_sum;
    } else {
      d.w_rw[w] = 0.0;
      d.w_s[w] = 0.0;
    }
  }
}

__device__ __forceinline__ bool schur_ok(double s) { return fabs(s) > 0.0 && isfinite(s); }

// ------------------------------------------------------------------ block reductions
// sum over the workgroup; every thread gets the result
__device__ __forceinline__ double block_sum(double v, double* sm) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) sm[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < nwv; ++k) s += sm[k];
  return s;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}
__device__ __forceinline__ double block_max(double v, double* sm) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = (blockDim.x + 63) >> 6;
  v = wave_max(v);
  __syncthreads();
  if (lane == 0) sm[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < nwv; ++k) s = fmax(s, sm[k]);
  return s;
}

// per tile: scaled family maxima, residual sums, FVF sums, pore volume, non-finite flag
constexpr int kConvVals = 8;
__global__ __launch_bounds__(kMaxThreads) void k_conv_partial(Dev d, const Tile* __restrict__ tiles) {
  __shared__ double sm[kMaxThreads / 64];
  const Tile t = tiles[blockIdx.x];
  if (!running(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const int N = d.Nt;
  const double dt = d.ctl[t.m].dt;
  double v[kConvVals] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    const double pi = d.p[g];
    double bo = base_bo<double>(M.P, pi);
    bo = bo > 1e-9 ? bo : 1e-9;
    const double ibw = inv_bw<double>(M.P, pi);
    const double bw = 1.0 / (ibw > 1e-9 ? ibw : 1e-9);
    const double pv = d.pv[g];
    double pvdt = fabs(pv / (dt > 1e-12 ? dt : 1e-12));
    pvdt = pvdt > 1.0 ? pvdt : 1.0;
    const double r0 = d.R[g], r1 = d.R[N + g];
    v[0] = fmax(v[0], fabs(r0) / (pvdt / bw));
    v[1] = fmax(v[1], fabs(r1) / (pvdt / bo));
    v[2] += r0;
    v[3] += r1;
    v[4] += bw;
    v[5] += bo;
    v[6] += pv > 1e-9 ? pv : 1e-9;
    if (!isfinite(r0) || !isfinite(r1) || !isfinite(pi) || !isfinite(d.sw[g])) v[7] = 1.0;
  }
  double out[kConvVals];
  out[0] = block_max(v[0], sm);
  out[1] = block_max(v[1], sm);
  for (int k = 2; k < 7; ++k) out[k] = block_sum(v[k], sm);
  out[7] = block_max(v[7], sm);
  if (threadIdx.x == 0)
    for (int k = 0; k < kConvVals; ++k) d.part[(size_t)blockIdx.x * d.pstride + k] = out[k];
}

// one thread per member: acceptance (rs/fim/newton/convergence.rs:352-385, residual_tol mode)
__global__ void k_conv_finish(Dev d) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= d.nM || !running(d, m)) return;
  const Member& M = d.mem[m];
  const RunDev& r = *d.run;
  double v[kConvVals] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = M.t0; b < M.t0 + M.nt; ++b) {
    const double* p = d.part + (size_t)b * d.pstride;
    v[0] = fmax(v[0], p[0]);
    v[1] = fmax(v[1], p[1]);
    for (int k = 2; k < 7; ++k) v[k] += p[k];
    v[7] = fmax(v[7], p[7]);
  }
  Ctl& c = d.ctl[m];
  const double nn = M.n > 1 ? M.n : 1;
  c.fam[0] = v[0];
  c.fam[1] = v[1];
  c.mb[0] = fabs(v[4] / nn * v[2]) / fmax(v[6], 1e-9);
  c.mb[1] = fabs(v[5] / nn * v[3]) / fmax(v[6], 1e-9);
  double wres = 0.0;
  bool bad = v[7] != 0.0;
  for (int w = M.w0; w < M.w0 + M.nw; ++w) {
    if ((d.w_flags[w] & 18) == 18) wres = fmax(wres, fabs(d.w_rw[w]));
    if (!isfinite(d.w_rw[w]) || !isfinite(d.w_bhp[w])) bad = true;
  }
  c.wres = wres;
  const bool accept = c.it >= r.min_newton && fmax(c.fam[0], c.fam[1]) <= r.tol && c.mb[0] <= r.mb_tol &&
                      c.mb[1] <= r.mb_tol && wres <= r.tol;
  if (bad) {
    c.phase = kFail;
  } else if (accept) {
    c.phase = kOk;
  } else if (c.it >= r.max_newton) {
    c.phase = kFail;
  } else {
    c.solve = 1;
  }
}

// ------------------------------------------------------------------ linear system
// b = -R; the Krylov solution starts at zero
__global__ __launch_bounds__(kMaxThreads) void k_rhs(Dev d, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  if (!solving(d, t.m)) return;
  const Member& M = d.mem[t.m];
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    d.b[g] = -d.R[g];
    d.b[d.Nt + g] = -d.R[d.Nt + g];
  }
}

// per member: + B r_w / s (reduced right-hand side), and the Schur fill on the stencil
// pattern into the preconditioner copy
__global__ void k_well_reduce(Dev d) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= d.nM || !solving(d, m)) return;
  const Member& M = d.mem[m];
  const Grid G = mgrid(M);
  const int N = d.Nt;
  for (int w = M.w0; w < M.w0 + M.nw; ++w) {
    const double s = d.w_s[w];
    if ((d.w_flags[w] & 18) != 18 || !schur_ok(s)) continue;
    const double rw = d.w_rw[w];
    for (int e = d.w_pf[w]; e < d.w_pf[w + 1]; ++e) {
      const int ce = d.pf_cell[e], g = M.off + ce;
      d.b[g] += d.pf_b[2 * e] * rw / s;
      d.b[N + g] += d.pf_b[2 * e + 1] * rw / s;
      int x, y, z;
      cell_ijk(G, ce, x, y, z);
      const int slot = M.off + cslot(M, ce);
      for (int f = d.w_pf[w]; f < d.w_pf[w + 1]; ++f) {
        const int cf = d.pf_cell[f];
        int sl = -1;
        if (cf == ce) {
          sl = 0;
        } else {
          for (int k = 1; k <= 6; ++k)
            if (slot_nbr(G, ce, k, x, y, z) == cf) sl = k;
        }
        if (sl < 0) continue;
        for (int r = 0; r < 2; ++r)
          for (int v = 0; v < 2; ++v) {
            const float val = float(-d.pf_b[2 * e + r] * d.pf_h[2 * f + v] / s);
            if (sl == 0) d.Dg[(size_t)(r * 2 + v) * N + slot] += val;
            else d.Mo[(size_t)((sl - 1) * 4 + r * 2 + v) * N + slot] += val;
          }
      }
    }
  }
}

// FP32 colour-compacted copy of the Jacobian's blocks (before the Schur fill)
__global__ __launch_bounds__(kMaxThreads) void k_pcopy(Dev d, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  if (!solving(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const int N = d.Nt;
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i, sl = M.off + cslot(M, i);
#pragma unroll
    for (int e = 0; e < 4; ++e) d.Dg[(size_t)e * N + sl] = float(d.J[(size_t)e * N + g]);
#pragma unroll
    for (int k = 0; k < 24; ++k) d.Mo[(size_t)k * N + sl] = float(d.J[(size_t)(4 + k) * N + g]);
  }
}

__device__ __forceinline__ int opposite(int s) { return (s & 1) ? s + 1 : s - 1; }

// red: D^-1 = A_rr^-1
__global__ __launch_bounds__(kMaxThreads) void k_factor_red(Dev d, const Tile* __restrict__ red) {
  const Tile t = red[blockIdx.x];
  if (!solving(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const int N = d.Nt;
  for (int k = t.first + threadIdx.x; k < t.first + t.count; k += blockDim.x) {
    const int sl = M.off + k;
    float o[4];
    inv2(d.Dg[sl], d.Dg[(size_t)N + sl], d.Dg[(size_t)2 * N + sl], d.Dg[(size_t)3 * N + sl], o);
#pragma unroll
    for (int e = 0; e < 4; ++e) d.Di[(size_t)e * N + sl] = o[e];
  }
}

// black: D = A_bb - sum_j A_bj D_j^-1 A_jb over the red neighbours j
__global__ __launch_bounds__(kMaxThreads) void k_factor_black(Dev d, const Tile* __restrict__ black) {
  const Tile t = black[blockIdx.x];
  if (!solving(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const Grid G = mgrid(M);
  const int N = d.Nt;
  for (int k = t.first + threadIdx.x; k < t.first + t.count; k += blockDim.x) {
    const int i = cell_of_slot(M, 1, k), sl = M.off + M.nred + k;
    int x, y, z;
    cell_ijk(G, i, x, y, z);
    float D[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) D[e] = d.Dg[(size_t)e * N + sl];
    for (int s = 1; s <= 6; ++s) {
      const int j = slot_nbr(G, i, s, x, y, z);
      if (j < 0) continue;
      const int sj = M.off + (j >> 1);  // red slot of j
      float a[4], di[4], c[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        a[e] = d.Mo[(size_t)((s - 1) * 4 + e) * N + sl];
        di[e] = d.Di[(size_t)e * N + sj];
        c[e] = d.Mo[(size_t)((opposite(s) - 1) * 4 + e) * N + sj];
      }
      // a * di * c
      const float t0 = a[0] * di[0] + a[1] * di[2], t1 = a[0] * di[1] + a[1] * di[3];
      const float t2 = a[2] * di[0] + a[3] * di[2], t3 = a[2] * di[1] + a[3] * di[3];
      D[0] -= t0 * c[0] + t1 * c[2];
      D[1] -= t0 * c[1] + t1 * c[3];
      D[2] -= t2 * c[0] + t3 * c[2];
      D[3] -= t2 * c[1] + t3 * c[3];
    }
    float o[4];
    inv2(D[0], D[1], D[2], D[3], o);
#pragma unroll
    for (int e = 0; e < 4; ++e) d.Di[(size_t)e * N + sl] = o[e];
  }
}

// z = M^-1 v, three passes: red w = D^-1 v; black z = D^-1 (v - sum A w); red
// z = w - D^-1 sum A z. v: FP64 natural (2 planes), z: FP32 natural.
__global__ __launch_bounds__(kMaxThreads) void k_ilu_red1(Dev d, const Tile* __restrict__ red, const double* __restrict__ v,
                                                          float* __restrict__ z) {
  const Tile t = red[blockIdx.x];
  if (!kactive(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const int N = d.Nt;
  for (int k = t.first + threadIdx.x; k < t.first + t.count; k += blockDim.x) {
    const int g = M.off + cell_of_slot(M, 0, k), sl = M.off + k;
    const float v0 = float(v[g]), v1 = float(v[N + g]);
    z[g] = d.Di[sl] * v0 + d.Di[(size_t)N + sl] * v1;
    z[N + g] = d.Di[(size_t)2 * N + sl] * v0 + d.Di[(size_t)3 * N + sl] * v1;
  }
}

__global__ __launch_bounds__(kMaxThreads) void k_ilu_black(Dev d, const Tile* __restrict__ black,
                                                           const double* __restrict__ v, float* __restrict__ z) {
  const Tile t = black[blockIdx.x];
  if (!kactive(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const Grid G = mgrid(M);
  const int N = d.Nt;
  for (int k = t.first + threadIdx.x; k < t.first + t.count; k += blockDim.x) {
    const int i = cell_of_slot(M, 1, k), g = M.off + i, sl = M.off + M.nred + k;
    int x, y, z_;
    cell_ijk(G, i, x, y, z_);
    float r0 = float(v[g]), r1 = float(v[N + g]);
    for (int s = 1; s <= 6; ++s) {
      const int j = slot_nbr(G, i, s, x, y, z_);
      if (j < 0) continue;
      const float w0 = z[M.off + j], w1 = z[N + M.off + j];
      r0 -= d.Mo[(size_t)((s - 1) * 4) * N + sl] * w0 + d.Mo[(size_t)((s - 1) * 4 + 1) * N + sl] * w1;
      r1 -= d.Mo[(size_t)((s - 1) * 4 + 2) * N + sl] * w0 + d.Mo[(size_t)((s - 1) * 4 + 3) * N + sl] * w1;
    }
    z[g] = d.Di[sl] * r0 + d.Di[(size_t)N + sl] * r1;
    z[N + g] = d.Di[(size_t)2 * N + sl] * r0 + d.Di[(size_t)3 * N + sl] * r1;
  }
}

__global__ __launch_bounds__(kMaxThreads) void k_ilu_red2(Dev d, const Tile* __restrict__ red, float* __restrict__ z) {
  const Tile t = red[blockIdx.x];
  if (!kactive(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const Grid G = mgrid(M);
  const int N = d.Nt;
  for (int k = t.first + threadIdx.x; k < t.first + t.count; k += blockDim.x) {
    const int i = cell_of_slot(M, 0, k), g = M.off + i, sl = M.off + k;
    int x, y, z_;
    cell_ijk(G, i, x, y, z_);
    float a0 = 0.0f, a1 = 0.0f;
    for (int s = 1; s <= 6; ++s) {
      const int j = slot_nbr(G, i, s, x, y, z_);
      if (j < 0) continue;
      const float w0 = z[M.off + j], w1 = z[N + M.off + j];
      a0 += d.Mo[(size_t)((s - 1) * 4) * N + sl] * w0 + d.Mo[(size_t)((s - 1) * 4 + 1) * N + sl] * w1;
      a1 += d.Mo[(size_t)((s - 1) * 4 + 2) * N + sl] * w0 + d.Mo[(size_t)((s - 1) * 4 + 3) * N + sl] * w1;
    }
    z[g] -= d.Di[sl] * a0 + d.Di[(size_t)N + sl] * a1;
    z[N + g] -= d.Di[(size_t)2 * N + sl] * a0 + d.Di[(size_t)3 * N + sl] * a1;
  }
}

// ------------------------------------------------------------------ CPR pressure stage
// per cell: quasi-IMPES weights w = first row of D^-1 (D the preconditioner's diagonal
// block: w^T D = e_p^T) and the coarse pressure row P_ij = w_i^T A_ij e_p
// (rs/fim/linear/gmres_block_jacobi.rs quasi-IMPES restriction)
__global__ __launch_bounds__(kMaxThreads) void k_cpr_setup(Dev d, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  if (!solving(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const int N = d.Nt;
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i, sl = M.off + cslot(M, i);
    const float a = d.Dg[sl], b = d.Dg[(size_t)N + sl], c = d.Dg[(size_t)2 * N + sl], e = d.Dg[(size_t)3 * N + sl];
    const float det = a * e - b * c;
    float w0 = 1.0f, w1 = 0.0f;
    if (fabsf(det) > 0.0f && isfinite(det)) {
      w0 = e / det;
      w1 = -b / det;
    }
    d.Wq[g] = w0;
    d.Wq[(size_t)N + g] = w1;
    d.Pp[g] = w0 * a + w1 * c;
    for (int s = 1; s <= 6; ++s)
      d.Pp[(size_t)s * N + g] = w0 * d.Mo[(size_t)((s - 1) * 4) * N + sl] + w1 * d.Mo[(size_t)((s - 1) * 4 + 2) * N + sl];
  }
}

// pressure ILU(0) in red-black order: red 1/P_rr, black 1/(P_bb - sum P_bj P_jb / P_jj)
__global__ __launch_bounds__(kMaxThreads) void k_pfactor(Dev d, const Tile* __restrict__ tiles, int colour) {
  const Tile t = tiles[blockIdx.x];
  if (!solving(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const Grid G = mgrid(M);
  const int N = d.Nt;
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    int x, y, z;
    cell_ijk(G, i, x, y, z);
    if (((x + y + z) & 1) != colour) continue;
    const int g = M.off + i;
    float dg = d.Pp[g];
    if (colour == 1)
      for (int s = 1; s <= 6; ++s) {
        const int j = slot_nbr(G, i, s, x, y, z);
        if (j >= 0) dg -= d.Pp[(size_t)s * N + g] * d.Pp[(size_t)opposite(s) * N + M.off + j] * d.pdi[M.off + j];
      }
    d.pdi[g] = (fabsf(dg) > 0.0f && isfinite(dg)) ? 1.0f / dg : 1.0f;
  }
}

// restriction rp = w^T v of the Krylov vector
__global__ __launch_bounds__(kMaxThreads) void k_cpr_restrict(Dev d, const Tile* __restrict__ tiles,
                                                              const double* __restrict__ v) {
  const Tile t = tiles[blockIdx.x];
  if (!kactive(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const int N = d.Nt;
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    d.rp[g] = d.Wq[g] * float(v[g]) + d.Wq[(size_t)N + g] * float(v[N + g]);
  }
}

constexpr int kPB = 1024;  // threads of the pressure BiCGStab workgroup

__device__ __forceinline__ double pb_sum(double v, double* sm) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) sm[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < kPB / 64; ++k) s += sm[k];
  return s;
}

// stencil offsets of slots 1..6 (x-, x+, y-, y+, z-, z+) in a member's natural order
// One member's pressure system seen by its BiCGStab workgroup: member-local pointers (the
// planes never overlap, so they are declared restrict and loads can run ahead of stores), the
// per-cell neighbour/colour mask and the stencil offsets of slots 1..6.
struct PbView {
  const float* __restrict__ P;  // plane s at P + s * N
  const float* __restrict__ di;
  const unsigned char* __restrict__ mk;
  int N, n;
  int o[7];
  __device__ PbView(const Dev& d, const Member& M)
      : P(d.Pp + M.off), di(d.pdi + M.off), mk(d.nmask + M.off), N(d.Nt), n(M.n) {
    const int nxy = M.nx * M.ny;
    o[0] = 0; o[1] = -1; o[2] = 1; o[3] = -M.nx; o[4] = M.nx; o[5] = -nxy; o[6] = nxy;
  }
};

// Cells per thread whose loads are issued together in the passes below: one workgroup
// carries a whole member, so without this a pass waits one memory latency per cell.
constexpr int kPU = 4;

// The stencil row of cell i: t0 + sum over existing neighbours (slots 1..6, in order) of
// sign * P_s x_j. Loads are unconditional (a missing neighbour reads the cell itself) and the
// terms are selected, so the sum is the one that skips them.
template <int SIGN>
__device__ __forceinline__ void pb_gather(const PbView& w, const float* __restrict__ x, int i, unsigned m, float c[6],
                                          float xs[6]) {
#pragma unroll
  for (int s = 1; s <= 6; ++s) {
    const bool has = (m >> (s - 1)) & 1u;
    c[s - 1] = w.P[(size_t)s * w.N + i];
    xs[s - 1] = x[has ? i + w.o[s] : i];
  }
}
template <int SIGN>
__device__ __forceinline__ float pb_sum_row(unsigned m, float a, const float c[6], const float xs[6]) {
#pragma unroll
  for (int s = 1; s <= 6; ++s) {
    const float t = SIGN > 0 ? a + c[s - 1] * xs[s - 1] : a - c[s - 1] * xs[s - 1];
    a = ((m >> (s - 1)) & 1u) ? t : a;
  }
  return a;
}

// out = ILU(0)^-1 in, red-black, over the member (in and out member-local)
__device__ void pb_ilu(const PbView& w, const float* __restrict__ in, float* __restrict__ out) {
  const int n = w.n;
  for (int i0 = threadIdx.x; i0 < n; i0 += kPU * kPB) {
    unsigned m[kPU];
    float a[kPU], dd[kPU];
#pragma unroll
    for (int u = 0; u < kPU; ++u) {
      const int i = min(i0 + u * kPB, n - 1);
      m[u] = w.mk[i];
      a[u] = in[i];
      dd[u] = w.di[i];
    }
#pragma unroll
    for (int u = 0; u < kPU; ++u) {
      const int i = i0 + u * kPB;
      if (i < n && !(m[u] & 64)) out[i] = a[u] * dd[u];
    }
  }
  __syncthreads();
  for (int i0 = threadIdx.x; i0 < n; i0 += kPU * kPB) {  // black: y = (in - L x_red) / d
    unsigned m[kPU];
    float a[kPU], dd[kPU], c[kPU][6], xs[kPU][6];
#pragma unroll
    for (int u = 0; u < kPU; ++u) {
      const int i = min(i0 + u * kPB, n - 1);
      m[u] = w.mk[i];
      a[u] = in[i];
      dd[u] = w.di[i];
      pb_gather<-1>(w, out, i, m[u], c[u], xs[u]);
    }
#pragma unroll
    for (int u = 0; u < kPU; ++u) {
      const int i = i0 + u * kPB;
      if (i < n && (m[u] & 64)) out[i] = pb_sum_row<-1>(m[u], a[u], c[u], xs[u]) * dd[u];
    }
  }
  __syncthreads();
  for (int i0 = threadIdx.x; i0 < n; i0 += kPU * kPB) {  // red: x = y - U x_black / d
    unsigned m[kPU];
    float y[kPU], dd[kPU], c[kPU][6], xs[kPU][6];
#pragma unroll
    for (int u = 0; u < kPU; ++u) {
      const int i = min(i0 + u * kPB, n - 1);
      m[u] = w.mk[i];
      y[u] = out[i];
      dd[u] = w.di[i];
      pb_gather<1>(w, out, i, m[u], c[u], xs[u]);
    }
#pragma unroll
    for (int u = 0; u < kPU; ++u) {
      const int i = i0 + u * kPB;
      if (i < n && !(m[u] & 64)) out[i] = y[u] - pb_sum_row<1>(m[u], 0.0f, c[u], xs[u]) * dd[u];
    }
  }
  __syncthreads();
}

// q = P x over the member, with dot = sum of q * a1 (and, with b given, also q * b and q * q):
// the partial sums of this thread in cell order
template <bool TWO>
__device__ void pb_rows(const PbView& w, const float* __restrict__ x, float* __restrict__ q,
                        const float* __restrict__ a1v, double& d1, double& d2) {
  const int n = w.n;
  for (int i0 = threadIdx.x; i0 < n; i0 += kPU * kPB) {
    unsigned m[kPU];
    float c0[kPU], x0[kPU], av[kPU], c[kPU][6], xs[kPU][6];
#pragma unroll
    for (int u = 0; u < kPU; ++u) {
      const int i = min(i0 + u * kPB, n - 1);
      m[u] = w.mk[i];
      c0[u] = w.P[i];
      x0[u] = x[i];
      av[u] = a1v[i];
      pb_gather<1>(w, x, i, m[u], c[u], xs[u]);
    }
#pragma unroll
    for (int u = 0; u < kPU; ++u) {
      const int i = i0 + u * kPB;
      if (i >= n) continue;
      const float r = pb_sum_row<1>(m[u], c0[u] * x0[u], c[u], xs[u]);
      q[i] = r;
      if (TWO) {  // t . s and t . t
        d1 += double(r) * av[u];
        d2 += double(r) * r;
      } else {  // rh . v
        d1 += double(av[u]) * r;
      }
    }
  }
}

// Pressure BiCGStab with ILU(0) right preconditioning (the recurrences and breakdown
// tests of the single-model path; rs/fim/linear/gmres_block_jacobi.rs:660-730): one
// workgroup per member, x = 0 start, stop at |r| <= p_tol |b| or p_max_it iterations.
// Work vectors in bw: r, rh, p, v, s, t, x, ph, sh (9 planes of N).
__global__ __launch_bounds__(kPB) void k_pbicg(Dev d) {
  __shared__ double sm[kPB / 64];
  const int m = blockIdx.x;
  if (!kactive(d, m)) return;
  const Member& M = d.mem[m];
  const PbView w(d, M);
  const RunDev& rd = *d.run;
  const int N = d.Nt, n = M.n, off = M.off;
  float* __restrict__ r = d.bw + off;
  float* __restrict__ rh = d.bw + (size_t)N + off;
  float* __restrict__ pv = d.bw + 2 * (size_t)N + off;
  float* __restrict__ v = d.bw + 3 * (size_t)N + off;
  float* __restrict__ sv = d.bw + 4 * (size_t)N + off;
  float* __restrict__ tv = d.bw + 5 * (size_t)N + off;
  float* __restrict__ xv = d.bw + 6 * (size_t)N + off;
  float* __restrict__ ph = d.bw + 7 * (size_t)N + off;
  float* __restrict__ sh = d.bw + 8 * (size_t)N + off;
  double b2 = 0.0;
  for (int i = threadIdx.x; i < n; i += kPB) {
    const float b = d.rp[off + i];
    r[i] = b;
    rh[i] = b;
    pv[i] = 0.0f;
    v[i] = 0.0f;
    xv[i] = 0.0f;
    b2 += double(b) * b;
  }
  b2 = pb_sum(b2, sm);
  double rho = 1.0, alpha = 1.0, omega = 1.0, rho_new = b2;
  const double stop2 = rd.p_tol * rd.p_tol * b2;
  if (b2 > 0.0 && isfinite(b2)) {
    for (int it = 0; it < rd.p_max_it; ++it) {
      const double beta = (rho_new / rho) * (alpha / omega);
      rho = rho_new;
#pragma unroll 4
      for (int i = threadIdx.x; i < n; i += kPB) pv[i] = r[i] + float(beta) * (pv[i] - float(omega) * v[i]);
      __syncthreads();
      pb_ilu(w, pv, ph);
      double a1 = 0.0, unused = 0.0;
      pb_rows<false>(w, ph, v, rh, a1, unused);
      a1 = pb_sum(a1, sm);
      if (!(fabs(a1) > 0.0) || !isfinite(a1)) break;
      alpha = rho / a1;
#pragma unroll 4
      for (int i = threadIdx.x; i < n; i += kPB) sv[i] = r[i] - float(alpha) * v[i];
      __syncthreads();
      pb_ilu(w, sv, sh);
      double ts = 0.0, tt = 0.0;
      pb_rows<true>(w, sh, tv, sv, ts, tt);
      ts = pb_sum(ts, sm);
      tt = pb_sum(tt, sm);
      omega = tt > 0.0 ? ts / tt : 0.0;
      double rr = 0.0, rn = 0.0;
#pragma unroll 4
      for (int i = threadIdx.x; i < n; i += kPB) {
        xv[i] += float(alpha) * ph[i] + float(omega) * sh[i];
        const float q = sv[i] - float(omega) * tv[i];
        r[i] = q;
        rr += double(q) * q;
        rn += double(rh[i]) * q;
      }
      rr = pb_sum(rr, sm);
      rn = pb_sum(rn, sm);
      rho_new = rn;
      if (rr <= stop2 || !(fabs(omega) > 0.0) || !(fabs(rho_new) > 0.0) || !isfinite(rr)) break;
    }
  }
  for (int i = threadIdx.x; i < n; i += kPB) d.xp[off + i] = xv[i];
}

// second-stage residual cv = v - A [xp, 0]: the stencil part (p-columns of J)
__global__ __launch_bounds__(kMaxThreads) void k_cpr_correct(Dev d, const Tile* __restrict__ tiles,
                                                             const double* __restrict__ v) {
  const Tile t = tiles[blockIdx.x];
  if (!kactive(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const Grid G = mgrid(M);
  const int N = d.Nt;
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    int x, y, z;
    cell_ijk(G, i, x, y, z);
    double a0 = 0.0, a1 = 0.0;
    for (int s = 0; s <= 6; ++s) {
      const int j = s == 0 ? i : slot_nbr(G, i, s, x, y, z);
      if (j < 0) continue;
      const double xq = d.xp[M.off + j];
      a0 += d.J[(size_t)(s * 4) * N + g] * xq;
      a1 += d.J[(size_t)(s * 4 + 2) * N + g] * xq;
    }
    d.cv[g] = v[g] - a0;
    d.cv[N + g] = v[N + g] - a1;
  }
}

// ... and the wells' rank-one part: cv += B (H_p . xp) / s
__global__ void k_cpr_well(Dev d) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= d.nM || !kactive(d, m)) return;
  const Member& M = d.mem[m];
  const int N = d.Nt;
  for (int wl = M.w0; wl < M.w0 + M.nw; ++wl) {
    const double s = d.w_s[wl];
    if ((d.w_flags[wl] & 18) != 18 || !schur_ok(s)) continue;
    double hz = 0.0;
    for (int e = d.w_pf[wl]; e < d.w_pf[wl + 1]; ++e) hz += d.pf_h[2 * e] * double(d.xp[M.off + d.pf_cell[e]]);
    const double f = hz / s;
    for (int e = d.w_pf[wl]; e < d.w_pf[wl + 1]; ++e) {
      const int g = M.off + d.pf_cell[e];
      d.cv[g] += d.pf_b[2 * e] * f;
      d.cv[N + g] += d.pf_b[2 * e + 1] * f;
    }
  }
}

// z += [xp, 0] (the pressure correction prolonged)
__global__ __launch_bounds__(kMaxThreads) void k_cpr_combine(Dev d, const Tile* __restrict__ tiles, float* __restrict__ z) {
  const Tile t = tiles[blockIdx.x];
  if (!kactive(d, t.m)) return;
  const Member& M = d.mem[t.m];
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    z[g] += d.xp[g];
  }
}

// w = J z (FP64 blocks, FP32 z)
__global__ __launch_bounds__(kMaxThreads) void k_spmv(Dev d, const Tile* __restrict__ tiles, const float* __restrict__ z,
                                                      double* __restrict__ w) {
  const Tile t = tiles[blockIdx.x];
  if (!kactive(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const Grid G = mgrid(M);
  const int N = d.Nt;
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    int x, y, z_;
    cell_ijk(G, i, x, y, z_);
    double a0 = 0.0, a1 = 0.0;
    for (int s = 0; s <= 6; ++s) {
      const int j = s == 0 ? i : slot_nbr(G, i, s, x, y, z_);
      if (j < 0) continue;
      const double x0 = z[M.off + j], x1 = z[N + M.off + j];
      a0 += d.J[(size_t)(s * 4) * N + g] * x0 + d.J[(size_t)(s * 4 + 1) * N + g] * x1;
      a1 += d.J[(size_t)(s * 4 + 2) * N + g] * x0 + d.J[(size_t)(s * 4 + 3) * N + g] * x1;
    }
    w[g] = a0;
    w[N + g] = a1;
  }
}

// the wells' rank-one terms of the reduced operator: w -= B (H . z) / s
__global__ void k_well_apply(Dev d, const float* __restrict__ z, double* __restrict__ w) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= d.nM || !kactive(d, m)) return;
  const Member& M = d.mem[m];
  const int N = d.Nt;
  for (int wl = M.w0; wl < M.w0 + M.nw; ++wl) {
    const double s = d.w_s[wl];
    if ((d.w_flags[wl] & 18) != 18 || !schur_ok(s)) continue;
    double hz = 0.0;
    for (int e = d.w_pf[wl]; e < d.w_pf[wl + 1]; ++e) {
      const int g = M.off + d.pf_cell[e];
      hz += d.pf_h[2 * e] * double(z[g]) + d.pf_h[2 * e + 1] * double(z[N + g]);
    }
    const double f = hz / s;
    for (int e = d.w_pf[wl]; e < d.w_pf[wl + 1]; ++e) {
      const int g = M.off + d.pf_cell[e];
      w[g] -= d.pf_b[2 * e] * f;
      w[N + g] -= d.pf_b[2 * e + 1] * f;
    }
  }
}

// per tile partial |x|^2 into part[tile * pstride]
__global__ __launch_bounds__(kMaxThreads) void k_norm_partial(Dev d, const Tile* __restrict__ tiles,
                                                              const double* __restrict__ x, int need_solve) {
  __shared__ double sm[kMaxThreads / 64];
  const Tile t = tiles[blockIdx.x];
  if (need_solve ? !solving(d, t.m) : !kactive(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const int N = d.Nt;
  double a = 0.0;
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    a += x[g] * x[g] + x[N + g] * x[N + g];
  }
  a = block_sum(a, sm);
  if (threadIdx.x == 0) d.part[(size_t)blockIdx.x * d.pstride] = a;
}

// beta = |b|, g = (beta, 0, ...), Krylov active when there is something to solve
__global__ void k_beta(Dev d) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= d.nM) return;
  Ctl& c = d.ctl[m];
  c.kact = 0;
  c.k = 0;
  if (!(c.phase == kIter && c.solve)) return;
  const Member& M = d.mem[m];
  double s = 0.0;
  for (int b = M.t0; b < M.t0 + M.nt; ++b) s += d.part[(size_t)b * d.pstride];
  const double beta = sqrt(s);
  c.beta = beta;
  c.red = 0.0;
  double* g = d.g + (size_t)m * (d.K + 1);
  for (int i = 0; i <= d.K; ++i) g[i] = 0.0;
  g[0] = beta;
  c.kact = (beta > 0.0 && isfinite(beta)) ? 1 : 0;
}

// dst = src * scale (scale = 1/beta at j = 0, else 1/h_{j+1,j})
__global__ __launch_bounds__(kMaxThreads) void k_scale(Dev d, const Tile* __restrict__ tiles, const double* __restrict__ src,
                                                       double* __restrict__ dst, int first) {
  const Tile t = tiles[blockIdx.x];
  const Ctl& c = d.ctl[t.m];
  if (c.kact != 1) return;
  const Member& M = d.mem[t.m];
  const int N = d.Nt;
  const double f = first ? 1.0 / c.beta : d.hc[(size_t)t.m * 2 * (d.K + 1) + 2 * d.K + 1];
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    dst[g] = src[g] * f;
    dst[N + g] = src[N + g] * f;
  }
}

// partial dots <w, V_i>, i <= j, per tile
__global__ __launch_bounds__(kMaxThreads) void k_mdot(Dev d, const Tile* __restrict__ tiles, int j) {
  __shared__ double sm[kMaxKrylov + 1][kMaxThreads / 64];
  const Tile t = tiles[blockIdx.x];
  if (!kactive(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const int N = d.Nt;
  const size_t vs = 2 * (size_t)N;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = (blockDim.x + 63) >> 6;
  for (int q = 0; q <= j; ++q) {
    const double* vq = d.V + q * vs;
    double a = 0.0;
    for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
      const int g = M.off + i;
      a += d.Wv[g] * vq[g] + d.Wv[N + g] * vq[N + g];
    }
    a = wave_sum(a);
    if (lane == 0) sm[q][wv] = a;
  }
  __syncthreads();
  for (int q = threadIdx.x; q <= j; q += blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < nwv; ++k) s += sm[q][k];
    d.part[(size_t)blockIdx.x * d.pstride + q] = s;
  }
}

// one workgroup per member: h_q = sum over its tiles (fixed order); pass 0 keeps it in
// hc, pass 1 adds the correction into H's column j (the coefficients of the second
// projection go to hc for the update)
__global__ void k_mdot_finish(Dev d, int j, int pass) {
  const int m = blockIdx.x;
  if (!kactive(d, m)) return;
  const Member& M = d.mem[m];
  double* hc = d.hc + (size_t)m * 2 * (d.K + 1);
  double* H = d.H + (size_t)m * (d.K + 1) * d.K;
  for (int q = threadIdx.x; q <= j; q += blockDim.x) {
    double s = 0.0;
    for (int b = M.t0; b < M.t0 + M.nt; ++b) s += d.part[(size_t)b * d.pstride + q];
    if (pass == 0) {
      hc[q] = s;
      H[(size_t)j * (d.K + 1) + q] = s;
    } else {
      hc[q] = s;
      H[(size_t)j * (d.K + 1) + q] += s;
    }
  }
}

// w -= sum_q hc_q V_q; with `norm` the partial |w|^2 per tile as well
__global__ __launch_bounds__(kMaxThreads) void k_orth(Dev d, const Tile* __restrict__ tiles, int j, int norm) {
  __shared__ double sm[kMaxThreads / 64];
  const Tile t = tiles[blockIdx.x];
  if (!kactive(d, t.m)) return;
  const Member& M = d.mem[t.m];
  const int N = d.Nt;
  const size_t vs = 2 * (size_t)N;
  const double* hc = d.hc + (size_t)t.m * 2 * (d.K + 1);
  double a = 0.0;
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    double w0 = d.Wv[g], w1 = d.Wv[N + g];
    for (int q = 0; q <= j; ++q) {
      const double h = hc[q];
      w0 -= h * d.V[q * vs + g];
      w1 -= h * d.V[q * vs + N + g];
    }
    d.Wv[g] = w0;
    d.Wv[N + g] = w1;
    a += w0 * w0 + w1 * w1;
  }
  if (norm) {
    a = block_sum(a, sm);
    if (threadIdx.x == 0) d.part[(size_t)blockIdx.x * d.pstride] = a;
  }
}

// one thread per member: h_{j+1,j}, the Givens rotations of column j, the residual
// estimate |g_{j+1}| and the stop test
__global__ void k_arnoldi(Dev d, int j) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= d.nM) return;
  Ctl& c = d.ctl[m];
  if (c.kact != 1) return;
  const Member& M = d.mem[m];
  const int K = d.K;
  double s = 0.0;
  for (int b = M.t0; b < M.t0 + M.nt; ++b) s += d.part[(size_t)b * d.pstride];
  const double hn = sqrt(s);
  double* H = d.H + (size_t)m * (K + 1) * K;
  double* h = H + (size_t)j * (K + 1);
  double* g = d.g + (size_t)m * (K + 1);
  double* cs = d.cs + (size_t)m * K;
  double* sn = d.sn + (size_t)m * K;
  h[j + 1] = hn;
  for (int i = 0; i < j; ++i) {
    const double a = cs[i] * h[i] + sn[i] * h[i + 1];
    h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
    h[i] = a;
  }
  const double r = sqrt(h[j] * h[j] + h[j + 1] * h[j + 1]);
  if (r > 0.0 && isfinite(r)) {
    cs[j] = h[j] / r;
    sn[j] = h[j + 1] / r;
  } else {
    cs[j] = 1.0;
    sn[j] = 0.0;
  }
  h[j] = cs[j] * h[j] + sn[j] * h[j + 1];
  h[j + 1] = 0.0;
  g[j + 1] = -sn[j] * g[j];
  g[j] = cs[j] * g[j];
  const double res = fabs(g[j + 1]);
  c.k = j + 1;
  c.red = res / c.beta;
  const RunDev& rd = *d.run;
  const bool breakdown = !(hn > 1e-300) || !isfinite(hn);
  d.hc[(size_t)m * 2 * (K + 1) + 2 * K + 1] = breakdown ? 0.0 : 1.0 / hn;  // scale of V_{j+1}
  if (res <= rd.lin_tol * c.beta || j + 1 >= K || breakdown || !isfinite(res)) c.kact = 0;  // stops after this j
}

// one thread per member: y from the triangular system, the linear verdict
__global__ void k_ysolve(Dev d) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= d.nM) return;
  Ctl& c = d.ctl[m];
  c.upd = 0;
  if (!(c.phase == kIter && c.solve)) return;
  const int K = d.K, k = c.k;
  const double* H = d.H + (size_t)m * (K + 1) * K;
  const double* g = d.g + (size_t)m * (K + 1);
  double* y = d.y + (size_t)m * K;
  bool ok = k > 0;
  for (int i = k - 1; i >= 0; --i) {
    double s = g[i];
    for (int l = i + 1; l < k; ++l) s -= H[(size_t)l * (K + 1) + i] * y[l];
    const double dgl = H[(size_t)i * (K + 1) + i];
    y[i] = (dgl != 0.0) ? s / dgl : 0.0;
    if (!isfinite(y[i])) ok = false;
  }
  const RunDev& rd = *d.run;
  c.linear += k;
  // relaxed linear acceptance (rs/fim/newton/convergence.rs:385-389)
  const bool lconv = c.red <= rd.lin_tol;
  if (!ok || !(lconv || (isfinite(c.red) && c.red < 0.01))) {
    c.phase = kFail;
    c.newton++;
    return;
  }
  c.upd = 1;
}

// dx = sum_{i<k} y_i z_i, then the chop and update of the cells
__global__ __launch_bounds__(kMaxThreads) void k_update(Dev d, const Tile* __restrict__ tiles) {
  const Tile t = tiles[blockIdx.x];
  const Ctl& c = d.ctl[t.m];
  if (!c.upd) return;
  const Member& M = d.mem[t.m];
  const int N = d.Nt, K = d.K, k = c.k;
  const size_t zs = 2 * (size_t)N;
  const double* y = d.y + (size_t)t.m * K;
  for (int i = t.first + threadIdx.x; i < t.first + t.count; i += blockDim.x) {
    const int g = M.off + i;
    double dp = 0.0, dsw = 0.0;
    for (int q = 0; q < k; ++q) {
      dp += y[q] * double(d.Z[q * zs + g]);
      dsw += y[q] * double(d.Z[q * zs + N + g]);
    }
    d.dx[g] = dp;
    d.dx[N + g] = dsw;
    // per-cell chop (two-phase: dso = -dsw) and raw update with the pressure floor
    const double mx = fabs(dsw);
    const double a = mx > 0.2 ? 0.2 / mx : 1.0;
    dsw *= a;
    const double pi = d.p[g];
    const double cap = 0.3 * fabs(pi);
    if (fabs(dp) > cap) dp = dp > 0.0 ? cap : -cap;
    const double np_ = pi + dp;
    d.p[g] = np_ > 1e-6 ? np_ : 1e-6;
    d.sw[g] += dsw;
  }
}

// wells: dbhp = (-r_w - H . dx) / s from the unchopped step, BHP step capped at |bhp| and
// floored at 1 bar, control bounds; the iteration count
__global__ void k_well_update(Dev d) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= d.nM) return;
  Ctl& c = d.ctl[m];
  if (!c.upd) return;
  const Member& M = d.mem[m];
  const int N = d.Nt;
  for (int w = M.w0; w < M.w0 + M.nw; ++w) {
    const double s = d.w_s[w];
    if ((d.w_flags[w] & 18) != 18 || !schur_ok(s)) continue;
    double hd = 0.0;
    for (int e = d.w_pf[w]; e < d.w_pf[w + 1]; ++e) {
      const int g = M.off + d.pf_cell[e];
      hd += d.pf_h[2 * e] * d.dx[g] + d.pf_h[2 * e + 1] * d.dx[N + g];
    }
    double db = (-d.w_rw[w] - hd) / s;
    const double bb = d.w_bhp[w], cap = fabs(bb);
    db = fabs(db) > cap ? (db > 0 ? cap : -cap) : db;
    if (bb + db < 1.0) db = 1.0 - bb;
    double nb = bb + db;
    if (d.w_flags[w] & 1) {
      nb = fmax(nb, 1e-6);
    } else {
      nb = fmax(nb, fmax(fmin(d.w_limit[w], M.bhp_min), 1e-6));
    }
    d.w_bhp[w] = nb;
  }
  c.it++;
  c.newton++;
  c.upd = 0;
}

__global__ void k_fill_f(long n, float* x, float v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}
__global__ void k_fill_d(long n, double* x, double v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// ------------------------------------------------------------------ host
struct Ensemble {
  rsf_ensemble_options opt{};
  hipStream_t stream = nullptr;
  int nM = 0, Nt = 0, nW = 0, nP = 0, K = 20;
  std::vector<Member> mh;
  std::vector<std::vector<int>> perf_order;  // per member: device completion -> completion index
  std::vector<int> n_wells_m;
  DevBuf<Member> d_mem;
  DevBuf<Ctl> d_ctl;
  DevBuf<RunDev> d_run;
  DevBuf<int> d_alive;
  DevBuf<double> p, sw, p0, sw0, acc0, pv, pcs, Tx, Ty, Tz, R, J, pack;
  DevBuf<int> w_flags, w_pf, pf_cell;
  DevBuf<double> w_target, w_limit, w_bhpt, w_bhp, w_bhp0, w_s, w_rw, pf_wi, pf_head, pf_b, pf_h, pf_q;
  DevBuf<float> Mo, Dg, Di, Z, Wq, Pp, pdi, rp, xp, bw;
  DevBuf<unsigned char> nmask;
  DevBuf<double> cv;
  DevBuf<double> V, Wv, b, dx, H, g, cs, sn, y, hc, part;
  DevBuf<double> replay_dt;
  DevBuf<int> replay_off;
  DevBuf<Tile> t_nat, t_red, t_blk;
  int nt_nat = 0, nt_red = 0, nt_blk = 0, pstride = 0;
  int threads = 256, cpt[3] = {1, 1, 1};
  int bounds[2] = {16384, 40000};
  double probe_ms[3][3][3];
  int* alive_h = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  long ticks = 0;
  size_t graph_nodes = 0;

  ~Ensemble() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (alive_h) (void)hipHostFree(alive_h);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int size_class(int n) const { return n < bounds[0] ? 0 : (n < bounds[1] ? 1 : 2); }

  Dev dev() {
    Dev d{};
    d.Nt = Nt; d.K = K; d.nM = nM; d.nW = nW; d.pstride = pstride;
    d.mem = d_mem.p; d.ctl = d_ctl.p; d.run = d_run.p; d.alive = d_alive.p;
    d.p = p.p; d.sw = sw.p; d.p0 = p0.p; d.sw0 = sw0.p; d.acc0 = acc0.p;
    d.pv = pv.p; d.pcs = pcs.p; d.Tx = Tx.p; d.Ty = Ty.p; d.Tz = Tz.p;
    d.R = R.p; d.J = J.p; d.pack = pack.p;
    d.w_flags = w_flags.p; d.w_pf = w_pf.p; d.w_target = w_target.p; d.w_limit = w_limit.p; d.w_bhpt = w_bhpt.p;
    d.w_bhp = w_bhp.p; d.w_bhp0 = w_bhp0.p; d.w_s = w_s.p; d.w_rw = w_rw.p;
    d.pf_cell = pf_cell.p; d.pf_wi = pf_wi.p; d.pf_head = pf_head.p; d.pf_b = pf_b.p; d.pf_h = pf_h.p; d.pf_q = pf_q.p;
    d.Mo = Mo.p; d.Dg = Dg.p; d.Di = Di.p;
    d.Wq = Wq.p; d.Pp = Pp.p; d.pdi = pdi.p; d.rp = rp.p; d.xp = xp.p; d.bw = bw.p; d.cv = cv.p;
    d.nmask = nmask.p;
    d.V = V.p; d.Wv = Wv.p; d.b = b.p; d.dx = dx.p; d.Z = Z.p;
    d.H = H.p; d.g = g.p; d.cs = cs.p; d.sn = sn.p; d.y = y.p; d.hc = hc.p; d.part = part.p;
    return d;
  }

  // tile maps: natural cells, red slots, black slots; cells per workgroup =
  // threads x cells-per-thread of the member's size class. `only` >= 0 keeps one class.
  void build_tiles(int thr, const int c[3], int only, std::vector<Tile>& nat, std::vector<Tile>& red,
                   std::vector<Tile>& blk, std::vector<Member>& mm) const {
    nat.clear(); red.clear(); blk.clear();
    mm = mh;
    for (int m = 0; m < nM; ++m) {
      const int cls = size_class(mh[m].n);
      mm[m].t0 = (int)nat.size();
      mm[m].nt = 0;
      if (only >= 0 && cls != only) continue;
      const int span = thr * c[cls];
      auto cut = [&](std::vector<Tile>& out, int first, int count) {
        for (int s = 0; s < count; s += span) out.push_back({m, first + s, std::min(span, count - s)});
      };
      cut(nat, 0, mh[m].n);
      mm[m].nt = (int)nat.size() - mm[m].t0;
      cut(red, 0, mh[m].nred);
      cut(blk, 0, mh[m].n - mh[m].nred);
    }
  }

  void install_tiles(int thr, const int c[3]) {
    std::vector<Tile> a, r, k;
    std::vector<Member> mm;
    build_tiles(thr, c, -1, a, r, k, mm);
    t_nat.upload(a, stream);
    t_red.upload(r, stream);
    t_blk.upload(k, stream);
    nt_nat = (int)a.size(); nt_red = (int)r.size(); nt_blk = (int)k.size();
    mh = mm;
    d_mem.upload(mh, stream);
    threads = thr;
    for (int i = 0; i < 3; ++i) cpt[i] = c[i];
  }

  int mblocks() const { return (nM + 255) / 256; }

  // one Newton iteration of every running member
  void tick_kernels() {
    const Dev d = dev();
    const dim3 T(threads), TM(256);
    hipLaunchKernelGGL(k_control, dim3(1), dim3(1024), 0, stream, d);
    hipLaunchKernelGGL(k_cellstate, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p);
    hipLaunchKernelGGL(k_cells, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p);
    hipLaunchKernelGGL(k_faces, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p);
    hipLaunchKernelGGL(k_wells, dim3(mblocks()), TM, 0, stream, d);
    hipLaunchKernelGGL(k_conv_partial, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p);
    hipLaunchKernelGGL(k_conv_finish, dim3(mblocks()), TM, 0, stream, d);
    hipLaunchKernelGGL(k_rhs, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p);
    hipLaunchKernelGGL(k_pcopy, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p);
    hipLaunchKernelGGL(k_well_reduce, dim3(mblocks()), TM, 0, stream, d);
    hipLaunchKernelGGL(k_factor_red, dim3(nt_red), T, 0, stream, d, (const Tile*)t_red.p);
    hipLaunchKernelGGL(k_factor_black, dim3(nt_blk), T, 0, stream, d, (const Tile*)t_blk.p);
    hipLaunchKernelGGL(k_cpr_setup, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p);
    hipLaunchKernelGGL(k_pfactor, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, 0);
    hipLaunchKernelGGL(k_pfactor, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, 1);
    hipLaunchKernelGGL(k_norm_partial, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, (const double*)b.p, 1);
    hipLaunchKernelGGL(k_beta, dim3(mblocks()), TM, 0, stream, d);
    hipLaunchKernelGGL(k_scale, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, (const double*)b.p, V.p, 1);
    const size_t vs = 2 * (size_t)Nt;
    for (int j = 0; j < K; ++j) {
      const double* vj = V.p + j * vs;
      float* zj = Z.p + j * vs;
      hipLaunchKernelGGL(k_cpr_restrict, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, vj);
      hipLaunchKernelGGL(k_pbicg, dim3(nM), dim3(kPB), 0, stream, d);
      hipLaunchKernelGGL(k_cpr_correct, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, vj);
      hipLaunchKernelGGL(k_cpr_well, dim3(mblocks()), TM, 0, stream, d);
      hipLaunchKernelGGL(k_ilu_red1, dim3(nt_red), T, 0, stream, d, (const Tile*)t_red.p, (const double*)cv.p, zj);
      hipLaunchKernelGGL(k_ilu_black, dim3(nt_blk), T, 0, stream, d, (const Tile*)t_blk.p, (const double*)cv.p, zj);
      hipLaunchKernelGGL(k_ilu_red2, dim3(nt_red), T, 0, stream, d, (const Tile*)t_red.p, zj);
      hipLaunchKernelGGL(k_cpr_combine, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, zj);
      hipLaunchKernelGGL(k_spmv, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, (const float*)zj, Wv.p);
      hipLaunchKernelGGL(k_well_apply, dim3(mblocks()), TM, 0, stream, d, (const float*)zj, Wv.p);
      hipLaunchKernelGGL(k_mdot, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, j);
      hipLaunchKernelGGL(k_mdot_finish, dim3(nM), dim3(256), 0, stream, d, j, 0);
      hipLaunchKernelGGL(k_orth, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, j, 0);
      hipLaunchKernelGGL(k_mdot, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, j);
      hipLaunchKernelGGL(k_mdot_finish, dim3(nM), dim3(256), 0, stream, d, j, 1);
      hipLaunchKernelGGL(k_orth, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, j, 1);
      hipLaunchKernelGGL(k_arnoldi, dim3(mblocks()), TM, 0, stream, d, j);
      if (j + 1 < K)
        hipLaunchKernelGGL(k_scale, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p, (const double*)Wv.p,
                           V.p + (j + 1) * vs, 0);
    }
    hipLaunchKernelGGL(k_ysolve, dim3(mblocks()), TM, 0, stream, d);
    hipLaunchKernelGGL(k_update, dim3(nt_nat), T, 0, stream, d, (const Tile*)t_nat.p);
    hipLaunchKernelGGL(k_well_update, dim3(mblocks()), TM, 0, stream, d);
  }

  void capture() {
    if (exec) { RSF_CHECK(hipGraphExecDestroy(exec)); exec = nullptr; }
    if (graph) { RSF_CHECK(hipGraphDestroy(graph)); graph = nullptr; }
    RSF_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    tick_kernels();
    RSF_CHECK(hipStreamEndCapture(stream, &graph));
    RSF_CHECK(hipGraphGetNodes(graph, nullptr, &graph_nodes));
    RSF_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }

  // the stencil kernels of a tick on one size class under a tile choice, timed
  double probe(int thr, const int c[3], int cls) {
    std::vector<Tile> a, r, k;
    std::vector<Member> mm;
    build_tiles(thr, c, cls, a, r, k, mm);
    if (a.empty()) return -1.0;
    DevBuf<Tile> ta, tr, tk;
    DevBuf<Member> dm;
    ta.upload(a, stream);
    tr.upload(r, stream);
    tk.upload(k, stream);
    dm.upload(mm, stream);
    Dev d = dev();
    d.mem = dm.p;
    const dim3 T(thr);
    const int jp = std::min(3, K);  // projections against V_0..V_jp (V holds K + 1 vectors)
    hipEvent_t e0, e1;
    RSF_CHECK(hipEventCreate(&e0));
    RSF_CHECK(hipEventCreate(&e1));
    auto once = [&]() {
      hipLaunchKernelGGL(k_cells, dim3(a.size()), T, 0, stream, d, (const Tile*)ta.p);
      hipLaunchKernelGGL(k_faces, dim3(a.size()), T, 0, stream, d, (const Tile*)ta.p);
      hipLaunchKernelGGL(k_ilu_red1, dim3(r.size()), T, 0, stream, d, (const Tile*)tr.p, (const double*)V.p, Z.p);
      hipLaunchKernelGGL(k_ilu_black, dim3(k.size()), T, 0, stream, d, (const Tile*)tk.p, (const double*)V.p, Z.p);
      hipLaunchKernelGGL(k_ilu_red2, dim3(r.size()), T, 0, stream, d, (const Tile*)tr.p, Z.p);
      hipLaunchKernelGGL(k_spmv, dim3(a.size()), T, 0, stream, d, (const Tile*)ta.p, (const float*)Z.p, Wv.p);
      hipLaunchKernelGGL(k_mdot, dim3(a.size()), T, 0, stream, d, (const Tile*)ta.p, jp);
      hipLaunchKernelGGL(k_orth, dim3(a.size()), T, 0, stream, d, (const Tile*)ta.p, jp, 1);
    };
    once();
    RSF_CHECK(hipEventRecord(e0, stream));
    for (int rep = 0; rep < 3; ++rep) once();
    RSF_CHECK(hipEventRecord(e1, stream));
    RSF_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    RSF_CHECK(hipEventElapsedTime(&ms, e0, e1));
    RSF_CHECK(hipEventDestroy(e0));
    RSF_CHECK(hipEventDestroy(e1));
    return ms / 3.0;
  }

  // per threads-per-workgroup choice the best cells per thread of each size class;
  // the threads choice with the smallest sum wins
  void autotune() {
    for (auto& a : probe_ms)
      for (auto& b : a)
        for (auto& c : b) c = -1.0;
    // every member running with a live Krylov state, so the probed kernels do their work
    std::vector<Ctl> c0(nM);
    d_ctl.download(c0.data(), nM, stream);
    RSF_CHECK(hipStreamSynchronize(stream));
    std::vector<Ctl> cp = c0;
    for (auto& c : cp) {
      c.phase = kIter;
      c.solve = 1;
      c.kact = 1;
      c.dt = 1.0;
    }
    d_ctl.upload(cp, stream);
    const int thr_c[3] = {128, 256, 512}, cpt_c[3] = {1, 2, 4};
    double best_total = INFINITY;
    int best_thr = 256, best_c[3] = {1, 1, 1};
    for (int ti = 0; ti < 3; ++ti) {
      double total = 0.0;
      int bc[3] = {1, 1, 1};
      for (int cls = 0; cls < 3; ++cls) {
        double bm = INFINITY;
        for (int ci = 0; ci < 3; ++ci) {
          int c[3] = {1, 1, 1};
          c[cls] = cpt_c[ci];
          const double ms = probe(thr_c[ti], c, cls);
          probe_ms[ti][cls][ci] = ms;
          if (ms >= 0.0 && ms < bm) { bm = ms; bc[cls] = cpt_c[ci]; }
        }
        if (std::isfinite(bm)) total += bm;
      }
      if (total < best_total) {
        best_total = total;
        best_thr = thr_c[ti];
        for (int k = 0; k < 3; ++k) best_c[k] = bc[k];
      }
    }
    d_ctl.upload(c0, stream);
    install_tiles(best_thr, best_c);
  }

  void zero_fill() {
    auto fd = [&](DevBuf<double>& x) {
      if (x.n) hipLaunchKernelGGL(k_fill_d, dim3((x.n + 255) / 256), dim3(256), 0, stream, (long)x.n, x.p, 0.0);
    };
    auto ff = [&](DevBuf<float>& x) {
      if (x.n) hipLaunchKernelGGL(k_fill_f, dim3((x.n + 255) / 256), dim3(256), 0, stream, (long)x.n, x.p, 0.0f);
    };
    fd(R); fd(J); fd(pack); fd(V); fd(Wv); fd(b); fd(dx); fd(H); fd(g); fd(cs); fd(sn); fd(y); fd(hc); fd(part);
    fd(pf_b); fd(pf_h); fd(pf_q); fd(w_s); fd(w_rw);
    ff(Mo); ff(Dg); ff(Di); ff(Z); ff(Wq); ff(Pp); ff(pdi); ff(rp); ff(xp); ff(bw);
    fd(cv);
    RSF_CHECK(hipGetLastError());
  }

  void create(rsf_sim* const* members, int count, const rsf_ensemble_options& o);
  int run(int mode, double report_dt, int n_reports, const double* dts, const int* offs);
};

// ------------------------------------------------------------------ set-up
void Ensemble::create(rsf_sim* const* members, int count, const rsf_ensemble_options& o) {
  opt = o;
  K = o.krylov_dimension;
  nM = count;
  RSF_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  RSF_CHECK(hipHostMalloc((void**)&alive_h, sizeof(int)));
  long total = 0;
  std::vector<int> wflags, wpf{0}, pcell;
  std::vector<double> wtarget, wlimit, wbhpt, wbhp, pwi, phead;
  mh.resize(nM);
  perf_order.resize(nM);
  n_wells_m.resize(nM);
  for (int m = 0; m < nM; ++m) {
    Sim& s = *members[m]->s;
    if (s.three_phase) throw std::invalid_argument("ensemble member " + std::to_string(m) + ": three-phase mode is not supported");
    if (s.gravity) throw std::invalid_argument("ensemble member " + std::to_string(m) + ": gravity is not supported");
    if (s.has_table) throw std::invalid_argument("ensemble member " + std::to_string(m) + ": a PVT table is not supported");
    total += s.n;
    if (total > (1L << 30) / 32)
      throw std::invalid_argument("ensemble: too many cells in total");
  }
  Nt = (int)total;
  // packed model arrays, filled member by member from the members' device copies
  p.alloc(Nt); sw.alloc(Nt); p0.alloc(Nt); sw0.alloc(Nt); acc0.alloc(2 * (size_t)Nt);
  pv.alloc(Nt); pcs.alloc(Nt); Tx.alloc(Nt); Ty.alloc(Nt); Tz.alloc(Nt);
  int off = 0;
  for (int m = 0; m < nM; ++m) {
    Sim& s = *members[m]->s;
    s.ensure_device();
    s.refresh_head_offsets();
    std::vector<double> bhp0, q0;
    s.init_substep(bhp0, q0);  // the member's own initial BHPs (rs/fim/wells.rs:431-504)
    const size_t bytes = s.n * sizeof(double);
    RSF_CHECK(hipMemcpyAsync(p.p + off, s.s_p.p, bytes, hipMemcpyDeviceToDevice, s.stream));
    RSF_CHECK(hipMemcpyAsync(sw.p + off, s.s_sw.p, bytes, hipMemcpyDeviceToDevice, s.stream));
    RSF_CHECK(hipMemcpyAsync(pv.p + off, s.d_pv.p, bytes, hipMemcpyDeviceToDevice, s.stream));
    RSF_CHECK(hipMemcpyAsync(pcs.p + off, s.d_pcs.p, bytes, hipMemcpyDeviceToDevice, s.stream));
    RSF_CHECK(hipMemcpyAsync(Tx.p + off, s.d_Tx.p, bytes, hipMemcpyDeviceToDevice, s.stream));
    RSF_CHECK(hipMemcpyAsync(Ty.p + off, s.d_Ty.p, bytes, hipMemcpyDeviceToDevice, s.stream));
    RSF_CHECK(hipMemcpyAsync(Tz.p + off, s.d_Tz.p, bytes, hipMemcpyDeviceToDevice, s.stream));
    s.sync();
    Member& M = mh[m];
    M.nx = s.nx; M.ny = s.ny; M.nz = s.nz; M.n = s.n; M.off = off;
    // red cells: one of each aligned pair, plus the last cell of an odd count when red
    M.nred = s.n / 2;
    if (s.n & 1) {
      const int i = s.n - 1, z = i / (s.nx * s.ny), r = i % (s.nx * s.ny), y = r / s.nx, x = r % s.nx;
      if (((x + y + z) & 1) == 0) M.nred++;
    }
    M.bhp_min = s.bhp_min;
    M.P = s.params(false);
    M.P.sat = nullptr; M.P.br_rs = nullptr; M.P.br_off = nullptr; M.P.br_len = nullptr; M.P.br_rows = nullptr;
    M.P.rs_base = nullptr; M.P.swof = nullptr; M.P.sgof = nullptr;
    M.P.n_swof = M.P.n_sgof = 0;
    M.w0 = (int)wflags.size();
    const int W = (int)s.topo.size();
    M.nw = W;
    n_wells_m[m] = W;
    for (int w = 0; w < W; ++w) {
      ControlState c = s.control(w);
      const bool fb = c.enabled && c.rate && c.has_target;
      wflags.push_back((s.topo[w].injector ? 1 : 0) | (fb ? 2 : 0) | (c.surface ? 4 : 0) | (c.enabled ? 16 : 0));
      wtarget.push_back(c.target);
      wlimit.push_back(c.limit);
      wbhpt.push_back(c.bhp_target);
      wbhp.push_back(bhp0[w]);
      for (int e : s.topo[w].perfs) {
        const auto& cp = s.wells[e];
        perf_order[m].push_back(e);
        pcell.push_back(s.idx(cp.i, cp.j, cp.k));
        double v = 0.0;
        s.well_index(e, v);
        pwi.push_back(v);
        phead.push_back(cp.head);
      }
      wpf.push_back((int)pcell.size());
    }
    off += s.n;
  }
  nW = (int)wflags.size();
  nP = (int)pcell.size();
  std::vector<int> zi{0};
  std::vector<double> zd{0.0};
  w_flags.upload(wflags.empty() ? zi : wflags, stream);
  w_pf.upload(wpf, stream);
  w_target.upload(wtarget.empty() ? zd : wtarget, stream);
  w_limit.upload(wlimit.empty() ? zd : wlimit, stream);
  w_bhpt.upload(wbhpt.empty() ? zd : wbhpt, stream);
  w_bhp.upload(wbhp.empty() ? zd : wbhp, stream);
  w_bhp0.upload(wbhp.empty() ? zd : wbhp, stream);
  w_s.alloc(std::max(nW, 1));
  w_rw.alloc(std::max(nW, 1));
  pf_cell.upload(pcell.empty() ? zi : pcell, stream);
  pf_wi.upload(pwi.empty() ? zd : pwi, stream);
  pf_head.upload(phead.empty() ? zd : phead, stream);
  pf_b.alloc(2 * std::max((size_t)nP, (size_t)1));
  pf_h.alloc(2 * std::max((size_t)nP, (size_t)1));
  pf_q.alloc(std::max((size_t)nP, (size_t)1));
  RSF_CHECK(hipMemcpyAsync(p0.p, p.p, Nt * sizeof(double), hipMemcpyDeviceToDevice, stream));
  RSF_CHECK(hipMemcpyAsync(sw0.p, sw.p, Nt * sizeof(double), hipMemcpyDeviceToDevice, stream));
  const size_t N = Nt;
  R.alloc(2 * N); J.alloc(28 * N); pack.alloc(3 * kPackQ * N);
  Mo.alloc(24 * N); Dg.alloc(4 * N); Di.alloc(4 * N);
  Wq.alloc(2 * N); Pp.alloc(7 * N); pdi.alloc(N); rp.alloc(N); xp.alloc(N); bw.alloc(9 * N); cv.alloc(2 * N);
  {
    std::vector<unsigned char> mk(N);
    for (const Member& M : mh)
      for (int k = 0; k < M.nz; ++k)
        for (int j = 0; j < M.ny; ++j)
          for (int i = 0; i < M.nx; ++i) {
            unsigned char b = 0;
            if (i > 0) b |= 1;
            if (i + 1 < M.nx) b |= 2;
            if (j > 0) b |= 4;
            if (j + 1 < M.ny) b |= 8;
            if (k > 0) b |= 16;
            if (k + 1 < M.nz) b |= 32;
            if ((i + j + k) & 1) b |= 64;
            mk[(size_t)M.off + i + M.nx * (j + (size_t)M.ny * k)] = b;
          }
    nmask.upload(mk, stream);
  }
  V.alloc((size_t)(K + 1) * 2 * N); Z.alloc((size_t)K * 2 * N);
  Wv.alloc(2 * N); b.alloc(2 * N); dx.alloc(2 * N);
  H.alloc((size_t)nM * (K + 1) * K); g.alloc((size_t)nM * (K + 1)); cs.alloc((size_t)nM * K); sn.alloc((size_t)nM * K);
  y.alloc((size_t)nM * K); hc.alloc((size_t)nM * 2 * (K + 1));
  // tile partials: the most tiles any choice makes (1 cell-per-thread at 128 threads)
  pstride = std::max(K + 1, kConvVals);
  long maxt = 0;
  for (auto& M : mh) maxt += (M.n + 127) / 128;
  part.alloc((size_t)maxt * pstride);
  d_alive.alloc(1);
  std::vector<Ctl> c0(nM);
  for (auto& c : c0) {
    std::memset(&c, 0, sizeof(Ctl));
    c.phase = kFinished;
  }
  d_ctl.upload(c0, stream);
  RunDev rd{};
  d_run.upload(&rd, 1, stream);
  zero_fill();
  const int c1[3] = {o.cells_per_thread[0], o.cells_per_thread[1], o.cells_per_thread[2]};
  install_tiles(o.threads_per_block, c1);
  if (o.autotune) autotune();
  zero_fill();
  RSF_CHECK(hipStreamSynchronize(stream));
  if (o.use_graph) capture();
}

int Ensemble::run(int mode, double report_dt, int n_reports, const double* dts, const int* offs) {
  std::vector<Ctl> c(nM);
  d_ctl.download(c.data(), nM, stream);
  RSF_CHECK(hipStreamSynchronize(stream));
  RunDev rd{};
  rd.mode = mode;
  rd.n_reports = n_reports;
  rd.max_retries = opt.max_retries;
  rd.max_newton = opt.max_newton_iterations;
  rd.min_newton = opt.min_newton_iteration_index;
  rd.report_dt = report_dt;
  rd.tol = opt.newton_residual_tol;
  rd.mb_tol = opt.newton_mb_tol;
  rd.lin_tol = opt.linear_rel_tol;
  rd.p_max_it = opt.pressure_max_iterations;
  rd.p_tol = opt.pressure_rel_tol;
  if (mode == 1) {
    std::vector<int> o(offs, offs + nM + 1);
    replay_dt.upload(dts, std::max(o[nM], 1), stream);
    replay_off.upload(o, stream);
    rd.replay_dt = replay_dt.p;
    rd.replay_off = replay_off.p;
  }
  d_run.upload(&rd, 1, stream);
  for (int m = 0; m < nM; ++m) {
    Ctl& x = c[m];
    if (x.phase == kDead) continue;
    x.phase = kStart;
    x.streak = 0;
    x.report = 0;
    x.replay_k = 0;
    x.t_rem = report_dt;
    if (mode == 1) {
      if (offs[m + 1] <= offs[m]) { x.phase = kFinished; continue; }
      x.dt = dts[offs[m]];
    } else {
      if (n_reports <= 0) { x.phase = kFinished; continue; }
      x.dt_next = report_dt;
      x.dt = report_dt;
    }
  }
  d_ctl.upload(c, stream);
  // every tick is one Newton iteration; a member needs at most max_newton + 1 ticks per
  // attempt, so this bound is never reached by a run that makes progress
  const long cap = 1000000;
  for (long t = 0; t < cap; ++t) {
    if (exec) {
      RSF_CHECK(hipGraphLaunch(exec, stream));
    } else {
      tick_kernels();
      RSF_CHECK(hipGetLastError());
    }
    ticks++;
    RSF_CHECK(hipMemcpyAsync(alive_h, d_alive.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    RSF_CHECK(hipStreamSynchronize(stream));
    if (*alive_h == 0) break;
  }
  d_ctl.download(c.data(), nM, stream);
  RSF_CHECK(hipStreamSynchronize(stream));
  for (auto& x : c)
    if (x.phase == kDead) return RSF_ERR_NOT_CONVERGED;
  return RSF_OK;
}

}  // namespace ens
}  // namespace rsf

// ======================================================================= C ABI
struct rsf_ensemble {
  rsf::ens::Ensemble* e;
};

namespace {
std::mutex g_ens_mu;
std::string g_ens_err;
void set_err(const std::string& s) {
  std::lock_guard<std::mutex> l(g_ens_mu);
  g_ens_err = s;
}

template <class F>
int eguard(rsf_ensemble* h, F&& f) {
  if (!h || !h->e) {
    set_err("invalid ensemble handle");
    return RSF_ERR_INVALID;
  }
  try {
    return f(*h->e);
  } catch (const rsf::HipError& x) {
    set_err(x.what());
    return RSF_ERR_HIP;
  } catch (const std::exception& x) {
    set_err(x.what());
    return RSF_ERR_INVALID;
  }
}
}  // namespace

extern "C" {

void rsf_ensemble_default_options(rsf_ensemble_options* o) {
  if (!o) return;
  o->max_newton_iterations = 20;
  o->min_newton_iteration_index = 1;
  o->newton_residual_tol = 1e-6;
  o->newton_mb_tol = 1e-7;
  o->linear_rel_tol = 5e-3;
  o->krylov_dimension = 20;
  o->pressure_max_iterations = 50;  // the reference's CPR (rs/fim/linear/gmres_block_jacobi.rs:25-26)
  o->pressure_rel_tol = 1e-6;
  o->max_retries = 12;
  o->autotune = 1;
  o->use_graph = 1;
  o->threads_per_block = 256;
  o->cells_per_thread[0] = o->cells_per_thread[1] = o->cells_per_thread[2] = 1;
}

int rsf_ensemble_create(rsf_sim* const* members, int count, const rsf_ensemble_options* opt, rsf_ensemble** out) {
  if (!out || !members || count <= 0) {
    set_err("rsf_ensemble_create: members, a positive count and an output pointer are required");
    return RSF_ERR_INVALID;
  }
  *out = nullptr;
  rsf_ensemble_options o;
  rsf_ensemble_default_options(&o);
  if (opt) o = *opt;
  auto bad = [](const char* m) {
    set_err(m);
    return RSF_ERR_INVALID;
  };
  if (o.krylov_dimension < 1 || o.krylov_dimension > rsf::ens::kMaxKrylov) return bad("krylov_dimension must be in 1..200");
  if (o.max_newton_iterations < 1 || o.min_newton_iteration_index < 0) return bad("invalid Newton iteration limits");
  if (o.pressure_max_iterations < 0 || !(o.pressure_rel_tol > 0)) return bad("invalid pressure solve limits");
  if (!(o.newton_residual_tol > 0) || !(o.newton_mb_tol > 0) || !(o.linear_rel_tol > 0))
    return bad("tolerances must be positive");
  if (o.threads_per_block != 128 && o.threads_per_block != 256 && o.threads_per_block != 512)
    return bad("threads_per_block must be 128, 256 or 512");
  for (int k = 0; k < 3; ++k)
    if (o.cells_per_thread[k] != 1 && o.cells_per_thread[k] != 2 && o.cells_per_thread[k] != 4)
      return bad("cells_per_thread must be 1, 2 or 4");
  for (int m = 0; m < count; ++m)
    if (!members[m] || !members[m]->s) return bad("rsf_ensemble_create: null member handle");
  auto* e = new rsf::ens::Ensemble();
  try {
    e->create(members, count, o);
  } catch (const rsf::HipError& x) {
    delete e;
    set_err(x.what());
    return RSF_ERR_HIP;
  } catch (const std::exception& x) {
    delete e;
    set_err(x.what());
    return RSF_ERR_INVALID;
  }
  *out = new rsf_ensemble{e};
  return RSF_OK;
}

void rsf_ensemble_destroy(rsf_ensemble* h) {
  if (!h) return;
  delete h->e;
  delete h;
}

int rsf_ensemble_run(rsf_ensemble* h, double report_dt_days, int n_reports) {
  return eguard(h, [&](rsf::ens::Ensemble& e) {
    if (!(report_dt_days > 0) || !std::isfinite(report_dt_days) || n_reports < 0) {
      set_err("rsf_ensemble_run: report_dt_days must be positive and n_reports >= 0");
      return RSF_ERR_INVALID;
    }
    return e.run(0, report_dt_days, n_reports, nullptr, nullptr);
  });
}

int rsf_ensemble_replay(rsf_ensemble* h, const double* dts, const int* offsets) {
  return eguard(h, [&](rsf::ens::Ensemble& e) {
    if (!offsets || offsets[0] != 0) {
      set_err("rsf_ensemble_replay: offsets must start at 0");
      return RSF_ERR_INVALID;
    }
    for (int m = 0; m < e.nM; ++m)
      if (offsets[m + 1] < offsets[m]) {
        set_err("rsf_ensemble_replay: offsets must not decrease");
        return RSF_ERR_INVALID;
      }
    if (offsets[e.nM] > 0 && !dts) {
      set_err("rsf_ensemble_replay: dts required");
      return RSF_ERR_INVALID;
    }
    for (int k = 0; k < offsets[e.nM]; ++k)
      if (!(dts[k] > 0) || !std::isfinite(dts[k])) {
        set_err("rsf_ensemble_replay: every substep must be positive");
        return RSF_ERR_INVALID;
      }
    return e.run(1, 1.0, 0, dts, offsets);
  });
}

int rsf_ensemble_member_state(rsf_ensemble* h, int m, double* pressure, double* sw, int n) {
  return eguard(h, [&](rsf::ens::Ensemble& e) {
    if (m < 0 || m >= e.nM || n != e.mh[m].n || !pressure || !sw) {
      set_err("rsf_ensemble_member_state: invalid member or size");
      return RSF_ERR_INVALID;
    }
    const int off = e.mh[m].off;
    RSF_CHECK(hipMemcpyAsync(pressure, e.p.p + off, n * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    RSF_CHECK(hipMemcpyAsync(sw, e.sw.p + off, n * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    RSF_CHECK(hipStreamSynchronize(e.stream));
    return RSF_OK;
  });
}

int rsf_ensemble_member_wells(rsf_ensemble* h, int m, double* bhp, int n_wells, double* q, int n_completions) {
  return eguard(h, [&](rsf::ens::Ensemble& e) {
    if (m < 0 || m >= e.nM || n_wells != e.mh[m].nw || n_completions != (int)e.perf_order[m].size()) {
      set_err("rsf_ensemble_member_wells: invalid member or sizes");
      return RSF_ERR_INVALID;
    }
    const rsf::ens::Member& M = e.mh[m];
    std::vector<int> flags(std::max(M.nw, 1)), pf(M.nw + 1);
    std::vector<double> b(std::max(M.nw, 1)), bt(std::max(M.nw, 1));
    if (M.nw) {
      RSF_CHECK(hipMemcpyAsync(flags.data(), e.w_flags.p + M.w0, M.nw * sizeof(int), hipMemcpyDeviceToHost, e.stream));
      RSF_CHECK(hipMemcpyAsync(pf.data(), e.w_pf.p + M.w0, (M.nw + 1) * sizeof(int), hipMemcpyDeviceToHost, e.stream));
      RSF_CHECK(hipMemcpyAsync(b.data(), e.w_bhp.p + M.w0, M.nw * sizeof(double), hipMemcpyDeviceToHost, e.stream));
      RSF_CHECK(hipMemcpyAsync(bt.data(), e.w_bhpt.p + M.w0, M.nw * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    }
    std::vector<double> qd(std::max(n_completions, 1));
    if (n_completions)
      RSF_CHECK(hipMemcpyAsync(qd.data(), e.pf_q.p + pf[0], n_completions * sizeof(double), hipMemcpyDeviceToHost,
                               e.stream));
    RSF_CHECK(hipStreamSynchronize(e.stream));
    for (int w = 0; w < M.nw; ++w)
      if (bhp) bhp[w] = (flags[w] & 2) ? b[w] : bt[w];
    if (q)
      for (int t = 0; t < n_completions; ++t) q[e.perf_order[m][t]] = qd[t];
    return RSF_OK;
  });
}

int rsf_ensemble_member_stats(rsf_ensemble* h, int m, rsf_ensemble_member_report* out) {
  return eguard(h, [&](rsf::ens::Ensemble& e) {
    if (m < 0 || m >= e.nM || !out) {
      set_err("rsf_ensemble_member_stats: invalid member");
      return RSF_ERR_INVALID;
    }
    rsf::ens::Ctl c;
    RSF_CHECK(hipMemcpyAsync(&c, e.d_ctl.p + m, sizeof(c), hipMemcpyDeviceToHost, e.stream));
    RSF_CHECK(hipStreamSynchronize(e.stream));
    out->status = c.phase == rsf::ens::kFinished ? 1 : (c.phase == rsf::ens::kDead ? 2 : 0);
    out->substeps = c.substeps;
    out->retries = c.retries;
    out->newton_iterations = c.newton;
    out->linear_iterations = c.linear;
    out->n_cells = e.mh[m].n;
    out->time_days = c.time;
    return RSF_OK;
  });
}

int rsf_ensemble_tuning_info(rsf_ensemble* h, rsf_ensemble_tuning* out) {
  return eguard(h, [&](rsf::ens::Ensemble& e) {
    if (!out) return RSF_ERR_INVALID;
    out->threads_per_block = e.threads;
    for (int k = 0; k < 3; ++k) out->cells_per_thread[k] = e.cpt[k];
    out->class_bounds[0] = e.bounds[0];
    out->class_bounds[1] = e.bounds[1];
    for (int k = 0; k < 3; ++k) out->class_members[k] = 0;
    for (auto& M : e.mh) out->class_members[e.size_class(M.n)]++;
    std::memcpy(out->probe_ms, e.probe_ms, sizeof(e.probe_ms));
    out->ticks = e.ticks;
    out->graph_nodes = (long)e.graph_nodes;
    out->n_members = e.nM;
    out->n_cells = e.Nt;
    return RSF_OK;
  });
}

const char* rsf_ensemble_last_error(void) {
  std::lock_guard<std::mutex> l(g_ens_mu);
  return g_ens_err.c_str();
}

}  // extern "C"
