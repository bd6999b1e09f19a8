// HIP kernels of the FIM Newton step on gfx950.
//
// Layout (N cells, natural index k*nx*ny + j*nx + i, rs/grid.rs:252-254):
//  * cell fields and 3N vectors are structure-of-arrays: v[c*N + i];
//  * the Jacobian of the well-reduced system is a 7-slot block-diagonal
//    ("stencil BSR") store A[(s*9 + e)*N + i] of 3x3 blocks, slot s in
//    {0 diag, 1 x-, 2 x+, 3 y-, 4 y+, 5 z-, 6 z+}, entry e = row*3 + col.
//    Neighbour columns are implicit, so lanes of a wavefront read
//    consecutive cells of every (slot, entry) plane: fully coalesced and no
//    index traffic. Fill between non-adjacent perforated cells of
//    rate-controlled wells lives in a small CSR side store ("extra blocks").
#pragma once
#include <hip/hip_runtime.h>

#include <climits>

#include "rsf_physics.hpp"

namespace rsf {

constexpr int kBlock = 256;
constexpr int kPack = 36;  // FaceCell<Dual<3>>: 9 quantities x (value + 3 partials)

struct Grid {
  int nx, ny, nz, n;
  // Preconditioner storage colour-compacted (red-black ordering, nx even): the pressure
  // matrix and its ILU diagonal, the preconditioner's copy of the block matrix and the
  // block-ILU D^-1. The Jacobian itself, the residual and every vector stay natural.
  int pcmp;
};

__device__ __forceinline__ void cell_ijk(const Grid& g, int i, int& x, int& y, int& z) {
  int nxy = g.nx * g.ny;
  z = i / nxy;
  int r = i - z * nxy;
  y = r / g.nx;
  x = r - y * g.nx;
}

// neighbour of cell i through slot s (1..6); -1 when outside the grid
__device__ __forceinline__ int slot_nbr(const Grid& g, int i, int s, int x, int y, int z) {
  switch (s) {
    case 1: return x > 0 ? i - 1 : -1;
    case 2: return x + 1 < g.nx ? i + 1 : -1;
    case 3: return y > 0 ? i - g.nx : -1;
    case 4: return y + 1 < g.ny ? i + g.nx : -1;
    case 5: return z > 0 ? i - g.nx * g.ny : -1;
    default: return z + 1 < g.nz ? i + g.nx * g.ny : -1;
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Renumbering them so that XCD k runs one contiguous range of tiles lets its L2
// hold the stencil neighbours (the z neighbours are nx*ny cells away) of that range.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  constexpr int kXcd = 8;
  const int q = nb / kXcd, r = nb % kXcd, k = b % kXcd, j = b / kXcd;
  return k * q + (k < r ? k : r) + j;
}
#define RSF_CELL_INDEX() (xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x)

// Storage index of cell i in the preconditioner planes (pressure matrix, pressure ILU
// diagonal, block-matrix copy, block D^-1): natural, or colour-compacted when g.pcmp is
// set (red cells first, then black, each in natural order; with nx even the k-th cell
// of a colour is i >> 1, so a wavefront of one colour reads contiguous memory).
__device__ __forceinline__ int pslot(const Grid& g, int i, int x, int y, int z) {
  return g.pcmp ? (i >> 1) + (((x + y + z) & 1) ? (g.n >> 1) : 0) : i;
}
// the same for a stencil neighbour j of cell i (opposite colour)
__device__ __forceinline__ int pslot_nbr(const Grid& g, int j, int x, int y, int z) {
  return g.pcmp ? (j >> 1) + (((x + y + z) & 1) ? 0 : (g.n >> 1)) : j;
}

using D3 = Dual<3>;

// --------------------------------------------------------------- assembly
// Previous-iterate inventory (a constant of the substep): acc0[c*N + i].
__global__ __launch_bounds__(kBlock) void k_inventory(PhysParams P, int n, const double* __restrict__ pv_ref, const double* __restrict__ p,
                            const double* __restrict__ sw, const double* __restrict__ hc, const int* __restrict__ reg,
                            double* __restrict__ acc0) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double inv3[3];
  inventory<double>(P, pv_ref[i], reg[i], p[i], sw[i], hc[i], inv3, rs_cap_at(P, i));
  acc0[i] = inv3[0];
  acc0[n + i] = inv3[1];
  acc0[2 * n + i] = inv3[2];
}

// Per cell: accumulation residual and its diagonal block (rs/fim/properties.rs:171-255),
// plus the face pack consumed by k_face_rows.
__global__ __launch_bounds__(kBlock) void k_cells(PhysParams P, int n, const double* __restrict__ pv_ref, const double* __restrict__ pc_scale,
                        const double* __restrict__ p, const double* __restrict__ sw, const double* __restrict__ hc,
                        const int* __restrict__ reg, const double* __restrict__ acc0, double* __restrict__ R,
                        double* __restrict__ A, double* __restrict__ pack) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  D3 dp = dvar<3>(p[i], 0), ds = dvar<3>(sw[i], 1), dh = dvar<3>(hc[i], 2);
  int rg = reg[i];
  D3 inv3[3];
  const double cap = rs_cap_at(P, i);
  inventory(P, pv_ref[i], rg, dp, ds, dh, inv3, cap);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    R[c * n + i] = inv3[c].v - acc0[c * n + i];
#pragma unroll
    for (int v = 0; v < 3; ++v) A[(size_t)(c * 3 + v) * n + i] = inv3[c].d[v];
  }
  FaceCell<D3> fc = face_cell(P, rg, dp, ds, dh, pc_scale[i], cap);
  const D3* q = &fc.mw;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    pack[(size_t)(4 * k) * n + i] = q[k].v;
#pragma unroll
    for (int v = 0; v < 3; ++v) pack[(size_t)(4 * k + 1 + v) * n + i] = q[k].d[v];
  }
}

// The face pack of one cell, read in place: entries are loaded where the flux
// needs them (the cell's own pack stays in L1 across its six faces), which keeps
// the face kernel's live state small enough to run without register spills.
struct PackCell {
  const double* __restrict__ pack;
  int n, i;
  __device__ __forceinline__ double operator()(int k) const { return pack[(size_t)k * n + i]; }
};

// quantity k of a pack as (value, d/dp, d/dsw, d/dhc)
#define PK(c, k, e) ((c)(4 * (k) + (e)))
enum { QMW = 0, QMO, QMG, QMS, QRW, QRO, QRG, QPCW, QPCG };

// TPFA flux of one face a -> b (a is the lower natural index), in sc m3/d, with
// partials against a's and b's primaries (rs/fim/flux.rs:42-118, upwinding on
// the value of each phase potential difference, >= 0 picks a).
__device__ __forceinline__ void face_flux(const PackCell& a, const PackCell& b, double pa, double pb, double T,
                                          double G, double F[3], double Ga[3][3], double Gb[3][3]) {
  // potential differences and their partials
  double dw = (pa - pb) - (PK(a, QPCW, 0) - PK(b, QPCW, 0)) - (PK(a, QRW, 0) + PK(b, QRW, 0)) * 0.5 * G;
  double dox = (pa - pb) - (PK(a, QRO, 0) + PK(b, QRO, 0)) * 0.5 * G;
  double dgx = (pa - pb) + (PK(a, QPCG, 0) - PK(b, QPCG, 0)) - (PK(a, QRG, 0) + PK(b, QRG, 0)) * 0.5 * G;
  double ddw_a[3], ddw_b[3], ddo_a[3], ddo_b[3], ddg_a[3], ddg_b[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    double e = v == 0 ? 1.0 : 0.0;
    ddw_a[v] = e - PK(a, QPCW, 1 + v) - 0.5 * G * PK(a, QRW, 1 + v);
    ddw_b[v] = -e + PK(b, QPCW, 1 + v) - 0.5 * G * PK(b, QRW, 1 + v);
    ddo_a[v] = e - 0.5 * G * PK(a, QRO, 1 + v);
    ddo_b[v] = -e - 0.5 * G * PK(b, QRO, 1 + v);
    ddg_a[v] = e + PK(a, QPCG, 1 + v) - 0.5 * G * PK(a, QRG, 1 + v);
    ddg_b[v] = -e - PK(b, QPCG, 1 + v) - 0.5 * G * PK(b, QRG, 1 + v);
  }
  bool uw = dw >= 0.0, uo = dox >= 0.0, ug = dgx >= 0.0;
  const PackCell& cw = uw ? a : b;
  const PackCell& co = uo ? a : b;
  const PackCell& cg = ug ? a : b;
  double mw = PK(cw, QMW, 0), mo = PK(co, QMO, 0), mg = PK(cg, QMG, 0), ms = PK(co, QMS, 0);
  F[0] = mw * dw * T;
  F[1] = mo * dox * T;
  F[2] = (mg * dgx + ms * dox) * T;
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    double mwa = uw ? PK(a, QMW, 1 + v) : 0.0, mwb = uw ? 0.0 : PK(b, QMW, 1 + v);
    double moa = uo ? PK(a, QMO, 1 + v) : 0.0, mob = uo ? 0.0 : PK(b, QMO, 1 + v);
    double msa = uo ? PK(a, QMS, 1 + v) : 0.0, msb = uo ? 0.0 : PK(b, QMS, 1 + v);
    double mga = ug ? PK(a, QMG, 1 + v) : 0.0, mgb = ug ? 0.0 : PK(b, QMG, 1 + v);
    Ga[0][v] = (mw * ddw_a[v] + mwa * dw) * T;
    Gb[0][v] = (mw * ddw_b[v] + mwb * dw) * T;
    Ga[1][v] = (mo * ddo_a[v] + moa * dox) * T;
    Gb[1][v] = (mo * ddo_b[v] + mob * dox) * T;
    Ga[2][v] = (mg * ddg_a[v] + mga * dgx + ms * ddo_a[v] + msa * dox) * T;
    Gb[2][v] = (mg * ddg_b[v] + mgb * dgx + ms * ddo_b[v] + msb * dox) * T;
  }
}

// Per cell: adds its six faces to its own residual row and diagonal block and
// writes its six off-diagonal blocks. Each face is evaluated by both of its
// cells in the same (lower -> upper) orientation, so the two rows receive
// exactly opposite fluxes without atomics (rs/fim/assembly_ad.rs:828-960).
// minimum waves per SIMD the face kernel is compiled for (register budget 512 / N)
#ifndef RSF_FACE_MIN_WAVES
#define RSF_FACE_MIN_WAVES 2
#endif
__global__ __launch_bounds__(kBlock, RSF_FACE_MIN_WAVES) void k_face_rows(Grid g, int gravity, double dt, const double* __restrict__ Tx,
                            const double* __restrict__ Ty, const double* __restrict__ Tz,
                            const double* __restrict__ depth_k, const double* __restrict__ p,
                            const double* __restrict__ pack, double* __restrict__ R, double* __restrict__ A) {
  int i = RSF_CELL_INDEX();
  int n = g.n;
  if (i >= n) return;
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  const PackCell me{pack, n, i};
  double pi = p[i];
  double r[3], d[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    r[c] = R[c * n + i];
#pragma unroll
    for (int v = 0; v < 3; ++v) d[c][v] = A[(size_t)(c * 3 + v) * n + i];
  }
#pragma unroll 1
  for (int s = 1; s <= 6; ++s) {
    int j = slot_nbr(g, i, s, x, y, z);
    double off[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (j >= 0) {
      bool minus = (s & 1) != 0;  // slots 1,3,5: neighbour below i
      int lo = minus ? j : i;
      const double* Tarr = s <= 2 ? Tx : (s <= 4 ? Ty : Tz);
      double T = Tarr[lo];
      if (T > 0.0) {
        const PackCell nb{pack, n, j};
        int zl = minus ? (s == 5 ? z - 1 : z) : z;
        int zu = minus ? z : (s == 6 ? z + 1 : z);
        double G = gravity ? kGravity * (depth_k[zl] - depth_k[zu]) * 1e-5 : 0.0;
        double F[3], Ga[3][3], Gb[3][3];
        // one evaluation in the lower -> upper orientation; this cell takes -F, -Gb
        // (its own block) and -Ga (the neighbour's) when it is the upper cell, else F, Ga
        // and Gb. A single inlined face_flux (instead of one per branch) leaves the
        // register budget without spills; the sign flips are exact, so the numbers do not change.
        const PackCell lo_c = minus ? nb : me, hi_c = minus ? me : nb;
        const double p_lo = minus ? p[j] : pi, p_hi = minus ? pi : p[j];
        face_flux(lo_c, hi_c, p_lo, p_hi, T, G, F, Ga, Gb);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          r[c] += (minus ? -F[c] : F[c]) * dt;
#pragma unroll
          for (int v = 0; v < 3; ++v) {
            d[c][v] += (minus ? -Gb[c][v] : Ga[c][v]) * dt;
            off[c][v] = (minus ? -Ga[c][v] : Gb[c][v]) * dt;
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int v = 0; v < 3; ++v) A[(size_t)(s * 9 + c * 3 + v) * n + i] = off[c][v];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    R[c * n + i] = r[c];
#pragma unroll
    for (int v = 0; v < 3; ++v) A[(size_t)(c * 3 + v) * n + i] = d[c][v];
  }
}

// --------------------------------------------------------------- wells
// One physical well per workgroup, one perforation per thread. Builds the
// perforation source terms, the tail equations (perforation rate consistency
// and the BHP / Fischer-Burmeister control row, rs/fim/wells_ad.rs:146-378),
// and eliminates the tail by its Schur complement (rs/fim/linear/well_schur.rs:70-209).
// The tail of one well is an arrowhead matrix [[alpha, beta^T], [gamma, I]], so
// its inverse is closed-form through s = alpha - beta.gamma.
struct WellDev {
  int n_wells, n_perf;
  const int* well_off;   // [W+1] perforation ranges per well
  const int* perf_cell;  // [P]
  const double* perf_wi;
  const double* perf_head;
  const int* perf_ref;   // [P] completion index (the reference's perforation order) of each perforation
  const int* well_flags;  // bit0 injector, bit1 rate-controlled (FB), bit2 surface target, bit3 gas fluid, bit4 enabled
  const double* well_target;
  const double* well_limit;
  const double* well_bhp_target;
  const int* clique_dest;   // per well: np*np destinations (slot < 7, else 7 + extra index); -1 diag
  const int* clique_off;    // [W+1] offsets into clique_dest
  double* extra_val;        // [9][n_extra]
  int n_extra;
  // per-Newton scratch (device)
  double* bhp;   // [W]
  double* q;     // [P]
  double* tail_R;     // [W + P] residual of tail rows (well rows then perforation rows)
  double* wk;    // per perforation, stride 16: a[3] g[3] h[3] gamma, d rate/dq, beta, connection rate
  double* wwell; // per well: alpha, s, rw = 4 values
};

template <class T>
__device__ __forceinline__ void perforation_terms(const PhysParams& P, bool inj, bool gas, const T& p, const T& sw,
                                                  const T& hc, int rg, T coef[3], T& cm, double rs_cap) {
  Fluids<T> f = cell_fluids(P, rg, p, sw, hc, rs_cap);
  Mobility<T> m = mobilities(P, sw, f.sg, p, f.rs);
  cm = floor_at(m.w + m.o + m.g, 0.0);
  if (inj) {
    if (!gas) {
      coef[0] = inv_bw(P, p);
      coef[1] = lift(p, 0.0);
      coef[2] = lift(p, 0.0);
    } else {
      coef[0] = lift(p, 0.0);
      coef[1] = lift(p, 0.0);
      coef[2] = inv(floor_at(f.bg, 1e-9));
    }
    return;
  }
  T lw = floor_at(m.w, 0.0), lo = floor_at(m.o, 0.0), lg = floor_at(m.g, 0.0);
  T lt = floor_at(lw + lo + lg, kEps);
  T fw = clamp01(lw / lt), fo = clamp01(lo / lt), fg = clamp01(lg / lt);
  T oc = fo / floor_at(f.bo, 1e-9);
  coef[0] = fw * inv_bw(P, p);
  coef[1] = oc;
  coef[2] = fg / floor_at(f.bg, 1e-9) + oc * f.rs;
}

// Perforation rate seen by a rate-controlled well's control row and its q derivative
// (rs/fim/wells_ad.rs:247-304): reservoir rates max(+-q, 0), surface rates scaled by the
// cell's 1/B factor `fac` (1/Bw or 1/Bg for an injector, fo/Bo for a producer).
__device__ __forceinline__ double control_rate(bool inj, bool surf, double q, double fac, double& drq) {
  if (!surf) {
    if (inj) {
      drq = -q >= 0.0 ? -1.0 : 0.0;
      return -q > 0.0 ? -q : 0.0;
    }
    drq = q >= 0.0 ? 1.0 : 0.0;
    return q > 0.0 ? q : 0.0;
  }
  const double cl = inj ? (-q >= 0.0 ? -q : 0.0) : (q >= 0.0 ? q : 0.0);
  const double dcl = inj ? (-q >= 0.0 ? -1.0 : 0.0) : (q >= 0.0 ? 1.0 : 0.0);
  drq = dcl * fac;
  return cl * fac;
}

// Fischer-Burmeister control row of a rate-controlled well (rs/fim/wells_ad.rs:538-572):
// residual phi(xa, xb) of the scaled BHP and rate slacks, alpha = d phi / d bhp and
// dfb = d phi / d xb; sr is the rate scale.
__device__ __forceinline__ void fb_control_row(bool inj, double bhp, double lim, double tgt, double actual,
                                               double& rw, double& alpha, double& dfb, double& sr) {
  const double sb = fabs(lim) > 1.0 ? fabs(lim) : 1.0;
  sr = fabs(tgt) > 1.0 ? fabs(tgt) : 1.0;
  const double xa = (inj ? lim - bhp : bhp - lim) / sb;
  const double xb = (tgt - actual) / sr;
  const double nrm = sqrt((xa * xa + xb * xb) + 2.0 * 1e-6 * 1e-6);
  rw = nrm - xa - xb;
  const double dfa = xa / nrm - 1.0;
  dfb = xb / nrm - 1.0;
  alpha = dfa * (inj ? -1.0 : 1.0) / sb;
}

constexpr int kWellBlock = 256;
// Per perforation (blockIdx.x: the well, blockIdx.y * 64 + threadIdx.x: the perforation in
// it; rs/fim/assembly_ad.rs:239-573): the source terms of its
// cell rows and their Jacobian, the perforation row q - conn(p, bhp) with the strict-sign
// clamp, and the factors the control row and the Schur step read (wk, stride 16; wk[13] is
// the perforation's rate seen by the control row). The cell physics alone, in its own
// kernel: together with the control row and the Schur fill it spilled to scratch.
__global__ __launch_bounds__(64) void k_wells_perf(PhysParams P, int n, double dt, WellDev W,
                                                   const double* __restrict__ p, const double* __restrict__ sw,
                                                   const double* __restrict__ hc, const int* __restrict__ reg,
                                                   double* __restrict__ R, double* __restrict__ A) {
  int w = blockIdx.x;
  int p0 = W.well_off[w], np = W.well_off[w + 1] - p0;
  int flags = W.well_flags[w];
  bool inj = flags & 1, fb = (flags & 2) != 0, surf = (flags & 4) != 0, gas = (flags & 8) != 0;
  double bhp = W.bhp[w];
  // one perforation per thread, no loop: in a strided loop the compiler hoists the property
  // tables' loads out of it and runs out of registers
  {
    const int t = blockIdx.y * blockDim.x + threadIdx.x;
    if (t >= np) return;
    int e = p0 + t;
    int c = W.perf_cell[e];
    D3 dp = dvar<3>(p[c], 0), ds = dvar<3>(sw[c], 1), dh = dvar<3>(hc[c], 2);
    D3 coef[3], cm;
    perforation_terms(P, inj, gas, dp, ds, dh, reg[c], coef, cm, rs_cap_at(P, c));
    double q = W.q[e];
    // source terms into the cell rows; J_RW column a = coef*dt
    double* wk = W.wk + 16 * e;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      R[r * n + c] += coef[r].v * q * dt;
#pragma unroll
      for (int v = 0; v < 3; ++v) A[(size_t)(r * 3 + v) * n + c] += coef[r].d[v] * q * dt;
      wk[r] = coef[r].v * dt;
    }
    // perforation row q - conn(p, bhp)   (strict-sign clamp)
    D3 raw = (cm * (dp - bhp - W.perf_head[e])) * W.perf_wi[e];
    bool live = inj ? raw.v < 0.0 : raw.v > 0.0;
    double conn = live ? raw.v : 0.0;
    W.tail_R[W.n_wells + e] = q - conn;
#pragma unroll
    for (int v = 0; v < 3; ++v) wk[3 + v] = live ? -raw.d[v] : 0.0;  // g = d row / d cell
    wk[9] = live ? cm.v * W.perf_wi[e] : 0.0;                         // gamma = d row / d bhp
    wk[12] = conn;                                                     // connection rate (adjoint: d/d ln k)
    // perforation rate seen by the control row (rs/fim/wells_ad.rs:247-304)
    double rate = 0.0, dr0 = 0.0, dr1 = 0.0, dr2 = 0.0, drq = 0.0;
    if (fb) {
      // 1/Bg, 1/Bw or fo/Bo, picked value by value with selects (picking a coef element by a
      // run-time index put the array in scratch memory)
      const bool g2 = inj && gas, g1 = !inj;
      const double fv = g2 ? coef[2].v : (g1 ? coef[1].v : coef[0].v);
      const double fd0 = g2 ? coef[2].d[0] : (g1 ? coef[1].d[0] : coef[0].d[0]);
      const double fd1 = g2 ? coef[2].d[1] : (g1 ? coef[1].d[1] : coef[0].d[1]);
      const double fd2 = g2 ? coef[2].d[2] : (g1 ? coef[1].d[2] : coef[0].d[2]);
      rate = control_rate(inj, surf, q, fv, drq);
      if (surf) {
        const double cl = inj ? (-q >= 0.0 ? -q : 0.0) : (q >= 0.0 ? q : 0.0);
        dr0 = cl * fd0;
        dr1 = cl * fd1;
        dr2 = cl * fd2;
      }
    }
    wk[6] = dr0;
    wk[7] = dr1;
    wk[8] = dr2;
    wk[10] = drq;
    wk[13] = rate;
  }
}

// Per well (one workgroup): the control row from its perforations' rates (k_wells_perf) and
// the arrowhead tail [[alpha, beta^T], [gamma, I]]: beta, the scaled h and s = alpha -
// beta.gamma (W.wwell). The Schur fill that uses them is k_wells_schur.
__global__ __launch_bounds__(kWellBlock) void k_wells(WellDev W) {
  int w = blockIdx.x;
  int p0 = W.well_off[w], np = W.well_off[w + 1] - p0;
  int flags = W.well_flags[w];
  bool inj = flags & 1, fb = (flags & 2) != 0;
  bool enabled = (flags & 16) != 0;
  double bhp = W.bhp[w];
  extern __shared__ double sh[];
  double* s_bg = sh;  // np (beta * gamma products)
  // control row (sequential, deterministic)
  __shared__ double s_alpha, s_rw, s_dfb, s_sr;
  if (threadIdx.x == 0) {
    double alpha, rw, dfb = 0.0, sr = 1.0;
    if (!enabled || !fb) {
      alpha = 1.0;
      rw = bhp - W.well_bhp_target[w];
    } else {
      double actual = 0.0;
      for (int t = 0; t < np; ++t) actual += W.wk[16 * (p0 + t) + 13];
      fb_control_row(inj, bhp, W.well_limit[w], W.well_target[w], actual, rw, alpha, dfb, sr);
    }
    s_alpha = alpha;
    s_rw = rw;
    s_dfb = dfb;
    s_sr = sr;
    W.tail_R[w] = rw;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    double* wk = W.wk + 16 * (p0 + t);
    double f = (enabled && fb) ? s_dfb * (-1.0 / s_sr) : 0.0;
    wk[11] = f * wk[10];  // beta = d control / d q_e
    for (int v = 0; v < 3; ++v) wk[6 + v] = f * wk[6 + v];  // h = d control / d cell
    s_bg[t] = wk[11] * wk[9];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sacc = s_alpha;
    for (int t = 0; t < np; ++t) sacc -= s_bg[t];
    W.wwell[4 * w] = s_alpha;
    W.wwell[4 * w + 1] = sacc;
  }
}

// Schur complement of the wells' tails into the cell blocks (rs/fim/linear/well_schur.rs:
// 70-209): S(ci,cj) = -a_i (x) (delta_ij g_j + gamma_i/s * u_j), u_j = beta_j g_j - h_j. One
// thread per completion pair (i, j) of well blockIdx.x, spread over blockIdx.y; every pair
// owns a distinct destination block (completions of a well sit in distinct cells).
__global__ __launch_bounds__(kWellBlock) void k_wells_schur(int n, WellDev W, double* __restrict__ A) {
  const int w = blockIdx.x;
  const int p0 = W.well_off[w], np = W.well_off[w + 1] - p0;
  const int k = blockIdx.y * blockDim.x + threadIdx.x;
  if (k >= np * np) return;
  const int flags = W.well_flags[w];
  const bool fb = (flags & 2) != 0, enabled = (flags & 16) != 0;
  const double s = W.wwell[4 * w + 1];
  const bool singular = !(fabs(s) > 0.0) || !isfinite(s);
  const bool couple = enabled && fb && !singular;
  const int t = k / np, u = k - t * np;
  if (u != t && !couple) return;
  const int* dest = W.clique_dest + W.clique_off[w];
  const int ci = W.perf_cell[p0 + t];
  const double* wi = W.wk + 16 * (p0 + t);
  const double* wj = W.wk + 16 * (p0 + u);
  const double gi = singular ? 0.0 : wi[9] / s;
  double vec[3];
  for (int v = 0; v < 3; ++v) {
    double uj = wj[11] * wj[3 + v] - wj[6 + v];
    vec[v] = (u == t ? wj[3 + v] : 0.0) + gi * uj;
  }
  const int dd = dest[t * np + u];
  for (int r = 0; r < 3; ++r)
    for (int v = 0; v < 3; ++v) {
      double val = -wi[r] * vec[v];
      if (u == t) {
        A[(size_t)(r * 3 + v) * n + ci] += val;
      } else if (dd >= 0 && dd < 7) {
        A[(size_t)(dd * 9 + r * 3 + v) * n + ci] += val;
      } else if (dd >= 7) {
        W.extra_val[(size_t)(r * 3 + v) * W.n_extra + (dd - 7)] = val;
      }
    }
}

// Newton right-hand side of the reduced system: b = -R_cells - J_RW T^-1 (-R_tail)
__global__ void k_wells_rhs(int n, WellDev W, const double* __restrict__ R, double* __restrict__ b) {
  int w = blockIdx.x;
  int p0 = W.well_off[w], np = W.well_off[w + 1] - p0;
  double s = W.wwell[4 * w + 1];
  bool singular = !(fabs(s) > 0.0) || !isfinite(s);
  double rw = -W.tail_R[w];
  double bq = 0.0;
  for (int t = 0; t < np; ++t) bq += W.wk[16 * (p0 + t) + 11] * (-W.tail_R[W.n_wells + p0 + t]);
  double al = W.wwell[4 * w];
  double dx_bhp = singular ? (fabs(al) > kEps ? rw / al : rw) : (rw - bq) / s;
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    int e = p0 + t;
    const double* wk = W.wk + 16 * e;
    double mq = -W.tail_R[W.n_wells + e] - (singular ? 0.0 : wk[9] * dx_bhp);
    int c = W.perf_cell[e];
    for (int r = 0; r < 3; ++r) b[r * n + c] -= wk[r] * mq;
  }
}

// Tail recovery dx_W = T^-1 (b_W - J_WR dx_R) (rs/fim/linear/well_schur.rs:382-408). The
// block gathers every perforation's operands (its cell's dx, g, h, gamma, beta and tail
// residual) into LDS in parallel; thread 0 then forms the sums in perforation order.
// LDS: 12 doubles per perforation.
__global__ __launch_bounds__(kWellBlock) void k_wells_recover(int n, WellDev W, const double* __restrict__ dx,
                                                              double* __restrict__ dx_tail) {
  extern __shared__ double sh[];
  int w = blockIdx.x;
  int p0 = W.well_off[w], np = W.well_off[w + 1] - p0;
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    const int e = p0 + t, c = W.perf_cell[e];
    const double* wk = W.wk + 16 * e;
    double* o = sh + 12 * t;
    for (int v = 0; v < 3; ++v) {
      o[v] = dx[v * n + c];
      o[3 + v] = wk[3 + v];
      o[6 + v] = wk[6 + v];
    }
    o[9] = wk[9];
    o[10] = wk[11];
    o[11] = W.tail_R[W.n_wells + e];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double s = W.wwell[4 * w + 1];
  bool singular = !(fabs(s) > 0.0) || !isfinite(s);
  double t0 = -W.tail_R[w];
  double bt = 0.0;
  for (int t = 0; t < np; ++t) {
    int e = p0 + t;
    const double* o = sh + 12 * t;
    double te = -o[11];
    for (int v = 0; v < 3; ++v) {
      t0 -= o[6 + v] * o[v];
      te -= o[3 + v] * o[v];
    }
    dx_tail[W.n_wells + e] = te;
    bt += o[10] * te;
  }
  double al = W.wwell[4 * w];
  double dbhp = singular ? (fabs(al) > kEps ? t0 / al : t0) : (t0 - bt) / s;
  dx_tail[w] = dbhp;
  for (int t = 0; t < np; ++t) {
    int e = p0 + t;
    if (!singular) dx_tail[W.n_wells + e] -= sh[12 * t + 9] * dbhp;
  }
}

// Nested per-well Newton (rs/fim/wells_inner.rs:523-598): every well converges its own rows
// (the control row and its perforations' rate consistency, the rows k_wells assembles) in
// its own unknowns (bhp, q of each perforation) with the cells frozen at the current
// iterate. One workgroup per well: the cell-dependent factors of each perforation are formed
// once, in parallel, by the functions k_wells uses; lane 0 then runs the chopped Newton
// loop. The local Jacobian is the arrowhead [[alpha, beta^T], [gamma, I]], solved through
// s = alpha - beta.gamma: dbhp = (-rw + beta.r) / s, dq = -r - gamma dbhp (the system is
// exactly singular when s = 0, where the reference's LU stops). Convergence: the scaled
// residual peak (rs/fim/scaling.rs:107-119) within `tol` and, for a pressure-controlled
// well, every perforation flowing its own way (rs/fim/wells_inner.rs:474-521). The BHP
// step is capped at dbhp_max_rel |bhp| and floored at 1 bar - 1 Pa (:432-436); q moves by
// the raw step. Report per well: rep_i[2w] converged, rep_i[2w + 1] iterations, rep_peak[w].
constexpr int kWellInnerArrays = 9;
__global__ __launch_bounds__(kWellBlock) void k_wells_inner(PhysParams P, WellDev W, int max_it, double tol,
                                                            double dbhp_max_rel, const double* __restrict__ p,
                                                            const double* __restrict__ sw, const double* __restrict__ hc,
                                                            const int* __restrict__ reg, int* __restrict__ rep_i,
                                                            double* __restrict__ rep_peak) {
  const int w = blockIdx.x;
  const int p0 = W.well_off[w], np = W.well_off[w + 1] - p0;
  const int flags = W.well_flags[w];
  const bool inj = flags & 1, fb = (flags & 2) != 0, surf = (flags & 4) != 0, gas = (flags & 8) != 0;
  extern __shared__ double sh[];
  double* s_cm = sh;            // total mobility of the perforated cell
  double* s_pc = sh + np;       // cell pressure
  double* s_head = sh + 2 * np; // head offset
  double* s_wi = sh + 3 * np;   // well index
  double* s_fac = sh + 4 * np;  // 1/B factor of the control rate
  double* s_q = sh + 5 * np;
  double* s_r = sh + 6 * np;    // perforation row residual
  double* s_g = sh + 7 * np;    // gamma = d row / d bhp
  double* s_dq = sh + 8 * np;   // d rate / d q
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    const int e = p0 + t, c = W.perf_cell[e];
    D3 dp = dvar<3>(p[c], 0), ds = dvar<3>(sw[c], 1), dh = dvar<3>(hc[c], 2);
    D3 coef[3], cm;
    perforation_terms(P, inj, gas, dp, ds, dh, reg[c], coef, cm, rs_cap_at(P, c));
    s_cm[t] = cm.v;
    s_pc[t] = p[c];
    s_head[t] = W.perf_head[e];
    s_wi[t] = W.perf_wi[e];
    s_fac[t] = inj ? (gas ? coef[2].v : coef[0].v) : coef[1].v;
    s_q[t] = W.q[e];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bhp = W.bhp[w], peak = 0.0;
    int its = 0;
    bool conv = false;
    while (true) {
      double rw, alpha = 1.0, f = 0.0;
      if (fb) {
        double actual = 0.0;
        for (int t = 0; t < np; ++t) actual += control_rate(inj, surf, s_q[t], s_fac[t], s_dq[t]);
        double dfb, sr;
        fb_control_row(inj, bhp, W.well_limit[w], W.well_target[w], actual, rw, alpha, dfb, sr);
        f = dfb * (-1.0 / sr);
      } else {
        rw = bhp - W.well_bhp_target[w];
      }
      peak = fabs(rw) / (fb ? 1.0 : fmax(fabs(bhp), 1.0));
      bool direction = true;
      double s = alpha, num = -rw;
      for (int t = 0; t < np; ++t) {
        const double q = s_q[t];
        const double raw = (s_cm[t] * ((s_pc[t] - bhp) - s_head[t])) * s_wi[t];
        const bool live = inj ? raw < 0.0 : raw > 0.0;
        const double r = q - (live ? raw : 0.0);
        const double g = live ? s_cm[t] * s_wi[t] : 0.0;
        peak = fmax(peak, fabs(r) / fmax(fabs(q), 1.0));
        direction = direction && (inj ? q <= 0.0 : q >= 0.0);
        const double beta = fb ? f * s_dq[t] : 0.0;
        s -= beta * g;
        num += beta * r;
        s_r[t] = r;
        s_g[t] = g;
      }
      conv = peak <= tol && (fb || direction);
      if (conv || its >= max_it) break;
      if (!(fabs(s) > 0.0) || !isfinite(s)) break;
      const double db = num / s;
      const double cap = fabs(bhp) * dbhp_max_rel;
      const double lim = db < -cap ? -cap : (db > cap ? cap : db);
      bhp = fmax(bhp + lim, 1.0 - 1.0e-5);
      for (int t = 0; t < np; ++t) s_q[t] += -s_r[t] - s_g[t] * db;
      its++;
    }
    W.bhp[w] = bhp;
    rep_i[2 * w] = conv ? 1 : 0;
    rep_i[2 * w + 1] = its;
    rep_peak[w] = peak;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < np; t += blockDim.x) W.q[p0 + t] = s_q[t];
}

// sum over the 64 lanes of a wavefront (result valid in lane 0)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// --------------------------------------------------------------- Newton record
// The Newton loop's bookkeeping stays on the device: the acceptance measures, the tail
// families, the well update with its bounds and relaxation. The host reads one record per
// Newton iteration. Reductions over block records run in block order with std::max's
// NaN behaviour ((a < b) ? b : a), as the host reduction they replace.
enum {
  kRecFam = 0,        // [3] scaled family maxima (water, oil, gas)
  kRecCnv = 3,        // [3] CNV
  kRecMb = 6,         // [3] MB
  kRecViolFrac = 9,   // violating pore-volume fraction
  kRecPvSum = 10,
  kRecTail = 11,      // [2] well and perforation row families
  kRecTail2 = 13,     // sum of squares of the tail residual
  kRecWellsOk = 14,   // every well passes its own convergence test
  kRecCells2 = 15,    // |R_cells|^2 (dot kernels)
  kRecBad = 16,       // non-finite state after the last update
  kRecPmax = 17,      // largest cell pressure after the last update
  kRecRaw = 18,       // [13] the reduced partial values
  kRecItem = 31,      // [5] the item of each family's peak: cell, cell, cell, well, completion
  kRecCount = 36
};
// Block records of k_conv_partial: 13 reduced values, then the first cell of each
// family's peak (-1: no positive value in the block)
constexpr int kConvStride = 16;

__device__ __forceinline__ double hmax(double a, double b) { return a < b ? b : a; }
__device__ __forceinline__ double hmin(double a, double b) { return b < a ? b : a; }

constexpr int kRecBlock = 256;  // threads of the record kernels

// Sum, or maximum with std::max's NaN behaviour, of r[stride b + k] over b < nb from
// `init`: strided per thread, then a fixed tree, so the result is the same on every run.
// Every thread of the kRecBlock-thread block calls it and gets the result.
template <bool kMax>
__device__ double block_records(const double* __restrict__ r, int stride, int k, int nb, double init) {
  __shared__ double sm[kRecBlock];
  double v = init;
  for (int b = threadIdx.x; b < nb; b += kRecBlock) {
    const double x = r[stride * b + k];
    v = kMax ? hmax(v, x) : v + x;
  }
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int o = kRecBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] = kMax ? hmax(sm[threadIdx.x], sm[threadIdx.x + o]) : sm[threadIdx.x] + sm[threadIdx.x + o];
    __syncthreads();
  }
  const double out = sm[0];
  __syncthreads();
  return out;
}

// Smallest of every thread's `v` over the kRecBlock-thread block (every thread calls it).
__device__ int block_min_int(int v) {
  __shared__ int sm[kRecBlock];
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int o = kRecBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] = min(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  const int out = sm[0];
  __syncthreads();
  return out;
}

// k_conv_partial records reduced (maxima for the first six values, sums for the rest);
// bavg = FVF sums / n. The cell of each family's peak is the first cell holding the
// family maximum (rs/fim/newton/convergence.rs:325-345 keeps a later row only when it is
// strictly larger): the smallest of the blocks' first cells among the blocks whose
// maximum equals it (0 when no value is positive, the reference's first row).
__global__ __launch_bounds__(kRecBlock) void k_conv_reduce(int nb, int n, const double* __restrict__ part,
                                                           double* __restrict__ bavg, double* __restrict__ rec) {
  for (int k = 0; k < 13; ++k) {
    const double v = k < 6 ? block_records<true>(part, kConvStride, k, nb, 0.0)
                           : block_records<false>(part, kConvStride, k, nb, 0.0);
    if (threadIdx.x == 0) {
      rec[kRecRaw + k] = v;
      if (k >= 9 && k < 12) bavg[k - 9] = v / (n > 1 ? n : 1);
    }
    if (k < 3) {
      int first = INT_MAX;
      for (int b = threadIdx.x; b < nb; b += kRecBlock) {
        const double* r = part + (size_t)kConvStride * b;
        const int c = (int)r[13 + k];
        if (r[k] == v && c >= 0) first = min(first, c);
      }
      first = block_min_int(first);
      if (threadIdx.x == 0) rec[kRecItem + k] = first == INT_MAX ? 0.0 : (double)first;
    }
  }
}

// Acceptance measures (rs/fim/newton/convergence.rs:352-385) from the reduced records and
// the violation partials; the tail families (rs/fim/scaling.rs:107-119), the tail's sum of
// squares and every well's own convergence test (rs/fim/wells_inner.rs:474-521) from the
// assembled tail rows.
__global__ __launch_bounds__(kRecBlock) void k_conv_finish(int nb, const double* __restrict__ vpart,
                                                           const double* __restrict__ bavg, WellDev W,
                                                           double* __restrict__ rec) {
  const double viol = block_records<false>(vpart, 1, 0, nb, 0.0);
  // Wells in parallel, each well's test in the sequential order. The family maxima start
  // from 0 (or +inf) and take hmax(running, value), which never admits a NaN, so they are
  // the maxima of the finite values whatever the order: a tree over the threads gives them.
  // The item of each peak is the first (well, or completion in the reference's order)
  // holding the maximum; -1 while no value is positive.
  double t0 = W.n_wells ? 0.0 : INFINITY, t1 = W.n_perf ? 0.0 : INFINITY;
  int i0 = -1, i1 = -1;
  bool ok = true;
  for (int w = threadIdx.x; w < W.n_wells; w += blockDim.x) {
    const int flags = W.well_flags[w];
    const bool fb = (flags & 2) != 0, inj = flags & 1;
    const double bhp = W.bhp[w];
    double peak = fabs(W.tail_R[w]) / (fb ? 1.0 : hmax(fabs(bhp), 1.0));
    if (t0 < peak) {
      t0 = peak;
      i0 = w;
    }
    bool direction = true;
    for (int e = W.well_off[w]; e < W.well_off[w + 1]; ++e) {
      const double q = W.q[e];
      const double a = fabs(W.tail_R[W.n_wells + e]) / hmax(fabs(q), 1.0);
      const int ref = W.perf_ref[e];
      if (t1 < a || (t1 == a && i1 >= 0 && ref < i1)) {
        t1 = a;
        i1 = ref;
      }
      peak = hmax(peak, a);
      direction = direction && (inj ? q <= 0.0 : q >= 0.0);
    }
    ok = ok && peak <= 1e-4 && (fb || direction);
  }
  __shared__ double s0[kRecBlock], s1[kRecBlock], stage[4 * kRecBlock];
  __shared__ int sok[kRecBlock], si0[kRecBlock], si1[kRecBlock];
  s0[threadIdx.x] = t0;
  s1[threadIdx.x] = t1;
  si0[threadIdx.x] = i0;
  si1[threadIdx.x] = i1;
  sok[threadIdx.x] = ok ? 1 : 0;
  __syncthreads();
  // a partner's peak replaces this thread's when larger, or equal with an earlier item
  // (equal positive values both carry an item; equal zeros carry none)
  for (int o = kRecBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const int u = threadIdx.x + o;
      if (s0[threadIdx.x] < s0[u] || (s0[threadIdx.x] == s0[u] && si0[u] >= 0 && si0[u] < si0[threadIdx.x])) {
        s0[threadIdx.x] = s0[u];
        si0[threadIdx.x] = si0[u];
      }
      if (s1[threadIdx.x] < s1[u] || (s1[threadIdx.x] == s1[u] && si1[u] >= 0 && si1[u] < si1[threadIdx.x])) {
        s1[threadIdx.x] = s1[u];
        si1[threadIdx.x] = si1[u];
      }
      sok[threadIdx.x] = sok[threadIdx.x] & sok[u];
    }
    __syncthreads();
  }
  // the tail's sum of squares in index order: chunks staged in LDS by every thread,
  // summed by thread 0
  double tail2 = 0.0;
  const int m = W.n_wells + W.n_perf;
  for (int c0 = 0; c0 < m; c0 += 4 * kRecBlock) {
    const int len = min(4 * kRecBlock, m - c0);
    for (int k = threadIdx.x; k < len; k += blockDim.x) stage[k] = W.tail_R[c0 + k];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int k = 0; k < len; ++k) tail2 += stage[k] * stage[k];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  t0 = s0[0];
  t1 = s1[0];
  ok = sok[0] != 0;
  const double* v = rec + kRecRaw;
  const double pvsum = v[12];
  for (int c = 0; c < 3; ++c) {
    rec[kRecFam + c] = v[c];
    rec[kRecCnv + c] = bavg[c] * v[3 + c];
    rec[kRecMb + c] = fabs(bavg[c] * v[6 + c]) / hmax(pvsum, 1e-9);
  }
  rec[kRecViolFrac] = viol / hmax(pvsum, 1e-9);
  rec[kRecPvSum] = pvsum;
  rec[kRecTail] = t0;
  rec[kRecTail + 1] = t1;
  rec[kRecItem + 3] = si0[0] < 0 ? 0.0 : (double)si0[0];
  rec[kRecItem + 4] = si1[0] < 0 ? 0.0 : (double)si1[0];
  rec[kRecTail2] = tail2;
  rec[kRecWellsOk] = ok ? 1.0 : 0.0;
}

// Newton update of the well unknowns from the recovered tail step (rs/fim/newton/
// convergence.rs:648-711): BHP step capped at |bhp| and floored at 1 bar, raw q step.
// One thread per well or perforation.
__global__ void k_well_update(WellDev W, const double* __restrict__ dx_tail, double relax) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < W.n_wells) {
    double d = dx_tail[k] * relax;
    const double bb = W.bhp[k], cap = fabs(bb);
    const double lim = fabs(d) > cap ? (d > 0 ? cap : -cap) : d;
    d = (bb + lim < 1.0) ? 1.0 - bb : lim;
    W.bhp[k] = bb + d;
  } else if (k < W.n_wells + W.n_perf) {
    const int e = k - W.n_wells;
    W.q[e] += dx_tail[W.n_wells + e] * relax;
  }
}

// Control bounds of every BHP (rs/fim/state.rs:433-459) with the largest cell pressure
// reduced from k_max_finite's records (first pass) or the recorded one (second pass), and
// the state's finiteness: cells from those records, wells checked here.
__global__ __launch_bounds__(kRecBlock) void k_well_bounds(WellDev W, int nb, const double* __restrict__ mf,
                                                           double bhp_min, double bhp_max, int second,
                                                           double* __restrict__ rec) {
  double pmax, bad;
  if (second) {
    pmax = rec[kRecPmax];
    bad = rec[kRecBad];
  } else {
    pmax = block_records<true>(mf, 2, 0, nb, -INFINITY);
    bad = block_records<true>(mf, 2, 1, nb, 0.0);
  }
  if (threadIdx.x != 0) return;
  const double pu = hmax(hmax(bhp_max, 1.0), pmax) + 500.0;
  for (int w = 0; w < W.n_wells; ++w) {
    const double lim = W.well_limit[w];
    double b = W.bhp[w];
    if (W.well_flags[w] & 1) {
      const double hi = hmax(pu, hmax(lim, bhp_max));
      b = hmin(hmax(b, 1e-6), hi);
    } else {
      const double lo = hmax(hmin(lim, bhp_min), 1e-6);
      b = hmin(hmax(b, lo), pu);
    }
    W.bhp[w] = b;
  }
  if (second) {
    for (int w = 0; w < W.n_wells; ++w)
      if (!isfinite(W.bhp[w])) bad = 1.0;
    for (int e = 0; e < W.n_perf; ++e)
      if (!isfinite(W.q[e])) bad = 1.0;
  }
  rec[kRecPmax] = pmax;
  rec[kRecBad] = bad;
}

// Relaxation of each well toward local consistency (rs/fim/state.rs:494-551) with the
// cells frozen at the updated iterate: the BHP blended 0.9 toward the consistent value
// (the control's BHP, or for a rate-controlled well the bisection of
// rs/fim/wells.rs:431-504 on its total rate) within 25 bar, then each q blended 0.75
// toward its connection rate within max(10 %, 250 m3/d), sign-clamped. One workgroup per
// well: the cell factors of its perforations are formed in parallel, lane 0 iterates.
constexpr int kWellRelaxArrays = 6;
__global__ __launch_bounds__(kWellBlock) void k_wells_relax(PhysParams P, WellDev W, const double* __restrict__ p,
                                                            const double* __restrict__ sw,
                                                            const double* __restrict__ hc,
                                                            const int* __restrict__ reg) {
  const int w = blockIdx.x;
  const int p0 = W.well_off[w], np = W.well_off[w + 1] - p0;
  const int flags = W.well_flags[w];
  const bool inj = flags & 1, rate = (flags & 2) != 0, surf = (flags & 4) != 0, gas = (flags & 8) != 0;
  const bool enabled = (flags & 16) != 0;
  extern __shared__ double sh[];
  double* s_cm = sh;            // total mobility
  double* s_pc = sh + np;       // cell pressure
  double* s_head = sh + 2 * np;
  double* s_wi = sh + 3 * np;
  double* s_a = sh + 4 * np;    // surface rate = (max(+-q, 0) a) / b
  double* s_b = sh + 5 * np;
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    const int e = p0 + t, c = W.perf_cell[e];
    const double pc = p[c];
    const Fluids<double> f = cell_fluids<double>(P, reg[c], pc, sw[c], hc[c], rs_cap_at(P, c));
    const Mobility<double> m = mobilities<double>(P, sw[c], f.sg, pc, f.rs);
    s_cm[t] = hmax(m.w + m.o + m.g, 0.0);
    s_pc[t] = pc;
    s_head[t] = W.perf_head[e];
    s_wi[t] = W.perf_wi[e];
    if (inj) {
      s_a[t] = gas ? 1.0 : inv_bw<double>(P, pc);
      s_b[t] = gas ? hmax(f.bg, 1e-9) : 1.0;
    } else {
      const double lw = hmax(m.w, 0.0), lo = hmax(m.o, 0.0), lg = hmax(m.g, 0.0);
      const double lt = hmax(lw + lo + lg, kEps);
      s_a[t] = hmin(hmax(lo / lt, 0.0), 1.0);
      s_b[t] = hmax(f.bo, 1e-9);
    }
  }
  __syncthreads();
  // the first wavefront runs the well: its lanes share the perforations, and every sum
  // over them is a fixed shuffle tree whose total all lanes hold
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  auto conn = [&](int t, double bhp) {  // rs/fim/wells.rs:1098-1127
    const double raw = s_wi[t] * s_cm[t] * (s_pc[t] - (bhp + s_head[t]));
    if (!isfinite(raw)) return 0.0;
    return inj ? hmin(raw, 0.0) : hmax(raw, 0.0);
  };
  auto total = [&](double bhp) {  // rs/fim/wells.rs:389-408
    double tot = 0.0;
    for (int t = lane; t < np; t += 64) {
      const double q = conn(t, bhp);
      const double cl = inj ? hmax(-q, 0.0) : hmax(q, 0.0);
      tot += surf ? (cl * s_a[t]) / s_b[t] : cl;
    }
    return __shfl(wave_sum(tot), 0, 64);
  };
  double cons = W.well_bhp_target[w];
  if (enabled && rate) {
    if (np == 0) return;
    double mn = INFINITY, mx = -INFINITY;  // every lane scans all perforations
    for (int t = 0; t < np; ++t) {
      const double z = s_pc[t] - s_head[t];
      mn = hmin(mn, z);
      mx = hmax(mx, z);
    }
    if (!isfinite(mn) || !isfinite(mx)) return;
    const double lim = W.well_limit[w], tgt = W.well_target[w];
    if (tgt >= total(lim) - 1e-9) {
      cons = lim;
    } else {
      double lo = inj ? hmin(mn, lim) : lim, hi = inj ? lim : hmax(mx, lim);
      for (int it = 0; it < 64; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double r = total(mid);
        if (inj) {
          if (r < tgt) lo = mid;
          else hi = mid;
        } else {
          if (r > tgt) lo = mid;
          else hi = mid;
        }
      }
      cons = 0.5 * (lo + hi);
    }
  }
  const double prop = W.bhp[w];
  const double bl = prop + 0.9 * (cons - prop);
  const double bhp = hmax(cons + hmin(hmax(bl - cons, -25.0), 25.0), 1e-6);
  if (lane == 0) W.bhp[w] = bhp;
  for (int t = lane; t < np; t += 64) {
    const double cq = enabled ? conn(t, bhp) : 0.0;
    const double pq = W.q[p0 + t];
    const double bq = pq + 0.75 * (cq - pq);
    const double tr = hmax(0.1 * fabs(cq), 250.0);
    const double v = cq + hmin(hmax(bq - cq, -tr), tr);
    W.q[p0 + t] = !enabled ? 0.0 : (inj ? hmin(v, 0.0) : hmax(v, 0.0));
  }
}

// --------------------------------------------------------------- linear algebra
// y = A x over the 7-slot store (value type V for x/y, matrix type M)
template <class M, class V, class Y>
__global__ void k_spmv(Grid g, const M* __restrict__ A, const V* __restrict__ x, Y* __restrict__ y) {
  int i = RSF_CELL_INDEX();
  int n = g.n;
  if (i >= n) return;
  int cx, cy, cz;
  cell_ijk(g, i, cx, cy, cz);
  double acc[3] = {0, 0, 0};
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    int j = s == 0 ? i : slot_nbr(g, i, s, cx, cy, cz);
    if (j < 0) continue;
    double xv[3] = {double(x[j]), double(x[n + j]), double(x[2 * n + j])};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int v = 0; v < 3; ++v) acc[c] += double(A[(size_t)(s * 9 + c * 3 + v) * n + i]) * xv[v];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) y[c * n + i] = Y(acc[c]);
}

// y = b - A x
template <class M, class V>
__global__ void k_residual(Grid g, const M* __restrict__ A, const V* __restrict__ x, const double* __restrict__ b,
                           double* __restrict__ y) {
  int i = RSF_CELL_INDEX();
  int n = g.n;
  if (i >= n) return;
  int cx, cy, cz;
  cell_ijk(g, i, cx, cy, cz);
  double acc[3] = {b[i], b[n + i], b[2 * n + i]};
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    int j = s == 0 ? i : slot_nbr(g, i, s, cx, cy, cz);
    if (j < 0) continue;
    double xv[3] = {double(x[j]), double(x[n + j]), double(x[2 * n + j])};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int v = 0; v < 3; ++v) acc[c] -= double(A[(size_t)(s * 9 + c * 3 + v) * n + i]) * xv[v];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) y[c * n + i] = acc[c];
}


// extra (clique) blocks: y[row] += sign * sum E x[col]; CSR over the rows that have
// extras, one 64-lane wavefront per row (a well clique row holds up to one entry per
// other completion of the well)
template <class M, class V, class Y>
__global__ void k_spmv_extra(int n, int n_rows, const int* __restrict__ rows, const int* __restrict__ ptr,
                             const int* __restrict__ col, const M* __restrict__ E, int n_extra, const V* __restrict__ x,
                             Y* __restrict__ y, double sign) {
  int t = blockIdx.x;
  if (t >= n_rows) return;
  int r = rows[t];
  double acc[3] = {0, 0, 0};
  for (int k = ptr[t] + threadIdx.x; k < ptr[t + 1]; k += 64) {
    int j = col[k];
    double xv[3] = {double(x[j]), double(x[n + j]), double(x[2 * n + j])};
    for (int c = 0; c < 3; ++c)
      for (int v = 0; v < 3; ++v) acc[c] += double(E[(size_t)(c * 3 + v) * n_extra + k]) * xv[v];
  }
  for (int c = 0; c < 3; ++c) acc[c] = wave_sum(acc[c]);
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) y[c * n + r] = Y(double(y[c * n + r]) + sign * acc[c]);
}

__device__ __forceinline__ void inv3x3(const double m[9], double o[9]) {
  // cofactor inverse; exactly singular or non-finite -> diagonal reciprocal
  // (rs/fim/linear/gmres_block_jacobi.rs:888-901)
  double c00 = m[4] * m[8] - m[5] * m[7];
  double c01 = m[5] * m[6] - m[3] * m[8];
  double c02 = m[3] * m[7] - m[4] * m[6];
  double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
  bool ok = det != 0.0 && isfinite(det);
  if (ok) {
    double id = 1.0 / det;
    o[0] = c00 * id;
    o[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    o[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    o[3] = c01 * id;
    o[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    o[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    o[6] = c02 * id;
    o[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    o[8] = (m[0] * m[4] - m[1] * m[3]) * id;
    for (int k = 0; k < 9; ++k) ok = ok && isfinite(o[k]);
  }
  if (!ok) {
    for (int k = 0; k < 9; ++k) o[k] = 0.0;
    for (int d = 0; d < 3; ++d) o[d * 4] = fabs(m[d * 4]) > kEps ? 1.0 / m[d * 4] : 1.0;
  }
}

// Quasi-IMPES restriction weights (row 0 of inv(diag block), max-abs normalised;
// rs/fim/linear/gmres_block_jacobi.rs:961-975, :123-133) and the coarse
// pressure matrix in the same 7-slot layout: P[s][i] = sum_l w_l A[s][l][0].
template <class PT>
__global__ void k_cpr_setup(Grid g, const double* __restrict__ A, double* __restrict__ w, PT* __restrict__ Pm,
                            PT* __restrict__ w_pt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n = g.n;
  if (i >= n) return;
  double m[9], o[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) m[k] = A[(size_t)k * n + i];
  inv3x3(m, o);
  double r[3] = {o[0], o[1], o[2]};
  double mx = fmax(fabs(r[0]), fmax(fabs(r[1]), fabs(r[2])));
  if (mx > kEps)
    for (int l = 0; l < 3; ++l) r[l] /= mx;
  for (int l = 0; l < 3; ++l) {
    w[l * n + i] = r[l];
    w_pt[l * n + i] = PT(r[l]);
  }
  int cx, cy, cz;
  cell_ijk(g, i, cx, cy, cz);
  const int pi = pslot(g, i, cx, cy, cz);
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    double acc = 0.0;
#pragma unroll
    for (int l = 0; l < 3; ++l)
      if (fabs(r[l]) > kEps) acc += r[l] * A[(size_t)(s * 9 + l * 3) * n + i];
    Pm[(size_t)s * n + pi] = PT(acc);
  }
}

template <class PT>
__global__ void k_cpr_extra(int n, int n_extra, const int* __restrict__ erow, const double* __restrict__ w,
                            const double* __restrict__ E, PT* __restrict__ Pe) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_extra) return;
  int r = erow[k];
  double acc = 0.0;
  for (int l = 0; l < 3; ++l) {
    double wl = w[l * n + r];
    if (fabs(wl) > kEps) acc += wl * E[(size_t)(l * 3) * n_extra + k];
  }
  Pe[k] = PT(acc);
}

// ------------------------------------------------ level-scheduled ILU0 (DILU)
// Ordering mode 0: natural (reference ordering, lower neighbours are slots 1,3,5);
// mode 1: red-black colouring by (x+y+z) parity (lower = all neighbours of a
// black cell). With a 7-point pattern ILU(0) only modifies diagonal blocks in
// both orderings, so the factor is D^-1 per cell and the off-diagonal factors
// are A's own blocks (rs/fim/linear/gmres_block_jacobi.rs:239-354).
__device__ __forceinline__ bool is_lower(int mode, int s, int x, int y, int z) {
  if (mode == 0) return (s & 1) != 0;
  return ((x + y + z) & 1) != 0;  // black cells see every neighbour as lower
}
__device__ __forceinline__ bool is_upper(int mode, int s, int x, int y, int z) {
  if (mode == 0) return (s & 1) == 0;
  return ((x + y + z) & 1) == 0;
}
__device__ __forceinline__ int opposite(int s) { return (s & 1) ? s + 1 : s - 1; }

// DILU factor arithmetic shared by the per-level kernels and the one-launch natural-order
// march (factorize_block_ilu0, rs/fim/linear/gmres_block_jacobi.rs:239-337, and
// factorize_pressure_ilu0, :746-808), so that both give the same numbers.
// Block: D -= (L_ij D_j^-1) U_ji for one lower neighbour j.
__device__ __forceinline__ void bilu_factor_sub(double (&D)[9], const double (&L)[9], const double (&Dj)[9],
                                                const double (&U)[9]) {
  double T1[9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) T1[a * 3 + b] = L[a * 3] * Dj[b] + L[a * 3 + 1] * Dj[3 + b] + L[a * 3 + 2] * Dj[6 + b];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
      D[a * 3 + b] -= T1[a * 3] * U[b] + T1[a * 3 + 1] * U[3 + b] + T1[a * 3 + 2] * U[6 + b];
}
// Pressure: u -= (p_ij d_j) p_ji for one lower neighbour j; the pivot's reciprocal, with a
// vanishing pivot replaced by the diagonal (or 1).
__device__ __forceinline__ void pilu_factor_sub(double& u, double pij, double dj, double pji) {
  const double lij = pij * dj;
  u -= lij * pji;
}
__device__ __forceinline__ double pilu_factor_inv(double a, double u) {
  if (fabs(u) <= kEps) u = fabs(a) > kEps ? a : 1.0;
  return 1.0 / u;
}

template <class FT>
__global__ void k_bilu_factor_level(Grid g, int mode, const int* __restrict__ rows, int cnt,
                                    const double* __restrict__ A, FT* __restrict__ Dinv) {
  int t = RSF_CELL_INDEX();
  if (t >= cnt) return;
  int n = g.n;
  int i = rows[t];
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  double D[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) D[k] = A[(size_t)k * n + i];
  for (int s = 1; s <= 6; ++s) {
    if (!is_lower(mode, s, x, y, z)) continue;
    int j = slot_nbr(g, i, s, x, y, z);
    if (j < 0) continue;
    const int pj = pslot_nbr(g, j, x, y, z);
    double L[9], Dj[9], U[9];
    for (int k = 0; k < 9; ++k) {
      L[k] = A[(size_t)(s * 9 + k) * n + i];
      Dj[k] = double(Dinv[(size_t)k * n + pj]);
      U[k] = A[(size_t)(opposite(s) * 9 + k) * n + j];
    }
    bilu_factor_sub(D, L, Dj, U);
  }
  double o[9];
  inv3x3(D, o);
  const int pi = pslot(g, i, x, y, z);
#pragma unroll
  for (int k = 0; k < 9; ++k) Dinv[(size_t)k * n + pi] = FT(o[k]);
}

// forward: y_i = r_i - sum_lower A_ij w_j ; w_i = D_i^-1 y_i  (w stored in out)
// (level-scheduled orderings: g.pcmp is 0, the factors are in natural order)
template <class FT, class V>
__global__ void k_bilu_fwd_level(Grid g, int mode, const int* __restrict__ rows, int cnt, const FT* __restrict__ A,
                                 const FT* __restrict__ Dinv, const V* __restrict__ r, V* __restrict__ out) {
  int t = RSF_CELL_INDEX();
  if (t >= cnt) return;
  int n = g.n;
  int i = rows[t];
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  double v[3] = {double(r[i]), double(r[n + i]), double(r[2 * n + i])};
  for (int s = 1; s <= 6; ++s) {
    if (!is_lower(mode, s, x, y, z)) continue;
    int j = slot_nbr(g, i, s, x, y, z);
    if (j < 0) continue;
    double wj[3] = {double(out[j]), double(out[n + j]), double(out[2 * n + j])};
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < 3; ++b) v[c] -= double(A[(size_t)(s * 9 + c * 3 + b) * n + i]) * wj[b];
  }
  for (int c = 0; c < 3; ++c) {
    double acc = 0.0;
    for (int b = 0; b < 3; ++b) acc += double(Dinv[(size_t)(c * 3 + b) * n + i]) * v[b];
    out[c * n + i] = V(acc);
  }
}

// backward: x_i = w_i - D_i^-1 sum_upper A_ij x_j (in place over w)
template <class FT, class V>
__global__ void k_bilu_bwd_level(Grid g, int mode, const int* __restrict__ rows, int cnt, const FT* __restrict__ A,
                                 const FT* __restrict__ Dinv, V* __restrict__ out) {
  int t = RSF_CELL_INDEX();
  if (t >= cnt) return;
  int n = g.n;
  int i = rows[t];
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  double v[3] = {0, 0, 0};
  bool any = false;
  for (int s = 1; s <= 6; ++s) {
    if (!is_upper(mode, s, x, y, z)) continue;
    int j = slot_nbr(g, i, s, x, y, z);
    if (j < 0) continue;
    any = true;
    double xj[3] = {double(out[j]), double(out[n + j]), double(out[2 * n + j])};
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < 3; ++b) v[c] += double(A[(size_t)(s * 9 + c * 3 + b) * n + i]) * xj[b];
  }
  if (!any) return;
  for (int c = 0; c < 3; ++c) {
    double acc = 0.0;
    for (int b = 0; b < 3; ++b) acc += double(Dinv[(size_t)(c * 3 + b) * n + i]) * v[b];
    out[c * n + i] = V(double(out[c * n + i]) - acc);
  }
}

// Red-black ordering (mode 1) applies the whole DILU solve in two launches.
// Red cells have no lower neighbours, so their forward value is w_r = D_r^-1 r_r;
// black cells have no upper neighbours, so their forward value is final:
//   black: x_b = D_b^-1 (r_b - sum_j A_bj w_j), the w_j formed on the fly;
//   red:   x_r = w_r - D_r^-1 sum_j A_rj x_j.
// Every value is rounded to V exactly where the level-by-level sweeps store it,
// so the result is identical to theirs.
// (Both kernels issue the loads of a group of slots unconditionally before using any, a
// slot outside the grid reading the cell's own entries with its terms not taken, as in
// bsr_row: a load under a branch would be waited for at the branch's end.)
template <class FT, class V>
__global__ __launch_bounds__(kBlock, sizeof(FT) == 4 ? 4 : 3) void k_bilu_rb_black(Grid g, const int* __restrict__ rows, int cnt,
                                                             const FT* __restrict__ A, const FT* __restrict__ Dinv,
                                                             const V* __restrict__ r, V* __restrict__ out) {
  int t = RSF_CELL_INDEX();
  if (t >= cnt) return;
  int n = g.n;
  int i = rows[t];
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  const int pi = pslot(g, i, x, y, z);
  double v[3] = {double(r[i]), double(r[n + i]), double(r[2 * n + i])};
#pragma unroll
  for (int gs = 1; gs <= 6; gs += 3) {  // slots 1-3, then 4-6
    FT dj[3][9], as[3][9];
    V rj[3][3];
    bool ok[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int s = gs + k;
      const int j = slot_nbr(g, i, s, x, y, z);
      ok[k] = j >= 0;
      const int jc = ok[k] ? j : i;
      const int pj = pslot_nbr(g, jc, x, y, z);
#pragma unroll
      for (int b = 0; b < 3; ++b) rj[k][b] = r[(size_t)b * n + jc];
#pragma unroll
      for (int e = 0; e < 9; ++e) {
        dj[k][e] = Dinv[(size_t)e * n + pj];
        as[k][e] = A[(size_t)(s * 9 + e) * n + pi];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double wj[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int b = 0; b < 3; ++b) acc += double(dj[k][c * 3 + b]) * double(rj[k][b]);
        wj[c] = double(V(acc));
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double tc = v[c];
#pragma unroll
        for (int b = 0; b < 3; ++b) tc -= double(as[k][c * 3 + b]) * wj[b];
        v[c] = ok[k] ? tc : v[c];
      }
    }
  }
  for (int c = 0; c < 3; ++c) {
    double acc = 0.0;
    for (int b = 0; b < 3; ++b) acc += double(Dinv[(size_t)(c * 3 + b) * n + pi]) * v[b];
    out[c * n + i] = V(acc);
  }
}

template <class FT, class V>
__global__ __launch_bounds__(kBlock, sizeof(FT) == 4 ? 4 : 3) void k_bilu_rb_red(Grid g, const int* __restrict__ rows, int cnt,
                                                           const FT* __restrict__ A, const FT* __restrict__ Dinv,
                                                           const V* __restrict__ r, V* __restrict__ out) {
  int t = RSF_CELL_INDEX();
  if (t >= cnt) return;
  int n = g.n;
  int i = rows[t];
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  const int pi = pslot(g, i, x, y, z);
  double ri[3] = {double(r[i]), double(r[n + i]), double(r[2 * n + i])}, wi[3];
  for (int c = 0; c < 3; ++c) {
    double acc = 0.0;
    for (int b = 0; b < 3; ++b) acc += double(Dinv[(size_t)(c * 3 + b) * n + pi]) * ri[b];
    wi[c] = double(V(acc));
  }
  double v[3] = {0, 0, 0};
  bool any = false;
#pragma unroll
  for (int gs = 1; gs <= 6; gs += 3) {  // slots 1-3, then 4-6
    FT as[3][9];
    V xj[3][3];
    bool ok[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int s = gs + k;
      const int j = slot_nbr(g, i, s, x, y, z);
      ok[k] = j >= 0;
      const int jc = ok[k] ? j : i;
#pragma unroll
      for (int b = 0; b < 3; ++b) xj[k][b] = out[(size_t)b * n + jc];
#pragma unroll
      for (int e = 0; e < 9; ++e) as[k][e] = A[(size_t)(s * 9 + e) * n + pi];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      any = any || ok[k];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double tc = v[c];
#pragma unroll
        for (int b = 0; b < 3; ++b) tc += double(as[k][c * 3 + b]) * double(xj[k][b]);
        v[c] = ok[k] ? tc : v[c];
      }
    }
  }
  for (int c = 0; c < 3; ++c) {
    double o = wi[c];
    if (any) {
      double acc = 0.0;
      for (int b = 0; b < 3; ++b) acc += double(Dinv[(size_t)(c * 3 + b) * n + pi]) * v[b];
      o = double(V(o)) - acc;
    }
    out[c * n + i] = V(o);
  }
}

template <class PT>
__global__ void k_pilu_rb_black(Grid g, const int* __restrict__ rows, int cnt, const PT* __restrict__ Pm,
                                const PT* __restrict__ dinv, const PT* __restrict__ r, PT* __restrict__ out,
                                const double* __restrict__ skip) {
  if (skip && *skip != 0.0) return;
  int t = RSF_CELL_INDEX();
  if (t >= cnt) return;
  int n = g.n;
  int i = rows[t];
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  const int pi = pslot(g, i, x, y, z);
  double v = double(r[i]);
  for (int s = 1; s <= 6; ++s) {
    int j = slot_nbr(g, i, s, x, y, z);
    if (j < 0) continue;
    v -= double(Pm[(size_t)s * n + pi]) * double(PT(double(r[j]) * double(dinv[pslot_nbr(g, j, x, y, z)])));
  }
  out[i] = PT(v * double(dinv[pi]));
}

template <class PT>
__global__ void k_pilu_rb_red(Grid g, const int* __restrict__ rows, int cnt, const PT* __restrict__ Pm,
                              const PT* __restrict__ dinv, const PT* __restrict__ r, PT* __restrict__ out,
                              const double* __restrict__ skip) {
  if (skip && *skip != 0.0) return;
  int t = RSF_CELL_INDEX();
  if (t >= cnt) return;
  int n = g.n;
  int i = rows[t];
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  const int pi = pslot(g, i, x, y, z);
  double di = double(dinv[pi]);
  PT wi = PT(double(r[i]) * di);
  double v = 0.0;
  bool any = false;
  for (int s = 1; s <= 6; ++s) {
    int j = slot_nbr(g, i, s, x, y, z);
    if (j < 0) continue;
    any = true;
    v += double(Pm[(size_t)s * n + pi]) * double(out[j]);
  }
  out[i] = any ? PT(double(wi) - di * v) : wi;
}

// scalar pressure ILU0 (DILU form; rs/fim/linear/gmres_block_jacobi.rs:746-808,
// zero-pivot fallback to the diagonal or 1)
template <class PT>
__global__ void k_pilu_factor_level(Grid g, int mode, const int* __restrict__ rows, int cnt,
                                    const PT* __restrict__ Pm, PT* __restrict__ dinv) {
  int t = RSF_CELL_INDEX();
  if (t >= cnt) return;
  int n = g.n;
  int i = rows[t];
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  const int pi = pslot(g, i, x, y, z);
  double a = double(Pm[pi]);
  double u = a;
  for (int s = 1; s <= 6; ++s) {
    if (!is_lower(mode, s, x, y, z)) continue;
    int j = slot_nbr(g, i, s, x, y, z);
    if (j < 0) continue;
    const int pj = pslot_nbr(g, j, x, y, z);
    pilu_factor_sub(u, double(Pm[(size_t)s * n + pi]), double(dinv[pj]), double(Pm[(size_t)opposite(s) * n + pj]));
  }
  dinv[pi] = PT(pilu_factor_inv(a, u));
}

template <class PT>
__global__ void k_pilu_fwd_level(Grid g, int mode, const int* __restrict__ rows, int cnt, const PT* __restrict__ Pm,
                                 const PT* __restrict__ dinv, const PT* __restrict__ r, PT* __restrict__ out,
                                 const double* __restrict__ skip) {
  if (skip && *skip != 0.0) return;
  int t = RSF_CELL_INDEX();
  if (t >= cnt) return;
  int n = g.n;
  int i = rows[t];
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  const int pi = pslot(g, i, x, y, z);
  double v = double(r[i]);
  for (int s = 1; s <= 6; ++s) {
    if (!is_lower(mode, s, x, y, z)) continue;
    int j = slot_nbr(g, i, s, x, y, z);
    if (j < 0) continue;
    v -= double(Pm[(size_t)s * n + pi]) * double(out[j]);
  }
  out[i] = PT(v * double(dinv[pi]));
}

template <class PT>
__global__ void k_pilu_bwd_level(Grid g, int mode, const int* __restrict__ rows, int cnt, const PT* __restrict__ Pm,
                                 const PT* __restrict__ dinv, PT* __restrict__ out, const double* __restrict__ skip) {
  if (skip && *skip != 0.0) return;
  int t = RSF_CELL_INDEX();
  if (t >= cnt) return;
  int n = g.n;
  int i = rows[t];
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  const int pi = pslot(g, i, x, y, z);
  double v = 0.0;
  bool any = false;
  for (int s = 1; s <= 6; ++s) {
    if (!is_upper(mode, s, x, y, z)) continue;
    int j = slot_nbr(g, i, s, x, y, z);
    if (j < 0) continue;
    any = true;
    v += double(Pm[(size_t)s * n + pi]) * double(out[j]);
  }
  if (!any) return;
  out[i] = PT(double(out[i]) - double(dinv[pi]) * v);
}

// --------------------------------------------------------------- reductions
// Deterministic two-pass dot: per-block partials, then one block sums them.
template <class V1, class V2>
__global__ void k_dot_partial(int n, const V1* __restrict__ a, const V2* __restrict__ b, double* __restrict__ part) {
  __shared__ double sm[kBlock];
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    acc += double(a[i]) * double(b[i]);
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// One modified Gram-Schmidt step fused with the next one's dot: w += sgn h w_q (as
// k_axpy_dev) and the block partials of next . w (as k_dot_partial over the same grid, so the
// partials are the same numbers); with next null, the partials of w . w (the norm).
__global__ void k_mgs_axpy_dot(int n, double sgn, const double* __restrict__ h, const double* __restrict__ vq,
                               double* __restrict__ w, const double* __restrict__ next, double* __restrict__ part) {
  __shared__ double sm[kBlock];
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double wi = double(w[i]) + sgn * (*h) * double(vq[i]);
    w[i] = wi;
    acc += (next ? double(next[i]) : wi) * wi;
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ void k_sum_final(int m, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sm[kBlock];
  double acc = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) acc += part[i];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

// --------------------------------------------------------------- vector ops
template <class X, class Y>
__global__ void k_copy(int n, const X* __restrict__ x, Y* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = Y(x[i]);
}
// The preconditioner's copy of the block matrix: plane k of A (63 planes, natural
// order) into plane k of out at the preconditioner slot of each cell.
template <class Y>
__global__ void k_copy_pslot(Grid g, const double* __restrict__ A, Y* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  int x, y, z;
  cell_ijk(g, i, x, y, z);
  const size_t k = blockIdx.y;
  out[k * g.n + pslot(g, i, x, y, z)] = Y(A[k * g.n + i]);
}

template <class X>
__global__ void k_fill(int n, X* __restrict__ x, X v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}
// y = a*x + y with a = sgn * (*a_num / *a_den) (device scalars; den may be null)
template <class X, class Y>
__global__ void k_axpy_dev(int n, double sgn, const double* __restrict__ a_num, const X* __restrict__ x,
                           Y* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = Y(double(y[i]) + sgn * (*a_num) * double(x[i]));
}
template <class X, class Y>
__global__ void k_axpy(int n, double a, const X* __restrict__ x, Y* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = Y(double(y[i]) + a * double(x[i]));
}
// y = x / sqrt(*nrm2)
template <class X>
__global__ void k_scale_inv_sqrt(int n, const double* __restrict__ nrm2, const X* __restrict__ x, X* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = X(double(x[i]) / sqrt(*nrm2));
}
// y = x / sqrt(*nrm2) when sqrt(*nrm2) > eps (the FGMRES basis vector's normalisation;
// a breakdown leaves it as it is)
template <class X>
__global__ void k_scale_inv_sqrt_if(int n, const double* __restrict__ nrm2, double eps, const X* __restrict__ x,
                                    X* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double nn = sqrt(*nrm2);
  if (i < n && nn > eps) y[i] = X(double(x[i]) / nn);
}

// FGMRES least-squares state on the device (rs/fim/linear/gmres_block_jacobi.rs:1575-1617):
// the Hessenberg matrix H[row * ldh + col] (ldh = restart), the Givens rotations cs, sn,
// the rotated right-hand side g and the solution y of the small triangular system.
// k_fgmres_begin opens a cycle with g = beta e0; k_fgmres_column takes the Gram-Schmidt
// column of inner step j (hc[0..j] the projections, hc[j+1] the squared norm of the new
// vector), applies the earlier rotations, forms and applies the new one, and solves the
// (j+1)-column triangular system by back substitution. One thread: j stays below the
// restart length.
__global__ void k_fgmres_begin(int m, int ldh, double beta, double* __restrict__ H, double* __restrict__ cs,
                               double* __restrict__ sn, double* __restrict__ g) {
  for (int k = threadIdx.x; k < (m + 1) * ldh; k += blockDim.x) H[k] = 0.0;
  for (int k = threadIdx.x; k <= m; k += blockDim.x) {
    g[k] = k == 0 ? beta : 0.0;
    if (k < m) {
      cs[k] = 0.0;
      sn[k] = 0.0;
    }
  }
}

__global__ void k_fgmres_column(int j, int ldh, double eps, const double* __restrict__ hc, double* __restrict__ H,
                                double* __restrict__ cs, double* __restrict__ sn, double* __restrict__ g,
                                double* __restrict__ y, double* __restrict__ nn_out) {
  if (threadIdx.x != 0) return;
  for (int q = 0; q <= j; ++q) H[q * ldh + j] = hc[q];
  const double nn = sqrt(hc[j + 1]);
  H[(j + 1) * ldh + j] = nn;
  *nn_out = nn;
  for (int q = 0; q < j; ++q) {
    const double a = H[q * ldh + j], c = H[(q + 1) * ldh + j];
    H[q * ldh + j] = cs[q] * a + sn[q] * c;
    H[(q + 1) * ldh + j] = -sn[q] * a + cs[q] * c;
  }
  const double a = H[j * ldh + j], c = H[(j + 1) * ldh + j];
  const double rad = hypot(a, c);
  double co = 1.0, si = 0.0;
  if (rad > eps) {
    co = a / rad;
    si = c / rad;
  }
  cs[j] = co;
  sn[j] = si;
  H[j * ldh + j] = co * a + si * c;
  H[(j + 1) * ldh + j] = -si * a + co * c;
  const double g0 = g[j], g1 = g[j + 1];
  g[j] = co * g0 + si * g1;
  g[j + 1] = -si * g0 + co * g1;
  for (int row = j; row >= 0; --row) {
    double s = g[row];
    for (int cc = row + 1; cc <= j; ++cc) s -= H[row * ldh + cc] * y[cc];
    const double dd = H[row * ldh + row];
    y[row] = fabs(dd) > eps ? s / dd : 0.0;
  }
}

// CPR restriction rp_i = sum_l w_l v_l (v in FP64 outer space)
template <class PT, class V>
__global__ void k_restrict(int n, const PT* __restrict__ w, const V* __restrict__ v, PT* __restrict__ rp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    rp[i] = PT(double(w[i]) * double(v[i]) + double(w[n + i]) * double(v[n + i]) + double(w[2 * n + i]) * double(v[2 * n + i]));
}

// CPR: corr = v - A(:, pressure columns) xp   (the residual after the prolonged pressure correction)
template <class M, class PT, class V, class O>
__global__ void k_cpr_corr(Grid g, const M* __restrict__ A, const PT* __restrict__ xp, const V* __restrict__ v,
                           O* __restrict__ corr) {
  int i = RSF_CELL_INDEX();
  int n = g.n;
  if (i >= n) return;
  int cx, cy, cz;
  cell_ijk(g, i, cx, cy, cz);
  const int ai = pslot(g, i, cx, cy, cz);  // A is the preconditioner's copy
  double acc[3] = {double(v[i]), double(v[n + i]), double(v[2 * n + i])};
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    int j = s == 0 ? i : slot_nbr(g, i, s, cx, cy, cz);
    if (j < 0) continue;
    double xv = double(xp[j]);
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] -= double(A[(size_t)(s * 9 + c * 3) * n + ai]) * xv;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) corr[c * n + i] = O(acc[c]);
}

// extras contribution to the CPR correction: corr[row] -= E(:, pressure col) xp[col]
template <class M, class PT, class O>
__global__ void k_cpr_corr_extra(int n, int n_rows, const int* __restrict__ rows, const int* __restrict__ ptr,
                                 const int* __restrict__ col, const M* __restrict__ E, int n_extra,
                                 const PT* __restrict__ xp, O* __restrict__ corr) {
  int t = blockIdx.x;
  if (t >= n_rows) return;
  int r = rows[t];
  double acc[3] = {0, 0, 0};
  for (int k = ptr[t] + threadIdx.x; k < ptr[t + 1]; k += 64) {
    double xv = double(xp[col[k]]);
    for (int c = 0; c < 3; ++c) acc[c] += double(E[(size_t)(c * 3) * n_extra + k]) * xv;
  }
  for (int c = 0; c < 3; ++c) acc[c] = wave_sum(acc[c]);
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) corr[c * n + r] = O(double(corr[c * n + r]) - acc[c]);
}

// z_0 += xp (prolongation e0 of the pressure correction)
template <class PT, class Z>
__global__ void k_prolong_add(int n, const PT* __restrict__ xp, Z* __restrict__ z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = Z(double(z[i]) + double(xp[i]));
}

// FGMRES candidate: out = origin + sum_c y_c Z_c   (coefficients in device memory)
template <class Z>
__global__ void k_combine(int n, int cols, const double* __restrict__ origin, const Z* const* __restrict__ Zs,
                          const double* __restrict__ y, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int c = 0; c < cols; ++c) acc += y[c] * double(Zs[c][i]);
  out[i] = origin[i] + acc;
}

// --------------------------------------------------------------- Newton helpers
// Per-cell convergence inputs: per-family scaled residual maxima, CNV maxima,
// residual sums, FVF sums and pore-volume sums (rs/fim/scaling.rs:139-187,
// rs/fim/newton/convergence.rs:478-612). Output per block: kConvStride partials, the
// last three the first cell of each family's block maximum (-1 when none is positive).
__global__ __launch_bounds__(kBlock) void k_conv_partial(PhysParams P, int n, double dt, const double* __restrict__ pv_ref,
                               const double* __restrict__ p, const double* __restrict__ rs_stored,
                               const double* __restrict__ R, double* __restrict__ part) {
  __shared__ double sm[13][kBlock];
  __shared__ int sfirst[3][kBlock];
  double v[13];
  int first[3] = {-1, -1, -1};
  for (int k = 0; k < 13; ++k) v[k] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double pi = p[i];
    double bo;
    if (P.has_table) {
      if (P.three_phase) {
        double mu;
        pvt_oil<double>(P, pi, rs_stored[i], bo, mu);
      } else {
        bo = pvt_saturated<double>(P, pi).bo;
      }
    } else {
      bo = base_bo<double>(P, pi);
    }
    bo = bo > 1e-9 ? bo : 1e-9;
    double bg = P.has_table ? pvt_saturated<double>(P, pi).bg : 1.0;
    bg = bg > 1e-9 ? bg : 1e-9;
    double ibw = inv_bw<double>(P, pi);
    double bw = 1.0 / (ibw > 1e-9 ? ibw : 1e-9);
    double pv = pv_ref[i];
    double pvdt = fabs(pv / (dt > 1e-12 ? dt : 1e-12));
    pvdt = pvdt > 1.0 ? pvdt : 1.0;
    double B[3] = {bw, bo, bg};
    double pvc = pv > 1e-9 ? pv : 1e-9;
    for (int c = 0; c < 3; ++c) {
      double r = R[c * n + i];
      const double sc = fabs(r) / (pvdt / B[c]);  // family scaled maxima and their first cell
      if (v[c] < sc) {
        v[c] = sc;
        first[c] = i;
      }
      v[3 + c] = fmax(v[3 + c], fabs(r) / pvc);      // CNV maxima
      v[6 + c] += r;                                   // residual sums
      v[9 + c] += B[c];                                // FVF sums
    }
    v[12] += pvc;
  }
  for (int k = 0; k < 13; ++k) sm[k][threadIdx.x] = v[k];
  for (int c = 0; c < 3; ++c) sfirst[c][threadIdx.x] = first[c];
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const int u = threadIdx.x + o;
      for (int c = 0; c < 3; ++c) {
        const double a = sm[c][threadIdx.x], b = sm[c][u];
        const int fb = sfirst[c][u];
        if (a < b || (a == b && fb >= 0 && fb < sfirst[c][threadIdx.x])) {
          sm[c][threadIdx.x] = b;
          sfirst[c][threadIdx.x] = fb;
        }
      }
      for (int k = 3; k < 13; ++k) {
        double a = sm[k][threadIdx.x], b = sm[k][u];
        sm[k][threadIdx.x] = (k < 6) ? fmax(a, b) : a + b;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int k = 0; k < 13; ++k) part[blockIdx.x * kConvStride + k] = sm[k][0];
    for (int c = 0; c < 3; ++c) part[blockIdx.x * kConvStride + 13 + c] = (double)sfirst[c][0];
  }
}

// violating pore volume with the field-average FVFs known
__global__ void k_conv_violation(int n, const double* __restrict__ pv_ref, const double* __restrict__ R,
                                 const double* __restrict__ bavg, double* __restrict__ part) {
  __shared__ double sm[kBlock];
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double pv = pv_ref[i] > 1e-9 ? pv_ref[i] : 1e-9;
    double cm = 0.0;
    for (int c = 0; c < 3; ++c) cm = fmax(cm, fabs(R[c * n + i]) * bavg[c] / pv);
    if (cm > 1e-2) acc += pv;
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// Per-cell OPM chop (rs/fim/newton/convergence.rs:648-711) and raw update with the
// pressure floor (rs/fim/state.rs:672-715)
__global__ void k_chop_update(int n, double relax, const int* __restrict__ reg, const double* __restrict__ dx,
                              double* __restrict__ p, double* __restrict__ sw, double* __restrict__ hc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double dp = dx[i] * relax, dsw = dx[n + i] * relax, dhc = dx[2 * n + i] * relax;
  bool sat = reg[i] == kSaturated;
  double dsg = sat ? dhc : 0.0, dso = sat ? -(dsw + dhc) : -dsw;
  double mx = fmax(fabs(dsw), fmax(fabs(dso), fabs(dsg)));
  double a = mx > 0.2 ? 0.2 / mx : 1.0;
  dsw *= a;
  if (sat) dhc *= a;
  else if (hc[i] + dhc < 0.0) dhc = -hc[i];
  double pi = p[i];
  double cap = 0.3 * fabs(pi);
  if (fabs(dp) > cap) dp = dp > 0.0 ? cap : -cap;
  double np_ = pi + dp;
  p[i] = np_ > 1e-6 ? np_ : 1e-6;
  sw[i] += dsw;
  hc[i] += dhc;
}

// max over cells (pressure upper bound of the well control clamp) and finiteness
__global__ void k_max_finite(int n, const double* __restrict__ p, const double* __restrict__ sw,
                             const double* __restrict__ hc, double* __restrict__ part) {
  __shared__ double sm[kBlock];
  __shared__ double bad[kBlock];
  double mx = -INFINITY, b = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    mx = fmax(mx, p[i]);
    if (!isfinite(p[i]) || !isfinite(sw[i]) || !isfinite(hc[i])) b = 1.0;
  }
  sm[threadIdx.x] = mx;
  bad[threadIdx.x] = b;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
      bad[threadIdx.x] = fmax(bad[threadIdx.x], bad[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sm[0];
    part[2 * blockIdx.x + 1] = bad[0];
  }
}

// Substep initialisation: regime classification and primaries (rs/fim/flash.rs:13-41,
// rs/fim/state.rs:109-145)
__global__ void k_classify(PhysParams P, int n, const double* __restrict__ sp, const double* __restrict__ ssw,
                           const double* __restrict__ ssg, const double* __restrict__ srs, double* __restrict__ p,
                           double* __restrict__ sw, double* __restrict__ hc, int* __restrict__ reg) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int r = kSaturated;
  if (P.three_phase && P.has_table && !(ssg[i] > 1e-9)) {
    double rsat = pvt_saturated<double>(P, sp[i]).rs;
    if (P.rs_base) rsat = fmin(rsat, srs[i]);  // rs/fim/flash.rs:31-33 (gas redissolution disabled)
    if (srs[i] < rsat - 1e-6) r = kUndersaturated;
  }
  p[i] = sp[i];
  sw[i] = ssw[i];
  hc[i] = r == kSaturated ? ssg[i] : srs[i];
  reg[i] = r;
}

// Commit: flash each cell to (So, Sg, Rs) (rs/fim/state.rs:899-914)
__global__ __launch_bounds__(kBlock) void k_write_back(PhysParams P, int n, const double* __restrict__ p, const double* __restrict__ sw,
                             const double* __restrict__ hc, const int* __restrict__ reg, double* __restrict__ sp,
                             double* __restrict__ ssw, double* __restrict__ sso, double* __restrict__ ssg,
                             double* __restrict__ srs) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fluids<double> f = cell_fluids<double>(P, reg[i], p[i], sw[i], hc[i], P.rs_base ? srs[i] : INFINITY);
  sp[i] = p[i];
  ssw[i] = sw[i];
  if (!P.three_phase) {
    ssg[i] = 0.0;
    sso[i] = fmax(1.0 - sw[i], 0.0);
    srs[i] = 0.0;
  } else {
    ssg[i] = f.sg;
    sso[i] = f.so;
    srs[i] = f.rs;
  }
}

// OPM relative change sums (rs/fim/timestep.rs:436-466): out[2] partial (delta, denom)
__global__ void k_rel_change(int n, const double* __restrict__ p0, const double* __restrict__ sw0,
                             const double* __restrict__ hc0, const int* __restrict__ reg0, const double* __restrict__ p,
                             const double* __restrict__ sw, const double* __restrict__ hc, const int* __restrict__ reg,
                             double* __restrict__ part) {
  __shared__ double a[kBlock], b[kBlock];
  double da = 0.0, db = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double dp = p[i] - p0[i];
    da += dp * dp;
    db += p[i] * p[i];
    double ns = reg[i] == kSaturated ? fmax(hc[i], 0.0) : 0.0;
    double os = reg0[i] == kSaturated ? fmax(hc0[i], 0.0) : 0.0;
    double nso = 1.0 - sw[i] - ns, oso = 1.0 - sw0[i] - os;
    double pr[3][2] = {{sw[i], sw0[i]}, {ns, os}, {nso, oso}};
    for (int k = 0; k < 3; ++k) {
      double ds = pr[k][0] - pr[k][1];
      da += ds * ds;
      db += pr[k][0] * pr[k][0];
    }
  }
  a[threadIdx.x] = da;
  b[threadIdx.x] = db;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      a[threadIdx.x] += a[threadIdx.x + o];
      b[threadIdx.x] += b[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a[0];
    part[2 * blockIdx.x + 1] = b[0];
  }
}

// gather perforated cells' iterate (for the host-side well relaxation)
// iterate of the completed cells plus their stored Rs (the redissolution cap)
__global__ void k_gather_cells(int m, const int* __restrict__ cells, const double* __restrict__ p,
                               const double* __restrict__ sw, const double* __restrict__ hc,
                               const int* __restrict__ reg, const double* __restrict__ srs, double* __restrict__ out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  int c = cells[t];
  out[5 * t] = p[c];
  out[5 * t + 1] = sw[c];
  out[5 * t + 2] = hc[c];
  out[5 * t + 3] = double(reg[c]);
  out[5 * t + 4] = srs[c];
}

}  // namespace rsf

namespace rsf {
// gather four per-cell fields at a list of cells: out[4*t + f]
// Report quantities of the stored state, block sums over cells (rs/fim/timestep.rs:1849-1888
// inventories, rs/reporting.rs:304-327 and 646-666 averages): [0] water, [1] oil and [2]
// gas (free + dissolved, three-phase only) inventory at surface conditions on the
// geometric pore volume, [3] pv p and [4] pv over cells with a positive finite pore
// volume, [5] Sw, [6] Sg. Fixed grid of kReportBlocks blocks; the host sums the block
// records in order.
constexpr int kReportBlocks = 256;
// Sweep sums over the stored state (rs/reporting.rs:52-123 compute_sweep_metrics): per
// block, the cells with sw above the threshold, the (i, j) columns holding one, and the
// mobile oil left, sum of max(so - sor, 0). One thread per column walks its layers.
__global__ __launch_bounds__(kBlock) void k_sweep_sums(int nx, int ny, int nz, double threshold, double sor,
                                                     const double* __restrict__ sw, const double* __restrict__ so,
                                                     double* __restrict__ part) {
  double a[3] = {0.0, 0.0, 0.0};
  const int nxy = nx * ny;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nxy; c += gridDim.x * blockDim.x) {
    double col = 0.0;
    for (int k = 0; k < nz; ++k) {
      const size_t i = (size_t)k * nxy + c;
      const double wgt = sw[i] > threshold ? 1.0 : 0.0;
      a[0] += wgt;
      col = col > wgt ? col : wgt;
      const double m = so[i] - sor;
      a[2] += m > 0.0 ? m : 0.0;
    }
    a[1] += col;
  }
  __shared__ double w[3][kBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double v = wave_sum(a[k]);
    if (lane == 0) w[k][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
    for (int j = 0; j < kBlock / 64; ++j) t += w[threadIdx.x][j];
    part[3 * blockIdx.x + threadIdx.x] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_report_sums(PhysParams P, int n, const double* __restrict__ pv,
                                                      const double* __restrict__ p, const double* __restrict__ sw,
                                                      const double* __restrict__ so, const double* __restrict__ sg,
                                                      const double* __restrict__ rs, double* __restrict__ part) {
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double pvi = pv[i], pi = p[i], pvf = pvi > 1e-9 ? pvi : 1e-9;
    double bo, mu;
    if (P.has_table) {  // rs/pvt.rs:769-779 get_b_o_cell at the stored Rs
      if (P.three_phase) pvt_oil<double>(P, pi, rs[i], bo, mu);
      else bo = pvt_saturated<double>(P, pi).bo;
    } else {
      bo = base_bo<double>(P, pi);
    }
    bo = bo > 1e-9 ? bo : 1e-9;
    a[0] += sw[i] * pvi * inv_bw<double>(P, pi);
    a[1] += so[i] * pvf / bo;
    if (P.three_phase) {
      double bg = P.has_table ? pvt_saturated<double>(P, pi).bg : 1.0;  // rs/pvt.rs:835-841
      bg = bg > 1e-9 ? bg : 1e-9;
      a[2] += sg[i] * pvf / bg + (P.has_table ? so[i] * pvf * rs[i] / bo : 0.0);
    }
    if (pvi > 0.0 && isfinite(pvi)) {
      a[3] += pi * pvi;
      a[4] += pvi;
    }
    a[5] += sw[i];
    a[6] += sg[i];
  }
  __shared__ double w[7][kBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const double v = wave_sum(a[k]);
    if (lane == 0) w[k][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    double t = 0.0;
    for (int j = 0; j < kBlock / 64; ++j) t += w[threadIdx.x][j];
    part[7 * blockIdx.x + threadIdx.x] = t;
  }
}

__global__ void k_gather4(int m, const int* __restrict__ cells, const double* __restrict__ a,
                          const double* __restrict__ b, const double* __restrict__ c, const double* __restrict__ d,
                          double* __restrict__ out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  int i = cells[t];
  out[4 * t] = a[i];
  out[4 * t + 1] = b[i];
  out[4 * t + 2] = c[i];
  out[4 * t + 3] = d[i];
}
}  // namespace rsf

namespace rsf {
// ------------------------------------------------ device-resident BiCGStab
// rs/fim/linear/gmres_block_jacobi.rs:660-730, right-preconditioned, x0 = 0.
// Iteration k is 8 launches: p update, 2 ILU, SpMV(+rh.v), s update(+|s|^2),
// 2 ILU, SpMV(+t.t, t.s), x/r update(+|r|^2, rh.r). The reducing kernels run on a
// fixed grid of at most kRGrid blocks of up to kRBlock threads and leave one partial
// pair per block. Each consuming kernel sums those pairs itself, in the same fixed
// order in every block, so all blocks derive bit-identical scalars (alpha, omega,
// beta, the stop decisions) without a separate reduction launch or a host round
// trip. Block 0 records them in `bs` for the kernels that follow. Once `done` is
// set every later kernel of the solve returns at once. A kernel never sets a flag
// that its own blocks test on entry unless every block would stop anyway; the
// closing kernel marks convergence on s in `conv_s`, which the next opening
// kernel turns into `done` (the host reads both).
constexpr int kRBlock = 1024;
constexpr int kRGrid = 512;
constexpr int kRUnroll = 2;  // rows per thread per sweep: loads of both are issued before use
// kBsK: iterations opened so far (the red-black kernels take the iteration from it)
// kBsRr: the residual norm the solve measured last (|b| at the start, |s| when it
// converges on s); with kBsTol the host forms the solve's residual reduction from it.
enum { kBsTol = 0, kBsRho0, kBsRho1, kBsAlpha, kBsOmega, kBsIts, kBsDone, kBsConvS, kBsRr, kBsK, kBsCount };
// partial-pair regions (each 2*kRGrid doubles): 0 |r|^2 and rh.r, 1 rh.v, 2 |s|^2, 3 t.t and t.s
constexpr int kPartStride = 2 * kRGrid;

// Block sums of two values per thread into part[2*block], part[2*block+1]; every
// thread of the block must call it (no early exits).
__device__ void block_partial2(double x1, double x2, double* __restrict__ part) {
  __shared__ double w1[16], w2[16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  x1 = wave_sum(x1);
  x2 = wave_sum(x2);
  if (lane == 0) {
    w1[wv] = x1;
    w2[wv] = x2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < nw; ++k) {
      a += w1[k];
      b += w2[k];
    }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// Totals of partial pairs, summed in a fixed order and returned to every thread:
// a = sum pa[2k], b = sum pa[2k+1] over m pairs, and with pc set c = sum pc[2k],
// d = sum pc[2k+1] over the same m (one pass, one barrier pair).
__device__ void sum_pairs(const double* __restrict__ pa, const double* __restrict__ pc, int m, double& a, double& b,
                          double& c, double& d) {
  __shared__ double w[4][16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double x0 = 0.0, x1 = 0.0, x2 = 0.0, x3 = 0.0;
  for (int k = threadIdx.x; k < m; k += blockDim.x) {
    x0 += pa[2 * k];
    x1 += pa[2 * k + 1];
    if (pc) {
      x2 += pc[2 * k];
      x3 += pc[2 * k + 1];
    }
  }
  x0 = wave_sum(x0);
  x1 = wave_sum(x1);
  if (pc) {
    x2 = wave_sum(x2);
    x3 = wave_sum(x3);
  }
  if (lane == 0) {
    w[0][wv] = x0;
    w[1][wv] = x1;
    w[2][wv] = x2;
    w[3][wv] = x3;
  }
  __syncthreads();
  a = b = c = d = 0.0;
  for (int k = 0; k < nw; ++k) {
    a += w[0][k];
    b += w[1][k];
    c += w[2][k];
    d += w[3][k];
  }
  __syncthreads();
}

// Entry test of a kernel that opens an iteration: true when the solve is over.
// Convergence on s, marked by the previous closing kernel, becomes `done` here.
__device__ __forceinline__ bool bicg_closed(double* __restrict__ bs) {
  if (bs[kBsDone] != 0.0) return true;
  if (bs[kBsConvS] != 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) bs[kBsDone] = 1.0;
    return true;
  }
  return false;
}

// Opening of iteration k from rr = |r|^2 and rho = rh.r (|b|^2 twice at k = 1,
// which also fixes the tolerance): stop tests, rho breakdown, beta. Every thread
// computes the same result; block 0 records rho and the iteration count (or done).
// Returns false when the solve stops.
__device__ bool bicg_open(double* __restrict__ bs, int k, double rr, double rho, double rel_tol, double& beta,
                          double& om) {
  const double eps = kEps;
  double tol;
  bool stop;
  beta = 0.0;
  om = 1.0;
  const double rn = sqrt(rr);
  if (k == 1) {
    tol = (rn > eps ? rn : eps) * rel_tol;
    stop = rn <= tol;
  } else {
    tol = bs[kBsTol];
    stop = rn <= tol;
    om = bs[kBsOmega];
    beta = (rho / bs[kBsRho0 + ((k - 1) & 1)]) * (bs[kBsAlpha] / om);
  }
  if (!stop && (!isfinite(rho) || fabs(rho) < eps)) stop = true;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (k == 1) bs[kBsTol] = tol;
    bs[kBsRr] = rn;
    if (stop) {
      bs[kBsDone] = 1.0;
    } else {
      bs[kBsRho0 + (k & 1)] = rho;
      bs[kBsIts] = double(k);
    }
  }
  return !stop;
}

// Sweeps of a grid-stride kernel over tiles of blockDim*kRUnroll rows (XCD-aware
// tile order); row of unroll slot u.
#define RSF_TILE_LOOP(base)                                                              \
  for (int base = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x * kRUnroll; base < n; \
       base += gridDim.x * blockDim.x * kRUnroll)
#define RSF_ROW(base, u) ((base) + (u) * (int)blockDim.x + (int)threadIdx.x)

// x = 0, r = rh = b, p = v = 0; partials |b|^2 and rh.r (= |b|^2) for iteration 1
template <class V>
__global__ __launch_bounds__(kRBlock, 8) void k_bicg_init(int n, const V* __restrict__ b, V* __restrict__ x,
                                                       V* __restrict__ r, V* __restrict__ rh, V* __restrict__ p,
                                                       V* __restrict__ v, double* __restrict__ bs,
                                                       double* __restrict__ part0) {
  double b2 = 0.0;
  RSF_TILE_LOOP(base) {
    V bv[kRUnroll];
#pragma unroll
    for (int u = 0; u < kRUnroll; ++u) {
      int i = RSF_ROW(base, u);
      bv[u] = i < n ? b[i] : V(0);
    }
#pragma unroll
    for (int u = 0; u < kRUnroll; ++u) {
      int i = RSF_ROW(base, u);
      if (i >= n) continue;
      x[i] = V(0);
      r[i] = bv[u];
      rh[i] = bv[u];
      p[i] = V(0);
      v[i] = V(0);
      b2 += double(bv[u]) * double(bv[u]);
    }
  }
  block_partial2(b2, b2, part0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    bs[kBsIts] = 0.0;
    bs[kBsDone] = 0.0;
    bs[kBsConvS] = 0.0;
  }
}

// Opens iteration k: stop tests on |r| (|b| at k = 1, which also fixes the
// tolerance), rho = rh.r with its breakdown test, beta; then p = r + beta (p - omega v).
template <class V>
__global__ __launch_bounds__(kRBlock, 8) void k_bicg_p(int n, int k, double* __restrict__ bs,
                                                    const double* __restrict__ part0, int m,
                                                    const V* __restrict__ r, const V* __restrict__ v,
                                                    V* __restrict__ p, double rel_tol) {
  if (bicg_closed(bs)) return;
  // first tile's loads go out before the partial sums, so the two latencies overlap
  const int base0 = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x * kRUnroll;
  V rv[kRUnroll], pv[kRUnroll], vv[kRUnroll];
  auto load = [&](int base) {
#pragma unroll
    for (int u = 0; u < kRUnroll; ++u) {
      int i = RSF_ROW(base, u);
      if (i < n) {
        rv[u] = r[i];
        pv[u] = p[i];
        vv[u] = v[i];
      }
    }
  };
  if (base0 < n) load(base0);
  double rr, rho, unused0, unused1;
  sum_pairs(part0, nullptr, m, rr, rho, unused0, unused1);
  double beta, om;
  if (!bicg_open(bs, k, rr, rho, rel_tol, beta, om)) return;
  RSF_TILE_LOOP(base) {
    if (base != base0) load(base);
#pragma unroll
    for (int u = 0; u < kRUnroll; ++u) {
      int i = RSF_ROW(base, u);
      if (i < n) p[i] = V(double(rv[u]) + beta * (double(pv[u]) - om * double(vv[u])));
    }
  }
}

// y = P x over the 7-point stencil plus the clique extras of well rows, grid-stride
// over tiles of blockDim*kRUnroll rows. A wavefront owns 64 consecutive rows; the extras rows among
// them (xwave[w] .. xwave[w+1] in the sorted list) are summed by the whole
// wavefront. With part set, the block partials of y.u (u = u_dot, or y when null) and y.w
// (w optional) follow. bs (optional) carries the solve's done flag.
template <class PT>
__global__ __launch_bounds__(kRBlock, 8) void k_pspmv_dot(Grid g, const PT* __restrict__ Pm, const PT* __restrict__ x,
                                                       PT* __restrict__ y, const int* __restrict__ xwave,
                                                       const int* __restrict__ xrows, const int* __restrict__ xptr,
                                                       const int* __restrict__ xcol, const PT* __restrict__ Pe,
                                                       const PT* __restrict__ u_dot, const PT* __restrict__ w,
                                                       const double* __restrict__ bs, double* __restrict__ part) {
  if (bs && bs[kBsDone] != 0.0) return;
  const int n = g.n, lane = threadIdx.x & 63;
  double d1 = 0.0, d2 = 0.0;
  RSF_TILE_LOOP(base) {
    double acc[kRUnroll];
#pragma unroll
    for (int u = 0; u < kRUnroll; ++u) {
      const int i = RSF_ROW(base, u);
      acc[u] = 0.0;
      if (i < n) {
        int cx, cy, cz;
        cell_ijk(g, i, cx, cy, cz);
        const int pi = pslot(g, i, cx, cy, cz);
        double a = double(Pm[pi]) * double(x[i]);
#pragma unroll
        for (int s = 1; s < 7; ++s) {
          int j = slot_nbr(g, i, s, cx, cy, cz);
          if (j < 0) continue;
          a += double(Pm[(size_t)s * n + pi]) * double(x[j]);
        }
        acc[u] = a;
      }
    }
    if (xwave) {
#pragma unroll
      for (int u = 0; u < kRUnroll; ++u) {
        const int i = RSF_ROW(base, u), wid = i >> 6;
        for (int t = xwave[wid]; t < xwave[wid + 1]; ++t) {
          double e = 0.0;
          for (int k = xptr[t] + lane; k < xptr[t + 1]; k += 64) e += double(Pe[k]) * double(x[xcol[k]]);
          e = __shfl(wave_sum(e), 0, 64);
          if (i == xrows[t]) acc[u] += e;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kRUnroll; ++u) {
      const int i = RSF_ROW(base, u);
      if (i >= n) continue;
      const PT yi = PT(acc[u]);
      y[i] = yi;
      double yd = double(yi);
      d1 += yd * (u_dot ? double(u_dot[i]) : yd);
      if (w) d2 += yd * double(w[i]);
    }
  }
  if (part) block_partial2(d1, d2, part);
}

// alpha = rho / rh.v (breakdown stops the solve), s = r - alpha v, partials |s|^2
template <class V>
__global__ __launch_bounds__(kRBlock, 8) void k_bicg_s(int n, int k, double* __restrict__ bs,
                                                    const double* __restrict__ part1, int m,
                                                    const V* __restrict__ r, const V* __restrict__ v,
                                                    V* __restrict__ s, double* __restrict__ part2) {
  if (bs[kBsDone] != 0.0) return;
  const int base0 = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x * kRUnroll;
  V rv[kRUnroll], vv[kRUnroll];
  auto load = [&](int base) {
#pragma unroll
    for (int u = 0; u < kRUnroll; ++u) {
      int i = RSF_ROW(base, u);
      if (i < n) {
        rv[u] = r[i];
        vv[u] = v[i];
      }
    }
  };
  if (base0 < n) load(base0);
  double vd, unused0, unused1, unused2;
  sum_pairs(part1, nullptr, m, vd, unused0, unused1, unused2);
  const bool stop = !isfinite(vd) || fabs(vd) < kEps;
  const double al = bs[kBsRho0 + (k & 1)] / vd;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (stop) bs[kBsDone] = 1.0;
    else bs[kBsAlpha] = al;
  }
  if (stop) return;
  double acc = 0.0;
  RSF_TILE_LOOP(base) {
    if (base != base0) load(base);
#pragma unroll
    for (int u = 0; u < kRUnroll; ++u) {
      int i = RSF_ROW(base, u);
      if (i >= n) continue;
      V si = V(double(rv[u]) - al * double(vv[u]));
      s[i] = si;
      acc += double(si) * double(si);
    }
  }
  block_partial2(acc, 0.0, part2);
}

// Closes iteration k. Converged on s: x += alpha ph and stop. Otherwise
// omega = t.s / t.t (breakdown stops the solve without an update),
// x += alpha ph + omega sh, r = s - omega t, partials |r|^2 and rh.r.
template <class V>
__global__ __launch_bounds__(kRBlock, 8) void k_bicg_update(int n, double* __restrict__ bs,
                                                         const double* __restrict__ part2,
                                                         const double* __restrict__ part3, int m,
                                                         const V* __restrict__ ph, const V* __restrict__ sh,
                                                         const V* __restrict__ s, const V* __restrict__ t,
                                                         const V* __restrict__ rh, V* __restrict__ x, V* __restrict__ r,
                                                         double* __restrict__ part0) {
  if (bs[kBsDone] != 0.0) return;
  const int base0 = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x * kRUnroll;
  V xv[kRUnroll], phv[kRUnroll], shv[kRUnroll], sv[kRUnroll], tv[kRUnroll], rhv[kRUnroll];
  auto load = [&](int base) {
#pragma unroll
    for (int u = 0; u < kRUnroll; ++u) {
      int i = RSF_ROW(base, u);
      if (i < n) {
        xv[u] = x[i];
        phv[u] = ph[i];
        shv[u] = sh[i];
        sv[u] = s[i];
        tv[u] = t[i];
        rhv[u] = rh[i];
      }
    }
  };
  if (base0 < n) load(base0);
  double ss, unused, tt, ts;
  sum_pairs(part2, part3, m, ss, unused, tt, ts);
  const double al = bs[kBsAlpha];
  const bool conv_s = sqrt(ss) <= bs[kBsTol];
  bool brk = false;
  double om = 0.0;
  if (!conv_s) {
    if (!isfinite(tt) || tt <= kEps) {
      brk = true;
    } else {
      om = ts / tt;
      brk = !isfinite(om) || fabs(om) < kEps;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (conv_s) {
      bs[kBsConvS] = 1.0;
      bs[kBsRr] = sqrt(ss);
    } else if (brk) {
      bs[kBsDone] = 1.0;  // every block stops here too
    } else {
      bs[kBsOmega] = om;
    }
  }
  if (brk) return;
  double a1 = 0.0, a2 = 0.0;
  RSF_TILE_LOOP(base) {
    if (base != base0) load(base);
#pragma unroll
    for (int u = 0; u < kRUnroll; ++u) {
      int i = RSF_ROW(base, u);
      if (i >= n) continue;
      if (conv_s) {
        x[i] = V(double(xv[u]) + al * double(phv[u]));
        continue;
      }
      x[i] = V(double(xv[u]) + al * double(phv[u]) + om * double(shv[u]));
      V ri = V(double(sv[u]) - om * double(tv[u]));
      r[i] = ri;
      a1 += double(ri) * double(ri);
      a2 += double(rhv[u]) * double(ri);
    }
  }
  if (!conv_s) block_partial2(a1, a2, part0);
}
}  // namespace rsf

#include "rsf_bicg_rb.hpp"
