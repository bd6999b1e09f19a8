// Pressure BiCGStab over the red-black, colour-compacted pressure system (g.pcmp set:
// red-black ordering, nx even). Same iteration as the general kernels
// (rs/fim/linear/gmres_block_jacobi.rs:660-730, right-preconditioned, x0 = 0), fused
// along the structure of red-black ILU0.
//
// Layout. Cells 2q and 2q+1 share a row and differ in colour; q is their pair. Every
// plane of the solve (matrix, ILU diagonal and all work vectors) is colour-compacted:
// the red cell of pair q at index q, the black one at n/2 + q.
//
// Mapping. A wavefront owns 64 consecutive pairs of one grid row (y, z). The row, the
// colour parity of the row and the offsets of the y and z neighbours are then the same
// for the whole wavefront (scalar registers); only the x ends of the row depend on the
// lane. Loads use a scalar base and a 32-bit lane offset.
//
// With red cells ordered first ILU0 has no red-red or black-black coupling, and
// z = M^-1 y is
//   black  z_b = d_b (y_b - sum_j P_bj w_j),  w_j = d_j y_j over the red neighbours,
//   red    z_r = w_r - d_r sum_j P_rj z_j     over the black neighbours,
// where the red sum is also the off-diagonal part of row r of P z.
//
// Iteration k in five launches:
//   A  the update of iteration k-1 (omega = t.s / t.t; x += alpha ph + omega sh;
//      r = s - omega t), the opening of iteration k (|r|^2 = |s|^2 - 2 omega t.s +
//      omega^2 t.t and rh.r = rh.s - omega rh.t from partials, stop tests, rho, beta),
//      p = r + beta (p - omega v), and the black half of ph = M^-1 p, forming each red
//      neighbour's new r and p from its s, t, p, v; partial chat.ph (black)
//   B  the red half of ph and the red rows of v = P ph, and the wells' clique extras of
//      v (a clique's red column formed on the fly); partial chat.ph (red)
//   C  alpha = rho / (chat.ph) with chat = P^T rh formed once (rh.v = chat.ph); the
//      black rows of v; s = r - alpha v; the black half of sh = M^-1 s, forming each
//      red neighbour's s on the fly; partials |s|^2 and rh.s
//   D  the red half of sh and the red rows of t = P sh
//   E  the black rows of t and the extras rows; partials t.t, t.s, rh.t
// A with k = max_it + 1 only applies the pending update of the last iteration.
//
// Every launch of an iteration is the same: the iteration number k lives in bs[kBsK]
// (A, B and C read it as bs[kBsK] + 1, E advances it), p alternates between two
// buffers by the parity of k (an iteration reads the previous p of a neighbour while it
// writes its own), and the solve ends on the device.
//
// Arithmetic is in the storage precision PT (scalars rounded to it once per launch);
// dot-product partials accumulate in FP64. Neighbour loops are branch-free: a slot
// outside the grid reads the pair's own entry with a zero coefficient, and the pair's
// partner is read back like any neighbour (its value, formed from the same inputs by
// the same operations, equals the thread's own).
//
// Flags: A's blocks test only `done` on entry, and A marks the end of the solve in
// `conv_s`, so every block of that A still applies its share of the update whenever it
// starts. Every later kernel treats either flag as the end; B and D turn conv_s into
// done, so a following A returns at once.
#pragma once

namespace rsf {

// Minimum waves per SIMD of the gather-heavy kernels A, B/D and C: every gather of a
// thread stays in flight, which in FP64 needs more than 64 registers.
template <class PT>
constexpr int rb_minw() {
  return sizeof(PT) == 4 ? 8 : 6;
}

// wavefronts over the pairs: ceil((nx/2) / 64) per row
__host__ __device__ inline int rb_waves(const Grid& g) {
  return ((g.nx / 2 + 63) / 64) * g.ny * g.nz;
}

struct RbWave {
  int q;     // this lane's pair
  int xp;    // pair column within the row
  int o;     // the black cell of each pair is at x = 2 xp + o (o = 1 - (y + z) % 2)
  int y, z;  // the row
  bool live;
};

// W must be the same for every lane of the wavefront
__device__ __forceinline__ RbWave rb_wave(const Grid& g, int W) {
  RbWave w;
  const int nxh = g.nx >> 1, wpr = (nxh + 63) >> 6;
  const int row = W / wpr, seg = W - row * wpr;
  w.z = row / g.ny;
  w.y = row - w.z * g.ny;
  w.o = 1 - ((w.y + w.z) & 1);
  w.xp = seg * 64 + (int)(threadIdx.x & 63);
  w.live = w.xp < nxh;
  w.q = row * nxh + w.xp;
  return w;
}

// Red neighbours of the pair's black cell by slot (x-, x+, y-, y+, z-, z+): compacted
// red index (the pair's own when outside the grid) and whether it exists.
__device__ __forceinline__ void rb_black_nbrs(const Grid& g, const RbWave& w, int (&j)[6], bool (&ok)[6]) {
  const int q = w.q, nxh = g.nx >> 1, nxyh = nxh * g.ny;
  ok[0] = w.o ? true : w.xp > 0;
  j[0] = w.o ? q : q - 1;
  ok[1] = w.o ? w.xp + 1 < nxh : true;
  j[1] = w.o ? q + 1 : q;
  ok[2] = w.y > 0;
  j[2] = q - nxh;
  ok[3] = w.y + 1 < g.ny;
  j[3] = q + nxh;
  ok[4] = w.z > 0;
  j[4] = q - nxyh;
  ok[5] = w.z + 1 < g.nz;
  j[5] = q + nxyh;
#pragma unroll
  for (int k = 0; k < 6; ++k) j[k] = ok[k] ? j[k] : q;
}

// Black neighbours of the pair's red cell: compacted black index minus n/2.
__device__ __forceinline__ void rb_red_nbrs(const Grid& g, const RbWave& w, int (&j)[6], bool (&ok)[6]) {
  const int q = w.q, nxh = g.nx >> 1, nxyh = nxh * g.ny;
  ok[0] = w.o ? w.xp > 0 : true;
  j[0] = w.o ? q - 1 : q;
  ok[1] = w.o ? true : w.xp + 1 < nxh;
  j[1] = w.o ? q : q + 1;
  ok[2] = w.y > 0;
  j[2] = q - nxh;
  ok[3] = w.y + 1 < g.ny;
  j[3] = q + nxh;
  ok[4] = w.z > 0;
  j[4] = q - nxyh;
  ok[5] = w.z + 1 < g.nz;
  j[5] = q + nxyh;
#pragma unroll
  for (int k = 0; k < 6; ++k) j[k] = ok[k] ? j[k] : q;
}

// element i of a (scalar base, 32-bit byte offset: a plane is under 4 GiB)
template <class T>
__device__ __forceinline__ T ldo(const T* __restrict__ a, int i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(a) + unsigned(i) * unsigned(sizeof(T)));
}
template <class T>
__device__ __forceinline__ void sto(T* __restrict__ a, int i, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(a) + unsigned(i) * unsigned(sizeof(T))) = v;
}

// r = s - omega t, p = r + beta (p - omega v), s = r - alpha v: one fixed sequence of
// fused operations wherever each is formed
template <class V>
__device__ __forceinline__ V rb_axpy(V a, V b, V c) {  // a + c b
  return fma(c, b, a);
}
template <class V>
__device__ __forceinline__ V rb_pnew(V r, V p, V v, V beta, V om) {
  return fma(beta, fma(-om, v, p), r);
}

// Block sums of NV values per thread into part[NV*block + k]; every thread of the
// block must call it.
template <int NV>
__device__ void block_partials(const double (&x)[NV], double* __restrict__ part) {
  __shared__ double w[NV][16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double s = wave_sum(x[k]);
    if (lane == 0) w[k][wv] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double a = 0.0;
    for (int j = 0; j < nw; ++j) a += w[threadIdx.x][j];
    part[NV * blockIdx.x + threadIdx.x] = a;
  }
}

// Totals of the records of two partial regions (NA and NB values per block, m blocks),
// summed in a fixed order and returned to every thread. Only the wavefronts that hold
// records reduce them.
template <int NA, int NB>
__device__ void sum_parts(const double* __restrict__ pa, const double* __restrict__ pb, int m, double (&a)[NA],
                          double (&b)[NB]) {
  __shared__ double w[NA + NB][16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nwv = min((m + 63) >> 6, (int)(blockDim.x >> 6));  // wavefronts holding records
  if (wv < nwv) {
    double xa[NA], xb[NB];
#pragma unroll
    for (int k = 0; k < NA; ++k) xa[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NB; ++k) xb[k] = 0.0;
    for (int j = threadIdx.x; j < m; j += nwv * 64) {
#pragma unroll
      for (int k = 0; k < NA; ++k) xa[k] += pa[NA * j + k];
#pragma unroll
      for (int k = 0; k < NB; ++k) xb[k] += pb[NB * j + k];
    }
#pragma unroll
    for (int k = 0; k < NA; ++k) {
      const double s = wave_sum(xa[k]);
      if (lane == 0) w[k][wv] = s;
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const double s = wave_sum(xb[k]);
      if (lane == 0) w[NA + k][wv] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    a[k] = 0.0;
    for (int j = 0; j < nwv; ++j) a[k] += w[k][j];
  }
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    b[k] = 0.0;
    for (int j = 0; j < nwv; ++j) b[k] += w[NA + k][j];
  }
  __syncthreads();
}

__device__ __forceinline__ bool rb_over(const double* __restrict__ bs) {
  return bs[kBsDone] != 0.0 || bs[kBsConvS] != 0.0;
}

// Wavefront loop: tiles of blockDim.x / 64 wavefronts in XCD-aware order; W is the
// wavefront's index (the same for all its lanes). The loop condition is uniform over
// a block.
#define RSF_WAVE_LOOP(g, W)                                                                          \
  for (int W##_t = xcd_tile(blockIdx.x, gridDim.x) * (int)(blockDim.x >> 6), W##_n = rb_waves(g); \
       W##_t < W##_n; W##_t += gridDim.x * (int)(blockDim.x >> 6))                                     \
    if (const int W = W##_t + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); W < W##_n)

// x = 0, r = rh = b (gathered into the compacted order), p = v = 0, chat = P^T b over
// the stencil (the extras follow in k_rb_extras_t); partials |b|^2 and rh.r (= |b|^2)
template <class PT>
__global__ __launch_bounds__(kRBlock, 8) void k_rb_init(Grid g, const PT* __restrict__ Pm, const PT* __restrict__ b,
                                                     PT* __restrict__ x, PT* __restrict__ r, PT* __restrict__ rh,
                                                     PT* __restrict__ p, PT* __restrict__ v, PT* __restrict__ chat,
                                                     double* __restrict__ bs, double* __restrict__ part0) {
  const int n = g.n, nh = n >> 1, nxy = g.nx * g.ny;
  double b2 = 0.0;
  RSF_WAVE_LOOP(g, W) {
    const RbWave w = rb_wave(g, W);
    if (!w.live) continue;
    const int q = w.q;
#pragma unroll
    for (int col = 0; col < 2; ++col) {  // red cell, then black
      const int xi = 2 * w.xp + (col ? w.o : 1 - w.o);
      const int i = (w.z * g.ny + w.y) * g.nx + xi, ci = col ? nh + q : q;
      const PT bv = ldo(b, i);
      sto(x, ci, PT(0));
      sto(r, ci, bv);
      sto(rh, ci, bv);
      sto(p, ci, PT(0));
      sto(v, ci, PT(0));
      b2 += double(bv) * double(bv);
      // column i of P: row j's entry towards i sits in row j's opposite slot
      int jc[6];
      bool ok[6];
      if (col) rb_black_nbrs(g, w, jc, ok);
      else rb_red_nbrs(g, w, jc, ok);
      const int jn[6] = {i - 1, i + 1, i - g.nx, i + g.nx, i - nxy, i + nxy};
      double a = double(ldo(Pm, ci)) * double(bv);
#pragma unroll
      for (int s = 1; s <= 6; ++s) {
        if (!ok[s - 1]) continue;
        const int cj = col ? jc[s - 1] : nh + jc[s - 1];
        a += double(ldo(Pm + (size_t)opposite(s) * n, cj)) * double(ldo(b, jn[s - 1]));
      }
      sto(chat, ci, PT(a));
    }
  }
  const double pr[2] = {b2, b2};
  block_partials<2>(pr, part0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    bs[kBsIts] = 0.0;
    bs[kBsDone] = 0.0;
    bs[kBsConvS] = 0.0;
    bs[kBsK] = 0.0;
  }
}

// iteration opened by the current A (and run by the B and C that follow it)
__device__ __forceinline__ int rb_iter(const double* __restrict__ bs) {
  return int(bs[kBsK]) + 1;
}

// chat[row] += sum_k E(col_k, row) rh[col_k] over the extras rows (E^T through the
// transpose map of the symmetric clique pattern), one wavefront per row; rows and
// columns in the compacted order
template <class PT>
__global__ void k_rb_extras_t(int n_rows, const int* __restrict__ rows, const int* __restrict__ ptr,
                              const int* __restrict__ col, const int* __restrict__ tperm, const PT* __restrict__ Pe,
                              const PT* __restrict__ rh, PT* __restrict__ chat) {
  const int t = blockIdx.x;
  if (t >= n_rows) return;
  double e = 0.0;
  for (int k = ptr[t] + threadIdx.x; k < ptr[t + 1]; k += 64) e += double(Pe[tperm[k]]) * double(rh[col[k]]);
  e = wave_sum(e);
  if (threadIdx.x == 0) chat[rows[t]] = PT(double(chat[rows[t]]) + e);
}

// A (see the header). p of iteration k goes to p0 (k even) or p1 (k odd).
template <class PT>
__global__ __launch_bounds__(kRBlock, rb_minw<PT>()) void k_rb_a(Grid g, int max_it, double* __restrict__ bs,
                                                  const double* __restrict__ part0, const double* __restrict__ part2,
                                                  const double* __restrict__ part3, int m, const PT* __restrict__ Pm,
                                                  const PT* __restrict__ d, const PT* __restrict__ chat,
                                                  PT* __restrict__ x, PT* __restrict__ r, const PT* __restrict__ s,
                                                  const PT* __restrict__ t, PT* __restrict__ ph,
                                                  const PT* __restrict__ sh, const PT* __restrict__ v,
                                                  PT* __restrict__ p0, PT* __restrict__ p1, double rel_tol,
                                                  double* __restrict__ part1) {
  if (bs[kBsDone] != 0.0) return;  // not conv_s: a block of this launch may set it
  const int n = g.n, nh = n >> 1;
  const int k = rb_iter(bs);
  const bool first = k == 1, final_only = k > max_it;
  const PT* __restrict__ pin = (k & 1) ? p0 : p1;
  PT* __restrict__ pout = (k & 1) ? p1 : p0;
  double al = 0.0, om = 1.0, beta = 0.0;
  bool upd_x = false, upd_sh = false, go = false;
  if (first && final_only) {  // no iterations allowed: x stays 0
    if (blockIdx.x == 0 && threadIdx.x == 0) bs[kBsConvS] = 1.0;
    return;
  }
  if (first) {
    double ra[2], unused[1];
    sum_parts<2, 1>(part0, part0, m, ra, unused);
    go = bicg_open(bs, 1, ra[0], ra[1], rel_tol, beta, om);
  } else {
    double c2[2], c3[3];  // (|s|^2, rh.s), (t.t, t.s, rh.t)
    sum_parts<2, 3>(part2, part3, m, c2, c3);
    al = bs[kBsAlpha];
    const double tol = bs[kBsTol];
    bool stop = true;
    double rn = sqrt(c2[0]);  // residual after this update: |s|, or |s - om t| below
    if (rn <= tol) {  // converged on s: x += alpha ph
      upd_x = true;
    } else if (isfinite(c3[0]) && c3[0] > kEps) {
      om = c3[1] / c3[0];
      if (isfinite(om) && fabs(om) >= kEps) {  // otherwise breakdown: no update
        upd_x = upd_sh = true;
        rn = sqrt(fmax(c2[0] - 2.0 * om * c3[1] + om * om * c3[0], 0.0));
        const double rho = c2[1] - om * c3[2];
        stop = final_only || rn <= tol || !isfinite(rho) || fabs(rho) < kEps;
        if (!stop) beta = (rho / bs[kBsRho0 + ((k - 1) & 1)]) * (al / om);
        if (!stop && blockIdx.x == 0 && threadIdx.x == 0) {
          bs[kBsRho0 + (k & 1)] = rho;
          bs[kBsOmega] = om;
          bs[kBsIts] = double(k);
        }
      }
    }
    go = !stop;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      if (upd_x) bs[kBsRr] = rn;
      if (stop) bs[kBsConvS] = 1.0;
    }
  }
  if (!go && !upd_x) return;
  const PT alp = PT(al), omp = PT(om), omx = PT(upd_sh ? om : 0.0), betp = PT(beta);
  double acc_dot = 0.0;
  RSF_WAVE_LOOP(g, W) {
    const RbWave w = rb_wave(g, W);
    if (!w.live) continue;
    const int q = w.q, qb = nh + q;
    PT rr_, rb_;
    if (first) {
      rr_ = ldo(r, q);
      rb_ = ldo(r, qb);
    } else {
      sto(x, q, PT(ldo(x, q) + alp * ldo(ph, q) + omx * ldo(sh, q)));
      sto(x, qb, PT(ldo(x, qb) + alp * ldo(ph, qb) + omx * ldo(sh, qb)));
      if (!upd_sh) continue;
      rr_ = rb_axpy(ldo(s, q), ldo(t, q), -omp);
      rb_ = rb_axpy(ldo(s, qb), ldo(t, qb), -omp);
      sto(r, q, rr_);
      sto(r, qb, rb_);
    }
    if (!go) continue;
    const PT pnr = rb_pnew(rr_, ldo(pin, q), ldo(v, q), betp, omp);
    const PT pnb = rb_pnew(rb_, ldo(pin, qb), ldo(v, qb), betp, omp);
    sto(pout, q, pnr);
    sto(pout, qb, pnb);
    int jn[6];
    bool ok[6];
    rb_black_nbrs(g, w, jn, ok);
    PT y = pnb;
#pragma unroll
    for (int sl = 1; sl <= 6; ++sl) {
      const int cj = jn[sl - 1];
      const PT rj = first ? ldo(r, cj) : rb_axpy(ldo(s, cj), ldo(t, cj), -omp);
      const PT pj = rb_pnew(rj, ldo(pin, cj), ldo(v, cj), betp, omp);
      const PT a = ok[sl - 1] ? ldo(Pm + (size_t)sl * n, qb) : PT(0);
      y -= a * (pj * ldo(d, cj));
    }
    const PT zb = y * ldo(d, qb);
    sto(ph, qb, zb);
    acc_dot += double(ldo(chat, qb)) * double(zb);
  }
  if (go) {
    const double pr[2] = {acc_dot, 0.0};
    block_partials<2>(pr, part1);
  }
}

// The pair q of any wavefront (row and column from q)
__device__ __forceinline__ RbWave rb_pair_at(const Grid& g, int q) {
  RbWave w;
  const int nxh = g.nx >> 1, row = q / nxh;
  w.xp = q - row * nxh;
  w.z = row / g.ny;
  w.y = row - w.z * g.ny;
  w.o = 1 - ((w.y + w.z) & 1);
  w.q = q;
  w.live = true;
  return w;
}

// Red half of z = M^-1 y at the pair of w, and its neighbour sum (the off-diagonal part
// of the red row of P z). One fixed sequence of fused operations, so that any thread
// forming it gets its owner's rounded value.
template <class PT>
__device__ __forceinline__ PT rb_red_z(const Grid& g, const RbWave& w, const PT* __restrict__ Pm,
                                       const PT* __restrict__ d, const PT* __restrict__ y,
                                       const PT* __restrict__ zb, PT& sum) {
  const int n = g.n, q = w.q;
  int jn[6];
  bool ok[6];
  rb_red_nbrs(g, w, jn, ok);
  const PT dr = ldo(d, q);
  const PT wr = ldo(y, q) * dr;
  sum = PT(0);
#pragma unroll
  for (int sl = 1; sl <= 6; ++sl) {
    const PT a = ok[sl - 1] ? ldo(Pm + (size_t)sl * n, q) : PT(0);
    sum = fma(a, ldo(zb, jn[sl - 1]), sum);
  }
  return fma(-dr, sum, wr);
}

// B and D: red half of z = M^-1 y and the red rows of o = P z. part (optional): partial
// chat.z over the red cells. With xwave set (B with well extras) the extras of o = P z
// are added too: the well rows among the wavefront's pairs (natural rows xrows[e],
// e in xwave[W] .. xwave[W+1]) are summed by the whole wavefront over their compacted
// columns, a red column's z formed on the fly (the black half of z is complete); the
// black half of o then holds the black rows' extras sums (0 elsewhere), to which C adds
// the stencil part.
template <class PT>
__global__ __launch_bounds__(kRBlock, rb_minw<PT>()) void k_rb_red(Grid g, double* __restrict__ bs,
                                                    const PT* __restrict__ Pm, const PT* __restrict__ d,
                                                    const PT* __restrict__ y0, const PT* __restrict__ y1,
                                                    PT* __restrict__ z, PT* __restrict__ o,
                                                    const PT* __restrict__ chat, double* __restrict__ part,
                                                    const int* __restrict__ xwave, const int* __restrict__ xrows,
                                                    const int* __restrict__ xptr, const int* __restrict__ xcolc,
                                                    const PT* __restrict__ Pe) {
  if (rb_over(bs)) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && bs[kBsConvS] != 0.0) bs[kBsDone] = 1.0;
    return;
  }
  const int n = g.n, nh = n >> 1, lane = threadIdx.x & 63;
  // y1 set: the input alternates like p (y1 for odd iterations)
  const PT* __restrict__ y = (y1 && (rb_iter(bs) & 1)) ? y1 : y0;
  const PT* __restrict__ zb = z + nh;
  double acc_dot = 0.0;
  RSF_WAVE_LOOP(g, W) {
    const RbWave w = rb_wave(g, W);
    const int q = w.q;
    PT zr = PT(0), sum = PT(0);
    if (w.live) zr = rb_red_z(g, w, Pm, d, y, zb, sum);
    double xr = 0.0, xb = 0.0;
    if (xwave) {  // every lane of the wavefront takes part
      const int ir = 2 * q + 1 - w.o, ib = 2 * q + w.o;
      for (int e = xwave[W]; e < xwave[W + 1]; ++e) {
        double acc = 0.0;
        for (int kk = xptr[e] + lane; kk < xptr[e + 1]; kk += 64) {
          const int c = xcolc[kk];
          PT zc, unused;
          if (c >= nh) zc = ldo(z, c);
          else zc = rb_red_z(g, rb_pair_at(g, c), Pm, d, y, zb, unused);
          acc += double(Pe[kk]) * double(zc);
        }
        acc = __shfl(wave_sum(acc), 0, 64);
        if (w.live && xrows[e] == ir) xr += acc;
        if (w.live && xrows[e] == ib) xb += acc;
      }
    }
    if (!w.live) continue;
    sto(z, q, zr);
    const PT orr = fma(ldo(Pm, q), zr, sum);
    sto(o, q, xwave ? PT(double(orr) + xr) : orr);
    if (xwave) sto(o, nh + q, PT(xb));
    if (part) acc_dot += double(ldo(chat, q)) * double(zr);
  }
  if (part) {
    const double pr[2] = {acc_dot, 0.0};
    block_partials<2>(pr, part);
  }
}

// C: alpha, the black rows of v (the black half of v holds the extras sums when
// has_extras), s = r - alpha v, the black half of sh; partials |s|^2 and rh.s
template <class PT>
__global__ __launch_bounds__(kRBlock, rb_minw<PT>()) void k_rb_c(Grid g, double* __restrict__ bs,
                                                  const double* __restrict__ part1a,
                                                  const double* __restrict__ part1b, int m,
                                                  const PT* __restrict__ Pm, const PT* __restrict__ d,
                                                  const PT* __restrict__ r, const PT* __restrict__ rh,
                                                  const PT* __restrict__ ph, PT* __restrict__ v, PT* __restrict__ s,
                                                  PT* __restrict__ sh, int has_extras, double* __restrict__ part2) {
  if (rb_over(bs)) return;
  const int n = g.n, nh = n >> 1;
  double ca[2], cb[2];
  sum_parts<2, 2>(part1a, part1b, m, ca, cb);
  const double cv = ca[0] + cb[0];
  const bool stop = !isfinite(cv) || fabs(cv) < kEps;
  const double al = bs[kBsRho0 + (rb_iter(bs) & 1)] / cv;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (stop) bs[kBsDone] = 1.0;
    else bs[kBsAlpha] = al;
  }
  if (stop) return;
  const PT alp = PT(al);
  double ss = 0.0, rs = 0.0;
  RSF_WAVE_LOOP(g, W) {
    const RbWave w = rb_wave(g, W);
    if (!w.live) continue;
    const int q = w.q, qb = nh + q;
    int jn[6];
    bool ok[6];
    rb_black_nbrs(g, w, jn, ok);
    const PT sr = rb_axpy(ldo(r, q), ldo(v, q), -alp);
    PT vb = ldo(Pm, qb) * ldo(ph, qb);
    if (has_extras) vb += ldo(v, qb);
    PT ys = PT(0);
#pragma unroll
    for (int sl = 1; sl <= 6; ++sl) {
      const int cj = jn[sl - 1];
      const PT a = ok[sl - 1] ? ldo(Pm + (size_t)sl * n, qb) : PT(0);
      vb += a * ldo(ph, cj);
      const PT sj = rb_axpy(ldo(r, cj), ldo(v, cj), -alp);
      ys += a * (sj * ldo(d, cj));
    }
    const PT sbv = rb_axpy(ldo(r, qb), vb, -alp);
    sto(v, qb, vb);
    sto(s, q, sr);
    sto(s, qb, sbv);
    ss += double(sr) * double(sr) + double(sbv) * double(sbv);
    rs += double(ldo(rh, q)) * double(sr) + double(ldo(rh, qb)) * double(sbv);
    sto(sh, qb, (sbv - ys) * ldo(d, qb));
  }
  const double pr[2] = {ss, rs};
  block_partials<2>(pr, part2);
}

// E: the black rows of t = P sh (the red half of t holds its red rows), the extras rows
// (natural rows xrows[e] of the wavefront's pairs, e in xwave[W] .. xwave[W+1], summed
// by the whole wavefront; compacted columns xcolc); partials t.t, t.s, rh.t
template <class PT>
__global__ __launch_bounds__(kRBlock, 8) void k_rb_e(Grid g, double* __restrict__ bs,
                                                  const PT* __restrict__ Pm, const PT* __restrict__ sh,
                                                  const PT* __restrict__ s, const PT* __restrict__ rh,
                                                  PT* __restrict__ t, const int* __restrict__ xwave,
                                                  const int* __restrict__ xrows, const int* __restrict__ xptr,
                                                  const int* __restrict__ xcolc, const PT* __restrict__ Pe,
                                                  double* __restrict__ part3) {
  if (rb_over(bs)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) bs[kBsK] += 1.0;  // no block of E reads it
  const int n = g.n, nh = n >> 1, lane = threadIdx.x & 63;
  double tt = 0.0, ts = 0.0, rt = 0.0;
  RSF_WAVE_LOOP(g, W) {
    const RbWave w = rb_wave(g, W);
    const int q = w.q, qb = nh + q;
    double tr = 0.0, tb = 0.0;
    if (w.live) {
      int jn[6];
      bool ok[6];
      rb_black_nbrs(g, w, jn, ok);
      PT tbp = ldo(Pm, qb) * ldo(sh, qb);
#pragma unroll
      for (int sl = 1; sl <= 6; ++sl) {
        const PT a = ok[sl - 1] ? ldo(Pm + (size_t)sl * n, qb) : PT(0);
        tbp += a * ldo(sh, jn[sl - 1]);
      }
      tr = double(ldo(t, q));
      tb = double(tbp);
    }
    bool touched = false;
    if (xwave) {  // every lane of the wavefront takes part
      const int ir = 2 * q + 1 - w.o, ib = 2 * q + w.o;
      for (int e = xwave[W]; e < xwave[W + 1]; ++e) {
        double acc = 0.0;
        for (int kk = xptr[e] + lane; kk < xptr[e + 1]; kk += 64) acc += double(Pe[kk]) * double(sh[xcolc[kk]]);
        acc = __shfl(wave_sum(acc), 0, 64);
        if (w.live && xrows[e] == ir) {
          tr += acc;
          touched = true;
        }
        if (w.live && xrows[e] == ib) tb += acc;
      }
    }
    if (!w.live) continue;
    const PT trp = PT(tr), tbp = PT(tb);
    if (touched) sto(t, q, trp);
    sto(t, qb, tbp);
    const double sr = double(ldo(s, q)), sbv = double(ldo(s, qb));
    tt += double(trp) * double(trp) + double(tbp) * double(tbp);
    ts += double(trp) * sr + double(tbp) * sbv;
    rt += double(ldo(rh, q)) * double(trp) + double(ldo(rh, qb)) * double(tbp);
  }
  const double pr[3] = {tt, ts, rt};
  block_partials<3>(pr, part3);
}

// solution out of the compacted order
template <class PT>
__global__ void k_rb_scatter(Grid g, const PT* __restrict__ xc, PT* __restrict__ x) {
  const int nh = g.n >> 1;
  RSF_WAVE_LOOP(g, W) {
    const RbWave w = rb_wave(g, W);
    if (!w.live) continue;
    const int i0 = 2 * w.q;  // natural index of the pair's first cell
    sto(x, i0 + 1 - w.o, ldo(xc, w.q));
    sto(x, i0 + w.o, ldo(xc, nh + w.q));
  }
}

}  // namespace rsf
