// Dense routes of the linear stack for small systems, as the reference takes them:
// - a direct solve when the full system has at most 512 rows (rs/fim/linear/mod.rs:40,
//   313-335; the reference factors the full system with a sparse LU, faer, falling back to
//   the iterative solver when that fails). Here the well tail is eliminated in closed form
//   first (the Schur complement is exact), the reduced cell system is solved densely and the
//   tail is recovered, which is the same solution to rounding;
// - the exact inverse of the CPR coarse pressure matrix when it has at most 299 rows
//   (rs/fim/linear/gmres_block_jacobi.rs:24, :903-917 invert_pressure_block, applied as a
//   dense product at :585-610).
// Both run Gauss-Jordan elimination with partial pivoting on an augmented matrix [M | R]
// (row-major, `cols` columns) in one workgroup: row c is the pivot row of column c, the
// largest |M[r][c]| over r >= c wins (the first one among equals), every other row has f
// times the scaled pivot row subtracted. At the end R holds M^-1 R. A pivot that is not
// above `tiny`, or not finite, marks the matrix singular in *flag.
#pragma once
#include <hip/hip_runtime.h>

#include "rsf_kernels.hpp"

namespace rsf {

constexpr int kDenseBlock = 1024;
constexpr int kDenseMaxRows = 512;   // direct route (rows of the full system)
constexpr int kCoarseDenseMaxRows = 299;

constexpr int kDenseMaxCols = 2 * kCoarseDenseMaxRows;  // widest augmented matrix (598 > 513)
constexpr int kDenseBatch = 8;                          // row updates in flight per thread

// Lane masks of a wavefront: bits below this lane.
__device__ __forceinline__ unsigned long long lanes_below(int lane) {
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// One pivot step c: (1) the pivot search over column c, carrying each candidate's signed
// value; (2) every thread picks the same row p from the wavefronts' candidates (largest
// |M[r][c]|, the first row among equals) with 1 / M[p][c]; (3) one pass swaps rows c and
// p, scales the new pivot row from column c on and copies it to LDS, and reads the factors
// f_r = M[r][c] of the other rows (row p's is the old diagonal); (4) the rows with
// f_r != 0 and the pivot row's nonzero columns are compacted; (5) every (row, column) pair
// of the two lists gets M[r][k] -= f_r * pivot[k], all loads of a batch issued before any
// store. An entry the lists leave out has f_r = 0 or pivot[k] = 0 with f_r finite, so the
// plain elimination would leave it as it is (up to the sign of a zero): the result is the
// plain elimination's. A non-finite factor keeps every column. One thread per row and
// per column: rows <= kDenseMaxRows and columns <= kDenseMaxCols, both within the block.
static_assert(kDenseMaxRows <= kDenseBlock && kDenseMaxCols <= kDenseBlock, "one thread per row and column");
__global__ __launch_bounds__(kDenseBlock) void k_gauss_jordan(int n, int cols, double* __restrict__ M, double tiny,
                                                               int* __restrict__ flag) {
  __shared__ double colc[kDenseMaxRows];
  __shared__ double prow[kDenseMaxCols];
  __shared__ int rlist[kDenseMaxRows], klist[kDenseMaxCols];
  __shared__ int wcr[kDenseBlock / 64], wck[kDenseBlock / 64];
  __shared__ double bv[kDenseBlock / 64], bs[kDenseBlock / 64];
  __shared__ int bi[kDenseBlock / 64];
  __shared__ double s_d;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  constexpr int nwv = kDenseBlock / 64;
  bool bad = false;
  for (int c = 0; c < n; ++c) {
    // (1) pivot search: NaN never wins; equal magnitudes keep the smallest row
    double best = -1.0, sbest = 0.0;
    int bidx = n;
    if (c + tid < n) {
      const double v = M[(size_t)(c + tid) * cols + c];
      if (tid == 0) s_d = v;  // the diagonal before the swap
      if (fabs(v) > best) { best = fabs(v); sbest = v; bidx = c + tid; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_down(best, o, 64), os = __shfl_down(sbest, o, 64);
      const int oi = __shfl_down(bidx, o, 64);
      if (ov > best || (ov == best && oi < bidx)) { best = ov; sbest = os; bidx = oi; }
    }
    if (lane == 0) { bv[wv] = best; bs[wv] = sbest; bi[wv] = bidx; }
    __syncthreads();
    // (2) the same decision in every thread
    double b = -1.0, pv0 = 0.0;
    int p = n;
#pragma unroll
    for (int w = 0; w < nwv; ++w)
      if (bv[w] > b || (bv[w] == b && bi[w] < p)) { b = bv[w]; pv0 = bs[w]; p = bi[w]; }
    const double d = s_d;
    // a NaN on the diagonal leaves the pivot undefined: singular, as a NaN best is
    if (d != d || !(b > tiny) || !isfinite(b) || p >= n) { bad = true; break; }
    const int width = cols - c;
    const double dinv = 1.0 / pv0;
    // (3) rows c and p swapped over every column, the new row c scaled from column c on
    bool kact = false;
    if (tid < cols) {
      const int k = tid;
      const double a = M[(size_t)c * cols + k], bk = M[(size_t)p * cols + k];
      const double nv = k >= c ? bk * dinv : bk;
      M[(size_t)c * cols + k] = nv;
      if (p != c) M[(size_t)p * cols + k] = a;
      if (k >= c) {
        prow[k - c] = nv;
        kact = nv != 0.0;
      }
    }
    double f = 0.0;
    bool ract = false;
    if (tid < n && tid != c) {
      f = tid == p ? d : M[(size_t)tid * cols + c];  // row p now holds the old row c
      colc[tid] = f;
      ract = f != 0.0;
    }
    // (4) rows with a nonzero factor, nonzero pivot-row columns
    const unsigned long long rm = __ballot(ract), km = __ballot(kact);
    if (lane == 0) {
      wcr[wv] = __popcll(rm);
      wck[wv] = __popcll(km);
    }
    const int anybad = __syncthreads_or(ract && !isfinite(f));
    int ro = 0, ko = 0, nr = 0, nk = 0;
#pragma unroll
    for (int w = 0; w < nwv; ++w) {
      ro += w < wv ? wcr[w] : 0;
      ko += w < wv ? wck[w] : 0;
      nr += wcr[w];
      nk += wck[w];
    }
    if (ract) rlist[ro + __popcll(rm & lanes_below(lane))] = tid;
    if (anybad) {
      if (tid >= c && tid < cols) klist[tid - c] = tid - c;
      nk = width;
    } else if (kact) {
      klist[ko + __popcll(km & lanes_below(lane))] = tid - c;
    }
    __syncthreads();
    // (5) the updates, kDenseBatch per thread in flight
    const int total = nr * nk;
    for (int base = tid; base < total; base += kDenseBlock * kDenseBatch) {
      int off[kDenseBatch];
      double fr[kDenseBatch], pv[kDenseBatch], v[kDenseBatch];
#pragma unroll
      for (int u = 0; u < kDenseBatch; ++u) {
        const int idx = base + u * kDenseBlock;
        const bool in = idx < total;
        const int ri = in ? idx / nk : 0, ki = in ? idx - ri * nk : 0;
        const int r = in ? rlist[ri] : 0, kk = in ? klist[ki] : 0;
        off[u] = in ? r * cols + c + kk : -1;
        fr[u] = colc[r];
        pv[u] = prow[kk];
        v[u] = M[in ? off[u] : 0];  // unconditional load (in range) so all stay in flight
      }
#pragma unroll
      for (int u = 0; u < kDenseBatch; ++u)
        if (off[u] >= 0) M[off[u]] = v[u] - fr[u] * pv[u];
    }
    __syncthreads();
  }
  if (tid == 0) *flag = bad ? 1 : 0;
}

// [P | I] of the CPR coarse matrix, n x 2n: the seven stencil planes (row i at pslot) and
// the wells' clique entries
template <class PT>
__global__ void k_pdense_fill(Grid g, const PT* __restrict__ Pm, int n_extra_rows, const int* __restrict__ xrows,
                              const int* __restrict__ xptr, const int* __restrict__ xcol, const PT* __restrict__ Pe,
                              double* __restrict__ M) {
  const int n = g.n, cols = 2 * n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* row = M + (size_t)i * cols;
  for (int k = 0; k < cols; ++k) row[k] = 0.0;
  int cx, cy, cz;
  cell_ijk(g, i, cx, cy, cz);
  const int pi = pslot(g, i, cx, cy, cz);
  row[i] = double(Pm[pi]);
  for (int s = 1; s < 7; ++s) {
    const int j = slot_nbr(g, i, s, cx, cy, cz);
    if (j >= 0) row[j] += double(Pm[(size_t)s * n + pi]);
  }
  for (int t = 0; t < n_extra_rows; ++t)
    if (xrows[t] == i)
      for (int k = xptr[t]; k < xptr[t + 1]; ++k) row[xcol[k]] += double(Pe[k]);
  row[n + i] = 1.0;
}

// xp = P^-1 rp with the inverse in columns n..2n-1 of M (row stride 2n); one wavefront per row
template <class PT>
__global__ void k_pdense_apply(int n, const double* __restrict__ M, const PT* __restrict__ rp, PT* __restrict__ xp) {
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (row >= n) return;
  const double* inv = M + (size_t)row * 2 * n + n;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += inv[j] * double(rp[j]);
  s = wave_sum(s);
  if (lane == 0) xp[row] = PT(s);
}

// [A | b] of the reduced cell system, 3n x (3n + 1), unknown 3 i + c: the stencil blocks
// (natural layout), the wells' clique blocks (9 planes of n_extra entries) and b (c*n + i)
__global__ void k_dense_system(Grid g, const double* __restrict__ A, int n_extra_rows, const int* __restrict__ xrows,
                               const int* __restrict__ xptr, const int* __restrict__ xcol,
                               const double* __restrict__ E, int n_extra, const double* __restrict__ b,
                               double* __restrict__ M) {
  const int n = g.n, m = 3 * n, cols = m + 1;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;  // q = 3 i + c
  if (q >= m) return;
  const int i = q / 3, c = q - 3 * i;
  double* row = M + (size_t)q * cols;
  for (int k = 0; k < cols; ++k) row[k] = 0.0;
  int cx, cy, cz;
  cell_ijk(g, i, cx, cy, cz);
  for (int s = 0; s < 7; ++s) {
    const int j = s == 0 ? i : slot_nbr(g, i, s, cx, cy, cz);
    if (j < 0) continue;
    for (int v = 0; v < 3; ++v) row[3 * j + v] += A[(size_t)(s * 9 + c * 3 + v) * n + i];
  }
  for (int t = 0; t < n_extra_rows; ++t)
    if (xrows[t] == i)
      for (int k = xptr[t]; k < xptr[t + 1]; ++k)
        for (int v = 0; v < 3; ++v) row[3 * xcol[k] + v] += E[(size_t)(c * 3 + v) * n_extra + k];
  row[m] = b[(size_t)c * n + i];
}

// the solution column of a reduced [A | b] after elimination, into x (c*n + i)
__global__ void k_dense_solution(int n, const double* __restrict__ M, double* __restrict__ x) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = 3 * n;
  if (q >= m) return;
  const int i = q / 3, c = q - 3 * i;
  x[(size_t)c * n + i] = M[(size_t)q * (m + 1) + m];
}

// Block partials of |b - y|^2 (part[blk]) and |b|^2 (part[stride + blk]): the true residual
// of a pressure solve, y = P x (rs/fim/linear/gmres_block_jacobi.rs:584-609)
template <class T>
__global__ __launch_bounds__(kBlock) void k_resid_norms_partial(int n, const T* __restrict__ b, const T* __restrict__ y,
                                                                double* __restrict__ part, int stride) {
  __shared__ double s1[kBlock], s2[kBlock];
  double a1 = 0.0, a2 = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double bi = double(b[i]), d = bi - double(y[i]);
    a1 += d * d;
    a2 += bi * bi;
  }
  s1[threadIdx.x] = a1;
  s2[threadIdx.x] = a2;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1[0];
    part[stride + blockIdx.x] = s2[0];
  }
}

}  // namespace rsf
