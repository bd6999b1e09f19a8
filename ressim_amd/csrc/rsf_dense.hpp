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

__global__ __launch_bounds__(kDenseBlock) void k_gauss_jordan(int n, int cols, double* __restrict__ M, double tiny,
                                                               int* __restrict__ flag) {
  __shared__ double colc[kDenseMaxRows];
  __shared__ double bv[kDenseBlock / 64];
  __shared__ int bi[kDenseBlock / 64];
  __shared__ int s_piv, s_bad;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) s_bad = 0;
  __syncthreads();
  for (int c = 0; c < n; ++c) {
    // pivot search: NaN never wins; equal magnitudes keep the smallest row
    double best = -1.0;
    int bidx = n;
    for (int r = c + tid; r < n; r += kDenseBlock) {
      const double v = fabs(M[(size_t)r * cols + c]);
      if (v > best) { best = v; bidx = r; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_down(best, o, 64);
      const int oi = __shfl_down(bidx, o, 64);
      if (ov > best || (ov == best && oi < bidx)) { best = ov; bidx = oi; }
    }
    if (lane == 0) { bv[wv] = best; bi[wv] = bidx; }
    __syncthreads();
    if (tid == 0) {
      double b = -1.0;
      int p = n;
      for (int w = 0; w < kDenseBlock / 64; ++w)
        if (bv[w] > b || (bv[w] == b && bi[w] < p)) { b = bv[w]; p = bi[w]; }
      const double d = M[(size_t)c * cols + c];
      // a NaN on the diagonal leaves the pivot undefined: singular, as a NaN best is
      if (d != d || !(b > tiny) || !isfinite(b) || p >= n) s_bad = 1;
      s_piv = p;
    }
    __syncthreads();
    if (s_bad) break;
    const int p = s_piv;
    if (p != c)
      for (int k = tid; k < cols; k += kDenseBlock) {
        const double a = M[(size_t)c * cols + k];
        M[(size_t)c * cols + k] = M[(size_t)p * cols + k];
        M[(size_t)p * cols + k] = a;
      }
    __syncthreads();
    const double dinv = 1.0 / M[(size_t)c * cols + c];
    __syncthreads();
    // columns left of c are already zero in the pivot row: they change nothing below
    for (int k = c + tid; k < cols; k += kDenseBlock) M[(size_t)c * cols + k] *= dinv;
    for (int r = tid; r < n; r += kDenseBlock) colc[r] = M[(size_t)r * cols + c];
    __syncthreads();
    // every row but c: row -= f * (scaled pivot row); one wavefront per row
    for (int r = wv; r < n; r += kDenseBlock / 64) {
      const double f = colc[r];
      if (r == c || f == 0.0) continue;
      double* dst = M + (size_t)r * cols;
      const double* src = M + (size_t)c * cols;
      for (int k = c + lane; k < cols; k += 64) dst[k] -= f * src[k];
    }
    __syncthreads();
  }
  if (tid == 0) *flag = s_bad;
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
