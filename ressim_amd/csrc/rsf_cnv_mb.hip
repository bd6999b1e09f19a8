// OPM CNV/MB acceptance measures of a Newton iterate on device arrays:
// cnv_mb_from_parts (rs/fim/newton/convergence.rs:478-612), as entry points that take the
// caller's device pointers and stream.
//
// Two passes over the cells, each a grid-stride loop over at most kCmMaxBlocks blocks:
//   k_cm_partial    per block: max |r_c|/pv, sum r_c, sum B_c (c = water, oil, gas), sum pv
//   k_cm_reduce     one block: the block records in a fixed tree order; b_avg = sum B / n
//   k_cm_violation  per block: pore volume of the cells whose max_c |r_c| b_avg_c / pv > CNV tol
//   k_cm_finish     one block: CNV, MB, violating fraction, the relaxation rule, the verdicts
// The block records are the scratch memory of the path: rsf_cnv_mb allocates them for the
// call; rsf_cnv_mb_noalloc uses a file-scope device array instead. Both launch the same
// kernels on the same grid, so both give the same bits.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/ressim_fim.h"

namespace {

constexpr int kCmBlock = 256;
constexpr int kCmMaxBlocks = 1024;
constexpr int kCmPart = 10;  // max coeff[3], r sums[3], FVF sums[3], pv sum
constexpr int kCmStage = 16; // reduced values [10], b_avg [3]
constexpr long kCmMaxCells = 1L << 28;
// scratch layout (doubles): block records, reduced values, violation records
constexpr int kCmScratchDoubles = kCmMaxBlocks * kCmPart + kCmStage + kCmMaxBlocks;

// rs/fim/newton/convergence.rs:352-356
constexpr double kTolCnv = 1e-2, kTolCnvRelaxed = 1.0, kTolMb = 1e-7, kTolMbRelaxed = 1e-6, kRelaxedPvFrac = 0.03;

}  // namespace

// The scratch of rsf_cnv_mb_noalloc: one copy per device, alive until the process ends.
// Host code never passes this name to a kernel: it takes the device address from
// hipGetSymbolAddress once per device (cm_scratch_address) and passes that.
static __device__ __align__(256) double scratch[kCmScratchDoubles];

namespace {

// larger of a and b with std::max's NaN behaviour: a NaN b never replaces a
__device__ __forceinline__ double cm_max(double a, double b) { return a < b ? b : a; }

// Sum, or maximum, of r[stride*b + k] over b < nb: strided per thread, then a fixed tree.
template <bool kMax>
__device__ double cm_records(const double* __restrict__ r, int stride, int k, int nb, double* sm) {
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += kCmBlock) {
    const double x = r[(long)stride * b + k];
    v = kMax ? cm_max(v, x) : v + x;
  }
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int o = kCmBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] = kMax ? cm_max(sm[threadIdx.x], sm[threadIdx.x + o]) : sm[threadIdx.x] + sm[threadIdx.x + o];
    __syncthreads();
  }
  const double out = sm[0];
  __syncthreads();
  return out;
}

// residual is the reference's cell rows, index 3 i + c; fvf is [B_w, B_o, B_g] per cell
__global__ __launch_bounds__(kCmBlock) void k_cm_partial(int n, const double* __restrict__ R, const double* __restrict__ pv,
                                                         const double* __restrict__ fvf, double* __restrict__ part) {
  __shared__ double sm[kCmPart][kCmBlock];
  double v[kCmPart];
  for (int k = 0; k < kCmPart; ++k) v[k] = 0.0;
  for (long i = (long)blockIdx.x * kCmBlock + threadIdx.x; i < n; i += (long)gridDim.x * kCmBlock) {
    const double p = cm_max(pv[i], 1e-9);
    for (int c = 0; c < 3; ++c) {
      const double r = R[3 * i + c];
      v[c] = cm_max(v[c], fabs(r) / p);
      v[3 + c] += r;
      v[6 + c] += fvf[3 * i + c];
    }
    v[9] += p;
  }
  for (int k = 0; k < kCmPart; ++k) sm[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int o = kCmBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < kCmPart; ++k) {
        const double a = sm[k][threadIdx.x], b = sm[k][threadIdx.x + o];
        sm[k][threadIdx.x] = k < 3 ? cm_max(a, b) : a + b;
      }
    __syncthreads();
  }
  if (threadIdx.x < kCmPart) part[blockIdx.x * kCmPart + threadIdx.x] = sm[threadIdx.x][0];
}

__global__ __launch_bounds__(kCmBlock) void k_cm_reduce(int nb, int n, const double* __restrict__ part,
                                                        double* __restrict__ stage) {
  __shared__ double sm[kCmBlock];
  for (int k = 0; k < kCmPart; ++k) {
    const double v = k < 3 ? cm_records<true>(part, kCmPart, k, nb, sm) : cm_records<false>(part, kCmPart, k, nb, sm);
    if (threadIdx.x == 0) {
      stage[k] = v;
      if (k >= 6 && k < 9) stage[kCmPart + k - 6] = v / (double)(n > 1 ? n : 1);
    }
  }
}

__global__ __launch_bounds__(kCmBlock) void k_cm_violation(int n, const double* __restrict__ R, const double* __restrict__ pv,
                                                           const double* __restrict__ stage, double* __restrict__ vpart) {
  __shared__ double sm[kCmBlock];
  const double b0 = stage[kCmPart], b1 = stage[kCmPart + 1], b2 = stage[kCmPart + 2];
  double acc = 0.0;
  for (long i = (long)blockIdx.x * kCmBlock + threadIdx.x; i < n; i += (long)gridDim.x * kCmBlock) {
    const double p = cm_max(pv[i], 1e-9);
    // f64::max ignores a NaN operand
    const double x0 = fabs(R[3 * i]) * b0 / p, x1 = fabs(R[3 * i + 1]) * b1 / p, x2 = fabs(R[3 * i + 2]) * b2 / p;
    const double cm = fmax(fmax(fmax(0.0, x0), x1), x2);
    if (cm > kTolCnv) acc += p;
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kCmBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) vpart[blockIdx.x] = sm[0];
}

// out: cnv[3], mb[3], violating pore-volume fraction, the 3 %-PV rule (0/1),
// would_accept (0/1), would_accept_strict (0/1)
__global__ __launch_bounds__(kCmBlock) void k_cm_finish(int nb, int relax_final, const double* __restrict__ stage,
                                                        const double* __restrict__ vpart, double* __restrict__ out) {
  __shared__ double sm[kCmBlock];
  const double viol = cm_records<false>(vpart, 1, 0, nb, sm);
  if (threadIdx.x != 0) return;
  const double pvsum = stage[9];
  const double pvd = cm_max(pvsum, 1e-9);
  const bool relaxes = viol < kRelaxedPvFrac * pvsum;
  const double cnv_tol = relax_final ? kTolCnvRelaxed : kTolCnv, mb_tol = relax_final ? kTolMbRelaxed : kTolMb;
  bool ok = true, strict = true;
  for (int c = 0; c < 3; ++c) {
    const double b = stage[kCmPart + c];
    const double cnv = b * stage[c];
    const double mb = fabs(b * stage[3 + c]) / pvd;
    out[c] = cnv;
    out[3 + c] = mb;
    ok = ok && (cnv <= cnv_tol || (relaxes && cnv <= kTolCnvRelaxed)) && mb <= mb_tol;
    strict = strict && cnv <= kTolCnv && mb <= kTolMb;
  }
  out[6] = viol / pvd;
  out[7] = relaxes ? 1.0 : 0.0;
  out[8] = ok ? 1.0 : 0.0;
  out[9] = strict ? 1.0 : 0.0;
}

thread_local std::string g_cm_err;

int cm_fail(int rc, const std::string& msg) {
  g_cm_err = msg;
  return rc;
}

bool on_device(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

int cm_check(const double* R, const double* pv, const double* fvf, long n, const double* out, long max_cells) {
  if (n < 0 || n > max_cells)
    return cm_fail(RSF_ERR_INVALID, "cnv_mb: n = " + std::to_string(n) + " cells is outside [0, " +
                                        std::to_string(max_cells) + "]");
  if (!out || !on_device(out)) return cm_fail(RSF_ERR_INVALID, "cnv_mb: out must be a device pointer");
  if (n > 0 && (!R || !pv || !fvf || !on_device(R) || !on_device(pv) || !on_device(fvf)))
    return cm_fail(RSF_ERR_INVALID, "cnv_mb: residual, pore volumes and FVFs must be device pointers");
  return RSF_OK;
}

// Enqueues the four kernels with `buf` (kCmScratchDoubles doubles of device memory) as
// their records. Every record is written before it is read, so nothing is cleared first.
int cm_launch(const double* R, const double* pv, const double* fvf, int n, int relax_final, double* out, double* buf,
              hipStream_t st) {
  const int nb = n > 0 ? (int)std::min<long>(((long)n + kCmBlock - 1) / kCmBlock, kCmMaxBlocks) : 1;
  double* part = buf;
  double* stage = buf + kCmMaxBlocks * kCmPart;
  double* vpart = stage + kCmStage;
  hipLaunchKernelGGL(k_cm_partial, dim3(nb), dim3(kCmBlock), 0, st, n, R, pv, fvf, part);
  hipLaunchKernelGGL(k_cm_reduce, dim3(1), dim3(kCmBlock), 0, st, nb, n, (const double*)part, stage);
  hipLaunchKernelGGL(k_cm_violation, dim3(nb), dim3(kCmBlock), 0, st, n, R, pv, (const double*)stage, vpart);
  hipLaunchKernelGGL(k_cm_finish, dim3(1), dim3(kCmBlock), 0, st, nb, relax_final, (const double*)stage,
                     (const double*)vpart, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cm_fail(RSF_ERR_HIP, std::string("cnv_mb: launch failed: ") + hipGetErrorString(e));
  return RSF_OK;
}

// Device address of `scratch` on the current device, looked up once per device.
constexpr int kCmMaxDevices = 64;
std::mutex g_cm_mu;
double* g_cm_scratch[kCmMaxDevices] = {};

int cm_scratch_address(double** ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return cm_fail(RSF_ERR_HIP, std::string("cnv_mb: hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kCmMaxDevices) return cm_fail(RSF_ERR_HIP, "cnv_mb: device index out of range");
  std::lock_guard<std::mutex> lock(g_cm_mu);
  if (!g_cm_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return cm_fail(RSF_ERR_HIP, std::string("cnv_mb: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_cm_scratch[dev] = static_cast<double*>(p);
  }
  *ptr = g_cm_scratch[dev];
  return RSF_OK;
}

}  // namespace

extern "C" {

int rsf_cnv_mb(const double* residual, const double* pore_volumes, const double* fvf, int n, int relax_final,
               double* out, void* stream) {
  int rc = cm_check(residual, pore_volumes, fvf, n, out, kCmMaxCells);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double* buf = nullptr;
  hipError_t e = hipMalloc(&buf, sizeof(double) * kCmScratchDoubles);
  if (e != hipSuccess) return cm_fail(RSF_ERR_HIP, std::string("cnv_mb: hipMalloc: ") + hipGetErrorString(e));
  rc = cm_launch(residual, pore_volumes, fvf, n, relax_final, out, buf, st);
  e = hipStreamSynchronize(st);
  (void)hipFree(buf);
  if (!rc && e != hipSuccess) rc = cm_fail(RSF_ERR_HIP, std::string("cnv_mb: ") + hipGetErrorString(e));
  return rc;
}

int rsf_cnv_mb_noalloc(const double* residual, const double* pore_volumes, const double* fvf, int n, int relax_final,
                       double* out, void* stream) {
  int rc = cm_check(residual, pore_volumes, fvf, n, out, RSF_CNV_MB_NOALLOC_MAX_CELLS);
  if (rc) return rc;
  double* buf = nullptr;
  rc = cm_scratch_address(&buf);
  if (rc) return rc;
  return cm_launch(residual, pore_volumes, fvf, n, relax_final, out, buf, static_cast<hipStream_t>(stream));
}

int rsf_cnv_mb_last_error(char* buf, int len) {
  if (!buf || len <= 0) return RSF_ERR_INVALID;
  std::snprintf(buf, len, "%s", g_cm_err.c_str());
  return RSF_OK;
}

}  // extern "C"
