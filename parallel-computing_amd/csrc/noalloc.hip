// noalloc.hip -- the allocation-free dot product (iga_dot_f64_noalloc).
//
// iga_dot_f64 allocates its partial-sum buffer on every call and returns the
// value on the host.  Callers that cannot allocate or synchronise (code under
// HIP graph capture, tight serving loops) use this form instead: its partial
// sums live in `scratch`, a file-scope device array of fixed size, and the
// result is written to device memory on the caller's stream.
//
// `scratch` is a host-side stand-in in host code, not a device address: its
// device address is resolved once per device with hipGetSymbolAddress (with
// that device current) and only that address is passed to kernels.  Every
// call on one GPU shares that GPU's array, so calls that overlap on one GPU
// race; each call writes every partial it folds before folding it, so nothing
// has to start at zero.
#include <hip/hip_runtime.h>

#include <mutex>

#include "iga_mf.h"

namespace {
constexpr int NA_RB = 256;               // threads per workgroup
constexpr int NA_U = 4;                  // values per thread
constexpr int NA_PER_WG = NA_RB * NA_U;  // values per partial sum (1,024)
constexpr int NA_SLOTS = 65536;          // partial sums the scratch array holds
constexpr int64_t NA_MAX_N = (int64_t)NA_SLOTS * NA_PER_WG;  // 67,108,864
constexpr int NA_MAXDEV = 64;
}  // namespace

// 512 KiB per GPU, held until the process ends
static __device__ __align__(256) double scratch[NA_SLOTS];

namespace {
std::mutex na_mu;
double *na_ptr[NA_MAXDEV] = {};  // device address of `scratch` per device

// Device address of `scratch` on the current device, resolved on the first
// call per device.
int scratch_on_current_device(double **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return IGA_ERR_HIP;
  if (dev < 0 || dev >= NA_MAXDEV) return IGA_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(na_mu);
  if (!na_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return IGA_ERR_HIP;
    }
    na_ptr[dev] = static_cast<double *>(p);
  }
  *out = na_ptr[dev];
  return IGA_OK;
}

// Sum over the workgroup in a fixed order (reproducible), valid in thread 0.
__device__ __forceinline__ double na_block_sum(double v) {
  __shared__ double ws[NA_RB / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < NA_RB / 64; ++w) s += ws[w];
  }
  return s;
}

// part[b] = sum of a[i] b[i] over the b-th block of NA_PER_WG values
__global__ __launch_bounds__(NA_RB) void k_na_dot_part(const double *__restrict__ a,
                                                       const double *__restrict__ b, int64_t n,
                                                       double *__restrict__ part) {
  const int64_t base = (int64_t)blockIdx.x * NA_PER_WG + threadIdx.x;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < NA_U; ++k) {
    const int64_t i = base + (int64_t)k * NA_RB;
    if (i < n) s += a[i] * b[i];
  }
  s = na_block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// *out = sum of part[0..nb) (one workgroup, fixed order)
__global__ __launch_bounds__(NA_RB) void k_na_dot_fold(const double *__restrict__ part, int nb,
                                                       double *__restrict__ out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += NA_RB) s += part[i];
  s = na_block_sum(s);
  if (threadIdx.x == 0) *out = s;
}
}  // namespace

extern "C" int iga_dot_f64_noalloc(const double *a, const double *b, int64_t n, double *out,
                                   void *stream) {
  if (!out || n < 0 || (n > 0 && (!a || !b))) return IGA_ERR_INVALID_ARGUMENT;
  if (n > NA_MAX_N) return IGA_ERR_SIZE_LIMIT;
  double *part = nullptr;
  const int rc = scratch_on_current_device(&part);
  if (rc) return rc;
  const int nb = n > 0 ? (int)((n + NA_PER_WG - 1) / NA_PER_WG) : 1;  // <= NA_SLOTS
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_na_dot_part, dim3(nb), dim3(NA_RB), 0, st, a, b, n, part);
  hipLaunchKernelGGL(k_na_dot_fold, dim3(1), dim3(NA_RB), 0, st, part, nb, out);
  return hipGetLastError() == hipSuccess ? IGA_OK : IGA_ERR_HIP;
}
