// Correlation energy on caller-owned device buffers without any allocation:
// the variant of ccgpu_energy (ccd_energy_save, ccm_energy.f90:4-71) for
// callers that cannot allocate device memory (code under hipGraph capture,
// serving loops).  Its block partials live in a file-scope device array,
// `scratch`, whose address is taken once per device with
// hipGetSymbolAddress; kernels receive that address as an argument.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/ccgpu.h"

namespace {

// Block partials of one call: at most kMaxParts complex partial sums, each
// over a fixed span of kSpan amplitudes (so at most kMaxParts * kSpan
// amplitudes per call).  Every call on a GPU shares that GPU's array: calls
// that overlap on one GPU race.  Every partial a call reads was written by
// the same call, so nothing has to be cleared.
constexpr int kMaxParts = 4096;
constexpr int64_t kSpan = 65536;
__device__ __align__(256) double2 scratch[kMaxParts];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double2 *g_scratch[kMaxDevices] = {};   // device address of `scratch` per device

// Partial sum of t * conj(r0) over one span per block, pairwise tree in the
// block (fixed order, so repeated calls give the same bits).
__global__ __launch_bounds__(256) void k_energy_spans(const double2 *__restrict__ t, const double2 *__restrict__ r0,
                                                      int64_t n, double2 *__restrict__ part) {
  const int64_t b = (int64_t)blockIdx.x * kSpan;
  const int64_t e = b + kSpan < n ? b + kSpan : n;
  double re = 0.0, im = 0.0;
  for (int64_t u = b + threadIdx.x; u < e; u += 256) {
    const double2 x = t[u], y = r0[u];
    re += x.x * y.x + x.y * y.y;
    im += x.y * y.x - x.x * y.y;
  }
  __shared__ double sr[256], si[256];
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      si[threadIdx.x] += si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sr[0], si[0]);
}

// Fixed-order sum of the partials into the caller's output.
__global__ __launch_bounds__(256) void k_energy_final(const double2 *__restrict__ part, int nparts,
                                                      double2 *__restrict__ out) {
  double re = 0.0, im = 0.0;
  for (int p = threadIdx.x; p < nparts; p += 256) {
    re += part[p].x;
    im += part[p].y;
  }
  __shared__ double sr[256], si[256];
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      si[threadIdx.x] += si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = make_double2(sr[0], si[0]);
}

// Device address of this GPU's `scratch`, taken once per device with that
// device current (the array's name is only a host-side stand-in).
double2 *scratch_of_current_device() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) throw std::string("hipGetDevice failed: ") + hipGetErrorString(e);
  if (dev < 0 || dev >= kMaxDevices) throw std::string("device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) throw std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
    g_scratch[dev] = (double2 *)p;
  }
  return g_scratch[dev];
}

}  // namespace

namespace ccgpu {
void set_last_error(const std::string &msg);   // ccgpu_last_error's message (api.cpp)
}

extern "C" {

int ccgpu_energy_noalloc(long long n, const double *t, const double *r0, double *ecorr, void *stream) {
  try {
    if (n < 0) throw std::string("negative amplitude count");
    if (n > (long long)kMaxParts * kSpan)
      throw std::string("amplitude count ") + std::to_string(n) + " above the maximum " +
          std::to_string((long long)kMaxParts * kSpan) + " of ccgpu_energy_noalloc";
    if (!ecorr || (n > 0 && (!t || !r0))) throw std::string("null device pointer");
    double2 *part = scratch_of_current_device();
    const hipStream_t st = (hipStream_t)stream;
    const int nparts = n == 0 ? 0 : (int)((n + kSpan - 1) / kSpan);
    if (nparts > 0)
      hipLaunchKernelGGL(k_energy_spans, dim3(nparts), dim3(256), 0, st, (const double2 *)t, (const double2 *)r0,
                         (int64_t)n, part);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, st, part, nparts, (double2 *)ecorr);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::string("kernel launch failed: ") + hipGetErrorString(e);
    return 0;
  } catch (const std::string &m) {
    ccgpu::set_last_error(m);
    return 1;
  }
}

}  // extern "C"
