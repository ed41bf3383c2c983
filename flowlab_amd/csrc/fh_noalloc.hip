// fh_solve_pressure_rbsor_noalloc: the red-black SOR pressure solve of
// fh_solve_pressure (pressureRedBlack.ts:88-113) on caller-owned device
// arrays, for callers that cannot allocate device memory.  Everything beyond
// the caller's p, div and label (the FP32 iterate, the FP32 right-hand side,
// the norm partials and sums) lives in one file-scope device array,
// `scratch`, of fixed size; each GPU holds its own copy until the process
// ends.  Nothing here calls hipMalloc, hipFree or any allocating routine.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/flowhip.h"
#include "fh_device.hpp"
#include "fh_launch.hpp"

using namespace fh;

namespace {

// Largest grid the scratch array holds: nx * ny <= kNoAllocMaxCells, and
// each side within the launch grids of the sweep and reduction kernels.
constexpr long kNoAllocMaxCells = 1L << 20;
constexpr int kNoAllocMaxSide = 1 << 16;
// the reduction grids' cap of 65536 blocks, 4 slots a block, and the fold area
constexpr size_t kPartialsDoubles = 4 * (65536 + kFoldBlocks);

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kOffX = 0;
constexpr size_t kOffB = kOffX + align256(kNoAllocMaxCells * sizeof(float));
constexpr size_t kOffPartials = kOffB + align256(kNoAllocMaxCells * sizeof(float));
constexpr size_t kOffSums = kOffPartials + align256(kPartialsDoubles * sizeof(double));
constexpr size_t kScratchBytes = kOffSums + align256(8 * sizeof(double));

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

thread_local std::string g_err;

int fail(int status, const std::string& msg) {
  g_err = msg;
  return status;
}

int fail_hip(hipError_t e, const char* what) {
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  g_err = buf;
  return (e == hipErrorIllegalAddress || e == hipErrorLaunchFailure) ? FH_ERR_DEVICE_LOST
                                                                     : FH_ERR_HIP;
}

#define NCHK(call)                                     \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return fail_hip(_e, #call);  \
  } while (0)

__global__ void k_zero_f64(double* __restrict__ a, long n) {
  long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k < n) a[k] = 0.0;
}

// Reads the 4 norm sums into host memory (a 64-byte copy into the caller's
// stack: nothing allocated) after the stream has reached them.
int read_sums(hipStream_t s, const double* d_sums, double* h) {
  NCHK(hipMemcpyAsync(h, d_sums, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  NCHK(hipStreamSynchronize(s));
  return FH_OK;
}

}  // namespace

// The scratch memory of this entry point on each device (~10 MB).  In host
// code `scratch` is only a stand-in for the symbol: its device address comes
// from hipGetSymbolAddress, once per device, and only that address is used.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// The device address of this device's scratch array (looked up on the first
// call on each device, with that device current).
int scratch_base(unsigned char** out) {
  int dev = 0;
  NCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(FH_ERR_BAD_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    NCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(FH_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch[dev] = ptr;
  }
  *out = static_cast<unsigned char*>(g_scratch[dev]);
  return FH_OK;
}

}  // namespace

extern "C" {

const char* fh_noalloc_last_error(void) { return g_err.c_str(); }

long long fh_noalloc_max_cells(void) { return kNoAllocMaxCells; }

int fh_solve_pressure_rbsor_noalloc(void* stream, double* p, const double* div,
                                    const uint8_t* label, int nx, int ny, double scale,
                                    int iterations, double omega, double tol, int* out_iters) {
  if (!p || !div || !label || !out_iters)
    return fail(FH_ERR_BAD_ARG, "solvePressureNoAlloc: null argument");
  if (nx < 2 || ny < 2)
    return fail(FH_ERR_BAD_ARG, "solvePressureNoAlloc: need nx, ny >= 2");
  if ((long)nx * ny > kNoAllocMaxCells || nx > kNoAllocMaxSide || ny > kNoAllocMaxSide) {
    char buf[160];
    snprintf(buf, sizeof buf,
             "solvePressureNoAlloc: %dx%d exceeds the scratch capacity (at most %ld cells, "
             "%d a side)",
             nx, ny, kNoAllocMaxCells, kNoAllocMaxSide);
    return fail(FH_ERR_BAD_ARG, buf);
  }
  if (iterations < 0) return fail(FH_ERR_BAD_ARG, "solvePressureNoAlloc: iterations < 0");
  unsigned char* base = nullptr;
  int st = scratch_base(&base);
  if (st != FH_OK) return st;
  float* x32 = reinterpret_cast<float*>(base + kOffX);
  float* b32 = reinterpret_cast<float*>(base + kOffB);
  double* partials = reinterpret_cast<double*>(base + kOffPartials);
  double* sums = reinterpret_cast<double*>(base + kOffSums);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int N = nx * ny;
  double h[4];

  // pressureRedBlack.ts:93-99: zero divergence -> p = 0, no sweeps.
  launch_norms(s, nullptr, p, div, label, nx, ny, scale, partials, sums);
  NCHK(hipGetLastError());
  if ((st = read_sums(s, sums, h)) != FH_OK) return st;
  if (h[2] == 0.0) {
    k_zero_f64<<<(N + 255) / 256, 256, 0, s>>>(p, N);
    NCHK(hipGetLastError());
    NCHK(hipStreamSynchronize(s));
    *out_iters = 0;
    return FH_OK;
  }
  const double divrms = std::sqrt(h[2] / h[3]);
  const double target = tol > 0 ? tol * divrms : -1;
  launch_f64_to_f32(s, p, x32, N);
  launch_rhs32(s, div, b32, N, 1.0);
  NCHK(hipGetLastError());
  int used = iterations;
  for (int k = 0; k < iterations; k++) {
    launch_rb_sweep(s, x32, b32, label, nx, ny, 0, (float)omega, (float)scale, 0);
    launch_rb_sweep(s, x32, b32, label, nx, ny, 1, (float)omega, (float)scale, 0);
    NCHK(hipGetLastError());
    if (target >= 0 && k % 12 == 11) {  // pressureRedBlack.ts:103-109
      launch_norms(s, x32, nullptr, div, label, nx, ny, scale, partials, sums);
      NCHK(hipGetLastError());
      if ((st = read_sums(s, sums, h)) != FH_OK) return st;
      if (h[1] > 0 && std::sqrt(h[0] / h[1]) <= target) {
        used = k + 1;
        break;
      }
    }
  }
  launch_f32_to_f64(s, x32, p, N);
  NCHK(hipGetLastError());
  NCHK(hipStreamSynchronize(s));
  *out_iters = used;
  return FH_OK;
}

}  // extern "C"
