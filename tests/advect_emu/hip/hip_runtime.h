// Host stand-ins for the HIP built-ins the LDS advection kernels use, so the
// kernels' bodies run on the CPU under tests/test_advect_emu.py.  Test only:
// nothing here is compiled into libflowhip.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
struct dim3 {
  unsigned x = 1, y = 1, z = 1;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
extern dim3 threadIdx, blockIdx, gridDim, blockDim;
struct EmuBarrier {};
extern int emu_phase, emu_barriers;
// Each emulated thread runs the kernel once per phase: in phase p up to its
// (p+1)-th barrier, so that everything the workgroup writes to LDS before
// that barrier is in place when any thread reads past it (the kernels' LDS
// writes between barriers are idempotent), and in the last phase to the end.
inline void __syncthreads() {
  if (++emu_barriers > emu_phase) throw EmuBarrier{};
}
inline int atomicExch(int* p, int v) {
  const int o = *p;
  *p = v;
  return o;
}
typedef void* hipStream_t;
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __restrict__
template <class T>
T __shfl_xor(T v, int, int) {
  return v;
}
// The device multiply uses the low 24 bits of each operand: an operand
// outside the signed 24-bit range would give a different product there.
inline int __mul24(int a, int b) {
  if (a < -(1 << 23) || a >= (1 << 23) || b < -(1 << 23) || b >= (1 << 23))
    throw std::runtime_error("__mul24 operand outside 24 bits");
  return a * b;
}
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
