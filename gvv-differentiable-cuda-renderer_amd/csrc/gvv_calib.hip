// gvv_calib.hip -- measurement support: the float-atomic throughput of this GPU, against which
// the rasterizer's atomic rate is quoted (include/gvv_renderer.h, gvv_calibrate_atomics).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/gvv_renderer.h"

namespace gvv {
int set_last_error(const std::string& m);
}

namespace {

__device__ __forceinline__ uint32_t mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Each thread issues `per_thread` atomicAdd(+1) into buf[0, n): scattered (hashed) addresses,
// or streaming ones where consecutive lanes hit consecutive floats.
__global__ void __launch_bounds__(256) k_atomic_rate(float* __restrict__ buf, uint32_t n, uint32_t per_thread,
                                                     int scattered) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = 0; i < per_thread; ++i) {
    const uint32_t a = scattered ? mix(t * per_thread + i) % n : (t + i * stride) % n;
    atomicAdd(buf + a, 1.0f);
  }
}

}  // namespace

extern "C" int gvv_calibrate_atomics(float* buf, int64_t n_floats, int64_t updates, int32_t scattered, void* stream) {
  if (!buf || n_floats <= 0 || n_floats > (int64_t)UINT32_MAX || updates <= 0)
    return gvv::set_last_error("calibrate atomics: need a buffer of 1 .. 2^32-1 floats and updates > 0");
  const uint32_t threads = 1u << 20;
  const uint32_t per = (uint32_t)((updates + threads - 1) / threads);
  hipLaunchKernelGGL(k_atomic_rate, dim3(threads / 256), dim3(256), 0, (hipStream_t)stream, buf, (uint32_t)n_floats,
                     per, scattered ? 1 : 0);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return gvv::set_last_error(std::string("calibrate atomics: ") + hipGetErrorString(e));
  return 0;
}
