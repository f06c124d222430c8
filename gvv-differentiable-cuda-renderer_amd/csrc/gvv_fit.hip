// gvv_fit.hip -- the fit harness's photometric loss: per-subject masked squared error of a
// render against a target, and its gradient with respect to the render (include/gvv_fit.h).
//
// Both are HBM-bound streams: the error reads render, target and mask once (28 B per pixel),
// the gradient reads them and writes the render gradient (40 B per pixel).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/gvv_fit.h"
#include "../../include/gvv_renderer.h"

namespace gvv {
int set_last_error(const std::string& m);
}

namespace {

constexpr int kThreads = 256;
constexpr int64_t kPixelsPerBlock = 16 * kThreads;   // 16 pixels per thread
constexpr int64_t kMaxBlocks = 1 << 16;             // partial sums per subject

int64_t blocks_for(int64_t P) {
  const int64_t n = (P + kPixelsPerBlock - 1) / kPixelsPerBlock;
  return n < 1 ? 1 : (n > kMaxBlocks ? kMaxBlocks : n);
}

// Sum over the workgroup in a fixed order: a butterfly within each wave, then wave 0 adds the
// four waves' sums.  Thread 0 returns the total.
__device__ __forceinline__ float block_sum(float v, float* s_w) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_w[wave] = v;
  __syncthreads();
  float t = 0.f;
  if (threadIdx.x == 0)
    for (int w = 0; w < kThreads / 64; ++w) t += s_w[w];
  return t;
}

// Block (j, b) sums pixels [j * chunk, min((j + 1) * chunk, P)) of subject b.
__global__ void __launch_bounds__(kThreads) k_masked_sq_partial(const float* __restrict__ render,
                                                               const float* __restrict__ target,
                                                               const float* __restrict__ mask, int64_t P,
                                                               int64_t chunk, int64_t nblk,
                                                               float* __restrict__ partial) {
  __shared__ float s_w[kThreads / 64];
  const int64_t b = blockIdx.y;
  const int64_t p0 = (int64_t)blockIdx.x * chunk;
  const int64_t p1 = p0 + chunk < P ? p0 + chunk : P;
  const float* r = render + b * P * 3;
  const float* t = target + b * P * 3;
  const float* m = mask + b * P;
  float acc = 0.f;
  for (int64_t p = p0 + threadIdx.x; p < p1; p += kThreads) {
    const float w = m[p];
    const float d0 = (r[3 * p] - t[3 * p]) * w, d1 = (r[3 * p + 1] - t[3 * p + 1]) * w,
                d2 = (r[3 * p + 2] - t[3 * p + 2]) * w;
    acc += d0 * d0 + d1 * d1 + d2 * d2;
  }
  const float s = block_sum(acc, s_w);
  if (threadIdx.x == 0) partial[b * nblk + blockIdx.x] = s;
}

// Block b adds subject b's partial sums in a fixed order.
__global__ void __launch_bounds__(kThreads) k_masked_sq_final(const float* __restrict__ partial, int64_t nblk,
                                                             float* __restrict__ sq) {
  __shared__ float s_w[kThreads / 64];
  const float* q = partial + (int64_t)blockIdx.x * nblk;
  float acc = 0.f;
  for (int64_t j = threadIdx.x; j < nblk; j += kThreads) acc += q[j];
  const float s = block_sum(acc, s_w);
  if (threadIdx.x == 0) sq[blockIdx.x] = s;
}

// One thread per pixel of the batch.
__global__ void __launch_bounds__(kThreads) k_masked_sq_grad(const float* __restrict__ render,
                                                            const float* __restrict__ target,
                                                            const float* __restrict__ mask,
                                                            const float* __restrict__ g_sq, int64_t P, int64_t n,
                                                            float* __restrict__ g_render) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float w = mask[i];
  const float s = 2.f * w * w * g_sq[i / P];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) g_render[3 * i + ch] = (render[3 * i + ch] - target[3 * i + ch]) * s;
}

bool check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return true;
  gvv::set_last_error(std::string(what) + ": " + hipGetErrorString(e));
  return false;
}

}  // namespace

extern "C" int64_t gvv_masked_sq_workspace_floats(int32_t B, int64_t P) {
  if (B <= 0 || B > 65535 || P <= 0) return -1;
  return (int64_t)B * blocks_for(P);
}

extern "C" int gvv_masked_sq_error(int32_t B, int64_t P, const float* render, const float* target,
                                   const float* mask, float* work, float* sq, void* stream) {
  if (B <= 0 || B > 65535 || P <= 0)
    return gvv::set_last_error("masked squared error: need 1 <= B <= 65535 subjects and P > 0 pixels");
  if (!render || !target || !mask || !work || !sq)
    return gvv::set_last_error("masked squared error: null pointer");
  const int64_t nblk = blocks_for(P);
  const int64_t chunk = (P + nblk - 1) / nblk;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_masked_sq_partial, dim3((unsigned)nblk, (unsigned)B), dim3(kThreads), 0, s, render, target,
                     mask, P, chunk, nblk, work);
  if (!check_launch("masked squared error")) return 1;
  hipLaunchKernelGGL(k_masked_sq_final, dim3((unsigned)B), dim3(kThreads), 0, s, work, nblk, sq);
  return check_launch("masked squared error") ? 0 : 1;
}

extern "C" int gvv_masked_sq_error_grad(int32_t B, int64_t P, const float* render, const float* target,
                                        const float* mask, const float* g_sq, float* g_render, void* stream) {
  if (B <= 0 || B > 65535 || P <= 0)
    return gvv::set_last_error("masked squared error gradient: need 1 <= B <= 65535 subjects and P > 0 pixels");
  if (!render || !target || !mask || !g_sq || !g_render)
    return gvv::set_last_error("masked squared error gradient: null pointer");
  const int64_t n = (int64_t)B * P;
  const int64_t grid = (n + kThreads - 1) / kThreads;
  if (grid > INT32_MAX) return gvv::set_last_error("masked squared error gradient: too many pixels");
  hipLaunchKernelGGL(k_masked_sq_grad, dim3((unsigned)grid), dim3(kThreads), 0, (hipStream_t)stream, render,
                     target, mask, g_sq, P, n, g_render);
  return check_launch("masked squared error gradient") ? 0 : 1;
}
