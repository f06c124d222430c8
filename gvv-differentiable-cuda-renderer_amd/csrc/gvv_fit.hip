// gvv_fit.hip -- kernels of the fit harness (include/gvv_fit.h).
//
// Photometric loss: per-subject masked squared error of a render against a target, and its
// gradient with respect to the render.  Both are HBM-bound streams: the error reads render,
// target and mask once (28 B per pixel), the gradient reads them and writes the render
// gradient (40 B per pixel).
//
// Shape model: posed vertices v = R (T + sum_k alpha_k B_k - c) + c + t of a linear shape basis
// under a rigid motion per subject, and the gradients of alpha and of (R, t).  The forward is
// one thread per vertex and group of 4 subjects (the coefficients and motions are wave-uniform
// scalar loads).  The backward sums over vertices:
// each workgroup takes a chunk of vertices, stages the chunk of the basis and the subjects'
// unrotated vertex gradients in LDS and forms every (subject, coefficient) partial dot
// product of the chunk; a second launch adds the chunks' partials in a fixed order.
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

// ---- shape model -------------------------------------------------------------------------
constexpr int kShapeSubjects = 16;     // subjects per backward workgroup (grid.y covers more)
constexpr int kShapeMaxSubjects = 65535 * 4;   // the forward's grid.y limit at 4 subjects per thread
constexpr int kShapeChunk = 64;        // vertices per backward workgroup
constexpr int kShapeRow = 3 * kShapeChunk + 1;   // padded LDS row: lanes of one column differ in bank
constexpr int kShapeMaxBasis = 128;
constexpr int kShapeRt = 12;           // R row-major (9) and t (3) per subject

// v[b][i] = R_b (T_i + sum_k alpha[b][k] basis[k][3i..3i+2] - c) + c + t_b and xc[b][i] = the
// unrotated, centred vertex (T_i + ... - c), kept for the backward.  Thread = (vertex, group of
// kShapeFwdSubjects subjects b0 .. b0 + nb - 1): enough wavefronts to cover the basis reads'
// latency, which one thread per vertex for all subjects leaves exposed.
constexpr int kShapeFwdSubjects = 4;

__global__ void __launch_bounds__(64) k_shape_forward(const float* __restrict__ alpha, const float* __restrict__ rt,
                                                      const float* __restrict__ templ, const float* __restrict__ basis,
                                                      const float* __restrict__ centre, int B, int N, int K,
                                                      float* __restrict__ v, float* __restrict__ xc) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int b0 = blockIdx.y * kShapeFwdSubjects;
  const int nb = B - b0 < kShapeFwdSubjects ? B - b0 : kShapeFwdSubjects;
  if (i >= N) return;
  float acc[kShapeFwdSubjects][3];
#pragma unroll
  for (int s = 0; s < kShapeFwdSubjects; ++s) acc[s][0] = acc[s][1] = acc[s][2] = 0.f;
  const size_t n3 = (size_t)3 * N;
  const float* bi = basis + 3 * (size_t)i;
#pragma unroll 4
  for (int k = 0; k < K; ++k) {
    const float e0 = bi[k * n3], e1 = bi[k * n3 + 1], e2 = bi[k * n3 + 2];
#pragma unroll
    for (int s = 0; s < kShapeFwdSubjects; ++s) {
      if (s < nb) {
        const float a = alpha[(size_t)(b0 + s) * K + k];
        acc[s][0] += a * e0;
        acc[s][1] += a * e1;
        acc[s][2] += a * e2;
      }
    }
  }
  const float c0 = centre[0], c1 = centre[1], c2 = centre[2];
  const float t0 = templ[3 * (size_t)i], t1 = templ[3 * (size_t)i + 1], t2 = templ[3 * (size_t)i + 2];
#pragma unroll
  for (int s = 0; s < kShapeFwdSubjects; ++s) {
    if (s < nb) {
      const float* m = rt + (size_t)(b0 + s) * kShapeRt;
      const float x = (t0 + acc[s][0]) - c0, y = (t1 + acc[s][1]) - c1, z = (t2 + acc[s][2]) - c2;
      const size_t o = ((size_t)(b0 + s) * N + i) * 3;
      xc[o] = x; xc[o + 1] = y; xc[o + 2] = z;
      v[o] = m[0] * x + m[1] * y + m[2] * z + c0 + m[9];
      v[o + 1] = m[3] * x + m[4] * y + m[5] * z + c1 + m[10];
      v[o + 2] = m[6] * x + m[7] * y + m[8] * z + c2 + m[11];
    }
  }
}

// Workgroup (chunk j, subject group): partial[j][b][0..K) = sum over the chunk's vertices of
// (R_b^T g_v[b][i]) . basis[k][3i..3i+2], and partial[j][b][K..K+12) the chunk's sums of
// g_v[b][i] (x) xc[b][i] (9, row-major) and of g_v[b][i] (3).  Dynamic LDS: (K + 16) padded
// rows of the chunk (basis rows, then the subjects' unrotated gradients).
__global__ void __launch_bounds__(256) k_shape_backward_partial(const float* __restrict__ g_v,
                                                                const float* __restrict__ xc,
                                                                const float* __restrict__ rt,
                                                                const float* __restrict__ basis, int B, int N, int K,
                                                                float* __restrict__ partial) {
  extern __shared__ float s_rows[];
  float* s_basis = s_rows;                       // [K][kShapeRow]
  float* s_g = s_rows + (size_t)K * kShapeRow;   // [16][kShapeRow]
  const int tid = threadIdx.x;
  const int j = blockIdx.x;
  const int i0 = j * kShapeChunk;
  const int nv = N - i0 < kShapeChunk ? N - i0 : kShapeChunk;
  const int b0 = blockIdx.y * kShapeSubjects;
  const int nb = B - b0 < kShapeSubjects ? B - b0 : kShapeSubjects;
  const int stride = K + kShapeRt;
  const size_t n3 = (size_t)3 * N;
  // the chunk of every basis row (coalesced along the row), four loads in flight per thread
  const int total = K * 3 * kShapeChunk;
  for (int q0 = tid; q0 < total; q0 += 4 * 256) {
    float e[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = q0 + u * 256;
      const int k = q / (3 * kShapeChunk), c = q - k * 3 * kShapeChunk;
      e[u] = q < total && c < 3 * nv ? basis[k * n3 + 3 * (size_t)i0 + c] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = q0 + u * 256;
      const int k = q / (3 * kShapeChunk), c = q - k * 3 * kShapeChunk;
      if (q < total) s_basis[k * kShapeRow + c] = e[u];
    }
  }
  // 16 lanes per subject, 4 vertices each: unrotated gradients into LDS, motion sums in registers
  {
    const int s = tid >> 4, l = tid & 15;
    float acc[kShapeRt];
#pragma unroll
    for (int q = 0; q < kShapeRt; ++q) acc[q] = 0.f;
    if (s < nb) {
      const float* m = rt + (size_t)(b0 + s) * kShapeRt;
      for (int vv = l; vv < kShapeChunk; vv += 16) {
        float g0 = 0.f, g1 = 0.f, g2 = 0.f, x0 = 0.f, x1 = 0.f, x2 = 0.f;
        if (vv < nv) {
          const size_t o = ((size_t)(b0 + s) * N + i0 + vv) * 3;
          g0 = g_v[o]; g1 = g_v[o + 1]; g2 = g_v[o + 2];
          x0 = xc[o]; x1 = xc[o + 1]; x2 = xc[o + 2];
        }
        s_g[s * kShapeRow + 3 * vv] = m[0] * g0 + m[3] * g1 + m[6] * g2;
        s_g[s * kShapeRow + 3 * vv + 1] = m[1] * g0 + m[4] * g1 + m[7] * g2;
        s_g[s * kShapeRow + 3 * vv + 2] = m[2] * g0 + m[5] * g1 + m[8] * g2;
        acc[0] += g0 * x0; acc[1] += g0 * x1; acc[2] += g0 * x2;
        acc[3] += g1 * x0; acc[4] += g1 * x1; acc[5] += g1 * x2;
        acc[6] += g2 * x0; acc[7] += g2 * x1; acc[8] += g2 * x2;
        acc[9] += g0; acc[10] += g1; acc[11] += g2;
      }
    }
    // sum over the subject's 16 lanes (a butterfly within each 16-lane group)
#pragma unroll
    for (int q = 0; q < kShapeRt; ++q) {
      float x = acc[q];
#pragma unroll
      for (int d = 8; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
      acc[q] = x;
    }
    if (l == 0 && s < nb) {
      float* dst = partial + ((size_t)j * B + b0 + s) * stride + K;
#pragma unroll
      for (int q = 0; q < kShapeRt; ++q) dst[q] = acc[q];
    }
  }
  __syncthreads();
  // every (subject, coefficient) dot product over the chunk
  for (int o = tid; o < nb * K; o += 256) {
    const int s = o / K, k = o - s * K;
    const float* gr = s_g + s * kShapeRow;
    const float* br = s_basis + k * kShapeRow;
    float a = 0.f;
    for (int c = 0; c < 3 * nv; ++c) a += gr[c] * br[c];
    partial[((size_t)j * B + b0 + s) * stride + k] = a;
  }
}

// Rotation of axis-angle w (rad), R = I + a [w]x + b [w]x^2 with a = sin(th)/th and
// b = (1 - cos(th))/th^2, and their series 1 - th^2/6 and 1/2 - th^2/24 below th^2 = 1e-8.
struct Rodrigues {
  float R[9], Kx[9], KK[9];
  float a, b, th2, den, cth;   // den = th^2 + 1e-24; cth = cos(th)
  bool small;
};

__device__ __forceinline__ Rodrigues rodrigues(float w0, float w1, float w2) {
  Rodrigues r;
  r.th2 = w0 * w0 + w1 * w1 + w2 * w2;
  r.den = r.th2 + 1e-24f;
  const float th = sqrtf(r.den);
  r.small = r.th2 < 1e-8f;
  r.cth = cosf(th);
  r.a = r.small ? 1.f - r.th2 / 6.f : sinf(th) / th;
  r.b = r.small ? 0.5f - r.th2 / 24.f : (1.f - r.cth) / r.den;
  const float K[9] = {0.f, -w2, w1, w2, 0.f, -w0, -w1, w0, 0.f};
#pragma unroll
  for (int q = 0; q < 9; ++q) r.Kx[q] = K[q];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.KK[3 * i + j] = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
#pragma unroll
  for (int q = 0; q < 9; ++q) r.R[q] = (q % 4 == 0 ? 1.f : 0.f) + r.a * r.Kx[q] + r.b * r.KK[q];
  return r;
}

// rt[b] = (R(pose[b][0..3]) row-major, tscale * pose[b][3..6]): thread per subject.
__global__ void __launch_bounds__(64) k_pose_motion(const float* __restrict__ pose, float tscale, int B,
                                                    float* __restrict__ rt) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= B) return;
  const float* p = pose + 6 * (size_t)b;
  const Rodrigues r = rodrigues(p[0], p[1], p[2]);
  float* o = rt + (size_t)b * kShapeRt;
#pragma unroll
  for (int q = 0; q < 9; ++q) o[q] = r.R[q];
#pragma unroll
  for (int q = 0; q < 3; ++q) o[9 + q] = tscale * p[3 + q];
}

// g_out[b] = (g_alpha (K), dL/dw (3) through the rotation, tscale * dL/dt (3)) from
// g_art[b] = (g_alpha, dL/dR row-major, dL/dt): thread per (b, q < K + 6).
__global__ void __launch_bounds__(256) k_pose_grad(const float* __restrict__ g_art, const float* __restrict__ pose,
                                                   float tscale, int B, int K, float* __restrict__ g_out) {
  const int o = blockIdx.x * 256 + threadIdx.x;
  const int width = K + 6;
  if (o >= B * width) return;
  const int b = o / width, q = o - b * width;
  const float* g = g_art + (size_t)b * (K + kShapeRt);
  if (q < K) {
    g_out[o] = g[q];
    return;
  }
  if (q >= K + 3) {
    g_out[o] = tscale * g[K + 9 + (q - K - 3)];
    return;
  }
  const int j = q - K;
  const float* p = pose + 6 * (size_t)b;
  const Rodrigues r = rodrigues(p[0], p[1], p[2]);
  const float wj = p[j];
  // d a / d w_j and d b / d w_j of the branch the forward took
  const float da = r.small ? -wj / 3.f : (r.cth - r.a) * wj / r.den;
  const float db = r.small ? -wj / 12.f : (r.a - 2.f * r.b) * wj / r.den;
  // E = d [w]x / d w_j
  float E[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (j == 0) { E[5] = -1.f; E[7] = 1.f; }
  else if (j == 1) { E[2] = 1.f; E[6] = -1.f; }
  else { E[1] = -1.f; E[3] = 1.f; }
  const float* G = g + K;
  float acc = 0.f;
#pragma unroll
  for (int rr = 0; rr < 3; ++rr)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float ek = 0.f;    // (E [w]x + [w]x E)[rr][c]
#pragma unroll
      for (int m = 0; m < 3; ++m) ek += E[3 * rr + m] * r.Kx[3 * m + c] + r.Kx[3 * rr + m] * E[3 * m + c];
      const int e = 3 * rr + c;
      acc += G[e] * (da * r.Kx[e] + r.a * E[e] + db * r.KK[e] + r.b * ek);
    }
  g_out[o] = acc;
}

// out[o] = sum over chunks j of partial[j][o], o = b * (K + 12) + q: a 1024-thread workgroup per
// 64 outputs; thread (g, l) sums chunks g, g + 16, ... of output l, then thread (0, l) adds the
// 16 group sums in order (deterministic).
constexpr int kFinalGroups = 16;
__global__ void __launch_bounds__(1024) k_shape_backward_final(const float* __restrict__ partial, int chunks, int B,
                                                               int stride, float* __restrict__ out) {
  __shared__ float s_sum[kFinalGroups][64];
  const int l = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int o = blockIdx.x * 64 + l;
  const int outs = B * stride;
  float a = 0.f;
  if (o < outs) {
#pragma unroll 4
    for (int j = g; j < chunks; j += kFinalGroups) a += partial[(size_t)j * outs + o];
  }
  s_sum[g][l] = a;
  __syncthreads();
  if (g == 0 && o < outs) {
    float t = 0.f;
#pragma unroll
    for (int q = 0; q < kFinalGroups; ++q) t += s_sum[q][l];
    out[o] = t;
  }
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

extern "C" int64_t gvv_shape_workspace_floats(int32_t B, int32_t N, int32_t K) {
  if (B <= 0 || B > kShapeMaxSubjects || N <= 0 || K <= 0 || K > kShapeMaxBasis) return -1;
  const int64_t chunks = (N + kShapeChunk - 1) / kShapeChunk;
  return (chunks + 1) * B * (K + kShapeRt);     // chunk partials, then their sums
}

extern "C" int gvv_shape_forward(int32_t B, int32_t N, int32_t K, const float* alpha, const float* pose,
                                 float tscale, const float* templ, const float* basis, const float* centre, float* v,
                                 float* xc, float* rt, void* stream) {
  if (B <= 0 || B > kShapeMaxSubjects || N <= 0 || K <= 0 || K > kShapeMaxBasis)
    return gvv::set_last_error("shape forward: need 1 <= B <= 262140 subjects, N >= 1 vertices and 1 <= K <= 128 basis vectors");
  if (!alpha || !pose || !templ || !basis || !centre || !v || !xc || !rt)
    return gvv::set_last_error("shape forward: null pointer");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pose_motion, dim3((unsigned)((B + 63) / 64)), dim3(64), 0, s, pose, tscale, B, rt);
  if (!check_launch("shape forward")) return 1;
  const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((B + kShapeFwdSubjects - 1) / kShapeFwdSubjects));
  hipLaunchKernelGGL(k_shape_forward, grid, dim3(64), 0, s, alpha, rt, templ, basis, centre, B, N, K, v, xc);
  return check_launch("shape forward") ? 0 : 1;
}

extern "C" int gvv_shape_backward(int32_t B, int32_t N, int32_t K, const float* g_v, const float* xc,
                                  const float* pose, const float* rt, float tscale, const float* basis, float* work,
                                  float* g_alpha_pose, void* stream) {
  if (B <= 0 || B > kShapeMaxSubjects || N <= 0 || K <= 0 || K > kShapeMaxBasis)
    return gvv::set_last_error("shape backward: need 1 <= B <= 262140 subjects, N >= 1 vertices and 1 <= K <= 128 basis vectors");
  if (!g_v || !xc || !pose || !rt || !basis || !work || !g_alpha_pose)
    return gvv::set_last_error("shape backward: null pointer");
  const int chunks = (N + kShapeChunk - 1) / kShapeChunk;
  const int stride = K + kShapeRt;
  const size_t lds = (size_t)(K + kShapeSubjects) * kShapeRow * sizeof(float);   // <= 111 KB at K = 128
  // dynamic LDS above 64 KB needs the attribute (set on the current device)
  const hipError_t e = hipFuncSetAttribute((const void*)k_shape_backward_partial,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return gvv::set_last_error(std::string("shape backward: ") + hipGetErrorString(e));
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)chunks, (unsigned)((B + kShapeSubjects - 1) / kShapeSubjects));
  hipLaunchKernelGGL(k_shape_backward_partial, grid, dim3(256), lds, s, g_v, xc, rt, basis, B, N, K, work);
  if (!check_launch("shape backward")) return 1;
  float* sums = work + (size_t)chunks * B * stride;
  const int64_t outs = (int64_t)B * stride;
  hipLaunchKernelGGL(k_shape_backward_final, dim3((unsigned)((outs + 63) / 64)), dim3(1024), 0, s, work, chunks, B,
                     stride, sums);
  if (!check_launch("shape backward")) return 1;
  const int64_t outs6 = (int64_t)B * (K + 6);
  hipLaunchKernelGGL(k_pose_grad, dim3((unsigned)((outs6 + 255) / 256)), dim3(256), 0, s, sums, pose, tscale, B, K,
                     g_alpha_pose);
  return check_launch("shape backward") ? 0 : 1;
}
