// gvv_conv.hip -- BF16 MFMA convolutions of the proposal network (declared in include/gvv_proposal.h).
//
// The reference has no learned component; these kernels serve the proposal network that
// warm-starts the analysis-by-synthesis fit.  Every convolution (and the linear head, as a
// 1x1 convolution over a 1x1 image) is an implicit GEMM on v_mfma_f32_16x16x32_bf16 with
// FP32 accumulation:
//
//   forward        O[q][p] = sum_k P[p][k] Q[q][k]   p = output channel, q = output pixel,
//                                                    k = (kh, kw, ci)
//   backward data  p = input channel, q = input pixel, k = (kh, kw, co); taps whose output
//                  pixel is off the stride grid contribute zero
//   backward wt    p = (kh, kw, ci), q = output channel, k = output pixel (split over the grid
//                  z dimension into FP32 partials, summed by k_split_sum)
//
// Activations are NHWC bf16 with the channel count a multiple of 8, so that 8 consecutive k
// are 8 consecutive channels of one tap: one 16-byte load per MFMA operand chunk.  For the
// weight gradient the reduction runs over pixels, which are strided in NHWC; each lane then
// gathers 8 pixels of one channel with 2-byte loads that are contiguous across the wave.
//
// MFMA lane layout (16x16x32, wave64): lane l supplies row l%16 of A and column l%16 of B with
// k = 8*(l/16) .. 8*(l/16)+7; it receives D[4*(l/16)+r][l%16] in accumulator r.  A is the P
// side, B the Q side, so each lane ends with 4 consecutive p of one q: one 8-byte (bf16) or
// 16-byte (fp32) store.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/gvv_proposal.h"

namespace gvv {
int set_last_error(const std::string& m);
}

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kBK = 32;           // k per main-loop step (one MFMA depth)
constexpr int kRow = kBK + 8;     // LDS row stride in bf16 (80 bytes: spreads rows over banks)
constexpr int kThreads = 256;     // 4 waves, 2 x 2 over the (P, Q) tile

enum Mode { kFwd = 0, kBwdData = 1, kBwdWeight = 2 };

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// float -> bf16, round to nearest even (NaN stays NaN)
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

struct ConvGeom {
  int B, H, W, Cin, OH, OW, Cout, KS, S, PAD;
};

struct ConvArgs {
  ConvGeom g;
  const uint16_t* x;       // [B][H][W][Cin]       forward input (bf16)
  const uint16_t* w;       // [Cout][KS][KS][Cin]  weights (bf16)
  const uint16_t* wt;      // [Cin][KS][KS][Cout]  transposed weights (bf16)
  const uint16_t* dz;      // [B][OH][OW][Cout]    gradient at the pre-activation output (bf16)
  const float* bias;       // [Cout]
  void* out;               // fwd: y (bf16 or f32); bwd data: dx (bf16); bwd wt: partials (f32)
  int out_f32, relu;
  const uint16_t* mask;    // bwd data: dx *= (mask > 0) where mask is the layer's forward input
  int64_t k_per_split;     // bwd wt: pixels per grid-z slice (multiple of kBK)
};

// GEMM extents of each mode
__host__ __device__ inline void gemm_dims(int mode, const ConvGeom& g, int64_t& P, int64_t& Q, int64_t& K) {
  const int64_t taps = int64_t(g.KS) * g.KS;
  if (mode == kFwd) { P = g.Cout; Q = int64_t(g.B) * g.OH * g.OW; K = taps * g.Cin; }
  else if (mode == kBwdData) { P = g.Cin; Q = int64_t(g.B) * g.H * g.W; K = taps * g.Cout; }
  else { P = taps * g.Cin; Q = g.Cout; K = int64_t(g.B) * g.OH * g.OW; }
}

__device__ __forceinline__ uint4 ld16(const uint16_t* p) { return *reinterpret_cast<const uint4*>(p); }

// Output-pixel index -> input element offset of tap (kh, kw), channel ci; -1 off the image.
__device__ __forceinline__ int64_t in_offset(const ConvGeom& g, int64_t pix, int kh, int kw, int ci) {
  const int ohw = g.OH * g.OW;
  const int img = int(pix / ohw);
  const int r = int(pix - int64_t(img) * ohw);
  const int oh = r / g.OW, ow = r - oh * g.OW;
  const int ih = oh * g.S - g.PAD + kh, iw = ow * g.S - g.PAD + kw;
  if (ih < 0 || ih >= g.H || iw < 0 || iw >= g.W) return -1;
  return ((int64_t(img) * g.H + ih) * g.W + iw) * g.Cin + ci;
}

template <int MODE, int BP, int BQ>
__global__ __launch_bounds__(kThreads) void k_conv(ConvArgs a) {
  static_assert(BP % 32 == 0 && BQ % 32 == 0, "tile must split over 2 x 2 waves of 16 x 16 MFMAs");
  constexpr int TP = BP / 32, TQ = BQ / 32;                 // MFMA tiles per wave
  constexpr int NP = (BP * 4 + kThreads - 1) / kThreads;    // P chunks per thread per step
  constexpr int NQ = (BQ * 4 + kThreads - 1) / kThreads;
  __shared__ __attribute__((aligned(16))) uint16_t sP[BP * kRow];
  __shared__ __attribute__((aligned(16))) uint16_t sQ[BQ * kRow];

  const ConvGeom g = a.g;
  int64_t P, Q, K;
  gemm_dims(MODE, g, P, Q, K);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t q0 = int64_t(blockIdx.x) * BQ;
  const int64_t p0 = int64_t(blockIdx.y) * BP;
  int64_t kbeg = 0, kend = K;
  if (MODE == kBwdWeight) {
    kbeg = int64_t(blockIdx.z) * a.k_per_split;
    kend = kbeg + a.k_per_split < K ? kbeg + a.k_per_split : K;
  }

  // Per-thread chunk coordinates: in forward and backward-data a chunk is (row, k-chunk) with
  // 4 consecutive threads on one row (64 contiguous bytes); in backward-weight consecutive
  // threads take consecutive rows so that the 2-byte gathers coalesce across the wave.
  auto p_of = [&](int t) { return MODE == kBwdWeight ? t % BP : t >> 2; };
  auto pk_of = [&](int t) { return MODE == kBwdWeight ? t / BP : t & 3; };
  auto q_of = [&](int t) { return MODE == kBwdWeight ? t % BQ : t >> 2; };
  auto qk_of = [&](int t) { return MODE == kBwdWeight ? t / BQ : t & 3; };

  // Q-row pixel decode (forward, backward data), fixed across k steps
  int q_img[NQ], q_y[NQ], q_x[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    const int t = tid + i * kThreads;
    const int64_t q = q0 + q_of(t);
    q_img[i] = -1; q_y[i] = 0; q_x[i] = 0;
    if (MODE != kBwdWeight && t < BQ * 4 && q < Q) {
      const int hh = MODE == kFwd ? g.OH : g.H, ww = MODE == kFwd ? g.OW : g.W;
      q_img[i] = int(q / (int64_t(hh) * ww));
      const int r = int(q - int64_t(q_img[i]) * hh * ww);
      q_y[i] = r / ww;
      q_x[i] = r - q_y[i] * ww;
    }
  }

  uint4 rP[NP], rQ[NQ];
  auto load_step = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int t = tid + i * kThreads;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (t < BP * 4) {
        const int64_t p = p0 + p_of(t);
        const int64_t k = k0 + pk_of(t) * 8;
        if (MODE == kFwd) {
          if (p < P && k < K) v = ld16(a.w + p * K + k);
        } else if (MODE == kBwdData) {
          if (p < P && k < K) v = ld16(a.wt + p * K + k);
        } else if (p < P) {
          const int tap = int(p / g.Cin), ci = int(p - int64_t(tap) * g.Cin);
          const int kh = tap / g.KS, kw = tap - kh * g.KS;
          uint16_t h[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            h[j] = 0;
            if (k + j < kend) {
              const int64_t off = in_offset(g, k + j, kh, kw, ci);
              if (off >= 0) h[j] = a.x[off];
            }
          }
          v = make_uint4(h[0] | (uint32_t(h[1]) << 16), h[2] | (uint32_t(h[3]) << 16),
                         h[4] | (uint32_t(h[5]) << 16), h[6] | (uint32_t(h[7]) << 16));
        }
      }
      rP[i] = v;
    }
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      const int t = tid + i * kThreads;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (t < BQ * 4) {
        const int64_t k = k0 + qk_of(t) * 8;
        if (MODE == kFwd) {
          if (q_img[i] >= 0 && k < K) {
            const int tap = int(k / g.Cin), ci = int(k - int64_t(tap) * g.Cin);
            const int kh = tap / g.KS, kw = tap - kh * g.KS;
            const int ih = q_y[i] * g.S - g.PAD + kh, iw = q_x[i] * g.S - g.PAD + kw;
            if (ih >= 0 && ih < g.H && iw >= 0 && iw < g.W)
              v = ld16(a.x + ((int64_t(q_img[i]) * g.H + ih) * g.W + iw) * g.Cin + ci);
          }
        } else if (MODE == kBwdData) {
          if (q_img[i] >= 0 && k < K) {
            const int tap = int(k / g.Cout), co = int(k - int64_t(tap) * g.Cout);
            const int kh = tap / g.KS, kw = tap - kh * g.KS;
            const int ty = q_y[i] + g.PAD - kh, tx = q_x[i] + g.PAD - kw;
            if (ty >= 0 && tx >= 0 && ty % g.S == 0 && tx % g.S == 0) {
              const int oh = ty / g.S, ow = tx / g.S;
              if (oh < g.OH && ow < g.OW)
                v = ld16(a.dz + ((int64_t(q_img[i]) * g.OH + oh) * g.OW + ow) * g.Cout + co);
            }
          }
        } else {
          const int64_t q = q0 + q_of(t);
          if (q < Q) {
            uint16_t h[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) h[j] = (k + j < kend) ? a.dz[(k + j) * g.Cout + q] : uint16_t(0);
            v = make_uint4(h[0] | (uint32_t(h[1]) << 16), h[2] | (uint32_t(h[3]) << 16),
                           h[4] | (uint32_t(h[5]) << 16), h[6] | (uint32_t(h[7]) << 16));
          }
        }
      }
      rQ[i] = v;
    }
  };

  f32x4 acc[TP][TQ];
#pragma unroll
  for (int i = 0; i < TP; ++i)
#pragma unroll
    for (int j = 0; j < TQ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int wp = wave & 1, wq = wave >> 1;
  const int frow = lane & 15, fchunk = lane >> 4;

  if (kbeg < kend) load_step(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += kBK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int t = tid + i * kThreads;
      if (t < BP * 4) *reinterpret_cast<uint4*>(&sP[p_of(t) * kRow + pk_of(t) * 8]) = rP[i];
    }
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      const int t = tid + i * kThreads;
      if (t < BQ * 4) *reinterpret_cast<uint4*>(&sQ[q_of(t) * kRow + qk_of(t) * 8]) = rQ[i];
    }
    __syncthreads();
    if (k0 + kBK < kend) load_step(k0 + kBK);   // next step's global loads overlap the MFMAs
    bf16x8 fa[TP], fb[TQ];
#pragma unroll
    for (int i = 0; i < TP; ++i)
      fa[i] = *reinterpret_cast<const bf16x8*>(&sP[(wp * (BP / 2) + i * 16 + frow) * kRow + fchunk * 8]);
#pragma unroll
    for (int j = 0; j < TQ; ++j)
      fb[j] = *reinterpret_cast<const bf16x8*>(&sQ[(wq * (BQ / 2) + j * 16 + frow) * kRow + fchunk * 8]);
#pragma unroll
    for (int i = 0; i < TP; ++i)
#pragma unroll
      for (int j = 0; j < TQ; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
  }

  // epilogue: lane holds D[p0 + .. + 4*(lane/16) + r][q0 + .. + lane%16], r = 0..3
#pragma unroll
  for (int i = 0; i < TP; ++i) {
    const int64_t pb = p0 + wp * (BP / 2) + i * 16 + 4 * fchunk;
#pragma unroll
    for (int j = 0; j < TQ; ++j) {
      const int64_t q = q0 + wq * (BQ / 2) + j * 16 + frow;
      if (q >= Q) continue;
      float v[4] = {acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
      if (MODE == kBwdWeight) {
        float* o = static_cast<float*>(a.out) + (int64_t(blockIdx.z) * Q + q) * P + pb;
        if (pb + 3 < P) {
          *reinterpret_cast<float4*>(o) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
          for (int r = 0; r < 4; ++r)
            if (pb + r < P) o[r] = v[r];
        }
        continue;
      }
      if (MODE == kFwd) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (pb + r < P) v[r] += a.bias[pb + r];
          if (a.relu) v[r] = v[r] > 0.f ? v[r] : 0.f;
        }
      } else if (a.mask) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (pb + r < P && !(bf2f(a.mask[q * P + pb + r]) > 0.f)) v[r] = 0.f;
      }
      if (a.out_f32) {
        float* o = static_cast<float*>(a.out) + q * P + pb;
        for (int r = 0; r < 4; ++r)
          if (pb + r < P) o[r] = v[r];
      } else {
        uint16_t* o = static_cast<uint16_t*>(a.out) + q * P + pb;
        if (pb + 3 < P) {
          *reinterpret_cast<uint2*>(o) = make_uint2(f2bf(v[0]) | (uint32_t(f2bf(v[1])) << 16),
                                                    f2bf(v[2]) | (uint32_t(f2bf(v[3])) << 16));
        } else {
          for (int r = 0; r < 4; ++r)
            if (pb + r < P) o[r] = f2bf(v[r]);
        }
      }
    }
  }
}

// out[i] = sum_s part[s][i], deterministic.  Block: 64 consecutive outputs x 4 split lanes;
// each lane sums every 4th split, then the four partial sums are added in fixed order.
__global__ __launch_bounds__(256) void k_split_sum(const float* __restrict__ part, int splits, int64_t n,
                                                   float* __restrict__ out) {
  __shared__ float red[4][64];
  const int64_t i = int64_t(blockIdx.x) * 64 + (threadIdx.x & 63);
  const int r = threadIdx.x >> 6;
  float s = 0.f;
  if (i < n)
    for (int k = r; k < splits; k += 4) s += part[int64_t(k) * n + i];
  red[r][threadIdx.x & 63] = s;
  __syncthreads();
  if (r == 0 && i < n) out[i] = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// Bias gradient partials: part[split][c] = sum over the split's pixels of dz[pix][c].
// Block: 64 channels x 4 pixel lanes; grid (ceil(C/64), splits).  Each lane keeps four
// independent running sums so that four loads are in flight.
__global__ __launch_bounds__(256) void k_bias_partial(const uint16_t* __restrict__ dz, int64_t pixels, int C,
                                                      int64_t per_split, float* __restrict__ part) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), r = threadIdx.x >> 6;
  const int64_t beg = int64_t(blockIdx.y) * per_split;
  const int64_t end = beg + per_split < pixels ? beg + per_split : pixels;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  if (c < C) {
    int64_t p = beg + r;
    for (; p + 12 < end; p += 16) {
      s0 += bf2f(dz[p * C + c]);
      s1 += bf2f(dz[(p + 4) * C + c]);
      s2 += bf2f(dz[(p + 8) * C + c]);
      s3 += bf2f(dz[(p + 12) * C + c]);
    }
    for (; p < end; p += 4) s0 += bf2f(dz[p * C + c]);
  }
  red[r][threadIdx.x & 63] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (r == 0 && c < C)
    part[int64_t(blockIdx.y) * C + c] = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// fp32 master weights [Cout][T][Cin] -> bf16 copy and bf16 transpose [Cin][T][Cout]
__global__ void k_prepare_weights(const float* __restrict__ w, int Cout, int T, int Cin, uint16_t* __restrict__ wb,
                                  uint16_t* __restrict__ wt) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n = int64_t(Cout) * T * Cin;
  if (i >= n) return;
  const int ci = int(i % Cin);
  const int64_t r = i / Cin;
  const int t = int(r % T), co = int(r / T);
  const uint16_t h = f2bf(w[i]);
  wb[i] = h;
  if (wt) wt[(int64_t(ci) * T + t) * Cout + co] = h;
}

// Renders [B][V][h][w][3] fp32 -> NHWC8 bf16 mosaic [B][rows*h/d][cols*w/d][8]: view v at tile
// (v / cols, v % cols), each output pixel the mean of a d x d block.  features == 0: channels
// 0..2 RGB, 3..7 zero.  features == 1: channels 0..2 RGB - 0.5, 3 the covered fraction of the
// block (pixels other than the (0, 1, 0) background), 4..5 the tile-local x and y in [-1, 1],
// 6..7 zero.  Empty tiles are zero.
__global__ void k_pack_mosaic(const float* __restrict__ render, int B, int V, int h, int w, int cols, int d,
                              int features, uint16_t* __restrict__ x) {
  const int rows = (V + cols - 1) / cols;
  const int th = h / d, tw = w / d;
  const int H = rows * th, W = cols * tw;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(B) * H * W) return;
  const int X = int(i % W);
  const int64_t r = i / W;
  const int Y = int(r % H), b = int(r / H);
  const int v = (Y / th) * cols + X / tw;
  uint16_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (v < V) {
    float acc[3] = {0.f, 0.f, 0.f};
    int covered = 0;
    const int y0 = (Y % th) * d, x0 = (X % tw) * d;
    for (int yy = 0; yy < d; ++yy) {
      const float* s = render + ((((int64_t(b) * V + v) * h) + y0 + yy) * w + x0) * 3;
      for (int xx = 0; xx < d; ++xx) {
        const float R = s[3 * xx], G = s[3 * xx + 1], Bc = s[3 * xx + 2];
        acc[0] += R;
        acc[1] += G;
        acc[2] += Bc;
        covered += !(R == 0.f && G == 1.f && Bc == 0.f);
      }
    }
    const float inv = 1.f / float(d * d);
    const float off = features ? 0.5f : 0.f;
    c[0] = f2bf(acc[0] * inv - off); c[1] = f2bf(acc[1] * inv - off); c[2] = f2bf(acc[2] * inv - off);
    if (features) {
      c[3] = f2bf(covered * inv);
      c[4] = f2bf(2.f * ((X % tw) + 0.5f) / float(tw) - 1.f);
      c[5] = f2bf(2.f * ((Y % th) + 0.5f) / float(th) - 1.f);
    }
  }
  *reinterpret_cast<uint4*>(x + i * 8) = make_uint4(c[0] | (uint32_t(c[1]) << 16), c[2] | (uint32_t(c[3]) << 16),
                                                    c[4] | (uint32_t(c[5]) << 16), 0);
}

bool geom_ok(const ConvGeom& g, std::string& why) {
  if (g.B <= 0 || g.H <= 0 || g.W <= 0 || g.Cin <= 0 || g.Cout <= 0 || g.KS <= 0 || g.S <= 0 || g.PAD < 0) {
    why = "conv: sizes must be positive (pad >= 0)";
    return false;
  }
  if (g.Cin % 8 || g.Cout % 8) {
    why = "conv: channel counts must be multiples of 8 (NHWC8 layout), got Cin=" + std::to_string(g.Cin) +
          " Cout=" + std::to_string(g.Cout);
    return false;
  }
  if (g.OH <= 0 || g.OW <= 0) {
    why = "conv: kernel larger than the padded input";
    return false;
  }
  if (int64_t(g.B) * g.H * g.W * g.Cin >= (int64_t(1) << 40)) {
    why = "conv: input too large";
    return false;
  }
  return true;
}

ConvGeom make_geom(int B, int H, int W, int Cin, int Cout, int ks, int stride, int pad) {
  ConvGeom g{B, H, W, Cin, 0, 0, Cout, ks, stride, pad};
  if (stride > 0) {
    g.OH = (H + 2 * pad - ks) / stride + 1;
    g.OW = (W + 2 * pad - ks) / stride + 1;
  }
  return g;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

int hip_status(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return gvv::set_last_error(std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

template <int MODE, int BP, int BQ>
void launch(const ConvArgs& a, int splits, hipStream_t s) {
  int64_t P, Q, K;
  gemm_dims(MODE, a.g, P, Q, K);
  dim3 grid(unsigned((Q + BQ - 1) / BQ), unsigned((P + BP - 1) / BP), unsigned(splits));
  hipLaunchKernelGGL((k_conv<MODE, BP, BQ>), grid, dim3(kThreads), 0, s, a);
}

// Tile choice: narrow P (32 channels) takes 32 x 128; otherwise 64 x 128 when Q is large
// enough to fill the chip, else 64 x 64.
template <int MODE>
void launch_tiled(const ConvArgs& a, int splits, hipStream_t s) {
  int64_t P, Q, K;
  gemm_dims(MODE, a.g, P, Q, K);
  if (P <= 32) launch<MODE, 32, 128>(a, splits, s);
  else if (Q >= 64 * 1024) launch<MODE, 64, 128>(a, splits, s);
  else launch<MODE, 64, 64>(a, splits, s);
}

// Split of the weight-gradient reduction: enough (P, Q) tiles x splits to fill 256 CUs a few
// times over, each split at least 8 k-steps (256 pixels) and at most 1024 splits; k_split_sum
// reduces the partials in parallel over outputs and splits.
void weight_split(const ConvGeom& g, int& splits, int64_t& per) {
  int64_t P, Q, K;
  gemm_dims(kBwdWeight, g, P, Q, K);
  const int BP = P <= 32 ? 32 : 64, BQ = P <= 32 ? 128 : (Q >= 64 * 1024 ? 128 : 64);
  const int64_t tiles = ((P + BP - 1) / BP) * ((Q + BQ - 1) / BQ);
  const int64_t steps = (K + kBK - 1) / kBK;
  int64_t want = (2048 + tiles - 1) / tiles;
  if (want > steps / 8) want = steps / 8;
  if (want > 1024) want = 1024;
  if (want < 1) want = 1;
  per = ((steps + want - 1) / want) * kBK;
  splits = int((K + per - 1) / per);
}

constexpr int64_t kBiasPer = 1024;   // pixels per bias-gradient partial

}  // namespace

extern "C" {

int gvv_conv_out_size(int32_t in, int32_t ks, int32_t stride, int32_t pad) {
  if (stride <= 0) return 0;
  return (in + 2 * pad - ks) / stride + 1;
}

int gvv_conv_forward(int32_t B, int32_t H, int32_t W, int32_t Cin, int32_t Cout, int32_t ks, int32_t stride,
                     int32_t pad, const void* x, const void* w, const float* bias, int32_t relu, void* y,
                     int32_t y_f32, void* stream) {
  const ConvGeom g = make_geom(B, H, W, Cin, Cout, ks, stride, pad);
  std::string why;
  if (!geom_ok(g, why)) return gvv::set_last_error(why);
  if (!x || !w || !bias || !y) return gvv::set_last_error("conv forward: null pointer");
  if (!aligned16(x) || !aligned16(w)) return gvv::set_last_error("conv forward: x and w must be 16-byte aligned");
  ConvArgs a{};
  a.g = g;
  a.x = static_cast<const uint16_t*>(x);
  a.w = static_cast<const uint16_t*>(w);
  a.bias = bias;
  a.out = y;
  a.out_f32 = y_f32 ? 1 : 0;
  a.relu = relu ? 1 : 0;
  launch_tiled<kFwd>(a, 1, static_cast<hipStream_t>(stream));
  return hip_status("conv forward");
}

int gvv_conv_backward_data(int32_t B, int32_t H, int32_t W, int32_t Cin, int32_t Cout, int32_t ks, int32_t stride,
                           int32_t pad, const void* dz, const void* wt, const void* x_mask, void* dx, void* stream) {
  const ConvGeom g = make_geom(B, H, W, Cin, Cout, ks, stride, pad);
  std::string why;
  if (!geom_ok(g, why)) return gvv::set_last_error(why);
  if (!dz || !wt || !dx) return gvv::set_last_error("conv backward data: null pointer");
  if (!aligned16(dz) || !aligned16(wt)) return gvv::set_last_error("conv backward data: dz and wt must be 16-byte aligned");
  ConvArgs a{};
  a.g = g;
  a.dz = static_cast<const uint16_t*>(dz);
  a.wt = static_cast<const uint16_t*>(wt);
  a.mask = static_cast<const uint16_t*>(x_mask);
  a.out = dx;
  launch_tiled<kBwdData>(a, 1, static_cast<hipStream_t>(stream));
  return hip_status("conv backward data");
}

int64_t gvv_conv_workspace_floats(int32_t B, int32_t H, int32_t W, int32_t Cin, int32_t Cout, int32_t ks,
                                  int32_t stride, int32_t pad) {
  const ConvGeom g = make_geom(B, H, W, Cin, Cout, ks, stride, pad);
  std::string why;
  if (!geom_ok(g, why)) return -1;
  int splits;
  int64_t per;
  weight_split(g, splits, per);
  const int64_t pixels = int64_t(B) * g.OH * g.OW;
  const int64_t bsplits = (pixels + kBiasPer - 1) / kBiasPer;
  return int64_t(splits) * Cout * ks * ks * Cin + bsplits * Cout;
}

int gvv_conv_backward_weight(int32_t B, int32_t H, int32_t W, int32_t Cin, int32_t Cout, int32_t ks, int32_t stride,
                             int32_t pad, const void* x, const void* dz, float* dw, float* db, float* workspace,
                             int64_t workspace_floats, void* stream) {
  const ConvGeom g = make_geom(B, H, W, Cin, Cout, ks, stride, pad);
  std::string why;
  if (!geom_ok(g, why)) return gvv::set_last_error(why);
  if (!x || !dz || !dw || !workspace) return gvv::set_last_error("conv backward weight: null pointer");
  const int64_t need = gvv_conv_workspace_floats(B, H, W, Cin, Cout, ks, stride, pad);
  if (workspace_floats < need)
    return gvv::set_last_error("conv backward weight: workspace holds " + std::to_string(workspace_floats) +
                               " floats, needs " + std::to_string(need));
  hipStream_t s = static_cast<hipStream_t>(stream);
  int splits;
  int64_t per;
  weight_split(g, splits, per);
  ConvArgs a{};
  a.g = g;
  a.x = static_cast<const uint16_t*>(x);
  a.dz = static_cast<const uint16_t*>(dz);
  a.out = workspace;
  a.out_f32 = 1;
  a.k_per_split = per;
  launch_tiled<kBwdWeight>(a, splits, s);
  const int64_t n = int64_t(Cout) * ks * ks * Cin;
  hipLaunchKernelGGL(k_split_sum, dim3(unsigned((n + 63) / 64)), dim3(256), 0, s, workspace, splits, n, dw);
  if (db) {
    const int64_t pixels = int64_t(B) * g.OH * g.OW;
    const int bsplits = int((pixels + kBiasPer - 1) / kBiasPer);
    float* bpart = workspace + int64_t(splits) * n;
    hipLaunchKernelGGL(k_bias_partial, dim3(unsigned((Cout + 63) / 64), unsigned(bsplits)), dim3(256), 0, s,
                       static_cast<const uint16_t*>(dz), pixels, Cout, kBiasPer, bpart);
    hipLaunchKernelGGL(k_split_sum, dim3(unsigned((Cout + 63) / 64)), dim3(256), 0, s, bpart, bsplits,
                       int64_t(Cout), db);
  }
  return hip_status("conv backward weight");
}

int gvv_conv_prepare_weights(int32_t Cout, int32_t taps, int32_t Cin, const float* w, void* w_bf16, void* wt_bf16,
                             void* stream) {
  if (Cout <= 0 || taps <= 0 || Cin <= 0) return gvv::set_last_error("prepare weights: sizes must be positive");
  if (!w || !w_bf16) return gvv::set_last_error("prepare weights: null pointer");
  const int64_t n = int64_t(Cout) * taps * Cin;
  hipLaunchKernelGGL(k_prepare_weights, dim3(unsigned((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     w, Cout, taps, Cin, static_cast<uint16_t*>(w_bf16), static_cast<uint16_t*>(wt_bf16));
  return hip_status("prepare weights");
}

int gvv_pack_mosaic(const float* render, int32_t B, int32_t V, int32_t h, int32_t w, int32_t cols, int32_t down,
                    int32_t features, void* x, void* stream) {
  if (B <= 0 || V <= 0 || h <= 0 || w <= 0 || cols <= 0 || down <= 0)
    return gvv::set_last_error("pack mosaic: sizes must be positive");
  if (h % down || w % down) return gvv::set_last_error("pack mosaic: view size must be a multiple of the downsampling");
  if (features != 0 && features != 1) return gvv::set_last_error("pack mosaic: features must be 0 or 1");
  if (!render || !x || !aligned16(x)) return gvv::set_last_error("pack mosaic: null or misaligned pointer");
  const int rows = (V + cols - 1) / cols;
  const int64_t n = int64_t(B) * rows * (h / down) * cols * (w / down);
  hipLaunchKernelGGL(k_pack_mosaic, dim3(unsigned((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     render, B, V, h, w, cols, down, features, static_cast<uint16_t*>(x));
  return hip_status("pack mosaic");
}

}  // extern "C"
