// gvv_tile.hip -- the tile-binned depth pass with the resolve fused at tile end (gfx950).
//
// Replaces the per-(view, face) bounding-box passes of the reference
// (cpp/src/Renderer/CUDABasedRasterization.cu:215-255 depth, :262-408 buffers) with two launches:
//   k_bin        one thread per (view, face): the screen tiles (16x16, 32x16 or 32x32 pixels) of
//                the face's bounding box that the conservative edge pre-test can reach get the
//                face id appended to their bin; the lanes of a wavefront that append to one
//                tile share one atomic.  A face whose box spans more than kBinMaxTiles tiles,
//                or that finds a bin full, goes to the large-face queue instead, which
//                k_raster_large drains into the global visibility keys, marking the tiles it
//                wrote.
//   k_tile       one 256-thread workgroup per tile.  Every thread computes the rays of its
//                pixels once, into LDS; the tile's faces are staged 64 per wavefront, walked with the
//                edge pre-test, compacted and given the exact ray/triangle test against the
//                LDS rays; the 64-bit (depth key | face id) minimum is kept in an LDS z-buffer
//                (merged with the global keys where the large-face pass marked the tile).
//                Then each thread resolves and shades its own pixel and writes every pixel
//                output.  No per-pixel visibility state goes through HBM.
// Results are bit-identical to the face-binned pass: the same binary32 ray, barycentrics and
// depth key per (face, pixel), and the same order-independent 64-bit minimum.
#include "gvv_kernels.hpp"
#include "gvv_raster.hpp"
#include "gvv_shade.hpp"

namespace gvv {

// LDS-resident state of one staged face of a tile: the scaled ray/triangle constants, the
// projected depths of the depth key, and the face id.
struct TileFace {
  TriSetup t;
  float z0, z1, z2;
  int f;
};

constexpr int kTileCand = 512;         // candidate list of one wavefront (32-bit words)

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Whether any pixel centre of tile (tx, ty) can pass the edge pre-test: for each edge, the
// largest value of its (pixel-normalised) edge function over the tile's pixel-centre
// rectangle, less a 1/64 px allowance for the rounding of this bound.
__device__ __forceinline__ bool tile_reachable(const EdgeTest& e, int tx, int ty, int W, int H, int sx, int sy) {
  if (!e.on) return true;
  const float x0 = (float)(tx << sx) + 0.5f, x1 = (float)min(((tx + 1) << sx) - 1, W - 1) + 0.5f;
  const float y0 = (float)(ty << sy) + 0.5f, y1 = (float)min(((ty + 1) << sy) - 1, H - 1) + 0.5f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float m = e.c[i] + (e.a[i] > 0.f ? e.a[i] * x1 : e.a[i] * x0) + (e.b[i] > 0.f ? e.b[i] * y1 : e.b[i] * y0);
    if (m < -0.015625f) return false;
  }
  return true;
}

// Next tile at or after box-relative position k (row-major over the tile range of width tw)
// that the face can reach; -1 when none is left.
__device__ __forceinline__ int next_tile(const EdgeTest& et, int k, int n, int tw, int tx0, int ty0, int W, int H,
                                         int sx, int sy) {
  for (; k < n; ++k) {
    const int ty = ty0 + k / tw, tx = tx0 + k % tw;
    if (tile_reachable(et, tx, ty, W, H, sx, sy)) return k;
  }
  return -1;
}

__global__ void __launch_bounds__(256) k_bin(const int32_t* __restrict__ faces, const float4* __restrict__ proj,
                                             unsigned* __restrict__ tile_count, unsigned* __restrict__ bins,
                                             unsigned int* __restrict__ large_count, int4* __restrict__ large, int B,
                                             int C, int N, int F, int W, int H, TileMarks tm, int mesh_id) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool in_range = i < (int64_t)B * C * F;
  bool to_queue = false;
  int f = 0, bc = 0, tx0 = 0, ty0 = 0, tw = 1, n = 0, k = -1;
  EdgeTest et;
  et.on = false;
  if (in_range) {
    f = (int)(i % F);
    bc = (int)(i / F);
    const int i0 = faces[3 * f], i1 = faces[3 * f + 1], i2 = faces[3 * f + 2];
    const float4* pc = proj + (size_t)bc * N;
    const float4 p0 = pc[i0], p1 = pc[i1], p2 = pc[i2];
    const Box bx = face_box(v3(p0.x, p0.y, p0.z), v3(p1.x, p1.y, p1.z), v3(p2.x, p2.y, p2.z), W, H);
    if (bx.x0 <= bx.x1 && bx.y0 <= bx.y1) {
      tx0 = bx.x0 >> tm.sx;
      ty0 = bx.y0 >> tm.sy;
      tw = (bx.x1 >> tm.sx) - tx0 + 1;
      n = tw * ((bx.y1 >> tm.sy) - ty0 + 1);
      if (n > kBinMaxTiles) {
        to_queue = true;
        n = 0;
      } else {
        et = edge_test(p0, p1, p2);
        k = next_tile(et, 0, n, tw, tx0, ty0, W, H, tm.sx, tm.sy);
      }
    }
  }
  // every lane's reachable tiles in turn; the lanes that append to one tile share one atomic
  const size_t view_tiles = (size_t)tm.tiles_x * tm.tiles_y;
  while (__ballot(k >= 0)) {
    const size_t t = k >= 0 ? (size_t)bc * view_tiles + (size_t)(ty0 + k / tw) * tm.tiles_x + (tx0 + k % tw) : 0;
    for (uint64_t pending = __ballot(k >= 0); pending;) {
      const int leader = __ffsll((unsigned long long)pending) - 1;
      const size_t tl = ((size_t)__builtin_amdgcn_readlane((unsigned)(t >> 32), leader) << 32) |
                        (size_t)__builtin_amdgcn_readlane((unsigned)t, leader);
      const bool mine = k >= 0 && t == tl;
      const uint64_t m = __ballot(mine);
      unsigned base = 0;
      if (lane == leader) base = atomicAdd(tile_count + tl, (unsigned)__popcll(m));
      base = __builtin_amdgcn_readlane(base, leader);
      if (mine) {
        const unsigned slot = base + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (slot < (unsigned)tm.bin_cap) bins[tl * tm.bin_cap + slot] = (unsigned)f;
        else to_queue = true;
      }
      pending &= ~m;
    }
    if (k >= 0) k = next_tile(et, k + 1, n, tw, tx0, ty0, W, H, tm.sx, tm.sy);
  }
  // the faces for the large-face pass, one queue atomic per wavefront
  const uint64_t qm = __ballot(to_queue);
  if (qm) {
    const int first = __ffsll((unsigned long long)qm) - 1;
    unsigned int base = 0;
    if (lane == first) base = atomicAdd(large_count, (unsigned)__popcll(qm));
    base = __builtin_amdgcn_readlane(base, first);
    if (to_queue) {
      const unsigned off = __builtin_amdgcn_mbcnt_hi((unsigned)(qm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)qm, 0u));
      large[base + off] = make_int4(bc, f, mesh_id, 0);
    }
  }
}

struct TileArgs {
  ResolveArgs r;                       // shading inputs and pixel outputs; global keys
  const TriSetup* tris;                // [B][F]
  const float4* proj;                  // [B][C][N]
  unsigned* tile_count;                // [B][C][tiles_y][tiles_x], reset here
  const unsigned* bins;                // [B][C][tiles_y][tiles_x][bin_cap]
  unsigned* dirty;                     // [B][C][tiles_y][tiles_x], reset here
  int F, tiles_x, tiles_y, bin_cap;
};

// Exact test of a wavefront's candidates against the LDS rays; 64-bit minimum into the LDS
// z-buffer.  Candidate words are (slot << 10 | pixel of the tile).
__device__ __forceinline__ void tile_flush(const TileFace* st, const unsigned* cand, int count, V3 os,
                                           const float* rx, const float* ry, const float* rz,
                                           unsigned long long* zbuf) {
  const int lane = threadIdx.x & 63;
  for (int k = lane; k < count; k += 64) {
    const unsigned c = cand[k];
    const int pix = (int)(c & 1023u);
    const TileFace& fs = st[c >> 10];
    const V3 abc = ray_barycentric(fs.t, os, v3(rx[pix], ry[pix], rz[pix]));
    if (!inside(abc)) continue;
    atomicMin(zbuf + pix, (unsigned long long)pack_key(depth_key(abc, fs.z0, fs.z1, fs.z2), fs.f));
  }
}

// One workgroup of 256 threads per TW x TH tile; thread tid owns pixels tid, tid + 256, ...
template <int ALBEDO, bool SHADE, int TW, int TH>
__global__ void __launch_bounds__(256) k_tile(TileArgs a) {
  constexpr int TP = TW * TH, PPT = TP / 256;
  static_assert(TP % 256 == 0 && TP <= 1024, "tile pixels: a multiple of 256, at most 1024");
  __shared__ unsigned long long s_key[TP];
  __shared__ float s_rx[TP], s_ry[TP], s_rz[TP];
  __shared__ TileFace s_face[4][64];
  __shared__ unsigned s_cand[4][kTileCand];
  const ResolveArgs& r = a.r;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int tiles = a.tiles_x * a.tiles_y;
  const int64_t t = xcd_contiguous(blockIdx.x, gridDim.x);
  const int bc = __builtin_amdgcn_readfirstlane((int)(t / tiles));
  const int tl = __builtin_amdgcn_readfirstlane((int)(t - (int64_t)bc * tiles));
  const int ty = tl / a.tiles_x, tx = tl - ty * a.tiles_x;
  const int ox = tx * TW, oy = ty * TH;
  const Camera& cam = r.cams[bc];
  const unsigned count = a.tile_count[t], dirty = a.dirty[t];
  const int n = (int)min(count, (unsigned)a.bin_cap);
  // the rays of this thread's pixels, once (bit-identical to the face-binned pass's
  // per-candidate rays), and the z-buffer's start: empty, or the large-face pass's keys
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const int pix = tid + 256 * q, px = ox + pix % TW, py = oy + pix / TW;
    const V3 d = pixel_ray(cam, px + 0.5f, py + 0.5f);
    s_rx[pix] = d.x; s_ry[pix] = d.y; s_rz[pix] = d.z;
    unsigned long long key = kEmptyKey;
    if (dirty && px < r.W && py < r.H) {
      const int64_t pi = ((int64_t)bc * r.H + py) * r.W + px;
      key = r.keys[pi];
      if (key != kEmptyKey) r.keys[pi] = kEmptyKey;
    }
    s_key[pix] = key;
  }
  __syncthreads();
  if (tid == 0) {
    a.tile_count[t] = 0u;
    a.dirty[t] = 0u;
  }
  if (n > 0) {
    const V3 os = cam_origin_scaled(cam);
    const int b = bc / r.C;
    const float4* pc = a.proj + (size_t)bc * r.N;
    TileFace* st = s_face[w];
    unsigned* cand = s_cand[w];
    const unsigned* bin = a.bins + (size_t)t * a.bin_cap;
    for (int base = 0; base < n; base += 256) {
      // stage this wavefront's faces: box clipped to the tile, edge pre-test
      const int j = base + w * 64 + lane;
      bool staged = false;
      EdgeTest et;
      et.on = false;
      int bx0 = 0, by0 = 0, bw = 0, bh = 0;
      if (j < n) {
        const int f = (int)bin[j];
        const int i0 = r.faces[3 * f], i1 = r.faces[3 * f + 1], i2 = r.faces[3 * f + 2];
        const float4 p0 = pc[i0], p1 = pc[i1], p2 = pc[i2];
        const Box bx = face_box(v3(p0.x, p0.y, p0.z), v3(p1.x, p1.y, p1.z), v3(p2.x, p2.y, p2.z), r.W, r.H);
        bx0 = max(bx.x0, ox) - ox; by0 = max(bx.y0, oy) - oy;
        bw = min(bx.x1, ox + TW - 1) - ox - bx0 + 1;
        bh = min(bx.y1, oy + TH - 1) - oy - by0 + 1;
        if (bw > 0 && bh > 0) {
          staged = true;
          TileFace s;
          s.t = a.tris[(size_t)b * a.F + f];
          s.z0 = p0.z; s.z1 = p1.z; s.z2 = p2.z;
          s.f = f;
          st[lane] = s;
          et = edge_test(p0, p1, p2);
        } else {
          bw = bh = 0;
        }
      }
      wave_sync();
      int cnt = 0;
      // boxes up to kTinyTileBox pixels of the tile: each lane walks its own, all lanes at once
      constexpr int kTinyTileBox = 64;
      const int my_area = staged && bw * bh <= kTinyTileBox ? bw * bh : 0;
      int dx = 0, dy = 0;
      for (int p = 0; __ballot(p < my_area); ++p) {
        if (cnt > kTileCand - 64) {
          wave_sync();
          tile_flush(st, cand, cnt, os, s_rx, s_ry, s_rz, s_key);
          wave_sync();
          cnt = 0;
        }
        const bool pass = p < my_area && et.pass(ox + bx0 + dx + 0.5f, oy + by0 + dy + 0.5f);
        const unsigned word = ((unsigned)lane << 10) | (unsigned)((by0 + dy) * TW + bx0 + dx);
        if (++dx == bw) { dx = 0; ++dy; }
        const uint64_t m = __ballot(pass);
        if (pass) cand[cnt + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = word;
        cnt += __popcll(m);
      }
      // larger clipped boxes: the whole wavefront walks one at a time, 64 pixels per step
      for (uint64_t todo = __ballot(staged && my_area == 0); todo; todo &= todo - 1) {
        const int jl = __ffsll((unsigned long long)todo) - 1;
        EdgeTest ej;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          ej.a[e] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(et.a[e]), jl));
          ej.b[e] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(et.b[e]), jl));
          ej.c[e] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(et.c[e]), jl));
        }
        ej.on = __builtin_amdgcn_readlane((int)et.on, jl) != 0;
        const int jx0 = __builtin_amdgcn_readlane(bx0, jl), jy0 = __builtin_amdgcn_readlane(by0, jl);
        const int jbw = __builtin_amdgcn_readlane(bw, jl), jarea = jbw * __builtin_amdgcn_readlane(bh, jl);
        for (int q0 = 0; q0 < jarea; q0 += 64) {
          if (cnt > kTileCand - 64) {
            wave_sync();
            tile_flush(st, cand, cnt, os, s_rx, s_ry, s_rz, s_key);
            wave_sync();
            cnt = 0;
          }
          const int q = q0 + lane;
          bool pass = false;
          unsigned word = 0;
          if (q < jarea) {
            const int qy = q / jbw, qx = q - qy * jbw;
            pass = ej.pass(ox + jx0 + qx + 0.5f, oy + jy0 + qy + 0.5f);
            word = ((unsigned)jl << 10) | (unsigned)((jy0 + qy) * TW + jx0 + qx);
          }
          const uint64_t m = __ballot(pass);
          if (pass) cand[cnt + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = word;
          cnt += __popcll(m);
        }
      }
      wave_sync();
      tile_flush(st, cand, cnt, os, s_rx, s_ry, s_rz, s_key);
      wave_sync();
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const int pix = tid + 256 * q, px = ox + pix % TW, py = oy + pix / TW;
    if (px >= r.W || py >= r.H) continue;
    const int64_t pi = ((int64_t)bc * r.H + py) * r.W + px;
    const unsigned long long key = s_key[pix];
    if (key == kEmptyKey) background_pixel(r, pi);
    else shade_pixel<ALBEDO, SHADE>(r, bc, pi, (int)(uint32_t)(key & 0xffffffffu), v3(s_rx[pix], s_ry[pix], s_rz[pix]));
  }
}

template <int ALBEDO, bool SHADE>
static void launch_tile_t(const TileArgs& a, int sx, int sy, int64_t n, hipStream_t s) {
  if (sx == 5 && sy == 5) hipLaunchKernelGGL((k_tile<ALBEDO, SHADE, 32, 32>), dim3((unsigned)n), dim3(256), 0, s, a);
  else if (sx == 5) hipLaunchKernelGGL((k_tile<ALBEDO, SHADE, 32, 16>), dim3((unsigned)n), dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_tile<ALBEDO, SHADE, 16, 16>), dim3((unsigned)n), dim3(256), 0, s, a);
}

hipError_t launch_bin(const ForwardPlan& p, hipStream_t s) {
  const int64_t n = (int64_t)p.B * p.C * p.F;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bin, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p.faces, p.proj, p.tile_count, p.bins,
                     p.counters, p.large, p.B, p.C, p.N, p.F, p.r.W, p.r.H, p.marks, p.mesh_id);
  return hipGetLastError();
}

hipError_t launch_tile(const ForwardPlan& p, hipStream_t s) {
  const int64_t n = (int64_t)p.B * p.C * p.marks.tiles_x * p.marks.tiles_y;
  if (n <= 0) return hipSuccess;
  if (n > (int64_t)UINT32_MAX) return hipErrorInvalidConfiguration;
  if (!((p.marks.sx == 4 && p.marks.sy == 4) || (p.marks.sx == 5 && (p.marks.sy == 4 || p.marks.sy == 5))))
    return hipErrorInvalidValue;
  TileArgs a;
  a.r = p.r;
  a.tris = p.tris; a.proj = p.proj;
  a.tile_count = p.tile_count; a.bins = p.bins; a.dirty = p.marks.dirty;
  a.F = p.F; a.tiles_x = p.marks.tiles_x; a.tiles_y = p.marks.tiles_y; a.bin_cap = p.marks.bin_cap;
  const int sx = p.marks.sx, sy = p.marks.sy;
  switch (p.albedo * 2 + (p.shade ? 1 : 0)) {
    case kVertexColor * 2 + 0: launch_tile_t<kVertexColor, false>(a, sx, sy, n, s); break;
    case kVertexColor * 2 + 1: launch_tile_t<kVertexColor, true>(a, sx, sy, n, s); break;
    case kTextured * 2 + 0: launch_tile_t<kTextured, false>(a, sx, sy, n, s); break;
    case kTextured * 2 + 1: launch_tile_t<kTextured, true>(a, sx, sy, n, s); break;
    case kNormal * 2 + 0: launch_tile_t<kNormal, false>(a, sx, sy, n, s); break;
    case kNormal * 2 + 1: launch_tile_t<kNormal, true>(a, sx, sy, n, s); break;
    case kLighting * 2 + 0: launch_tile_t<kLighting, false>(a, sx, sy, n, s); break;
    case kLighting * 2 + 1: launch_tile_t<kLighting, true>(a, sx, sy, n, s); break;
    case kForegroundMask * 2 + 0: launch_tile_t<kForegroundMask, false>(a, sx, sy, n, s); break;
    default: launch_tile_t<kForegroundMask, true>(a, sx, sy, n, s); break;
  }
  return hipGetLastError();
}

}  // namespace gvv
