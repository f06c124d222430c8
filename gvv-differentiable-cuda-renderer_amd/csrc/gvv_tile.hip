// gvv_tile.hip -- the tile-binned depth pass with the resolve fused at tile end (gfx950).
//
// Replaces the per-(view, face) bounding-box passes of the reference
// (cpp/src/Renderer/CUDABasedRasterization.cu:215-255 depth, :262-408 buffers) with two launches:
//   k_bin        one thread per (view, face): the 16x16-pixel screen tiles of the face's
//                bounding box that the conservative edge pre-test can reach get the face id
//                appended to their bin (kBinCap entries per tile).  A face whose box spans
//                more than kBinMaxTiles tiles, or that finds a bin full, goes to the large-face
//                queue instead, which k_raster_large drains into the global visibility keys,
//                marking the tiles it wrote.
//   k_tile       one 256-thread workgroup per tile.  Every thread computes the ray of its pixel
//                once, into LDS; the tile's faces are staged 64 per wavefront, walked with the
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
__device__ __forceinline__ bool tile_reachable(const EdgeTest& e, int tx, int ty, int W, int H) {
  if (!e.on) return true;
  const float x0 = (float)(tx * kScreenTile) + 0.5f, x1 = (float)min(tx * kScreenTile + kScreenTile - 1, W - 1) + 0.5f;
  const float y0 = (float)(ty * kScreenTile) + 0.5f, y1 = (float)min(ty * kScreenTile + kScreenTile - 1, H - 1) + 0.5f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float m = e.c[i] + (e.a[i] > 0.f ? e.a[i] * x1 : e.a[i] * x0) + (e.b[i] > 0.f ? e.b[i] * y1 : e.b[i] * y0);
    if (m < -0.015625f) return false;
  }
  return true;
}

__global__ void __launch_bounds__(256) k_bin(const int32_t* __restrict__ faces, const float4* __restrict__ proj,
                                             unsigned* __restrict__ tile_count, unsigned* __restrict__ bins,
                                             unsigned int* __restrict__ large_count, int4* __restrict__ large, int B,
                                             int C, int N, int F, int W, int H, int tiles_x, int tiles_y, int mesh_id) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool in_range = i < (int64_t)B * C * F;
  bool to_queue = false;
  int f = 0, bc = 0;
  if (in_range) {
    f = (int)(i % F);
    bc = (int)(i / F);
    const int i0 = faces[3 * f], i1 = faces[3 * f + 1], i2 = faces[3 * f + 2];
    const float4* pc = proj + (size_t)bc * N;
    const float4 p0 = pc[i0], p1 = pc[i1], p2 = pc[i2];
    const Box bx = face_box(v3(p0.x, p0.y, p0.z), v3(p1.x, p1.y, p1.z), v3(p2.x, p2.y, p2.z), W, H);
    if (bx.x0 <= bx.x1 && bx.y0 <= bx.y1) {
      const int tx0 = bx.x0 / kScreenTile, tx1 = bx.x1 / kScreenTile;
      const int ty0 = bx.y0 / kScreenTile, ty1 = bx.y1 / kScreenTile;
      if ((tx1 - tx0 + 1) * (ty1 - ty0 + 1) > kBinMaxTiles) {
        to_queue = true;
      } else {
        const EdgeTest et = edge_test(p0, p1, p2);
        for (int ty = ty0; ty <= ty1; ++ty)
          for (int tx = tx0; tx <= tx1; ++tx) {
            if (!tile_reachable(et, tx, ty, W, H)) continue;
            const size_t t = ((size_t)bc * tiles_y + ty) * tiles_x + tx;
            const unsigned slot = atomicAdd(tile_count + t, 1u);
            if (slot < (unsigned)kBinCap) bins[t * kBinCap + slot] = (unsigned)f;
            else to_queue = true;
          }
      }
    }
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
  const unsigned* bins;                // [B][C][tiles_y][tiles_x][kBinCap]
  unsigned* dirty;                     // [B][C][tiles_y][tiles_x], reset here
  int F, tiles_x, tiles_y;
};

// Exact test of a wavefront's candidates against the LDS rays; 64-bit minimum into the LDS
// z-buffer.  Candidate words are (slot << 8 | pixel of the tile).
__device__ __forceinline__ void tile_flush(const TileFace* st, const unsigned* cand, int count, V3 os,
                                           const float* rx, const float* ry, const float* rz,
                                           unsigned long long* zbuf) {
  const int lane = threadIdx.x & 63;
  for (int k = lane; k < count; k += 64) {
    const unsigned c = cand[k];
    const int pix = (int)(c & 255u);
    const TileFace& fs = st[c >> 8];
    const V3 abc = ray_barycentric(fs.t, os, v3(rx[pix], ry[pix], rz[pix]));
    if (!inside(abc)) continue;
    atomicMin(zbuf + pix, (unsigned long long)pack_key(depth_key(abc, fs.z0, fs.z1, fs.z2), fs.f));
  }
}

template <int ALBEDO, bool SHADE>
__global__ void __launch_bounds__(256) k_tile(TileArgs a) {
  __shared__ unsigned long long s_key[kScreenTile * kScreenTile];
  __shared__ float s_rx[kScreenTile * kScreenTile], s_ry[kScreenTile * kScreenTile], s_rz[kScreenTile * kScreenTile];
  __shared__ TileFace s_face[4][64];
  __shared__ unsigned s_cand[4][kTileCand];
  const ResolveArgs& r = a.r;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int tiles = a.tiles_x * a.tiles_y;
  const int64_t t = xcd_contiguous(blockIdx.x, gridDim.x);
  const int bc = __builtin_amdgcn_readfirstlane((int)(t / tiles));
  const int tl = __builtin_amdgcn_readfirstlane((int)(t - (int64_t)bc * tiles));
  const int ty = tl / a.tiles_x, tx = tl - ty * a.tiles_x;
  const int ox = tx * kScreenTile, oy = ty * kScreenTile;
  const int px = ox + (tid & (kScreenTile - 1)), py = oy + tid / kScreenTile;
  const bool in_img = px < r.W && py < r.H;
  const int64_t pi = ((int64_t)bc * r.H + (in_img ? py : 0)) * r.W + (in_img ? px : 0);
  const Camera& cam = r.cams[bc];
  const unsigned count = a.tile_count[t], dirty = a.dirty[t];
  const int n = (int)min(count, (unsigned)kBinCap);
  // this pixel's ray, once (bit-identical to the face-binned pass's per-candidate ray)
  const V3 d = pixel_ray(cam, px + 0.5f, py + 0.5f);
  s_rx[tid] = d.x; s_ry[tid] = d.y; s_rz[tid] = d.z;
  unsigned long long key = kEmptyKey;
  if (dirty && in_img) {
    key = r.keys[pi];
    if (key != kEmptyKey) r.keys[pi] = kEmptyKey;
  }
  s_key[tid] = key;
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
    const unsigned* bin = a.bins + (size_t)t * kBinCap;
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
        bw = min(bx.x1, ox + kScreenTile - 1) - ox - bx0 + 1;
        bh = min(bx.y1, oy + kScreenTile - 1) - oy - by0 + 1;
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
        const unsigned word = ((unsigned)lane << 8) | (unsigned)((by0 + dy) * kScreenTile + bx0 + dx);
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
            word = ((unsigned)jl << 8) | (unsigned)((jy0 + qy) * kScreenTile + jx0 + qx);
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
  if (!in_img) return;
  key = s_key[tid];
  if (key == kEmptyKey) background_pixel(r, pi);
  else shade_pixel<ALBEDO, SHADE>(r, bc, pi, (int)(uint32_t)(key & 0xffffffffu), d);
}

template <int ALBEDO, bool SHADE>
static void launch_tile_t(const TileArgs& a, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL((k_tile<ALBEDO, SHADE>), dim3((unsigned)n), dim3(256), 0, s, a);
}

hipError_t launch_bin(const ForwardPlan& p, hipStream_t s) {
  const int64_t n = (int64_t)p.B * p.C * p.F;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bin, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p.faces, p.proj, p.tile_count, p.bins,
                     p.counters, p.large, p.B, p.C, p.N, p.F, p.r.W, p.r.H, p.marks.tiles_x, p.marks.tiles_y,
                     p.mesh_id);
  return hipGetLastError();
}

hipError_t launch_tile(const ForwardPlan& p, hipStream_t s) {
  const int64_t n = (int64_t)p.B * p.C * p.marks.tiles_x * p.marks.tiles_y;
  if (n <= 0) return hipSuccess;
  if (n > (int64_t)UINT32_MAX) return hipErrorInvalidConfiguration;
  TileArgs a;
  a.r = p.r;
  a.tris = p.tris; a.proj = p.proj;
  a.tile_count = p.tile_count; a.bins = p.bins; a.dirty = p.marks.dirty;
  a.F = p.F; a.tiles_x = p.marks.tiles_x; a.tiles_y = p.marks.tiles_y;
  switch (p.albedo * 2 + (p.shade ? 1 : 0)) {
    case kVertexColor * 2 + 0: launch_tile_t<kVertexColor, false>(a, n, s); break;
    case kVertexColor * 2 + 1: launch_tile_t<kVertexColor, true>(a, n, s); break;
    case kTextured * 2 + 0: launch_tile_t<kTextured, false>(a, n, s); break;
    case kTextured * 2 + 1: launch_tile_t<kTextured, true>(a, n, s); break;
    case kNormal * 2 + 0: launch_tile_t<kNormal, false>(a, n, s); break;
    case kNormal * 2 + 1: launch_tile_t<kNormal, true>(a, n, s); break;
    case kLighting * 2 + 0: launch_tile_t<kLighting, false>(a, n, s); break;
    case kLighting * 2 + 1: launch_tile_t<kLighting, true>(a, n, s); break;
    case kForegroundMask * 2 + 0: launch_tile_t<kForegroundMask, false>(a, n, s); break;
    default: launch_tile_t<kForegroundMask, true>(a, n, s); break;
  }
  return hipGetLastError();
}

}  // namespace gvv
