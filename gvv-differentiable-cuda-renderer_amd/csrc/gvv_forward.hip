// gvv_forward.hip -- forward kernels of the rasterizer (gfx950).
//
// Replaces renderBuffersGPU (reference cpp/src/Renderer/CUDABasedRasterization.cu:449-473):
//   k_setup         one launch for the short stages: per-view inverse matrices and ray origin
//                   (.cu:23-67); unnormalised world-space face normals and the scaled
//                   ray/triangle constants, once per mesh and face rather than per view
//                   (.cu:122-141); zero fills (the backward's gradient buffers)
//   k_vertices      vertex normals by CSR gather + per-view projection  (.cu:98-115,148-174)
//   k_raster        per (view, face) bounding-box walk; 64-bit (depth key | face id)
//                   atomicMin into the visibility buffer               (.cu:184-255)
//   k_resolve       one thread per pixel: decode the winner, recompute its barycentrics,
//                   shade, write all pixel outputs, reset the visibility key (.cu:74-91,262-408)
//   k_normal_map    texel-space normal map                             (.cu:415-445)
#include <algorithm>

#include "gvv_kernels.hpp"
#include "gvv_raster.hpp"
#include "gvv_shade.hpp"

namespace gvv {

// The setup stages of one call as one launch (SetupArgs): blocks [0, cb) set up cameras (block
// 0 also clears the large-face counters), blocks [cb, cb + fb) set up faces, and the rest
// zero-fill the listed buffers, 16 bytes per lane where a buffer is 16-byte aligned.
__device__ __forceinline__ void setup_body(const SetupGrid& g, int blk) {
  const SetupArgs& a = g.a;
  const int t = threadIdx.x;
  if (blk < g.cam_blocks) {
    if (blk == 0 && a.counters && t < 2) a.counters[t] = 0u;
    const int i = blk * 256 + t;
    if (i < a.n_cams) {
      Camera c;
      camera_setup(a.extr + 12 * (size_t)i, a.intr + 9 * (size_t)i, c);
      a.cams[i] = c;
    }
    return;
  }
  blk -= g.cam_blocks;
  if (blk < g.face_blocks) {
    const int64_t i = (int64_t)blk * 256 + t;
    if (i >= (int64_t)a.B * a.F) return;
    const int b = (int)(i / a.F), f = (int)(i % a.F);
    const float* vp = a.vpos + (size_t)b * a.N * 3;
    const V3 v0 = ld3(vp + 3 * a.faces[3 * f]), v1 = ld3(vp + 3 * a.faces[3 * f + 1]),
             v2 = ld3(vp + 3 * a.faces[3 * f + 2]);
    if (a.fnormal) {
      const V3 n = cross(v1 - v0, v2 - v0);
      float* o = a.fnormal + 3 * i;
      o[0] = n.x; o[1] = n.y; o[2] = n.z;
    }
    if (a.tris) a.tris[i] = tri_setup(v0, v1, v2);
    if (a.face_grad) {
      FaceGrad r;
      r.t = tri_setup(v0, v1, v2);
      r.g = bary_face_setup(v0, v1, v2);
      a.face_grad[i] = r;
    }
    return;
  }
  blk -= g.face_blocks;
  const int64_t start = (int64_t)blk * 256 + t, stride = (int64_t)g.fill_blocks * 256;
  for (int k = 0; k < kFillSegs; ++k) {
    float* p = a.fill[k];
    const int64_t n = a.fill_n[k];
    if (!p || n <= 0) continue;
    int64_t head = 0;
    if (((uintptr_t)p & 15) == 0) {
      const int64_t n4 = n >> 2;
      float4* p4 = reinterpret_cast<float4*>(p);
      for (int64_t i = start; i < n4; i += stride) p4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      head = n4 << 2;
    }
    for (int64_t i = head + start; i < n; i += stride) p[i] = 0.f;
  }
}

__global__ void __launch_bounds__(256) k_setup(SetupGrid g) { setup_body(g, blockIdx.x); }

__global__ void __launch_bounds__(256) k_setup_multi(MeshTable<SetupGrid> t) {
  const MeshSlice m = t.find(blockIdx.x);
  const SetupGrid g = t.load(m.mesh);
  setup_body(g, (int)m.local);
}

// One thread per (b, c, vertex): vertex normal as the CSR-ordered sum of incident face
// normals (first term assigned, the rest added), written per view; projection into view c.
// `blk` is the mesh's own 256-thread block index.
__device__ __forceinline__ void vertices_body(const VertexArgs& va, int64_t blk) {
  const float* __restrict__ vpos = va.vpos;
  const float* __restrict__ fnormal = va.fnormal;
  const int32_t* __restrict__ csr_start = va.csr_start;
  const int32_t* __restrict__ csr_count = va.csr_count;
  const int32_t* __restrict__ csr_list = va.csr_list;
  const Camera* __restrict__ cams = va.cams;
  float* __restrict__ vnormal = va.vnormal;
  float4* __restrict__ proj = va.proj;
  const int B = va.B, C = va.C, N = va.N, F = va.F;
  // one thread per (view, vertex)
  const int64_t i = blk * 256 + threadIdx.x;
  const bool valid = i < (int64_t)B * C * N;
  const int v = valid ? (int)(i % N) : 0;
  const int bc = valid ? (int)(i / N) : 0;
  const int b = bc / C;
  const float* fn = fnormal + (size_t)b * F * 3;
  const int s = valid ? csr_start[v] : 0, n = valid ? csr_count[v] : 0;
  // Vertices of up to kHeavy incident faces sum them in their own lane, loading the face ids
  // and then the normals of kBatch faces at a time.  More incident faces (the centre of a fan:
  // hundreds) would make one lane's chain of loads the whole kernel's time, so the wavefront
  // loads their normals 64 at a time and folds them in list order with readlane.  Both keep
  // the reference's summation order.
  constexpr int kBatch = 8, kHeavy = 2 * kBatch;
  V3 acc = v3(0.f, 0.f, 0.f);
  if (n <= kHeavy) {
    for (int k0 = 0; k0 < n; k0 += kBatch) {
      int id[kBatch];
#pragma unroll
      for (int j = 0; j < kBatch; ++j) id[j] = k0 + j < n ? csr_list[s + k0 + j] : 0;
      V3 q[kBatch];
#pragma unroll
      for (int j = 0; j < kBatch; ++j) q[j] = k0 + j < n ? ld3(fn + 3 * id[j]) : v3(0.f, 0.f, 0.f);
#pragma unroll
      for (int j = 0; j < kBatch; ++j) {
        if (k0 + j >= n) break;
        if (k0 + j == 0) acc = q[j];
        else { acc.x = acc.x + q[j].x; acc.y = acc.y + q[j].y; acc.z = acc.z + q[j].z; }
      }
    }
  }
  const int lane = threadIdx.x & 63;
  __shared__ float s_fold[4][64 * 3];   // one row per wavefront of the 256-thread workgroup
  float* fold = s_fold[threadIdx.x >> 6];
  for (uint64_t heavy = __ballot(n > kHeavy); heavy; heavy &= heavy - 1) {
    const int j = __ffsll((unsigned long long)heavy) - 1;
    const int sj = __builtin_amdgcn_readlane(s, j), nj = __builtin_amdgcn_readlane(n, j);
    const float* fnj = fnormal + (size_t)__builtin_amdgcn_readlane(b, j) * F * 3;
    V3 sum = v3(0.f, 0.f, 0.f);
    // 256 entries per pass: the face ids of all four 64-entry chunks, then their normals, are
    // loaded at once (entries past the end repeat the last and are not summed), then folded
    for (int k0 = 0; k0 < nj; k0 += 256) {
      int id[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) id[c] = csr_list[sj + min(k0 + 64 * c + lane, nj - 1)];
      V3 q[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) q[c] = ld3(fnj + 3 * id[c]);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int cnt = min(64, nj - (k0 + 64 * c));
        if (cnt <= 0) break;
        // the chunk goes through this wavefront's LDS row and every lane folds it in list
        // order from there (uniform-address reads pipeline; a readlane per value does not)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        fold[3 * lane + 0] = q[c].x; fold[3 * lane + 1] = q[c].y; fold[3 * lane + 2] = q[c].z;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 8
        for (int t = 0; t < cnt; ++t) {
          const float qx = fold[3 * t], qy = fold[3 * t + 1], qz = fold[3 * t + 2];
          if (k0 + 64 * c + t == 0) sum = v3(qx, qy, qz);
          else { sum.x = sum.x + qx; sum.y = sum.y + qy; sum.z = sum.z + qz; }
        }
      }
    }
    if (lane == j) acc = sum;
  }
  if (!valid) return;
  float* o = vnormal + 3 * i;
  o[0] = acc.x; o[1] = acc.y; o[2] = acc.z;
  const V3 p = project(cams[bc], ld3(vpos + ((size_t)b * N + v) * 3));
  proj[i] = make_float4(p.x, p.y, p.z, 0.f);
}

__global__ void __launch_bounds__(256) k_vertices(VertexArgs va) { vertices_body(va, blockIdx.x); }

// The multi-mesh kernels copy the mesh's arguments out of the table first, so that they stay in
// scalar registers across the kernel's stores.
__global__ void __launch_bounds__(256) k_vertices_multi(MeshTable<VertexArgs> t) {
  const MeshSlice m = t.find(blockIdx.x);
  const VertexArgs a = t.load(m.mesh);
  vertices_body(a, m.local);
}

constexpr int kLargeWorkgroups = 4096;

struct RasterFace {
  int f, i0, i1, i2;
  float4 p0, p1, p2;
  Box bx;
};

__device__ __forceinline__ RasterFace load_face(const int32_t* __restrict__ faces, const float4* __restrict__ pc,
                                                int f, int W, int H) {
  RasterFace r;
  r.f = f;
  r.i0 = faces[3 * f]; r.i1 = faces[3 * f + 1]; r.i2 = faces[3 * f + 2];
  r.p0 = pc[r.i0]; r.p1 = pc[r.i1]; r.p2 = pc[r.i2];
  r.bx = face_box(v3(r.p0.x, r.p0.y, r.p0.z), v3(r.p1.x, r.p1.y, r.p1.z), v3(r.p2.x, r.p2.y, r.p2.z), W, H);
  return r;
}

// Per-face constants of the exact test, staged in LDS for the candidate pass: 64 bytes, so
// that 64 staged faces and the candidate list take 5 KB and eight wavefronts fit per SIMD.
// N.N is recomputed per candidate; the face id, view and box stay in registers (StageOwner).
// Packed as four 16-byte words: (v0, v1.x), (v1.yz, v2.xy), (v2.z, N), (v0.N, z0, z1, z2), with
// v0..v2 and N the scaled vertices and face normal of the ray/triangle test (TriSetup) and z0..z2
// the projected depths.
struct alignas(16) FaceStage {
  float4 q[4];
};

__device__ __forceinline__ void stage_face(FaceStage* dst, const TriSetup& t, float z0, float z1, float z2) {
  dst->q[0] = make_float4(t.v0.x, t.v0.y, t.v0.z, t.v1.x);
  dst->q[1] = make_float4(t.v1.y, t.v1.z, t.v2.x, t.v2.y);
  dst->q[2] = make_float4(t.v2.z, t.N.x, t.N.y, t.N.z);
  dst->q[3] = make_float4(t.dv0N, z0, z1, z2);
}

// Registers of the face a lane staged (its staging slot is that lane), read by the lanes that
// test its candidates with a cross-lane permute.  The large-face pass holds its one face in
// every lane.
struct StageOwner {
  int f, bc, x0, y0, bw;
};

// Candidate list of one wavefront: 32-bit words naming a staged face (its slot) and a pixel.
// Images up to 8192 px on a side pack the pixel itself (slot << 26 | v << 13 | u); larger ones
// keep (slot << 24 | dy * bw + dx) within the face's box, where the depth pass's boxes stay
// below 2^24 px (large_box <= 2^20) and the large-face pass, whose one face is slot 0, gives
// all 32 bits to the box index.
constexpr int kCandCap = 256;        // depth pass
constexpr int kCandCapLarge = 512;   // large-face pass (one staged face, LDS is not the limit)
__device__ __forceinline__ bool cand_packed(int W, int H) { return W <= 8192 && H <= 8192; }
__device__ __forceinline__ unsigned cand_word(bool packed, bool single, int slot, int u, int v, int dx, int dy,
                                              int bw) {
  if (packed) return ((unsigned)slot << 26) | ((unsigned)v << 13) | (unsigned)u;
  const unsigned p = (unsigned)dy * (unsigned)bw + (unsigned)dx;
  return single ? p : (((unsigned)slot << 24) | p);
}

// Exact test of every staged candidate (reference uv2barycentric + inside test + depth key,
// CUDABasedRasterization.cu:237-250), then the 64-bit minimum into the visibility buffer.
// Each lane takes candidates lane, lane+64, ...; all lanes run the expensive test together.
// `single`: the candidates of the large-face pass, which stages one face (slot 0).  Called in
// wave-uniform control flow: the owner reads need every lane active.
__device__ __forceinline__ void flush_candidates(const FaceStage* st, const StageOwner& own, const unsigned* cand,
                                                 int count, const Camera* __restrict__ cams, bool one_view, int bc0,
                                                 unsigned long long* __restrict__ keys, int W, int H, bool single,
                                                 TileMarks marks = TileMarks{}) {
  const int lane = threadIdx.x & 63;
  const bool packed = cand_packed(W, H);
  for (int base = 0; base < count; base += 64) {
    const int i = base + lane;
    const bool valid = i < count;
    const unsigned c = valid ? cand[i] : 0u;
    int slot, u, v;
    if (packed) {
      slot = single ? 0 : (int)(c >> 26);
      v = (int)((c >> 13) & 0x1fffu);
      u = (int)(c & 0x1fffu);
    } else {
      slot = single ? 0 : (int)(c >> 24);
      const unsigned p = single ? c : (c & 0xffffffu);
      const int x0 = __shfl(own.x0, slot, 64), y0 = __shfl(own.y0, slot, 64);
      const unsigned w = (unsigned)__shfl(own.bw, slot, 64);
      const unsigned dy = valid ? p / w : 0u;
      u = x0 + (int)(p - dy * w);
      v = y0 + (int)dy;
    }
    const int f = __shfl(own.f, slot, 64);
    const int bc = one_view ? bc0 : __shfl(own.bc, slot, 64);
    if (!valid) continue;
    const float4 q0 = st[slot].q[0], q1 = st[slot].q[1], q2 = st[slot].q[2], q3 = st[slot].q[3];
    const Camera& cam = cams[bc];
    const V3 d = pixel_ray(cam, u + 0.5f, v + 0.5f);
    TriSetup t;
    t.v0 = v3(q0.x, q0.y, q0.z);
    t.v1 = v3(q0.w, q1.x, q1.y);
    t.v2 = v3(q1.z, q1.w, q2.x);
    t.N = v3(q2.y, q2.z, q2.w);
    t.dv0N = q3.x;
    t.den = dot(t.N, t.N);      // as tri_setup forms it
    const V3 abc = ray_barycentric(t, cam_origin_scaled(cam), d);
    if (!inside(abc)) continue;
    const uint64_t k = pack_key(depth_key(abc, q3.y, q3.z, q3.w), f);
    atomicMin(keys + ((size_t)bc * H + v) * W + u, (unsigned long long)k);
    // screen-tile depth pass: the tile kernel merges this tile's visibility keys
    if (marks.dirty) marks.dirty[((size_t)bc * marks.tiles_y + (v >> marks.sy)) * marks.tiles_x + (u >> marks.sx)] = 1u;
  }
}

// Pre-test the box of staged face `slot` lane-parallel (64 pixels per round) and append the
// passing pixels to the candidate list; flushes first whenever a round could overflow it.
__device__ __forceinline__ void collect_face(const EdgeTest& et_in, int slot, int x0, int y0, int bw, int bh,
                                             FaceStage* st, const StageOwner& own, unsigned* cand, int& count,
                                             const Camera* __restrict__ cams, bool one_view, int bc0,
                                             unsigned long long* __restrict__ keys, int W, int H) {
  const int lane = threadIdx.x & 63;
  const int area = bw * bh;
  for (int base = 0; base < area; base += 64) {
    if (count > kCandCap - 64) {
      __syncthreads();
      flush_candidates(st, own, cand, count, cams, one_view, bc0, keys, W, H, false);
      __syncthreads();
      count = 0;
    }
    const int p = base + lane;
    bool pass = false;
    int dx = 0, dy = 0;
    if (p < area) {
      dy = p / bw;
      dx = p - dy * bw;
      pass = et_in.pass(x0 + dx + 0.5f, y0 + dy + 0.5f);
    }
    const uint64_t m = __ballot(pass);
    if (pass) {
      const int off = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
      cand[count + off] = cand_word(cand_packed(W, H), false, slot, x0 + dx, y0 + dy, dx, dy, bw);
    }
    count += __popcll(m);
  }
}

__device__ __forceinline__ EdgeTest readlane_edges(const EdgeTest& e, int j) {
  EdgeTest r;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    r.a[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.a[i]), j));
    r.b[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.b[i]), j));
    r.c[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.c[i]), j));
  }
  r.on = __builtin_amdgcn_readlane((int)e.on, j) != 0;
  return r;
}

// Boxes up to this many pixels are walked by their own lane (all lanes at once, pixel by
// pixel along the box rows); larger staged boxes are walked by the whole wavefront, one face
// at a time.  Per-lane walks share each iteration among all the wavefront's small faces, while
// the wavefront walk pays a per-face setup: at 100k triangles over 8 x 1024^2 views (most
// boxes 17-128 px) 64 was the fastest of 16, 32, 64, 96, 128 and 256.
constexpr int kTinyBox = 64;

// One wavefront per 64 consecutive (b, c, face) items (wg: the mesh's own logical block).
// Faces with bounding boxes up to large_box pixels are staged and rasterised here (pre-test,
// compaction, exact test); larger ones are appended to `large` for k_raster_large.
__device__ __forceinline__ void raster_body(const RasterArgs& ra, int64_t wg) {
  const TriSetup* __restrict__ tris = ra.tris;
  const int32_t* __restrict__ faces = ra.faces;
  const float4* __restrict__ proj = ra.proj;
  const Camera* __restrict__ cams = ra.cams;
  unsigned long long* __restrict__ keys = ra.keys;
  unsigned int* __restrict__ large_count = ra.large_count;
  int4* __restrict__ large = ra.large;
  const int B = ra.B, C = ra.C, N = ra.N, F = ra.F, W = ra.W, H = ra.H, large_box = ra.large_box, mesh_id = ra.mesh_id;
  __shared__ FaceStage st[64];
  __shared__ unsigned cand[kCandCap];
  const int lane = threadIdx.x;
  const int64_t i = wg * 64 + lane;
  const bool in_range = i < (int64_t)B * C * F;
  const int f = in_range ? (int)(i % F) : 0;
  const int bc = in_range ? (int)(i / F) : 0;
  bool small = false, to_queue = false;
  EdgeTest et;
  et.on = false;
  int bw = 0, bh = 0, fx0 = 0, fy0 = 0;
  if (in_range) {
    const RasterFace r = load_face(faces, proj + (size_t)bc * N, f, W, H);
    if (r.bx.x0 <= r.bx.x1 && r.bx.y0 <= r.bx.y1) {
      bw = r.bx.x1 - r.bx.x0 + 1;
      bh = r.bx.y1 - r.bx.y0 + 1;
      if ((int64_t)bw * bh > large_box) {
        to_queue = true;
      } else {
        small = true;
        stage_face(&st[lane], tris[(size_t)(bc / C) * F + f], r.p0.z, r.p1.z, r.p2.z);
        et = edge_test(r.p0, r.p1, r.p2);
        fx0 = r.bx.x0;
        fy0 = r.bx.y0;
      }
    }
  }
  // large faces go to the queue with one atomic per wavefront (the cursor is a single word
  // that every wavefront of the launch would otherwise hit once per face)
  if (!ra.queue_large) to_queue = false;
  const uint64_t qm = __ballot(to_queue);
  if (qm) {
    unsigned int base = 0;
    if (lane == __ffsll((unsigned long long)qm) - 1) base = atomicAdd(large_count, (unsigned)__popcll(qm));
    base = __builtin_amdgcn_readlane(base, __ffsll((unsigned long long)qm) - 1);
    if (to_queue) {
      const unsigned off = __builtin_amdgcn_mbcnt_hi((unsigned)(qm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)qm, 0u));
      large[base + off] = make_int4(bc, f, mesh_id, 0);
    }
  }
  if (!__ballot(small)) return;
  const int bc0 = __builtin_amdgcn_readfirstlane(bc);
  const bool one_view = __ballot(bc != bc0 && in_range) == 0;
  const StageOwner own{f, bc, fx0, fy0, bw};
  __syncthreads();
  int count = 0;
  // tiny boxes: each lane walks its own box; pixel p of every lane's face per iteration
  const int my_area = small && bw * bh <= kTinyBox ? bw * bh : 0;
  int dx = 0, dy = 0;     // pixel p of this lane's box, stepped along its rows
  EdgeWalk walk;
  walk.start(et, fx0, fy0, bw);
  for (int p = 0; __ballot(p < my_area); ++p) {
    if (count > kCandCap - 64) {
      __syncthreads();
      flush_candidates(st, own, cand, count, cams, one_view, bc0, keys, W, H, false);
      __syncthreads();
      count = 0;
    }
    const bool pass = p < my_area && walk.pass(et.on);
    const unsigned word = cand_word(cand_packed(W, H), false, lane, fx0 + dx, fy0 + dy, dx, dy, bw);
    const bool row_end = ++dx == bw;
    if (row_end) { dx = 0; ++dy; }
    walk.step(et, row_end);
    const uint64_t m = __ballot(pass);
    if (pass) {
      const int off = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
      cand[count + off] = word;
    }
    count += __popcll(m);
  }
  // larger staged boxes: the whole wavefront walks one box at a time
  uint64_t todo = __ballot(small && my_area == 0);
  while (todo) {
    const int j = __ffsll((unsigned long long)todo) - 1;
    todo &= todo - 1;
    const EdgeTest ej = readlane_edges(et, j);
    const int x0 = __builtin_amdgcn_readlane(fx0, j), y0 = __builtin_amdgcn_readlane(fy0, j);
    const int jbw = __builtin_amdgcn_readlane(bw, j), jbh = __builtin_amdgcn_readlane(bh, j);
    collect_face(ej, j, x0, y0, jbw, jbh, st, own, cand, count, cams, one_view, bc0, keys, W, H);
  }
  __syncthreads();
  flush_candidates(st, own, cand, count, cams, one_view, bc0, keys, W, H, false);
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) k_raster(RasterArgs ra) { raster_body(ra, xcd_contiguous(blockIdx.x, gridDim.x)); }

// The depth pass's large-face test alone, so that the large-face pass can run beside the depth
// pass: the (view, face) items whose bounding boxes exceed the threshold are appended to the
// queue (one atomic per wavefront), by the same box computation as k_raster, which then skips
// them.
__global__ void __launch_bounds__(64) k_classify_large(RasterArgs ra) {
  const int lane = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  bool to_queue = false;
  int f = 0, bc = 0;
  if (i < (int64_t)ra.B * ra.C * ra.F) {
    f = (int)(i % ra.F);
    bc = (int)(i / ra.F);
    const RasterFace r = load_face(ra.faces, ra.proj + (size_t)bc * ra.N, f, ra.W, ra.H);
    if (r.bx.x0 <= r.bx.x1 && r.bx.y0 <= r.bx.y1) {
      const int bw = r.bx.x1 - r.bx.x0 + 1, bh = r.bx.y1 - r.bx.y0 + 1;
      to_queue = (int64_t)bw * bh > ra.large_box;
    }
  }
  const uint64_t qm = __ballot(to_queue);
  if (!qm) return;
  unsigned int base = 0;
  if (lane == __ffsll((unsigned long long)qm) - 1) base = atomicAdd(ra.large_count, (unsigned)__popcll(qm));
  base = __builtin_amdgcn_readlane(base, __ffsll((unsigned long long)qm) - 1);
  if (to_queue) {
    const unsigned off = __builtin_amdgcn_mbcnt_hi((unsigned)(qm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)qm, 0u));
    ra.large[base + off] = make_int4(bc, f, ra.mesh_id, 0);
  }
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) k_raster_multi(MeshTable<RasterArgs> t) {
  const MeshSlice m = t.find((unsigned)xcd_contiguous(blockIdx.x, gridDim.x));
  const RasterArgs a = t.load(m.mesh);
  raster_body(a, m.local);
}

// One large face (bc, f) rasterised by the whole wavefront.  Instead of the whole bounding box
// the wavefront walks, 64 rows at a time, the conservative span of each row (row_span; a wave
// prefix sum turns spans into pixel indices), applies the edge pre-test, compacts and runs the
// exact test as k_raster does.
__device__ __forceinline__ void raster_large_face(const TriSetup* __restrict__ tris, const int32_t* __restrict__ faces,
                                                  const float4* __restrict__ proj, const Camera* __restrict__ cams,
                                                  unsigned long long* __restrict__ keys, int C, int N, int F, int W,
                                                  int H, int bc, int f, FaceStage* st, unsigned* cand, int* s_xl,
                                                  int* s_start, TileMarks marks) {
  const RasterFace r = load_face(faces, proj + (size_t)bc * N, f, W, H);
  const int bw = r.bx.x1 - r.bx.x0 + 1, bh = r.bx.y1 - r.bx.y0 + 1;
  __syncthreads();
  if (threadIdx.x == 0) stage_face(&st[0], tris[(size_t)(bc / C) * F + f], r.p0.z, r.p1.z, r.p2.z);
  const StageOwner own{f, bc, r.bx.x0, r.bx.y0, bw};
  const EdgeTest et = edge_test(r.p0, r.p1, r.p2);
  __syncthreads();
  int count = 0;
  const int lane = threadIdx.x;
  for (int rbase = 0; rbase < bh; rbase += 64) {
    // this lane's row: span and length; inclusive scan of the lengths over the wave
    const int y = r.bx.y0 + rbase + lane;
    int xl = 0, xr = -1;
    if (rbase + lane < bh) row_span(et, y, r.bx.x0, r.bx.x1, xl, xr);
    const int len = xr >= xl ? xr - xl + 1 : 0;
    int incl = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(incl, d, 64);
      if (lane >= d) incl += t;
    }
    const int total = __shfl(incl, 63, 64);
    if (total == 0) continue;
    __syncthreads();
    s_xl[lane] = xl;
    s_start[lane] = incl - len;
    __syncthreads();
    for (int q0 = 0; q0 < total; q0 += 64) {
      if (count > kCandCapLarge - 64) {
        __syncthreads();
        flush_candidates(st, own, cand, count, cams, true, bc, keys, W, H, true, marks);
        __syncthreads();
        count = 0;
      }
      const int q = q0 + lane;
      bool pass = false;
      int p = 0;
      if (q < total) {
        // the row holding span pixel q: the last row whose start is <= q
        int row = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
          if (row + step < 64 && s_start[row + step] <= q) row += step;
        const int x = s_xl[row] + (q - s_start[row]);
        const int yy = r.bx.y0 + rbase + row;
        pass = et.pass(x + 0.5f, yy + 0.5f);
        p = (int)cand_word(cand_packed(W, H), true, 0, x, yy, x - r.bx.x0, yy - r.bx.y0, bw);
      }
      const uint64_t m = __ballot(pass);
      if (pass) {
        const int off = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        cand[count + off] = (unsigned)p;
      }
      count += __popcll(m);
    }
  }
  __syncthreads();
  flush_candidates(st, own, cand, count, cams, true, bc, keys, W, H, true, marks);
}

// Persistent wavefronts drain the large-face queue: wavefront w rasterises faces w, w + G,
// w + 2G, ... (G wavefronts) up to the queue length written by k_raster (raster_large_face).
// A shared atomic cursor would balance better, but every wavefront taking every item from one
// word serialises on that word's memory-side atomics: 4k cursor atomics took longer than the
// faces themselves.
__global__ void __launch_bounds__(64) k_raster_large(const TriSetup* __restrict__ tris, const int32_t* __restrict__ faces,
                                                     const float4* __restrict__ proj, const Camera* __restrict__ cams,
                                                     unsigned long long* __restrict__ keys,
                                                     unsigned int* __restrict__ counters, const int4* __restrict__ large,
                                                     int C, int N, int F, int W, int H, TileMarks marks) {
  __shared__ FaceStage st[1];
  __shared__ unsigned cand[kCandCapLarge];
  __shared__ int s_xl[64], s_start[64];
  const unsigned int n_large = counters[0];
  for (unsigned int item = blockIdx.x; item < n_large; item += gridDim.x)
    raster_large_face(tris, faces, proj, cams, keys, C, N, F, W, H, large[item].x, large[item].y, st, cand, s_xl,
                      s_start, marks);
}

// The same over a queue shared by several meshes: each entry names its mesh, whose raster
// state comes from the descriptor table.  One launch drains the large faces of every mesh.
__global__ void __launch_bounds__(64) k_raster_large_multi(const RasterMesh* __restrict__ meshes,
                                                           unsigned int* __restrict__ counters,
                                                           const int4* __restrict__ large) {
  __shared__ FaceStage st[1];
  __shared__ unsigned cand[kCandCapLarge];
  __shared__ int s_xl[64], s_start[64];
  const unsigned int n_large = counters[0];
  for (unsigned int item = blockIdx.x; item < n_large; item += gridDim.x) {
    const int4 e = large[item];
    const RasterMesh m = meshes[e.z];
    raster_large_face(m.tris, m.faces, m.proj, m.cams, m.keys, m.C, m.N, m.F, m.W, m.H, e.x, e.y, st, cand, s_xl,
                      s_start, m.marks);
  }
}

__global__ void k_zero_u32(unsigned int* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

template <int ALBEDO, bool SHADE>
__device__ __forceinline__ void resolve_body(const ResolveArgs& a, unsigned wg) {
  // one workgroup per 256-pixel chunk of one image row: the view, row and chunk follow from the
  // workgroup index alone (uniform arithmetic), so no lane divides its pixel index
  const unsigned chunks = ((unsigned)a.W + 255u) >> 8;
  const unsigned per_view = chunks * (unsigned)a.H;
  const int bc = (int)(wg / per_view);
  const unsigned rem = wg - (unsigned)bc * per_view;
  const int py = (int)(rem / chunks);
  const int px = (int)(rem - (unsigned)py * chunks) * 256 + (int)threadIdx.x;
  if (bc >= a.B * a.C || px >= a.W) return;
  const int64_t i = ((int64_t)bc * a.H + py) * a.W + px;
  unsigned long long key = a.keys[i];
  if (key == kEmptyKey) {
    background_pixel(a, i);
    return;
  }
  st_out(a.keys + i, (unsigned long long)kEmptyKey);   // next read by the next call's depth pass
  shade_pixel<ALBEDO, SHADE>(a, bc, i, (int)(uint32_t)(key & 0xffffffffu), pixel_ray(a.cams[bc], px + 0.5f, py + 0.5f));
}

template <int ALBEDO, bool SHADE>
__global__ void __launch_bounds__(256) k_resolve(ResolveArgs a) {
  resolve_body<ALBEDO, SHADE>(a, (unsigned)xcd_contiguous(blockIdx.x, gridDim.x));
}

template <int ALBEDO, bool SHADE>
__global__ void __launch_bounds__(256) k_resolve_multi(MeshTable<ResolveArgs> t) {
  const MeshSlice m = t.find((unsigned)xcd_contiguous(blockIdx.x, gridDim.x));
  const ResolveArgs a = t.load(m.mesh);
  resolve_body<ALBEDO, SHADE>(a, (unsigned)m.local);
}

__global__ void __launch_bounds__(256) k_normal_map(const UvTexel* __restrict__ table, const int32_t* __restrict__ faces,
                                                    const float* __restrict__ vnormal, float* __restrict__ nmap,
                                                    int B, int C, int N, int T) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * T) return;
  int b = (int)(i / T), t = (int)(i % T);
  UvTexel e = table[t];
  const float* vn = vnormal + (size_t)b * C * N * 3;   // view 0
  int f = e.face;
  V3 n = ld3(vn + 3 * faces[3 * f]) * e.a + ld3(vn + 3 * faces[3 * f + 1]) * e.b + ld3(vn + 3 * faces[3 * f + 2]) * e.c;
  float len = length(n);
  if (len != 0.f) n = n / len;
  float* o = nmap + 3 * i;
  o[0] = (n.x + 1.f) / 2.f; o[1] = (n.y + 1.f) / 2.f; o[2] = (n.z + 1.f) / 2.f;
}

__global__ void __launch_bounds__(256) k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_f32(float* p, int64_t n, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------------------
// host-side launchers
// ------------------------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(n)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, int64_t n, float v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(n)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

SetupGrid setup_grid(const SetupArgs& a) {
  SetupGrid g;
  g.a = a;
  g.cam_blocks = a.n_cams > 0 || a.counters ? (int)blocks_for(a.n_cams > 0 ? a.n_cams : 1) : 0;
  g.face_blocks = a.faces ? (int)blocks_for((int64_t)a.B * a.F) : 0;
  int64_t most = 0;
  for (int k = 0; k < kFillSegs; ++k)
    if (a.fill[k] && a.fill_n[k] > most) most = a.fill_n[k];
  g.fill_blocks = most > 0 ? (int)std::min<int64_t>(2048, (int64_t)blocks_for((most + 3) / 4)) : 0;
  return g;
}

hipError_t launch_setup(const SetupArgs& a, hipStream_t s) {
  const SetupGrid g = setup_grid(a);
  const int total = g.cam_blocks + g.face_blocks + g.fill_blocks;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup, dim3(total), dim3(256), 0, s, g);
  return hipGetLastError();
}

template <int ALBEDO, bool SHADE>
static void launch_resolve_t(const ResolveArgs& a, hipStream_t s) {
  const int64_t n = (int64_t)a.B * a.C * a.H * ((a.W + 255) / 256);   // workgroups: row chunks
  if (n <= 0) return;
  hipLaunchKernelGGL((k_resolve<ALBEDO, SHADE>), dim3((unsigned)n), dim3(256), 0, s, a);
}

SetupArgs forward_setup_args(const ForwardPlan& p) {
  SetupArgs sa;
  sa.extr = p.extr; sa.intr = p.intr; sa.cams = p.cams; sa.n_cams = p.B * p.C;
  sa.vpos = p.vpos; sa.faces = p.faces; sa.fnormal = p.fnormal; sa.tris = p.tris;
  sa.B = p.B; sa.N = p.N; sa.F = p.F;
  sa.counters = p.shared_queue ? nullptr : p.counters;
  return sa;
}

VertexArgs vertex_args(const ForwardPlan& p) {
  VertexArgs a;
  a.vpos = p.vpos; a.fnormal = p.fnormal; a.csr_start = p.csr_start; a.csr_count = p.csr_count;
  a.csr_list = p.csr_list; a.cams = p.cams; a.vnormal = p.vnormal_out; a.proj = p.proj;
  a.B = p.B; a.C = p.C; a.N = p.N; a.F = p.F;
  return a;
}

RasterArgs raster_args(const ForwardPlan& p) {
  RasterArgs a;
  a.tris = p.tris; a.faces = p.faces; a.proj = p.proj; a.cams = p.cams; a.keys = p.r.keys;
  a.large_count = p.counters; a.large = p.large;
  a.B = p.B; a.C = p.C; a.N = p.N; a.F = p.F; a.W = p.r.W; a.H = p.r.H;
  a.large_box = p.large_box; a.mesh_id = p.mesh_id;
  a.queue_large = 1;
  return a;
}

RasterMesh raster_mesh(const ForwardPlan& p) {
  RasterMesh m;
  m.tris = p.tris; m.faces = p.faces; m.proj = p.proj; m.cams = p.cams; m.keys = p.r.keys;
  m.C = p.C; m.N = p.N; m.F = p.F; m.W = p.r.W; m.H = p.r.H;
  m.marks = p.tiled ? p.marks : TileMarks{};
  return m;
}

hipError_t launch_clear_counters(unsigned int* counters, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, counters, 2);
  return hipGetLastError();
}

hipError_t launch_raster_large_multi(const RasterMesh* meshes, unsigned int* counters, const int4* large,
                                     hipStream_t s) {
  hipLaunchKernelGGL(k_raster_large_multi, dim3(kLargeWorkgroups), dim3(64), 0, s, meshes, counters, large);
  return hipGetLastError();
}

hipError_t launch_forward_raster(const ForwardPlan& p, hipStream_t s) {
  const int B = p.B, C = p.C, N = p.N, F = p.F;
  {
    // cameras, face setup and (for a queue of its own) the large-face counters
    ProfScope ps(p.prof, "k_setup", s);
    hipError_t e = launch_setup(forward_setup_args(p), s);
    if (e != hipSuccess) return e;
  }
  {
    ProfScope ps(p.prof, "k_vertices", s);
    hipLaunchKernelGGL(k_vertices, dim3(blocks_for((int64_t)B * C * N)), dim3(256), 0, s, vertex_args(p));
  }
  if (p.normal_map) {
    ProfScope ps(p.prof, "k_normal_map", s);
    int T = p.r.texW * p.r.texH;
    hipLaunchKernelGGL(k_normal_map, dim3(blocks_for((int64_t)B * T)), dim3(256), 0, s, p.uv_table, p.faces,
                       p.vnormal_out, p.normal_map_out, B, C, N, T);
    return hipGetLastError();
  }
  // the large faces classified and rasterised on the handle's second stream, beside the depth
  // pass (both take the 64-bit minimum of the same keys, so the order is free); in line when
  // the caller's stream is being captured into a graph
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  if (p.side && hipStreamIsCapturing(s, &capture) != hipSuccess) capture = hipStreamCaptureStatusActive;
  if (!p.tiled && !p.shared_queue && p.side && capture == hipStreamCaptureStatusNone) {
    hipError_t e = hipEventRecord(p.fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(p.side, p.fork, 0);
    if (e != hipSuccess) return e;
    // from here on the second stream is joined back whatever fails in between
    RasterArgs ra = raster_args(p);
    {
      ProfScope ps(p.prof, "k_classify_large", p.side);
      hipLaunchKernelGGL(k_classify_large, dim3(blocks_for((int64_t)B * C * F, 64)), dim3(64), 0, p.side, ra);
    }
    {
      ProfScope ps(p.prof, "k_raster_large", p.side);
      hipLaunchKernelGGL(k_raster_large, dim3(kLargeWorkgroups), dim3(64), 0, p.side, p.tris, p.faces, p.proj, p.cams,
                         p.r.keys, p.counters, p.large, C, N, F, p.r.W, p.r.H, TileMarks{});
    }
    const hipError_t e_side = hipGetLastError();
    ra.queue_large = 0;
    {
      ProfScope ps(p.prof, "k_raster", s);
      hipLaunchKernelGGL(k_raster, dim3(blocks_for((int64_t)B * C * F, 64)), dim3(64), 0, s, ra);
    }
    const hipError_t e_main = hipGetLastError();
    e = hipEventRecord(p.join, p.side);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, p.join, 0);
    if (e != hipSuccess) return e;
    return e_side != hipSuccess ? e_side : e_main;
  }
  if (p.tiled) {
    ProfScope ps(p.prof, "k_bin", s);
    hipError_t e = launch_bin(p, s);
    if (e != hipSuccess) return e;
  } else {
    ProfScope ps(p.prof, "k_raster", s);
    hipLaunchKernelGGL(k_raster, dim3(blocks_for((int64_t)B * C * F, 64)), dim3(64), 0, s, raster_args(p));
  }
  if (!p.shared_queue) {
    ProfScope ps(p.prof, "k_raster_large", s);
    hipLaunchKernelGGL(k_raster_large, dim3(kLargeWorkgroups), dim3(64), 0, s, p.tris, p.faces, p.proj, p.cams,
                       p.r.keys, p.counters, p.large, C, N, F, p.r.W, p.r.H, p.tiled ? p.marks : TileMarks{});
  }
  return hipGetLastError();
}

hipError_t launch_forward_resolve(const ForwardPlan& p, hipStream_t s) {
  // In normal-map mode the visibility buffer is empty, so this writes the background.
  // One workgroup per 256-pixel row chunk: the count must fit a 32-bit grid dimension.
  if (p.tiled && !p.normal_map) {
    ProfScope ps(p.prof, "k_tile", s);
    return launch_tile(p, s);
  }
  if ((int64_t)p.r.B * p.r.C * p.r.H * ((p.r.W + 255) / 256) > (int64_t)UINT32_MAX) return hipErrorInvalidConfiguration;
  ProfScope ps(p.prof, "k_resolve", s);
  switch (p.albedo * 2 + (p.shade ? 1 : 0)) {
    case kVertexColor * 2 + 0: launch_resolve_t<kVertexColor, false>(p.r, s); break;
    case kVertexColor * 2 + 1: launch_resolve_t<kVertexColor, true>(p.r, s); break;
    case kTextured * 2 + 0: launch_resolve_t<kTextured, false>(p.r, s); break;
    case kTextured * 2 + 1: launch_resolve_t<kTextured, true>(p.r, s); break;
    case kNormal * 2 + 0: launch_resolve_t<kNormal, false>(p.r, s); break;
    case kNormal * 2 + 1: launch_resolve_t<kNormal, true>(p.r, s); break;
    case kLighting * 2 + 0: launch_resolve_t<kLighting, false>(p.r, s); break;
    case kLighting * 2 + 1: launch_resolve_t<kLighting, true>(p.r, s); break;
    case kForegroundMask * 2 + 0: launch_resolve_t<kForegroundMask, false>(p.r, s); break;
    default: launch_resolve_t<kForegroundMask, true>(p.r, s); break;
  }
  return hipGetLastError();
}

hipError_t launch_setup_multi(const MeshTable<SetupGrid>& t, unsigned blocks, hipStream_t s) {
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_setup_multi, dim3(blocks), dim3(256), 0, s, t);
  return hipGetLastError();
}

hipError_t launch_vertices_multi(const MeshTable<VertexArgs>& t, unsigned blocks, hipStream_t s) {
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vertices_multi, dim3(blocks), dim3(256), 0, s, t);
  return hipGetLastError();
}

hipError_t launch_raster_multi(const MeshTable<RasterArgs>& t, unsigned blocks, hipStream_t s) {
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_raster_multi, dim3(blocks), dim3(64), 0, s, t);
  return hipGetLastError();
}

hipError_t launch_resolve_multi(const MeshTable<ResolveArgs>& t, unsigned blocks, int albedo, bool shade,
                                hipStream_t s) {
  if (blocks == 0) return hipSuccess;
  switch (albedo * 2 + (shade ? 1 : 0)) {
    case kVertexColor * 2 + 0: hipLaunchKernelGGL((k_resolve_multi<kVertexColor, false>), dim3(blocks), dim3(256), 0, s, t); break;
    case kVertexColor * 2 + 1: hipLaunchKernelGGL((k_resolve_multi<kVertexColor, true>), dim3(blocks), dim3(256), 0, s, t); break;
    case kTextured * 2 + 0: hipLaunchKernelGGL((k_resolve_multi<kTextured, false>), dim3(blocks), dim3(256), 0, s, t); break;
    case kTextured * 2 + 1: hipLaunchKernelGGL((k_resolve_multi<kTextured, true>), dim3(blocks), dim3(256), 0, s, t); break;
    case kForegroundMask * 2 + 0:
      hipLaunchKernelGGL((k_resolve_multi<kForegroundMask, false>), dim3(blocks), dim3(256), 0, s, t);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

unsigned resolve_blocks(const ResolveArgs& a) { return (unsigned)((int64_t)a.B * a.C * a.H * ((a.W + 255) / 256)); }

hipError_t launch_forward(const ForwardPlan& p, hipStream_t s) {
  hipError_t e = launch_forward_raster(p, s);
  if (e != hipSuccess) return e;
  return launch_forward_resolve(p, s);
}

}  // namespace gvv
