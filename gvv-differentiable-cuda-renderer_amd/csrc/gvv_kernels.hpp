// gvv_kernels.hpp -- launch plans shared by the forward/backward kernels and the C ABI.
#pragma once
#include <string>
#include <vector>

#include "gvv_bary_grad.hpp"
#include "gvv_math.hpp"

namespace gvv {

// Optional per-launch timing with HIP events on the launch stream.  When disabled, begin/end
// do nothing.  collect() waits for the recorded events and accumulates milliseconds by name.
struct Profiler {
  bool enabled = false;
  struct Mark { const char* name; hipEvent_t a, b; };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> pool;
  std::vector<std::string> names;
  std::vector<double> total_ms;
  std::vector<long long> count;

  hipEvent_t take() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  void begin(const char* name, hipStream_t s) {
    if (!enabled) return;
    Mark m{name, take(), take()};
    (void)hipEventRecord(m.a, s);
    marks.push_back(m);
  }
  void end(hipStream_t s) {
    if (!enabled || marks.empty()) return;
    (void)hipEventRecord(marks.back().b, s);
  }
  void collect() {
    for (auto& m : marks) {
      (void)hipEventSynchronize(m.b);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, m.a, m.b);
      size_t k = 0;
      while (k < names.size() && names[k] != m.name) ++k;
      if (k == names.size()) { names.push_back(m.name); total_ms.push_back(0.0); count.push_back(0); }
      total_ms[k] += ms; count[k] += 1;
      pool.push_back(m.a); pool.push_back(m.b);
    }
    marks.clear();
  }
  void reset() { collect(); names.clear(); total_ms.clear(); count.clear(); }
  ~Profiler() {
    for (auto& m : marks) { (void)hipEventDestroy(m.a); (void)hipEventDestroy(m.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

struct ProfScope {
  Profiler* p; hipStream_t s;
  ProfScope(Profiler* p_, const char* name, hipStream_t s_) : p(p_), s(s_) { if (p) p->begin(name, s); }
  ~ProfScope() { if (p) p->end(s); }
};

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each with its own L2.
// xcd_contiguous maps hardware workgroup id wg of n to a logical work index such that every
// XCD receives one contiguous range of logical indices (a bijection on [0, n)); neighbouring
// work -- pixels of one view, faces of one view -- then shares one L2 instead of eight.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_contiguous(int64_t wg, int64_t n) {
  const int64_t x = wg % kXcds, j = wg / kXcds;
  const int64_t q = n / kXcds, r = n % kXcds;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + j;
}

// One launch over several meshes: mesh i owns the blocks [start[i], start[i+1]) of the grid
// (in the logical order, after any XCD remapping) and takes args[i].  find() is a binary search
// over the uniform prefix table.
// The table is read through the constant address space, so that its workgroup-uniform values
// are scalar loads into scalar registers: through a generic pointer the compiler cannot prove
// the table unchanged by the kernels' own stores, and keeps the loaded arguments in vector
// registers, where they crowd out the kernel's working set.
typedef const __attribute__((address_space(4))) int* ConstWords;
struct MeshSlice { int mesh; int64_t local; };
template <class Args>
struct MeshTable {
  const Args* args;                  // [n]
  const unsigned* start;             // [n + 1]
  int n;
  __device__ __forceinline__ MeshSlice find(unsigned wg) const {
    int lo = 0, hi = n - 1;
    const ConstWords st = (ConstWords)start;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if ((unsigned)st[mid] <= wg) lo = mid;
      else hi = mid - 1;
    }
    MeshSlice r;
    r.mesh = lo;
    r.local = (int64_t)(wg - (unsigned)st[lo]);
    return r;
  }
  // the mesh's arguments, by scalar loads
  __device__ __forceinline__ Args load(int mesh) const {
    static_assert(sizeof(Args) % 4 == 0, "argument records are whole 32-bit words");
    Args a;
    const ConstWords src = (ConstWords)(args + mesh);
    int* dst = reinterpret_cast<int*>(&a);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(Args) / 4); ++k) dst[k] = src[k];
    return a;
  }
};

// Albedo and shading modes, numbered as the reference enums
// (cpp/src/Renderer/CUDABasedRasterizationInput.h:25-35).
enum Albedo { kVertexColor = 0, kTextured = 1, kNormal = 2, kLighting = 3, kForegroundMask = 4 };
enum Shading { kShaded = 0, kShadeless = 1 };

// One entry of the texel -> (face, barycentrics) table used by the normal map.
struct UvTexel { int32_t face; float a, b, c; };

// Screen tiles of the tile-binned depth pass: (1 << sx) x (1 << sy) pixels, one workgroup
// each, with a bin of bin_cap face ids.  `dirty` flags ([B][C][tiles_y][tiles_x]) the tiles
// whose pixels the large-face pass wrote into the global visibility keys; null when the
// face-binned depth pass runs.
constexpr int kBinMaxTiles = 16;     // faces whose boxes span more tiles go to the large-face pass
struct TileMarks {
  unsigned* dirty = nullptr;
  int tiles_x = 0, tiles_y = 0;
  int sx = 4, sy = 4;                // log2 of the tile width and height
  int bin_cap = 512;                 // faces per tile bin (a face finding its bin full goes to the large-face pass)
};

// The raster state of one mesh as the large-face kernel needs it.
struct RasterMesh {
  const TriSetup* tris;              // [B][F]
  const int32_t* faces;              // [F][3]
  const float4* proj;                // [B][C][N]
  const Camera* cams;                // [B][C]
  unsigned long long* keys;          // [B][C][H][W]
  int C, N, F, W, H;
  TileMarks marks;
};

struct VertexArgs {
  const float* vpos;                 // [B][N][3]
  const float* fnormal;              // [B][F][3]
  const int32_t* csr_start;
  const int32_t* csr_count;
  const int32_t* csr_list;
  const Camera* cams;                // [B][C]
  float* vnormal;                    // [B][C][N][3]
  float4* proj;                      // [B][C][N]
  int B, C, N, F;
};

struct RasterArgs {
  const TriSetup* tris;              // [B][F]
  const int32_t* faces;              // [F][3]
  const float4* proj;                // [B][C][N]
  const Camera* cams;                // [B][C]
  unsigned long long* keys;          // [B][C][H][W]
  unsigned int* large_count;         // queue length
  int4* large;                       // large-face queue
  int B, C, N, F, W, H, large_box, mesh_id;
  int queue_large;                   // 1: append large faces to the queue; 0: skip them (queued
                                     // already by k_classify_large)
};

struct ResolveArgs {
  int B, C, N, W, H, texW, texH;
  unsigned long long* keys;          // [B][C][H][W]; reset to empty after reading
  const float* vpos;                 // [B][N][3]
  const float* vcol;                 // [B][N][3]
  const float* tex;                  // [B][texH][texW][3]
  const float* sh;                   // [B][C][27]
  const float* uv;                   // [F][6]
  const int32_t* faces;              // [F][3]
  const Camera* cams;                // [B][C]
  const float* vnormal;              // [B][C][N][3]
  int32_t* face;                     // [B][C][H][W]
  float* bary;                       // [B][C][H][W][2]
  float* render;                     // [B][C][H][W][3]
};

struct ForwardPlan {
  int B, C, N, F;
  int albedo;
  bool shade;
  bool normal_map;
  const float* vpos;
  const float* extr;
  const float* intr;
  const int32_t* faces;
  const int32_t* csr_start;
  const int32_t* csr_count;
  const int32_t* csr_list;
  const UvTexel* uv_table;
  Camera* cams;
  float* fnormal;                    // [B][F][3]
  TriSetup* tris;                    // [B][F] scaled ray/triangle constants
  float4* proj;                      // [B][C][N]
  float* vnormal_out;                // [B][C][N][3]
  float* normal_map_out;             // [B][texH][texW][3]
  unsigned int* counters;            // [2]: large-face queue length, (unused)
  int4* large;                       // large-face queue of (b*C + c, face, mesh, 0)
  int mesh_id;                       // written into the queue entries
  bool shared_queue;                 // the queue is shared by several meshes: the caller clears
                                     // the counters and drains it (k_raster_large_multi)
  int large_box;                     // bounding boxes above this many pixels go to the queue
  // tile-binned depth pass (when tiled): per-tile face bins, their counts and the large-face
  // marks; counts and marks are zero between calls (the tile kernel resets them)
  bool tiled;
  unsigned* tile_count;              // [B][C][tiles_y][tiles_x]
  unsigned* bins;                    // [B][C][tiles_y][tiles_x][bin_cap] face ids
  TileMarks marks;
  ResolveArgs r;
  Profiler* prof;
  // the large-face pass beside the depth pass: a second stream of the handle and the events
  // that fork it from the caller's stream after k_vertices and join it back (null: in line)
  hipStream_t side;
  hipEvent_t fork, join;
};

struct BackwardArgs {
  int B, C, N, F, W, H, texW, texH, image_fs;
  int tile;                          // pixel tile edge of k_pixel_grad: 8, 16 or 32
  int tiles_x, tiles_y;
  const float* g_render;             // [B][C][H][W][3]
  const float* g_target;             // [B][C][H][W][3] or null (no model-to-data term)
  const float* target;               // [B][C][H][W][3]
  const float* vpos;
  const float* vcol;
  const float* tex;
  const float* sh;
  const float* vnormal;
  const float* bary;                 // [B][C][H][W][2] or null (recomputed from the face id)
  const int32_t* face;
  const int32_t* faces;
  const float* uv;
  const Camera* cams;
  const int32_t* csr_start;          // vertex -> incident faces (the finishing gather)
  const int32_t* csr_count;
  const int32_t* csr_list;
  const FaceGrad* face_grad;         // [B][F] ray/triangle and Jacobian constants per face
  float* g_pos;                      // [B][N][3]
  float* g_col;                      // [B][N][3]
  float* g_tex;                      // [B][texH][texW][3]
  float* g_pv;                       // [B][N][6] scratch (shaded modes): per vertex the pixels'
                                     // position terms (xyz), then dL/d(vertex normal) (xyz)
  float* sh_partial;                 // [B*C][tiles][27] scratch
  float* g_sh;                       // [B][C][27]
};

// The reductions after the pixel kernel (k_grad_finish): fj_blocks face-Jacobian blocks, then
// one SH block per (view, coefficient).
struct FinishArgs {
  const float* vpos;
  const int32_t* faces;
  const int32_t* csr_start;
  const int32_t* csr_count;
  const int32_t* csr_list;
  const float* g_pv;                 // [B][N][6]: position terms, normal gradient
  float* g_pos;
  const float* sh_partial;
  float* g_sh;
  int B, N, F, tiles, vg_blocks;
};
FinishArgs finish_args(const BackwardArgs& a);

struct BackwardPlan {
  int albedo;
  bool shaded;
  BackwardArgs a;
  const float* extr;
  const float* intr;
  Camera* cams;
  Profiler* prof;
};

hipError_t launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s);
hipError_t launch_fill_f32(float* p, int64_t n, float v, hipStream_t s);

// The short setup stages of a call in one launch: per-view cameras, the per-face setup
// (when faces is set), zero fills of up to kFillSegs float buffers, and the two large-face
// counters (when set).
constexpr int kFillSegs = 5;
struct SetupArgs {
  const float* extr = nullptr;
  const float* intr = nullptr;
  Camera* cams = nullptr;
  int n_cams = 0;
  const float* vpos = nullptr;
  const int32_t* faces = nullptr;
  float* fnormal = nullptr;
  TriSetup* tris = nullptr;
  FaceGrad* face_grad = nullptr;     // the backward's per-face constants
  int B = 0, N = 0, F = 0;
  float* fill[kFillSegs] = {};
  int64_t fill_n[kFillSegs] = {};
  unsigned int* counters = nullptr;
};
// The block layout of one setup launch (k_setup): camera, face and fill blocks.
struct SetupGrid { SetupArgs a; int cam_blocks, face_blocks, fill_blocks; };
SetupGrid setup_grid(const SetupArgs& a);
hipError_t launch_setup(const SetupArgs& a, hipStream_t s);
hipError_t launch_forward(const ForwardPlan& p, hipStream_t s);
// The two halves of launch_forward: up to and including the depth pass (with the large-face
// queue unless the plan's queue is shared), then resolve.
hipError_t launch_forward_raster(const ForwardPlan& p, hipStream_t s);
hipError_t launch_forward_resolve(const ForwardPlan& p, hipStream_t s);
RasterMesh raster_mesh(const ForwardPlan& p);
SetupArgs forward_setup_args(const ForwardPlan& p);
VertexArgs vertex_args(const ForwardPlan& p);
RasterArgs raster_args(const ForwardPlan& p);
hipError_t launch_clear_counters(unsigned int* counters, hipStream_t s);
// Drains a large-face queue shared by several meshes (entries carry the mesh index).
hipError_t launch_raster_large_multi(const RasterMesh* meshes, unsigned int* counters, const int4* large,
                                     hipStream_t s);
hipError_t launch_backward(const BackwardPlan& p, hipStream_t s);
// The backward's pixel kernel alone (no setup, no finishing reductions).
hipError_t launch_pixel_grad(const BackwardPlan& p, hipStream_t s);

// Stage launches over several meshes (the multi-mesh runtime): one launch per stage, each
// workgroup finding its mesh in the table.  `blocks` is the table's total block count.
SetupArgs backward_setup_args(const BackwardPlan& p);
unsigned resolve_blocks(const ResolveArgs& a);
unsigned finish_blocks(const FinishArgs& f, int C);
hipError_t launch_setup_multi(const MeshTable<SetupGrid>& t, unsigned blocks, hipStream_t s);
hipError_t launch_vertices_multi(const MeshTable<VertexArgs>& t, unsigned blocks, hipStream_t s);
hipError_t launch_raster_multi(const MeshTable<RasterArgs>& t, unsigned blocks, hipStream_t s);
hipError_t launch_resolve_multi(const MeshTable<ResolveArgs>& t, unsigned blocks, int albedo, bool shade,
                                hipStream_t s);
hipError_t launch_grad_finish_multi(const MeshTable<FinishArgs>& t, unsigned blocks, hipStream_t s);
// The tile-binned depth pass (gvv_tile.hip): binning (after k_vertices), and the fused
// per-tile raster + resolve (after the large-face pass).
hipError_t launch_bin(const ForwardPlan& p, hipStream_t s);
hipError_t launch_tile(const ForwardPlan& p, hipStream_t s);

}  // namespace gvv
