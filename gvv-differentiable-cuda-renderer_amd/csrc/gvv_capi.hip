// gvv_capi.hip -- C ABI of the MI355X rasterizer (declared in include/gvv_renderer.h).
//
// One handle corresponds to one reference op instance: the topology and attributes of
// REGISTER_OP("CudaRendererGpu") / ("CudaRendererGradGpu")
// (reference cpp/src/TensorflowOperators/CudaRenderer/CudaRenderer.cpp:5-33,
//  CudaRendererGrad.cpp:6-39).  The handle owns device copies of the faces, UVs and the
// vertex->face CSR, the texel table of the normal map, and scratch buffers that grow with
// the batch size.  Tensor arguments are borrowed device pointers; every launch goes to the
// caller's stream.  Errors return non-zero and leave a message for gvv_last_error().
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gvv_handle.hpp"
#include "../../include/gvv_renderer.h"

namespace {
thread_local std::string g_err;
}  // namespace

namespace gvv {
int set_last_error(const std::string& m) { g_err = m; return 1; }
}  // namespace gvv

namespace {

using namespace gvv;

int fail(const std::string& m) { return set_last_error(m); }

#define GVV_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// Vertex->face lists in increasing face order, each face once per vertex
// (reference CUDABasedRasterization.cpp:125-154), built by counting sort in O(N + F).
void build_csr(int N, const std::vector<int32_t>& faces, std::vector<int32_t>& start,
               std::vector<int32_t>& count, std::vector<int32_t>& list) {
  const int F = (int)(faces.size() / 3);
  start.assign(N, 0); count.assign(N, 0);
  for (int f = 0; f < F; ++f) {
    int a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
    count[a]++;
    if (b != a) count[b]++;
    if (c != a && c != b) count[c]++;
  }
  int run = 0;
  for (int v = 0; v < N; ++v) { start[v] = run; run += count[v]; }
  list.assign(run, 0);
  std::vector<int32_t> fill(N, 0);
  for (int f = 0; f < F; ++f) {
    int a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
    list[start[a] + fill[a]++] = f;
    if (b != a) list[start[b] + fill[b]++] = f;
    if (c != a && c != b) list[start[c] + fill[c]++] = f;
  }
}

// Texel -> (face, barycentrics) table (reference CUDABasedRasterization.cpp:237-298): each
// face's UV triangle is intersected in texel space by the ray from (x+.5, y+.5, 1) along -z;
// later faces overwrite earlier ones; uncovered texels keep (face 0, 0, 0, 0).
void build_uv_table(const Handle& h, int texW, int texH, std::vector<UvTexel>& table) {
  table.assign((size_t)texW * texH, UvTexel{0, 0.f, 0.f, 0.f});
  const float* uv = h.uv.data();
  const V3 dir = v3(0.f, 0.f, -1.f);
  for (int f = 0; f < h.F; ++f) {
    V3 t0 = v3((float)texW * uv[f * 6 + 0], (float)texH * (1.f - uv[f * 6 + 1]), 0.f);
    V3 t1 = v3((float)texW * uv[f * 6 + 2], (float)texH * (1.f - uv[f * 6 + 3]), 0.f);
    V3 t2 = v3((float)texW * uv[f * 6 + 4], (float)texH * (1.f - uv[f * 6 + 5]), 0.f);
    int xMin = sat_i32(fmaxf(fminf(t0.x, fminf(t1.x, t2.x)) - 2.f, 0.f));
    int xMax = sat_i32(fminf(fmaxf(t0.x, fmaxf(t1.x, t2.x)) + 2.f, (float)texW));
    int yMin = sat_i32(fmaxf(fminf(t0.y, fminf(t1.y, t2.y)) - 2.f, 0.f));
    int yMax = sat_i32(fminf(fmaxf(t0.y, fmaxf(t1.y, t2.y)) + 2.f, (float)texH));
    TriSetup ts = tri_setup(t0, t1, t2);
    for (int x = xMin; x < xMax; ++x)
      for (int y = yMin; y < yMax; ++y) {
        V3 o = v3(x + 0.5f, y + 0.5f, 1.f);
        V3 abc = ray_barycentric(ts, o / 1000.f, dir);
        if (abc.x != -1.f && abc.y != -1.f && abc.z != -1.f)
          table[(size_t)y * texW + x] = UvTexel{f, abc.x, abc.y, abc.z};
      }
  }
}

inline int tiles(int n, int tile) { return (n + tile - 1) / tile; }

}  // namespace

namespace gvv {

int plan_forward(Handle* h, int B, int texH, int texW, const float* vpos, const float* vcol, const float* tex,
                 const float* sh, const float* extr, const float* intr, float* bary, int32_t* face, float* render,
                 float* vnormal, float* normal_map, hipStream_t s, ForwardPlan& p) {
  if (B <= 0) return fail("batch size must be > 0");
  if (texH <= 0 || texW <= 0) return fail("texture height/width must be > 0");
  if (!vpos || !vcol || !tex || !sh || !extr || !intr || !bary || !face || !render || !vnormal)
    return fail("null tensor pointer");
  if (h->normal_map && !normal_map) return fail("compute_normal_map is set but normal_map is null");
  const int64_t pix = (int64_t)B * h->C * h->W * h->H;
  size_t old_keys = h->keys.n;
  GVV_CHECK(h->cams.reserve((size_t)B * h->C));
  GVV_CHECK(h->fnormal.reserve((size_t)B * h->F * 3));
  GVV_CHECK(h->tris.reserve((size_t)B * h->F));
  GVV_CHECK(h->proj.reserve((size_t)B * h->C * h->N));
  GVV_CHECK(h->keys.reserve((size_t)pix));
  GVV_CHECK(h->counters.reserve(2));
  GVV_CHECK(h->large.reserve((size_t)B * h->C * h->F));
  if (h->keys.n != old_keys || old_keys == 0)
    GVV_CHECK(launch_fill_u64(h->keys.p, (int64_t)h->keys.n, kEmptyKey, s));
  const bool tiled = h->depth_pass != 0;
  // screen tiles: 16 x 16 (mode 1), 32 x 16 (mode 2) or 32 x 32 (mode 3) pixels; bins hold
  // twice a tile's pixel count
  const int sx = h->depth_pass >= 2 ? 5 : 4, sy = h->depth_pass == 3 ? 5 : 4;
  const int tiles_x = tiles(h->W, 1 << sx), tiles_y = tiles(h->H, 1 << sy);
  const int bin_cap = 2 << (sx + sy);
  if (tiled) {
    const size_t nt = (size_t)B * h->C * tiles_x * tiles_y;
    const size_t old_t = h->tile_count.n;
    GVV_CHECK(h->tile_count.reserve(nt));
    GVV_CHECK(h->tile_dirty.reserve(nt));
    GVV_CHECK(h->bins.reserve(nt * bin_cap));
    if (h->tile_count.n != old_t || old_t == 0) {
      // counts and marks start at zero; the tile kernel resets the ones it read
      GVV_CHECK(launch_fill_f32(reinterpret_cast<float*>(h->tile_count.p), (int64_t)h->tile_count.n, 0.f, s));
      GVV_CHECK(launch_fill_f32(reinterpret_cast<float*>(h->tile_dirty.p), (int64_t)h->tile_dirty.n, 0.f, s));
    }
  }
  if (h->normal_map && (h->uvW != texW || h->uvH != texH)) {
    std::vector<UvTexel> table;
    build_uv_table(*h, texW, texH, table);
    GVV_CHECK(h->d_uv_table.reserve(table.size()));
    GVV_CHECK(hipMemcpyAsync(h->d_uv_table.p, table.data(), sizeof(UvTexel) * table.size(), hipMemcpyHostToDevice, s));
    GVV_CHECK(hipStreamSynchronize(s));
    h->uvW = texW; h->uvH = texH;
  }
  p.B = B; p.C = h->C; p.N = h->N; p.F = h->F;
  p.albedo = h->albedo;
  p.shade = (h->shading == kShaded && h->albedo != kNormal) || h->albedo == kLighting;
  p.normal_map = h->normal_map != 0;
  p.vpos = vpos; p.extr = extr; p.intr = intr;
  p.faces = h->d_faces.p; p.csr_start = h->d_csr_start.p; p.csr_count = h->d_csr_count.p; p.csr_list = h->d_csr_list.p;
  p.uv_table = h->d_uv_table.p;
  p.cams = h->cams.p; p.fnormal = h->fnormal.p; p.tris = h->tris.p; p.proj = h->proj.p;
  p.counters = h->counters.p; p.large = h->large.p; p.large_box = h->large_box;
  p.mesh_id = 0; p.shared_queue = false;
  p.tiled = tiled;
  p.tile_count = tiled ? h->tile_count.p : nullptr;
  p.bins = tiled ? h->bins.p : nullptr;
  p.marks.dirty = tiled ? h->tile_dirty.p : nullptr;
  p.marks.tiles_x = tiles_x;
  p.marks.tiles_y = tiles_y;
  p.marks.sx = sx;
  p.marks.sy = sy;
  p.marks.bin_cap = bin_cap;
  p.vnormal_out = vnormal;
  p.normal_map_out = normal_map;
  p.side = nullptr; p.fork = p.join = nullptr;      // the single-mesh entry point sets them
  ResolveArgs& r = p.r;
  r.B = B; r.C = h->C; r.N = h->N; r.W = h->W; r.H = h->H; r.texW = texW; r.texH = texH;
  r.keys = h->keys.p; r.vpos = vpos; r.vcol = vcol; r.tex = tex; r.sh = sh; r.uv = h->d_uv.p;
  r.faces = h->d_faces.p; r.cams = h->cams.p; r.vnormal = vnormal;
  r.face = face; r.bary = bary; r.render = render;
  p.prof = nullptr;
  return 0;
}

int plan_backward_into(Handle* h, int B, int texH, int texW, const float* g_render, const float* g_target,
                       const float* target, const float* vpos, const float* vcol, const float* tex, const float* sh,
                       const float* vnormal, const float* bary, const int32_t* face, const float* extr,
                       const float* intr, float* g_pos, float* g_col, float* g_tex, float* g_sh, Camera* cams,
                       float* g_pv, float* sh_partial, FaceGrad* face_grad, BackwardPlan& p) {
  if (B <= 0) return fail("batch size must be > 0");
  if (texH <= 0 || texW <= 0) return fail("texture height/width must be > 0");
  if (!g_pos || !g_tex || !g_sh) return fail("null gradient output pointer");   // g_col may be NULL
  if (h->albedo == kNormal || h->albedo == kLighting) return fail("normal and lighting modes have no gradients");
  if (!g_render || !vpos || !vcol || !tex || !sh || !vnormal || !face || !extr || !intr)
    return fail("null tensor pointer");
  if (g_target && !target) return fail("target_buffer_grad given without target_image");
  const int tx = tiles(h->W, h->tile), ty = tiles(h->H, h->tile);
  p.albedo = h->albedo;
  p.shaded = h->shading == kShaded;
  p.extr = extr; p.intr = intr; p.cams = cams;
  p.prof = nullptr;
  BackwardArgs& a = p.a;
  a.B = B; a.C = h->C; a.N = h->N; a.F = h->F; a.W = h->W; a.H = h->H; a.texW = texW; a.texH = texH;
  a.image_fs = h->image_fs; a.tile = h->tile; a.tiles_x = tx; a.tiles_y = ty;
  a.g_render = g_render; a.g_target = g_target; a.target = target;
  a.vpos = vpos; a.vcol = vcol; a.tex = tex; a.sh = sh; a.vnormal = vnormal; a.bary = bary; a.face = face;
  a.faces = h->d_faces.p; a.uv = h->d_uv.p; a.cams = cams; a.face_grad = face_grad;
  a.csr_start = h->d_csr_start.p; a.csr_count = h->d_csr_count.p; a.csr_list = h->d_csr_list.p;
  a.g_pos = g_pos; a.g_col = g_col; a.g_tex = g_tex; a.g_pv = g_pv; a.sh_partial = sh_partial; a.g_sh = g_sh;
  return 0;
}

BackwardScratch backward_scratch(const Handle* h, int B) {
  BackwardScratch s;
  s.cams = (size_t)B * h->C;
  s.g_pv = (size_t)B * h->N * 6;
  s.sh_partial = (size_t)B * h->C * tiles(h->W, h->tile) * tiles(h->H, h->tile) * 27;
  s.face_grad = (size_t)B * h->F;
  return s;
}

int plan_backward(Handle* h, int B, int texH, int texW, const float* g_render, const float* g_target,
                  const float* target, const float* vpos, const float* vcol, const float* tex, const float* sh,
                  const float* vnormal, const float* bary, const int32_t* face, const float* extr,
                  const float* intr, float* g_pos, float* g_col, float* g_tex, float* g_sh, BackwardPlan& p) {
  if (B > 0) {
    const BackwardScratch n = backward_scratch(h, B);
    GVV_CHECK(h->cams.reserve(n.cams));
    GVV_CHECK(h->g_pv.reserve(n.g_pv));
    GVV_CHECK(h->sh_partial.reserve(n.sh_partial));
    GVV_CHECK(h->face_grad.reserve(n.face_grad));
  }
  return plan_backward_into(h, B, texH, texW, g_render, g_target, target, vpos, vcol, tex, sh, vnormal, bary, face,
                            extr, intr, g_pos, g_col, g_tex, g_sh, h->cams.p, h->g_pv.p, h->sh_partial.p,
                            h->face_grad.p, p);
}

int check_device(const Handle* h) {
  int dev = -1;
  GVV_CHECK(hipGetDevice(&dev));
  if (dev != h->device) return fail("handle was created on device " + std::to_string(h->device) +
                                    " but the current device is " + std::to_string(dev));
  return 0;
}

}  // namespace gvv

extern "C" {

const char* gvv_last_error(void) { return g_err.c_str(); }

int gvv_version(void) { return GVV_RENDERER_ABI_VERSION; }

int gvv_create(const int32_t* faces, int64_t n_face_idx, const float* texcoords, int64_t n_texcoords,
               int32_t n_vertices, int32_t n_cameras, int32_t width, int32_t height, int32_t albedo_mode,
               int32_t shading_mode, int32_t image_filter_size, int32_t texture_filter_size,
               int32_t compute_normal_map, void** handle) {
  if (!handle) return fail("handle pointer is null");
  *handle = nullptr;
  if (!faces || n_face_idx <= 0 || n_face_idx % 3 != 0) return fail("faces must hold 3*F > 0 vertex indices");
  const int64_t F = n_face_idx / 3;
  if (F > (1ll << 30)) return fail("too many faces");
  if (!texcoords || n_texcoords != 6 * F) return fail("texture_coordinates must hold 6*F values (u,v per face corner)");
  if (n_vertices <= 0) return fail("number_of_vertices must be > 0");
  if (n_cameras <= 0) return fail("number_of_cameras must be > 0");
  if (width <= 0 || height <= 0) return fail("render_resolution_u/v must be > 0");
  if (albedo_mode < 0 || albedo_mode > 4) return fail("invalid albedo mode");
  if (shading_mode < 0 || shading_mode > 1) return fail("invalid shading mode");
  if (image_filter_size <= 0) return fail("image_filter_size must be > 0");
  if (texture_filter_size <= 0) return fail("texture_filter_size must be > 0");
  for (int64_t i = 0; i < n_face_idx; ++i)
    if (faces[i] < 0 || faces[i] >= n_vertices)
      return fail("face index " + std::to_string(faces[i]) + " out of range [0, " + std::to_string(n_vertices) + ")");
  Handle* h = new Handle();
  h->N = n_vertices; h->F = (int)F; h->C = n_cameras; h->W = width; h->H = height;
  h->albedo = albedo_mode;
  // foregroundMask forces shadeless (CudaRenderer.cpp:72-76, CudaRendererGrad.cpp:78-82)
  h->shading = albedo_mode == kForegroundMask ? kShadeless : shading_mode;
  h->image_fs = image_filter_size; h->texture_fs = texture_filter_size;
  h->normal_map = compute_normal_map ? 1 : 0;
  h->faces.assign(faces, faces + n_face_idx);
  h->uv.assign(texcoords, texcoords + n_texcoords);
  std::vector<int32_t> start, count, list;
  build_csr(h->N, h->faces, start, count, list);
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = h->d_faces.reserve(h->faces.size());
  if (e == hipSuccess) e = h->d_uv.reserve(h->uv.size());
  if (e == hipSuccess) e = h->d_csr_start.reserve(start.size());
  if (e == hipSuccess) e = h->d_csr_count.reserve(count.size());
  if (e == hipSuccess) e = h->d_csr_list.reserve(list.size());
  if (e == hipSuccess) e = hipMemcpy(h->d_faces.p, h->faces.data(), sizeof(int32_t) * h->faces.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->d_uv.p, h->uv.data(), sizeof(float) * h->uv.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->d_csr_start.p, start.data(), sizeof(int32_t) * start.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->d_csr_count.p, count.data(), sizeof(int32_t) * count.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !list.empty())
    e = hipMemcpy(h->d_csr_list.p, list.data(), sizeof(int32_t) * list.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete h;
    return fail(std::string("gvv_create: ") + hipGetErrorString(e));
  }
  *handle = h;
  return 0;
}

void gvv_destroy(void* handle) { delete (Handle*)handle; }

int gvv_forward(void* handle, int32_t B, int32_t texH, int32_t texW, const float* vpos, const float* vcol,
                const float* tex, const float* sh, const float* extr, const float* intr, float* bary,
                int32_t* face, float* render, float* vnormal, float* normal_map, void* hip_stream) {
  Handle* h = (Handle*)handle;
  if (!h) return fail("null handle");
  std::lock_guard<std::mutex> lock(h->mu);
  if (check_device(h)) return 1;
  hipStream_t s = (hipStream_t)hip_stream;
  ForwardPlan p;
  if (plan_forward(h, B, texH, texW, vpos, vcol, tex, sh, extr, intr, bary, face, render, vnormal, normal_map, s, p))
    return 1;
  // the handle's second stream for the large-face pass, made at its first forward (the
  // multi-mesh runtime's meshes share one queue and never use one)
  if (!h->side) {
    GVV_CHECK(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    GVV_CHECK(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
    GVV_CHECK(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
  }
  p.side = h->side; p.fork = h->ev_fork; p.join = h->ev_join;
  p.prof = h->prof.enabled ? &h->prof : nullptr;
  GVV_CHECK(launch_forward(p, s));
  return 0;
}

int gvv_backward(void* handle, int32_t B, int32_t texH, int32_t texW, const float* g_render, const float* g_target,
                 const float* target, const float* vpos, const float* vcol, const float* tex, const float* sh,
                 const float* vnormal, const float* bary, const int32_t* face, const float* extr,
                 const float* intr, float* g_pos, float* g_col, float* g_tex, float* g_sh, void* hip_stream) {
  Handle* h = (Handle*)handle;
  if (!h) return fail("null handle");
  std::lock_guard<std::mutex> lock(h->mu);
  if (check_device(h)) return 1;
  if (B <= 0) return fail("batch size must be > 0");
  if (texH <= 0 || texW <= 0) return fail("texture height/width must be > 0");
  if (!g_pos || !g_tex || !g_sh) return fail("null gradient output pointer");   // g_col may be NULL
  hipStream_t s = (hipStream_t)hip_stream;
  const int64_t nv = (int64_t)B * h->N * 3;
  if (h->albedo == kNormal || h->albedo == kLighting) {
    // the reference routes these modes to zero gradients (python/CudaRenderer.py:207-213)
    GVV_CHECK(launch_fill_f32(g_pos, nv, 0.f, s));
    if (g_col) GVV_CHECK(launch_fill_f32(g_col, nv, 0.f, s));
    GVV_CHECK(launch_fill_f32(g_tex, (int64_t)B * texH * texW * 3, 0.f, s));
    GVV_CHECK(launch_fill_f32(g_sh, (int64_t)B * h->C * 27, 0.f, s));
    return 0;
  }
  BackwardPlan p;
  if (plan_backward(h, B, texH, texW, g_render, g_target, target, vpos, vcol, tex, sh, vnormal, bary, face, extr, intr,
                    g_pos, g_col, g_tex, g_sh, p))
    return 1;
  p.prof = h->prof.enabled ? &h->prof : nullptr;
  GVV_CHECK(launch_backward(p, s));
  return 0;
}

int gvv_set_tuning(void* handle, int32_t tile, int32_t large_box) {
  Handle* h = (Handle*)handle;
  if (!h) return fail("null handle");
  if (tile != 8 && tile != 16 && tile != 32) return fail("tile must be 8, 16 or 32");
  if (large_box < 16 || large_box > (1 << 20)) return fail("large_box must be in [16, 2^20]");
  std::lock_guard<std::mutex> lock(h->mu);
  h->tile = tile;
  h->large_box = large_box;
  return 0;
}

int gvv_set_depth_pass(void* handle, int32_t mode) {
  Handle* h = (Handle*)handle;
  if (!h) return fail("null handle");
  if (mode < 0 || mode > 3)
    return fail("depth pass must be 0 (faces binned by size) or 1, 2, 3 (screen tiles of 16x16, 32x16, 32x32 px)");
  std::lock_guard<std::mutex> lock(h->mu);
  h->depth_pass = mode;
  return 0;
}

int gvv_get_depth_pass(void* handle, int32_t* mode) {
  Handle* h = (Handle*)handle;
  if (!h || !mode) return fail("null pointer");
  std::lock_guard<std::mutex> lock(h->mu);
  *mode = h->depth_pass;
  return 0;
}

int gvv_get_tuning(void* handle, int32_t* tile, int32_t* large_box) {
  Handle* h = (Handle*)handle;
  if (!h || !tile || !large_box) return fail("null pointer");
  std::lock_guard<std::mutex> lock(h->mu);
  *tile = h->tile;
  *large_box = h->large_box;
  return 0;
}

int gvv_profile_enable(void* handle, int32_t enable) {
  Handle* h = (Handle*)handle;
  if (!h) return fail("null handle");
  std::lock_guard<std::mutex> lock(h->mu);
  h->prof.reset();
  h->prof.enabled = enable != 0;
  return 0;
}

int64_t gvv_profile_report(void* handle, char* buf, int64_t cap) {
  Handle* h = (Handle*)handle;
  if (!h) return fail("null handle"), -1;
  std::lock_guard<std::mutex> lock(h->mu);
  h->prof.collect();
  std::string out;
  for (size_t k = 0; k < h->prof.names.size(); ++k) {
    char line[256];
    snprintf(line, sizeof(line), "%s %.6f %lld\n", h->prof.names[k].c_str(), h->prof.total_ms[k], h->prof.count[k]);
    out += line;
  }
  if (buf && cap > 0) {
    size_t n = std::min<size_t>(out.size(), (size_t)cap - 1);
    memcpy(buf, out.data(), n);
    buf[n] = 0;
  }
  return (int64_t)out.size();
}

}  // extern "C"
