// gvv_handle.hpp -- the renderer handle behind the C ABI and the plan builders shared by the
// single-mesh entry points (gvv_capi.hip) and the multi-mesh runtime (gvv_multi.hip).
#pragma once
#include <mutex>
#include <string>
#include <vector>

#include "gvv_kernels.hpp"

namespace gvv {

int set_last_error(const std::string& m);

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t reserve(size_t count) {
    if (count <= n) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr; n = 0;
    hipError_t e = hipMalloc((void**)&p, sizeof(T) * (count ? count : 1));
    if (e == hipSuccess) n = count;
    return e;
  }
  void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

struct Handle {
  int N = 0, F = 0, C = 0, W = 0, H = 0;
  int albedo = 0, shading = 0, image_fs = 1, texture_fs = 1, normal_map = 0;
  int device = 0;
  int tile = 16;                    // k_pixel_grad tile edge (8, 16, 32)
  int large_box = 256;              // k_raster: bounding boxes above this go to the large-face queue
  int depth_pass = 0;               // 0: faces binned by size (k_raster + k_resolve); 1: screen tiles (k_bin + k_tile)
  std::vector<int32_t> faces;
  std::vector<float> uv;
  DevBuf<int32_t> d_faces, d_csr_start, d_csr_count, d_csr_list;
  DevBuf<float> d_uv;
  int uvW = -1, uvH = -1;
  DevBuf<UvTexel> d_uv_table;
  // scratch
  DevBuf<Camera> cams;
  DevBuf<float> fnormal;
  DevBuf<TriSetup> tris;
  DevBuf<float4> proj;
  DevBuf<unsigned long long> keys;
  DevBuf<unsigned int> counters;
  DevBuf<int4> large;
  DevBuf<unsigned> tile_count, tile_dirty, bins;
  DevBuf<float> g_pv;
  DevBuf<float> sh_partial;
  DevBuf<FaceGrad> face_grad;          // the backward's per-face constants [B][F]
  Profiler prof;
  std::mutex mu;
  // the forward's second stream (default priority) for the large-face pass beside the depth
  // pass, and its fork/join events; created at the first forward
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;

  ~Handle() {
    if (side) (void)hipStreamSynchronize(side);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
    d_faces.release(); d_csr_start.release(); d_csr_count.release(); d_csr_list.release();
    d_uv.release(); d_uv_table.release(); cams.release(); fnormal.release(); tris.release(); proj.release();
    keys.release(); counters.release(); tile_count.release(); tile_dirty.release(); bins.release(); large.release(); g_pv.release(); sh_partial.release();
    face_grad.release();
  }
};

// Validates the arguments of a forward call, sizes the handle's scratch for batch B (this
// may allocate, and clears a new visibility buffer on stream s) and fills the launch plan.
// Returns 0, or 1 with the message set.  The caller holds h->mu.
int plan_forward(Handle* h, int B, int texH, int texW, const float* vpos, const float* vcol, const float* tex,
                 const float* sh, const float* extr, const float* intr, float* bary, int32_t* face, float* render,
                 float* vnormal, float* normal_map, hipStream_t s, ForwardPlan& p);

// Same for the backward.  Modes without gradients (normal, lighting) are rejected here: the
// single-mesh entry point handles them before planning.
int plan_backward(Handle* h, int B, int texH, int texW, const float* g_render, const float* g_target,
                  const float* target, const float* vpos, const float* vcol, const float* tex, const float* sh,
                  const float* vnormal, const float* bary, const int32_t* face, const float* extr,
                  const float* intr, float* g_pos, float* g_col, float* g_tex, float* g_sh, BackwardPlan& p);

// Element counts of the backward's scratch buffers for batch B: cameras, dL/d(vertex normal),
// the SH tile partials and the per-face constants.
struct BackwardScratch { size_t cams, g_pv, sh_partial, face_grad; };
BackwardScratch backward_scratch(const Handle* h, int B);

// plan_backward with caller-provided scratch (sized by backward_scratch): validates and fills
// the plan, allocates nothing.
int plan_backward_into(Handle* h, int B, int texH, int texW, const float* g_render, const float* g_target,
                       const float* target, const float* vpos, const float* vcol, const float* tex, const float* sh,
                       const float* vnormal, const float* bary, const int32_t* face, const float* extr,
                       const float* intr, float* g_pos, float* g_col, float* g_tex, float* g_sh, Camera* cams,
                       float* g_pv, float* sh_partial, FaceGrad* face_grad, BackwardPlan& p);

// 0 when the current HIP device is the one the handle was created on; else 1 with the message set.
int check_device(const Handle* h);

}  // namespace gvv
