/* gvv_multi.h -- C ABI of the multi-mesh runtime (libgvv_amd.so).
 *
 * No counterpart in the reference: its op takes one topology per op instance (the faces are
 * op attributes, cpp/src/TensorflowOperators/CudaRenderer/CudaRenderer.cpp:23) and loops over
 * the batch on the host (CudaRenderer.cpp:309-328).  This runtime renders many meshes of
 * uneven triangle counts, view counts and resolutions -- each with its own topology, inputs
 * and outputs in the single-mesh layouts of gvv_renderer.h -- with one hipGraph launch, and
 * tunes the launch settings per size class.
 *
 * Conventions as in gvv_renderer.h: device pointers, 0 on success, gvv_last_error() on failure.
 */
#ifndef GVV_MULTI_H
#define GVV_MULTI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Topology and render target of one mesh (host pointers, copied by gvv_multi_create). */
typedef struct gvv_mesh_desc {
  const int32_t* faces;      /* [3F] vertex indices */
  int64_t n_face_idx;        /* 3F */
  const float* texcoords;    /* [6F] (u, v) per face corner */
  int32_t n_vertices, n_cameras, width, height;
} gvv_mesh_desc;

/* Per-call tensors of one mesh, device pointers in the single-mesh layouts (B = batch). */
typedef struct gvv_mesh_io {
  int32_t batch, tex_h, tex_w;
  const float* vpos;         /* [B][N][3] */
  const float* vcol;         /* [B][N][3] */
  const float* tex;          /* [B][tex_h][tex_w][3] */
  const float* sh;           /* [B][C][27] */
  const float* extr;         /* [B][C*12] */
  const float* intr;         /* [B][C*9] */
  float* bary;               /* [B][C][H][W][2]   forward outputs */
  int32_t* face;             /* [B][C][H][W] */
  float* render;             /* [B][C][H][W][3] */
  float* vnormal;            /* [B][C][N][3] */
  const float* g_render;     /* [B][C][H][W][3]   backward input (NULL: forward only) */
  float* g_pos;              /* [B][N][3]          backward outputs */
  float* g_col;              /* [B][N][3] */
  float* g_tex;              /* [B][tex_h][tex_w][3] */
  float* g_sh;               /* [B][C][27] */
} gvv_mesh_io;

/* Creates a runtime over n meshes sharing one albedo/shading mode (vertexColor, textured or
 * foregroundMask: the modes with gradients). */
int gvv_multi_create(const gvv_mesh_desc* meshes, int32_t n, int32_t albedo_mode, int32_t shading_mode,
                     void** handle);
void gvv_multi_destroy(void* handle);

/* Forward of every mesh, then (if backward != 0) the backward of every mesh against
 * io[i].g_render.  With graphs on (the default) all launches of all meshes are captured once
 * into one hipGraph, meshes spread over parallel branches, and replayed on `stream` while the
 * io table stays the same; a changed table or tuning is captured again.  The large faces of
 * all meshes go to one queue that one persistent launch drains, so runs of one handle must not
 * overlap on different streams.  Results equal those of gvv_forward / gvv_backward on each mesh
 * with the same tuning. */
int gvv_multi_run(void* handle, const gvv_mesh_io* io, int32_t backward, void* stream);

/* 1: hipGraph replay (default); 0: plain launches in order on `stream`. */
int gvv_multi_use_graph(void* handle, int32_t on);

/* 1: one launch per stage for all meshes -- setup, vertices, depth pass, the shared
 * large-face pass, resolve, and in the backward setup and the finishing reductions -- each
 * workgroup finding its mesh in a device table of per-mesh arguments; the backward's pixel
 * kernels stay per mesh, fanned out over parallel graph branches.  0 (default): every mesh's own
 * launches for every stage, fanned out over the branches.  Results are identical. */
int gvv_multi_set_batched(void* handle, int32_t on);

/* Autotuner: for every size class present (triangle-count bucket x pixel-count bucket) times
 * forward+backward of the class's largest mesh under each candidate setting (backward tile
 * 8/16/32 x large-face threshold 64/256/1024 px, best of `reps`) and applies the fastest to
 * every mesh of the class.  io must be valid for a backward run.  Overwrites the outputs. */
int gvv_multi_autotune(void* handle, const gvv_mesh_io* io, int32_t reps, void* stream);

/* Sets the settings of mesh i (as gvv_set_tuning), e.g. from a saved autotuner result. */
int gvv_multi_set_tuning(void* handle, int32_t mesh, int32_t tile, int32_t large_box);

/* Settings and size class of mesh i. */
int gvv_multi_tuning(void* handle, int32_t mesh, int32_t* tile, int32_t* large_box, int32_t* size_class);

#ifdef __cplusplus
}
#endif

#endif /* GVV_MULTI_H */
