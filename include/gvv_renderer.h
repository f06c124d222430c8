/* gvv_renderer.h -- C ABI of the MI355X differentiable mesh rasterizer.
 *
 * Drop-in for the reference TensorFlow op pair (paths relative to the reference repository):
 *   forward   REGISTER_OP("CudaRendererGpu")      cpp/src/TensorflowOperators/CudaRenderer/CudaRenderer.cpp:5-33
 *             + its kernel CudaRenderer::Compute   CudaRenderer.cpp:298-335
 *   backward  REGISTER_OP("CudaRendererGradGpu")  cpp/src/TensorflowOperators/CudaRenderer/CudaRendererGrad.cpp:6-39
 *             + its kernel CudaRendererGrad::Compute CudaRendererGrad.cpp:252-293
 *   Python    python/CudaRenderer.py:30-215 (mirrored by CudaRenderer.py in the package)
 *
 * Conventions
 *   - All tensor arguments are borrowed, contiguous, row-major device pointers: float32 or
 *     int32, in exactly the reference op layouts listed per function.
 *   - Every function returns 0 on success and non-zero on error; gvv_last_error() then holds a
 *     thread-local message.  Nothing calls exit().
 *   - Work is enqueued on `hip_stream` (a hipStream_t; NULL = default stream).  A handle may be
 *     used from one stream at a time; calls on one handle are serialised by a mutex.
 *   - Albedo modes: 0 vertexColor, 1 textured, 2 normal, 3 lighting, 4 foregroundMask
 *     (cpp/src/Renderer/CUDABasedRasterizationInput.h:25-28).  Shading: 0 shaded, 1 shadeless
 *     (:33-35).  foregroundMask forces shadeless (CudaRenderer.cpp:72-76).
 *   - Geometry is expected in millimetres: the reference's ray/triangle thresholds are
 *     absolute (cpp/src/Utils/RendererUtil.h:29-47).
 */
#ifndef GVV_RENDERER_H
#define GVV_RENDERER_H

#include <stdint.h>

#define GVV_RENDERER_ABI_VERSION 1

#ifdef __cplusplus
extern "C" {
#endif

/* Thread-local message of the last failed call on this thread. */
const char* gvv_last_error(void);

/* ABI version (GVV_RENDERER_ABI_VERSION). */
int gvv_version(void);

/* Creates a renderer for one mesh topology and attribute set.  Replaces the op attributes and
 * the constructors CudaRenderer::CudaRenderer (CudaRenderer.cpp:37-155) and
 * CUDABasedRasterization::CUDABasedRasterization (cpp/src/Renderer/CUDABasedRasterization.cpp:11-106).
 *   faces        [3F] int32 vertex indices (attr `faces`)
 *   texcoords    [6F] float, per face per corner (u, v) (attr `texture_coordinates`)
 *   n_vertices, n_cameras, width (=render_resolution_u), height (=render_resolution_v),
 *   albedo_mode, shading_mode, image_filter_size, texture_filter_size, compute_normal_map
 * Host pointers; copied.  Rejects sizes <= 0, face indices out of range and texcoord counts
 * other than 6F (the reference printed a warning and continued). */
int gvv_create(const int32_t* faces, int64_t n_face_idx, const float* texcoords, int64_t n_texcoords,
               int32_t n_vertices, int32_t n_cameras, int32_t width, int32_t height, int32_t albedo_mode,
               int32_t shading_mode, int32_t image_filter_size, int32_t texture_filter_size,
               int32_t compute_normal_map, void** handle);

/* Releases a handle and its device buffers. */
void gvv_destroy(void* handle);

/* Forward render.  Replaces CudaRenderer::Compute (CudaRenderer.cpp:298-335) and
 * renderBuffersGPU (cpp/src/Renderer/CUDABasedRasterization.cu:449-473).
 *   inputs   vpos [B][N][3], vcol [B][N][3], tex [B][texH][texW][3], sh [B][C][27],
 *            extr [B][C*12] (row-major 3x4 world->camera), intr [B][C*9] (row-major K)
 *   outputs  bary [B][C][H][W][2], face [B][C][H][W] (int32, -1 = background),
 *            render [B][C][H][W][3], vnormal [B][C][N][3],
 *            normal_map [B][texH][texW][3] (written only when compute_normal_map; may be NULL
 *            otherwise).
 * The op's target_image_out output is the target input itself (an alias); it is not handled
 * here.  Visibility ties between faces at the same integer depth resolve to the lowest face
 * id (the reference resolves them by an unsynchronised last write).
 * Work is ordered on the caller's stream.  Outside stream capture, the large-face part of the
 * depth pass runs on a second stream the handle owns, forked from and joined back to the
 * caller's stream with events inside the call, so the call's results are complete in the
 * caller's stream order as usual; under capture everything stays on the caller's stream. */
int gvv_forward(void* handle, int32_t B, int32_t texH, int32_t texW, const float* vpos, const float* vcol,
                const float* tex, const float* sh, const float* extr, const float* intr, float* bary,
                int32_t* face, float* render, float* vnormal, float* normal_map, void* hip_stream);

/* Backward.  Replaces CudaRendererGrad::Compute (CudaRendererGrad.cpp:252-293) and
 * renderBuffersGradGPU (cpp/src/Renderer/CUDABasedRasterizationGrad.cu:624-635).
 *   g_render  [B][C][H][W][3]  gradient of the render buffer (op input render_buffer_grad)
 *   g_target  [B][C][H][W][3]  gradient of target_image_out (target_buffer_grad), or NULL for
 *                              zero, which skips the model-to-data term
 *   target    [B][C][H][W][3]  target image (required when g_target is given)
 *   vpos, vcol, tex, sh, extr, intr as in gvv_forward; vnormal and face are forward outputs
 *   bary      forward barycentrics, or NULL to recompute them from the face id (bit-identical
 *             to the forward's values)
 *   outputs   g_pos [B][N][3], g_col [B][N][3], g_tex [B][texH][texW][3], g_sh [B][C][27]
 *             g_col may be NULL (an extension; the reference op always computes it): the
 *             vertex-colour gradient is then neither formed nor written, which spares the
 *             vertexColor backward a third of its per-face terms when colours are not fitted
 * In the normal and lighting modes every gradient is zero, as the reference's gradient
 * routing returns (python/CudaRenderer.py:207-213). */
int gvv_backward(void* handle, int32_t B, int32_t texH, int32_t texW, const float* g_render, const float* g_target,
                 const float* target, const float* vpos, const float* vcol, const float* tex, const float* sh,
                 const float* vnormal, const float* bary, const int32_t* face, const float* extr,
                 const float* intr, float* g_pos, float* g_col, float* g_tex, float* g_sh, void* hip_stream);

/* gvv_backward without any device allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments (g_col may be NULL), results and
 * zero-gradient modes as gvv_backward.  The backward's scratch -- per-view cameras, the per-vertex position and
 * normal terms, the SH tile partials and the per-face constants, B*C*156 + B*N*24
 * + B*C*ceil(W/tile)*ceil(H/tile)*108 + B*F*128 bytes, each part rounded up to 256 -- lives in
 * one fixed 32 MiB device array per GPU that the library holds until the process ends; a call
 * that would need more is rejected with a message (use gvv_backward or a smaller batch).
 * For one mesh of 50k vertices and 100k faces in 8 views of 1024^2 (tile 16) that is 17 MiB.
 * Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other streams, other host
 * threads, other handles) race and must be serialised by the caller. */
int gvv_backward_noalloc(void* handle, int32_t B, int32_t texH, int32_t texW, const float* g_render,
                         const float* g_target, const float* target, const float* vpos, const float* vcol,
                         const float* tex, const float* sh, const float* vnormal, const float* bary,
                         const int32_t* face, const float* extr, const float* intr, float* g_pos, float* g_col,
                         float* g_tex, float* g_sh, void* hip_stream);

/* Launch settings of one handle (no counterpart in the reference; chosen by the autotuner of
 * the multi-mesh runtime, gvv_multi.h).  tile: pixel tile edge of the backward pixel kernel,
 * 8, 16 or 32 (default 16).  large_box: bounding-box size in pixels above which the raster
 * kernel hands a face to its persistent large-face queue, 16 .. 2^20 (default 256).  The
 * forward results do not depend on either; backward results change only in the order of
 * floating-point additions. */
int gvv_set_tuning(void* handle, int32_t tile, int32_t large_box);
int gvv_get_tuning(void* handle, int32_t* tile, int32_t* large_box);

/* Depth pass of one handle (no counterpart in the reference, whose two per-(view, face)
 * bounding-box passes both are replacements for, CUDABasedRasterization.cu:215-408).
 *   0  faces binned by bounding-box size: wavefront-staged small faces and a persistent
 *      large-face queue, a 64-bit (depth key | face id) atomicMin into a global visibility
 *      buffer, then a per-pixel resolve;
 *   1  16x16-pixel screen tiles: faces binned per tile, each tile rasterised by one workgroup
 *      against rays computed once per pixel, with its z-buffer in LDS and the resolve and
 *      shading fused at tile end (faces spanning many tiles still take the large-face queue).
 * Both give bit-identical outputs. */
int gvv_set_depth_pass(void* handle, int32_t mode);
int gvv_get_depth_pass(void* handle, int32_t* mode);

/* Measurement support (no counterpart in the reference): enqueues 2^20 threads x
 * ceil(updates / 2^20) float atomicAdd(+1.0) into buf[0, n_floats) -- scattered (hashed)
 * addresses if `scattered`, else consecutive lanes on consecutive floats -- so that the caller
 * can time the GPU's float-atomic throughput.  Overwrites buf. */
int gvv_calibrate_atomics(float* buf, int64_t n_floats, int64_t updates, int32_t scattered, void* stream);

/* Kernel timing (no counterpart in the reference, which has no timing code).  When enabled,
 * every kernel launch of this handle is bracketed by HIP events on its stream; enabling or
 * disabling clears the accumulated figures. */
int gvv_profile_enable(void* handle, int32_t enable);

/* Waits for the recorded events and writes one line per kernel, "name total_ms launches\n",
 * into buf (NUL-terminated, truncated to cap).  Returns the full report length, -1 on error. */
int64_t gvv_profile_report(void* handle, char* buf, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* GVV_RENDERER_H */
