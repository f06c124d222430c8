/* gvv_fit.h -- C ABI of the analysis-by-synthesis fit harness's own kernels: the photometric
 * loss and the shape model that feeds the renderer its vertices.
 *
 * No counterpart in the reference op pair.  The reference's use case fits parameters by
 * rendering and comparing with a target under the foreground mask of the rendered face buffer
 * (the mask of python/CudaRenderer.py:134-140, getModelMaskTF), with any rigid motion applied
 * to the vertices before the renderer (its extrinsics receive no gradient,
 * python/CudaRenderer.py:215).  These functions compute the masked squared error and its
 * gradient with respect to the render in one pass each, per subject of a batch, and the posed
 * vertices of a linear shape basis with their gradients, so that the fit's step does not go
 * through chains of broadcast element-wise kernels, a reduction over millions of entries per
 * subject and small library GEMMs.
 *
 * Layouts (row-major, contiguous, float32 device pointers):
 *   render, target  [B][P][3]   P pixels per subject (all views of one subject together)
 *   mask            [B][P]      one weight per pixel (1 foreground, 0 background)
 *   sq, g_sq        [B]
 *   alpha           [B][K]      shape coefficients (K <= 128)
 *   pose            [B][6]      per subject: axis-angle rotation (rad), then translation units
 *   rt              [B][12]     per subject: rotation R (row-major 3x3), then translation t
 *   templ           [N][3]      template vertices T; basis [K][N][3]; centre c [3]
 *   v, xc, g_v      [B][N][3]
 * Returns 0 on success; otherwise gvv_last_error() (gvv_renderer.h) holds the message.
 * Work goes to `stream` (hipStream_t, NULL = default stream).
 */
#ifndef GVV_FIT_H
#define GVV_FIT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Floats of workspace gvv_masked_sq_error needs for B subjects of P pixels (-1 if invalid). */
int64_t gvv_masked_sq_workspace_floats(int32_t B, int64_t P);

/* sq[b] = sum over pixels p and channels ch of ((render - target)[b][p][ch] * mask[b][p])^2.
 * Deterministic: every workgroup sums a fixed pixel range in a fixed order and a second launch
 * adds the workgroups' sums in a fixed order.  `work` holds gvv_masked_sq_workspace_floats. */
int gvv_masked_sq_error(int32_t B, int64_t P, const float* render, const float* target, const float* mask,
                        float* work, float* sq, void* stream);

/* g_render[b][p][ch] = 2 (render - target)[b][p][ch] mask[b][p]^2 g_sq[b], the gradient of
 * sum_b g_sq[b] sq[b] with respect to the render. */
int gvv_masked_sq_error_grad(int32_t B, int64_t P, const float* render, const float* target, const float* mask,
                             const float* g_sq, float* g_render, void* stream);

/* Posed vertices v[b][i] = R_b (T_i + sum_k alpha[b][k] basis[k][i] - c) + c + t_b for the pose
 * pose[b] = (w, p) [B][6]: R_b the rotation of axis-angle w (rad; Rodrigues' formula, with its
 * series below |w|^2 = 1e-8) and t_b = tscale * p.  Also writes xc[b][i] = T_i + sum_k
 * alpha[b][k] basis[k][i] - c (the unrotated centred vertex) and rt[b] = (R_b row-major,
 * t_b) [B][12], which gvv_shape_backward takes back. */
int gvv_shape_forward(int32_t B, int32_t N, int32_t K, const float* alpha, const float* pose, float tscale,
                      const float* templ, const float* basis, const float* centre, float* v, float* xc, float* rt,
                      void* stream);

/* Floats of workspace gvv_shape_backward needs (-1 if invalid). */
int64_t gvv_shape_workspace_floats(int32_t B, int32_t N, int32_t K);

/* Gradients of sum g_v . v: g_alpha_pose [B][K + 6] holds per subject dL/dalpha (K), dL/dw (3,
 * through Rodrigues' formula from dL/dR = sum_i g_v[b][i] xc[b][i]^T) and dL/dp (3).
 * Deterministic: fixed-order sums over chunks of 64 vertices.  `work` holds
 * gvv_shape_workspace_floats. */
int gvv_shape_backward(int32_t B, int32_t N, int32_t K, const float* g_v, const float* xc, const float* pose,
                       const float* rt, float tscale, const float* basis, float* work, float* g_alpha_pose,
                       void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GVV_FIT_H */
