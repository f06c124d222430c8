/* gvv_fit.h -- C ABI of the analysis-by-synthesis fit harness's photometric loss.
 *
 * No counterpart in the reference op pair.  The reference's use case fits parameters by
 * rendering and comparing with a target under the foreground mask of the rendered face buffer
 * (the mask of python/CudaRenderer.py:134-140, getModelMaskTF); these functions compute that
 * masked squared error and its gradient with respect to the render in one pass each, per
 * subject of a batch, so that the fit's loss does not go through chains of broadcast
 * element-wise kernels and a reduction over millions of entries per subject.
 *
 * Layouts (row-major, contiguous, float32 device pointers):
 *   render, target  [B][P][3]   P pixels per subject (all views of one subject together)
 *   mask            [B][P]      one weight per pixel (1 foreground, 0 background)
 *   sq, g_sq        [B]
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

#ifdef __cplusplus
}
#endif

#endif /* GVV_FIT_H */
