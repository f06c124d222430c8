/* gvv_proposal.h -- C ABI of the proposal network's BF16 MFMA kernels (libgvv_amd.so).
 *
 * These entry points have no counterpart in the reference, which has no learned component.
 * They serve the proposal network that predicts shape, SH and pose coefficients from renders
 * to warm-start the analysis-by-synthesis fit (the reference's fit loop:
 * python/test_gradients_VertexColor.py:93-138).
 *
 * Conventions
 *   - Activations are NHWC bf16 (raw 16-bit storage) with channel counts that are multiples
 *     of 8.  Weights are [Cout][ks][ks][Cin]; their bf16 copy has the same layout and the
 *     transposed copy is [Cin][ks][ks][Cout].  Biases and gradients of weights/biases are fp32.
 *   - Output size per spatial dim: (in + 2*pad - ks) / stride + 1 (gvv_conv_out_size).
 *   - Pointers are device pointers; x, w, dz and wt must be 16-byte aligned.
 *   - Returns 0 on success; otherwise gvv_last_error() (gvv_renderer.h) holds the message.
 *   - Work goes to `stream` (hipStream_t, NULL = default stream).
 */
#ifndef GVV_PROPOSAL_H
#define GVV_PROPOSAL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int gvv_conv_out_size(int32_t in, int32_t ks, int32_t stride, int32_t pad);

/* y = act(conv(x, w) + bias); act = ReLU if relu != 0.  y is bf16 NHWC, or fp32 if y_f32. */
int gvv_conv_forward(int32_t B, int32_t H, int32_t W, int32_t Cin, int32_t Cout, int32_t ks, int32_t stride,
                     int32_t pad, const void* x, const void* w, const float* bias, int32_t relu, void* y,
                     int32_t y_f32, void* stream);

/* dx = conv_transpose(dz, w) [* (x_mask > 0) elementwise if x_mask != NULL], bf16 NHWC
 * [B][H][W][Cin].  wt is the transposed bf16 weight copy.  Passing the layer's forward input
 * as x_mask yields the gradient at the previous layer's pre-activation (ReLU backward). */
int gvv_conv_backward_data(int32_t B, int32_t H, int32_t W, int32_t Cin, int32_t Cout, int32_t ks, int32_t stride,
                           int32_t pad, const void* dz, const void* wt, const void* x_mask, void* dx, void* stream);

/* Floats of workspace gvv_conv_backward_weight needs for this geometry (-1 if invalid). */
int64_t gvv_conv_workspace_floats(int32_t B, int32_t H, int32_t W, int32_t Cin, int32_t Cout, int32_t ks,
                                  int32_t stride, int32_t pad);

/* dw [Cout][ks][ks][Cin] = sum over pixels of dz (x) x-patch; db [Cout] = sum over pixels of dz
 * (skipped if db == NULL).  Both fp32, overwritten (not accumulated).  Deterministic. */
int gvv_conv_backward_weight(int32_t B, int32_t H, int32_t W, int32_t Cin, int32_t Cout, int32_t ks, int32_t stride,
                             int32_t pad, const void* x, const void* dz, float* dw, float* db, float* workspace,
                             int64_t workspace_floats, void* stream);

/* fp32 weights [Cout][taps][Cin] -> bf16 copy (same layout) and, if wt_bf16 != NULL, the bf16
 * transpose [Cin][taps][Cout]. */
int gvv_conv_prepare_weights(int32_t Cout, int32_t taps, int32_t Cin, const float* w, void* w_bf16, void* wt_bf16,
                             void* stream);

/* fp32 renders [B][V][h][w][3] -> bf16 NHWC8 mosaic [B][ceil(V/cols)*h/down][cols*w/down][8]:
 * view v at tile (v / cols, v % cols), each pixel the mean of a down x down block of the view.
 * features 0: channels 0..2 RGB, 3..7 zero.  features 1: channels 0..2 RGB - 0.5, 3 the covered
 * fraction of the block (pixels other than the (0, 1, 0) background), 4..5 the tile-local x, y
 * in [-1, 1], 6..7 zero.  Empty tiles are zero.  h and w must be multiples of down. */
int gvv_pack_mosaic(const float* render, int32_t B, int32_t V, int32_t h, int32_t w, int32_t cols, int32_t down,
                    int32_t features, void* x, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GVV_PROPOSAL_H */
