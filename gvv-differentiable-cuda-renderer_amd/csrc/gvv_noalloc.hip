// gvv_noalloc.hip -- gvv_backward_noalloc: the backward with its scratch in a fixed device
// array, for callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
//
// The backward's scratch -- per-view cameras, dL/d(vertex normal), the SH tile partials and
// the per-face constants -- lives in `scratch`, a file-scope device array of kScratchBytes that every GPU the library runs
// on holds until the process ends.  In host code the array's name is only a host-side stand-in:
// its device address is taken with hipGetSymbolAddress on the first call on each device (with
// that device current), checked, and kept per device; only that address is passed to kernels.
// Every call on one GPU shares that GPU's array, so calls that overlap on one GPU (different
// streams, different host threads) race and must be serialised by the caller.  Everything the
// kernels read from the array is written earlier in the same call: the cameras, the per-face
// constants and the zero fill of the normal gradients by the setup launch, every SH partial by
// the pixel kernel.
#include <mutex>
#include <string>

#include "gvv_handle.hpp"
#include "../../include/gvv_renderer.h"

namespace {

using namespace gvv;

constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
int scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return set_last_error("device index out of range for the no-alloc scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return 0;
}

inline size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

extern "C" int gvv_backward_noalloc(void* handle, int32_t B, int32_t texH, int32_t texW, const float* g_render,
                                    const float* g_target, const float* target, const float* vpos,
                                    const float* vcol, const float* tex, const float* sh, const float* vnormal,
                                    const float* bary, const int32_t* face, const float* extr, const float* intr,
                                    float* g_pos, float* g_col, float* g_tex, float* g_sh, void* hip_stream) {
  Handle* h = (Handle*)handle;
  if (!h) return set_last_error("null handle");
  std::lock_guard<std::mutex> lock(h->mu);
  if (check_device(h)) return 1;
  if (B <= 0) return set_last_error("batch size must be > 0");
  if (texH <= 0 || texW <= 0) return set_last_error("texture height/width must be > 0");
  if (!g_pos || !g_tex || !g_sh) return set_last_error("null gradient output pointer");   // g_col may be NULL
  hipStream_t s = (hipStream_t)hip_stream;
  hipError_t e = hipSuccess;
  if (h->albedo == kNormal || h->albedo == kLighting) {
    // zero gradients, as gvv_backward (python/CudaRenderer.py:207-213); no scratch needed
    const int64_t nv = (int64_t)B * h->N * 3;
    if (e == hipSuccess) e = launch_fill_f32(g_pos, nv, 0.f, s);
    if (e == hipSuccess && g_col) e = launch_fill_f32(g_col, nv, 0.f, s);
    if (e == hipSuccess) e = launch_fill_f32(g_tex, (int64_t)B * texH * texW * 3, 0.f, s);
    if (e == hipSuccess) e = launch_fill_f32(g_sh, (int64_t)B * h->C * 27, 0.f, s);
    return e == hipSuccess ? 0 : set_last_error(std::string("gvv_backward_noalloc: ") + hipGetErrorString(e));
  }
  const BackwardScratch n = backward_scratch(h, B);
  const size_t off_vn = align256(n.cams * sizeof(Camera));
  const size_t off_sh = off_vn + align256(n.g_pv * sizeof(float));
  const size_t off_face = off_sh + align256(n.sh_partial * sizeof(float));
  const size_t need = off_face + n.face_grad * sizeof(FaceGrad);
  if (need > kScratchBytes)
    return set_last_error("gvv_backward_noalloc: batch " + std::to_string(B) + " needs " + std::to_string(need) +
                          " bytes of scratch, more than its fixed " + std::to_string(kScratchBytes) +
                          " bytes; use gvv_backward or a smaller batch");
  unsigned char* base = nullptr;
  if (scratch_address(h->device, &base)) return 1;
  BackwardPlan p;
  if (plan_backward_into(h, B, texH, texW, g_render, g_target, target, vpos, vcol, tex, sh, vnormal, bary, face, extr,
                         intr, g_pos, g_col, g_tex, g_sh, reinterpret_cast<Camera*>(base),
                         reinterpret_cast<float*>(base + off_vn), reinterpret_cast<float*>(base + off_sh),
                         reinterpret_cast<FaceGrad*>(base + off_face), p))
    return 1;
  p.prof = nullptr;
  e = launch_backward(p, s);
  return e == hipSuccess ? 0 : set_last_error(std::string("gvv_backward_noalloc: ") + hipGetErrorString(e));
}
