// gvv_shade.hpp -- the per-pixel outputs of the forward for one pixel whose visible face is
// known: face id, the barycentrics recomputed for that face (bit-identical to the depth pass),
// and the shaded colour (reference renderBuffersDevice, CUDABasedRasterization.cu:262-408).
// Shared by the resolve after the face-binned depth pass and by the screen-tile kernel.
#pragma once
#include "gvv_kernels.hpp"

namespace gvv {

// The pixel outputs are written once and not read again by the forward: streaming
// (non-temporal) stores, which keep them from displacing the vertex, texel and key lines the
// resolve reads (k_resolve 0.104 -> 0.090 ms at 8 x 1024^2).
template <class T>
__device__ __forceinline__ void st_out(T* p, T v) { __builtin_nontemporal_store(v, p); }

// Background: face -1, barycentrics 0, render (0, 1, 0) (CUDABasedRasterization.cu:80-89).
__device__ __forceinline__ void background_pixel(const ResolveArgs& a, int64_t i) {
  st_out(a.face + i, -1);
  st_out(a.bary + 2 * i, 0.f); st_out(a.bary + 2 * i + 1, 0.f);
  st_out(a.render + 3 * i, 0.f); st_out(a.render + 3 * i + 1, 1.f); st_out(a.render + 3 * i + 2, 0.f);
}

// Pixel i of view bc with visible face f and unit ray d through the pixel centre.
template <int ALBEDO, bool SHADE>
__device__ __forceinline__ void shade_pixel(const ResolveArgs& a, int bc, int64_t i, int f, V3 d) {
  const int b = bc / a.C;
  const int i0 = a.faces[3 * f], i1 = a.faces[3 * f + 1], i2 = a.faces[3 * f + 2];
  // recomputed from the vertices: cheaper here than a 64-byte gather of the face's constants
  const float* vp = a.vpos + (size_t)b * a.N * 3;
  const TriSetup t = tri_setup(ld3(vp + 3 * i0), ld3(vp + 3 * i1), ld3(vp + 3 * i2));
  const Camera& cam = a.cams[bc];
  const V3 abc = ray_barycentric(t, cam_origin_scaled(cam), d);
  st_out(a.face + i, f);
  st_out(a.bary + 2 * i, abc.x); st_out(a.bary + 2 * i + 1, abc.y);

  const float* vn = a.vnormal + (size_t)bc * a.N * 3;
  V3 n = pixel_normal(ld3(vn + 3 * i0), ld3(vn + 3 * i1), ld3(vn + 3 * i2), abc);
  if (dot(n, d) > 0.f) n = -n;
  V3 col;
  if (ALBEDO == kTextured) {
    TexCoord tc = tex_coord(a.uv + 6 * (size_t)f, abc, a.texW, a.texH);
    int row = min(max(tc.LV, 0), a.texH - 1), cl = min(max(tc.LU, 0), a.texW - 1);
    col = ld3(a.tex + (size_t)b * a.texH * a.texW * 3 + 3 * ((size_t)a.texW * row + cl));
  } else if (ALBEDO == kVertexColor) {
    const float* vc = a.vcol + (size_t)b * a.N * 3;
    V3 c0 = ld3(vc + 3 * i0), c1 = ld3(vc + 3 * i1), c2 = ld3(vc + 3 * i2);
    col = v3(c0.x * abc.x + c1.x * abc.y + c2.x * abc.z, c0.y * abc.x + c1.y * abc.y + c2.y * abc.z,
             c0.z * abc.x + c1.z * abc.y + c2.z * abc.z);
  } else if (ALBEDO == kNormal) {
    col = v3((1.f + n.x) / 2.f, (1.f + n.y) / 2.f, (1.f + n.z) / 2.f);
  } else {
    col = v3(1.f, 1.f, 1.f);
  }
  if (SHADE) {
    V3 L = sh_illum(n, a.sh + (size_t)bc * 27);
    col = v3(L.x * col.x, L.y * col.y, L.z * col.z);
  }
  st_out(a.render + 3 * i, col.x); st_out(a.render + 3 * i + 1, col.y); st_out(a.render + 3 * i + 2, col.z);
}

}  // namespace gvv
