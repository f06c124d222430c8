// gvv_forward.hip -- forward kernels of the rasterizer (gfx950).
//
// Replaces renderBuffersGPU (reference cpp/src/Renderer/CUDABasedRasterization.cu:449-473):
//   k_cameras       per-view inverse matrices and ray origin           (.cu:23-67)
//   k_face_normals  unnormalised world-space face normals, once per mesh (.cu:122-141)
//   k_vertices      vertex normals by CSR gather + per-view projection  (.cu:98-115,148-174)
//   k_raster        per (view, face) bounding-box walk; 64-bit (depth key | face id)
//                   atomicMin into the visibility buffer               (.cu:184-255)
//   k_resolve       one thread per pixel: decode the winner, recompute its barycentrics,
//                   shade, write all pixel outputs, reset the visibility key (.cu:74-91,262-408)
//   k_normal_map    texel-space normal map                             (.cu:415-445)
#include "gvv_kernels.hpp"

namespace gvv {

__global__ void __launch_bounds__(256) k_cameras(const float* __restrict__ extr, const float* __restrict__ intr,
                                                 Camera* __restrict__ cams, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Camera c;
  camera_setup(extr + 12 * (size_t)i, intr + 9 * (size_t)i, c);
  cams[i] = c;
}

__global__ void __launch_bounds__(256) k_face_normals(const float* __restrict__ vpos, const int32_t* __restrict__ faces,
                                                      float* __restrict__ fnormal, int B, int N, int F) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * F) return;
  int b = (int)(i / F), f = (int)(i % F);
  const float* vp = vpos + (size_t)b * N * 3;
  V3 v0 = ld3(vp + 3 * faces[3 * f]), v1 = ld3(vp + 3 * faces[3 * f + 1]), v2 = ld3(vp + 3 * faces[3 * f + 2]);
  V3 n = cross(v1 - v0, v2 - v0);
  float* o = fnormal + 3 * i;
  o[0] = n.x; o[1] = n.y; o[2] = n.z;
}

// One thread per (b, c, vertex): vertex normal as the CSR-ordered sum of incident face
// normals (first term assigned, the rest added), written per view; projection into view c.
__global__ void __launch_bounds__(256) k_vertices(const float* __restrict__ vpos, const float* __restrict__ fnormal,
                                                  const int32_t* __restrict__ csr_start, const int32_t* __restrict__ csr_count,
                                                  const int32_t* __restrict__ csr_list, const Camera* __restrict__ cams,
                                                  float* __restrict__ vnormal, float4* __restrict__ proj,
                                                  int B, int C, int N, int F) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * C * N) return;
  int v = (int)(i % N);
  int bc = (int)(i / N);
  int b = bc / C;
  const float* fn = fnormal + (size_t)b * F * 3;
  int s = csr_start[v], n = csr_count[v];
  V3 acc = v3(0.f, 0.f, 0.f);
  if (n > 0) {
    acc = ld3(fn + 3 * csr_list[s]);
    for (int k = 1; k < n; ++k) {
      V3 q = ld3(fn + 3 * csr_list[s + k]);
      acc.x = acc.x + q.x; acc.y = acc.y + q.y; acc.z = acc.z + q.z;
    }
  }
  float* o = vnormal + 3 * i;
  o[0] = acc.x; o[1] = acc.y; o[2] = acc.z;
  V3 p = project(cams[bc], ld3(vpos + ((size_t)b * N + v) * 3));
  proj[i] = make_float4(p.x, p.y, p.z, 0.f);
}

// One thread per (b, c, face): walk the triangle's clamped bounding box (u outer, v inner as
// the reference) and keep the smallest (depth key, face id) per pixel.
__global__ void __launch_bounds__(256) k_raster(const float* __restrict__ vpos, const int32_t* __restrict__ faces,
                                                const float4* __restrict__ proj, const Camera* __restrict__ cams,
                                                unsigned long long* __restrict__ keys, int B, int C, int N, int F,
                                                int W, int H) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * C * F) return;
  int f = (int)(i % F);
  int bc = (int)(i / F);
  int b = bc / C;
  int i0 = faces[3 * f], i1 = faces[3 * f + 1], i2 = faces[3 * f + 2];
  const float4* pc = proj + (size_t)bc * N;
  float4 p0 = pc[i0], p1 = pc[i1], p2 = pc[i2];
  Box bx = face_box(v3(p0.x, p0.y, p0.z), v3(p1.x, p1.y, p1.z), v3(p2.x, p2.y, p2.z), W, H);
  if (bx.x0 > bx.x1 || bx.y0 > bx.y1) return;
  const float* vp = vpos + (size_t)b * N * 3;
  TriSetup t = tri_setup(ld3(vp + 3 * i0), ld3(vp + 3 * i1), ld3(vp + 3 * i2));
  const Camera& cam = cams[bc];
  V3 o_s = cam_origin(cam) / 1000.f;
  unsigned long long* kc = keys + (size_t)bc * W * H;
  for (int u = bx.x0; u <= bx.x1; ++u)
    for (int v = bx.y0; v <= bx.y1; ++v) {
      V3 d = pixel_ray(cam, u + 0.5f, v + 0.5f);
      V3 abc = ray_barycentric(t, o_s, d);
      if (!inside(abc)) continue;
      uint64_t k = pack_key(depth_key(abc, p0.z, p1.z, p2.z), f);
      atomicMin(kc + (size_t)v * W + u, (unsigned long long)k);
    }
}

template <int ALBEDO, bool SHADE>
__global__ void __launch_bounds__(256) k_resolve(ResolveArgs a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t P = (int64_t)a.W * a.H;
  if (i >= (int64_t)a.B * a.C * P) return;
  unsigned long long key = a.keys[i];
  if (key == kEmptyKey) {
    a.face[i] = -1;
    a.bary[2 * i] = 0.f; a.bary[2 * i + 1] = 0.f;
    a.render[3 * i] = 0.f; a.render[3 * i + 1] = 1.f; a.render[3 * i + 2] = 0.f;
    return;
  }
  a.keys[i] = kEmptyKey;
  int bc = (int)(i / P);
  int pix = (int)(i % P);
  int b = bc / a.C, c = bc % a.C;
  int py = pix / a.W, px = pix % a.W;
  int f = (int)(uint32_t)(key & 0xffffffffu);
  int i0 = a.faces[3 * f], i1 = a.faces[3 * f + 1], i2 = a.faces[3 * f + 2];
  const float* vp = a.vpos + (size_t)b * a.N * 3;
  TriSetup t = tri_setup(ld3(vp + 3 * i0), ld3(vp + 3 * i1), ld3(vp + 3 * i2));
  const Camera& cam = a.cams[bc];
  V3 d = pixel_ray(cam, px + 0.5f, py + 0.5f);
  V3 abc = ray_barycentric(t, cam_origin(cam) / 1000.f, d);
  a.face[i] = f;
  a.bary[2 * i] = abc.x; a.bary[2 * i + 1] = abc.y;

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
  a.render[3 * i] = col.x; a.render[3 * i + 1] = col.y; a.render[3 * i + 2] = col.z;
}

__global__ void __launch_bounds__(256) k_normal_map(const UvTexel* __restrict__ table, const int32_t* __restrict__ faces,
                                                    const float* __restrict__ vnormal, float* __restrict__ nmap,
                                                    int B, int C, int N, int T) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * T) return;
  int b = (int)(i / T), t = (int)(i % T);
  UvTexel e = table[t];
  const float* vn = vnormal + (size_t)b * C * N * 3;   // view 0
  int f = e.face;
  V3 n = ld3(vn + 3 * faces[3 * f]) * e.a + ld3(vn + 3 * faces[3 * f + 1]) * e.b + ld3(vn + 3 * faces[3 * f + 2]) * e.c;
  float len = length(n);
  if (len != 0.f) n = n / len;
  float* o = nmap + 3 * i;
  o[0] = (n.x + 1.f) / 2.f; o[1] = (n.y + 1.f) / 2.f; o[2] = (n.z + 1.f) / 2.f;
}

__global__ void __launch_bounds__(256) k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_f32(float* p, int64_t n, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------------------
// host-side launchers
// ------------------------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(n)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, int64_t n, float v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(n)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_cameras(const float* extr, const float* intr, Camera* cams, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_cameras, dim3(blocks_for(n)), dim3(256), 0, s, extr, intr, cams, n);
  return hipGetLastError();
}

template <int ALBEDO, bool SHADE>
static void launch_resolve_t(const ResolveArgs& a, hipStream_t s) {
  int64_t n = (int64_t)a.B * a.C * a.W * a.H;
  hipLaunchKernelGGL((k_resolve<ALBEDO, SHADE>), dim3(blocks_for(n)), dim3(256), 0, s, a);
}

hipError_t launch_forward(const ForwardPlan& p, hipStream_t s) {
  const int B = p.B, C = p.C, N = p.N, F = p.F;
  {
    ProfScope ps(p.prof, "k_cameras", s);
    hipError_t e = launch_cameras(p.extr, p.intr, p.cams, B * C, s);
    if (e != hipSuccess) return e;
  }
  {
    ProfScope ps(p.prof, "k_face_normals", s);
    hipLaunchKernelGGL(k_face_normals, dim3(blocks_for((int64_t)B * F)), dim3(256), 0, s, p.vpos, p.faces,
                       p.fnormal, B, N, F);
  }
  {
    ProfScope ps(p.prof, "k_vertices", s);
    hipLaunchKernelGGL(k_vertices, dim3(blocks_for((int64_t)B * C * N)), dim3(256), 0, s, p.vpos, p.fnormal,
                       p.csr_start, p.csr_count, p.csr_list, p.cams, p.vnormal_out, p.proj, B, C, N, F);
  }
  if (p.normal_map) {
    ProfScope ps(p.prof, "k_normal_map", s);
    int T = p.r.texW * p.r.texH;
    hipLaunchKernelGGL(k_normal_map, dim3(blocks_for((int64_t)B * T)), dim3(256), 0, s, p.uv_table, p.faces,
                       p.vnormal_out, p.normal_map_out, B, C, N, T);
  } else {
    ProfScope ps(p.prof, "k_raster", s);
    hipLaunchKernelGGL(k_raster, dim3(blocks_for((int64_t)B * C * F)), dim3(256), 0, s, p.vpos, p.faces, p.proj,
                       p.cams, p.r.keys, B, C, N, F, p.r.W, p.r.H);
  }
  // In normal-map mode the visibility buffer is empty, so this writes the background.
  ProfScope ps(p.prof, "k_resolve", s);
  switch (p.albedo * 2 + (p.shade ? 1 : 0)) {
    case kVertexColor * 2 + 0: launch_resolve_t<kVertexColor, false>(p.r, s); break;
    case kVertexColor * 2 + 1: launch_resolve_t<kVertexColor, true>(p.r, s); break;
    case kTextured * 2 + 0: launch_resolve_t<kTextured, false>(p.r, s); break;
    case kTextured * 2 + 1: launch_resolve_t<kTextured, true>(p.r, s); break;
    case kNormal * 2 + 0: launch_resolve_t<kNormal, false>(p.r, s); break;
    case kNormal * 2 + 1: launch_resolve_t<kNormal, true>(p.r, s); break;
    case kLighting * 2 + 0: launch_resolve_t<kLighting, false>(p.r, s); break;
    case kLighting * 2 + 1: launch_resolve_t<kLighting, true>(p.r, s); break;
    case kForegroundMask * 2 + 0: launch_resolve_t<kForegroundMask, false>(p.r, s); break;
    default: launch_resolve_t<kForegroundMask, true>(p.r, s); break;
  }
  return hipGetLastError();
}

}  // namespace gvv
