// gvv_math.hpp -- binary32 device math for the renderer's visibility stage.
//
// Every routine here evaluates IEEE binary32 operations in a fixed order, with the library
// compiled under -ffp-contract=off and correctly rounded division/sqrt, so triangle IDs,
// barycentrics and vertex normals are bit-reproducible.  The operand order follows the
// reference (citations relative to the reference repository root):
//   vector ops           cpp/thirdParty/Shared/cutil/inc/cutil_math.h:881,1123,1166,1295
//   camera inverse       cpp/src/Utils/float4x4.h:78-101,160-285
//   projection / ray     cpp/src/Utils/CameraUtil.h:141-186,223-258
//   ray/triangle         cpp/src/Utils/RendererUtil.h:26-128
// `normalize` multiplies by a correctly rounded 1/sqrt instead of the reference's approximate
// rsqrtf, and float->int conversion saturates with NaN->0.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GVV_HD __host__ __device__ __forceinline__

namespace gvv {

struct V3 { float x, y, z; };
GVV_HD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
GVV_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
GVV_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
GVV_HD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
GVV_HD V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
GVV_HD V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
GVV_HD V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
GVV_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
GVV_HD V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
GVV_HD float length(V3 a) { return sqrtf(dot(a, a)); }
GVV_HD V3 normalize(V3 a) { float inv = 1.0f / sqrtf(dot(a, a)); return a * inv; }
GVV_HD V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

GVV_HD int32_t sat_i32(float z) {
  if (!(z == z)) return 0;
  if (z >= 2147483648.0f) return INT32_MAX;
  if (z <= -2147483648.0f) return INT32_MIN;
  return (int32_t)z;
}

// Visibility key: ascending in (depth key, face id); all-ones marks an empty pixel.
GVV_HD uint64_t pack_key(int32_t zq, int32_t face) {
  return ((uint64_t)((uint32_t)zq ^ 0x80000000u) << 32) | (uint64_t)(uint32_t)face;
}
constexpr uint64_t kEmptyKey = ~0ull;

// Per-view constants: rows 0..2 of (K E)^-1, the ray origin (and the origin / 1000 that the
// scaled ray/triangle test uses), and the raw E and K.
struct Camera {
  float invP[12];
  float o[3];
  float os[3];
  float E[12];
  float K[9];
};

// Six signed triple products per entry of the 4x4 cofactor inverse, in the reference's
// evaluation order (row-major entries).
struct CofTerm { int8_t s, a, b, c; };

GVV_HD void camera_setup(const float* extr12, const float* intr9, Camera& cam) {
  const CofTerm kCof[16][6] = {
    {{+1, 5, 10, 15}, {-1, 5, 11, 14}, {-1, 9, 6, 15}, {+1, 9, 7, 14}, {+1, 13, 6, 11}, {-1, 13, 7, 10}},
    {{-1, 1, 10, 15}, {+1, 1, 11, 14}, {+1, 9, 2, 15}, {-1, 9, 3, 14}, {-1, 13, 2, 11}, {+1, 13, 3, 10}},
    {{+1, 1, 6, 15}, {-1, 1, 7, 14}, {-1, 5, 2, 15}, {+1, 5, 3, 14}, {+1, 13, 2, 7}, {-1, 13, 3, 6}},
    {{-1, 1, 6, 11}, {+1, 1, 7, 10}, {+1, 5, 2, 11}, {-1, 5, 3, 10}, {-1, 9, 2, 7}, {+1, 9, 3, 6}},
    {{-1, 4, 10, 15}, {+1, 4, 11, 14}, {+1, 8, 6, 15}, {-1, 8, 7, 14}, {-1, 12, 6, 11}, {+1, 12, 7, 10}},
    {{+1, 0, 10, 15}, {-1, 0, 11, 14}, {-1, 8, 2, 15}, {+1, 8, 3, 14}, {+1, 12, 2, 11}, {-1, 12, 3, 10}},
    {{-1, 0, 6, 15}, {+1, 0, 7, 14}, {+1, 4, 2, 15}, {-1, 4, 3, 14}, {-1, 12, 2, 7}, {+1, 12, 3, 6}},
    {{+1, 0, 6, 11}, {-1, 0, 7, 10}, {-1, 4, 2, 11}, {+1, 4, 3, 10}, {+1, 8, 2, 7}, {-1, 8, 3, 6}},
    {{+1, 4, 9, 15}, {-1, 4, 11, 13}, {-1, 8, 5, 15}, {+1, 8, 7, 13}, {+1, 12, 5, 11}, {-1, 12, 7, 9}},
    {{-1, 0, 9, 15}, {+1, 0, 11, 13}, {+1, 8, 1, 15}, {-1, 8, 3, 13}, {-1, 12, 1, 11}, {+1, 12, 3, 9}},
    {{+1, 0, 5, 15}, {-1, 0, 7, 13}, {-1, 4, 1, 15}, {+1, 4, 3, 13}, {+1, 12, 1, 7}, {-1, 12, 3, 5}},
    {{-1, 0, 5, 11}, {+1, 0, 7, 9}, {+1, 4, 1, 11}, {-1, 4, 3, 9}, {-1, 8, 1, 7}, {+1, 8, 3, 5}},
    {{-1, 4, 9, 14}, {+1, 4, 10, 13}, {+1, 8, 5, 14}, {-1, 8, 6, 13}, {-1, 12, 5, 10}, {+1, 12, 6, 9}},
    {{+1, 0, 9, 14}, {-1, 0, 10, 13}, {-1, 8, 1, 14}, {+1, 8, 2, 13}, {+1, 12, 1, 10}, {-1, 12, 2, 9}},
    {{-1, 0, 5, 14}, {+1, 0, 6, 13}, {+1, 4, 1, 14}, {-1, 4, 2, 13}, {-1, 12, 1, 6}, {+1, 12, 2, 5}},
    {{+1, 0, 5, 10}, {-1, 0, 6, 9}, {-1, 4, 1, 10}, {+1, 4, 2, 9}, {+1, 8, 1, 6}, {-1, 8, 2, 5}},
  };
  float K4[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  float E4[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) K4[r * 4 + c] = intr9[r * 3 + c];
    K4[r * 4 + 3] = 0.f;
    for (int c = 0; c < 4; ++c) E4[r * 4 + c] = extr12[r * 4 + c];
  }
  float M[2][16];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c)
      M[1][r * 4 + c] = K4[r * 4 + 0] * E4[0 * 4 + c] + K4[r * 4 + 1] * E4[1 * 4 + c] +
                        K4[r * 4 + 2] * E4[2 * 4 + c] + K4[r * 4 + 3] * E4[3 * 4 + c];
  for (int k = 0; k < 16; ++k) M[0][k] = E4[k];
  float out[2][16];
  for (int m = 0; m < 2; ++m) {
    const float* e = M[m];
    float inv[16];
    for (int k = 0; k < 16; ++k) {
      const CofTerm* t = kCof[k];
      float acc = (e[t[0].a] * e[t[0].b]) * e[t[0].c];
      if (t[0].s < 0) acc = -acc;
      for (int j = 1; j < 6; ++j) {
        float p = (e[t[j].a] * e[t[j].b]) * e[t[j].c];
        acc = t[j].s > 0 ? acc + p : acc - p;
      }
      inv[k] = acc;
    }
    float det = e[0] * inv[0] + e[1] * inv[4] + e[2] * inv[8] + e[3] * inv[12];
    float detr = 1.0f / det;
    for (int k = 0; k < 16; ++k) out[m][k] = inv[k] * detr;
  }
  for (int k = 0; k < 12; ++k) cam.invP[k] = out[1][k];
  float w = out[0][15];
  cam.o[0] = out[0][3] / w; cam.o[1] = out[0][7] / w; cam.o[2] = out[0][11] / w;
  for (int k = 0; k < 3; ++k) cam.os[k] = cam.o[k] / 1000.f;
  for (int k = 0; k < 12; ++k) cam.E[k] = extr12[k];
  for (int k = 0; k < 9; ++k) cam.K[k] = intr9[k];
}

GVV_HD V3 cam_origin(const Camera& cam) { return v3(cam.o[0], cam.o[1], cam.o[2]); }
GVV_HD V3 cam_origin_scaled(const Camera& cam) { return v3(cam.os[0], cam.os[1], cam.os[2]); }

// Unit ray direction through pixel position (px, py), back-projected at depth 1000.
GVV_HD V3 pixel_ray(const Camera& cam, float px, float py) {
  const float z = 1000.f;
  float tx = px * z, ty = py * z, tz = z;
  const float* P = cam.invP;
  V3 t = v3(P[0] * tx + P[1] * ty + P[2] * tz + P[3] * 1.f,
            P[4] * tx + P[5] * ty + P[6] * tz + P[7] * 1.f,
            P[8] * tx + P[9] * ty + P[10] * tz + P[11] * 1.f);
  return normalize(t - cam_origin(cam));
}

// World -> pixel (x/z, y/z, z) with the reference's z clamp.
GVV_HD V3 project(const Camera& cam, V3 p) {
  const float* E = cam.E;
  V3 c = v3(E[0] * p.x + E[1] * p.y + E[2] * p.z + E[3] * 1.f,
            E[4] * p.x + E[5] * p.y + E[6] * p.z + E[7] * 1.f,
            E[8] * p.x + E[9] * p.y + E[10] * p.z + E[11] * 1.f);
  const float* K = cam.K;
  float x = c.x * K[0] + c.y * K[1] + c.z * K[2];
  float y = c.x * K[3] + c.y * K[4] + c.z * K[5];
  float z = c.x * K[6] + c.y * K[7] + c.z * K[8];
  if (!(z > 0.0000001f)) z = 0.00001f;
  return v3(x / z, y / z, z);
}

// x / 1000, correctly rounded, as a multiply and two FMAs (the remainder x - q*1000 of the
// first quotient q = x * RN(1/1000) is exact; one Newton correction then rounds correctly).
// Checked bit-for-bit against the IEEE division for every binary32 x with |x| >= 2^-120; below
// that (subnormal quotients, signed zero) the division itself is used.
GVV_HD float div1000(float x) {
  const float y = 0.001f;
  if (!(fabsf(x) >= 0x1p-120f)) return x / 1000.f;
  const float q = x * y;
  const float r = fmaf(-q, 1000.f, x);
  return fmaf(r, y, q);
}
GVV_HD V3 div1000(V3 a) { return v3(div1000(a.x), div1000(a.y), div1000(a.z)); }

// Triangle constants of the scaled ray/triangle test, shared by every pixel and every view of
// the triangle (64 bytes: four 16-byte loads).
struct alignas(16) TriSetup {
  V3 v0, v1, v2;   // vertices / 1000
  V3 N;            // cross(v1 - v0, v2 - v0)
  float dv0N;      // dot(v0, N)
  float den;       // dot(N, N)
  float pad[2];
};

GVV_HD TriSetup tri_setup(V3 v0, V3 v1, V3 v2) {
  TriSetup t;
  t.v0 = div1000(v0); t.v1 = div1000(v1); t.v2 = div1000(v2);
  t.N = cross(t.v1 - t.v0, t.v2 - t.v0);
  t.dv0N = dot(t.v0, t.N);
  t.den = dot(t.N, t.N);
  t.pad[0] = t.pad[1] = 0.f;
  return t;
}

// Barycentrics (a -> v0, b -> v1, c -> v2) of the ray hit, or (-1,-1,-1) on a miss.
// o_s is the ray origin already divided by 1000.
GVV_HD V3 ray_barycentric(const TriSetup& t, V3 o_s, V3 d) {
  const V3 miss = v3(-1.f, -1.f, -1.f);
  float ndr = dot(d, t.N);
  if (fabsf(ndr) < 0.0000001f) return miss;
  float tt = (t.dv0N - dot(o_s, t.N)) / ndr;
  if (tt < 0) return miss;
  V3 P = o_s + tt * d;
  if (dot(t.N, cross(t.v1 - t.v0, P - t.v0)) < 0) return miss;
  float a = dot(t.N, cross(t.v2 - t.v1, P - t.v1));
  if (a < 0) return miss;
  float b = dot(t.N, cross(t.v0 - t.v2, P - t.v2));
  if (b < 0) return miss;
  a /= t.den; b /= t.den;
  return v3(a, b, 1.f - a - b);
}

GVV_HD bool inside(V3 abc) {
  return (abc.x >= -0.001f) && (abc.y >= -0.001f) && (abc.z >= -0.001f) &&
         (abc.x <= 1.001f) && (abc.y <= 1.001f) && (abc.z <= 1.001f);
}

GVV_HD int32_t depth_key(V3 abc, float z0, float z1, float z2) {
  float z = 1.f / (abc.x / z0 + abc.y / z1 + abc.z / z2);
  z *= 10000.f;
  return sat_i32(z);
}

// Interpolated, normalised vertex normal (before any flip).
GVV_HD V3 pixel_normal(V3 n0, V3 n1, V3 n2, V3 abc) {
  V3 n = n0 * abc.x + n1 * abc.y + n2 * abc.z;
  return n / length(n);
}

struct TexCoord { float U0, V0; int LU, LV; };

// Per-face (u, 1-v) corners interpolated, scaled to texels, clamped; nearest texel index.
GVV_HD TexCoord tex_coord(const float* uv6, V3 abc, int texW, int texH) {
  float t0x = uv6[0], t0y = 1.f - uv6[1];
  float t1x = uv6[2], t1y = 1.f - uv6[3];
  float t2x = uv6[4], t2y = 1.f - uv6[5];
  float fx = t0x * abc.x + t1x * abc.y + t2x * abc.z;
  float fy = t0y * abc.x + t1y * abc.y + t2y * abc.z;
  fx = fx * (float)texW;
  fy = fy * (float)texH;
  fx = fmaxf(fx, 0.f); fx = fminf(fx, (float)(texW - 1));
  fy = fmaxf(fy, 0.f); fy = fminf(fy, (float)(texH - 1));
  TexCoord t;
  t.U0 = fx; t.V0 = fy;
  t.LU = sat_i32(fx - 0.5f);
  t.LV = sat_i32(fy - 0.5f);
  return t;
}

// Screen bounding box of a projected triangle: +-0.5 px, clamped, truncated.
struct Box { int x0, y0, x1, y1; };
GVV_HD Box face_box(V3 a, V3 b, V3 c, int W, int H) {
  Box r;
  r.x0 = sat_i32(fmaxf(fminf(a.x, fminf(b.x, c.x)) - 0.5f, 0.f));
  r.y0 = sat_i32(fmaxf(fminf(a.y, fminf(b.y, c.y)) - 0.5f, 0.f));
  r.x1 = sat_i32(fminf(fmaxf(a.x, fmaxf(b.x, c.x)) + 0.5f, (float)(W - 1)));
  r.y1 = sat_i32(fminf(fmaxf(a.y, fmaxf(b.y, c.y)) + 0.5f, (float)(H - 1)));
  return r;
}

// 9-term SH illumination, channel c uses sh[c*9 .. c*9+8] (RendererUtil.h:135-214).
GVV_HD V3 sh_illum(V3 n, const float* sh) {
  float L[3];
  for (int ch = 0; ch < 3; ++ch) {
    const float* s = sh + 9 * ch;
    float l = s[0];
    l += s[1] * n.y;
    l += s[2] * n.z;
    l += s[3] * n.x;
    l += s[4] * (n.x * n.y);
    l += s[5] * (n.z * n.y);
    l += s[6] * (3.f * (n.z * n.z) - 1.f);
    l += s[7] * (n.x * n.z);
    l += s[8] * ((n.x * n.x) - (n.y * n.y));
    L[ch] = l;
  }
  return v3(L[0], L[1], L[2]);
}

}  // namespace gvv
