// gvv_oracle.cpp -- CPU restatement of the reference rasterizer, forward and backward.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP renderer is compared
// against, and the CPU baseline that bench.py times.  The product path (the HIP library
// and the Python drop-in) never links, loads or calls it.
//
// Numeric contract (all citations relative to the reference repository root):
//   * Visibility stage in IEEE binary32, compiled with -ffp-contract=off: camera inverses,
//     projection, bounding boxes, pixel rays, ray/triangle barycentrics, inside test, the
//     integer depth key, the normal-flip decision, the nearest-texel index, and the early-exit
//     tests of the barycentric Jacobian.  Every float operation follows the reference's
//     operand order, so triangle IDs and barycentrics are reproducible bit for bit.
//     Two deliberate replacements: `normalize` uses v * (1/sqrtf(v.v)) where the reference
//     uses the approximate rsqrtf (cpp/thirdParty/Shared/cutil/inc/cutil_math.h:1184-1188),
//     and float->int conversion saturates (NaN->0) as the GPU conversion instruction does.
//   * Visibility ties: the reference lets every triangle whose depth key equals the minimum
//     write its fields without synchronisation (cpp/src/Renderer/CUDABasedRasterization.cu:295-404).
//     Here the lowest face ID among equal keys wins (64-bit (key<<32|id) minimum).
//   * Shading, the normal map and every gradient are evaluated in FP64 from the binary32
//     inputs and the binary32 discrete decisions above.
//
// Parity pins (tests/test_oracle_*.py): closed-form single-triangle renders, SH basis values
// at known normals, brute-force visibility on tiny scenes, camera-inverse identities, and
// central finite differences of this oracle's own forward for every gradient branch.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;
int fail(const std::string& m) { g_err = m; return 1; }

enum Albedo { kVertexColor = 0, kTextured = 1, kNormal = 2, kLighting = 3, kForegroundMask = 4 };
enum Shading { kShaded = 0, kShadeless = 1 };

int g_threads = 0;
int num_threads() {
  if (g_threads > 0) return g_threads;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

// ------------------------------------------------------------------------------------------
// binary32 vector algebra, operand order as cutil_math.h:881,1123,1166,1295
// ------------------------------------------------------------------------------------------
struct f3 { float x, y, z; };
inline f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
inline f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
inline f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline f3 normalize(f3 v) { float inv = 1.0f / std::sqrt(dot(v, v)); return v * inv; }
inline float length(f3 v) { return std::sqrt(dot(v, v)); }
inline f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

inline int32_t sat_i32(float z) {
  if (!(z == z)) return 0;
  if (z >= 2147483648.0f) return INT32_MAX;
  if (z <= -2147483648.0f) return INT32_MIN;
  return (int32_t)z;
}

inline uint64_t pack_key(int32_t zq, int32_t face) {
  return ((uint64_t)((uint32_t)zq ^ 0x80000000u) << 32) | (uint64_t)(uint32_t)face;
}

// ------------------------------------------------------------------------------------------
// FP64 vector algebra for shading and gradients
// ------------------------------------------------------------------------------------------
struct d3 { double x, y, z; };
inline d3 dk(double x, double y, double z) { d3 r; r.x = x; r.y = y; r.z = z; return r; }
inline d3 D(f3 a) { return dk(a.x, a.y, a.z); }
inline d3 operator+(d3 a, d3 b) { return dk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline d3 operator-(d3 a, d3 b) { return dk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline d3 operator-(d3 a) { return dk(-a.x, -a.y, -a.z); }
inline d3 operator*(d3 a, double s) { return dk(a.x * s, a.y * s, a.z * s); }
inline d3 operator*(double s, d3 a) { return dk(s * a.x, s * a.y, s * a.z); }
inline d3 operator/(d3 a, double s) { return dk(a.x / s, a.y / s, a.z / s); }
inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline d3 cross(d3 a, d3 b) { return dk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline double comp(d3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
inline d3 unit(int i, double s) { return dk(i == 0 ? s : 0.0, i == 1 ? s : 0.0, i == 2 ? s : 0.0); }

// ------------------------------------------------------------------------------------------
// Camera set-up: float4x4 cofactor inverse (cpp/src/Utils/float4x4.h:160-285) and product
// (float4x4.h:78-101), driven by initializeCamerasDevice (CUDABasedRasterization.cu:23-67).
// kCofactor[k] lists the six signed triple products of inverse entry k in the evaluation
// order of the reference expression; entries are row-major.
// ------------------------------------------------------------------------------------------
struct Term { int8_t s, a, b, c; };
const Term kCofactor[16][6] = {
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

void inverse4(const float* e, float* out) {
  float inv[16];
  for (int k = 0; k < 16; ++k) {
    const Term* t = kCofactor[k];
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
  for (int k = 0; k < 16; ++k) out[k] = inv[k] * detr;
}

void matmul4(const float* a, const float* b, float* out) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c)
      out[r * 4 + c] = a[r * 4 + 0] * b[0 * 4 + c] + a[r * 4 + 1] * b[1 * 4 + c] +
                       a[r * 4 + 2] * b[2 * 4 + c] + a[r * 4 + 3] * b[3 * 4 + c];
}

struct Camera {
  float E[12];     // world->camera [R|t], row-major 3x4
  float K[9];      // intrinsics, row-major 3x3
  float invE[16];  // (E4)^-1
  float invP[16];  // (K4 E4)^-1
  f3 o;            // ray origin: column 3 of invE divided by its w (CameraUtil.h:251-258)
};

Camera make_camera(const float* extr12, const float* intr9) {
  Camera cam;
  std::memcpy(cam.E, extr12, sizeof(cam.E));
  std::memcpy(cam.K, intr9, sizeof(cam.K));
  float K4[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  float E4[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) K4[r * 4 + c] = intr9[r * 3 + c];
    K4[r * 4 + 3] = 0.f;
    for (int c = 0; c < 4; ++c) E4[r * 4 + c] = extr12[r * 4 + c];
  }
  inverse4(E4, cam.invE);
  float KE[16];
  matmul4(K4, E4, KE);
  inverse4(KE, cam.invP);
  float w = cam.invE[15];
  cam.o = mk(cam.invE[3] / w, cam.invE[7] / w, cam.invE[11] / w);
  return cam;
}

// getRayCuda2 + backprojectPixelCuda(p, invProj) (CameraUtil.h:223-236, 251-258)
f3 pixel_ray(const Camera& cam, float px, float py) {
  const float z = 1000.f;
  float tx = px * z, ty = py * z, tz = z;
  const float* P = cam.invP;
  f3 t = mk(P[0] * tx + P[1] * ty + P[2] * tz + P[3] * 1.f,
            P[4] * tx + P[5] * ty + P[6] * tz + P[7] * 1.f,
            P[8] * tx + P[9] * ty + P[10] * tz + P[11] * 1.f);
  return normalize(t - cam.o);
}

// getCamSpacePoint + projectPointFloat3 (CameraUtil.h:141-186)
f3 project(const Camera& cam, f3 p) {
  const float* E = cam.E;
  f3 c = mk(E[0] * p.x + E[1] * p.y + E[2] * p.z + E[3] * 1.f,
            E[4] * p.x + E[5] * p.y + E[6] * p.z + E[7] * 1.f,
            E[8] * p.x + E[9] * p.y + E[10] * p.z + E[11] * 1.f);
  const float* K = cam.K;
  float x = c.x * K[0] + c.y * K[1] + c.z * K[2];
  float y = c.x * K[3] + c.y * K[4] + c.z * K[5];
  float z = c.x * K[6] + c.y * K[7] + c.z * K[8];
  if (!(z > 0.0000001f)) z = 0.00001f;
  return mk(x / z, y / z, z);
}

// rayTriangleIntersect (RendererUtil.h:26-101) + uv2barycentric (RendererUtil.h:108-128).
// Vertices are scaled by 1/1000 inside; returns (a,b,c) or (-1,-1,-1).
f3 ray_barycentric(f3 o, f3 d, f3 v0, f3 v1, f3 v2) {
  v0 = v0 / 1000.f; v1 = v1 / 1000.f; v2 = v2 / 1000.f; o = o / 1000.f;
  f3 N = cross(v1 - v0, v2 - v0);
  float ndr = dot(d, N);
  f3 miss = mk(-1.f, -1.f, -1.f);
  if (std::fabs(ndr) < 0.0000001f) return miss;
  float t = (dot(v0, N) - dot(o, N)) / ndr;
  if (t < 0) return miss;
  f3 P = o + t * d;
  if (dot(N, cross(v1 - v0, P - v0)) < 0) return miss;
  float a = dot(N, cross(v2 - v1, P - v1));
  if (a < 0) return miss;
  float b = dot(N, cross(v0 - v2, P - v2));
  if (b < 0) return miss;
  float den = dot(N, N);
  a /= den; b /= den;
  return mk(a, b, 1.f - a - b);
}

inline bool inside(f3 abc) {
  return (abc.x >= -0.001f) && (abc.y >= -0.001f) && (abc.z >= -0.001f) &&
         (abc.x <= 1.001f) && (abc.y <= 1.001f) && (abc.z <= 1.001f);
}

// Depth key: CUDABasedRasterization.cu:247-250 (two roundings: 1/sum, then *1e4)
inline int32_t depth_key(f3 abc, float z0, float z1, float z2) {
  float z = 1.f / (abc.x / z0 + abc.y / z1 + abc.z / z2);
  z *= 10000.f;
  return sat_i32(z);
}

// Normal-flip decision in binary32 (CUDABasedRasterization.cu:308-319; ...Grad.cu:231-240)
inline bool flip_decision(f3 vn0, f3 vn1, f3 vn2, f3 abc, f3 d) {
  f3 n = vn0 * abc.x + vn1 * abc.y + vn2 * abc.z;
  n = n / length(n);
  return dot(n, d) > 0.f;
}

// Texture coordinate in binary32 (CUDABasedRasterization.cu:326-345)
struct TexCoord { float U0, V0; int LU, LV; };
inline TexCoord tex_coord(const float* uv6, f3 abc, int texW, int texH) {
  float t0x = uv6[0], t0y = 1.f - uv6[1];
  float t1x = uv6[2], t1y = 1.f - uv6[3];
  float t2x = uv6[4], t2y = 1.f - uv6[5];
  float fx = t0x * abc.x + t1x * abc.y + t2x * abc.z;
  float fy = t0y * abc.x + t1y * abc.y + t2y * abc.z;
  fx = fx * (float)texW;
  fy = fy * (float)texH;
  fx = std::fmax(fx, 0.f); fx = std::fmin(fx, (float)(texW - 1));
  fy = std::fmax(fy, 0.f); fy = std::fmin(fy, (float)(texH - 1));
  TexCoord t;
  t.U0 = fx; t.V0 = fy;
  t.LU = sat_i32(fx - 0.5f);
  t.LV = sat_i32(fy - 0.5f);
  return t;
}

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// SH basis, 9 unnormalised terms (RendererUtil.h:135-214, 351-364)
inline void sh_basis(d3 n, double Y[9]) {
  Y[0] = 1.0; Y[1] = n.y; Y[2] = n.z; Y[3] = n.x; Y[4] = n.x * n.y; Y[5] = n.z * n.y;
  Y[6] = 3.0 * n.z * n.z - 1.0; Y[7] = n.x * n.z; Y[8] = n.x * n.x - n.y * n.y;
}
inline d3 sh_illum(d3 n, const float* sh27) {
  double Y[9]; sh_basis(n, Y);
  double L[3];
  for (int ch = 0; ch < 3; ++ch) {
    double s = 0.0;
    for (int k = 0; k < 9; ++k) s += (double)sh27[ch * 9 + k] * Y[k];
    L[ch] = s;
  }
  return dk(L[0], L[1], L[2]);
}

// ------------------------------------------------------------------------------------------
// Mesh topology: vertex->face CSR (CUDABasedRasterization.cpp:125-154).  Faces appear in
// increasing index order per vertex, each face once per vertex.
// ------------------------------------------------------------------------------------------
struct Topology {
  int N = 0, F = 0;
  std::vector<int32_t> faces;   // 3F
  std::vector<float> uv;        // 6F
  std::vector<int32_t> start, count, list;
};

void build_csr(Topology& t) {
  t.start.assign(t.N, 0); t.count.assign(t.N, 0); t.list.clear();
  for (int f = 0; f < t.F; ++f) {
    int a = t.faces[3 * f], b = t.faces[3 * f + 1], c = t.faces[3 * f + 2];
    t.count[a]++;
    if (b != a) t.count[b]++;
    if (c != a && c != b) t.count[c]++;
  }
  int run = 0;
  for (int v = 0; v < t.N; ++v) { t.start[v] = run; run += t.count[v]; }
  t.list.assign(run, 0);
  std::vector<int> fill(t.N, 0);
  for (int f = 0; f < t.F; ++f) {
    int a = t.faces[3 * f], b = t.faces[3 * f + 1], c = t.faces[3 * f + 2];
    t.list[t.start[a] + fill[a]++] = f;
    if (b != a) t.list[t.start[b] + fill[b]++] = f;
    if (c != a && c != b) t.list[t.start[c] + fill[c]++] = f;
  }
}

struct Handle {
  Topology topo;
  int C = 0, W = 0, H = 0;
  int albedo = 0, shading = 0, image_fs = 1, texture_fs = 1, normal_map = 0;
  // texel -> (face, a, b, c) table for the normal map (CUDABasedRasterization.cpp:241-298)
  int uvW = -1, uvH = -1;
  std::vector<int32_t> uv_face;
  std::vector<float> uv_abc;
};

// Host UV rasterisation (CUDABasedRasterization.cpp:158-298): the texel ray starts at
// (x+.5, y+.5, 1) and points along -z; later faces overwrite earlier ones.
void build_uv_table(Handle& h, int texW, int texH) {
  if (h.uvW == texW && h.uvH == texH) return;
  h.uvW = texW; h.uvH = texH;
  h.uv_face.assign((size_t)texW * texH, 0);
  h.uv_abc.assign((size_t)texW * texH * 3, 0.f);
  const float* uv = h.topo.uv.data();
  for (int f = 0; f < h.topo.F; ++f) {
    f3 t0 = mk((float)texW * uv[f * 6 + 0], (float)texH * (1.f - uv[f * 6 + 1]), 0.f);
    f3 t1 = mk((float)texW * uv[f * 6 + 2], (float)texH * (1.f - uv[f * 6 + 3]), 0.f);
    f3 t2 = mk((float)texW * uv[f * 6 + 4], (float)texH * (1.f - uv[f * 6 + 5]), 0.f);
    int xMin = sat_i32(std::fmax(std::fmin(t0.x, std::fmin(t1.x, t2.x)) - 2.f, 0.f));
    int xMax = sat_i32(std::fmin(std::fmax(t0.x, std::fmax(t1.x, t2.x)) + 2.f, (float)texW));
    int yMin = sat_i32(std::fmax(std::fmin(t0.y, std::fmin(t1.y, t2.y)) - 2.f, 0.f));
    int yMax = sat_i32(std::fmin(std::fmax(t0.y, std::fmax(t1.y, t2.y)) + 2.f, (float)texH));
    for (int x = xMin; x < xMax; ++x)
      for (int y = yMin; y < yMax; ++y) {
        f3 abc = ray_barycentric(mk(x + 0.5f, y + 0.5f, 1.f), mk(0.f, 0.f, -1.f), t0, t1, t2);
        if (abc.x != -1.f && abc.y != -1.f && abc.z != -1.f) {
          size_t i = (size_t)y * texW + x;
          h.uv_face[i] = f;
          h.uv_abc[3 * i] = abc.x; h.uv_abc[3 * i + 1] = abc.y; h.uv_abc[3 * i + 2] = abc.z;
        }
      }
  }
}

template <class Fn>
void parallel_tasks(int ntasks, Fn fn) {
  int T = std::min(num_threads(), std::max(ntasks, 1));
  if (T <= 1) { for (int i = 0; i < ntasks; ++i) fn(i, 0); return; }
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t]() { for (int i = t; i < ntasks; i += T) fn(i, t); });
  for (auto& x : th) x.join();
}

struct Box { int x0, y0, x1, y1; };

// projectFacesDevice (CUDABasedRasterization.cu:184-208): +-0.5 px, clamp, truncate.
Box face_box(f3 a, f3 b, f3 c, int W, int H) {
  Box r;
  r.x0 = sat_i32(std::fmax(std::fmin(a.x, std::fmin(b.x, c.x)) - 0.5f, 0.f));
  r.y0 = sat_i32(std::fmax(std::fmin(a.y, std::fmin(b.y, c.y)) - 0.5f, 0.f));
  r.x1 = sat_i32(std::fmin(std::fmax(a.x, std::fmax(b.x, c.x)) + 0.5f, (float)(W - 1)));
  r.y1 = sat_i32(std::fmin(std::fmax(a.y, std::fmax(b.y, c.y)) + 0.5f, (float)(H - 1)));
  return r;
}

// ------------------------------------------------------------------------------------------
// Forward: renderBuffersGPU (CUDABasedRasterization.cu:449-473) for one batch element.
// ------------------------------------------------------------------------------------------
void forward_one(Handle& h, int texH, int texW, const float* vpos, const float* vcol,
                 const float* tex, const float* sh, const float* extr, const float* intr,
                 float* bary, int32_t* face, double* render, float* vnormal, double* nmap) {
  const Topology& T = h.topo;
  const int C = h.C, W = h.W, H = h.H, N = T.N, F = T.F;
  const size_t P = (size_t)W * H;
  std::vector<Camera> cams(C);
  for (int c = 0; c < C; ++c) cams[c] = make_camera(extr + 12 * c, intr + 9 * c);

  // initializeDevice (.cu:74-91)
  for (size_t i = 0; i < P * C; ++i) {
    face[i] = -1; bary[2 * i] = 0.f; bary[2 * i + 1] = 0.f;
    render[3 * i] = 0.0; render[3 * i + 1] = 1.0; render[3 * i + 2] = 0.0;
  }

  // projectVerticesDevice (.cu:98-115)
  std::vector<f3> proj((size_t)C * N);
  for (int c = 0; c < C; ++c)
    for (int v = 0; v < N; ++v) proj[(size_t)c * N + v] = project(cams[c], ld3(vpos + 3 * v));

  // renderFaceNormalDevice (.cu:122-141) + renderVertexNormalDevice (.cu:148-174)
  std::vector<f3> fn(F), vn(N, mk(0.f, 0.f, 0.f));
  for (int f = 0; f < F; ++f) {
    f3 v0 = ld3(vpos + 3 * T.faces[3 * f]), v1 = ld3(vpos + 3 * T.faces[3 * f + 1]),
       v2 = ld3(vpos + 3 * T.faces[3 * f + 2]);
    fn[f] = cross(v1 - v0, v2 - v0);
  }
  for (int v = 0; v < N; ++v) {
    int s = T.start[v], n = T.count[v];
    if (n == 0) continue;
    f3 acc = fn[T.list[s]];
    for (int k = 1; k < n; ++k) {
      f3 q = fn[T.list[s + k]];
      acc.x = acc.x + q.x; acc.y = acc.y + q.y; acc.z = acc.z + q.z;
    }
    vn[v] = acc;
  }
  for (int c = 0; c < C; ++c)
    for (int v = 0; v < N; ++v) {
      vnormal[((size_t)c * N + v) * 3 + 0] = vn[v].x;
      vnormal[((size_t)c * N + v) * 3 + 1] = vn[v].y;
      vnormal[((size_t)c * N + v) * 3 + 2] = vn[v].z;
    }

  if (h.normal_map) {
    // renderNormalMapDevice (.cu:415-445), camera-0 normals
    for (size_t i = 0; i < (size_t)texW * texH; ++i) {
      int f = h.uv_face[i];
      const float* abc = &h.uv_abc[3 * i];
      d3 n = D(vn[T.faces[3 * f]]) * (double)abc[0] + D(vn[T.faces[3 * f + 1]]) * (double)abc[1] +
             D(vn[T.faces[3 * f + 2]]) * (double)abc[2];
      double len = std::sqrt(dot(n, n));
      if (len != 0.0) n = n / len;
      nmap[3 * i + 0] = (n.x + 1.0) / 2.0;
      nmap[3 * i + 1] = (n.y + 1.0) / 2.0;
      nmap[3 * i + 2] = (n.z + 1.0) / 2.0;
    }
    return;
  }

  std::vector<Box> boxes((size_t)C * F);
  for (int c = 0; c < C; ++c)
    for (int f = 0; f < F; ++f) {
      const f3* pc = &proj[(size_t)c * N];
      boxes[(size_t)c * F + f] =
          face_box(pc[T.faces[3 * f]], pc[T.faces[3 * f + 1]], pc[T.faces[3 * f + 2]], W, H);
    }

  std::vector<uint64_t> key(P * C, ~0ull);
  const int bands = std::max(1, std::min(H, 4 * num_threads()));
  auto band_rows = [&](int band, int& r0, int& r1) {
    r0 = (int)((int64_t)H * band / bands);
    r1 = (int)((int64_t)H * (band + 1) / bands) - 1;
  };

  // Pass 1: renderDepthBufferDevice (.cu:215-255), per (view, face) bbox loop.
  parallel_tasks(C * bands, [&](int task, int) {
    int c = task / bands, band = task % bands, r0, r1;
    band_rows(band, r0, r1);
    const Camera& cam = cams[c];
    const f3* pc = &proj[(size_t)c * N];
    uint64_t* kc = &key[(size_t)c * P];
    for (int f = 0; f < F; ++f) {
      const Box& bx = boxes[(size_t)c * F + f];
      int y0 = std::max(bx.y0, r0), y1 = std::min(bx.y1, r1);
      if (y0 > y1 || bx.x0 > bx.x1) continue;
      int i0 = T.faces[3 * f], i1 = T.faces[3 * f + 1], i2 = T.faces[3 * f + 2];
      f3 v0 = ld3(vpos + 3 * i0), v1 = ld3(vpos + 3 * i1), v2 = ld3(vpos + 3 * i2);
      for (int u = bx.x0; u <= bx.x1; ++u)
        for (int v = y0; v <= y1; ++v) {
          f3 d = pixel_ray(cam, u + 0.5f, v + 0.5f);
          f3 abc = ray_barycentric(cam.o, d, v0, v1, v2);
          if (!inside(abc)) continue;
          uint64_t k = pack_key(depth_key(abc, pc[i0].z, pc[i1].z, pc[i2].z), f);
          uint64_t& slot = kc[(size_t)v * W + u];
          if (k < slot) slot = k;
        }
    }
  });

  // Pass 2: renderBuffersDevice (.cu:262-408): the winning triangle writes ID, bary, colour.
  parallel_tasks(C * bands, [&](int task, int) {
    int c = task / bands, band = task % bands, r0, r1;
    band_rows(band, r0, r1);
    const Camera& cam = cams[c];
    const f3* pc = &proj[(size_t)c * N];
    const float* shc = sh + 27 * c;
    for (int f = 0; f < F; ++f) {
      const Box& bx = boxes[(size_t)c * F + f];
      int y0 = std::max(bx.y0, r0), y1 = std::min(bx.y1, r1);
      if (y0 > y1 || bx.x0 > bx.x1) continue;
      int i0 = T.faces[3 * f], i1 = T.faces[3 * f + 1], i2 = T.faces[3 * f + 2];
      f3 v0 = ld3(vpos + 3 * i0), v1 = ld3(vpos + 3 * i1), v2 = ld3(vpos + 3 * i2);
      for (int u = bx.x0; u <= bx.x1; ++u)
        for (int v = y0; v <= y1; ++v) {
          size_t pix = (size_t)c * P + (size_t)v * W + u;
          f3 d = pixel_ray(cam, u + 0.5f, v + 0.5f);
          f3 abc = ray_barycentric(cam.o, d, v0, v1, v2);
          if (!inside(abc)) continue;
          uint64_t k = pack_key(depth_key(abc, pc[i0].z, pc[i1].z, pc[i2].z), f);
          if (k != key[pix]) continue;
          face[pix] = f;
          bary[2 * pix] = abc.x;
          bary[2 * pix + 1] = abc.y;
          bool flip = flip_decision(vn[i0], vn[i1], vn[i2], abc, d);
          d3 n = D(vn[i0]) * (double)abc.x + D(vn[i1]) * (double)abc.y + D(vn[i2]) * (double)abc.z;
          n = n / std::sqrt(dot(n, n));
          if (flip) n = -n;
          d3 col = dk(0, 0, 0);
          if (h.albedo == kTextured) {
            TexCoord tc = tex_coord(&T.uv[6 * f], abc, texW, texH);
            int row = clampi(tc.LV, 0, texH - 1), cl = clampi(tc.LU, 0, texW - 1);
            const float* t = tex + 3 * ((size_t)texW * row + cl);
            col = dk(t[0], t[1], t[2]);
          } else if (h.albedo == kVertexColor) {
            col = D(ld3(vcol + 3 * i0)) * (double)abc.x + D(ld3(vcol + 3 * i1)) * (double)abc.y +
                  D(ld3(vcol + 3 * i2)) * (double)abc.z;
          } else if (h.albedo == kNormal) {
            col = dk((1.0 + n.x) / 2.0, (1.0 + n.y) / 2.0, (1.0 + n.z) / 2.0);
          } else {
            col = dk(1, 1, 1);
          }
          if ((h.shading == kShaded && h.albedo != kNormal) || h.albedo == kLighting) {
            d3 L = sh_illum(n, shc);
            col = dk(col.x * L.x, col.y * L.y, col.z * L.z);
          }
          render[3 * pix] = col.x; render[3 * pix + 1] = col.y; render[3 * pix + 2] = col.z;
        }
    }
  });
}

// ------------------------------------------------------------------------------------------
// Backward helpers (FP64)
// ------------------------------------------------------------------------------------------
// imageGradient (RendererUtil.h:566-620) at integer pixel coordinates.
void image_gradient(const float* img, int px, int py, int W, int H, int fs, d3& du, d3& dv) {
  du = dk(0, 0, 0); dv = dk(0, 0, 0);
  if (!(px >= fs + 1 && py >= fs + 1 && px < W - (fs + 1) && py < H - (fs + 1))) return;
  double norm = 0.0;
  for (int y = -fs; y <= fs; ++y)
    for (int x = -fs; x <= fs; ++x) {
      const float* I = img + 3 * ((size_t)(py + y) * W + (px + x));
      double den = (double)(x * x + y * y), gu = 0.0, gv = 0.0;
      if (den != 0.0) { gu = x / den; gv = y / den; }
      du = du + dk(I[0], I[1], I[2]) * gu;
      dv = dv + dk(I[0], I[1], I[2]) * gv;
      norm += std::fabs(gu);
    }
  du = du / norm; dv = dv / norm;
}

// getJProjection (RendererUtil.h:275-332): d(pixel)/d(world point) = dDivide * K * R.
void projection_jacobian(const float* K, const float* E, d3 X, double J[2][3]) {
  double KE[3][4];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c)
      KE[r][c] = (double)K[r * 3 + 0] * E[0 * 4 + c] + (double)K[r * 3 + 1] * E[1 * 4 + c] +
                 (double)K[r * 3 + 2] * E[2 * 4 + c];
  double P[3];
  for (int r = 0; r < 3; ++r) P[r] = KE[r][0] * X.x + KE[r][1] * X.y + KE[r][2] * X.z + KE[r][3];
  for (int r = 0; r < 2; ++r) for (int c = 0; c < 3; ++c) J[r][c] = 0.0;
  if (!(std::fabs(P[2]) > 0.0001)) return;
  double dd[2][3] = {{1.0 / P[2], 0.0, -P[0] / (P[2] * P[2])}, {0.0, 1.0 / P[2], -P[1] / (P[2] * P[2])}};
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c) J[r][c] = dd[r][0] * KE[0][c] + dd[r][1] * KE[1][c] + dd[r][2] * KE[2][c];
}

// dJBCDVerpos (RendererUtil.h:670-861): d(a,b,c)/d(v0,v1,v2) holding the ray fixed.
// The early exit is decided in binary32 from the binary32 ray and vertices.
void bary_jacobian(f3 of, f3 df, f3 v0f, f3 v1f, f3 v2f, double J[3][9]) {
  for (int r = 0; r < 3; ++r) for (int k = 0; k < 9; ++k) J[r][k] = 0.0;
  {
    f3 Nf = cross(v1f - v0f, v2f - v0f);
    float den = dot(Nf, Nf);
    if (std::fabs(dot(normalize(df), normalize(Nf))) < 0.001f || std::fabs(den * den) < 0.001f) return;
  }
  d3 o = D(of), d = D(df), v0 = D(v0f), v1 = D(v1f), v2 = D(v2f);
  d3 N = cross(v1 - v0, v2 - v0);
  double den = dot(N, N), nd = dot(N, d);
  double t = (dot(v0, N) - dot(o, N)) / nd;
  d3 P = o + t * d;
  d3 E[2] = {v2 - v1, v0 - v2};
  d3 vp[2] = {P - v1, P - v2};
  d3 Cc[2] = {cross(E[0], vp[0]), cross(E[1], vp[1])};
  d3 dN[9];
  for (int k = 0; k < 3; ++k) {
    d3 ek = unit(k, 1.0), nek = unit(k, -1.0);
    dN[k] = cross(v1 - v0, nek) + cross(nek, v2 - v0);   // getJ_vi
    dN[3 + k] = cross(ek, v2 - v0);                      // getJ_vj
    dN[6 + k] = cross(v1 - v0, ek);                      // getJ_vk
  }
  for (int var = 0; var < 9; ++var) {
    int vtx = var / 3, ax = var % 3;
    d3 dV[3] = {dk(0, 0, 0), dk(0, 0, 0), dk(0, 0, 0)};
    dV[vtx] = unit(ax, 1.0);
    d3 dE[2] = {dV[2] - dV[1], dV[0] - dV[2]};
    double dt = (-1.0 / (nd * nd)) * dot(dN[var], d) * dot(v0 - o, N) +
                (1.0 / nd) * (dot(dV[0], N) + dot(v0 - o, dN[var]));
    for (int r = 0; r < 2; ++r) {
      double tmp = dot(N, Cc[r]);
      double dtmp = dot(dN[var], Cc[r]) + dot(N, cross(dE[r], vp[r]) + cross(E[r], dt * d - dV[r + 1]));
      J[r][var] = dtmp / den + tmp * (-1.0 / (den * den)) * (2.0 * dot(N, dN[var]));
    }
    J[2][var] = -J[0][var] - J[1][var];
  }
}

struct GradAcc {
  std::vector<double> pos, col, tex, sh;
  void init(int N, int texH, int texW, int C) {
    pos.assign((size_t)N * 3, 0.0); col.assign((size_t)N * 3, 0.0);
    tex.assign((size_t)texH * texW * 3, 0.0); sh.assign((size_t)C * 27, 0.0);
  }
};

inline void add3(std::vector<double>& g, int v, d3 x) {
  g[3 * (size_t)v] += x.x; g[3 * (size_t)v + 1] += x.y; g[3 * (size_t)v + 2] += x.z;
}

// ------------------------------------------------------------------------------------------
// Backward: renderBuffersGradDevice (CUDABasedRasterizationGrad.cu:114-617) for one pixel.
// ------------------------------------------------------------------------------------------
void backward_pixel(const Handle& h, const Camera& cam, int c, int idh, int idw, int texH, int texW,
                    const float* g_render, const float* g_target, const float* target,
                    const float* vpos, const float* vcol, const float* tex, const float* sh,
                    const float* vnormal, const float* bary, const int32_t* face, GradAcc& acc) {
  const Topology& T = h.topo;
  const int W = h.W, H = h.H, N = T.N;
  const size_t idx = ((size_t)c * H + idh) * W + idw;
  const int idf = face[idx];
  if (idf == -1) return;
  const bool shaded = h.shading == kShaded;

  f3 df = pixel_ray(cam, idw + 0.5f, idh + 0.5f);
  f3 of = cam.o;
  float ba = bary[2 * idx], bb = bary[2 * idx + 1];
  f3 bccf = mk(ba, bb, 1.f - ba - bb);
  d3 bcc = dk(ba, bb, 1.0 - (double)ba - (double)bb);
  double w3[3] = {bcc.x, bcc.y, bcc.z};
  int vid[3] = {T.faces[3 * idf], T.faces[3 * idf + 1], T.faces[3 * idf + 2]};
  const float* shc = sh + 27 * c;
  f3 vnf[3];
  d3 vp[3], vc[3], vnd[3];
  for (int r = 0; r < 3; ++r) {
    vp[r] = D(ld3(vpos + 3 * vid[r]));
    vc[r] = D(ld3(vcol + 3 * vid[r]));
    vnf[r] = ld3(vnormal + 3 * ((size_t)c * N + vid[r]));
    vnd[r] = D(vnf[r]);
  }
  d3 frag = vp[0] * bcc.x + vp[1] * bcc.y + vp[2] * bcc.z;
  d3 un = vnd[0] * bcc.x + vnd[1] * bcc.y + vnd[2] * bcc.z;
  double unl = std::sqrt(dot(un, un));
  d3 n = un / unl;
  bool flipped = flip_decision(vnf[0], vnf[1], vnf[2], bccf, df);
  if (flipped) n = -n;

  TexCoord tc;
  tc.U0 = tc.V0 = 0.f; tc.LU = tc.LV = 0;
  if (h.albedo == kTextured) tc = tex_coord(&T.uv[6 * idf], bccf, texW, texH);

  d3 L = sh_illum(n, shc);
  d3 jcoal = shaded ? L : dk(1, 1, 1);   // diagonal of JCoAl
  d3 alb = dk(0, 0, 0);                  // diagonal of JCoLi
  if (h.albedo == kVertexColor) {
    alb = vc[0] * bcc.x + vc[1] * bcc.y + vc[2] * bcc.z;
  } else if (h.albedo == kTextured) {
    // CUDABasedRasterizationGrad.cu:282-312 (not the forward's nearest texel)
    double U0 = tc.U0, V0 = tc.V0;
    double LU = tc.LU + 0.5, HU = tc.LU + 1.5, LV = tc.LV + 0.5, HV = tc.LV + 1.5;
    int lu = clampi(tc.LU, 0, texW - 1), hu = clampi(tc.LU + 1, 0, texW - 1);
    int lv = clampi(tc.LV, 0, texH - 1), hv = clampi(tc.LV + 1, 0, texH - 1);
    auto texel = [&](int row, int col) {
      const float* t = tex + 3 * ((size_t)texW * row + col);
      return dk(t[0], t[1], t[2]);
    };
    alb = (V0 - LV) * (((U0 - LU) * texel(lv, lu)) + ((HU - U0) * texel(lv, hu))) +
          (HV - V0) * (((U0 - LU) * texel(hv, lu)) + ((HU - U0) * texel(hv, hu)));
  }

  const float* gr = g_render + 3 * idx;
  d3 g = dk(gr[0], gr[1], gr[2]);

  // Albedo gradients (Grad.cu:327-395)
  if (h.albedo == kVertexColor) {
    d3 ga = dk(g.x * jcoal.x, g.y * jcoal.y, g.z * jcoal.z);
    for (int r = 0; r < 3; ++r) add3(acc.col, vid[r], ga * w3[r]);
  } else if (h.albedo == kTextured) {
    if (!flipped) {
      d3 ga = dk(g.x * jcoal.x, g.y * jcoal.y, g.z * jcoal.z);
      int row = tc.LV, col = tc.LU;
      if (row >= 0 && row < texH && col >= 0 && col < texW) add3(acc.tex, row * texW + col, ga);
    }
  }

  // Lighting gradients (Grad.cu:402-436)
  d3 gl = dk(g.x * alb.x, g.y * alb.y, g.z * alb.z);   // g * JCoLi
  if (shaded) {
    double Y[9]; sh_basis(n, Y);
    double glc[3] = {gl.x, gl.y, gl.z};
    for (int ch = 0; ch < 3; ++ch)
      for (int k = 0; k < 9; ++k) acc.sh[(size_t)c * 27 + ch * 9 + k] += glc[ch] * Y[k];
  }

  // Position gradients through the normal (Grad.cu:458-525)
  d3 gu = dk(0, 0, 0);
  if (shaded) {
    double JLiNo[3][3];
    for (int i = 0; i < 3; ++i) {
      const float* s = shc + 9 * i;
      JLiNo[i][0] = s[3] + s[4] * n.y + s[7] * n.z + s[8] * 2.0 * n.x;
      JLiNo[i][1] = s[1] + s[4] * n.x + s[5] * n.z + s[8] * -2.0 * n.y;
      JLiNo[i][2] = s[2] + s[5] * n.y + s[6] * 6.0 * n.z + s[7] * n.x;
    }
    double glc[3] = {gl.x, gl.y, gl.z}, gn[3];
    for (int k = 0; k < 3; ++k) gn[k] = glc[0] * JLiNo[0][k] + glc[1] * JLiNo[1][k] + glc[2] * JLiNo[2][k];
    // JNoNu (RendererUtil.h:398-415) from the unflipped, unnormalised normal
    double p2 = unl * unl, p3 = p2 * unl, u[3] = {un.x, un.y, un.z}, gv[3];
    for (int k = 0; k < 3; ++k) {
      double s = 0.0;
      for (int j = 0; j < 3; ++j) s += gn[j] * ((j == k ? p2 : 0.0) - u[j] * u[k]) / p3;
      gv[k] = s;
    }
    gu = dk(gv[0], gv[1], gv[2]);
    double JB[3][9];
    bary_jacobian(of, df, ld3(vpos + 3 * vid[0]), ld3(vpos + 3 * vid[1]), ld3(vpos + 3 * vid[2]), JB);
    double gb[3] = {dot(gu, vnd[0]), dot(gu, vnd[1]), dot(gu, vnd[2])};   // gu * JNoBc
    double gvp[9];
    for (int k = 0; k < 9; ++k) gvp[k] = gb[0] * JB[0][k] + gb[1] * JB[1][k] + gb[2] * JB[2][k];
    for (int r = 0; r < 3; ++r) add3(acc.pos, vid[r], dk(gvp[3 * r], gvp[3 * r + 1], gvp[3 * r + 2]));
  }

  // Model-to-data term (Grad.cu:531-555)
  {
    const float* gt = g_target + 3 * idx;
    d3 gT = dk(gt[0], gt[1], gt[2]);
    d3 du, dv;
    image_gradient(target + 3 * (size_t)c * W * H, idw, idh, W, H, h.image_fs, du, dv);
    double q0 = dot(gT, du), q1 = dot(gT, dv);
    double JP[2][3];
    projection_jacobian(cam.K, cam.E, frag, JP);
    d3 gx = dk(q0 * JP[0][0] + q1 * JP[1][0], q0 * JP[0][1] + q1 * JP[1][1], q0 * JP[0][2] + q1 * JP[1][2]);
    for (int r = 0; r < 3; ++r) add3(acc.pos, vid[r], gx * w3[r]);
  }

  // Normal -> position through the vertex-face adjacency (Grad.cu:559-615)
  if (shaded) {
    for (int i = 0; i < 3; ++i) {
      d3 gni = gu * w3[i];
      int idv = vid[i];
      for (int j = T.start[idv]; j < T.start[idv] + T.count[idv]; ++j) {
        int fj = T.list[j];
        int a0 = T.faces[3 * fj], a1 = T.faces[3 * fj + 1], a2 = T.faces[3 * fj + 2];
        d3 vi = D(ld3(vpos + 3 * a0)), vj = D(ld3(vpos + 3 * a1)), vk = D(ld3(vpos + 3 * a2));
        // row vector gni times the 3x3 Jacobian whose column k is dN/d(v.e_k)
        d3 gi, gj, gk;
        double ri[3], rj[3], rk[3];
        for (int k = 0; k < 3; ++k) {
          d3 ek = unit(k, 1.0), nek = unit(k, -1.0);
          ri[k] = dot(gni, cross(vj - vi, nek) + cross(nek, vk - vi));
          rj[k] = dot(gni, cross(ek, vk - vi));
          rk[k] = dot(gni, cross(vj - vi, ek));
        }
        gi = dk(ri[0], ri[1], ri[2]); gj = dk(rj[0], rj[1], rj[2]); gk = dk(rk[0], rk[1], rk[2]);
        add3(acc.pos, a0, gi); add3(acc.pos, a1, gj); add3(acc.pos, a2, gk);
      }
    }
  }
}

void backward_one(const Handle& h, int texH, int texW, const float* g_render, const float* g_target,
                  const float* target, const float* vpos, const float* vcol, const float* tex,
                  const float* sh, const float* vnormal, const float* bary, const int32_t* face,
                  const float* extr, const float* intr, double* g_pos, double* g_col, double* g_tex,
                  double* g_sh) {
  const int C = h.C, W = h.W, H = h.H, N = h.topo.N;
  std::vector<Camera> cams(C);
  for (int c = 0; c < C; ++c) cams[c] = make_camera(extr + 12 * c, intr + 9 * c);
  const int nrows = C * H;
  const int T = std::max(1, std::min(num_threads(), nrows));
  std::vector<GradAcc> accs(T);
  for (auto& a : accs) a.init(N, texH, texW, C);
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t]() {
      for (int r = t; r < nrows; r += T) {
        int c = r / H, row = r % H;
        for (int col = 0; col < W; ++col)
          backward_pixel(h, cams[c], c, row, col, texH, texW, g_render, g_target, target, vpos, vcol,
                         tex, sh, vnormal, bary, face, accs[t]);
      }
    });
  for (auto& x : th) x.join();
  auto reduce = [&](std::vector<double> GradAcc::*m, double* out, size_t n) {
    for (size_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int t = 0; t < T; ++t) s += (accs[t].*m)[i];
      out[i] = s;
    }
  };
  reduce(&GradAcc::pos, g_pos, (size_t)N * 3);
  reduce(&GradAcc::col, g_col, (size_t)N * 3);
  reduce(&GradAcc::tex, g_tex, (size_t)texH * texW * 3);
  reduce(&GradAcc::sh, g_sh, (size_t)C * 27);
}

}  // namespace

// ==========================================================================================
// C ABI (host pointers).  Same argument meaning as the HIP library's gvv_* entry points;
// continuous outputs are FP64.
// ==========================================================================================
extern "C" {

const char* gvvo_last_error(void) { return g_err.c_str(); }

int gvvo_set_threads(int n) { g_threads = n; return num_threads(); }

int gvvo_create(const int32_t* faces, int64_t n_face_idx, const float* texcoords, int64_t n_texcoords,
                int32_t n_vertices, int32_t n_cameras, int32_t width, int32_t height,
                int32_t albedo_mode, int32_t shading_mode, int32_t image_filter_size,
                int32_t texture_filter_size, int32_t compute_normal_map, void** handle) {
  if (!handle) return fail("handle pointer is null");
  *handle = nullptr;
  if (n_face_idx <= 0 || n_face_idx % 3 != 0) return fail("faces must hold 3*F > 0 indices");
  int64_t F = n_face_idx / 3;
  if (n_texcoords != 6 * F) return fail("texture_coordinates must hold 6*F values");
  if (n_vertices <= 0) return fail("number_of_vertices must be > 0");
  if (n_cameras <= 0) return fail("number_of_cameras must be > 0");
  if (width <= 0 || height <= 0) return fail("render resolution must be > 0");
  if (albedo_mode < 0 || albedo_mode > 4) return fail("invalid albedo mode");
  if (shading_mode < 0 || shading_mode > 1) return fail("invalid shading mode");
  if (image_filter_size <= 0 || texture_filter_size <= 0) return fail("filter sizes must be > 0");
  for (int64_t i = 0; i < n_face_idx; ++i)
    if (faces[i] < 0 || faces[i] >= n_vertices) return fail("face index out of range");
  Handle* h = new Handle();
  h->topo.N = n_vertices; h->topo.F = (int)F;
  h->topo.faces.assign(faces, faces + n_face_idx);
  h->topo.uv.assign(texcoords, texcoords + n_texcoords);
  build_csr(h->topo);
  h->C = n_cameras; h->W = width; h->H = height;
  h->albedo = albedo_mode;
  h->shading = albedo_mode == kForegroundMask ? kShadeless : shading_mode;
  h->image_fs = image_filter_size; h->texture_fs = texture_filter_size;
  h->normal_map = compute_normal_map ? 1 : 0;
  *handle = h;
  return 0;
}

void gvvo_destroy(void* handle) { delete (Handle*)handle; }

int gvvo_forward(void* handle, int32_t B, int32_t texH, int32_t texW, const float* vpos, const float* vcol,
                 const float* tex, const float* sh, const float* extr, const float* intr, float* bary,
                 int32_t* face, double* render, float* vnormal, double* normal_map) {
  Handle* h = (Handle*)handle;
  if (!h) return fail("null handle");
  if (B <= 0 || texH <= 0 || texW <= 0) return fail("batch and texture sizes must be > 0");
  build_uv_table(*h, texW, texH);
  const size_t N = h->topo.N, C = h->C, P = (size_t)h->W * h->H, T = (size_t)texH * texW;
  for (int b = 0; b < B; ++b)
    forward_one(*h, texH, texW, vpos + b * N * 3, vcol + b * N * 3, tex + b * T * 3, sh + b * C * 27,
                extr + b * C * 12, intr + b * C * 9, bary + b * C * P * 2, face + b * C * P,
                render + b * C * P * 3, vnormal + b * C * N * 3, normal_map ? normal_map + b * T * 3 : nullptr);
  return 0;
}

int gvvo_backward(void* handle, int32_t B, int32_t texH, int32_t texW, const float* g_render,
                  const float* g_target, const float* target, const float* vpos, const float* vcol,
                  const float* tex, const float* sh, const float* vnormal, const float* bary,
                  const int32_t* face, const float* extr, const float* intr, double* g_pos,
                  double* g_col, double* g_tex, double* g_sh) {
  Handle* h = (Handle*)handle;
  if (!h) return fail("null handle");
  if (B <= 0 || texH <= 0 || texW <= 0) return fail("batch and texture sizes must be > 0");
  const size_t N = h->topo.N, C = h->C, P = (size_t)h->W * h->H, T = (size_t)texH * texW;
  if (h->albedo == kNormal || h->albedo == kLighting) {
    // CudaRenderer.py:207-213: zero gradients in these modes
    std::memset(g_pos, 0, sizeof(double) * B * N * 3);
    std::memset(g_col, 0, sizeof(double) * B * N * 3);
    std::memset(g_tex, 0, sizeof(double) * B * T * 3);
    std::memset(g_sh, 0, sizeof(double) * B * C * 27);
    return 0;
  }
  for (int b = 0; b < B; ++b)
    backward_one(*h, texH, texW, g_render + b * C * P * 3, g_target + b * C * P * 3, target + b * C * P * 3,
                 vpos + b * N * 3, vcol + b * N * 3, tex + b * T * 3, sh + b * C * 27, vnormal + b * C * N * 3,
                 bary + b * C * P * 2, face + b * C * P, extr + b * C * 12, intr + b * C * 9, g_pos + b * N * 3,
                 g_col + b * N * 3, g_tex + b * T * 3, g_sh + b * C * 27);
  return 0;
}

// Reference-faithful vertex->face lists (CUDABasedRasterization.cpp:125-154), including the
// compaction that skips isolated vertices.  Returns the number of (start,count) pairs.
int gvvo_reference_vertex_faces(const int32_t* faces, int64_t n_face_idx, int32_t n_vertices,
                                int32_t* out_list, int64_t list_cap, int32_t* out_pairs, int64_t pair_cap,
                                int64_t* out_list_len) {
  std::vector<int32_t> list, pairs;
  for (int v = 0; v < n_vertices; ++v) {
    int start = (int)list.size();
    for (int64_t j = 0; j < n_face_idx; j += 3)
      if (v == faces[j] || v == faces[j + 1] || v == faces[j + 2]) list.push_back((int)(j / 3));
    int n = (int)list.size() - start;
    if (n > 0) { pairs.push_back(start); pairs.push_back(n); }
  }
  if ((int64_t)list.size() > list_cap || (int64_t)pairs.size() > pair_cap) return -1;
  std::copy(list.begin(), list.end(), out_list);
  std::copy(pairs.begin(), pairs.end(), out_pairs);
  *out_list_len = (int64_t)list.size();
  return (int)(pairs.size() / 2);
}

// Camera set-up exposed for tests: invE[16], invP[16], origin[3].
void gvvo_camera(const float* extr12, const float* intr9, float* invE, float* invP, float* origin) {
  Camera c = make_camera(extr12, intr9);
  std::memcpy(invE, c.invE, sizeof(c.invE));
  std::memcpy(invP, c.invP, sizeof(c.invP));
  origin[0] = c.o.x; origin[1] = c.o.y; origin[2] = c.o.z;
}

// Pixel ray and triangle barycentrics exposed for tests.
void gvvo_pixel_barycentric(const float* extr12, const float* intr9, float px, float py, const float* v9,
                            float* out_dir3, float* out_abc3) {
  Camera c = make_camera(extr12, intr9);
  f3 d = pixel_ray(c, px, py);
  f3 abc = ray_barycentric(c.o, d, ld3(v9), ld3(v9 + 3), ld3(v9 + 6));
  out_dir3[0] = d.x; out_dir3[1] = d.y; out_dir3[2] = d.z;
  out_abc3[0] = abc.x; out_abc3[1] = abc.y; out_abc3[2] = abc.z;
}

void gvvo_bary_jacobian(const float* o3, const float* d3_, const float* v9, double* out27) {
  double J[3][9];
  bary_jacobian(ld3(o3), ld3(d3_), ld3(v9), ld3(v9 + 3), ld3(v9 + 6), J);
  for (int r = 0; r < 3; ++r) for (int k = 0; k < 9; ++k) out27[r * 9 + k] = J[r][k];
}

}  // extern "C"
