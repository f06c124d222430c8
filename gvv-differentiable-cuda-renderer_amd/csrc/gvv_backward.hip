// gvv_backward.hip -- backward kernels of the rasterizer (gfx950).
//
// Replaces renderBuffersGradGPU (reference cpp/src/Renderer/CUDABasedRasterizationGrad.cu:624-635)
// and its per-pixel Jacobian chains (Grad.cu:114-617, RendererUtil.h:241-914).
//
//   k_pixel_grad   one wavefront per TILE x TILE pixel tile (TILE 8/16/32, in 8x8 rounds).
//                  Each covered pixel forms its albedo, SH, barycentric-position,
//                  model-to-data and normal terms, from the face's ray/triangle and Jacobian
//                  constants (FaceGrad, one 128-byte line per face, written by the setup
//                  launch).  Per-face terms are summed in LDS slots (one per distinct face of
//                  the tile) before lane-parallel global atomics flush them to the face's three
//                  vertices; in the shaded modes a vertex's position and normal terms share one
//                  24-byte row of a scratch buffer, so that they cost one atomic request rather
//                  than two.  SH terms accumulate in registers over the tile and are reduced
//                  once per tile (a transposed DPP sum) into a partial buffer.
//   k_grad_finish  one launch of two parts.  Position gradient: the normal -> position chain.
//                  Instead of walking every incident face of every corner for every pixel
//                  (Grad.cu:559-615), the pixels accumulate dL/d(vertex normal) per vertex;
//                  each vertex then gathers, over its incident faces, the face's share: with G
//                  the sum of the face's vertices' normal gradients, G x (vj - vi) to corner k,
//                  (vk - vi) x G to corner j and minus both to corner i -- the same sum,
//                  regrouped -- and adds the pixels' position terms from the scratch rows;
//                  plain stores, no atomics.  SH: fixed-order reduction of the SH partials
//                  (deterministic).
#include "gvv_bary_grad.hpp"
#include "gvv_kernels.hpp"

namespace gvv {

template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, BOUND_ZERO));
}

// One halving step of a transposed wave sum: of the 2H values r[0..2H), a lane keeps the half
// that its lane bit MASK selects (upper half when set) and adds the partner's (lane ^ MASK)
// copy of that half, which the partner gives away.  XOR_FROM reads lane ^ MASK (same-row
// permutations: a DPP control; across rows: a shuffle).
template <int H, class XorFrom>
__device__ __forceinline__ void halve(float* r, bool upper, XorFrom xor_from) {
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const float keep = upper ? r[H + j] : r[j];
    const float give = upper ? r[j] : r[H + j];
    r[j] = keep + xor_from(give);
  }
}

// The sums over the 64 lanes of the 27 values v[0..27), transposed: lane L returns the total of
// value L >> 1 (for L < 54; lanes L and L ^ 1 hold the same total).  Five halving exchanges
// (lane bits 5 to 1) leave each lane one value summed over 32 lanes; one more adds the pair.
// All 64 lanes must be active.  The summation order is fixed, so the result is deterministic.
__device__ __forceinline__ float wave_sums27(const float* v, int lane) {
  float r[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) r[k] = k < 27 ? v[k] : 0.f;
  halve<16>(r, lane & 32, [](float x) { return __shfl_xor(x, 32); });
  halve<8>(r, lane & 16, [](float x) { return __shfl_xor(x, 16); });
  halve<4>(r, lane & 8, [](float x) { return dpp_mov<0x128, 0xf, false>(x); });       // row_ror:8
  const bool low4 = (lane & 4) == 0;
  halve<2>(r, !low4, [low4](float x) {
    const float up = dpp_mov<0x104, 0xf, false>(x), down = dpp_mov<0x114, 0xf, false>(x);   // row_shl:4 / row_shr:4
    return low4 ? up : down;
  });
  halve<1>(r, lane & 2, [](float x) { return dpp_mov<0x4E, 0xf, false>(x); });        // quad_perm [2,3,0,1]
  return r[0] + dpp_mov<0xB1, 0xf, false>(r[0]);                                     // quad_perm [1,0,3,2]
}

// Merges the per-face terms of lane pairs (lane, lane ^ MASK) that add to the same face slot:
// the lower lane takes the sum and the upper one drops out (slot -1), so fewer lanes add to
// one LDS address in the atomics that follow (same-address atomics serialise).  UP is the DPP
// control by which a lower lane reads lane ^ MASK, DOWN the one by which an upper lane does
// (the same control for the self-inverse permutations); every lane must be active.
template <int UP, int DOWN, int MASK, int NT>
__device__ __forceinline__ void merge_pairs(float* t, int& slot, int lane) {
  const bool low = (lane & MASK) == 0;
  const int up = __builtin_amdgcn_update_dpp(-1, slot, UP, 0xf, 0xf, false);
  const int down = UP == DOWN ? up : __builtin_amdgcn_update_dpp(-1, slot, DOWN, 0xf, 0xf, false);
  const bool merge = slot >= 0 && (low ? up : down) == slot;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    const float w = dpp_mov<UP, 0xf, false>(t[k]);
    if (merge && low) t[k] += w;
  }
  if (merge && !low) slot = -1;
}

// d(a,b,c)/d(v0,v1,v2) for a fixed ray, contracted with the row vector gb (gradient w.r.t. the
// three barycentrics), added to out[OFF .. OFF+8] (gvv_bary_grad.hpp; reference dJBCDVerpos,
// RendererUtil.h:670-861).  Binary32 throughout: the ill-conditioned plane hit is formed with
// error-free transformations.
// `f` holds the face's constants (bary_face_setup of the same vertices), v0 its first vertex.
template <int OFF>
__device__ __forceinline__ void bary_position_grad(V3 o, V3 d, V3 v0, const BaryFace& f, V3 gb, float* out) {
  const BaryGradSetup g = bary_grad_pixel(f, o, d, v0);
  if (!g.live) return;
  bary_grad_apply(g, d, gb, __builtin_amdgcn_rcpf(g.den), __builtin_amdgcn_rcpf(g.nd), out + OFF);
}
#pragma clang fp contract(fast)
// SH basis of a unit normal (the reference's order of the nine terms) and the illumination
// it gives under the coefficients of one view, for the continuous terms of the backward.
GVV_HD void sh_basis(V3 n, float Y[9]) {
  Y[0] = 1.f; Y[1] = n.y; Y[2] = n.z; Y[3] = n.x; Y[4] = n.x * n.y; Y[5] = n.z * n.y;
  Y[6] = 3.f * n.z * n.z - 1.f; Y[7] = n.x * n.z; Y[8] = n.x * n.x - n.y * n.y;
}
GVV_HD V3 sh_light(const float Y[9], const float* sh) {
  float L[3];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    float l = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) l += sh[9 * ch + k] * Y[k];
    L[ch] = l;
  }
  return v3(L[0], L[1], L[2]);
}
// The reference's four-texel blend of the backward's albedo (Grad.cu:282-312).
GVV_HD V3 blend4(float su, float sv, float tu, float tv, V3 cLL, V3 cHL, V3 cLH, V3 cHH) {
  const F3 r = sv * (su * f3(cLL) + tu * f3(cHL)) + tv * (su * f3(cLH) + tu * f3(cHH));
  return v3(r.x, r.y, r.z);
}
#pragma clang fp contract(off)

// Per-pixel gradient terms of one covered pixel (reference Grad.cu:205-555).  Adds the SH
// terms to sh_acc; returns the per-face terms (position, vertex-normal gradient and, when
// G_COL >= 0, vertex-colour rows for the three corners) in t and, textured, the texture-gradient term in
// tex_val for texel tex_key of the view's batch entry (left at -1 when there is none).
// `cam` and `shc` are the view's camera and SH coefficients, held in registers by the caller.
template <int ALBEDO, bool SHADED, bool M2D, bool HAVE_BARY, int G_POS, int G_VN, int G_COL>
__device__ __forceinline__ void pixel_terms(const BackwardArgs& a, const Camera& cam, const float* shc, int bc, int b,
                                            int px, int py, size_t idx, int idf, int i0, int i1, int i2, float* t,
                                            float sh_acc[27], int& tex_key, float tex_val[3]) {
  // Every load that does not depend on the barycentrics is issued before they are computed,
  // so that all of them share one round trip to memory; only the texels wait for a second.
  const float* vp = a.vpos + (size_t)b * a.N * 3;
  const V3 p0 = ld3(vp + 3 * i0);
  V3 p1, p2;
  if (M2D) { p1 = ld3(vp + 3 * i1); p2 = ld3(vp + 3 * i2); }
  // the face's ray/triangle and Jacobian constants (one 128-byte line, written by the setup)
  const FaceGrad* fr = a.face_grad + (size_t)b * a.F + idf;
  TriSetup ts;
  if (!HAVE_BARY) ts = fr->t;
  BaryFace bf;
  if (SHADED) bf = fr->g;
  const float* vnb = a.vnormal + (size_t)bc * a.N * 3;
  const V3 n0 = ld3(vnb + 3 * i0), n1 = ld3(vnb + 3 * i1), n2 = ld3(vnb + 3 * i2);
  const float* gr = a.g_render + 3 * idx;
  const V3 g = v3(gr[0], gr[1], gr[2]);
  float uv6[6];
  if (ALBEDO == kTextured) {
#pragma unroll
    for (int k = 0; k < 6; ++k) uv6[k] = a.uv[6 * (size_t)idf + k];
  }
  V3 c0, c1, c2;
  if (ALBEDO == kVertexColor) {
    const float* vc = a.vcol + (size_t)b * a.N * 3;
    c0 = ld3(vc + 3 * i0); c1 = ld3(vc + 3 * i1); c2 = ld3(vc + 3 * i2);
  }
  float ba = 0.f, bb = 0.f;
  if (HAVE_BARY) { ba = a.bary[2 * idx]; bb = a.bary[2 * idx + 1]; }

  const V3 o = cam_origin(cam);
  const V3 d = pixel_ray(cam, px + 0.5f, py + 0.5f);
  V3 abc;
  if (HAVE_BARY) abc = v3(ba, bb, 1.f - ba - bb);
  else abc = ray_barycentric(ts, cam_origin_scaled(cam), d);
  const float w3[3] = {abc.x, abc.y, abc.z};
  // textured: the four texels of the backward's blend (Grad.cu:282-312; not the forward's
  // nearest texel) are loaded before the normal and the illumination are formed, which do not
  // need them
  TexCoord tc;
  V3 cLL, cHL, cLH, cHH;
  if (ALBEDO == kTextured) {
    tc = tex_coord(uv6, abc, a.texW, a.texH);
    const float* tb = a.tex + (size_t)b * a.texH * a.texW * 3;
    const int lu = min(max(tc.LU, 0), a.texW - 1), hu = min(max(tc.LU + 1, 0), a.texW - 1);
    const int lv = min(max(tc.LV, 0), a.texH - 1), hv = min(max(tc.LV + 1, 0), a.texH - 1);
    cLL = ld3(tb + 3 * ((size_t)lv * a.texW + lu)); cHL = ld3(tb + 3 * ((size_t)lv * a.texW + hu));
    cLH = ld3(tb + 3 * ((size_t)hv * a.texW + lu)); cHH = ld3(tb + 3 * ((size_t)hv * a.texW + hu));
  }
  const V3 un = n0 * abc.x + n1 * abc.y + n2 * abc.z;
  const float unl = length(un);
  V3 n = un / unl;         // IEEE division: the flip decision below is part of the binary32 contract
  const bool flipped = dot(n, d) > 0.f;
  if (flipped) n = -n;

  V3 L = v3(1.f, 1.f, 1.f);
  float Y[9];
  if (SHADED) {
    sh_basis(n, Y);
    L = sh_light(Y, shc);
  }

  V3 alb = v3(0.f, 0.f, 0.f);
  if (ALBEDO == kVertexColor) {
    alb = c0 * abc.x + c1 * abc.y + c2 * abc.z;
    if (G_COL >= 0) {
      const V3 ga = v3(g.x * L.x, g.y * L.y, g.z * L.z);
#pragma unroll
      for (int r = 0; r < 3; ++r) { t[G_COL + 3 * r] = ga.x * w3[r]; t[G_COL + 3 * r + 1] = ga.y * w3[r]; t[G_COL + 3 * r + 2] = ga.z * w3[r]; }
    }
  } else if (ALBEDO == kTextured) {
    const float LU = tc.LU + 0.5f, HU = tc.LU + 1.5f, LV = tc.LV + 0.5f, HV = tc.LV + 1.5f;
    const float U0 = tc.U0, V0 = tc.V0;
    alb = blend4(U0 - LU, V0 - LV, HU - U0, HV - V0, cLL, cHL, cLH, cHH);
    if (!flipped && tc.LV >= 0 && tc.LV < a.texH && tc.LU >= 0 && tc.LU < a.texW) {
      tex_key = tc.LV * a.texW + tc.LU;
      tex_val[0] = g.x * L.x;
      tex_val[1] = g.y * L.y;
      tex_val[2] = g.z * L.z;
    }
  }

  if (SHADED) {
    // continuous terms only (the flip decision is taken): products and sums may fuse
#pragma clang fp contract(fast)
    const V3 gl = v3(g.x * alb.x, g.y * alb.y, g.z * alb.z);
    const float glc[3] = {gl.x, gl.y, gl.z};
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
#pragma unroll
      for (int k = 0; k < 9; ++k) sh_acc[ch * 9 + k] += glc[ch] * Y[k];

    // d light / d n (RendererUtil.h:371-391), then through the normalisation of the
    // unflipped interpolated normal (RendererUtil.h:398-415).
    float gn[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float* sc = shc + 9 * i;
      gn[0] += glc[i] * (sc[3] + sc[4] * n.y + sc[7] * n.z + sc[8] * 2.f * n.x);
      gn[1] += glc[i] * (sc[1] + sc[4] * n.x + sc[5] * n.z + sc[8] * -2.f * n.y);
      gn[2] += glc[i] * (sc[2] + sc[5] * n.y + sc[6] * 6.f * n.z + sc[7] * n.x);
    }
    const float unl2 = unl * unl, inv_unl3 = __builtin_amdgcn_rcpf(unl2 * unl);
    const float gnu = gn[0] * un.x + gn[1] * un.y + gn[2] * un.z;
    const V3 gu = v3((gn[0] * unl2 - un.x * gnu) * inv_unl3, (gn[1] * unl2 - un.y * gnu) * inv_unl3,
                     (gn[2] * unl2 - un.z * gnu) * inv_unl3);
#pragma unroll
    for (int r = 0; r < 3; ++r) { t[G_VN + 3 * r] = gu.x * w3[r]; t[G_VN + 3 * r + 1] = gu.y * w3[r]; t[G_VN + 3 * r + 2] = gu.z * w3[r]; }
    const V3 gb = v3(dot(f3(gu), f3(n0)), dot(f3(gu), f3(n1)), dot(f3(gu), f3(n2)));
    bary_position_grad<G_POS>(o, d, p0, bf, gb, t);
  }

  if (M2D) {
    // Model-to-data term: target image gradient at integer pixel coordinates
    // (RendererUtil.h:566-620) times the projection Jacobian of the interpolated surface
    // point (RendererUtil.h:275-332).
    const float* gtp = a.g_target + 3 * idx;
    const V3 gT = v3(gtp[0], gtp[1], gtp[2]);
    const int fs = a.image_fs;
    float q0 = 0.f, q1 = 0.f;
    if (px >= fs + 1 && py >= fs + 1 && px < a.W - (fs + 1) && py < a.H - (fs + 1)) {
      const float* img = a.target + (size_t)bc * a.W * a.H * 3;
      V3 du = v3(0.f, 0.f, 0.f), dv = v3(0.f, 0.f, 0.f);
      float norm = 0.f;
      for (int y = -fs; y <= fs; ++y)
        for (int x = -fs; x <= fs; ++x) {
          const float den = (float)(x * x + y * y);
          float gu_ = 0.f, gv_ = 0.f;
          if (den != 0.f) { gu_ = (float)x / den; gv_ = (float)y / den; }
          const V3 I = ld3(img + 3 * ((size_t)(py + y) * a.W + (px + x)));
          du = du + I * gu_;
          dv = dv + I * gv_;
          norm += fabsf(gu_);
        }
      q0 = dot(gT, du) / norm;
      q1 = dot(gT, dv) / norm;
    }
    if (q0 != 0.f || q1 != 0.f) {
      const V3 X = p0 * abc.x + p1 * abc.y + p2 * abc.z;
      const float* K = cam.K;
      const float* E = cam.E;
      float KE[3][4];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) KE[r][cc] = K[r * 3] * E[cc] + K[r * 3 + 1] * E[4 + cc] + K[r * 3 + 2] * E[8 + cc];
      float P[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) P[r] = KE[r][0] * X.x + KE[r][1] * X.y + KE[r][2] * X.z + KE[r][3];
      if (fabsf(P[2]) > 0.0001f) {
        const float iz = 1.f / P[2];
        const float dd0[3] = {iz, 0.f, -P[0] * iz * iz}, dd1[3] = {0.f, iz, -P[1] * iz * iz};
        float gx[3];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          const float j0 = dd0[0] * KE[0][cc] + dd0[2] * KE[2][cc];
          const float j1 = dd1[1] * KE[1][cc] + dd1[2] * KE[2][cc];
          gx[cc] = q0 * j0 + q1 * j1;
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          t[G_POS + 3 * r] += gx[0] * w3[r]; t[G_POS + 3 * r + 1] += gx[1] * w3[r]; t[G_POS + 3 * r + 2] += gx[2] * w3[r];
        }
      }
    }
  }
}

// One wavefront per TILE x TILE pixel tile of one (b, c) image (TILE = 8, 16 or 32), in
// (TILE/8)^2 rounds of 8x8 pixels.
// Per-face terms: every distinct face of the tile gets a slot in LDS (every lane finds or
// inserts its face in an LDS hash table at once, and the inserting lanes number the new slots
// in lane order); lanes add their terms into their face's slot with LDS atomics, and the
// slots are flushed with lane-parallel global atomics (one per face, vertex and component).
// SH terms accumulate in registers over the tile and are reduced once into the tile's
// partial slots.
constexpr int kFaceSlots = 96;
constexpr int kHashCap = 128;

// COL_GRAD: the vertex-colour gradient is formed (vertexColor mode with a g_col output).
template <int ALBEDO, bool SHADED, bool M2D, bool HAVE_BARY, int TILE, bool COL_GRAD>
__device__ __forceinline__ void pixel_grad_body(const BackwardArgs& a, int64_t work) {
  static_assert(TILE == 8 || TILE == 16 || TILE == 32, "tile edge must be 8, 16 or 32");
  constexpr int RT = TILE / 8;                            // 8x8 rounds per tile edge
  constexpr bool NEED_POS = SHADED || M2D;
  constexpr bool NEED_COL = ALBEDO == kVertexColor && COL_GRAD;
  constexpr int G_POS = 0, G_VN = NEED_POS ? 9 : 0, G_COL = G_VN + (SHADED ? 9 : 0);
  constexpr int NT = G_COL + (NEED_COL ? 9 : 0);         // per-face terms
  constexpr int NT1 = NT > 0 ? NT : 1;
  __shared__ float s_acc[kFaceSlots][NT1];
  __shared__ int s_face[kFaceSlots];
  // face id -> slot: open addressing with linear probing, keys face + 1 (0: free)
  __shared__ int2 s_hash[kHashCap];
  constexpr int TS = ALBEDO == kTextured ? 64 : 1;
  __shared__ int s_tkey[TS];          // a round's texture-gradient terms, compacted
  __shared__ float s_tval[TS * 3];
  const int lane = threadIdx.x;
  const int tiles = a.tiles_x * a.tiles_y;
  // one (view, tile) per workgroup
  // the grid has fewer than 2^32 workgroups: 32-bit quotient and remainder
  const unsigned w32 = (unsigned)work;
  const int tile = __builtin_amdgcn_readfirstlane((int)(w32 % (unsigned)tiles));
  const int bc = __builtin_amdgcn_readfirstlane((int)(w32 / (unsigned)tiles));
  const int b = __builtin_amdgcn_readfirstlane(bc / a.C);
  const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
  // the view's camera and SH coefficients: uniform, loaded once before any store of the kernel
  // (so into scalar registers), not per pixel
  Camera cam;
  {
    const Camera& src = a.cams[bc];
#pragma unroll
    for (int k = 0; k < 12; ++k) cam.invP[k] = src.invP[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) { cam.o[k] = src.o[k]; cam.os[k] = src.os[k]; }
    if (M2D) {
#pragma unroll
      for (int k = 0; k < 12; ++k) cam.E[k] = src.E[k];
#pragma unroll
      for (int k = 0; k < 9; ++k) cam.K[k] = src.K[k];
    }
  }
  float shv[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) shv[k] = SHADED ? a.sh[(size_t)bc * 27 + k] : 0.f;
  float sh_acc[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) sh_acc[k] = 0.f;
  bool any = false;
  int nslots = 0;
  static_assert(kHashCap % 64 == 0 && kHashCap > kFaceSlots, "hash table: whole lanes, spare capacity");
#pragma unroll
  for (int q = 0; q < kHashCap / 64; ++q) s_hash[lane + 64 * q].x = 0;
  __syncthreads();

  auto flush = [&]() {
    __syncthreads();
    for (int i = lane; i < nslots * NT; i += 64) {
      const int slot = i / NT, k = i % NT;
      const float v = s_acc[slot][k];
      if (v == 0.f) continue;
      const int grp = k / 9, kk = k % 9;
      const int vtx = a.faces[3 * s_face[slot] + kk / 3];
      const size_t row = (size_t)b * a.N + vtx;
      float* dst;
      if (SHADED && grp < 2) dst = a.g_pv + row * 6 + 3 * grp + kk % 3;   // position, normal: one 24-byte row
      else dst = (grp * 9 == G_POS && NEED_POS ? a.g_pos : a.g_col) + row * 3 + kk % 3;
      atomicAdd(dst, v);
    }
#pragma unroll
    for (int q = 0; q < kHashCap / 64; ++q) s_hash[lane + 64 * q].x = 0;
    __syncthreads();
    nslots = 0;
  };

  // Software pipeline over the rounds: the next round's face id, and then its vertex ids, are
  // loaded while this round computes.
  auto round_pixel = [&](int round, int& px, int& py, size_t& idx) {
    px = tx * TILE + (round % RT) * 8 + (lane & 7);
    py = ty * TILE + (round / RT) * 8 + (lane >> 3);
    const bool valid = px < a.W && py < a.H;
    idx = ((size_t)bc * a.H + (valid ? py : 0)) * a.W + (valid ? px : 0);
    int f = valid ? a.face[idx] : -1;
    return f >= a.F ? -1 : f;   // ids outside the mesh are treated as background
  };
  int npx, npy;
  size_t nidx;
  int nidf = round_pixel(0, npx, npy, nidx);
  int ni0 = 0, ni1 = 0, ni2 = 0;
  if (nidf >= 0) { ni0 = a.faces[3 * nidf]; ni1 = a.faces[3 * nidf + 1]; ni2 = a.faces[3 * nidf + 2]; }
  for (int round = 0; round < RT * RT; ++round) {
    const int px = npx, py = npy, idf = nidf, i0 = ni0, i1 = ni1, i2 = ni2;
    const size_t idx = nidx;
    if (round + 1 < RT * RT) nidf = round_pixel(round + 1, npx, npy, nidx);
    uint64_t pending = __ballot(idf >= 0);
    if (pending) {
      any = true;
    }
    float t[NT1 + 9];     // +9: unused groups of pixel_terms write past NT1 only in dead code
#pragma unroll
    for (int k = 0; k < NT1 + 9; ++k) t[k] = 0.f;
    int tex_key = -1;
    float tex_val[3] = {0.f, 0.f, 0.f};
    if (idf >= 0)
      pixel_terms<ALBEDO, SHADED, M2D, HAVE_BARY, G_POS, G_VN, NEED_COL ? G_COL : -1>(a, cam, shv, bc, b, px, py, idx, idf, i0, i1, i2, t,
                                                                     sh_acc, tex_key, tex_val);
    if (round + 1 < RT * RT && nidf >= 0) {
      ni0 = a.faces[3 * nidf]; ni1 = a.faces[3 * nidf + 1]; ni2 = a.faces[3 * nidf + 2];
    }
    if (ALBEDO == kTextured && pending) {
      // texture-gradient terms: lanes of one texel are merged first (about 1.8 pixels share a
      // texel in a round at one texel per pixel), then one atomic per texel and channel
      merge_pairs<0xB1, 0xB1, 1, 3>(tex_val, tex_key, lane);
      merge_pairs<0x128, 0x128, 8, 3>(tex_val, tex_key, lane);
      merge_pairs<0x4E, 0x4E, 2, 3>(tex_val, tex_key, lane);
      merge_pairs<0x104, 0x114, 4, 3>(tex_val, tex_key, lane);
      // the surviving texels are compacted into LDS and added with lane q taking channel q % 3
      // of texel q / 3, so the three channels of a texel (and neighbouring texels) are adjacent
      // addresses of one atomic instruction, not one address each of three
      const uint64_t sm = __ballot(tex_key >= 0);
      const int nsurv = __popcll(sm);
      if (tex_key >= 0) {
        const int r = __builtin_amdgcn_mbcnt_hi((unsigned)(sm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)sm, 0u));
        s_tkey[r] = tex_key;
        s_tval[3 * r + 0] = tex_val[0];
        s_tval[3 * r + 1] = tex_val[1];
        s_tval[3 * r + 2] = tex_val[2];
      }
      __syncthreads();
      float* gtb = a.g_tex + (size_t)b * a.texH * a.texW * 3;
      for (int q = lane; q < 3 * nsurv; q += 64) {
        const int sidx = q / 3;
        atomicAdd(gtb + 3 * (size_t)s_tkey[sidx] + (q - 3 * sidx), s_tval[q]);
      }
      __syncthreads();
    }
    if (!pending) continue;
    if (NT == 0) continue;
    // assign each distinct face of this round a slot (reusing slots of earlier rounds).  The
    // round adds at most as many faces as it has runs of equal ids along the lanes (a run
    // starts where a lane's face differs from the previous lane's, or at a 16-lane row start),
    // so flushing first when fewer slots remain keeps every slot of this round valid until its
    // terms are added, without flushing a tile whose faces all fit
    {
      const int prev = __builtin_amdgcn_update_dpp(-2, idf, 0x111, 0xf, 0xf, false);   // row_shr:1
      const int runs = __popcll(__ballot(idf >= 0 && idf != prev));
      if (nslots + runs > kFaceSlots) flush();
    }
    // every lane finds or inserts its face in the hash table at once; the lanes that inserted a
    // face number their new slots in lane order
    int myslot = -1, h = 0;
    bool inserted = false;
    if (idf >= 0) {
      const int key = idf + 1;
      h = (int)(((unsigned)idf * 2654435761u) >> 25) & (kHashCap - 1);
      for (;;) {
        const int old = atomicCAS(&s_hash[h].x, 0, key);
        if (old == 0) { inserted = true; break; }
        if (old == key) break;
        h = (h + 1) & (kHashCap - 1);
      }
    }
    const uint64_t ins = __ballot(inserted);
    const int nnew = __popcll(ins);
    if (inserted) {
      const int slot = nslots + __builtin_amdgcn_mbcnt_hi((unsigned)(ins >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)ins, 0u));
      s_hash[h].y = slot;
      s_face[slot] = idf;
    }
    for (int q = lane; q < nnew * NT; q += 64) s_acc[nslots + q / NT][q % NT] = 0.f;
    nslots += nnew;
    __syncthreads();
    if (idf >= 0) myslot = s_hash[h].y;
    // lanes are 8x8 pixels (x = lane & 7, y = lane >> 3): merge x neighbours (quad_perm
    // [1,0,3,2]), y neighbours (row_ror:8), x pairs two apart (quad_perm [2,3,0,1]), then
    // four apart (row_shl:4 / row_shr:4); survivors hold the sums of up to 16 lanes
    merge_pairs<0xB1, 0xB1, 1, NT>(t, myslot, lane);
    merge_pairs<0x128, 0x128, 8, NT>(t, myslot, lane);
    merge_pairs<0x4E, 0x4E, 2, NT>(t, myslot, lane);
    merge_pairs<0x104, 0x114, 4, NT>(t, myslot, lane);
    __syncthreads();
    if (myslot >= 0) {
#pragma unroll
      for (int k = 0; k < NT; ++k)
        if (t[k] != 0.f) atomicAdd(&s_acc[myslot][k], t[k]);
    }
  }
  if (NT > 0 && nslots > 0) flush();

  if (SHADED) {
    // lane 2k stores the tile's total of coefficient k
    const float s = any ? wave_sums27(sh_acc, lane) : 0.f;
    const int k = lane >> 1;
    if ((lane & 1) == 0 && k < 27) a.sh_partial[((size_t)bc * 27 + k) * tiles + tile] = s;
  }
}

template <int ALBEDO, bool SHADED, bool M2D, bool HAVE_BARY, int TILE, bool COL_GRAD>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_pixel_grad(BackwardArgs a) {
  pixel_grad_body<ALBEDO, SHADED, M2D, HAVE_BARY, TILE, COL_GRAD>(a, xcd_contiguous(blockIdx.x, gridDim.x));
}

// Fixed-order reduction of the SH tile partials ([b*C + c][27][tiles]): one workgroup per
// (b, c, coefficient), so the result does not depend on scheduling.
__device__ __forceinline__ void sh_reduce_block(const float* __restrict__ partial, float* __restrict__ g_sh, int tiles,
                                                int blk, float* red) {
  const int bc = blk / 27, k = blk % 27;
  const float* p = partial + ((size_t)bc * 27 + k) * tiles;
  float s = 0.f;
  for (int t = threadIdx.x; t < tiles; t += 256) s += p[t];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) g_sh[(size_t)bc * 27 + k] = red[0];
}

// Face f's share of vertex v's position gradient through the vertex normals: with G the sum
// of the normal gradients of f's distinct vertices, G x (vj - vi) goes to corner k,
// (vk - vi) x G to corner j and minus both to corner i; v takes the shares of every corner it
// holds.
__device__ __forceinline__ V3 normal_chain_share(const float* __restrict__ vp, const float* __restrict__ gv,
                                                 const int32_t* __restrict__ faces, int f, int v) {
  const int a0 = faces[3 * f], a1 = faces[3 * f + 1], a2 = faces[3 * f + 2];
  V3 G = ld3(gv + 6 * a0);
  if (a1 != a0) G = G + ld3(gv + 6 * a1);
  if (a2 != a0 && a2 != a1) G = G + ld3(gv + 6 * a2);
  V3 r = v3(0.f, 0.f, 0.f);
  if (G.x == 0.f && G.y == 0.f && G.z == 0.f) return r;
  const V3 vi = ld3(vp + 3 * a0), vj = ld3(vp + 3 * a1), vk = ld3(vp + 3 * a2);
  const V3 gk = cross(G, vj - vi);
  const V3 gj = cross(vk - vi, G);
  if (a0 == v) r = r - (gj + gk);
  if (a1 == v) r = r + gj;
  if (a2 == v) r = r + gk;
  return r;
}

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m);
  return x;
}

// The position gradient of the workgroup's 256 (batch entry, vertex) items, gathered: the
// pixels' position terms from the scratch rows plus every incident face's normal-chain share
// (over the vertex -> face lists, in list order), written with plain stores -- no atomics.  A
// vertex with more than 16 incident faces (the centre of a cap fan has 250) is gathered by
// its whole wavefront.
__device__ __forceinline__ void pos_gather_block(const FinishArgs& f, int blk) {
  constexpr int kHeavy = 16;
  const int64_t i = (int64_t)blk * 256 + threadIdx.x;
  const bool valid = i < (int64_t)f.B * f.N;
  const int b = valid ? (int)(i / f.N) : 0, v = valid ? (int)(i % f.N) : 0;
  const int s = valid ? f.csr_start[v] : 0, n = valid ? f.csr_count[v] : 0;
  const float* vp = f.vpos + (size_t)b * f.N * 3;
  const float* gv = f.g_pv + (size_t)b * f.N * 6 + 3;   // the normal-gradient part of each row
  V3 acc = v3(0.f, 0.f, 0.f);
  if (valid) acc = ld3(f.g_pv + (size_t)i * 6);
  if (n <= kHeavy)
    for (int k = 0; k < n; ++k) acc = acc + normal_chain_share(vp, gv, f.faces, f.csr_list[s + k], v);
  const int lane = threadIdx.x & 63;
  for (uint64_t heavy = __ballot(n > kHeavy); heavy; heavy &= heavy - 1) {
    const int j = __ffsll((unsigned long long)heavy) - 1;
    const int sj = __builtin_amdgcn_readlane(s, j), nj = __builtin_amdgcn_readlane(n, j);
    const int vj = __builtin_amdgcn_readlane(v, j), bj = __builtin_amdgcn_readlane(b, j);
    const float* vpj = f.vpos + (size_t)bj * f.N * 3;
    const float* gvj = f.g_pv + (size_t)bj * f.N * 6 + 3;
    V3 part = v3(0.f, 0.f, 0.f);
    for (int k = lane; k < nj; k += 64) part = part + normal_chain_share(vpj, gvj, f.faces, f.csr_list[sj + k], vj);
    const V3 tot = v3(wave_sum(part.x), wave_sum(part.y), wave_sum(part.z));
    if (lane == j) acc = acc + tot;
  }
  if (!valid) return;
  float* o = f.g_pos + (size_t)i * 3;
  o[0] = acc.x; o[1] = acc.y; o[2] = acc.z;
}

// The reductions after the pixel kernel in one launch: blocks [0, vg_blocks) gather the
// position gradient, the rest run the SH reduction (one per view and coefficient).
__device__ __forceinline__ void grad_finish_body(const FinishArgs& f, int blk) {
  __shared__ float smem[256];
  if (blk < f.vg_blocks) {
    pos_gather_block(f, blk);
    return;
  }
  sh_reduce_block(f.sh_partial, f.g_sh, f.tiles, blk - f.vg_blocks, smem);
}

__global__ void __launch_bounds__(256) k_grad_finish(FinishArgs f) { grad_finish_body(f, blockIdx.x); }

__global__ void __launch_bounds__(256) k_grad_finish_multi(MeshTable<FinishArgs> t) {
  const MeshSlice m = t.find(blockIdx.x);
  const FinishArgs f = t.load(m.mesh);
  grad_finish_body(f, (int)m.local);
}

FinishArgs finish_args(const BackwardArgs& a) {
  FinishArgs f;
  f.vpos = a.vpos; f.faces = a.faces; f.g_pv = a.g_pv; f.g_pos = a.g_pos; f.sh_partial = a.sh_partial; f.g_sh = a.g_sh;
  f.csr_start = a.csr_start; f.csr_count = a.csr_count; f.csr_list = a.csr_list;
  f.B = a.B; f.N = a.N; f.F = a.F; f.tiles = a.tiles_x * a.tiles_y;
  f.vg_blocks = (int)(((int64_t)a.B * a.N + 255) / 256);
  return f;
}

template <int ALBEDO, bool SHADED, bool M2D, bool HAVE_BARY, bool COL_GRAD>
static void launch_pixel_grad_t(const BackwardArgs& a, hipStream_t s) {
  dim3 grid((unsigned)((int64_t)a.tiles_x * a.tiles_y * a.B * a.C));
  if (a.tile == 8)
    hipLaunchKernelGGL((k_pixel_grad<ALBEDO, SHADED, M2D, HAVE_BARY, 8, COL_GRAD>), grid, dim3(64), 0, s, a);
  else if (a.tile == 32)
    hipLaunchKernelGGL((k_pixel_grad<ALBEDO, SHADED, M2D, HAVE_BARY, 32, COL_GRAD>), grid, dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((k_pixel_grad<ALBEDO, SHADED, M2D, HAVE_BARY, 16, COL_GRAD>), grid, dim3(64), 0, s, a);
}

template <int ALBEDO, bool SHADED, bool COL_GRAD>
static void launch_pixel_grad_c(const BackwardArgs& a, hipStream_t s) {
  const bool m2d = a.g_target != nullptr, hb = a.bary != nullptr;
  if (m2d && hb) launch_pixel_grad_t<ALBEDO, SHADED, true, true, COL_GRAD>(a, s);
  else if (m2d) launch_pixel_grad_t<ALBEDO, SHADED, true, false, COL_GRAD>(a, s);
  else if (hb) launch_pixel_grad_t<ALBEDO, SHADED, false, true, COL_GRAD>(a, s);
  else launch_pixel_grad_t<ALBEDO, SHADED, false, false, COL_GRAD>(a, s);
}

// The vertex-colour gradient's terms are formed only in vertexColor mode with a g_col output.
template <int ALBEDO, bool SHADED>
static void launch_pixel_grad_m(const BackwardArgs& a, hipStream_t s) {
  if (ALBEDO == kVertexColor && !a.g_col) launch_pixel_grad_c<ALBEDO, SHADED, false>(a, s);
  else launch_pixel_grad_c<ALBEDO, SHADED, true>(a, s);
}

static inline unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

unsigned finish_blocks(const FinishArgs& f, int C) { return (unsigned)(f.vg_blocks + f.B * C * 27); }

hipError_t launch_grad_finish_multi(const MeshTable<FinishArgs>& t, unsigned blocks, hipStream_t s) {
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_grad_finish_multi, dim3(blocks), dim3(256), 0, s, t);
  return hipGetLastError();
}

SetupArgs backward_setup_args(const BackwardPlan& p) {
  const BackwardArgs& a = p.a;
  const int64_t nv = (int64_t)a.B * a.N * 3;
  SetupArgs sa;
  sa.extr = p.extr; sa.intr = p.intr; sa.cams = p.cams; sa.n_cams = a.B * a.C;
  // the per-face constants of the pixel kernel, once per face and batch entry
  sa.vpos = a.vpos; sa.faces = a.faces; sa.face_grad = const_cast<FaceGrad*>(a.face_grad);
  sa.B = a.B; sa.N = a.N; sa.F = a.F;
  sa.fill[0] = a.g_pos; sa.fill_n[0] = nv;
  sa.fill[1] = a.g_col; sa.fill_n[1] = nv;
  sa.fill[2] = a.g_tex; sa.fill_n[2] = (int64_t)a.B * a.texH * a.texW * 3;
  if (p.shaded) { sa.fill[3] = a.g_pv; sa.fill_n[3] = 2 * nv; }
  else { sa.fill[3] = a.g_sh; sa.fill_n[3] = (int64_t)a.B * a.C * 27; }
  return sa;
}

hipError_t launch_pixel_grad(const BackwardPlan& p, hipStream_t s) {
  const BackwardArgs& a = p.a;
  switch (p.albedo * 2 + (p.shaded ? 1 : 0)) {
    case kVertexColor * 2 + 1: launch_pixel_grad_m<kVertexColor, true>(a, s); break;
    case kVertexColor * 2 + 0: launch_pixel_grad_m<kVertexColor, false>(a, s); break;
    case kTextured * 2 + 1: launch_pixel_grad_m<kTextured, true>(a, s); break;
    case kTextured * 2 + 0: launch_pixel_grad_m<kTextured, false>(a, s); break;
    case kForegroundMask * 2 + 1: launch_pixel_grad_m<kForegroundMask, true>(a, s); break;
    default: launch_pixel_grad_m<kForegroundMask, false>(a, s); break;
  }
  return hipGetLastError();
}

hipError_t launch_backward(const BackwardPlan& p, hipStream_t s) {
  const BackwardArgs& a = p.a;
  {
    // cameras and the zero fills of the gradient outputs in one launch
    ProfScope ps(p.prof, "k_setup", s);
    hipError_t e = launch_setup(backward_setup_args(p), s);
    if (e != hipSuccess) return e;
  }
  {
    ProfScope ps(p.prof, "k_pixel_grad", s);
    hipError_t e = launch_pixel_grad(p, s);
    if (e != hipSuccess) return e;
  }
  if (p.shaded) {
    ProfScope ps(p.prof, "k_grad_finish", s);
    const FinishArgs f = finish_args(a);
    hipLaunchKernelGGL(k_grad_finish, dim3(finish_blocks(f, a.C)), dim3(256), 0, s, f);
  }
  return hipGetLastError();
}

}  // namespace gvv
