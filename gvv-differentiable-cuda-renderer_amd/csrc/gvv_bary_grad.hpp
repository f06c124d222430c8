// gvv_bary_grad.hpp -- derivative of the ray/triangle barycentrics with respect to the three
// vertex positions, contracted with the gradient of the barycentrics (the reference's
// dJBCDVerpos, cpp/src/Utils/RendererUtil.h:670-861), in binary32 only.  Host- and
// device-callable so that its accuracy can be checked on the CPU against the FP64 oracle.
//
// Regrouped form of the reference's per-variable loop.  With e1 = v1-v0, e2 = v2-v0,
// N = e1 x e2, q0 = P - v0 (P the plane hit of the ray), q1 = q0 - e1, q2 = q0 - e2, the edges
// E0 = v2-v1 and E1 = v0-v2, C_r = E_r x q_{r+1}, and w_r = gb_r - gb_2 (the third row is minus
// the sum of the first two), the contribution of vertex v with A_0 = e1-e2, A_1 = e2, A_2 = -e1
// (dN/d(v . e_k) = e_k x A_v) is
//   [A_v x (w0 C0 + w1 C1) + Q_v + dt_v (w0 N.(E0 x d) + w1 N.(E1 x d)) - D_v] / |N|^2
//   - 2 (A_v x N) (w0 N.C0 + w1 N.C1) / |N|^4
// with dt_v = (delta_v0 N - A_v x q0) / (N.d) the derivative of the ray parameter,
// Q_0 = w1 q2 x N, Q_1 = -w0 q1 x N, Q_2 = (w0 q1 - w1 q2) x N from the edge derivatives,
// and D_1 = w0 N x E0, D_2 = w1 N x E1 (D_0 = 0) from the vertex derivatives.
//
// Precision.  The plane hit is ill-conditioned: q0 = t d - w with w = v0 - o (metres from the
// camera) and t = (w.N)/(d.N), so a rounding of w, of w.N or of d.N (which cancels near
// grazing incidence) moves q0 by about eps |w| / cos, hundreds of times the triangle's own
// scale; with plain binary32 that showed as 2.8e-4 max-norm error in the position gradient of
// a 5k-face sphere.  The setup therefore keeps those quantities exact or compensated, in
// binary32 error-free transformations (Knuth's TwoSum, the FMA TwoProd; Ogita, Rump and
// Oishi, "Accurate sum and dot product", SIAM J. Sci. Comput. 26, 2005):
//   - w is the exact two-term sum v0 - o (TwoSum);
//   - the edges are exact two-term differences and N = e1 x e2 is formed from them to twice
//     the working precision (a binary32 N tilts the plane by ~eps, which moves the hit of a
//     grazing ray by |q0| eps / cos);
//   - w.N and d.N are compensated dot products (as if in twice the precision, then rounded);
//   - t is a two-term quotient, and q0 = t d - w is assembled from the exact product and sum
//     residuals, so its error is a few ulp of |q0| plus eps^2 |w|.
// Everything after q0 is well conditioned and plain binary32.
#pragma once
#include "gvv_math.hpp"

namespace gvv {

struct BaryGradSetup {
  V3 e1, e2, N, q0;
  float nd;     // N.d
  float den;    // N.N
  bool live;    // false: the reference's early exit (zero derivative)
};

// a + b = s + e exactly (Knuth's TwoSum)
GVV_HD void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// a * b = p + e exactly (FMA TwoProd)
GVV_HD void two_prod(float a, float b, float& p, float& e) {
  p = a * b;
  e = fmaf(a, b, -p);
}

// Cross product of two vectors given as exact two-term sums ah + al and bh + bl, returned as
// Nh + Nl to within a few ulp of eps^2 |a||b| per component: every product of the high parts is
// split exactly (TwoProd), the cross terms with the low parts are second order.
GVV_HD void cross2(V3 ah, V3 al, V3 bh, V3 bl, V3& Nh, V3& Nl) {
  const float a[3] = {ah.x, ah.y, ah.z}, b[3] = {bh.x, bh.y, bh.z};
  const float la[3] = {al.x, al.y, al.z}, lb[3] = {bl.x, bl.y, bl.z};
  float h[3], l[3];
  for (int k = 0; k < 3; ++k) {
    const int i = (k + 1) % 3, j = (k + 2) % 3;     // N_k = a_i b_j - a_j b_i
    float p1, r1, p2, r2, s, se;
    two_prod(a[i], b[j], p1, r1);
    two_prod(a[j], b[i], p2, r2);
    two_sum(p1, -p2, s, se);
    const float lo = se + ((r1 - r2) + ((a[i] * lb[j] + la[i] * b[j]) - (a[j] * lb[i] + la[j] * b[i])));
    two_sum(s, lo, h[k], l[k]);
  }
  Nh = v3(h[0], h[1], h[2]);
  Nl = v3(l[0], l[1], l[2]);
}

// Compensated dot product of three terms (Ogita-Rump-Oishi Dot2): hi + lo approximates a.b as
// if computed in twice the working precision.
GVV_HD void dot2(V3 a, V3 b, float& hi, float& lo) {
  float p, s, h, r, q;
  two_prod(a.x, b.x, p, s);
  two_prod(a.y, b.y, h, r);
  two_sum(p, h, p, q);
  s = s + (q + r);
  two_prod(a.z, b.z, h, r);
  two_sum(p, h, p, q);
  s = s + (q + r);
  two_sum(p, s, hi, lo);
}

// The per-face constants of the Jacobian, shared by every pixel and view of the face (64 bytes:
// four 16-byte loads).  den < 0 marks a face the reference's early exit always takes.
struct alignas(16) BaryFace {
  V3 e1, e2;     // edges v1 - v0, v2 - v0 (binary32)
  V3 Nh, Nl;     // e1 x e2 to twice the working precision, from the exact edges
  float den;     // Nh . Nh, or -1 for a degenerate face
  V3 nN;         // unit binary32 normal normalize(e1 x e2), for the early exit's angle test
};
static_assert(sizeof(BaryFace) == 64, "four 16-byte loads");

GVV_HD BaryFace bary_face_setup(V3 v0, V3 v1, V3 v2) {
  BaryFace f;
  // the reference's early-exit quantities in binary32 (decisions)
  const V3 e1 = v1 - v0, e2 = v2 - v0;
  const V3 N = cross(e1, e2);
  const float denf = dot(N, N);
  const bool degenerate = fabsf(denf * denf) < 0.001f;
  f.nN = degenerate ? v3(0.f, 0.f, 0.f) : normalize(N);
  // the edges exactly (e = eh + el; eh is the binary32 edge) and their cross product to twice
  // the working precision: the plane's tilt enters the hit as |q0| |dN| / (|N| cos), so N must
  // be far more accurate than binary32 for the hit to be accurate to binary32
  V3 e1l, e2l;
  two_sum(v1.x, -v0.x, f.e1.x, e1l.x);
  two_sum(v1.y, -v0.y, f.e1.y, e1l.y);
  two_sum(v1.z, -v0.z, f.e1.z, e1l.z);
  two_sum(v2.x, -v0.x, f.e2.x, e2l.x);
  two_sum(v2.y, -v0.y, f.e2.y, e2l.y);
  two_sum(v2.z, -v0.z, f.e2.z, e2l.z);
  cross2(f.e1, e1l, f.e2, e2l, f.Nh, f.Nl);
  f.den = degenerate ? -1.f : dot(f.Nh, f.Nh);
  return f;
}

// The plane hit of a ray from o along d on the face with constants f and first vertex v0.
GVV_HD BaryGradSetup bary_grad_pixel(const BaryFace& f, V3 o, V3 d, V3 v0) {
  BaryGradSetup g;
  g.live = false;
  // the reference's binary32 early exit (a decision): grazing ray or degenerate face
  if (f.den < 0.f || fabsf(dot(normalize(d), f.nN)) < 0.001f) return g;
  g.live = true;
  g.e1 = f.e1;
  g.e2 = f.e2;
  g.N = f.Nh;
  g.den = f.den;
  const V3 Nl = f.Nl;
  // w = v0 - o exactly, as wh + wl
  V3 wh, wl;
  two_sum(v0.x, -o.x, wh.x, wl.x);
  two_sum(v0.y, -o.y, wh.y, wl.y);
  two_sum(v0.z, -o.z, wh.z, wl.z);
  // numerator w.N and denominator d.N, compensated
  float nh, nl, dh, dl;
  dot2(wh, g.N, nh, nl);
  nl = nl + (dot(wl, g.N) + dot(wh, Nl));
  dot2(d, g.N, dh, dl);
  dl = dl + dot(d, Nl);
  // t = (nh + nl) / (dh + dl) as th + tl
  const float th = nh / dh;
  float ph, pl;
  two_prod(th, dh, ph, pl);
  const float tl = (((nh - ph) - pl) + (nl - th * dl)) / dh;
  // q0 = t d - w, component by component from the exact residuals
  float q[3];
  const float dv[3] = {d.x, d.y, d.z}, whv[3] = {wh.x, wh.y, wh.z}, wlv[3] = {wl.x, wl.y, wl.z};
  for (int k = 0; k < 3; ++k) {
    float p, pe;
    two_prod(th, dv[k], p, pe);
    float s, se;
    two_sum(p, -whv[k], s, se);
    q[k] = s + (se + (pe + (fmaf(tl, dv[k], -wlv[k]))));
  }
  g.q0 = v3(q[0], q[1], q[2]);
  g.nd = dh + dl;
  return g;
}

// The plane hit and the triangle constants of the Jacobian for a ray from o along d.
GVV_HD BaryGradSetup bary_grad_setup(V3 o, V3 d, V3 v0, V3 v1, V3 v2) {
  return bary_grad_pixel(bary_face_setup(v0, v1, v2), o, d, v0);
}

// Everything the backward's pixel kernel needs of one face besides its vertex ids: the scaled
// ray/triangle constants that give the barycentrics, and the Jacobian's face constants (one
// 128-byte line per face).
struct alignas(16) FaceGrad {
  TriSetup t;
  BaryFace g;
};
static_assert(sizeof(FaceGrad) == 128, "one 128-byte line per face");

// Arithmetic of continuous gradient terms.  Binary32 rounding is part of the renderer's
// contract only for the visibility and decision values (barycentrics, texel indices, the
// normal flip, early exits); terms that only flow into gradients are checked against the FP64
// oracle with a relative tolerance, so in this type's operations products and sums may fuse
// into FMAs.
#pragma clang fp contract(fast)
struct F3 { float x, y, z; };
GVV_HD F3 f3(V3 v) { F3 r; r.x = v.x; r.y = v.y; r.z = v.z; return r; }
GVV_HD F3 f3(float x, float y, float z) { F3 r; r.x = x; r.y = y; r.z = z; return r; }
GVV_HD F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
GVV_HD F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
GVV_HD F3 operator-(F3 a) { return f3(-a.x, -a.y, -a.z); }
GVV_HD F3 operator*(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
GVV_HD F3 operator*(F3 a, float s) { return f3(s * a.x, s * a.y, s * a.z); }
GVV_HD float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
GVV_HD F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// Adds gb . d(a,b,c)/d(v0,v1,v2) (a 9-vector: vertex-major, xyz) to out[0..8].  inv_den and
// inv_nd are reciprocals of g.den and g.nd (the caller's choice of reciprocal).
//
// The three vertices' terms share their cross products.  Every A_v x X is linear in A_v, and
// A_0 = e1 - e2, A_1 = e2, A_2 = -e1, so the A x Cw, A x q0 (through dt_v) and A x N terms
// of all three come from e1 x M and e2 x M with M = (Cw - s q0) / |N|^2 + c2 N, s = NEdw / N.d.
// With K_r = q_r x N the edge and vertex terms collapse to Q_0 - D_0 = w1 K_2,
// Q_1 - D_1 = -w0 K_2 and Q_2 - D_2 = w0 K_1 - w1 K_0 (q1 - E0 = q2, q2 - E1 = q0).
// NEdw = N . (Ew x d) with Ew = w0 E0 + w1 E1.
GVV_HD void bary_grad_apply(const BaryGradSetup& g, V3 dv, V3 gb, float inv_den, float inv_nd, float* out) {
  const F3 e1 = f3(g.e1), e2 = f3(g.e2), N = f3(g.N), q0 = f3(g.q0), d = f3(dv);
  const F3 q1 = q0 - e1, q2 = q0 - e2;
  const F3 E0 = e2 - e1, E1 = -e2;
  const float w0 = gb.x - gb.z, w1 = gb.y - gb.z;
  const F3 Cw = w0 * cross(E0, q1) + w1 * cross(E1, q2);
  const float NCw = dot(N, Cw);
  const float NEdw = dot(N, cross(w0 * E0 + w1 * E1, d));
  const float c2 = -2.f * NCw * inv_den * inv_den;
  const float s = NEdw * inv_nd;
  const F3 M = inv_den * (Cw - s * q0) + c2 * N;
  const F3 X1 = cross(e1, M), X2 = cross(e2, M);
  const F3 K0 = cross(q0, N), K1 = cross(q1, N), K2 = cross(q2, N);
  const F3 g0 = (X1 - X2) + inv_den * (w1 * K2 + s * N);
  const F3 g1 = X2 - (inv_den * w0) * K2;
  const F3 g2 = inv_den * (w0 * K1 - w1 * K0) - X1;
  out[0] += g0.x; out[1] += g0.y; out[2] += g0.z;
  out[3] += g1.x; out[4] += g1.y; out[5] += g1.z;
  out[6] += g2.x; out[7] += g2.y; out[8] += g2.z;
}
#pragma clang fp contract(off)

}  // namespace gvv
