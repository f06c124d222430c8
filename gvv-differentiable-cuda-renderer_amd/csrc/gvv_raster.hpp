// gvv_raster.hpp -- screen-space pre-test of the depth pass: the conservative edge test and
// the per-row spans it implies.  Host- and device-callable so that its conservativeness can be
// checked on the CPU.
#pragma once
#include "gvv_math.hpp"

namespace gvv {

// Conservative screen-space pre-test.  A pixel passes the exact ray/triangle test only if its
// centre lies inside the projected triangle up to the float error of that test (ray/triangle
// barycentrics below zero are rejected outright), so every pixel farther than kEdgeMargin px
// outside an edge can be skipped without changing any result.  Over 2e8 pixel tests of random
// triangles at the renderer's scale the farthest accepted pixel lay 7e-4 px outside an edge
// (tests/edge_margin_check.hip), so 1/32 px leaves a factor of 45.  The pre-test is disabled
// (every bounding-box pixel takes the exact test) when a vertex is near or behind the camera
// plane, when coordinates are huge, or when the projected triangle is degenerate.
constexpr float kEdgeMargin = 0.03125f;

struct EdgeTest {
  float a[3], b[3], c[3];
  bool on;
  GVV_HD bool pass(float x, float y) const {
    if (!on) return true;
    return (a[0] * x + b[0] * y + c[0] >= 0.f) && (a[1] * x + b[1] * y + c[1] >= 0.f) &&
           (a[2] * x + b[2] * y + c[2] >= 0.f);
  }
};

GVV_HD EdgeTest edge_test(float4 p0, float4 p1, float4 p2) {
  EdgeTest e;
  e.on = false;
  const float px[3] = {p0.x, p1.x, p2.x}, py[3] = {p0.y, p1.y, p2.y}, pz[3] = {p0.z, p1.z, p2.z};
  for (int i = 0; i < 3; ++i)
    if (!(pz[i] >= 1.0f) || !(fabsf(px[i]) <= 1e6f) || !(fabsf(py[i]) <= 1e6f)) return e;
  const float area2 = (px[1] - px[0]) * (py[2] - py[0]) - (py[1] - py[0]) * (px[2] - px[0]);
  if (!(fabsf(area2) >= 1e-3f)) return e;
  const float s = area2 > 0.f ? 1.f : -1.f;
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3;
    const float ex = px[j] - px[i], ey = py[j] - py[i];
    const float len = sqrtf(ex * ex + ey * ey);
    if (!(len > 0.f)) return e;
    // signed distance of (x, y) to edge i, positive inside: s * cross(e, p - p_i) / |e|
    e.a[i] = -s * ey / len;
    e.b[i] = s * ex / len;
    e.c[i] = -(e.a[i] * px[i] + e.b[i] * py[i]) + kEdgeMargin;
  }
  e.on = true;
  return e;
}

// The pre-test walked along the rows of a box: the three edge values at the current pixel are
// kept and advanced by a (one pixel along a row) or, at a row end, by b - (bw - 1) a (back to
// the row's first pixel, one row down), instead of being evaluated anew at every pixel.  The
// stepped values differ from the direct ones by rounding only (at most 64 steps of values
// under 100 px), far inside the margin; tests/edge_margin_check.hip runs this walk over every
// box it draws and counts the accepted pixels that it would reject (none).
struct EdgeWalk {
  float e[3];     // edge values at the current pixel
  float w[3];     // row-end step b - (bw - 1) a
  GVV_HD void start(const EdgeTest& t, int x0, int y0, int bw) {
    const float x = x0 + 0.5f, y = y0 + 0.5f;
    for (int i = 0; i < 3; ++i) {
      e[i] = t.a[i] * x + t.b[i] * y + t.c[i];
      w[i] = t.b[i] - (float)(bw - 1) * t.a[i];
    }
  }
  GVV_HD bool pass(bool on) const { return !on || (e[0] >= 0.f && e[1] >= 0.f && e[2] >= 0.f); }
  GVV_HD void step(const EdgeTest& t, bool row_end) {
    for (int i = 0; i < 3; ++i) e[i] += row_end ? w[i] : t.a[i];
  }
};

// Conservative x-range [xl, xr] of bounding-box row y (pixel centres at y + 0.5) inside the
// three half-planes of the edge pre-test, widened by one pixel on each side so that float
// rounding of the bounds never drops a pixel the pre-test would pass.  Without a usable edge
// test the whole row is returned.
GVV_HD void row_span(const EdgeTest& e, int y, int x0, int x1, int& xl, int& xr) {
  xl = x0;
  xr = x1;
  if (!e.on) return;
  const float yc = y + 0.5f;
  float lo = (float)x0, hi = (float)x1;    // range of pixel indices x (centre x + 0.5)
  for (int i = 0; i < 3; ++i) {
    const float rhs = -(e.b[i] * yc + e.c[i]);      // need a * xc >= rhs
    if (e.a[i] > 0.f) lo = fmaxf(lo, rhs / e.a[i] - 1.5f);
    else if (e.a[i] < 0.f) hi = fminf(hi, rhs / e.a[i] + 0.5f);
    else if (rhs > 1.f) hi = lo - 1.f;              // horizontal edge with the row well outside
  }
  if (!(lo <= hi)) { xr = x0 - 1; return; }         // empty (also catches NaN)
  xl = (int)floorf(lo);
  xr = (int)ceilf(hi);
  if (xl < x0) xl = x0;
  if (xr > x1) xr = x1;
}

}  // namespace gvv
