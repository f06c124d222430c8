// CPU check of the depth pass's screen-space pre-test (gvv_raster.hpp): for many random
// projected triangles -- slivers, near-horizontal and near-vertical edges, tiny and huge ones,
// partly off-screen -- every bounding-box pixel that the edge pre-test passes must lie inside
// its row's span (row_span), so that walking spans instead of the box drops nothing.  Also
// reports how much of the box the spans cover.  Runs on the host only.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../gvv-differentiable-cuda-renderer_amd/csrc/gvv_raster.hpp"

using namespace gvv;

int main(int argc, char** argv) {
  const int n = argc > 1 ? atoi(argv[1]) : 20000;
  std::mt19937 rng(12345);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  const int W = 640, H = 480;
  long long passed = 0, box_px = 0, span_px = 0, missed = 0, tested = 0;
  for (int t = 0; t < n; ++t) {
    float4 p[3];
    const int kind = t % 5;
    const float cx = -100.f + U(rng) * (W + 200.f), cy = -100.f + U(rng) * (H + 200.f);
    for (int i = 0; i < 3; ++i) {
      float sx = 1.f, sy = 1.f;
      if (kind == 0) { sx = sy = 3.f; }                          // tiny
      else if (kind == 1) { sx = sy = 40.f; }                    // medium
      else if (kind == 2) { sx = 400.f; sy = 400.f; }            // huge, partly off-screen
      else if (kind == 3) { sx = 300.f; sy = 2.f; }              // horizontal sliver
      else { sx = 2.f; sy = 300.f; }                             // vertical sliver
      p[i] = make_float4(cx + (U(rng) - 0.5f) * sx, cy + (U(rng) - 0.5f) * sy, 1.f + 10.f * U(rng), 0.f);
    }
    if (t % 7 == 0) p[1].y = p[0].y;                             // exactly horizontal edge
    if (t % 11 == 0) p[2].x = p[0].x;                            // exactly vertical edge
    const Box bx = face_box(v3(p[0].x, p[0].y, p[0].z), v3(p[1].x, p[1].y, p[1].z), v3(p[2].x, p[2].y, p[2].z), W, H);
    if (bx.x0 > bx.x1 || bx.y0 > bx.y1) continue;
    const EdgeTest e = edge_test(p[0], p[1], p[2]);
    ++tested;
    for (int y = bx.y0; y <= bx.y1; ++y) {
      int xl, xr;
      row_span(e, y, bx.x0, bx.x1, xl, xr);
      if (xr >= xl) span_px += xr - xl + 1;
      for (int x = bx.x0; x <= bx.x1; ++x) {
        ++box_px;
        if (!e.pass(x + 0.5f, y + 0.5f)) continue;
        ++passed;
        if (x < xl || x > xr) {
          if (missed < 5) printf("MISS tri %d pixel (%d,%d) span [%d,%d]\n", t, x, y, xl, xr);
          ++missed;
        }
      }
    }
  }
  printf("triangles %lld box_px %lld span_px %lld passed %lld missed %lld\n", tested, box_px, span_px, passed, missed);
  return missed == 0 ? 0 : 1;
}
