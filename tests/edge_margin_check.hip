// How far outside the depth pass's screen-space edges (before the pre-test's margin) can a
// pixel lie that the exact ray/triangle test accepts?  Random triangles of 0.3-100 mm, a third
// of them slivers, seen by pinhole cameras (f = 1.2 W, 1024^2) from 0.6-5 m at every
// orientation; prints the most negative raw edge distance of an accepted pixel, which the
// pre-test's margin (gvv_raster.hpp kEdgeMargin) must exceed.  Boxes of up to 64 pixels are
// also walked as the depth pass walks them (EdgeWalk, edge values stepped along the rows), and
// every pixel the exact test accepts but the walk would reject is counted (walk_missed).
// Host only.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <cmath>
#include "../gvv-differentiable-cuda-renderer_amd/csrc/gvv_raster.hpp"
using namespace gvv;
static void look_at(const double pos[3], double E[12]) {
  double z[3] = {-pos[0], -pos[1], -pos[2]}; double n = std::sqrt(z[0]*z[0]+z[1]*z[1]+z[2]*z[2]);
  for (int i = 0; i < 3; ++i) z[i] /= n;
  double up[3] = {0, 1, 0}; if (std::fabs(z[1]) > 0.99) { up[0] = 1; up[1] = 0; }
  double x[3] = {up[1]*z[2]-up[2]*z[1], up[2]*z[0]-up[0]*z[2], up[0]*z[1]-up[1]*z[0]};
  n = std::sqrt(x[0]*x[0]+x[1]*x[1]+x[2]*x[2]); for (int i = 0; i < 3; ++i) x[i] /= n;
  double y[3] = {z[1]*x[2]-z[2]*x[1], z[2]*x[0]-z[0]*x[2], z[0]*x[1]-z[1]*x[0]};
  const double* R[3] = {x, y, z};
  for (int r = 0; r < 3; ++r) { for (int c = 0; c < 3; ++c) E[r*4+c] = R[r][c]; E[r*4+3] = -(R[r][0]*pos[0]+R[r][1]*pos[1]+R[r][2]*pos[2]); }
}
int main(int argc, char** argv) {
  const int ntri = argc > 1 ? atoi(argv[1]) : 20000;
  std::mt19937 rng(7);
  std::uniform_real_distribution<double> U(0, 1);
  std::normal_distribution<double> G(0, 1);
  const int W = 1024, H = 1024;
  double worst = 1e9; long long accepted = 0, tested = 0;
  double worst_size = 0;
  long long walk_boxes = 0, walk_missed = 0;
  double walk_dev = 0;
  for (int t = 0; t < ntri; ++t) {
    double pos[3] = {G(rng), G(rng) * 0.3, G(rng)}; double n = std::sqrt(pos[0]*pos[0]+pos[1]*pos[1]+pos[2]*pos[2]);
    const double dist = (t % 4 == 0 ? 600.0 : 2000.0) + 3000.0 * U(rng);
    for (int i = 0; i < 3; ++i) pos[i] = pos[i] / n * dist;
    double Ed[12]; look_at(pos, Ed);
    float E[12], K[9] = {1.2f * W, 0, W / 2.f, 0, 1.2f * W, H / 2.f, 0, 0, 1};
    for (int i = 0; i < 12; ++i) E[i] = (float)Ed[i];
    Camera cam; camera_setup(E, K, cam);
    const double size = std::pow(10.0, -0.5 + 2.5 * U(rng));      // 0.3 .. 100 mm
    double c[3] = {G(rng) * 300, G(rng) * 300, G(rng) * 300};
    V3 v[3]; float4 p[3];
    const int kind = t % 3;
    for (int k = 0; k < 3; ++k) {
      double s = size, s2 = kind == 2 ? size * 0.02 : size;          // slivers
      v[k] = v3((float)(c[0] + G(rng) * s), (float)(c[1] + G(rng) * (k == 2 ? s2 : s)), (float)(c[2] + G(rng) * s));
      const V3 q = project(cam, v[k]);
      p[k] = make_float4(q.x, q.y, q.z, 0.f);
    }
    const Box bx = face_box(v3(p[0].x, p[0].y, p[0].z), v3(p[1].x, p[1].y, p[1].z), v3(p[2].x, p[2].y, p[2].z), W, H);
    if (bx.x0 > bx.x1 || bx.y0 > bx.y1) continue;
    if ((long long)(bx.x1 - bx.x0 + 1) * (bx.y1 - bx.y0 + 1) > 40000) continue;
    const EdgeTest et = edge_test(p[0], p[1], p[2]);
    if (!et.on) continue;
    const TriSetup ts = tri_setup(v[0], v[1], v[2]);
    const int bw = bx.x1 - bx.x0 + 1;
    const bool walked = (long long)bw * (bx.y1 - bx.y0 + 1) <= 64;
    EdgeWalk walk;
    walk.start(et, bx.x0, bx.y0, bw);
    if (walked) ++walk_boxes;
    for (int y = bx.y0; y <= bx.y1; ++y)
      for (int x = bx.x0; x <= bx.x1; ++x) {
        ++tested;
        const bool walk_pass = walk.pass(et.on);
        if (walked) {
          for (int i = 0; i < 3; ++i) {
            const double dev = std::fabs((double)walk.e[i] - ((double)et.a[i] * (x + 0.5) + (double)et.b[i] * (y + 0.5) + (double)et.c[i]));
            if (dev > walk_dev) walk_dev = dev;
          }
        }
        walk.step(et, x == bx.x1);
        const V3 d = pixel_ray(cam, x + 0.5f, y + 0.5f);
        const V3 abc = ray_barycentric(ts, cam_origin_scaled(cam), d);
        if (!inside(abc)) continue;
        ++accepted;
        if (walked && !walk_pass) ++walk_missed;
        for (int i = 0; i < 3; ++i) {
          const double e = (double)et.a[i] * (x + 0.5) + (double)et.b[i] * (y + 0.5) + (double)et.c[i] - kEdgeMargin;
          if (e < worst) { worst = e; worst_size = size; }
        }
      }
  }
  printf("tested %lld accepted %lld worst %.6e margin %.6e size %.2f walk_boxes %lld walk_missed %lld walk_dev %.3e\n",
         tested, accepted, worst, (double)kEdgeMargin, worst_size, walk_boxes, walk_missed, walk_dev);
  return 0;
}
