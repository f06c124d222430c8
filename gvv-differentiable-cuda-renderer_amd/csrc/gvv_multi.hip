// gvv_multi.hip -- multi-mesh runtime: many meshes of uneven size in one hipGraph launch, with
// an autotuner over the per-mesh launch settings (C ABI in include/gvv_multi.h).
//
// Each mesh keeps its own renderer handle (topology, CSR, scratch).  The one thing the meshes
// share is the large-face queue: every mesh's depth pass (k_raster) appends its faces with big
// bounding boxes, tagged with the mesh index, to one queue, and one persistent launch
// (k_raster_large_multi) drains the large faces of all meshes together, so a mesh with few
// large faces leaves no idle tail and the launch count does not grow with the mesh count.
// A run plans every mesh on the host (this is where scratch grows, outside any capture), then
// either launches the stages in order or captures them once into a hipGraph of three stages:
// the per-mesh depth passes on parallel branches (one per hardware queue, meshes
// longest-first onto the least-loaded branch, LPT), the shared large-face launch, then the
// per-mesh resolve and backward chains on the branches again.  The graph is replayed while
// the io table and the settings are unchanged.
//
// The reference renders one topology per op and serialises the batch on the host
// (cpp/src/TensorflowOperators/CudaRenderer/CudaRenderer.cpp:23, :309-328); nothing here has a
// counterpart there.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gvv_handle.hpp"
#include "../../include/gvv_multi.h"
#include "../../include/gvv_renderer.h"

namespace {

using namespace gvv;

// graph branches: twice the process's 4 hardware queues (32 uneven meshes, 103 Mpix per run:
// 7.75-7.84 ms against 7.97-8.00 with 4 branches, 8.0-8.1 with 12 or 16, 9.2 with 2)
constexpr int kLanes = 8;
constexpr int kTiles[3] = {8, 16, 32};          // backward pixel-tile edges
constexpr int kLargeBoxes[3] = {64, 256, 1024}; // raster large-face thresholds (pixels)

int fail(const std::string& m) { return set_last_error(m); }

#define MULTI_CHECK(expr)                                                            \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// Size class: triangle-count bucket (<8k, <32k, <128k, larger) x pixels per view
// (<=256^2, <=512^2, larger).
int size_class(const Handle* h) {
  const int tb = h->F < 8192 ? 0 : h->F < 32768 ? 1 : h->F < 131072 ? 2 : 3;
  const int64_t px = (int64_t)h->W * h->H;
  const int pb = px <= 256 * 256 ? 0 : px <= 512 * 512 ? 1 : 2;
  return tb * 3 + pb;
}

// The per-stage argument tables of a batched run (one launch per stage across all meshes):
// for each stage the meshes' argument records and the prefix of their block counts, packed
// into one device buffer (uploaded when it changes).
enum Stage { kFwdSetup, kVertices, kRaster, kResolve, kBwdSetup, kFinish, kStages };
struct StageTables {
  std::vector<unsigned char> host;
  DevBuf<unsigned char> dev;
  size_t off_args[kStages] = {}, off_start[kStages] = {};
  int n[kStages] = {};
  unsigned blocks[kStages] = {};
  template <class A>
  MeshTable<A> table(int st) const {
    MeshTable<A> t;
    t.args = reinterpret_cast<const A*>(dev.p + off_args[st]);
    t.start = reinterpret_cast<const unsigned*>(dev.p + off_start[st]);
    t.n = n[st];
    return t;
  }
};

struct Multi {
  std::vector<Handle*> meshes;
  std::vector<int> cls;
  int albedo = 0, shading = 0, device = 0;
  bool use_graph = true;
  bool batched = false;                         // one launch per stage across meshes
  StageTables stages;
  hipStream_t lanes[kLanes] = {};
  hipEvent_t fork[2] = {};
  hipEvent_t join[2][kLanes] = {};
  // the shared large-face queue and the raster state of every mesh it names
  DevBuf<unsigned int> counters;                // [2]: queue length, (unused)
  DevBuf<int4> large;                           // sum over meshes of B*C*F entries
  DevBuf<RasterMesh> table;                     // [meshes]
  std::vector<RasterMesh> host_table;           // what `table` holds
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<char> key;      // what the graph was captured for
  std::mutex mu;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key.clear();
  }
  ~Multi() {
    drop_graph();
    for (auto s : lanes)
      if (s) (void)hipStreamDestroy(s);
    for (int st = 0; st < 2; ++st) {
      if (fork[st]) (void)hipEventDestroy(fork[st]);
      for (auto e : join[st])
        if (e) (void)hipEventDestroy(e);
    }
    counters.release();
    large.release();
    table.release();
    stages.dev.release();
    for (auto* h : meshes) gvv_destroy(h);
  }
};

double mesh_cost(const Handle* h, int B) {
  return (double)B * h->C * ((double)h->F + 0.5 * (double)h->W * h->H);
}

int plan_mesh(Multi* m, int i, const gvv_mesh_io& io, bool backward, hipStream_t s, ForwardPlan& fp, BackwardPlan& bp) {
  Handle* h = m->meshes[i];
  const std::string where = "mesh " + std::to_string(i) + ": ";
  if (plan_forward(h, io.batch, io.tex_h, io.tex_w, io.vpos, io.vcol, io.tex, io.sh, io.extr, io.intr, io.bary,
                   io.face, io.render, io.vnormal, nullptr, s, fp))
    return fail(where + gvv_last_error());
  if (backward) {
    if (!io.g_render) return fail(where + "g_render is null");
    if (plan_backward(h, io.batch, io.tex_h, io.tex_w, io.g_render, nullptr, nullptr, io.vpos, io.vcol, io.tex, io.sh,
                      io.vnormal, nullptr, io.face, io.extr, io.intr, io.g_pos, io.g_col, io.g_tex, io.g_sh, bp))
      return fail(where + gvv_last_error());
  }
  return 0;
}

hipError_t enqueue_mesh(const ForwardPlan& fp, const BackwardPlan& bp, bool backward, hipStream_t s) {
  hipError_t e = launch_forward(fp, s);
  if (e == hipSuccess && backward) e = launch_backward(bp, s);
  return e;
}

hipError_t enqueue_tail(const ForwardPlan& fp, const BackwardPlan& bp, bool backward, hipStream_t s) {
  hipError_t e = launch_forward_resolve(fp, s);
  if (e == hipSuccess && backward) e = launch_backward(bp, s);
  return e;
}

int check_device(const Multi* m) {
  int dev = -1;
  MULTI_CHECK(hipGetDevice(&dev));
  if (dev != m->device) return fail("runtime was created on device " + std::to_string(m->device));
  return 0;
}

template <class A>
void put_stage(StageTables& st, std::vector<unsigned char>& blob, int stage, const std::vector<A>& args,
               const std::vector<unsigned>& blocks) {
  auto align = [&](size_t a) { blob.resize((blob.size() + a - 1) / a * a); };
  align(256);
  st.off_args[stage] = blob.size();
  const unsigned char* a = reinterpret_cast<const unsigned char*>(args.data());
  blob.insert(blob.end(), a, a + sizeof(A) * args.size());
  align(256);
  st.off_start[stage] = blob.size();
  std::vector<unsigned> start(args.size() + 1, 0u);
  for (size_t i = 0; i < args.size(); ++i) start[i + 1] = start[i] + blocks[i];
  const unsigned char* b = reinterpret_cast<const unsigned char*>(start.data());
  blob.insert(blob.end(), b, b + sizeof(unsigned) * start.size());
  st.n[stage] = (int)args.size();
  st.blocks[stage] = start.back();
}

// The per-stage tables of a batched run from the meshes' plans (uploaded on `s` when changed).
int build_stages(Multi* m, const std::vector<ForwardPlan>& fp, const std::vector<BackwardPlan>& bp, bool backward,
                 hipStream_t s) {
  const int n = (int)fp.size();
  StageTables& st = m->stages;
  std::vector<unsigned char> blob;
  std::vector<SetupGrid> setup(n);
  std::vector<VertexArgs> vert(n);
  std::vector<RasterArgs> rast(n);
  std::vector<ResolveArgs> res(n);
  std::vector<unsigned> b_setup(n), b_vert(n), b_rast(n), b_res(n);
  for (int i = 0; i < n; ++i) {
    setup[i] = setup_grid(forward_setup_args(fp[i]));
    b_setup[i] = (unsigned)(setup[i].cam_blocks + setup[i].face_blocks + setup[i].fill_blocks);
    vert[i] = vertex_args(fp[i]);
    b_vert[i] = (unsigned)(((int64_t)fp[i].B * fp[i].C * fp[i].N + 255) / 256);
    rast[i] = raster_args(fp[i]);
    b_rast[i] = (unsigned)(((int64_t)fp[i].B * fp[i].C * fp[i].F + 63) / 64);
    res[i] = fp[i].r;
    b_res[i] = resolve_blocks(fp[i].r);
  }
  put_stage(st, blob, kFwdSetup, setup, b_setup);
  put_stage(st, blob, kVertices, vert, b_vert);
  put_stage(st, blob, kRaster, rast, b_rast);
  put_stage(st, blob, kResolve, res, b_res);
  for (int k = kBwdSetup; k < kStages; ++k) { st.n[k] = 0; st.blocks[k] = 0; }
  if (backward) {
    std::vector<SetupGrid> bsetup(n);
    std::vector<unsigned> b_bsetup(n);
    std::vector<FinishArgs> fin;
    std::vector<unsigned> b_fin;
    for (int i = 0; i < n; ++i) {
      bsetup[i] = setup_grid(backward_setup_args(bp[i]));
      b_bsetup[i] = (unsigned)(bsetup[i].cam_blocks + bsetup[i].face_blocks + bsetup[i].fill_blocks);
      if (bp[i].shaded) {
        fin.push_back(finish_args(bp[i].a));
        b_fin.push_back(finish_blocks(fin.back(), bp[i].a.C));
      }
    }
    put_stage(st, blob, kBwdSetup, bsetup, b_bsetup);
    put_stage(st, blob, kFinish, fin, b_fin);
  }
  if (blob != st.host || !st.dev.p) {
    // the previous contents may still be read by queued work; the copy source must outlive it
    MULTI_CHECK(hipStreamSynchronize(s));
    MULTI_CHECK(st.dev.reserve(blob.size()));
    st.host = std::move(blob);
    MULTI_CHECK(hipMemcpyAsync(st.dev.p, st.host.data(), st.host.size(), hipMemcpyHostToDevice, s));
  }
  return 0;
}

// One launch per stage across all meshes: clear the queue, setup, vertices, depth pass, the
// shared large-face launch, resolve; then the backward's setup, pixel kernels and finishing
// reductions.  The pixel kernels are the one stage still launched per mesh (through `pixel`,
// which may fan them out over the graph's branches): across meshes they would keep the mesh's
// arguments in registers the single-mesh kernel reloads from its argument segment, and spill.
template <class PixelStage>
hipError_t enqueue_batched(Multi* m, bool backward, hipStream_t s, PixelStage&& pixel) {
  const StageTables& st = m->stages;
  const bool shade = (m->shading == kShaded) && m->albedo != kNormal;
  hipError_t e = launch_clear_counters(m->counters.p, s);
  if (e == hipSuccess) e = launch_setup_multi(st.table<SetupGrid>(kFwdSetup), st.blocks[kFwdSetup], s);
  if (e == hipSuccess) e = launch_vertices_multi(st.table<VertexArgs>(kVertices), st.blocks[kVertices], s);
  if (e == hipSuccess) e = launch_raster_multi(st.table<RasterArgs>(kRaster), st.blocks[kRaster], s);
  if (e == hipSuccess) e = launch_raster_large_multi(m->table.p, m->counters.p, m->large.p, s);
  if (e == hipSuccess) e = launch_resolve_multi(st.table<ResolveArgs>(kResolve), st.blocks[kResolve], m->albedo, shade, s);
  if (!backward) return e;
  if (e == hipSuccess) e = launch_setup_multi(st.table<SetupGrid>(kBwdSetup), st.blocks[kBwdSetup], s);
  if (e == hipSuccess) e = pixel();
  if (e == hipSuccess) e = launch_grad_finish_multi(st.table<FinishArgs>(kFinish), st.blocks[kFinish], s);
  return e;
}

// Plans every mesh onto the shared large-face queue and brings the device copy of the mesh
// table up to date (on `s`, after the work already queued there, which may still read it).
int plan_all(Multi* m, const gvv_mesh_io* io, bool backward, hipStream_t s, std::vector<ForwardPlan>& fp,
             std::vector<BackwardPlan>& bp) {
  const int n = (int)m->meshes.size();
  fp.resize(n);
  bp.resize(n);
  size_t queue = 0;
  for (int i = 0; i < n; ++i) {
    if (plan_mesh(m, i, io[i], backward, s, fp[i], bp[i])) return 1;
    queue += (size_t)io[i].batch * m->meshes[i]->C * m->meshes[i]->F;
  }
  if (m->large.n < queue || !m->counters.p || !m->table.p) MULTI_CHECK(hipStreamSynchronize(s));
  MULTI_CHECK(m->counters.reserve(2));
  MULTI_CHECK(m->large.reserve(queue));
  MULTI_CHECK(m->table.reserve(n));
  std::vector<RasterMesh> t(n);
  for (int i = 0; i < n; ++i) {
    fp[i].counters = m->counters.p;
    fp[i].large = m->large.p;
    fp[i].mesh_id = i;
    fp[i].shared_queue = true;
    t[i] = raster_mesh(fp[i]);
  }
  if (m->host_table.size() != t.size() ||
      memcmp(m->host_table.data(), t.data(), t.size() * sizeof(RasterMesh)) != 0) {
    // the previous contents may still be read by queued work; the copy source must outlive it
    MULTI_CHECK(hipStreamSynchronize(s));
    m->host_table = t;
    MULTI_CHECK(hipMemcpyAsync(m->table.p, m->host_table.data(), t.size() * sizeof(RasterMesh),
                               hipMemcpyHostToDevice, s));
  }
  if (m->batched) return build_stages(m, fp, bp, backward, s);
  return 0;
}

// Every launch of a run in stream order on one stream: clear the queue, the depth passes, the
// shared large-face launch, then each mesh's resolve and backward.
hipError_t enqueue_all(Multi* m, const std::vector<ForwardPlan>& fp, const std::vector<BackwardPlan>& bp,
                       bool backward, hipStream_t s) {
  hipError_t e = launch_clear_counters(m->counters.p, s);
  for (size_t i = 0; i < fp.size() && e == hipSuccess; ++i) e = launch_forward_raster(fp[i], s);
  if (e == hipSuccess) e = launch_raster_large_multi(m->table.p, m->counters.p, m->large.p, s);
  for (size_t i = 0; i < fp.size() && e == hipSuccess; ++i) e = enqueue_tail(fp[i], bp[i], backward, s);
  return e;
}

std::vector<char> graph_key(const Multi* m, const gvv_mesh_io* io, bool backward) {
  const size_t n = m->meshes.size();
  const size_t head = n * sizeof(gvv_mesh_io) + 1;
  std::vector<char> k(head + n * 2 * sizeof(int) + 3 * sizeof(void*));
  memcpy(k.data(), io, n * sizeof(gvv_mesh_io));
  k[n * sizeof(gvv_mesh_io)] = backward ? 1 : 0;
  char* t = k.data() + head;
  for (size_t i = 0; i < n; ++i) {
    memcpy(t + 8 * i, &m->meshes[i]->tile, sizeof(int));
    memcpy(t + 8 * i + 4, &m->meshes[i]->large_box, sizeof(int));
  }
  // the shared queue's buffers and the stage tables are baked into the graph
  const void* bufs[3] = {m->counters.p, m->large.p, m->table.p};
  memcpy(t + 8 * n, bufs, sizeof(bufs));
  const void* stage_buf = m->batched ? (const void*)m->stages.dev.p : nullptr;
  const char* sb = reinterpret_cast<const char*>(&stage_buf);
  k.insert(k.end(), sb, sb + sizeof(stage_buf));
  k.push_back(m->batched ? 1 : 0);
  return k;
}

// Fans the meshes out over the lanes (longest first onto the least-loaded lane), each lane
// starting after event `fork` on lane 0 and joining back into lane 0.
template <class Enqueue>
hipError_t fan_out(Multi* m, const std::vector<int>& order, const gvv_mesh_io* io, int stage, Enqueue&& enqueue) {
  hipError_t e = hipEventRecord(m->fork[stage], m->lanes[0]);
  for (int k = 1; k < kLanes && e == hipSuccess; ++k) e = hipStreamWaitEvent(m->lanes[k], m->fork[stage], 0);
  double load[kLanes] = {};
  for (size_t j = 0; j < order.size() && e == hipSuccess; ++j) {
    const int i = order[j];
    int lane = 0;
    for (int k = 1; k < kLanes; ++k)
      if (load[k] < load[lane]) lane = k;
    load[lane] += mesh_cost(m->meshes[i], io[i].batch);
    e = enqueue(i, m->lanes[lane]);
  }
  for (int k = 1; k < kLanes && e == hipSuccess; ++k) {
    e = hipEventRecord(m->join[stage][k], m->lanes[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(m->lanes[0], m->join[stage][k], 0);
  }
  return e;
}

// Captures a run into one graph: clear the queue; depth passes fanned out over the lanes;
// the shared large-face launch; resolve + backward fanned out again.
int capture(Multi* m, const gvv_mesh_io* io, bool backward, const std::vector<ForwardPlan>& fp,
            const std::vector<BackwardPlan>& bp) {
  const int n = (int)m->meshes.size();
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    return mesh_cost(m->meshes[a], io[a].batch) > mesh_cost(m->meshes[b], io[b].batch);
  });
  m->drop_graph();
  MULTI_CHECK(hipStreamBeginCapture(m->lanes[0], hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  if (m->batched) {
    e = enqueue_batched(m, backward, m->lanes[0], [&]() {
      return fan_out(m, order, io, 1, [&](int i, hipStream_t s) { return launch_pixel_grad(bp[i], s); });
    });
  } else {
    e = launch_clear_counters(m->counters.p, m->lanes[0]);
    if (e == hipSuccess)
      e = fan_out(m, order, io, 0, [&](int i, hipStream_t s) { return launch_forward_raster(fp[i], s); });
    if (e == hipSuccess) e = launch_raster_large_multi(m->table.p, m->counters.p, m->large.p, m->lanes[0]);
    if (e == hipSuccess)
      e = fan_out(m, order, io, 1, [&](int i, hipStream_t s) { return enqueue_tail(fp[i], bp[i], backward, s); });
  }
  hipGraph_t g = nullptr;
  const hipError_t e_end = hipStreamEndCapture(m->lanes[0], &g);
  if (e != hipSuccess || e_end != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e_end));
  }
  m->graph = g;
  MULTI_CHECK(hipGraphInstantiate(&m->exec, g, nullptr, nullptr, 0));
  m->key = graph_key(m, io, backward);
  return 0;
}

}  // namespace

extern "C" {

int gvv_multi_create(const gvv_mesh_desc* meshes, int32_t n, int32_t albedo_mode, int32_t shading_mode,
                     void** handle) {
  if (!handle) return fail("handle pointer is null");
  *handle = nullptr;
  if (!meshes || n <= 0) return fail("need at least one mesh");
  if (albedo_mode != kVertexColor && albedo_mode != kTextured && albedo_mode != kForegroundMask)
    return fail("multi-mesh runtime supports the vertexColor, textured and foregroundMask modes");
  Multi* m = new Multi();
  m->albedo = albedo_mode;
  m->shading = shading_mode;
  for (int i = 0; i < n; ++i) {
    const gvv_mesh_desc& d = meshes[i];
    void* h = nullptr;
    if (gvv_create(d.faces, d.n_face_idx, d.texcoords, d.n_face_idx * 2, d.n_vertices, d.n_cameras, d.width,
                   d.height, albedo_mode, shading_mode, 1, 1, 0, &h)) {
      const std::string msg = "mesh " + std::to_string(i) + ": " + gvv_last_error();
      delete m;
      return fail(msg);
    }
    m->meshes.push_back((Handle*)h);
    m->cls.push_back(size_class((Handle*)h));
  }
  hipError_t e = hipGetDevice(&m->device);
  for (int k = 0; k < kLanes && e == hipSuccess; ++k) e = hipStreamCreateWithFlags(&m->lanes[k], hipStreamNonBlocking);
  for (int st = 0; st < 2; ++st) {
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->fork[st], hipEventDisableTiming);
    for (int k = 0; k < kLanes && e == hipSuccess; ++k)
      e = hipEventCreateWithFlags(&m->join[st][k], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    delete m;
    return fail(std::string("gvv_multi_create: ") + hipGetErrorString(e));
  }
  *handle = m;
  return 0;
}

void gvv_multi_destroy(void* handle) { delete (Multi*)handle; }

int gvv_multi_use_graph(void* handle, int32_t on) {
  Multi* m = (Multi*)handle;
  if (!m) return fail("null handle");
  std::lock_guard<std::mutex> lock(m->mu);
  m->use_graph = on != 0;
  return 0;
}

int gvv_multi_set_batched(void* handle, int32_t on) {
  Multi* m = (Multi*)handle;
  if (!m) return fail("null handle");
  std::lock_guard<std::mutex> lock(m->mu);
  m->batched = on != 0;
  m->drop_graph();
  return 0;
}

int gvv_multi_run(void* handle, const gvv_mesh_io* io, int32_t backward, void* stream) {
  Multi* m = (Multi*)handle;
  if (!m) return fail("null handle");
  if (!io) return fail("null io table");
  std::lock_guard<std::mutex> lock(m->mu);
  if (check_device(m)) return 1;
  hipStream_t s = (hipStream_t)stream;
  std::vector<ForwardPlan> fp;
  std::vector<BackwardPlan> bp;
  if (plan_all(m, io, backward != 0, s, fp, bp)) return 1;
  if (!m->use_graph) {
    if (m->batched) {
      MULTI_CHECK(enqueue_batched(m, backward != 0, s, [&]() {
        hipError_t e = hipSuccess;
        for (size_t i = 0; i < bp.size() && e == hipSuccess; ++i) e = launch_pixel_grad(bp[i], s);
        return e;
      }));
    } else {
      MULTI_CHECK(enqueue_all(m, fp, bp, backward != 0, s));
    }
    return 0;
  }
  if (!m->exec || m->key != graph_key(m, io, backward != 0))
    if (capture(m, io, backward != 0, fp, bp)) return 1;
  MULTI_CHECK(hipGraphLaunch(m->exec, s));
  return 0;
}

int gvv_multi_autotune(void* handle, const gvv_mesh_io* io, int32_t reps, void* stream) {
  Multi* m = (Multi*)handle;
  if (!m) return fail("null handle");
  if (!io) return fail("null io table");
  if (reps <= 0) return fail("reps must be > 0");
  std::lock_guard<std::mutex> lock(m->mu);
  if (check_device(m)) return 1;
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t a = nullptr, b = nullptr;
  MULTI_CHECK(hipEventCreate(&a));
  MULTI_CHECK(hipEventCreate(&b));
  int rc = 0;
  const int n = (int)m->meshes.size();
  for (int c = 0; c < 12 && rc == 0; ++c) {
    int rep = -1;
    for (int i = 0; i < n; ++i)
      if (m->cls[i] == c && (rep < 0 || mesh_cost(m->meshes[i], io[i].batch) > mesh_cost(m->meshes[rep], io[rep].batch)))
        rep = i;
    if (rep < 0) continue;
    Handle* h = m->meshes[rep];
    float best = 1e30f;
    int best_tile = h->tile, best_box = h->large_box;
    for (int ti = 0; ti < 3 && rc == 0; ++ti)
      for (int bi = 0; bi < 3 && rc == 0; ++bi) {
        h->tile = kTiles[ti];
        h->large_box = kLargeBoxes[bi];
        ForwardPlan fp;
        BackwardPlan bp;
        if (plan_mesh(m, rep, io[rep], true, s, fp, bp)) { rc = 1; break; }
        hipError_t e = enqueue_mesh(fp, bp, true, s);   // warm-up
        float t_min = 1e30f;
        for (int r = 0; r < reps && e == hipSuccess; ++r) {
          e = hipEventRecord(a, s);
          if (e == hipSuccess) e = enqueue_mesh(fp, bp, true, s);
          if (e == hipSuccess) e = hipEventRecord(b, s);
          if (e == hipSuccess) e = hipEventSynchronize(b);
          float ms = 0.f;
          if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
          if (e == hipSuccess) t_min = std::min(t_min, ms);
        }
        if (e != hipSuccess) { rc = fail(std::string("autotune: ") + hipGetErrorString(e)); break; }
        if (t_min < best) { best = t_min; best_tile = kTiles[ti]; best_box = kLargeBoxes[bi]; }
      }
    for (int i = 0; i < n; ++i)
      if (m->cls[i] == c) { m->meshes[i]->tile = best_tile; m->meshes[i]->large_box = best_box; }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  m->drop_graph();
  return rc;
}

int gvv_multi_set_tuning(void* handle, int32_t mesh, int32_t tile, int32_t large_box) {
  Multi* m = (Multi*)handle;
  if (!m) return fail("null handle");
  std::lock_guard<std::mutex> lock(m->mu);
  if (mesh < 0 || mesh >= (int32_t)m->meshes.size()) return fail("mesh index out of range");
  if (gvv_set_tuning(m->meshes[mesh], tile, large_box)) return 1;
  m->drop_graph();
  return 0;
}

int gvv_multi_tuning(void* handle, int32_t mesh, int32_t* tile, int32_t* large_box, int32_t* size_class) {
  Multi* m = (Multi*)handle;
  if (!m) return fail("null handle");
  std::lock_guard<std::mutex> lock(m->mu);
  if (mesh < 0 || mesh >= (int32_t)m->meshes.size()) return fail("mesh index out of range");
  if (tile) *tile = m->meshes[mesh]->tile;
  if (large_box) *large_box = m->meshes[mesh]->large_box;
  if (size_class) *size_class = m->cls[mesh];
  return 0;
}

}  // extern "C"
