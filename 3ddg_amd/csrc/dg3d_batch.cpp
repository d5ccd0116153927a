// C ABI of the batched path: many independent blocks ("solves") of different element counts and
// orders, advanced by LSRK54. Per stage, every order present gets one launch over all its solves'
// element groups (or, under schedule 4, one persistent launch covers every order); the launches
// of a stage run under a tuned schedule, and whole steps are captured in HIP graphs and replayed.
//
// Each solve is what one dg3d_ctx holds (Grid.cpp:176-375 builds one such block per rank): its
// own mesh, material, operators and state. The per-node arithmetic is the fused stage of
// dg3d_step, so a solve of the batch follows the same trajectory as a context of its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "basis.hpp"
#include "dg3d.h"
#include "dg3d_device.hpp"

namespace {

thread_local std::string g_batch_create_error;

template <typename T>
hipError_t bmalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * count);
}

struct HostSolve {
  int n = 0, np = 0, es = 0;
  int nE[3] = {0, 0, 0};
  int periodic[3] = {0, 0, 0};
  int64_t E = 0, fs = 0, slot0 = 0;
  double dt = 0.0;
  bool has_halo[6] = {false, false, false, false, false, false};
};

struct OrderOps {
  double* opd[3] = {nullptr, nullptr, nullptr};
  float* opf[3] = {nullptr, nullptr, nullptr};
};

}  // namespace

struct dg3d_batch {
  int precision = DG3D_MIXED;
  std::vector<HostSolve> solves;
  std::map<int, OrderOps> ops;         // per order
  std::vector<double*> allocs;         // per-solve geometry, material and halo arrays
  double* arena[2] = {nullptr, nullptr};
  void* du = nullptr;
  int64_t total_slots = 0;
  int cur = 0;                         // arena holding the state
  bool halos_pending = true;           // the next stage is the first of the run
  dg3d::RhsLaunch* d_records = nullptr;  // [solve][kBatchRecords]
  // boundary sources: entries for the solves that have one, the solves' clocks (time of the
  // current step), time steps and stage offsets C_s dt
  dg3d::BatchSource* d_sources = nullptr;
  int nsources = 0;
  int64_t source_items = 0;  // face nodes of the largest source face
  double* d_clocks = nullptr;
  double* d_dts = nullptr;
  double* d_offsets = nullptr;
  // per order present (descending n): task lists, grids and the tuned shape
  struct Group {
    int n = 0;
    dg3d::BatchTask* d_tasks[2] = {nullptr, nullptr};  // per elements-per-workgroup shape
    int ntasks[2] = {0, 0};
    int grid[dg3d::kNumBatchVariants] = {};
    int variant = -1;
    int64_t nodes = 0;
    hipStream_t stream = nullptr;      // branch of the stage's fork/join
    hipEvent_t done = nullptr;
    int branch = -1;                   // stream index this order runs on (-1: the batch stream)
  };
  std::vector<Group> groups;
  // schedule 4: every order's element groups in one task list (largest order first) for one
  // persistent launch, and that launch's grid
  dg3d::BatchTask* d_all_tasks = nullptr;
  int all_ntasks = 0;
  int all_grid = 1;
  bool tuned = false;
  int schedule = 0;  // how a stage runs the orders (kScheduleBranches, kScheduleSerial, kScheduleTwo)
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, fork = nullptr;
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // one step starting from arena 0 / arena 1
  float last_ms = 0.f;
  int last_launches = 0;
  std::string err;

  int fail(int code, const std::string& msg) {
    err = msg;
    return code;
  }
  int hip_fail(hipError_t e, const char* what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return DG3D_E_HIP;
  }
};

namespace {

void drop_graphs(dg3d_batch* b) {
  for (auto& g : b->graph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

void free_batch(dg3d_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  for (double* p : b->allocs) (void)hipFree(p);
  for (auto& kv : b->ops) {
    for (int i = 0; i < 3; ++i) {
      (void)hipFree(kv.second.opd[i]);
      (void)hipFree(kv.second.opf[i]);
    }
  }
  (void)hipFree(b->arena[0]);
  (void)hipFree(b->arena[1]);
  (void)hipFree(b->du);
  (void)hipFree(b->d_records);
  (void)hipFree(b->d_sources);
  (void)hipFree(b->d_clocks);
  (void)hipFree(b->d_dts);
  (void)hipFree(b->d_offsets);
  (void)hipFree(b->d_all_tasks);
  for (auto& g : b->groups) {
    for (auto* t : g.d_tasks) (void)hipFree(t);
    if (g.done) (void)hipEventDestroy(g.done);
    if (g.stream) (void)hipStreamDestroy(g.stream);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

dg3d::BatchLaunch stage_launch(const dg3d_batch* b, const dg3d_batch::Group& g, int variant, int stage, bool halo,
                               int from) {
  dg3d::BatchLaunch B{};
  B.records = b->d_records;
  B.tasks = g.d_tasks[variant & 1];
  B.ntasks = g.ntasks[variant & 1];
  B.record = (halo && stage == 0) ? 10 + from : 2 * stage + from;
  B.mixed = (b->precision == DG3D_MIXED) ? 1 : 0;
  return B;
}

// Stage schedules: every order on a branch of its own (the orders share the device); all orders
// one after another on the batch stream, largest first (each has the device to itself); two
// branches with about the same number of nodes; the largest order on one branch and the others
// one after another on a second, so that they run beside it and fill the device while its last
// workgroups drain; or every order in one persistent launch (k_rhs_batch_all).
constexpr int kScheduleBranches = 0, kScheduleSerial = 1, kScheduleTwo = 2, kScheduleBeside = 3, kScheduleOne = 4,
              kNumSchedules = 5;

void set_schedule(dg3d_batch* b, int schedule) {
  b->schedule = schedule;
  int64_t load[2] = {0, 0};
  for (size_t i = 0; i < b->groups.size(); ++i) {
    auto& g = b->groups[i];
    if (schedule == kScheduleBranches) g.branch = (int)i;
    else if (schedule == kScheduleSerial || schedule == kScheduleOne) g.branch = -1;
    else if (schedule == kScheduleBeside) g.branch = (i == 0) ? 0 : 1;  // groups are in descending order
    else if (schedule == kScheduleTwo) {
      // groups are in descending order; each goes to the branch with fewer nodes so far
      g.branch = (load[1] < load[0]) ? 1 : 0;
      load[g.branch] += g.nodes;
    }
  }
}

// One stage of every solve under the batch's schedule (fork/join on events for the branches).
hipError_t launch_stage(dg3d_batch* b, int stage, bool halo, int from) {
  const size_t ng = b->groups.size();
  std::vector<char> used(ng, 0);
  for (const auto& g : b->groups)
    if (g.branch >= 0) used[g.branch] = 1;
  // the source faces' exterior traces at every solve's stage time, before the stage kernels
  hipError_t e = dg3d::launch_batch_sources(b->d_sources, b->nsources, b->source_items, b->d_clocks, b->d_offsets,
                                            stage, halo ? 1 : 0, b->stream);
  if (b->schedule == kScheduleOne) {
    if (e != hipSuccess) return e;
    dg3d::BatchLaunch B{};
    B.records = b->d_records;
    B.tasks = b->d_all_tasks;
    B.ntasks = b->all_ntasks;
    B.record = (halo && stage == 0) ? 10 + from : 2 * stage + from;
    B.mixed = (b->precision == DG3D_MIXED) ? 1 : 0;
    return dg3d::launch_rhs_batch_all(B, b->all_grid, b->stream);
  }
  if (e == hipSuccess) e = hipEventRecord(b->fork, b->stream);
  for (size_t i = 0; i < ng && e == hipSuccess; ++i)
    if (used[i]) e = hipStreamWaitEvent(b->groups[i].stream, b->fork, 0);
  for (auto& g : b->groups) {
    if (e != hipSuccess) break;
    hipStream_t s = g.branch >= 0 ? b->groups[g.branch].stream : b->stream;
    e = dg3d::launch_rhs_batch(stage_launch(b, g, g.variant, stage, halo, from), g.n, g.variant, g.grid[g.variant], s);
  }
  for (size_t i = 0; i < ng && e == hipSuccess; ++i) {
    if (!used[i]) continue;
    e = hipEventRecord(b->groups[i].done, b->groups[i].stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->groups[i].done, 0);
  }
  return e;
}

// One LSRK54 step from arena `from` (5 stages; the state ends in the other arena).
hipError_t launch_step(dg3d_batch* b, int from, bool first_halo) {
  int cur = from;
  for (int st = 0; st < 5; ++st) {
    hipError_t e = launch_stage(b, st, first_halo && st == 0, cur);
    if (e != hipSuccess) return e;
    cur = 1 - cur;
  }
  // the solves with a source advance their clocks by their own dt
  if (b->nsources > 0) return dg3d::launch_batch_clocks(b->d_clocks, b->d_dts, (int)b->solves.size(), b->stream);
  return hipSuccess;
}

hipError_t build_graphs(dg3d_batch* b) {
  drop_graphs(b);
  for (int from = 0; from < 2; ++from) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t le = launch_step(b, from, false);
    e = hipStreamEndCapture(b->stream, &g);
    if (le != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return le;
    }
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&b->graph[from], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Per order, times every variant (shape x persistent or one task per workgroup x early or late
// RK-register loads) on a later-stage launch (reads and writes
// du) that writes only arena 1 - cur and du; du is cleared first, and the run's next stage 0
// overwrites both.
int batch_autotune(dg3d_batch* b) {
  hipError_t e = hipMemsetAsync(b->du, 0, (b->precision == DG3D_MIXED ? 4 : 8) * (size_t)b->total_slots, b->stream);
  if (e != hipSuccess) return b->hip_fail(e, "batch autotune");
  for (auto& g : b->groups) {
    // warm-up, then two sweeps of which each variant keeps the faster
    const dg3d::BatchLaunch W = stage_launch(b, g, 0, 1, false, b->cur);
    for (int r = 0; r < 4 && e == hipSuccess; ++r) e = dg3d::launch_rhs_batch(W, g.n, 0, g.grid[0], b->stream);
    float best[dg3d::kNumBatchVariants];
    for (float& x : best) x = 1e30f;
    for (int sweep = 0; sweep < 2; ++sweep) {
      for (int v = 0; v < dg3d::kNumBatchVariants; ++v) {
        const dg3d::BatchLaunch B = stage_launch(b, g, v, 1, false, b->cur);
        if (e == hipSuccess) e = dg3d::launch_rhs_batch(B, g.n, v, g.grid[v], b->stream);
        if (e == hipSuccess) e = hipEventRecord(b->ev0, b->stream);
        for (int r = 0; r < 3 && e == hipSuccess; ++r) e = dg3d::launch_rhs_batch(B, g.n, v, g.grid[v], b->stream);
        if (e == hipSuccess) e = hipEventRecord(b->ev1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
        if (e != hipSuccess) return b->hip_fail(e, "batch autotune");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
        best[v] = std::min(best[v], ms);
      }
    }
    g.variant = (int)(std::min_element(best, best + dg3d::kNumBatchVariants) - best);
  }
  // then the schedule, on whole stages of the same launches
  float sbest[kNumSchedules];
  for (float& x : sbest) x = 1e30f;
  for (int sweep = 0; sweep < 2; ++sweep) {
    for (int sc = 0; sc < kNumSchedules; ++sc) {
      set_schedule(b, sc);
      e = launch_stage(b, 1, false, b->cur);
      if (e == hipSuccess) e = hipEventRecord(b->ev0, b->stream);
      for (int r = 0; r < 3 && e == hipSuccess; ++r) e = launch_stage(b, 1, false, b->cur);
      if (e == hipSuccess) e = hipEventRecord(b->ev1, b->stream);
      if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
      if (e != hipSuccess) return b->hip_fail(e, "batch autotune");
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
      sbest[sc] = std::min(sbest[sc], ms);
    }
  }
  set_schedule(b, (int)(std::min_element(sbest, sbest + kNumSchedules) - sbest));
  b->tuned = true;
  return DG3D_OK;
}

}  // namespace

extern "C" {

const char* dg3d_batch_create_error(void) { return g_batch_create_error.c_str(); }

const char* dg3d_batch_last_error(const dg3d_batch* b) { return b ? b->err.c_str() : "null batch"; }

int dg3d_batch_create(int count, const dg3d_solve_desc* descs, int precision, dg3d_batch** out) {
  if (!out) return DG3D_E_INVALID;
  *out = nullptr;
  if (count < 1 || !descs) {
    g_batch_create_error = "batch: count must be >= 1";
    return DG3D_E_INVALID;
  }
  if (precision != DG3D_FP64 && precision != DG3D_MIXED) {
    g_batch_create_error = "batch: bad precision";
    return DG3D_E_INVALID;
  }
  dg3d_batch* b = new (std::nothrow) dg3d_batch();
  if (!b) return DG3D_E_NOMEM;
  b->precision = precision;
  auto bail = [&](int code, const std::string& msg) {
    g_batch_create_error = msg;
    free_batch(b);
    return code;
  };
  // host-side validation and layout
  b->solves.resize(count);
  int64_t slot = 0;
  for (int i = 0; i < count; ++i) {
    const dg3d_solve_desc& d = descs[i];
    HostSolve& h = b->solves[i];
    const std::string tag = "batch solve " + std::to_string(i) + ": ";
    if (d.order < 1 || d.order > 7) return bail(DG3D_E_UNSUPPORTED, tag + "order must be 1..7");
    for (int a = 0; a < 3; ++a) {
      if (d.num_elements[a] < 1) return bail(DG3D_E_INVALID, tag + "num_elements must be >= 1");
      if (!d.widths[a]) return bail(DG3D_E_INVALID, tag + "widths missing");
    }
    if (!d.mass || !d.stiff || !d.c0 || !d.rho0 || !d.state) return bail(DG3D_E_INVALID, tag + "missing array");
    if (!(d.dt > 0.0) || !std::isfinite(d.dt)) return bail(DG3D_E_INVALID, tag + "dt must be positive");
    if (d.source_face >= 0) {
      if (d.source_face > 5) return bail(DG3D_E_INVALID, tag + "source face must be 0..5");
      if (d.periodic[2 - d.source_face / 2]) return bail(DG3D_E_INVALID, tag + "source on a periodic face");
      if (!(d.source_tau > 0.0)) return bail(DG3D_E_INVALID, tag + "source tau must be positive");
    }
    h.n = d.order + 1;
    h.np = h.n * h.n * h.n;
    h.es = dg3d::elem_stride(h.n);
    for (int a = 0; a < 3; ++a) {
      h.nE[a] = d.num_elements[a];
      h.periodic[a] = d.periodic[a] ? 1 : 0;
    }
    h.E = (int64_t)h.nE[0] * h.nE[1] * h.nE[2];
    h.fs = dg3d::field_slots(h.E, h.n);
    if (h.fs >= (1ll << 29)) return bail(DG3D_E_INVALID, tag + "a field takes 2^29 or more node slots");
    h.slot0 = slot;
    slot += 4 * h.fs;
    h.dt = d.dt;
  }
  b->total_slots = slot;
  hipError_t e = hipStreamCreate(&b->stream);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e != hipSuccess) return bail(DG3D_E_HIP, std::string("HIP init: ") + hipGetErrorString(e));
  // arenas
  if (bmalloc(&b->arena[0], slot) != hipSuccess || bmalloc(&b->arena[1], slot) != hipSuccess)
    return bail(DG3D_E_NOMEM, "batch: state arena allocation failed");
  const size_t du_bytes = (precision == DG3D_MIXED ? 4 : 8) * (size_t)slot;
  if (hipMalloc(&b->du, du_bytes) != hipSuccess) return bail(DG3D_E_NOMEM, "batch: du arena allocation failed");
  hipError_t up = hipSuccess;  // first failure of the uploads below, reported after them
  auto upload = [&up](hipError_t r) {
    if (up == hipSuccess) up = r;
  };
  upload(hipMemset(b->arena[0], 0, sizeof(double) * slot));
  upload(hipMemset(b->arena[1], 0, sizeof(double) * slot));
  upload(hipMemset(b->du, 0, du_bytes));
  // operators per order
  for (int i = 0; i < count; ++i) {
    const int n = b->solves[i].n;
    if (b->ops.count(n)) continue;
    dg3d::LineOperator op;
    if (!dg3d::make_line_operator(n, descs[i].mass, descs[i].stiff, &op))
      return bail(DG3D_E_INVALID, "batch: singular mass matrix for order " + std::to_string(n - 1));
    OrderOps& oo = b->ops[n];
    const std::vector<double>* srcs[3] = {&op.A, &op.L0, &op.LN};
    for (int k = 0; k < 3; ++k) {
      const size_t cnt = srcs[k]->size();
      std::vector<float> f(srcs[k]->begin(), srcs[k]->end());
      if (bmalloc(&oo.opd[k], cnt) != hipSuccess || bmalloc(&oo.opf[k], cnt) != hipSuccess)
        return bail(DG3D_E_NOMEM, "batch: operator allocation failed");
      upload(hipMemcpy(oo.opd[k], srcs[k]->data(), sizeof(double) * cnt, hipMemcpyHostToDevice));
      upload(hipMemcpy(oo.opf[k], f.data(), sizeof(float) * cnt, hipMemcpyHostToDevice));
    }
  }
  // per-solve device data and launch records
  std::vector<dg3d::RhsLaunch> recs((size_t)count * dg3d::kBatchRecords);
  std::vector<dg3d::BatchSource> sources;
  int64_t source_items = 0;
  std::vector<double> stage;
  for (int i = 0; i < count; ++i) {
    const dg3d_solve_desc& d = descs[i];
    HostSolve& h = b->solves[i];
    dg3d::RhsLaunch base{};
    dg3d::MeshArgs& m = base.mesh;
    for (int a = 0; a < 3; ++a) {
      m.nE[a] = h.nE[a];
      m.periodic[a] = h.periodic[a];
      std::vector<double> iw(h.nE[a]);
      for (int k = 0; k < h.nE[a]; ++k) {
        if (!(d.widths[a][k] > 0.0) || !std::isfinite(d.widths[a][k]))
          return bail(DG3D_E_INVALID, "batch solve " + std::to_string(i) + ": element widths must be positive");
        iw[k] = 1.0 / d.widths[a][k];
      }
      double* p = nullptr;
      if (bmalloc(&p, iw.size()) != hipSuccess) return bail(DG3D_E_NOMEM, "batch: geometry allocation failed");
      b->allocs.push_back(p);
      upload(hipMemcpy(p, iw.data(), sizeof(double) * iw.size(), hipMemcpyHostToDevice));
      m.inv_w[a] = p;
    }
    m.E = h.E;
    m.fs = h.fs;
    m.es = h.es;
    m.grp = dg3d::group_elems(h.n);
    m.stride[0] = (int64_t)h.nE[1] * h.nE[2];
    m.stride[1] = h.nE[2];
    m.stride[2] = 1;
    m.div12 = dg3d::make_div32((uint32_t)(h.nE[1] * h.nE[2]));
    m.div2 = dg3d::make_div32((uint32_t)h.nE[2]);
    std::vector<double> rec((size_t)h.E * dg3d::kMatStride, 0.0);
    for (int64_t k = 0; k < h.E; ++k) {
      double* r = &rec[(size_t)k * dg3d::kMatStride];
      if (!(d.c0[k] > 0.0) || !(d.rho0[k] > 0.0))
        return bail(DG3D_E_INVALID, "batch solve " + std::to_string(i) + ": c0 and rho0 must be positive");
      r[dg3d::MAT_C] = d.c0[k];
      r[dg3d::MAT_RHO0] = d.rho0[k];
      r[dg3d::MAT_BA] = d.b_over_a ? d.b_over_a[k] : 0.0;
      r[dg3d::MAT_SIGMA] = d.sigma ? d.sigma[k] : 0.0;
      r[dg3d::MAT_IRHO0] = 1.0 / d.rho0[k];
    }
    double* mat = nullptr;
    if (bmalloc(&mat, rec.size()) != hipSuccess) return bail(DG3D_E_NOMEM, "batch: material allocation failed");
    b->allocs.push_back(mat);
    upload(hipMemcpy(mat, rec.data(), sizeof(double) * rec.size(), hipMemcpyHostToDevice));
    m.mat = mat;
    const double* traces[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int f = 0; f < 6; ++f) {
      if (!d.halo[f] || h.periodic[2 - f / 2]) continue;
      const int axis = 2 - f / 2;
      const int64_t fe = (int64_t)h.nE[axis == 0 ? 1 : 0] * h.nE[axis == 2 ? 1 : 2];
      const size_t cnt = (size_t)fe * h.n * h.n * 3;
      double* p = nullptr;
      if (bmalloc(&p, cnt) != hipSuccess) return bail(DG3D_E_NOMEM, "batch: halo allocation failed");
      b->allocs.push_back(p);
      upload(hipMemcpy(p, d.halo[f], sizeof(double) * cnt, hipMemcpyHostToDevice));
      traces[f] = p;
      h.has_halo[f] = true;
    }
    const OrderOps& oo = b->ops[h.n];
    base.n = h.n;
    base.mixed = (precision == DG3D_MIXED) ? 1 : 0;
    base.epilogue = dg3d::EPI_LSRK;
    base.mask = dg3d::kTermAll;
    base.ext.src_face = -1;
    for (int k = 0; k < 3; ++k) {
      base.opd[k] = oo.opd[k];
      base.opf[k] = oo.opf[k];
    }
    base.dt = h.dt;
    char* du0 = static_cast<char*>(b->du) + (precision == DG3D_MIXED ? 4 : 8) * (size_t)h.slot0;
    base.du = du0;
    // a boundary source: its face's trace buffer, which the source kernel fills every stage
    double* src_out = nullptr;
    if (d.source_face >= 0) {
      const int f = d.source_face, axis = 2 - f / 2;
      const int64_t fe = (int64_t)h.nE[axis == 0 ? 1 : 0] * h.nE[axis == 2 ? 1 : 2];
      const int64_t items = fe * h.n * h.n;
      if (bmalloc(&src_out, (size_t)items * 3) != hipSuccess) return bail(DG3D_E_NOMEM, "batch: source allocation failed");
      b->allocs.push_back(src_out);
      double* lead = nullptr;
      if (d.source_lead) {
        if (bmalloc(&lead, (size_t)items) != hipSuccess) return bail(DG3D_E_NOMEM, "batch: source allocation failed");
        b->allocs.push_back(lead);
        upload(hipMemcpy(lead, d.source_lead, sizeof(double) * items, hipMemcpyHostToDevice));
      }
      dg3d::BatchSource bs{};
      bs.x.src_face = f;
      bs.x.amp = d.source_amplitude;
      bs.x.freq = d.source_frequency;
      bs.x.phase = d.source_phase;
      bs.x.t0 = d.source_t0;
      bs.x.tau = d.source_tau;
      bs.x.lead = lead;
      bs.m = m;
      bs.n = h.n;
      bs.solve = i;
      bs.halo = traces[f];
      bs.out = src_out;
      sources.push_back(bs);
      source_items = std::max(source_items, items);
    }
    for (int r = 0; r < dg3d::kBatchRecords; ++r) {
      const int from = r & 1, st = r < 10 ? r / 2 : 0;
      dg3d::RhsLaunch L = base;
      L.q_in = b->arena[from] + h.slot0;
      L.q_out = b->arena[1 - from] + h.slot0;
      L.a = dg3d::kLsA[st];
      L.b = dg3d::kLsB[st];
      for (int f = 0; f < 6; ++f) L.ext.trace[f] = (r >= 10) ? traces[f] : nullptr;
      if (src_out) L.ext.trace[d.source_face] = src_out;
      recs[(size_t)i * dg3d::kBatchRecords + r] = L;
    }
    // initial state: host [4][E][np] -> group-blocked slots (dg3d::reg_slot), padding zero
    stage.assign((size_t)4 * h.fs, 0.0);
    for (int f = 0; f < 4; ++f)
      for (int64_t el = 0; el < h.E; ++el)
        std::memcpy(&stage[dg3d::reg_slot(h.fs, h.es, dg3d::group_elems(h.n), f, el, 0)], d.state + ((size_t)f * h.E + el) * h.np,
                    sizeof(double) * h.np);
    e = hipMemcpy(b->arena[0] + h.slot0, stage.data(), sizeof(double) * stage.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(DG3D_E_HIP, std::string("batch: state upload: ") + hipGetErrorString(e));
  }
  if (!sources.empty()) {
    // clocks from 0; stage offsets C_s dt computed here, as dg3d_step computes them
    std::vector<double> dts(count), offs((size_t)count * 5);
    for (int i = 0; i < count; ++i) {
      dts[i] = b->solves[i].dt;
      for (int st = 0; st < 5; ++st) offs[(size_t)i * 5 + st] = 0.0 + dg3d::kLsC[st] * dts[i];
    }
    if (bmalloc(&b->d_sources, sources.size()) != hipSuccess || bmalloc(&b->d_clocks, (size_t)count) != hipSuccess ||
        bmalloc(&b->d_dts, (size_t)count) != hipSuccess || bmalloc(&b->d_offsets, offs.size()) != hipSuccess)
      return bail(DG3D_E_NOMEM, "batch: source allocation failed");
    upload(hipMemcpy(b->d_sources, sources.data(), sizeof(dg3d::BatchSource) * sources.size(), hipMemcpyHostToDevice));
    upload(hipMemset(b->d_clocks, 0, sizeof(double) * count));
    upload(hipMemcpy(b->d_dts, dts.data(), sizeof(double) * count, hipMemcpyHostToDevice));
    upload(hipMemcpy(b->d_offsets, offs.data(), sizeof(double) * offs.size(), hipMemcpyHostToDevice));
    b->nsources = (int)sources.size();
    b->source_items = source_items;
  }
  if (bmalloc(&b->d_records, recs.size()) != hipSuccess) return bail(DG3D_E_NOMEM, "batch: record allocation failed");
  upload(hipMemcpy(b->d_records, recs.data(), sizeof(dg3d::RhsLaunch) * recs.size(), hipMemcpyHostToDevice));
  // one group per order present, largest first; task lists for both shapes
  std::vector<int> orders;
  for (const HostSolve& h : b->solves)
    if (std::find(orders.begin(), orders.end(), h.n) == orders.end()) orders.push_back(h.n);
  std::sort(orders.rbegin(), orders.rend());
  for (int n : orders) {
    dg3d_batch::Group g;
    g.n = n;
    for (int i = 0; i < count; ++i)
      if (b->solves[i].n == n) g.nodes += b->solves[i].E * b->solves[i].np;
    for (int v = 0; v < 2; ++v) {
      const int epb = dg3d::batch_epb(v, n);
      std::vector<dg3d::BatchTask> tasks;
      for (int i = 0; i < count; ++i) {
        if (b->solves[i].n != n) continue;
        for (int64_t e0 = 0; e0 < b->solves[i].E; e0 += epb) tasks.push_back(dg3d::BatchTask{i, (int)e0});
      }
      if (tasks.size() >= (size_t)INT32_MAX) return bail(DG3D_E_INVALID, "batch: too many element groups");
      if (bmalloc(&g.d_tasks[v], tasks.size()) != hipSuccess) {
        b->groups.push_back(g);
        return bail(DG3D_E_NOMEM, "batch: task list allocation failed");
      }
      upload(hipMemcpy(g.d_tasks[v], tasks.data(), sizeof(dg3d::BatchTask) * tasks.size(), hipMemcpyHostToDevice));
      g.ntasks[v] = (int)tasks.size();
    }
    for (int v = 0; v < dg3d::kNumBatchVariants; ++v) {
      const int resident = dg3d::batch_max_grid(n, v, precision == DG3D_MIXED ? 1 : 0);
      g.grid[v] = std::max(1, std::min(g.ntasks[v & 1], resident > 0 ? resident : 1024));
    }
    g.variant = 0;
    e = hipStreamCreate(&g.stream);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g.done, hipEventDisableTiming);
    b->groups.push_back(g);
    if (e != hipSuccess) return bail(DG3D_E_HIP, std::string("batch streams: ") + hipGetErrorString(e));
  }
  {
    // schedule 4's task list: every order's element groups, largest order first
    std::vector<dg3d::BatchTask> tasks;
    for (int n : orders) {
      const int epb = dg3d::batch_all_epb(n);
      for (int i = 0; i < count; ++i) {
        if (b->solves[i].n != n) continue;
        for (int64_t e0 = 0; e0 < b->solves[i].E; e0 += epb) tasks.push_back(dg3d::BatchTask{i, (int)e0});
      }
    }
    if (tasks.size() >= (size_t)INT32_MAX) return bail(DG3D_E_INVALID, "batch: too many element groups");
    if (bmalloc(&b->d_all_tasks, tasks.size()) != hipSuccess) return bail(DG3D_E_NOMEM, "batch: task list allocation failed");
    upload(hipMemcpy(b->d_all_tasks, tasks.data(), sizeof(dg3d::BatchTask) * tasks.size(), hipMemcpyHostToDevice));
    b->all_ntasks = (int)tasks.size();
    const int resident = dg3d::batch_all_max_grid(precision == DG3D_MIXED ? 1 : 0);
    b->all_grid = std::max(1, std::min(b->all_ntasks, resident > 0 ? resident : 1024));
  }
  set_schedule(b, kScheduleBranches);
  e = hipDeviceSynchronize();
  if (e == hipSuccess) e = up;
  if (e != hipSuccess) return bail(DG3D_E_HIP, std::string("batch setup: ") + hipGetErrorString(e));
  *out = b;
  return DG3D_OK;
}

int dg3d_batch_destroy(dg3d_batch* b) {
  free_batch(b);
  return DG3D_OK;
}

int dg3d_batch_set_schedule(dg3d_batch* b, int schedule) {
  if (!b) return DG3D_E_INVALID;
  if (schedule < 0 || schedule >= kNumSchedules) return b->fail(DG3D_E_INVALID, "batch: schedule out of range");
  set_schedule(b, schedule);
  drop_graphs(b);
  return DG3D_OK;
}

int dg3d_batch_get_schedule(const dg3d_batch* b, int* schedule) {
  if (!b || !schedule) return DG3D_E_INVALID;
  *schedule = b->schedule;
  return DG3D_OK;
}

int dg3d_batch_set_variant(dg3d_batch* b, int variant) {
  if (!b) return DG3D_E_INVALID;
  if (variant < 0 || variant >= dg3d::kNumBatchVariants) return b->fail(DG3D_E_INVALID, "batch: variant out of range");
  for (auto& g : b->groups) g.variant = variant;
  b->tuned = true;
  drop_graphs(b);
  return DG3D_OK;
}

int dg3d_batch_autotune(dg3d_batch* b, int* variants) {
  if (!b) return DG3D_E_INVALID;
  const int rc = batch_autotune(b);
  if (rc) return rc;
  drop_graphs(b);
  if (variants)
    for (size_t i = 0; i < b->groups.size(); ++i) variants[i] = b->groups[i].variant;
  return DG3D_OK;
}

int dg3d_batch_step(dg3d_batch* b, int nsteps) {
  if (!b) return DG3D_E_INVALID;
  if (nsteps < 0) return b->fail(DG3D_E_INVALID, "batch step: nsteps < 0");
  if (nsteps == 0) return DG3D_OK;
  if (!b->tuned) {
    const int rc = batch_autotune(b);
    if (rc) return rc;
  }
  hipError_t e = hipSuccess;
  if (!b->graph[0] || !b->graph[1]) {
    e = build_graphs(b);
    if (e != hipSuccess) return b->hip_fail(e, "batch graph capture");
  }
  e = hipEventRecord(b->ev0, b->stream);
  int s = 0;
  if (e == hipSuccess && b->halos_pending) {
    // the run's first stage reads the initial halo traces: launched directly, not from a graph
    e = launch_step(b, b->cur, true);
    b->cur = 1 - b->cur;
    b->halos_pending = false;
    s = 1;
  }
  for (; s < nsteps && e == hipSuccess; ++s) {
    e = hipGraphLaunch(b->graph[b->cur], b->stream);
    b->cur = 1 - b->cur;
  }
  if (e == hipSuccess) e = hipEventRecord(b->ev1, b->stream);
  if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
  if (e != hipSuccess) return b->hip_fail(e, "batch step");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
  b->last_ms = ms;
  b->last_launches = 5 * nsteps;
  return DG3D_OK;
}

int dg3d_batch_get_state(dg3d_batch* b, int solve, double* host) {
  if (!b || !host) return DG3D_E_INVALID;
  if (solve < 0 || solve >= (int)b->solves.size()) return b->fail(DG3D_E_INVALID, "batch: solve out of range");
  const HostSolve& h = b->solves[solve];
  hipError_t e = hipStreamSynchronize(b->stream);
  std::vector<double> stage((size_t)4 * h.fs);
  if (e == hipSuccess)
    e = hipMemcpy(stage.data(), b->arena[b->cur] + h.slot0, sizeof(double) * stage.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return b->hip_fail(e, "batch get_state");
  for (int f = 0; f < 4; ++f)
    for (int64_t el = 0; el < h.E; ++el)
      std::memcpy(host + ((size_t)f * h.E + el) * h.np, &stage[dg3d::reg_slot(h.fs, h.es, dg3d::group_elems(h.n), f, el, 0)],
                  sizeof(double) * h.np);
  return DG3D_OK;
}

int dg3d_batch_info(const dg3d_batch* b, int64_t* total_nodes, int* num_orders, int* orders, int* variants,
                    int* tasks, int* grids) {
  if (!b) return DG3D_E_INVALID;
  int64_t nodes = 0;
  for (const HostSolve& h : b->solves) nodes += h.E * h.np;
  if (total_nodes) *total_nodes = nodes;
  if (num_orders) *num_orders = (int)b->groups.size();
  for (size_t i = 0; i < b->groups.size(); ++i) {
    const auto& g = b->groups[i];
    const int v = g.variant < 0 ? 0 : g.variant;
    if (orders) orders[i] = g.n - 1;
    if (variants) variants[i] = g.variant;
    if (tasks) tasks[i] = g.ntasks[v & 1];
    if (grids) grids[i] = g.grid[v];
  }
  return DG3D_OK;
}

int dg3d_batch_last_step_timing(const dg3d_batch* b, double* avg_stage_ms, int* launches) {
  if (!b) return DG3D_E_INVALID;
  if (avg_stage_ms) *avg_stage_ms = b->last_launches ? b->last_ms / b->last_launches : 0.0;
  if (launches) *launches = b->last_launches;
  return DG3D_OK;
}

}  // extern "C"
