// C ABI of the HIP implementation (include/dg3d.h). Owns all device memory of a block:
// state registers (the reference's Grid copies), per-element material, per-axis widths,
// the precomputed line operator, halo face traces and the work registers of the fused step.
#include "dg3d.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "basis.hpp"
#include "dg3d_device.hpp"

namespace {

// LSRK(5,4) 2N-storage coefficients, Carpenter & Kennedy (1994), NASA TM-109112.
using dg3d::kLsA;
using dg3d::kLsB;
using dg3d::kLsC;

thread_local std::string g_create_error;

// A captured run of `steps` steps of dg3d_step (each: stages, probe, clock), keyed by the
// registers it starts from, dt and the step count; everything else it reads is fixed until a
// setter drops it.
struct StepGraph {
  hipGraphExec_t exec = nullptr;
  int scheme = -1;
  int steps = 0;
  double dt = 0.0;
  double* cur = nullptr;    // state register the step reads
  double* other = nullptr;  // LSRK54: the second state buffer
  double* cur_end = nullptr;
  double* other_end = nullptr;
};
constexpr int kMaxStepGraphs = 16;
// dg3d_step replays runs of kGraphSteps steps, and one shorter run (a power of two each) for the
// remainder.
constexpr int kGraphSteps = 16;

struct Register {
  double* d = nullptr;          // nfields * fs doubles
  double* halo[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  bool halo_valid[6] = {false, false, false, false, false, false};
};

}  // namespace

struct dg3d_ctx {
  int N = 0, n = 0, np = 0, es = 0;  // es: node slots per element in device memory
  int grp = 1;                        // elements per group of the register layout
  int nE[3] = {0, 0, 0};
  int periodic[3] = {0, 0, 0};
  int64_t E = 0, fs = 0;
  int nfields = 4;
  int precision = DG3D_FP64, contraction = DG3D_CONTRACT_VALU, xcd = 1;
  hipStream_t stream = nullptr;
  std::vector<Register> regs;
  double* work_q = nullptr;
  void* work_du = nullptr;
  double* work_k1 = nullptr;
  double* work_k2 = nullptr;
  double* d_inv_w[3] = {nullptr, nullptr, nullptr};
  std::vector<double> h_w[3];
  double corner[3] = {0, 0, 0};
  double* d_mat = nullptr;
  std::vector<double> h_mat;
  double* d_diff = nullptr;
  std::vector<double> h_diff;
  bool has_diff = false;  // some element has D != 0: the fused stages add the LDG term
  double* d_opd[3] = {nullptr, nullptr, nullptr};
  float* d_opf[3] = {nullptr, nullptr, nullptr};
  std::vector<double> nodes, weights;
  double* d_wq = nullptr;
  dg3d_source src{};
  double* d_lead = nullptr;
  double* d_src_trace = nullptr;  // exterior traces of the source face, one slot per stage of a step
  double* d_clock = nullptr;    // dg3d_step's time at the start of the current step
  int* d_probe_step = nullptr;  // probe row the next step writes
  bool use_graphs = true;
  StepGraph graphs[kMaxStepGraphs];
  int64_t* d_probe_nodes = nullptr;
  double* d_probe_buf = nullptr;
  int probe_count = 0, probe_cap = 0, probe_steps = 0;
  double* d_part = nullptr;
  std::vector<double> h_part;
  // the allocation-free reductions: this device's copy of the library's scratch array, resolved
  // when the context is created (null, with the error kept, if that failed)
  int device = 0;
  double* scratch = nullptr;
  hipError_t scratch_err = hipSuccess;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int tune_variant[2] = {-1, -1};  // per scheme: chosen kernel variant (-1 = not tuned yet)
  int tune_xcd[2] = {1, 1};
  double last_ms = 0.0;
  int last_launches = 0;
  std::string err;

  int fail(int code, const std::string& msg) {
    err = msg;
    return code;
  }
  // a message written into err's reserved capacity, without a temporary string (the
  // allocation-free entry points report through this)
  int fail_text(int code, const char* a, const char* b = "") {
    err.assign(a);
    err.append(b);
    return code;
  }
  int hip_fail(hipError_t e, const char* what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return DG3D_E_HIP;
  }
  int64_t face_elems(int face) const {
    const int axis = 2 - face / 2;
    const int b = (axis == 0) ? 1 : 0, c = (axis == 2) ? 1 : 2;
    return (int64_t)nE[b] * nE[c];
  }
  dg3d::MeshArgs mesh() const {
    dg3d::MeshArgs m;
    for (int a = 0; a < 3; ++a) {
      m.nE[a] = nE[a];
      m.periodic[a] = periodic[a];
      m.inv_w[a] = d_inv_w[a];
    }
    m.E = E;
    m.fs = fs;
    m.es = es;
    m.grp = grp;
    m.stride[0] = (int64_t)nE[1] * nE[2];
    m.stride[1] = nE[2];
    m.stride[2] = 1;
    m.mat = d_mat;
    m.div12 = dg3d::make_div32((uint32_t)(nE[1] * nE[2]));
    m.div2 = dg3d::make_div32((uint32_t)nE[2]);
    return m;
  }
  dg3d::ExtArgs ext_for(int reg, double t) const {
    dg3d::ExtArgs x{};
    for (int f = 0; f < 6; ++f) x.trace[f] = (reg >= 0 && regs[reg].halo_valid[f]) ? regs[reg].halo[f] : nullptr;
    x.src_face = src.face;
    x.amp = src.amplitude;
    x.freq = src.frequency;
    x.phase = src.phase;
    x.t0 = src.t0;
    x.tau = src.tau;
    x.lead = d_lead;
    x.t = t;
    x.clock = nullptr;
    return x;
  }
  bool reg_ok(int r) const { return r >= 0 && r < (int)regs.size(); }
  // Exterior arguments of the RHS kernel: with a source, a small kernel first writes the source
  // face's trace (halo + source at time t) and the RHS kernel reads it as that face's halo.
  // With dev_clock, t is the stage's offset from the device clock of dg3d_step. With src_slot >= 0
  // the step has already written the trace of this stage into that slot (launch_step).
  hipError_t ext_rhs(int reg, double t, dg3d::ExtArgs* out, bool dev_clock = false, int src_slot = -1) const {
    dg3d::ExtArgs x = ext_for(reg, t);
    if (dev_clock) x.clock = d_clock;
    if (src.face >= 0) {
      if (src_slot < 0) {
        hipError_t e = dg3d::launch_source_trace(x, mesh(), n, x.trace[src.face], d_src_trace, stream);
        if (e != hipSuccess) return e;
      }
      x.trace[src.face] = d_src_trace + (src_slot > 0 ? src_slot : 0) * dg3d::source_trace_count(src.face, mesh(), n) * 3;
      x.src_face = -1;
    }
    *out = x;
    return hipSuccess;
  }
};

namespace {

template <typename T>
hipError_t dmalloc(T** p, size_t count) {
  return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (count ? count : 1));
}

void drop_graphs(dg3d_ctx* c) {
  for (StepGraph& g : c->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = StepGraph{};
  }
}

void free_ctx(dg3d_ctx* c) {
  if (!c) return;
  drop_graphs(c);
  for (auto& r : c->regs) {
    (void)hipFree(r.d);
    for (auto* h : r.halo) (void)hipFree(h);
  }
  (void)hipFree(c->work_q);
  (void)hipFree(c->work_du);
  (void)hipFree(c->work_k1);
  (void)hipFree(c->work_k2);
  for (int a = 0; a < 3; ++a) {
    (void)hipFree(c->d_inv_w[a]);
    (void)hipFree(c->d_opd[a]);
    (void)hipFree(c->d_opf[a]);
  }
  (void)hipFree(c->d_mat);
  (void)hipFree(c->d_diff);
  (void)hipFree(c->d_wq);
  (void)hipFree(c->d_lead);
  (void)hipFree(c->d_src_trace);
  (void)hipFree(c->d_probe_nodes);
  (void)hipFree(c->d_probe_buf);
  (void)hipFree(c->d_clock);
  (void)hipFree(c->d_probe_step);
  (void)hipFree(c->d_part);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int upload_material(dg3d_ctx* c) {
  std::vector<double> rec((size_t)c->E * dg3d::kMatStride, 0.0);
  for (int64_t i = 0; i < c->E; ++i) {
    double* r = &rec[(size_t)i * dg3d::kMatStride];
    r[dg3d::MAT_C] = c->h_mat[i * 4 + 0];
    r[dg3d::MAT_RHO0] = c->h_mat[i * 4 + 1];
    r[dg3d::MAT_BA] = c->h_mat[i * 4 + 2];
    r[dg3d::MAT_SIGMA] = c->h_mat[i * 4 + 3];
    r[dg3d::MAT_IRHO0] = 1.0 / c->h_mat[i * 4 + 1];
    r[dg3d::MAT_DIFF] = c->h_diff[i];
  }
  c->has_diff = std::any_of(c->h_diff.begin(), c->h_diff.end(), [](double d) { return d != 0.0; });
  hipError_t e = hipMemcpyAsync(c->d_mat, rec.data(), sizeof(double) * rec.size(), hipMemcpyHostToDevice,
                                c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "material upload");
  e = hipMemcpyAsync(c->d_diff, c->h_diff.data(), sizeof(double) * c->h_diff.size(), hipMemcpyHostToDevice,
                     c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "material upload");
  e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "material upload");
  return DG3D_OK;
}

int ensure_work(dg3d_ctx* c, int scheme) {
  hipError_t e = hipSuccess;
  const bool fresh_q = scheme == DG3D_LSRK54 && !c->work_q;
  const bool fresh_k = scheme != DG3D_LSRK54 && !c->work_k1;
  if (scheme == DG3D_LSRK54) {
    if (!c->work_q) e = dmalloc(&c->work_q, (size_t)c->nfields * c->fs);
    if (e == hipSuccess && !c->work_du) {
      const size_t bytes = (c->precision == DG3D_MIXED ? sizeof(float) : sizeof(double)) * 4 * (size_t)c->fs;
      e = hipMalloc(&c->work_du, bytes);
    }
  } else {
    if (!c->work_k1) e = dmalloc(&c->work_k1, (size_t)c->nfields * c->fs);
    if (e == hipSuccess && !c->work_k2) e = dmalloc(&c->work_k2, (size_t)c->nfields * c->fs);
  }
  if (e != hipSuccess) return c->fail(DG3D_E_NOMEM, std::string("work register allocation: ") + hipGetErrorString(e));
  // the kernels write only element slots; the slots past the last element must read as zeros
  // once a work register becomes register 0 (LSRK54 hands the state over by pointer), since the
  // reductions cover every slot
  if (fresh_q || fresh_k) {
    const size_t reg_bytes = sizeof(double) * (size_t)c->nfields * (size_t)c->fs;
    if (fresh_q) e = hipMemsetAsync(c->work_q, 0, reg_bytes, c->stream);
    if (e == hipSuccess && fresh_k) e = hipMemsetAsync(c->work_k1, 0, reg_bytes, c->stream);
    if (e == hipSuccess && fresh_k) e = hipMemsetAsync(c->work_k2, 0, reg_bytes, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "work register clear");
  }
  return DG3D_OK;
}

// Fields 0 .. nfields-1 between the host layout ([field][E][n^3]) and a register's group-blocked
// slots (dg3d::reg_slot), in chunks of whole groups: a chunk's state fields are one contiguous
// run of the register, and so are its Q fields. Writing only some of a block's fields reads the
// chunk first; otherwise the gaps and padding slots are written as zeros.
hipError_t transfer_state(const dg3d_ctx* c, double* dev, double* host, int nfields, bool to_device) {
  constexpr int64_t P = dg3d::kRunSlots;
  const int es = c->es, np = c->np;
  const int64_t G = c->grp, groups = (c->E + G - 1) / G;
  const int64_t chunk = std::max<int64_t>(1, ((int64_t)1 << 22) / (4 * P));  // groups per chunk
  std::vector<double> buf;
  hipError_t e = hipSuccess;
  // fields f_lo .. f_lo+f_n-1 held in blocks of blk runs from slot d_base, groups g0 .. g1-1
  auto part = [&](double* d_base, int blk, int f_lo, int f_n, int64_t g0, int64_t g1) {
    const size_t count = (size_t)(g1 - g0) * blk * P;
    double* d = d_base + (size_t)g0 * blk * P;
    if (to_device && f_n == blk) {
      buf.assign(count, 0.0);
    } else {
      buf.resize(count);
      e = hipMemcpy(buf.data(), d, sizeof(double) * count, hipMemcpyDeviceToHost);
      if (e != hipSuccess) return;
    }
    for (int64_t el = g0 * G; el < std::min(c->E, g1 * G); ++el)
      for (int k = 0; k < f_n; ++k) {
        double* slot = &buf[(size_t)((el / G - g0) * blk + k) * P + (size_t)(el % G) * es];
        double* h = host + ((size_t)(f_lo + k) * c->E + el) * np;
        if (to_device) std::memcpy(slot, h, sizeof(double) * np);
        else std::memcpy(h, slot, sizeof(double) * np);
      }
    if (to_device) e = hipMemcpy(d, buf.data(), sizeof(double) * count, hipMemcpyHostToDevice);
  };
  const int sf = std::min(nfields, 4), qf = nfields - sf;
  for (int64_t g0 = 0; g0 < groups && e == hipSuccess; g0 += chunk) {
    const int64_t g1 = std::min(groups, g0 + chunk);
    part(dev, 4, 0, sf, g0, g1);
    if (e == hipSuccess && qf > 0) part(dev + 4 * c->fs, 3, 4, qf, g0, g1);
  }
  return e;
}

dg3d::RhsLaunch base_launch(const dg3d_ctx* c) {
  dg3d::RhsLaunch L{};
  L.n = c->n;
  L.mesh = c->mesh();
  L.xcd_remap = c->xcd;
  L.variant = 0;
  L.contraction = (c->precision == DG3D_MIXED) ? c->contraction : 0;
  for (int i = 0; i < 3; ++i) {
    L.opd[i] = c->d_opd[i];
    L.opf[i] = c->d_opf[i];
  }
  return L;
}

// Time of a stage of one step from the step's start: LSRK54's c_s dt; RK33's 0, dt, dt/2 (the
// times of Grid.cpp:705, 765-766, 826-827)
double stage_offset(int scheme, int stage, double dt) {
  if (scheme == DG3D_LSRK54) return kLsC[stage] * dt;
  return stage == 0 ? 0.0 : (stage == 1 ? dt : 0.5 * dt);
}

// One fused stage. With dev_clock, t is 0 and the source trace reads the step's start time from
// the device clock, so that the launches do not depend on the step being run. With src_slot >= 0
// the source trace of the stage is already in that slot of the trace buffer.
int run_stage(dg3d_ctx* c, int scheme, int stage, double dt, double t, double** cur, double** other, bool first,
              bool dev_clock = false, int src_slot = -1) {
  dg3d::RhsLaunch L = base_launch(c);
  if (c->tune_variant[scheme] >= 0) {
    L.variant = c->tune_variant[scheme];
    L.xcd_remap = c->tune_xcd[scheme];
  }
  L.mixed = (c->precision == DG3D_MIXED) ? 1 : 0;
  L.mask = dg3d::kTermAll;
  L.dt = dt;
  if (scheme == DG3D_LSRK54) {
    L.epilogue = dg3d::EPI_LSRK;
    L.q_in = *cur;
    L.q_out = *other;
    L.du = c->work_du;
    L.a = kLsA[stage];
    L.b = kLsB[stage];
    hipError_t e = c->ext_rhs(first ? 0 : -1, t + stage_offset(scheme, stage, dt), &L.ext, dev_clock, src_slot);
    if (e != hipSuccess) return c->hip_fail(e, "source trace");
  } else {
    L.epilogue = dg3d::EPI_RK33;
    double* u0 = c->regs[0].d;
    hipError_t e;
    if (stage == 0) {
      L.q_in = u0; L.out = c->work_k1; L.old = nullptr; L.a = 0.0; L.b = 1.0;
    } else if (stage == 1) {
      L.q_in = c->work_k1; L.out = c->work_k2; L.old = u0; L.a = 0.75; L.b = 0.25;
    } else {
      L.q_in = c->work_k2; L.out = u0; L.old = u0; L.a = 1.0 / 3.0; L.b = 2.0 / 3.0;
    }
    e = c->ext_rhs(first && stage == 0 ? 0 : -1, t + stage_offset(scheme, stage, dt), &L.ext, dev_clock, src_slot);
    if (e != hipSuccess) return c->hip_fail(e, "source trace");
  }
  if (c->has_diff && c->nfields == 7) {
    // Grid.cpp:655-667: Q of the stage input from its rho' (in place), then the stage adds
    // OperateDRhoFromQ for the three axes
    dg3d::LdgLaunch Q{};
    Q.n = c->n;
    Q.mesh = c->mesh();
    Q.ext = L.ext;
    for (int i = 0; i < 3; ++i) Q.opd[i] = c->d_opd[i];
    Q.in = L.q_in;
    Q.out = const_cast<double*>(L.q_in);
    hipError_t e = dg3d::launch_ldg_q(Q, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "LDG kernel launch");
    L.ldg = 1;
  }
  hipError_t e = dg3d::launch_rhs(L, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "stage kernel launch");
  if (scheme == DG3D_LSRK54) std::swap(*cur, *other);
  return DG3D_OK;
}

// The launches of one step of dg3d_step: the source traces of all its stages (one launch), the
// stages, then the probe row and the device clock. Only the first stage of a call may read
// register 0's halo traces (`first`).
int launch_step(dg3d_ctx* c, int scheme, double dt, double** cur, double** other, bool first) {
  const int ns = (scheme == DG3D_LSRK54) ? 5 : 3;
  if (c->src.face >= 0) {
    dg3d::ExtArgs x = c->ext_for(first ? 0 : -1, 0.0);
    x.clock = c->d_clock;
    dg3d::StageTimes offs{};
    for (int st = 0; st < ns; ++st) offs.v[st] = stage_offset(scheme, st, dt);
    const hipError_t e =
        dg3d::launch_source_trace_step(x, c->mesh(), c->n, x.trace[c->src.face], c->d_src_trace, offs, ns, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "source trace");
  }
  for (int st = 0; st < ns; ++st) {
    const int rc = run_stage(c, scheme, st, dt, 0.0, cur, other, first && st == 0, true, c->src.face >= 0 ? st : -1);
    if (rc) return rc;
  }
  hipError_t e = hipSuccess;
  if (c->probe_count > 0)
    e = dg3d::launch_probe(*cur, c->mesh(), c->n, c->d_probe_nodes, c->probe_count, c->d_probe_buf, c->d_probe_step,
                           c->probe_cap, c->stream);
  if (e == hipSuccess && (c->src.face >= 0 || c->probe_count > 0))
    e = dg3d::launch_step_end(c->d_clock, dt, c->d_probe_step, c->probe_cap, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "step");
  return DG3D_OK;
}

// The captured form of `steps` launch_step calls for (scheme, dt, cur, other), captured on first
// use.
int step_graph(dg3d_ctx* c, int scheme, double dt, int steps, double* cur, double* other, StepGraph** out) {
  for (StepGraph& g : c->graphs)
    if (g.exec && g.scheme == scheme && g.steps == steps && g.dt == dt && g.cur == cur && g.other == other) {
      *out = &g;
      return DG3D_OK;
    }
  StepGraph* slot = nullptr;
  for (StepGraph& g : c->graphs)
    if (!g.exec) { slot = &g; break; }
  if (!slot) {
    drop_graphs(c);
    slot = &c->graphs[0];
  }
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return c->hip_fail(e, "step graph capture");
  double* a = cur;
  double* b = other;
  int rc = DG3D_OK;
  for (int i = 0; i < steps && !rc; ++i) rc = launch_step(c, scheme, dt, &a, &b, false);
  e = hipStreamEndCapture(c->stream, &graph);
  if (rc || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc ? rc : c->hip_fail(e, "step graph capture");
  }
  e = hipGraphInstantiate(&slot->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    slot->exec = nullptr;
    return c->hip_fail(e, "step graph instantiate");
  }
  slot->scheme = scheme;
  slot->steps = steps;
  slot->dt = dt;
  slot->cur = cur;
  slot->other = other;
  slot->cur_end = a;
  slot->other_end = b;
  *out = slot;
  return DG3D_OK;
}

int reduce_partials(dg3d_ctx* c, int per_block, int blocks) {
  c->h_part.resize((size_t)per_block * blocks);
  hipError_t e = hipMemcpyAsync(c->h_part.data(), c->d_part, sizeof(double) * c->h_part.size(),
                                hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "reduction");
  return DG3D_OK;
}

// Times every (kernel variant, workgroup order) pair on launches of a stage that reads and writes
// the full register set (LSRK54 stage 2; RK33 stage 2 with register 0 as both inputs), after a
// warm-up, in two sweeps of which each pair keeps the faster; the four fastest are timed again in
// longer alternating sweeps, and the three best of those in alternating windows of about 0.25 s, of
// which the faster is kept. The launches write only the work
// registers, which the next step's first stage overwrites without reading, so tuning leaves the
// solution unchanged.
int autotune(dg3d_ctx* c, int scheme) {
  drop_graphs(c);
  int rc = ensure_work(c, scheme);
  if (rc) return rc;
  if (scheme == DG3D_LSRK54) {  // give the du register defined contents before it is read
    hipError_t e = hipMemsetAsync(c->work_du, 0, (c->precision == DG3D_MIXED ? 4 : 8) * 4 * (size_t)c->fs, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "autotune");
  }
  // The tuning launches pass the state back and forth between two work registers, as the stages
  // of a step do (each stage reads what the previous one wrote), starting from a pseudo-random
  // state with every value a normal number. Launches that re-read one register, or that started
  // from an initial condition whose pulse tails are zero or subnormal over most of the mesh, ranked
  // the workgroup orders differently than long runs did (64^3 N=7 split-BF16: order 0 ahead by
  // 1.5 % on the untouched pulse, 2 % behind once it has spread). For LSRK54 the second register
  // is a temporary allocation.
  const size_t reg_bytes = sizeof(double) * (size_t)c->nfields * (size_t)c->fs;
  double* ping = scheme == DG3D_LSRK54 ? c->work_q : c->work_k1;
  double* pong = c->work_k2;
  double* tmp = nullptr;
  hipError_t e = hipSuccess;
  if (scheme == DG3D_LSRK54) {
    e = hipMalloc(reinterpret_cast<void**>(&tmp), reg_bytes);
    if (e != hipSuccess) return c->hip_fail(e, "autotune");
    pong = tmp;
  }
  struct TmpFree {  // on every return: wait for the launches that use it, then free it
    double* p;
    hipStream_t s;
    ~TmpFree() {
      if (!p) return;
      (void)hipStreamSynchronize(s);  // teardown: status unchecked
      (void)hipFree(p);
    }
  } tmp_free{tmp, c->stream};
  e = dg3d::launch_fill_pattern(ping, 1.0e-3, (int64_t)c->nfields * c->fs, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "autotune");
  int parity = 0, stage_count = 0;
  auto tuning_launch = [&](int v, int x) {
    dg3d::RhsLaunch L = base_launch(c);
    L.variant = v;
    L.xcd_remap = x;
    L.mixed = (c->precision == DG3D_MIXED) ? 1 : 0;
    L.mask = dg3d::kTermAll;
    L.dt = 1e-30;
    L.q_in = parity ? pong : ping;
    double* out = parity ? ping : pong;
    parity ^= 1;
    if (scheme == DG3D_LSRK54) {
      // the five stages' coefficients in turn, as a step has them (stage 1 skips the du reads)
      const int st = stage_count++ % 5;
      L.epilogue = dg3d::EPI_LSRK;
      L.q_out = out;
      L.du = c->work_du;
      L.a = kLsA[st];
      L.b = kLsB[st];
    } else {
      L.epilogue = dg3d::EPI_RK33;
      L.old = c->regs[0].d;
      L.out = out;
      L.a = 0.75;
      L.b = 0.25;
    }
    L.ext = c->ext_for(-1, 0.0);
    return L;
  };
  // bring the clocks up before anything is timed
  for (int r = 0; r < 8 && e == hipSuccess; ++r) e = dg3d::launch_rhs(tuning_launch(0, c->xcd), c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "autotune");
  // two sweeps over every (variant, order) pair; each pair keeps its faster sweep
  constexpr int kPairs = dg3d::kNumVariants * dg3d::kNumOrders;
  std::vector<float> best(kPairs, 1e30f);
  const int reps = 4;
  for (int sweep = 0; sweep < 2; ++sweep) {
    for (int v = 0; v < dg3d::kNumVariants; ++v) {
      for (int x = 0; x < dg3d::kNumOrders; ++x) {
        e = dg3d::launch_rhs(tuning_launch(v, x), c->stream);  // first use loads the code object
        if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
        for (int r = 0; r < reps && e == hipSuccess; ++r) e = dg3d::launch_rhs(tuning_launch(v, x), c->stream);
        if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
        if (e != hipSuccess) return c->hip_fail(e, "autotune");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        float& b = best[v * dg3d::kNumOrders + x];
        b = std::min(b, ms);
      }
    }
  }
  // the four fastest pairs again, in three alternating sweeps of sixteen launches, ranked by their
  // mean: at 64^3 N=7 the workgroup orders differ by up to 9 % in long runs, and one noisy first
  // round could pick an order several percent slower than the best
  constexpr int kFinal = 4, kFinalSweeps = 3, kFinalReps = 16;
  std::vector<int> idx(kPairs);
  for (int i = 0; i < kPairs; ++i) idx[i] = i;
  std::partial_sort(idx.begin(), idx.begin() + kFinal, idx.end(), [&](int a, int b) { return best[a] < best[b]; });
  float total[kFinal] = {0.f, 0.f, 0.f, 0.f};
  for (int sweep = 0; sweep < kFinalSweeps; ++sweep) {
    for (int f = 0; f < kFinal; ++f) {
      const int fv = idx[f] / dg3d::kNumOrders, fx = idx[f] % dg3d::kNumOrders;
      e = dg3d::launch_rhs(tuning_launch(fv, fx), c->stream);
      if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
      for (int r = 0; r < kFinalReps && e == hipSuccess; ++r) e = dg3d::launch_rhs(tuning_launch(fv, fx), c->stream);
      if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
      if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
      if (e != hipSuccess) return c->hip_fail(e, "autotune");
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
      total[f] += ms;
    }
  }
  // the three best of those in long alternating windows of about 0.25 s each: over short windows
  // the 64^3 N=7 split-BF16 stage ranked workgroup order 0 first on some boxes, while every long
  // run had order 2 ahead by 6 % (the chip's clocks settle over longer runs than a few launches);
  // with only two in the windows, the 48^3 N=5 split-BF16 stage's sweeps of 16 launches often
  // left out its best order, 2 % ahead in long runs on either data
  int order2[kFinal];
  for (int f = 0; f < kFinal; ++f) order2[f] = f;
  std::sort(order2, order2 + kFinal, [&](int a, int b) { return total[a] < total[b]; });
  const float launch_ms = total[order2[0]] / (kFinalSweeps * kFinalReps);
  const int long_reps = std::max(16, std::min(512, (int)(250.f / std::max(launch_ms, 1e-3f))));
  constexpr int kWindows = 3;
  float window[kWindows] = {0.f, 0.f, 0.f};
  for (int sweep = 0; sweep < 2; ++sweep) {
    for (int f = 0; f < kWindows; ++f) {
      const int fv = idx[order2[f]] / dg3d::kNumOrders, fx = idx[order2[f]] % dg3d::kNumOrders;
      e = dg3d::launch_rhs(tuning_launch(fv, fx), c->stream);
      if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
      for (int r = 0; r < long_reps && e == hipSuccess; ++r) e = dg3d::launch_rhs(tuning_launch(fv, fx), c->stream);
      if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
      if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
      if (e != hipSuccess) return c->hip_fail(e, "autotune");
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
      window[f] += ms;
    }
  }
  int wbest = 0;
  for (int f = 1; f < kWindows; ++f)
    if (window[f] < window[wbest]) wbest = f;
  const int k = idx[order2[wbest]];
  c->tune_variant[scheme] = k / dg3d::kNumOrders;
  c->tune_xcd[scheme] = k % dg3d::kNumOrders;
  // the pattern also filled the slots past the last element, which no stage writes
  e = hipMemsetAsync(ping, 0, reg_bytes, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "autotune");
  return DG3D_OK;
}

}  // namespace

extern "C" {

int dg3d_autotune(dg3d_ctx* c, int scheme, int* variant, int* xcd_remap) {
  if (!c || (scheme != DG3D_LSRK54 && scheme != DG3D_RK33)) return c ? c->fail(DG3D_E_INVALID, "autotune: bad scheme") : DG3D_E_INVALID;
  const int rc = autotune(c, scheme);
  if (rc) return rc;
  if (variant) *variant = c->tune_variant[scheme];
  if (xcd_remap) *xcd_remap = c->tune_xcd[scheme];
  return DG3D_OK;
}

int dg3d_set_variant(dg3d_ctx* c, int scheme, int variant, int xcd_remap) {
  if (!c || (scheme != DG3D_LSRK54 && scheme != DG3D_RK33)) return c ? c->fail(DG3D_E_INVALID, "set_variant: bad scheme") : DG3D_E_INVALID;
  if (variant < 0 || variant >= dg3d::kNumVariants) return c->fail(DG3D_E_INVALID, "set_variant: variant out of range");
  if (xcd_remap < 0 || xcd_remap >= dg3d::kNumOrders) return c->fail(DG3D_E_INVALID, "set_variant: order out of range");
  drop_graphs(c);
  c->tune_variant[scheme] = variant;
  c->tune_xcd[scheme] = xcd_remap;
  return DG3D_OK;
}

int dg3d_set_graphs(dg3d_ctx* c, int enable) {
  if (!c) return DG3D_E_INVALID;
  drop_graphs(c);
  c->use_graphs = enable != 0;
  return DG3D_OK;
}

int dg3d_group_order(int order, const int* num_elements, int elems_per_group, int64_t groups, int64_t* out) {
  if (!num_elements || !out || order < 0 || order >= dg3d::kNumOrders || elems_per_group < 1 || groups < 1 ||
      groups >= (1ll << 31))
    return DG3D_E_INVALID;
  const dg3d::GroupMap g = dg3d::make_group_map(order, (uint32_t)groups, num_elements, elems_per_group);
  for (int64_t w = 0; w < groups; ++w) out[w] = dg3d::map_group((uint32_t)w, g);
  return DG3D_OK;
}

int dg3d_abi_version(void) { return DG3D_ABI_VERSION; }
const char* dg3d_create_error(void) { return g_create_error.c_str(); }
const char* dg3d_last_error(const dg3d_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int dg3d_create(const dg3d_desc* d, dg3d_ctx** out) {
  if (!out) return DG3D_E_INVALID;
  *out = nullptr;
  g_create_error.clear();
  if (!d) { g_create_error = "null descriptor"; return DG3D_E_INVALID; }
  if (d->order < 1 || d->order > 7) { g_create_error = "order must be 1..7"; return DG3D_E_UNSUPPORTED; }
  if (!d->mass || !d->stiff || !d->lgl_nodes || !d->lgl_weights) {
    g_create_error = "mass, stiff, lgl_nodes and lgl_weights are required";
    return DG3D_E_INVALID;
  }
  for (int a = 0; a < 3; ++a)
    if (d->num_elements[a] < 1) { g_create_error = "num_elements must be >= 1"; return DG3D_E_INVALID; }
  if (d->order >= 1 && d->order <= 7 &&
      dg3d::field_slots((int64_t)d->num_elements[0] * d->num_elements[1] * d->num_elements[2], d->order + 1) >=
          (1ll << 29)) {
    // the reductions and the probe index a register's state fields with 32-bit slot numbers
    g_create_error = "a field of the block takes 2^29 or more node slots (the nodes with their padding and group "
                     "gaps): split the domain into blocks (NUM_BLOCKS)";
    return DG3D_E_INVALID;
  }
  if (d->num_registers < 1 || d->num_registers > 64) { g_create_error = "num_registers must be 1..64"; return DG3D_E_INVALID; }
  if (d->precision != DG3D_FP64 && d->precision != DG3D_MIXED) { g_create_error = "bad precision"; return DG3D_E_INVALID; }
  if (d->contraction < DG3D_CONTRACT_VALU || d->contraction > DG3D_CONTRACT_MFMA_BF16) {
    g_create_error = "bad contraction";
    return DG3D_E_INVALID;
  }
  dg3d_ctx* c = new (std::nothrow) dg3d_ctx();
  if (!c) return DG3D_E_NOMEM;
  try {
    c->N = d->order;
    c->n = d->order + 1;
    c->np = c->n * c->n * c->n;
    for (int a = 0; a < 3; ++a) {
      c->nE[a] = d->num_elements[a];
      c->periodic[a] = d->periodic[a] ? 1 : 0;
    }
    c->E = (int64_t)c->nE[0] * c->nE[1] * c->nE[2];
    c->es = dg3d::elem_stride(c->n);
    c->grp = dg3d::group_elems(c->n);
    c->fs = dg3d::field_slots(c->E, c->n);
    c->nfields = d->with_q ? 7 : 4;
    c->precision = d->precision;
    c->contraction = d->contraction;
    c->xcd = (d->xcd_remap >= 0 && d->xcd_remap < dg3d::kNumOrders) ? d->xcd_remap : 0;
    c->nodes.assign(d->lgl_nodes, d->lgl_nodes + c->n);
    c->weights.assign(d->lgl_weights, d->lgl_weights + c->n);
    dg3d::LineOperator op;
    if (!dg3d::make_line_operator(c->n, d->mass, d->stiff, &op)) {
      g_create_error = "mass matrix is singular";
      free_ctx(c);
      return DG3D_E_INVALID;
    }
    c->err.reserve(256);
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);
    if (e == hipSuccess) {
      // once per device, with the context's device current: the scratch array's address
      c->scratch_err = dg3d::scratch_address(c->device, &c->scratch);
      if (c->scratch_err != hipSuccess) c->scratch = nullptr;
    }
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
      g_create_error = std::string("HIP init: ") + hipGetErrorString(e);
      free_ctx(c);
      return DG3D_E_HIP;
    }
    hipError_t up = hipSuccess;  // first failure of the uploads below, reported after them
    auto upload = [&up](hipError_t r) {
      if (up == hipSuccess) up = r;
    };
    const std::vector<double>* srcs[3] = {&op.A, &op.L0, &op.LN};
    for (int i = 0; i < 3; ++i) {
      const size_t cnt = srcs[i]->size();
      std::vector<float> f(srcs[i]->begin(), srcs[i]->end());
      if (dmalloc(&c->d_opd[i], cnt) != hipSuccess || dmalloc(&c->d_opf[i], cnt) != hipSuccess) {
        g_create_error = "operator allocation failed";
        free_ctx(c);
        return DG3D_E_NOMEM;
      }
      upload(hipMemcpy(c->d_opd[i], srcs[i]->data(), sizeof(double) * cnt, hipMemcpyHostToDevice));
      upload(hipMemcpy(c->d_opf[i], f.data(), sizeof(float) * cnt, hipMemcpyHostToDevice));
    }
    for (int a = 0; a < 3; ++a) {
      c->h_w[a].assign(c->nE[a], 1.0);
      std::vector<double> iw(c->nE[a], 1.0);
      if (dmalloc(&c->d_inv_w[a], c->nE[a]) != hipSuccess) { g_create_error = "geometry allocation failed"; free_ctx(c); return DG3D_E_NOMEM; }
      upload(hipMemcpy(c->d_inv_w[a], iw.data(), sizeof(double) * iw.size(), hipMemcpyHostToDevice));
    }
    c->h_mat.assign((size_t)c->E * 4, 0.0);
    for (int64_t i = 0; i < c->E; ++i) {
      c->h_mat[i * 4 + 0] = 1500.0;
      c->h_mat[i * 4 + 1] = 1000.0;
    }
    c->h_diff.assign((size_t)c->E, 0.0);
    if (dmalloc(&c->d_mat, (size_t)c->E * dg3d::kMatStride) != hipSuccess || dmalloc(&c->d_diff, c->h_diff.size()) != hipSuccess ||
        dmalloc(&c->d_wq, c->n) != hipSuccess || dmalloc(&c->d_part, 1024 * 16) != hipSuccess ||
        dmalloc(&c->d_clock, 1) != hipSuccess || dmalloc(&c->d_probe_step, 1) != hipSuccess) {
      g_create_error = "material allocation failed";
      free_ctx(c);
      return DG3D_E_NOMEM;
    }
    upload(hipMemcpy(c->d_wq, c->weights.data(), sizeof(double) * c->n, hipMemcpyHostToDevice));
    upload(hipMemset(c->d_clock, 0, sizeof(double)));
    upload(hipMemset(c->d_probe_step, 0, sizeof(int)));
    if (upload_material(c) != DG3D_OK) { g_create_error = c->err; free_ctx(c); return DG3D_E_HIP; }
    c->regs.resize(d->num_registers);
    for (auto& r : c->regs) {
      if (dmalloc(&r.d, (size_t)c->nfields * c->fs) != hipSuccess) {
        g_create_error = "register allocation failed (" + std::to_string((double)c->nfields * c->fs * 8 / 1e9) + " GB each)";
        free_ctx(c);
        return DG3D_E_NOMEM;
      }
      upload(hipMemsetAsync(r.d, 0, sizeof(double) * c->nfields * c->fs, c->stream));
    }
    c->src.face = -1;
    e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = up;
    if (e != hipSuccess) { g_create_error = std::string("HIP: ") + hipGetErrorString(e); free_ctx(c); return DG3D_E_HIP; }
  } catch (const std::exception& ex) {
    g_create_error = ex.what();
    free_ctx(c);
    return DG3D_E_NOMEM;
  }
  *out = c;
  return DG3D_OK;
}

void dg3d_destroy(dg3d_ctx* ctx) {
  if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_ctx(ctx);
}

int dg3d_set_geometry(dg3d_ctx* c, const double* corner, const double* wx, const double* wy, const double* wz) {
  if (!c) return DG3D_E_INVALID;
  const double* w[3] = {wx, wy, wz};
  for (int a = 0; a < 3; ++a) {
    if (corner) c->corner[a] = corner[a];
    if (!w[a]) continue;
    std::vector<double> iw(c->nE[a]);
    for (int i = 0; i < c->nE[a]; ++i) {
      if (!(w[a][i] > 0.0) || !std::isfinite(w[a][i])) return c->fail(DG3D_E_INVALID, "element widths must be positive");
      c->h_w[a][i] = w[a][i];
      iw[i] = 1.0 / w[a][i];
    }
    hipError_t e = hipMemcpy(c->d_inv_w[a], iw.data(), sizeof(double) * iw.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return c->hip_fail(e, "geometry upload");
  }
  return DG3D_OK;
}

int dg3d_set_material(dg3d_ctx* c, const double* c0, const double* rho0, const double* ba, const double* sigma,
                      const double* diff) {
  if (!c) return DG3D_E_INVALID;
  for (int64_t i = 0; i < c->E; ++i) {
    if (c0) c->h_mat[i * 4 + 0] = c0[i];
    if (rho0) c->h_mat[i * 4 + 1] = rho0[i];
    if (ba) c->h_mat[i * 4 + 2] = ba[i];
    if (sigma) c->h_mat[i * 4 + 3] = sigma[i];
    if (diff) c->h_diff[i] = diff[i];
    if (!(c->h_mat[i * 4 + 0] > 0.0) || !(c->h_mat[i * 4 + 1] > 0.0))
      return c->fail(DG3D_E_INVALID, "c0 and rho0 must be positive");
  }
  drop_graphs(c);  // whether the stages run the LDG term depends on D
  return upload_material(c);
}

int dg3d_set_state(dg3d_ctx* c, int reg, const double* host, int nfields) {
  if (!c || !c->reg_ok(reg) || !host) return c ? c->fail(DG3D_E_INVALID, "set_state: bad register or pointer") : DG3D_E_INVALID;
  if (nfields < 1 || nfields > c->nfields) return c->fail(DG3D_E_INVALID, "set_state: nfields exceeds the register");
  const hipError_t e = transfer_state(c, c->regs[reg].d, const_cast<double*>(host), nfields, true);
  if (e != hipSuccess) return c->hip_fail(e, "set_state");
  return DG3D_OK;
}

int dg3d_get_state(dg3d_ctx* c, int reg, double* host, int nfields) {
  if (!c || !c->reg_ok(reg) || !host) return c ? c->fail(DG3D_E_INVALID, "get_state: bad register or pointer") : DG3D_E_INVALID;
  if (nfields < 1 || nfields > c->nfields) return c->fail(DG3D_E_INVALID, "get_state: nfields exceeds the register");
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = transfer_state(c, c->regs[reg].d, host, nfields, false);
  if (e != hipSuccess) return c->hip_fail(e, "get_state");
  return DG3D_OK;
}

int dg3d_set_halo(dg3d_ctx* c, int reg, int face, const double* trace) {
  if (!c || !c->reg_ok(reg) || face < 0 || face > 5) return c ? c->fail(DG3D_E_INVALID, "set_halo: bad register or face") : DG3D_E_INVALID;
  Register& r = c->regs[reg];
  if (!trace) {
    r.halo_valid[face] = false;
    return DG3D_OK;
  }
  const size_t cnt = (size_t)c->face_elems(face) * c->n * c->n * 3;
  if (!r.halo[face] && dmalloc(&r.halo[face], cnt) != hipSuccess) return c->fail(DG3D_E_NOMEM, "halo allocation failed");
  hipError_t e = hipMemcpy(r.halo[face], trace, sizeof(double) * cnt, hipMemcpyHostToDevice);
  if (e != hipSuccess) return c->hip_fail(e, "set_halo");
  r.halo_valid[face] = true;
  return DG3D_OK;
}

int dg3d_zero(dg3d_ctx* c, int reg) {
  if (!c || !c->reg_ok(reg)) return c ? c->fail(DG3D_E_INVALID, "zero: bad register") : DG3D_E_INVALID;
  Register& r = c->regs[reg];
  hipError_t e = hipMemsetAsync(r.d, 0, sizeof(double) * c->nfields * c->fs, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "zero");
  for (bool& v : r.halo_valid) v = false;
  return DG3D_OK;
}

int dg3d_copy(dg3d_ctx* c, int dst, int src) {
  if (!c || !c->reg_ok(dst) || !c->reg_ok(src)) return c ? c->fail(DG3D_E_INVALID, "copy: bad register") : DG3D_E_INVALID;
  if (dst == src) return DG3D_OK;
  Register& d = c->regs[dst];
  const Register& s = c->regs[src];
  hipError_t e = hipMemcpyAsync(d.d, s.d, sizeof(double) * c->nfields * c->fs, hipMemcpyDeviceToDevice, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "copy");
  for (int f = 0; f < 6; ++f) {
    d.halo_valid[f] = s.halo_valid[f];
    if (!s.halo_valid[f]) continue;
    const size_t cnt = (size_t)c->face_elems(f) * c->n * c->n * 3;
    if (!d.halo[f] && dmalloc(&d.halo[f], cnt) != hipSuccess) return c->fail(DG3D_E_NOMEM, "halo allocation failed");
    e = hipMemcpyAsync(d.halo[f], s.halo[f], sizeof(double) * cnt, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "copy halo");
  }
  return DG3D_OK;
}

int dg3d_axpby(dg3d_ctx* c, int out, double self_scalar, int in, double other_scalar) {
  if (!c || !c->reg_ok(out) || !c->reg_ok(in)) return c ? c->fail(DG3D_E_INVALID, "axpby: bad register") : DG3D_E_INVALID;
  hipError_t e = dg3d::launch_axpby(c->regs[out].d, self_scalar, c->regs[in].d, other_scalar,
                                    (int64_t)c->nfields * c->fs, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "axpby");
  return DG3D_OK;
}

int dg3d_operate(dg3d_ctx* c, int op, int axis, int mode, int in_reg, int out_reg) {
  if (!c || !c->reg_ok(in_reg) || !c->reg_ok(out_reg)) return c ? c->fail(DG3D_E_INVALID, "operate: bad register") : DG3D_E_INVALID;
  if (op == DG3D_OP_UFROMV || op == DG3D_OP_SPONGE) {
    if (op == DG3D_OP_UFROMV) {
      if (axis < 0 || axis > 2 || mode < 0 || mode > 1) return c->fail(DG3D_E_INVALID, "operate: axis 0..2, mode 0..1");
      if (in_reg == out_reg) return c->fail(DG3D_E_INVALID, "operate: UFromV writes a register other than the one it reads");
    }
    dg3d::RhsLaunch L = base_launch(c);
    L.mixed = 0;
    L.epilogue = dg3d::EPI_ACCUM;
    L.mask = (op == DG3D_OP_SPONGE) ? dg3d::kTermSponge : (1u << (2 * axis + mode));
    L.q_in = c->regs[in_reg].d;
    L.out = c->regs[out_reg].d;
    hipError_t e = c->ext_rhs(in_reg, 0.0, &L.ext);
    if (e == hipSuccess) e = dg3d::launch_rhs(L, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "operate");
    return DG3D_OK;
  }
  if (op == DG3D_OP_QFROMDRHO || op == DG3D_OP_DRHOFROMQ) {
    if (c->nfields < 7) return c->fail(DG3D_E_INVALID, "operate: LDG operators need with_q");
    if (axis < 0 || axis > 2) return c->fail(DG3D_E_INVALID, "operate: axis 0..2");
    if (op == DG3D_OP_DRHOFROMQ && in_reg == out_reg)
      return c->fail(DG3D_E_INVALID, "operate: DRhoFromQ writes a register other than the one it reads");
    dg3d::LdgLaunch L{};
    L.n = c->n;
    L.which = (op == DG3D_OP_QFROMDRHO) ? 0 : 1;
    L.axis = axis;
    L.mesh = c->mesh();
    L.ext = c->ext_for(in_reg, 0.0);
    L.diff = c->d_diff;
    for (int i = 0; i < 3; ++i) L.opd[i] = c->d_opd[i];
    L.in = c->regs[in_reg].d;
    L.out = c->regs[out_reg].d;
    hipError_t e = dg3d::launch_ldg(L, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "operate");
    return DG3D_OK;
  }
  return c->fail(DG3D_E_INVALID, "operate: unknown operator");
}

int dg3d_halo(dg3d_ctx* c, int reg) {
  if (!c || !c->reg_ok(reg)) return c ? c->fail(DG3D_E_INVALID, "halo: bad register") : DG3D_E_INVALID;
  return DG3D_OK;
}

int dg3d_set_source(dg3d_ctx* c, const dg3d_source* s) {
  if (!c) return DG3D_E_INVALID;
  const hipError_t se = hipStreamSynchronize(c->stream);  // launches in flight may read the old trace
  if (se != hipSuccess) return c->hip_fail(se, "set_source");
  drop_graphs(c);
  (void)hipFree(c->d_lead);
  (void)hipFree(c->d_src_trace);
  c->d_lead = nullptr;
  c->d_src_trace = nullptr;
  if (!s || s->face < 0) {
    c->src = dg3d_source{};
    c->src.face = -1;
    return DG3D_OK;
  }
  if (s->face > 5) return c->fail(DG3D_E_INVALID, "source face must be 0..5");
  if (c->periodic[2 - s->face / 2]) return c->fail(DG3D_E_INVALID, "source on a periodic face");
  if (!(s->tau > 0.0)) return c->fail(DG3D_E_INVALID, "source tau must be positive");
  c->src = *s;
  c->src.lead = nullptr;
  if (dmalloc(&c->d_src_trace, (size_t)c->face_elems(s->face) * c->n * c->n * 3 * dg3d::kMaxStages) != hipSuccess)
    return c->fail(DG3D_E_NOMEM, "source allocation failed");
  if (s->lead) {
    const size_t cnt = (size_t)c->face_elems(s->face) * c->n * c->n;
    if (dmalloc(&c->d_lead, cnt) != hipSuccess) return c->fail(DG3D_E_NOMEM, "source allocation failed");
    hipError_t e = hipMemcpy(c->d_lead, s->lead, sizeof(double) * cnt, hipMemcpyHostToDevice);
    if (e != hipSuccess) return c->hip_fail(e, "source upload");
  }
  return DG3D_OK;
}

int dg3d_stage(dg3d_ctx* c, int scheme, int stage, double dt, double t) {
  if (!c) return DG3D_E_INVALID;
  const int ns = (scheme == DG3D_LSRK54) ? 5 : 3;
  if ((scheme != DG3D_LSRK54 && scheme != DG3D_RK33) || stage < 0 || stage >= ns)
    return c->fail(DG3D_E_INVALID, "stage: bad scheme or stage");
  int rc = ensure_work(c, scheme);
  if (rc) return rc;
  double* cur = c->regs[0].d;
  double* other = c->work_q;
  const bool first = (stage == 0);
  rc = run_stage(c, scheme, stage, dt, t, &cur, &other, first);
  if (rc) return rc;
  if (scheme == DG3D_LSRK54) {
    // the new state becomes register 0; the old buffer becomes the work register
    if (c->nfields > 4) {
      const hipError_t e = hipMemcpyAsync(cur + 4 * c->fs, c->regs[0].d + 4 * c->fs, sizeof(double) * 3 * c->fs,
                                          hipMemcpyDeviceToDevice, c->stream);
      if (e != hipSuccess) return c->hip_fail(e, "stage");
    }
    c->work_q = c->regs[0].d;
    c->regs[0].d = cur;
  }
  if (first)
    for (bool& v : c->regs[0].halo_valid) v = false;
  return DG3D_OK;
}

int dg3d_step(dg3d_ctx* c, int scheme, double dt, int nsteps, double* t) {
  if (!c || !t) return DG3D_E_INVALID;
  if (scheme != DG3D_LSRK54 && scheme != DG3D_RK33) return c->fail(DG3D_E_INVALID, "step: unknown scheme");
  if (nsteps < 0) return c->fail(DG3D_E_INVALID, "step: nsteps < 0");
  if (!(dt > 0.0) || !std::isfinite(dt)) return c->fail(DG3D_E_INVALID, "step: dt must be positive");
  if (c->has_diff && c->nfields < 7)
    return c->fail(DG3D_E_INVALID, "step: diffusion (D != 0) needs a context created with with_q");
  int rc = ensure_work(c, scheme);
  if (rc) return rc;
  if (nsteps == 0) return DG3D_OK;
  if (c->tune_variant[scheme] < 0) {
    rc = autotune(c, scheme);
    if (rc) return rc;
  }
  const int ns = (scheme == DG3D_LSRK54) ? 5 : 3;
  double* cur = c->regs[0].d;
  double* other = c->work_q;
  double tt = *t;
  hipError_t e = hipEventRecord(c->ev0, c->stream);
  if (e == hipSuccess && c->src.face >= 0) e = dg3d::launch_clock_set(c->d_clock, tt, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "step");
  bool halos = false;
  for (bool v : c->regs[0].halo_valid) halos = halos || v;
  for (int s = 0; s < nsteps;) {
    int run = 1;
    if (!c->use_graphs || (s == 0 && halos)) {
      // the first stage reads the halo traces set on register 0 (initial-condition halos)
      rc = launch_step(c, scheme, dt, &cur, &other, s == 0);
      if (rc) return rc;
      if (s == 0)
        for (bool& v : c->regs[0].halo_valid) v = false;
    } else {
      run = kGraphSteps;
      while (run > nsteps - s) run /= 2;
      StepGraph* g = nullptr;
      rc = step_graph(c, scheme, dt, run, cur, other, &g);
      if (rc) return rc;
      e = hipGraphLaunch(g->exec, c->stream);
      if (e != hipSuccess) return c->hip_fail(e, "step graph launch");
      cur = g->cur_end;
      other = g->other_end;
    }
    for (int i = 0; i < run; ++i) {
      tt += dt;
      if (c->probe_count > 0 && c->probe_steps < c->probe_cap) ++c->probe_steps;
    }
    s += run;
  }
  e = hipEventRecord(c->ev1, c->stream);
  if (e == hipSuccess && scheme == DG3D_LSRK54 && cur != c->regs[0].d) {
    if (c->nfields > 4)  // Q rides along in register 0 (Element.cpp:288-307)
      e = hipMemcpyAsync(cur + 4 * c->fs, c->regs[0].d + 4 * c->fs, sizeof(double) * 3 * c->fs,
                         hipMemcpyDeviceToDevice, c->stream);
    c->work_q = c->regs[0].d;
    c->regs[0].d = cur;
  }
  if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
  if (e != hipSuccess) return c->hip_fail(e, "step");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->last_ms = ms;
  c->last_launches = nsteps * ns;
  *t = tt;
  return DG3D_OK;
}

int dg3d_last_step_timing(dg3d_ctx* c, double* avg_stage_ms, int* launches) {
  if (!c) return DG3D_E_INVALID;
  if (avg_stage_ms) *avg_stage_ms = c->last_launches ? c->last_ms / c->last_launches : 0.0;
  if (launches) *launches = c->last_launches;
  return DG3D_OK;
}

int dg3d_cfl_dt(dg3d_ctx* c, double cfl, double* dt) {
  if (!c || !dt) return DG3D_E_INVALID;
  double dxi = 2.0;
  for (int i = 0; i + 1 < c->n; ++i) dxi = std::min(dxi, c->nodes[i + 1] - c->nodes[i]);
  hipError_t e = dg3d::launch_cfl(c->mesh(), 0.5 * dxi, c->d_part, c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "cfl");
  const int blocks = dg3d::reduce_blocks(c->E);
  int rc = reduce_partials(c, 1, blocks);
  if (rc) return rc;
  double m = 1e300;
  for (int b = 0; b < blocks; ++b) m = std::min(m, c->h_part[b]);
  *dt = cfl * m;
  return DG3D_OK;
}

namespace {

// Per-block values a reduction kind writes, and the combine kind of dg3d::launch_reduce_combine.
bool reduce_shape(int kind, int* per_block, int* combine) {
  if (kind == DG3D_REDUCE_MAXABS) { *per_block = 4; *combine = 0; return true; }
  if (kind == DG3D_REDUCE_ENERGY) { *per_block = 2; *combine = 1; return true; }
  if (kind == DG3D_REDUCE_RELL2 || kind == DG3D_REDUCE_DIFFL2) { *per_block = 8; *combine = 2; return true; }
  return false;
}

hipError_t launch_reduction(dg3d_ctx* c, int kind, int reg, int reg2, double* part) {
  if (kind == DG3D_REDUCE_MAXABS) return dg3d::launch_maxabs(c->regs[reg].d, c->fs, part, c->stream);
  if (kind == DG3D_REDUCE_ENERGY) return dg3d::launch_energy(c->regs[reg].d, c->mesh(), c->n, c->d_wq, part, c->stream);
  return dg3d::launch_l2diff(c->regs[reg].d, c->regs[reg2].d, c->fs, part, c->stream);
}

// The outputs of a reduction kind from its combined values (v: 4 maxima, the 2 energy sums, or
// the 4 squared-difference sums then the 4 squared-reference sums).
int finish_reduce(dg3d_ctx* c, int kind, const double* v, double* out) {
  if (kind == DG3D_REDUCE_MAXABS) {
    for (int f = 0; f < 4; ++f) out[f] = v[f];
    return DG3D_OK;
  }
  if (kind == DG3D_REDUCE_ENERGY) {
    out[0] = v[0];
    out[1] = v[1];
    if (!std::isfinite(out[0])) return c->fail_text(DG3D_E_NAN, "energy is not finite");
    return DG3D_OK;
  }
  for (int f = 0; f < 4; ++f) {
    if (kind == DG3D_REDUCE_RELL2) out[f] = (v[4 + f] > 0.0) ? std::sqrt(v[f] / v[4 + f]) : std::sqrt(v[f]);
    else {
      out[f] = std::sqrt(v[f]);
      out[4 + f] = std::sqrt(v[4 + f]);
    }
  }
  return DG3D_OK;
}

}  // namespace

int dg3d_reduce(dg3d_ctx* c, int kind, int reg, int reg2, double* out) {
  if (!c || !c->reg_ok(reg) || !out) return c ? c->fail(DG3D_E_INVALID, "reduce: bad register") : DG3D_E_INVALID;
  int per_block = 0, combine = 0;
  if (!reduce_shape(kind, &per_block, &combine)) return c->fail(DG3D_E_INVALID, "reduce: unknown kind");
  if (combine == 2 && !c->reg_ok(reg2)) return c->fail(DG3D_E_INVALID, "reduce: bad second register");
  const int blocks = dg3d::reduce_blocks(c->fs);
  hipError_t e = launch_reduction(c, kind, reg, reg2, c->d_part);
  if (e != hipSuccess) return c->hip_fail(e, "reduce");
  int rc = reduce_partials(c, per_block, blocks);
  if (rc) return rc;
  // combined on the host in block order (dg3d::launch_reduce_combine combines in the same order)
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double* h = c->h_part.data();
  if (combine == 0) {
    for (int f = 0; f < 4; ++f) {
      double m = 0.0;
      for (int b = 0; b < blocks; ++b) {
        const double x = h[b * 4 + f];
        if (std::isnan(x) || std::isnan(m)) m = std::nan("");
        else m = std::max(m, x);
      }
      v[f] = m;
    }
  } else if (combine == 1) {
    for (int b = 0; b < blocks; ++b) {
      v[0] += h[b * 2];
      v[1] += h[b * 2 + 1];
    }
  } else {
    for (int b = 0; b < blocks; ++b)
      for (int f = 0; f < 4; ++f) {
        v[f] += h[(b * 4 + f) * 2];
        v[4 + f] += h[(b * 4 + f) * 2 + 1];
      }
  }
  return finish_reduce(c, kind, v, out);
}

int dg3d_reduce_noalloc(dg3d_ctx* c, int kind, int reg, int reg2, double* out) {
  if (!c) return DG3D_E_INVALID;
  if (!c->reg_ok(reg) || !out) return c->fail_text(DG3D_E_INVALID, "reduce_noalloc: bad register or output");
  int per_block = 0, combine = 0;
  if (!reduce_shape(kind, &per_block, &combine)) return c->fail_text(DG3D_E_INVALID, "reduce_noalloc: unknown kind");
  if (combine == 2 && !c->reg_ok(reg2)) return c->fail_text(DG3D_E_INVALID, "reduce_noalloc: bad second register");
  if (!c->scratch)
    return c->fail_text(DG3D_E_HIP, "reduce_noalloc: the device scratch could not be resolved: ",
                        hipGetErrorString(c->scratch_err));
  const int blocks = dg3d::reduce_blocks(c->fs);
  if ((int64_t)blocks * per_block > dg3d::kScratchPartials)
    return c->fail_text(DG3D_E_INVALID, "reduce_noalloc: the reduction needs more partials than the scratch holds");
  double* part = c->scratch;
  double* vals = c->scratch + dg3d::kScratchPartials;
  hipError_t e = launch_reduction(c, kind, reg, reg2, part);
  if (e == hipSuccess) e = dg3d::launch_reduce_combine(part, blocks, combine, vals, c->stream);
  double v[dg3d::kScratchValues];
  if (e == hipSuccess) e = hipMemcpyAsync(v, vals, sizeof(v), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->fail_text(DG3D_E_HIP, "reduce_noalloc: ", hipGetErrorString(e));
  return finish_reduce(c, kind, v, out);
}

int dg3d_set_probes(dg3d_ctx* c, int count, const int64_t* nodes, int capacity_steps) {
  if (!c || count < 0 || capacity_steps < 0) return DG3D_E_INVALID;
  const hipError_t se = hipStreamSynchronize(c->stream);  // launches in flight may write the old buffer
  if (se != hipSuccess) return c->hip_fail(se, "set_probes");
  drop_graphs(c);
  hipError_t e0 = hipMemsetAsync(c->d_probe_step, 0, sizeof(int), c->stream);
  if (e0 != hipSuccess) return c->hip_fail(e0, "set_probes");
  (void)hipFree(c->d_probe_nodes);
  (void)hipFree(c->d_probe_buf);
  c->d_probe_nodes = nullptr;
  c->d_probe_buf = nullptr;
  c->probe_count = 0;
  c->probe_cap = 0;
  c->probe_steps = 0;
  if (count == 0) return DG3D_OK;
  std::vector<int64_t> slots(count);
  for (int i = 0; i < count; ++i) {
    if (nodes[i] < 0 || nodes[i] >= c->E * c->np) return c->fail(DG3D_E_INVALID, "probe node out of range");
    slots[i] = nodes[i];  // host numbering; the probe kernel finds the slot
  }
  if (dmalloc(&c->d_probe_nodes, count) != hipSuccess ||
      dmalloc(&c->d_probe_buf, (size_t)count * std::max(capacity_steps, 1)) != hipSuccess)
    return c->fail(DG3D_E_NOMEM, "probe allocation failed");
  const hipError_t pe = hipMemcpy(c->d_probe_nodes, slots.data(), sizeof(int64_t) * count, hipMemcpyHostToDevice);
  if (pe != hipSuccess) return c->hip_fail(pe, "set_probes");
  c->probe_count = count;
  c->probe_cap = capacity_steps;
  return DG3D_OK;
}

int dg3d_get_probes(dg3d_ctx* c, double* out, int* steps) {
  if (!c) return DG3D_E_INVALID;
  if (steps) *steps = c->probe_steps;
  if (out && c->probe_steps > 0) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess)
      e = hipMemcpy(out, c->d_probe_buf, sizeof(double) * c->probe_count * c->probe_steps, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "get_probes");
  }
  return DG3D_OK;
}

int dg3d_sync(dg3d_ctx* c) {
  if (!c) return DG3D_E_INVALID;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "sync");
  return DG3D_OK;
}

}  // extern "C"
