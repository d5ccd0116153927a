// Device-side parameter blocks and launcher declarations shared by the C ABI (dg3d_capi.cpp)
// and the kernels (dg3d_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace dg3d {

constexpr int kMaxN = 8;  // n = N+1 <= 8 (orders 1..7)

// LSRK(5,4) of Carpenter & Kennedy (NASA TM-109112, 1994), 2N-storage: per stage
// du = A_s du + dt rhs(t + C_s dt); q = q + B_s du.
inline constexpr double kLsA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                                   -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
inline constexpr double kLsB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                                   1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                                   2277821191437.0 / 14882151754819.0};
inline constexpr double kLsC[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                                   2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

// Term mask bits of the RHS kernel: UFromV(axis, mode) = bit (2*axis + mode), sponge = bit 6.
constexpr unsigned kTermSponge = 1u << 6;
constexpr unsigned kTermAll = 0x7Fu;

enum Epilogue : int {
  EPI_ACCUM = 0,  // out += rhs over the masked fields (Block::Operate* into a register)
  EPI_LSRK = 1,   // du = a du + dt rhs; q_out = q + b du
  EPI_RK33 = 2,   // out = (rhs dt beta + q beta) + old alpha
};

// Division of n < 2^31 by a runtime divisor d without a divide instruction:
// q = (n * m) >> (31 + s) with s = ceil(log2 d), m = ceil(2^(31+s) / d) < 2^32.
struct Div32 {
  uint32_t m;
  uint32_t s;
};
inline Div32 make_div32(uint32_t d) {
  uint32_t s = 0;
  while ((1ull << s) < d) ++s;
  const uint64_t num = 1ull << (31 + s);
  return Div32{(uint32_t)((num + d - 1) / d), s};
}
__host__ __device__ inline uint32_t div32(uint32_t n, Div32 d) {
  return (uint32_t)(((uint64_t)n * d.m) >> (31 + d.s));
}

// Workgroup -> element-group map of the stage kernels. The hardware dispatches workgroup w to
// XCD w mod 8, and each XCD has its own L2, so the order decides which neighbour traces an XCD
// finds in its L2:
//   0: hardware order (XCD x takes groups x, x+8, ...);
//   1: XCD x takes one contiguous slab of G/8 groups;
//   2: XCD x takes runs of C consecutive groups (runs x, x+8, ... in time), with C one y-row of
//      groups split over the 8 XCDs, so that the groups one row (y) and one plane (x) away fall
//      on the same XCD while the 8 XCDs stay in one compact window of memory;
//   3: order 1's slabs over a traversal tiled in blocks of TJ element rows (y) and TK groups (z),
//      i outermost inside a block, then j, then the group: the groups one plane, one row and
//      one element away are at most TJ*TK, TK and 1 positions apart. Falls back to order 1
//      when the mesh does not divide.
// The host prepares every divisor (G < 2^31 groups), so a workgroup spends a few scalar
// instructions on it rather than 64-bit division sequences before its first load.
constexpr int kTileJ = 8, kTileK = 8;
struct GroupMap {
  int mode;
  uint32_t per, rem;          // orders 1, 3: G / 8, G % 8
  uint32_t C, full;           // order 2: run length, groups covered by whole 8-run rounds
  Div32 dC;
  int tiled;                  // order 3: the mesh divides into tiles
  uint32_t per_jb, per_kb, tt, tk, tj, kg, nE1;
  Div32 d_jb, d_kb, d_tt, d_tk;
};

inline GroupMap make_group_map(int mode, uint32_t G, const int* nE, int epb) {
  GroupMap g{};
  g.mode = mode;
  g.per = G / 8;
  g.rem = G % 8;
  if (mode == 2) {
    uint32_t C = 4;
    if (nE[2] % epb == 0) {
      const uint32_t sy = (uint32_t)(nE[2] / epb);
      C = (sy % 8 == 0) ? sy / 8 : sy;
    }
    g.C = C;
    g.full = G / (8 * C) * (8 * C);
    g.dC = make_div32(C);
  }
  if (mode == 3 && nE[2] % epb == 0) {
    const uint32_t kg = (uint32_t)(nE[2] / epb);
    const uint32_t tk = kg < (uint32_t)kTileK ? kg : kTileK;
    const uint32_t tj = (uint32_t)nE[1] < (uint32_t)kTileJ ? (uint32_t)nE[1] : kTileJ;
    if (kg % tk == 0 && nE[1] % tj == 0) {
      g.tiled = 1;
      g.kg = kg;
      g.tk = tk;
      g.tj = tj;
      g.nE1 = (uint32_t)nE[1];
      g.per_jb = tj * (uint32_t)nE[0] * kg;
      g.per_kb = (uint32_t)nE[0] * tj * tk;
      g.tt = tj * tk;
      g.d_jb = make_div32(g.per_jb);
      g.d_kb = make_div32(g.per_kb);
      g.d_tt = make_div32(g.tt);
      g.d_tk = make_div32(tk);
    }
  }
  return g;
}

__host__ __device__ inline uint32_t map_group(uint32_t w, const GroupMap& g) {
  if (g.mode == 1 || g.mode == 3) {
    const uint32_t x = w & 7u, i = w >> 3;
    const uint32_t s = x * g.per + (x < g.rem ? x : g.rem) + i;
    if (g.mode == 1 || !g.tiled) return s;
    const uint32_t jb = div32(s, g.d_jb);
    uint32_t r = s - jb * g.per_jb;
    const uint32_t kb = div32(r, g.d_kb);
    r -= kb * g.per_kb;
    const uint32_t ii = div32(r, g.d_tt);
    r -= ii * g.tt;
    const uint32_t jj = div32(r, g.d_tk), kk = r - jj * g.tk;
    return (ii * g.nE1 + jb * g.tj + jj) * g.kg + kb * g.tk + kk;
  }
  if (g.mode == 2) {
    if (w >= g.full) return w;
    const uint32_t x = w & 7u, i = w >> 3;
    const uint32_t q = div32(i, g.dC);
    return q * (8 * g.C) + x * g.C + (i - q * g.C);
  }
  return w;
}

// Device layout of a register: group-blocked.
//   * Each element has ES(n) node slots per field: n^3 rounded up to a multiple of 32 slots
//     (n = 3, 5, 6, 7), so every element's field starts on a 128-byte line and a workgroup's
//     elements cover whole lines: no line is written in part by two workgroups. The padding slots
//     hold zeros. n = 2 keeps 8 slots: its workgroups take 16 or 32 elements, whole lines too.
//   * Consecutive elements form groups of G(n) elements. A group's state fields rho', u, v, w are
//     one block of 4 runs of kRunSlots slots: field f of element e at
//     (e / G) 4 P + f P + (e % G) ES + node, P = kRunSlots. G ES <= P; the slots past a group's
//     elements in each run are a gap that nothing reads. The LDG fields Q_x, Q_y, Q_z of a
//     7-field register follow from slot 4 fs in blocks of 3 runs: field 4 + a at
//     4 fs + (e / G) 3 P + a P + (e % G) ES + node.
//   * Why runs of 4096 FP64 values: a probe of the LSRK stage's stream (state read, RK register
//     read and written in place, state written; 64^3 N=4 FP64) ran at 83-84 % of the HBM peak
//     with runs of 4096 slots at a spacing of 4096, at 79-83 % with shorter runs of 3520-4032
//     slots at that spacing, but at 74-76 % with unpadded runs of 3584-6144 slots, 71-73 % with
//     runs of one element (128 slots) or with every field its own array, and at 68-70 % whenever
//     the spacing was not a power of two. Every workgroup's elements lie in one group (G is a
//     multiple of every elements-per-workgroup shape of the order), so a workgroup streams one
//     piece of each of the 4 runs of its group.
//   * A register takes nfields * fs slots, fs = ceil(E / G) P (the gaps and the slots of a last,
//     partial group hold zeros).
constexpr int kRunSlots = 4096;
constexpr int elem_stride(int n) {
  return (n == 2 || (n * n * n) % 32 == 0) ? n * n * n : (n * n * n + 31) / 32 * 32;
}
// elements per group: as many as fit a run, rounded down to a multiple of the order's
// elements-per-workgroup shapes (16 and 32 at n = 2, 8 and 16 at n = 3, 4 and 8 at n = 4, 2 and 4
// at n = 5, 2 and 1 at n = 6 and 7, 1 and 2 at n = 8)
constexpr int group_elems(int n) {
  return n == 2 ? 512 : n == 3 ? 128 : n == 4 ? 64 : n == 5 ? 32 : n == 6 ? 18 : n == 7 ? 10 : 8;
}
inline int64_t field_slots(int64_t E, int n) {
  const int64_t g = group_elems(n);
  return (E + g - 1) / g * kRunSlots;
}
// Slot of field f (0..6) of element e, node slot `node`, in a register of field unit fs with es
// slots per element and g elements per group.
__host__ __device__ inline int64_t reg_slot(int64_t fs, int es, int g, int f, int64_t e, int node) {
  const int64_t grp = e / g, eg = e - grp * g;
  return f < 4 ? grp * 4 * kRunSlots + (int64_t)f * kRunSlots + eg * es + node
               : 4 * fs + grp * 3 * kRunSlots + (int64_t)(f - 4) * kRunSlots + eg * es + node;
}

constexpr int kMatStride = 8;  // per-element material record (doubles)
enum MatSlot { MAT_C = 0, MAT_RHO0 = 1, MAT_BA = 2, MAT_SIGMA = 3, MAT_IRHO0 = 4, MAT_DIFF = 5 };

struct MeshArgs {
  int nE[3];
  int periodic[3];
  int64_t E;          // interior elements (< 2^31)
  int64_t fs;         // field unit of a register: field_slots(E, n) (layout: reg_slot)
  int es;             // node slots per element: elem_stride(n)
  int grp;            // elements per group: group_elems(n)
  int64_t stride[3];  // element index stride per axis: e = (i*nE1 + j)*nE2 + k
  Div32 div12;        // divide by nE1*nE2
  Div32 div2;         // divide by nE2
  const double* inv_w[3];  // 1/width per element index along each axis
  const double* mat;       // [E][kMatStride] = c0, rho0, B/A, sigma, 1/rho0, D, -, -
};

struct ExtArgs {
  const double* trace[6];  // halo face traces [faceElem][n*n][3] (rho', u_axis, Q_axis) or null
  int src_face;            // -1 none
  double amp, freq, phase, t0, tau;
  const double* lead;      // [faceElem][n*n] or null
  double t;                // stage time; with `clock`, the offset of the stage from clock[0]
  const double* clock;     // device copy of the step's start time, or null
};

template <typename CT>
struct OpArgs {
  const CT* A;   // n*n
  const CT* L0;  // n
  const CT* LN;  // n
};

struct RhsLaunch {
  int n;            // nodes per axis
  int mixed;        // 0: FP64 compute, 1: FP32 compute
  int epilogue;     // Epilogue
  unsigned mask;    // term mask
  int xcd_remap;    // workgroup order 0..kNumOrders-1 (autotuned)
  int variant;      // kernel variant 0..kNumVariants-1 (autotuned)
  int contraction;  // mixed path: 0 VALU, 1 MFMA FP32, 2 MFMA split-BF16
  int ldg;          // 1: add the LDG diffusion term from the Q fields of q_in (fields 4..6)
  MeshArgs mesh;
  ExtArgs ext;
  const double* opd[3];  // A, L0, LN in FP64
  const float* opf[3];   // A, L0, LN in FP32
  const double* q_in;    // state read (and neighbour traces)
  double* out;           // EPI_ACCUM / EPI_RK33 output register
  const double* old;     // EPI_RK33 old register (alpha != 0)
  double* q_out;         // EPI_LSRK new state
  void* du;              // EPI_LSRK register (double or float per `mixed`)
  double a, b, dt;       // LSRK: a, b, dt; RK33: alpha, beta, dt
};

hipError_t launch_rhs(const RhsLaunch& L, hipStream_t s);
// The workgroup order of L.xcd_remap for a stage launch of `groups` element groups.
inline GroupMap group_map_for(const RhsLaunch& L, int64_t groups, int epb) {
  return make_group_map(L.xcd_remap, (uint32_t)groups, L.mesh.nE, epb);
}
constexpr int kNumVariants = 8;
constexpr int kNumOrders = 4;

// Batched stages: many independent blocks (solves) of different sizes and orders. Each solve owns
// a region of 4 * fs slots in the state arenas; its stage launches are described by launch records
// (RhsLaunch with the solve's mesh, operators and field pointers), kBatchRecords per solve:
// record 2 * stage + from for LSRK54 stage 0..4 reading arena `from`, and 10 + from for stage 0
// with the first-stage halo traces. The solves' element groups are tasks that workgroups take.
constexpr int kBatchRecords = 12;
struct BatchTask {
  int solve;
  int e0;  // first element of the group
};
struct BatchLaunch {
  const RhsLaunch* records;  // [solve][kBatchRecords]
  const BatchTask* tasks;    // this order's groups
  int ntasks;
  int record;                // record index of this stage
  int mixed;
};
// Per order, autotuned: bit 0 picks the elements-per-workgroup shape (the stage kernel's EPB0 or
// EPB1), bit 1 one task per workgroup instead of the persistent loop, bit 2 late RK-register loads.
constexpr int kNumBatchVariants = 8;
int batch_epb(int variant, int n);
// One persistent launch over the tasks of the solves of order n-1 (B.tasks holds only those).
hipError_t launch_rhs_batch(const BatchLaunch& B, int n, int variant, int grid, hipStream_t s);
int batch_max_grid(int n, int variant, int mixed);  // resident workgroups on the device
// Every order in one persistent launch (batch schedule 4): B.tasks holds the element groups of all
// solves, each group batch_all_epb(n) elements of its solve's order n.
int batch_all_epb(int n);
hipError_t launch_rhs_batch_all(const BatchLaunch& B, int grid, hipStream_t s);
int batch_all_max_grid(int mixed);  // resident workgroups of that launch on the device

// Exterior trace of the source face at the stage time (x.t, or clock[0] + x.t) into `out`
// (halo trace + boundary source), [faceElem][n*n][3]; the RHS kernel then reads it as that face's
// halo.
hipError_t launch_source_trace(const ExtArgs& x, const MeshArgs& m, int n, const double* halo, double* out,
                               hipStream_t s);

// Values per stage of the source face's trace: [faceElem][n*n][3].
inline int64_t source_trace_count(int face, const MeshArgs& m, int n) {
  const int axis = 2 - face / 2;
  const int64_t nb = (axis == 0) ? m.nE[1] : m.nE[0];
  const int64_t nc = (axis == 2) ? m.nE[1] : m.nE[2];
  return nb * nc * n * n;
}
constexpr int kMaxStages = 5;
struct StageTimes {
  double v[kMaxStages];
};
// The source face's traces of every stage of one step in one launch: stage s at the step's start
// on the device clock (x.clock; absolute times without one) plus offs.v[s], into
// out + s * source_trace_count * 3. Stage 0 adds halo0 (the halo traces a call's first stage
// reads, or null); the later stages have no halo.
hipError_t launch_source_trace_step(const ExtArgs& x, const MeshArgs& m, int n, const double* halo0, double* out,
                                    const StageTimes& offs, int stages, hipStream_t s);

// Boundary sources of the batched path: one entry per solve with a source. A stage's launch
// computes every entry's exterior trace at the solve's stage time clocks[solve] +
// offsets[solve * 5 + stage] (the offsets C_s dt prepared on the host, as dg3d_step's).
struct BatchSource {
  ExtArgs x;           // src_face, amp .. tau, lead (t and clock unused)
  MeshArgs m;          // nE and mat of the solve
  int n;
  int solve;
  const double* halo;  // the face's first-stage halo trace, or null
  double* out;         // the solve's trace buffer for that face
};
hipError_t launch_batch_sources(const BatchSource* list, int count, int64_t max_items, const double* clocks,
                                const double* offsets, int stage, int first, hipStream_t s);
// clocks[i] += dts[i] for every solve
hipError_t launch_batch_clocks(double* clocks, const double* dts, int count, hipStream_t s);

// LDG operators (Element.cpp:623-758) for one axis, FP64.
struct LdgLaunch {
  int n;
  int which;  // 0: QFromDRho (writes Q_axis of out), 1: DRhoFromQ (adds to rho' of out)
  int axis;
  MeshArgs mesh;
  ExtArgs ext;
  const double* diff;  // per element D
  const double* opd[3];
  const double* in;    // register read (7 fields)
  double* out;         // register written
};
hipError_t launch_ldg(const LdgLaunch& L, hipStream_t s);
// Element::OperateQFromDRho for the three axes in one pass (Grid.cpp:655-667): Q_a of `in`
// (fields 4..6) from rho' of `in` and the exterior traces, in place. FP64.
hipError_t launch_ldg_q(const LdgLaunch& L, hipStream_t s);

// out = out*a + in*b over nfields fields of length fs.
hipError_t launch_axpby(double* out, double a, const double* in, double b, int64_t count, hipStream_t s);

// out[i] = pseudo-random normal numbers of magnitude amp/16..amp and either sign, i < count.
hipError_t launch_fill_pattern(double* out, double amp, int64_t count, hipStream_t s);

// Reductions: partial results per block into `partials` (size returned by reduce_blocks). The
// max |.| and L2 reductions cover every slot of the four state fields' region (4 fs slots, the
// gaps and the slots past the last element included): slot i belongs to field (i / P) mod 4.
int reduce_blocks(int64_t count);
hipError_t launch_maxabs(const double* reg, int64_t fs, double* partials, hipStream_t s);
hipError_t launch_l2diff(const double* a, const double* b, int64_t fs, double* partials, hipStream_t s);
hipError_t launch_energy(const double* reg, const MeshArgs& m, int n, const double* wq, double* partials,
                         hipStream_t s);
hipError_t launch_cfl(const MeshArgs& m, double dxi_half, double* partials, hipStream_t s);

// The allocation-free reductions' scratch: a file-scope device array of kScratchPartials
// per-block partials (8 per block at the most blocks reduce_blocks gives) and kScratchValues
// combined values, one copy per device. scratch_address resolves its device address for
// `device`, which must be the current device, once per device (later calls return the kept
// address). launch_reduce_combine combines `blocks` partials of a reduction kind (0 max |.| of
// 4 fields, 1 energy sums, 2 L2 sums of 4 fields) in block order into out[0..7].
constexpr int kScratchPartials = 1024 * 8;
constexpr int kScratchValues = 8;
hipError_t scratch_address(int device, double** out);
hipError_t launch_reduce_combine(const double* partials, int blocks, int kind, double* out, hipStream_t s);
// Probe values into row *step of `out` ([cap][count]); nothing once *step reaches cap. `nodes`
// holds host node indices (element * n^3 + node).
hipError_t launch_probe(const double* reg, const MeshArgs& m, int n, const int64_t* nodes, int count,
                        double* out, const int* step, int cap, hipStream_t s);
// Device clock of dg3d_step: clock[0] = t; the step counter of the probes is left alone.
hipError_t launch_clock_set(double* clock, double t, hipStream_t s);
// End of a step: clock[0] += dt, and the probe step counter advances while below cap.
hipError_t launch_step_end(double* clock, double dt, int* probe_step, int cap, hipStream_t s);

}  // namespace dg3d
