// HIP kernels for gfx950: the per-stage right-hand side of the first-order acoustic system
// on tensor-product LGL hexahedra, fused with the RK update.
//
// One workgroup owns EPB whole elements, one lane per node. A stage is:
//   1. each lane loads its node's state (rho', u, v, w) — element-major SoA, so a wave reads
//      64 consecutive doubles per field — and writes the fluxes P/rho0, m_x, m_y, m_z to LDS
//      (kernelFuncs.cpp:13-30);
//   2. lanes take the 6 n^2 face nodes of the workgroup's elements, read the neighbour's
//      trace straight from its state in HBM (the on-device replacement of the halo exchange,
//      Grid.cpp:487-565) or the block's exterior trace, and store the upwind flux pair
//      (Element.cpp:540-600) in LDS;
//   3. each lane contracts the three lines through its node with A = M^{-1} S, lifts the face
//      fluxes with M^{-1} e_0 / M^{-1} e_N (the algebraic form of Element.cpp:602), applies
//      the element sponge (Element.cpp:760-781) and the RK update in registers.
// The same kernel, with a term mask and an accumulate epilogue, implements Block::Operate*.
#include <mutex>
#include <type_traits>

#include "dg3d_device.hpp"

namespace dg3d {
namespace {

// Workgroup shape: EPB whole elements of NN^3 nodes, one lane per node, padded to whole waves.
template <int NN, int EPB_>
struct Geo {
  static constexpr int NP = NN * NN * NN;
  static constexpr int NF2 = NN * NN;
  static constexpr int EPB = EPB_;
  static constexpr int ES = elem_stride(NN);  // node slots per element in a field run
  static constexpr int NT = ((EPB * ES + 63) / 64) * 64;
  static constexpr int G = group_elems(NN);     // elements per group of the register layout
  static_assert(G % EPB_ == 0 && G * ES <= kRunSlots, "a workgroup's elements lie in one group");
};

// Register slots of the group-blocked layout (dg3d_device.hpp): field 0 of element e (< 2^31) at
// node slot `node`; field f is f * kRunSlots further on. The Q fields: field 4 + a of element e.
template <int NN>
__device__ __forceinline__ int64_t state_slot(uint32_t e, int node) {
  constexpr uint32_t G = group_elems(NN);
  const uint32_t g = e / G;
  return (int64_t)g * 4 * kRunSlots + (int64_t)(e - g * G) * elem_stride(NN) + node;
}
template <int NN>
__device__ __forceinline__ int64_t q_slot(int64_t fs, uint32_t e, int a, int node) {
  constexpr uint32_t G = group_elems(NN);
  const uint32_t g = e / G;
  return 4 * fs + (int64_t)g * 3 * kRunSlots + (int64_t)a * kRunSlots + (int64_t)(e - g * G) * elem_stride(NN) + node;
}
// Default shape per order (whole waves with < 10% idle lanes); used by the LDG kernel.
template <int NN>
struct Shape : Geo<NN, (NN == 2) ? 32 : (NN == 3) ? 8 : (NN == 4) ? 4 : (NN == 5) ? 2 : (NN == 6) ? 2 : (NN == 7) ? 2 : 1> {};

// Tunable variants of the stage kernel (chosen at run time by the autotuner of the C ABI):
// bit 0 selects the alternative elements-per-workgroup, bit 1 the occupancy target (minimum
// waves per SIMD the register allocation must allow), bit 2 late RK-register loads.
template <int NN>
struct Variants {
  static constexpr int EPB0 = Shape<NN>::EPB;
  static constexpr int EPB1 = (NN == 2) ? 16 : (NN == 3) ? 16 : (NN == 4) ? 8 : (NN == 5) ? 4 : (NN == 8) ? 2 : 1;
  static constexpr int MINW1 = 6;
};

// kernelFuncs.cpp:17-20: P/rho0 = (rho' + rho'^2 * (B/A) * 0.5 / rho0) * c^2 / rho0, written
// with the element constants k8 = (B/A) * 0.5 / rho0 and k9 = c^2 / rho0.
// kernelFuncs.cpp:26-30: rho0 u if B/A == 0 exactly, else u (rho' + rho0), written as
// u (rho' k11 + k10) with k10 = rho0 and k11 = 0 or 1 (the same value for finite rho').
// The stage evaluates both through these functions for its own nodes (phase 1) and for gathered
// neighbour traces (phase 2), so that a neighbour read from the workgroup's LDS and one gathered
// from HBM give the same bits.
template <typename CT>
__device__ __forceinline__ CT mom_k8(CT ba, CT ir0) { return ba * CT(0.5) * ir0; }
template <typename CT>
__device__ __forceinline__ CT mom_k9(CT c, CT ir0) { return c * c * ir0; }
template <typename CT>
__device__ __forceinline__ CT mass_k11(CT ba) { return (ba == CT(0)) ? CT(0) : CT(1); }
template <typename CT>
__device__ __forceinline__ CT momentum_k(CT rho, CT k8, CT k9) { return (rho + rho * rho * k8) * k9; }
template <typename CT>
__device__ __forceinline__ CT mass_k(CT u, CT rho, CT k10, CT k11) { return u * (rho * k11 + k10); }

__device__ __forceinline__ double source_value(const ExtArgs& x, double t) {
  const double a = (t - x.t0) / x.tau;
  return x.amp * sin(2.0 * M_PI * x.freq * t + x.phase) * exp(-a * a);
}

// Element coordinates (i, j, k) of element e (< 2^31) by magic-number division.
__device__ __forceinline__ void elem_coords(const MeshArgs& m, uint32_t e, int* c) {
  const uint32_t i = div32(e, m.div12);
  const uint32_t r = e - i * (uint32_t)(m.nE[1] * m.nE[2]);
  const uint32_t j = div32(r, m.div2);
  c[0] = (int)i;
  c[1] = (int)j;
  c[2] = (int)(r - j * (uint32_t)m.nE[2]);
}

// Neighbour element along `axis` on `side` given the element's coordinates, or -1 for the
// exterior (halo) of the block.
[[maybe_unused]] __device__ __forceinline__ int neighbour(const MeshArgs& m, int e, const int* ec, int axis, int side) {
  const int st = (int)m.stride[axis];
  const int pos = ec[axis];
  if (side == 0) {
    if (pos > 0) return e - st;
    if (m.periodic[axis]) return e + (m.nE[axis] - 1) * st;
  } else {
    if (pos < m.nE[axis] - 1) return e + st;
    if (m.periodic[axis]) return e - (m.nE[axis] - 1) * st;
  }
  return -1;
}

// Index of an element among the elements of a face normal to `axis` (the other two element
// coordinates in (i, j, k) order).
[[maybe_unused]] __device__ __forceinline__ int face_elem(const MeshArgs& m, const int* ec, int axis) {
  if (axis == 0) return ec[1] * m.nE[2] + ec[2];
  if (axis == 1) return ec[0] * m.nE[2] + ec[2];
  return ec[0] * m.nE[1] + ec[1];
}

// Node index of the face node (b, c) at position `ia` along `axis`.
template <int NN>
__device__ __forceinline__ int face_node(int axis, int ia, int b, int c) {
  if (axis == 0) return (ia * NN + b) * NN + c;
  if (axis == 1) return (b * NN + ia) * NN + c;
  return (b * NN + c) * NN + ia;
}

// Buffer-resource access. A stage workgroup's resources start at its first element's slot and
// span its elements' pieces of the register (base, size in bytes): the loads and stores of all
// fields then share one small 32-bit offset register, and an access beyond the range reads 0 /
// is dropped.
using Rsrc = __amdgpu_buffer_rsrc_t;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int kPolicyNT = 2;  // nontemporal
__device__ __forceinline__ Rsrc make_rsrc(const void* base, uint32_t bytes) {
  // dword 3 = 0x00020000: 32-bit data format, raw (stride 0) addressing
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
template <typename T, int POL = 0>
__device__ __forceinline__ T bload(Rsrc r, uint32_t i) {
  if constexpr (std::is_same<T, double>::value) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, i * 8u, 0, POL));
  } else {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, i * 4u, 0, POL));
  }
}
template <typename T, int POL = 0>
__device__ __forceinline__ void bstore(T v, Rsrc r, uint32_t i) {
  if constexpr (std::is_same<T, double>::value) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, i * 8u, 0, POL);
  } else {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, i * 4u, 0, POL);
  }
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// The four FP32 fields of one node of a node-major register at byte offset off, in one 16-byte
// access (the whole vector is reinterpreted; no swizzle is bit-cast)
template <int POL = 0>
__device__ __forceinline__ void bload4f(Rsrc r, uint32_t off, double out[4]) {
  const f32x4 a = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, POL));
#pragma unroll
  for (int k = 0; k < 4; ++k) out[k] = (double)a[k];
}
// Two FP64 fields of one node of a pair-major register (16 bytes)
template <int POL = 0>
__device__ __forceinline__ void bload2d(Rsrc r, uint32_t off, double& x0, double& x1) {
  const f64x2 a = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, POL));
  x0 = a[0];
  x1 = a[1];
}
template <int POL = 0>
__device__ __forceinline__ void bstore2d(double x0, double x1, Rsrc r, uint32_t off) {
  f64x2 a;
  a[0] = x0;
  a[1] = x1;
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, a), r, off, 0, POL);
}
template <int POL = 0>
__device__ __forceinline__ void bstore4f(const float v[4], Rsrc r, uint32_t off) {
  f32x4 a;
#pragma unroll
  for (int k = 0; k < 4; ++k) a[k] = v[k];
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, a), r, off, 0, POL);
}

template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void st_stream(T v, T* p) { __builtin_nontemporal_store(v, p); }

// Per (element, face) flux coefficients. With mom = P/rho0 and mass = m_axis of self (s) and
// neighbour (o), the upwind pair of Element.cpp:540-600 scaled by the lift factors of
// Element.cpp:602 (1/rho0 for the momentum equation, rho0 for the mass equation) is
//   F0 = K0 mom_o + K1 mom_s + K2 mass_o + K3 mass_s
//   F1 = K4 mass_o + K5 mass_s + K6 mom_o + K7 mom_s
// and the neighbour's kernel functions are mom_o = (rho + rho^2 K8) K9, mass_o = u (rho K11 + K10).
// K12 = rho0_s / rho0_o converts a pressure shift expressed as P/rho0 of self into the
// neighbour's P/rho0 (the FP32 stage's shifted lines, below); the FP64 stage, which does not
// shift its lines, keeps 12 coefficients per face in LDS.
constexpr int kNK = 13;
template <typename CT>
constexpr int nk_of() {
  return std::is_same<CT, double>::value ? 12 : 13;
}
template <typename CT>
__device__ __forceinline__ void face_coeffs(int side, CT c1, CT r1, CT ir1, CT c2, CT r2, CT ba2, CT ir2, CT* K) {
  const CT z1 = r1 * c1, z2 = r2 * c2;
  const CT iz = CT(1) / (z1 + z2);
  const CT sg = side ? iz : -iz;  // the previous face carries the leading minus sign
  const CT s0 = ir1, s1 = r1;
  K[0] = sg * z1 * r2 * s0;
  K[1] = sg * z2 * r1 * s0;
  K[2] = -iz * z1 * c2 * s0;
  K[3] = iz * z2 * c1 * s0;
  K[4] = sg * c2 * s1;
  K[5] = sg * c1 * s1;
  K[6] = -iz * r2 * s1;
  K[7] = iz * r1 * s1;
  K[8] = mom_k8(ba2, ir2);
  K[9] = mom_k9(c2, ir2);
  K[10] = r2;
  K[11] = mass_k11(ba2);
}

// A value known to be equal on every lane, as such to the compiler (bit-exact: the two halves
// of the double are moved separately)
[[maybe_unused]] __device__ __forceinline__ double wave_uniform(double v) {
  const u32x2 u = __builtin_bit_cast(u32x2, v);
  u32x2 r;
  r[0] = __builtin_amdgcn_readfirstlane(u[0]);
  r[1] = __builtin_amdgcn_readfirstlane(u[1]);
  return __builtin_bit_cast(double, r);
}

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef short bf16x4 __attribute__((ext_vector_type(4)));

// x = hi + mid + lo as three bf16 pieces (hardware conversion, round to nearest even); both
// residuals are exact in FP32 and the representation error is below 2^-26 |x|.
[[maybe_unused]] __device__ __forceinline__ void split3(float x, short& h, short& m, short& l) {
  const __bf16 a = (__bf16)x;
  const float r1 = x - (float)a;
  const __bf16 b = (__bf16)r1;
  const float r2 = r1 - (float)b;
  h = __builtin_bit_cast(short, a);
  m = __builtin_bit_cast(short, b);
  l = __builtin_bit_cast(short, (__bf16)r2);
}

typedef float float2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));

// split3 of two values at once: each conversion fills both halves of one register, in the
// operand's order (element 0 in the low half), with the same bits as split3 per element.
[[maybe_unused]] __device__ __forceinline__ void split3x2(float2v x, bf16x2v& h, bf16x2v& m, bf16x2v& l) {
  h = __builtin_convertvector(x, bf16x2v);
  const float2v r1 = x - __builtin_convertvector(h, float2v);
  m = __builtin_convertvector(r1, bf16x2v);
  const float2v r2 = r1 - __builtin_convertvector(m, float2v);
  l = __builtin_convertvector(r2, bf16x2v);
}

[[maybe_unused]] __device__ __forceinline__ bf16x4 join4(bf16x2v a, bf16x2v b) {
  return __builtin_bit_cast(bf16x4, __builtin_shufflevector(a, b, 0, 1, 2, 3));
}

// Line of a contraction column: a column is (element, axis, field, face node) with the field
// P/rho0 (f=0) or m_axis (f=1). fo/oo: LDS offset of the line's first node in sF and in sOut,
// ls: node stride along the line.
template <int NN>
__device__ __forceinline__ void column_line(int col, int& fo, int& oo, int& ls) {
  constexpr int NP = NN * NN * NN, NF2 = NN * NN;
  const int le = col / (6 * NF2);
  const int rr = col - le * 6 * NF2;
  const int af = rr / NF2;
  const int fn = rr - af * NF2;
  const int a = af >> 1, f = af & 1;
  const int b = fn / NN, c = fn - b * NN;
  const int base = (a == 0) ? b * NN + c : (a == 1 ? b * NF2 + c : b * NF2 + c * NN);
  ls = (a == 0) ? NF2 : (a == 1 ? NN : 1);
  fo = le * 4 * NP + (f == 0 ? 0 : (1 + a)) * NP + base;
  oo = (le * 6 + af) * NP + base;
}

// One contraction column's line, mapped once per tile: loads and stores of its node m, which
// read as 0 and are dropped past the line's end (m >= NN) and past the last column.
// A P/rho0 column (field 0) reads its line shifted by the value at the line's middle node (see
// kShiftPressureLines in stage_body); mass columns read the line as stored.
template <int NN>
struct TileLine {
  int fo, oo, ls;
  bool ok;
  float sh;
  __device__ __forceinline__ TileLine(int col, int C, const float* sF) {
    ok = col < C;
    column_line<NN>(ok ? col : 0, fo, oo, ls);  // branch-free: column 0 stands in past the end
    const int rr = (ok ? col : 0) % (6 * NN * NN);
    const bool pressure = ((rr / (NN * NN)) & 1) == 0;
    const float v0 = sF[fo + (NN / 2) * ls];
    sh = pressure ? v0 : 0.f;
  }
  __device__ __forceinline__ float load(const float* sF, int m) const {
    const bool in = ok && m < NN;
    const float v = sF[in ? fo + m * ls : 0];
    return in ? v - sh : 0.f;
  }
  __device__ __forceinline__ void store(float* sOut, int m, float v) const {
    if (ok && m < NN) sOut[oo + m * ls] = v;
  }
};

// Orders whose split-BF16 contraction takes one line per lane (contract_lines_bf16); at 8 nodes per
// line the 16x16x16 tiles of contract_mfma_bf16, whose two 8-row blocks fill each tile, ran 1.7 %
// faster (64^3 N=7), while at 5 and 6 nodes per line the lines ran 2.8 and 2.4 % faster than
// those tiles (64^3 N=4, 48^3 N=5).
template <int NN>
constexpr bool kLinesPerLane = NN < 8;

// Shape of the split-BF16 contraction of a workgroup (contract_lines_bf16): C lines, in groups of
// 64 (one line per lane of a wave), G groups taken round-robin by the W waves, at most GMAX per
// wave; RB row blocks and KB K blocks of 4 cover a line of NN nodes.
template <int NN, int EPB, int NT>
struct LineShape {
  static constexpr int C = EPB * 6 * NN * NN;
  static constexpr int G = (C + 63) / 64;
  static constexpr int W = NT / 64;
  static constexpr int GMAX = (G + W - 1) / W;
  static constexpr int RB = NN > 4 ? 2 : 1;
  static constexpr int KB = RB;
};

// Line descriptors of the split-BF16 contraction, one 32-bit word per line of the workgroup, padded
// to whole groups of 64 (a line is one (element, axis, field, face node) with the field P/rho0 or
// m_axis; the mapping of column_line, found at compile time): bits 0-11 the line's first node in
// sF, bits 12-24 in sOut, bits 25-26 the axis (node stride NN^2, NN or 1), bit 27 a P/rho0 line
// (entered shifted by its middle value), bit 31 a line that exists (0: padding).
template <int NN, int EPB>
struct LineTable {
  static constexpr int NP = NN * NN * NN, NF2 = NN * NN, C = EPB * 6 * NF2;
  static constexpr int N = (C + 63) / 64 * 64;
  static_assert(EPB * 4 * NP <= 4096 && EPB * 6 * NP <= 8192, "descriptor fields");
  uint32_t v[N];
  constexpr LineTable() : v() {
    for (int col = 0; col < C; ++col) {
      const int le = col / (6 * NF2), rr = col % (6 * NF2), af = rr / NF2, fn = rr % NF2;
      const int a = af >> 1, f = af & 1, b = fn / NN, c = fn % NN;
      const int base = (a == 0) ? b * NN + c : (a == 1 ? b * NF2 + c : b * NF2 + c * NN);
      const int fo = le * 4 * NP + (f == 0 ? 0 : 1 + a) * NP + base;
      const int oo = (le * 6 + af) * NP + base;
      v[col] = (uint32_t)fo | ((uint32_t)oo << 12) | ((uint32_t)a << 25) | ((uint32_t)(f == 0) << 27) | (1u << 31);
    }
  }
};
template <int NN, int EPB>
__device__ inline constexpr LineTable<NN, EPB> kLineTable{};

// The descriptors of the lines a lane takes in its wave's groups (0 past the last group), read with
// the stage's other global loads
template <int NN, int EPB, int NT>
__device__ __forceinline__ void line_descriptors(int tid, uint32_t* d) {
  using M = LineShape<NN, EPB, NT>;
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int k = 0; k < M::GMAX; ++k) {
    const int g = wave + k * M::W;
    d[k] = g < M::G ? kLineTable<NN, EPB>.v[g * 64 + lane] : 0u;
  }
}

[[maybe_unused]] __device__ __forceinline__ floatx4 mfma4(bf16x4 a, bf16x4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a, b, c, 0, 0, 0);
}

// The split-BF16 derivative contraction on the matrix cores, one line per lane: out = A line for
// every line of the workgroup (P/rho0 lines entered shifted by their middle value, as in the VALU
// contraction of the FP32 stage). v_mfma_f32_4x4x4bf16_1k computes 16 independent 4x4x4 products;
// its lane layout, checked on the device before use: lane 4b + i holds row i of block b's A operand
// (its 4 K values), lane 4b + j holds column j of block b's B operand (4 K values), and register r
// of lane 4b + j receives row r, column j of block b's result. Every block takes the same 4x4 piece
// of A, so each lane's B column is a K block of its own line and its results are rows of its own
// line: one piece for n <= 4, four for n > 4 (row blocks R and K blocks of 4, zero past n). A
// (in sAp, split once per workgroup) and the lines are split into three bf16 pieces each,
// hi + mid + lo, exactly; per row block the five products of combined order 1 and 2 are accumulated
// in FP32, smallest first, through the matrix core's accumulator, and the hi x hi products of each
// K block come from a zero accumulator and are added with FP32 adds (fed the small terms as its
// accumulator input, the matrix core biased the sum).
template <int NN, int EPB, int NT>
__device__ __forceinline__ void contract_lines_bf16(const float* sF, const bf16x4* sAp, float* sOut, int tid,
                                                    const uint32_t* ldesc) {
  using M = LineShape<NN, EPB, NT>;
  constexpr int RB = M::RB, KB = M::KB;
  const int wave = tid >> 6, ia = tid & 3;
#pragma unroll
  for (int k = 0; k < M::GMAX; ++k) {
    if (wave + k * M::W >= M::G) break;  // wave-uniform
    const uint32_t d = ldesc[k];
    const bool ok = (d >> 31) != 0u;
    const int fo = ok ? (int)(d & 0xFFFu) : 0, oo = (int)((d >> 12) & 0x1FFFu);
    const uint32_t ax = (d >> 25) & 3u;
    const int ls = ax == 0u ? NN * NN : (ax == 1u ? NN : 1);
    float x[4 * KB];
#pragma unroll
    for (int m = 0; m < 4 * KB; ++m) x[m] = m < NN ? sF[fo + m * ls] : 0.f;
    const float sh = ((d >> 27) & 1u) ? x[NN / 2] : 0.f;
#pragma unroll
    for (int m = 0; m < NN; ++m) x[m] = ok ? x[m] - sh : 0.f;
    bf16x4 bh[KB], bm[KB], bl[KB];
#pragma unroll
    for (int K = 0; K < KB; ++K) {
      bf16x2v h0, m0, l0, h1, m1, l1;
      split3x2(float2v{x[4 * K], x[4 * K + 1]}, h0, m0, l0);
      split3x2(float2v{x[4 * K + 2], x[4 * K + 3]}, h1, m1, l1);
      bh[K] = join4(h0, h1);
      bm[K] = join4(m0, m1);
      bl[K] = join4(l0, l1);
    }
#pragma unroll
    for (int R = 0; R < RB; ++R) {
      const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
      floatx4 acc = zero, hh = zero;
#pragma unroll
      for (int K = 0; K < KB; ++K) {
        const bf16x4* ap = sAp + (R * KB + K) * 12 + ia;  // pieces hi, mid, lo 4 apart
        const bf16x4 ah = ap[0], am = ap[4], al = ap[8];
        acc = mfma4(ah, bl[K], acc);
        acc = mfma4(am, bm[K], acc);
        acc = mfma4(al, bh[K], acc);
        acc = mfma4(ah, bm[K], acc);
        acc = mfma4(am, bh[K], acc);
        const floatx4 h = mfma4(ah, bh[K], zero);
        if (K == 0) hh = h;
        else hh = hh + h;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (ok && 4 * R + r < NN) sOut[oo + (4 * R + r) * ls] = hh[r] + acc[r];
    }
  }
}

// The FP32 derivative contraction on the matrix cores, one line per lane: out = A line for every
// line of the workgroup (P/rho0 lines entered shifted by their middle value, as in the VALU
// contraction of the FP32 stage). v_mfma_f32_4x4x1_16b_f32 computes 16 independent 4x1 by 1x4
// products; its lane layout, checked on the device before use: lane 4b + i holds row i of block
// b's A operand, lane 4b + j column j of its B operand, and register r of lane 4b + j receives
// row r, column j of block b's result. Every block takes the same column of A, so each lane's B
// value is node k of its own line and its results are rows of its own line: NN products per row
// block of 4, accumulated in FP32 in node order, as the VALU contraction sums them.
template <int NN, int EPB, int NT>
__device__ __forceinline__ void contract_lines_f32(const float* sF, const float* sA, float* sOut, int tid,
                                                   const uint32_t* ldesc) {
  using M = LineShape<NN, EPB, NT>;
  constexpr int RB = (NN + 3) / 4;
  const int wave = tid >> 6, ia = tid & 3;
  // the operator entries this lane feeds: rows 4R + ia (zero past the line), every column
  float a[RB][NN];
#pragma unroll
  for (int R = 0; R < RB; ++R)
#pragma unroll
    for (int k = 0; k < NN; ++k) a[R][k] = (4 * R + ia < NN) ? sA[(4 * R + ia) * NN + k] : 0.f;
#pragma unroll
  for (int g = 0; g < M::GMAX; ++g) {
    if (wave + g * M::W >= M::G) break;  // wave-uniform
    const uint32_t d = ldesc[g];
    const bool ok = (d >> 31) != 0u;
    const int fo = ok ? (int)(d & 0xFFFu) : 0, oo = (int)((d >> 12) & 0x1FFFu);
    const uint32_t ax = (d >> 25) & 3u;
    const int ls = ax == 0u ? NN * NN : (ax == 1u ? NN : 1);
    float x[NN];
#pragma unroll
    for (int m = 0; m < NN; ++m) x[m] = sF[fo + m * ls];
    const float sh = ((d >> 27) & 1u) ? x[NN / 2] : 0.f;
#pragma unroll
    for (int m = 0; m < NN; ++m) x[m] = ok ? x[m] - sh : 0.f;
#pragma unroll
    for (int R = 0; R < RB; ++R) {
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < NN; ++k) acc = __builtin_amdgcn_mfma_f32_4x4x1f32(a[R][k], x[k], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (ok && 4 * R + r < NN) sOut[oo + (4 * R + r) * ls] = acc[r];
    }
  }
}

// The split-BF16 derivative contractions of a workgroup as one GEMM on the matrix cores, used at 8
// nodes per line (other orders take contract_lines_bf16):
// out[i][col] = sum_m A[i][m] line_col[m], one column per (element, axis, field, face node);
// P/rho0 lines enter shifted by their middle value, as in the VALU contraction of the FP32 stage.
// Two copies of A sit block-diagonally in the 16x16 operand of v_mfma_f32_16x16x16bf16_1k, block
// b on rows and K indices [8b, 8b + 8), so the 4 rows a lane holds (4g..4g+3, in both the K
// operand and the result) lie in one block. A tile covers 16 columns per block (columns
// col0 + 16 b), and waves take tiles round-robin. A and the lines are split into three bf16
// pieces each; the five products of combined order 1 and 2 are accumulated in FP32, smallest
// first, and the hi x hi products from zero, then added to them.
// Lane layout (lane l, g = l/16, j = l%16): A[j][k = 4g + t], B[k = 4g + t][j], D[4g + r][j].
template <int NN, int EPB, int NT>
__device__ __forceinline__ void contract_mfma_bf16(const float* sF, const float* sAf, float* sOut, int tid) {
  constexpr int BS = 8;
  constexpr int P = 16 / BS;
  constexpr int C = EPB * 6 * NN * NN;
  constexpr int TC = 16 * P;
  constexpr int T = (C + TC - 1) / TC;
  constexpr int W = NT / 64;
  static_assert(NT % 64 == 0, "whole waves");
  static_assert(NN > 4 && NN <= 8, "one block per 8 rows");
  const int lane = tid & 63, wave = tid >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int bj = j / BS, mj = j - bj * BS;    // block and line node of operand row j
  const int bd = (4 * g) / BS;                // block of the lane's result rows 4g..4g+3
  const int md = 4 * g - bd * BS;             // their first line node
  {
    // split-BF16 on v_mfma_f32_16x16x16bf16_1k (used at 8 nodes per line, where the two 8-row
    // blocks fill the tile): A and the lines are split into three bf16 pieces each; the five
    // products of combined order 1 and 2 are accumulated in FP32, smallest first, and the hi x hi
    // products from a zero accumulator, then added to them
    bf16x4 ah, am, al;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int mk = md + q;
      short h, m, l;
      split3((bj == bd && mj < NN && mk < NN) ? sAf[mj * NN + mk] : 0.f, h, m, l);
      ah[q] = h;
      am[q] = m;
      al[q] = l;
    }
    for (int t = wave; t < T; t += W) {
      const TileLine<NN> ln(t * TC + 16 * bd + j, C, sF);
      bf16x2v h0, m0, l0, h1, m1, l1;
      split3x2(float2v{ln.load(sF, md), ln.load(sF, md + 1)}, h0, m0, l0);
      split3x2(float2v{ln.load(sF, md + 2), ln.load(sF, md + 3)}, h1, m1, l1);
      const bf16x4 bh = join4(h0, h1), bm = join4(m0, m1), bl = join4(l0, l1);
      const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
      floatx4 acc = zero;
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, bm, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bm, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, bh, acc, 0, 0, 0);
      // the dominant hi x hi group from a zero accumulator, then one round-to-nearest add: fed
      // the small terms as its accumulator input, the matrix core biased the sum (by about half
      // its mean rounding error, always the same way), which grew linearly over a run
      const floatx4 hh = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, zero, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) ln.store(sOut, md + r, hh[r] + acc[r]);
    }
  }
}

template <typename T>
__device__ __forceinline__ T sel3(int a, T x0, T x1, T x2) {
  return a == 0 ? x0 : (a == 1 ? x1 : x2);
}

// Across face `dir` of element he with coordinates ec: the neighbour element (periodic wrap
// included), or -1 - (index of he among the block's elements on that face) for the exterior.
__device__ __forceinline__ int face_neighbour(const MeshArgs& m, int he, const int* ec, int dir) {
  const int axis = 2 - (dir >> 1);
  const int pos = sel3(axis, ec[0], ec[1], ec[2]);
  const int ne = sel3(axis, m.nE[0], m.nE[1], m.nE[2]);
  const int st = (int)sel3(axis, m.stride[0], m.stride[1], m.stride[2]);
  const int per = sel3(axis, m.periodic[0], m.periodic[1], m.periodic[2]);
  const int fe = sel3(axis, ec[1] * m.nE[2] + ec[2], ec[0] * m.nE[2] + ec[2], ec[0] * m.nE[1] + ec[1]);
  if (dir & 1) return (pos < ne - 1) ? he + st : (per ? he - (ne - 1) * st : -1 - fe);
  return (pos > 0) ? he - st : (per ? he + (ne - 1) * st : -1 - fe);
}

// LDS of one stage workgroup: EPB elements of order NN-1.
template <int NN, int EPB, typename CT, int CONTRACT, bool LDG = false>
struct StageSmem {
  static constexpr int NP = NN * NN * NN, NF2 = NN * NN;
  CT sF[EPB * 4 * NP];                    // P/rho0, m_x, m_y, m_z per node
  CT sFl[EPB * 6 * NF2 * 2];              // scaled flux pair per face node
  CT sK[EPB * 6 * nk_of<CT>()];           // flux coefficients per face
  CT sIw[EPB * 3];                        // 1/width per element and axis
  float sOut[CONTRACT ? EPB * 6 * NP : 1];  // matrix-core contraction results
  // split-BF16 contraction: the pieces of A, [row block][K block][hi, mid, lo][row in block]
  bf16x4 sAp[CONTRACT == 2 && NN < 8 ? (NN > 4 ? 4 : 1) * 12 : 1];
  CT sQ[LDG ? EPB * 3 * NP : 1];          // LDG: Q_x, Q_y, Q_z per node
  CT sFq[LDG ? EPB * 6 * NF2 : 1];        // LDG: central trace of Q_axis per face node
  CT sA[NN * NN];
  CT sL0[NN];
  CT sLN[NN];
};

constexpr unsigned kItemExterior = 1u << 31;  // face item flag: zero exterior state
constexpr unsigned kItemLocal = 1u << 30;     // face item flag: the neighbour is in the workgroup

// One workgroup's share of a stage: elements e0 .. e0+EPB-1 of the block described by m (field
// arrays in L; traces: the 6 exterior face traces in memory, or null for none). NT is the
// workgroup size (>= EPB * ES); every lane must call it, since it holds barriers. LATE issues the
// RK-register loads (du / old, needed only by the epilogue) after the first barrier instead of
// with the rest: fewer registers live at the peak, hence more waves, against later loads. LDG
// adds Element::OperateDRhoFromQ for the three axes (Element.cpp:691-758) from the Q fields of
// q_in, which k_ldg_q filled for this stage: rho' += -D (1/w_a) (A Q_a + L0 Fq_prev + LN Fq_next),
// with the central traces Fq_prev = (Q_o + Q_s)/2 and Fq_next = -(Q_s + Q_o)/2.
template <int NN, int EPB, int NT, typename CT, int EPI, typename DUT, int CONTRACT, bool LATE = false,
          bool LDG = false>
__device__ __forceinline__ void stage_body(const RhsLaunch& L, const MeshArgs& m, const double* const* traces,
                                           const int e0, const int tid, StageSmem<NN, EPB, CT, CONTRACT, LDG>& sm) {
  using S = Geo<NN, EPB>;
  constexpr int NP = S::NP, NF2 = S::NF2;
  static_assert(NT >= EPB * S::ES, "workgroup too small for the element group");
  static_assert(EPB <= 32 && NN * NN * NN <= 512, "face item encoding: local element in 5 bits, node in 9");
  // FP32 stage: every line's P/rho0 values enter the contraction and the flux pair shifted by
  // the value c at the line's middle node (NN/2 along the axis), and the neighbour's by
  // c rho0_s/rho0_o, i.e. the same pressure. A constant pressure drops out of the line's equation exactly: the
  // upwind flux's pressure weights sum to one (z1 + z2 over z1 + z2), so it moves both face
  // fluxes by the constant, which cancels the constant's volume term, and the jump terms see
  // no change. In FP32 the shift removes the cancellation between a line's volume term and its
  // lifted fluxes: on a line along which the pressure is (nearly) uniform, e.g. across a plane
  // wave, the unshifted sum leaves rounding of the order of P/rho0 times the operator's norm
  // in the transverse velocity, where the shifted one leaves (nearly) none. FP64 is unshifted.
  constexpr bool kShiftPressureLines = !std::is_same<CT, double>::value;
  constexpr int NITEMS = EPB * 6 * NF2;            // face work items of the workgroup
  constexpr int R = (NITEMS + NT - 1) / NT;        // face items per lane
  CT* const sF = sm.sF;
  CT* const sFl = sm.sFl;
  CT* const sK = sm.sK;
  CT* const sIw = sm.sIw;
  float* const sOut = sm.sOut;
  CT* const sA = sm.sA;
  CT* const sL0 = sm.sL0;
  CT* const sLN = sm.sLN;
  const unsigned mask = L.mask;
  const int E = (int)m.E;
  const int64_t fs = m.fs;
  // the epilogue's scalars, read before the first global store: a launch record in global
  // memory (the batched kernels) would otherwise be re-read, with vector loads, after every
  // store that might alias it
  const double ep_a = L.a, ep_b = L.b, ep_dt = L.dt;
  double* const ep_out = L.out;
  double* const ep_qout = L.q_out;

  // ---- phase 0: issue every global load of the workgroup -------------------------------
  // The node's state and RK register, the neighbour traces of the lane's face items, the
  // material on both sides of each face, the line operators and 1/width. Neighbours are found
  // by index arithmetic, so no load waits on another and all are in flight together.
  // lane -> (element, slot); slots n^3..ES-1 are padding: loaded and stored (as zeros) so the
  // workgroup writes whole lines, but they take no part in the operator
  constexpr int ES = S::ES;
  const int le = tid / ES;
  const int node = tid - le * ES;
  const int e = e0 + le;
  const bool io = (le < EPB) && (e < E);
  // When the element groups lie in z columns (EPB divides nE_z, as groups start at multiples of
  // EPB), element e0 + l has the coordinates of e0 with k + l: they are found once, on the
  // scalar unit, instead of by two divisions per face item and helper lane
  const bool col_groups = (m.nE[2] % EPB) == 0;
  int gc[3];
  elem_coords(m, (uint32_t)e0, gc);
  auto coords_of = [&](int l, int* c) {
    if (col_groups) {
      c[0] = gc[0];
      c[1] = gc[1];
      c[2] = gc[2] + l;
    } else {
      elem_coords(m, (uint32_t)(e0 + l), c);
    }
  };
  const bool valid = io && node < NP;
  // lanes past the workgroup's elements load a slot of its last element, a line that its own
  // lanes read anyway, rather than one fixed element that every workgroup would read
  const int ge = io ? e : min(e0 + EPB - 1, E - 1);
  // Every register access of the workgroup goes through buffer resources that start at its first
  // element's slot in field 0 and span its elements' pieces of the 4 runs of their group (the
  // workgroup lies in one group), so the offsets are small whatever the register's size. gi: the
  // lane's slot from there; field f is P = kRunSlots slots further on.
  constexpr int P = kRunSlots;
  const int ne = (E - e0 < EPB) ? E - e0 : EPB;
  const uint32_t gi = (uint32_t)((ge - e0) * ES + node);
  const uint32_t wb64 = (uint32_t)((3 * P + ne * ES) * 8);
  const size_t wofs = (size_t)state_slot<NN>((uint32_t)e0, 0);
  const Rsrc rq = make_rsrc(L.q_in + wofs, wb64);
  double q[4], pre[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int f = 0; f < 4; ++f) q[f] = bload<double>(rq, gi + f * P);
  // The LSRK register is never read across a face, so its layout within a group's block of 4 P
  // values is free. A lane moves its node's share in 16-byte accesses: the FP64 register
  // pair-major (the group's (rho', u) pairs in the block's first half, its (v, w) pairs in the
  // second: two accesses), the FP32 register node-major (four fields in one access). Laid out
  // like the state instead, it ran 1.7-4 % slower in FP64 (64^3 N=4 and N=7, 48^3 N=6) and
  // 0.8-2.3 % slower in FP32.
  constexpr bool kDuNodeMajor = std::is_same<DUT, float>::value;
  constexpr bool kDuPairMajor = std::is_same<DUT, double>::value;
  Rsrc rdu;
  uint32_t du_off = 0;
  if constexpr (EPI == EPI_LSRK) {
    const DUT* d0 = static_cast<const DUT*>(L.du) + (size_t)(e0 / S::G) * 4 * P;  // the group's block
    const int eg = e0 % S::G;
    if constexpr (kDuNodeMajor) {
      rdu = make_rsrc(d0 + (size_t)eg * ES * 4, (uint32_t)(ne * ES * 16));
      du_off = gi * 16u;
    } else if constexpr (kDuPairMajor) {
      rdu = make_rsrc(d0 + (size_t)eg * ES * 2, (uint32_t)((2 * P + ne * ES * 2) * 8));
      du_off = gi * 16u;  // pair 1: + 2 P values
    } else {
      rdu = make_rsrc(d0 + (size_t)eg * ES, (uint32_t)((3 * P + ne * ES) * sizeof(DUT)));
    }
  }
  auto load_pre = [&]() {
    if constexpr (EPI == EPI_LSRK) {
      if (L.a != 0.0) {
        if constexpr (kDuNodeMajor) {
          bload4f<kPolicyNT>(rdu, du_off, pre);
        } else if constexpr (kDuPairMajor) {
          bload2d<kPolicyNT>(rdu, du_off, pre[0], pre[1]);
          bload2d<kPolicyNT>(rdu, du_off + (uint32_t)P * 16u, pre[2], pre[3]);
        } else {
#pragma unroll
          for (int f = 0; f < 4; ++f) pre[f] = (double)bload<DUT, kPolicyNT>(rdu, gi + f * P);
        }
      }
    }
    if constexpr (EPI == EPI_RK33) {
      if (L.a != 0.0) {
        const Rsrc ro = make_rsrc(L.old + wofs, wb64);
#pragma unroll
        for (int f = 0; f < 4; ++f) pre[f] = bload<double, kPolicyNT>(ro, gi + f * P);
      }
    }
  };
  if constexpr (!LATE) load_pre();
  const Rsrc rm = make_rsrc(m.mat, (uint32_t)(m.E * kMatStride * 8));
  // the lane's element material. When the element slots fill whole waves (ES a multiple of 64),
  // or the workgroup has one element, the element is the same on every lane of a wave that
  // holds any of its slots (lanes past the workgroup's elements return before the epilogue),
  // and its record is read once per wave through the scalar cache, which keeps these loads off
  // the vector memory pipeline that the state and the neighbour gathers use. The FP32 one-element
  // shape of 7 nodes per line keeps the vector loads: it ran 1.5 % slower with the scalar ones
  // (FP64: 4 % faster)
  double mat[6];
  if constexpr (ES % 64 == 0 || (EPB == 1 && (NN != 7 || std::is_same<CT, double>::value))) {
    typedef const __attribute__((address_space(4))) double* ConstPtr;  // read-only in the kernel
    const int lw = __builtin_amdgcn_readfirstlane(le);
    const ConstPtr pm = (ConstPtr)(m.mat + (size_t)((lw < EPB && e0 + lw < E) ? e0 + lw : 0) * kMatStride);
#pragma unroll
    for (int k = 0; k < 6; ++k) mat[k] = (k < 5 || LDG) ? wave_uniform(pm[k]) : 0.0;
  } else {
    const uint32_t mi = (uint32_t)ge * kMatStride;  // padding slots read their element's record
#pragma unroll
    for (int k = 0; k < 5; ++k) mat[k] = bload<double>(rm, mi + k);
    mat[MAT_DIFF] = LDG ? bload<double>(rm, mi + MAT_DIFF) : 0.0;
  }
  const CT c = (CT)mat[MAT_C], r0 = (CT)mat[MAT_RHO0];
  const CT ba = (CT)mat[MAT_BA], sg = (CT)mat[MAT_SIGMA];
  const CT ir0 = (CT)mat[MAT_IRHO0];
  // the split-BF16 contraction's line descriptors (a small table every workgroup reads)
  constexpr bool kLines = CONTRACT == 2 && kLinesPerLane<NN>;
  constexpr bool kLineDesc = kLines || CONTRACT == 1;
  uint32_t ldesc[kLineDesc ? LineShape<NN, EPB, NT>::GMAX : 1] = {};
  if constexpr (kLineDesc) line_descriptors<NN, EPB, NT>(tid, ldesc);
  double qq[3] = {0.0, 0.0, 0.0};
  CT dcoef = CT(0);
  if constexpr (LDG) {
    // the Q fields: blocks of 3 runs from slot 4 fs
    const Rsrc rqq = make_rsrc(L.q_in + q_slot<NN>(fs, (uint32_t)e0, 0, 0), (uint32_t)((2 * P + ne * ES) * 8));
#pragma unroll
    for (int a = 0; a < 3; ++a) qq[a] = bload<double>(rqq, gi + a * P);
    dcoef = (CT)mat[MAT_DIFF];
  }

  // helper lanes, one per (element, face): material of both sides and 1/width. These and the
  // line operators are issued before the neighbour gathers: phase 1 needs them, and the wait
  // for them before the first barrier then leaves the gathers in flight until phase 2
  // The helper lanes sit in the workgroup's last wave and the operator lanes in its second, when
  // they fit there, so that no wave issues both sets of loads on top of its own
  constexpr int HB = ((NT / 64 - 1) * 64 + EPB * 6 <= NT) ? (NT / 64 - 1) * 64 : 0;
  constexpr int OB = (NT >= 128 && 64 + NN * NN <= NT) ? 64 : 0;
  const int ht = tid - HB, ot = tid - OB;
  const bool helper = ht >= 0 && ht < EPB * 6;
  const int hle = ht / 6, hdir = ht - (ht / 6) * 6;
  const int he = e0 + hle;
  double ms_c = 1.0, ms_r = 1.0, ms_ir = 1.0, mo_c = 1.0, mo_r = 1.0, mo_ba = 0.0, mo_ir = 1.0, hiw = 0.0;
  if (helper && he < E) {
    int ec[3];
    coords_of(hle, ec);
    const int nb = face_neighbour(m, he, ec, hdir);
    const uint32_t ms = (uint32_t)he * kMatStride;
    // the halo element has the adjacent element's material (Block.cpp:167-248)
    const uint32_t mo = (uint32_t)(nb >= 0 ? nb : he) * kMatStride;
    ms_c = bload<double>(rm, ms + MAT_C);
    ms_r = bload<double>(rm, ms + MAT_RHO0);
    ms_ir = bload<double>(rm, ms + MAT_IRHO0);
    mo_c = bload<double>(rm, mo + MAT_C);
    mo_r = bload<double>(rm, mo + MAT_RHO0);
    mo_ba = bload<double>(rm, mo + MAT_BA);
    mo_ir = bload<double>(rm, mo + MAT_IRHO0);
    const int axis = 2 - (hdir >> 1);
    if ((hdir & 1) == 0) hiw = sel3(axis, m.inv_w[0], m.inv_w[1], m.inv_w[2])[sel3(axis, ec[0], ec[1], ec[2])];
  }
  CT av = CT(0), l0v = CT(0), lnv = CT(0);
  // split-BF16 contraction: entry (row, col) of A on the grid of its 4x4 pieces, zero past n
  constexpr int kPW = 4 * LineShape<NN, EPB, NT>::KB;  // grid width
  float apv = 0.f;
  if constexpr (kLines) {
    const int row = ot / kPW, col = ot - (ot / kPW) * kPW;
    if (ot >= 0 && row < NN && col < NN) apv = L.opf[0][row * NN + col];
  }
  if (ot >= 0 && ot < NN * NN) {
    if constexpr (std::is_same<CT, double>::value) av = L.opd[0][ot];
    else av = L.opf[0][ot];
  }
  if (ot >= 0 && ot < NN) {
    if constexpr (std::is_same<CT, double>::value) {
      l0v = L.opd[1][ot];
      lnv = L.opd[2][ot];
    } else {
      l0v = L.opf[1][ot];
      lnv = L.opf[2][ot];
    }
  }

  // face items (element, direction, face node) of this lane; item = fle | dir << 6 | fn << 9 |
  // self node << 16 | kItemExterior when the other side is the block's exterior without a trace
  // (zero state) | kItemLocal | local element << 25 when the neighbour element is one of this
  // workgroup's, or ~0u when the item has nothing to do
  double nrho[R], nu[R], nq[R];
  unsigned item[R];
  // the exterior trace pointers are the same for every lane: read them once, as scalars, and
  // select per lane (a per-lane indexed read would be a vector load whose wait holds back the
  // neighbour loads behind every load already in flight)
  const double* trace_ptr[6];
#pragma unroll
  for (int d = 0; d < 6; ++d) trace_ptr[d] = traces ? traces[d] : nullptr;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int f = tid + r * NT;
    nrho[r] = 0.0;
    nu[r] = 0.0;
    nq[r] = 0.0;
    item[r] = ~0u;
    if (f < NITEMS) {
      const int fle = f / (6 * NF2);
      const int rr = f - fle * (6 * NF2);
      const int dir = rr / NF2;
      const int fn = rr - dir * NF2;
      const int axis = 2 - (dir >> 1);
      const int he = e0 + fle;
      if (he < E && (LDG || ((mask >> (2 * axis)) & 3u) != 0u)) {
        const int b = fn / NN, cc = fn - (fn / NN) * NN;
        const int sn = face_node<NN>(axis, (dir & 1) ? NN - 1 : 0, b, cc);
        const int on = face_node<NN>(axis, (dir & 1) ? 0 : NN - 1, b, cc);
        item[r] = (unsigned)fle | ((unsigned)dir << 6) | ((unsigned)fn << 9) | ((unsigned)sn << 16);
        int ec[3];
        coords_of(fle, ec);
        const int nb = face_neighbour(m, he, ec, dir);
        const double* tr = trace_ptr[0];
#pragma unroll
        for (int d = 1; d < 6; ++d)
          if (dir == d) tr = trace_ptr[d];
        // one address per value, chosen without a branch: the neighbour's node in the state, or
        // the exterior trace; an exterior face without a trace reads the lane's own node and is
        // zeroed in phase 2 (kItemExterior)
        const bool inner = nb >= 0;
        const bool state = inner || tr == nullptr;
        item[r] |= (!inner && tr == nullptr) ? kItemExterior : 0u;
        // a neighbour in this workgroup: phase 2 reads its kernel functions from LDS, no load
        const bool local = inner && nb >= e0 && nb < e0 + EPB;
        item[r] |= local ? (kItemLocal | ((unsigned)(nb - e0) << 25)) : 0u;
        if (!local) {
          const uint32_t oe = (uint32_t)(inner ? nb : he);  // the element whose node is read
          const int onode = inner ? on : sn;
          const int64_t oi = state_slot<NN>(oe, onode);
          const int64_t ti = ((int64_t)(-1 - nb) * NF2 + fn) * 3;
          const double* prho = state ? L.q_in + oi : tr + ti;
          const double* pu = state ? L.q_in + oi + (1 + axis) * P : tr + ti + 1;
          nrho[r] = *prho;
          nu[r] = *pu;
          if constexpr (LDG) {
            const double* pq = state ? L.q_in + q_slot<NN>(fs, oe, axis, onode) : tr + ti + 2;
            nq[r] = *pq;
          }
        }
      }
    }
  }

  // ---- phase 1: kernel functions, flux coefficients and operators into LDS --------------
  if (valid) {
    const CT rho = (CT)q[0];
    CT* F = sF + le * 4 * NP;
    F[node] = momentum_k(rho, mom_k8(ba, ir0), mom_k9(c, ir0));
    const CT k11 = mass_k11(ba);
#pragma unroll
    for (int a = 0; a < 3; ++a) F[(1 + a) * NP + node] = mass_k((CT)q[1 + a], rho, r0, k11);
  }
  if constexpr (LDG) {
    if (valid) {
#pragma unroll
      for (int a = 0; a < 3; ++a) sm.sQ[(le * 3 + a) * NP + node] = (CT)qq[a];
    }
  }
  if (helper) {
    CT K[kNK];
#pragma unroll
    for (int k = 0; k < kNK; ++k) K[k] = CT(0);
    if (he < E) {
      face_coeffs<CT>(hdir & 1, (CT)ms_c, (CT)ms_r, (CT)ms_ir, (CT)mo_c, (CT)mo_r, (CT)mo_ba, (CT)mo_ir, K);
      if constexpr (!std::is_same<CT, double>::value) K[12] = (CT)(ms_r / mo_r);  // 1 for equal rho0
      if ((hdir & 1) == 0) sIw[hle * 3 + 2 - (hdir >> 1)] = (CT)hiw;
    }
#pragma unroll
    for (int k = 0; k < nk_of<CT>(); ++k) sK[ht * nk_of<CT>() + k] = K[k];
  }
  if (ot >= 0 && ot < NN * NN) sA[ot] = av;
  if (ot >= 0 && ot < NN) {
    sL0[ot] = l0v;
    sLN[ot] = lnv;
  }
  if constexpr (kLines) {
    if (ot >= 0 && ot < kPW * kPW) {
      const int row = ot / kPW, col = ot - (ot / kPW) * kPW;
      short pc[3];
      split3(apv, pc[0], pc[1], pc[2]);
      short* sp = reinterpret_cast<short*>(sm.sAp);
#pragma unroll
      for (int pi = 0; pi < 3; ++pi)
        sp[((((row >> 2) * (kPW / 4) + (col >> 2)) * 3 + pi) * 4 + (row & 3)) * 4 + (col & 3)] = pc[pi];
    }
  }
  __syncthreads();

  if constexpr (LATE) load_pre();

  // ---- phase 2: upwind flux pair of every face node -------------------------------------
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned it = item[r];
    if (it == ~0u) continue;
    const int fle = it & 63u, dir = (it >> 6) & 7u, fn = (it >> 9) & 127u;
    const int sn = (it >> 16) & 0x1FFu;
    const int axis = 2 - (dir >> 1);
    const CT* F = sF + fle * 4 * NP;
    const CT mom_s = F[sn];
    const CT mass_s = F[(1 + axis) * NP + sn];
    const CT* K = sK + (fle * 6 + dir) * nk_of<CT>();
    const bool zero = (it & kItemExterior) != 0u;  // exterior face without a trace
    double vr = nrho[r], vu = nu[r];
    // FP32 compute: an opaque use keeps the conversions here; folded back to the loads, they
    // would make each wave wait for its neighbour loads as soon as they are issued
    if constexpr (!std::is_same<CT, double>::value) asm volatile("" : "+v"(vr), "+v"(vu));
    const CT ro = zero ? CT(0) : (CT)vr, uo = zero ? CT(0) : (CT)vu;
    CT mom_o = momentum_k(ro, K[8], K[9]);
    CT mass_o = mass_k(uo, ro, K[10], K[11]);
    if (it & kItemLocal) {
      // the neighbour element is in this workgroup: phase 1 computed the same values
      const int b = fn / NN, cc = fn - (fn / NN) * NN;
      const int on = face_node<NN>(axis, (dir & 1) ? 0 : NN - 1, b, cc);
      const CT* Fo = sF + ((it >> 25) & 31u) * 4 * NP;
      mom_o = Fo[on];
      mass_o = Fo[(1 + axis) * NP + on];
    }
    CT* Fl = sFl + ((fle * 6 + dir) * NF2 + fn) * 2;
    if constexpr (kShiftPressureLines) {
      // the shift of the line through this face node along the face's axis: its middle node
      const int b = fn / NN, cc = fn - (fn / NN) * NN;
      const CT c0 = F[face_node<NN>(axis, NN / 2, b, cc)];
      mom_o -= c0 * K[12];
      const CT mom_sd = mom_s - c0;
      Fl[0] = K[0] * mom_o + K[1] * mom_sd + K[2] * mass_o + K[3] * mass_s;
      Fl[1] = K[4] * mass_o + K[5] * mass_s + K[6] * mom_o + K[7] * mom_sd;
    } else {
      Fl[0] = K[0] * mom_o + K[1] * mom_s + K[2] * mass_o + K[3] * mass_s;
      Fl[1] = K[4] * mass_o + K[5] * mass_s + K[6] * mom_o + K[7] * mom_s;
    }
    if constexpr (LDG) {
      double vq = nq[r];
      if constexpr (!std::is_same<CT, double>::value) asm volatile("" : "+v"(vq));
      CT qo = zero ? CT(0) : (CT)vq;
      if (it & kItemLocal) {
        const int b = fn / NN, cc = fn - (fn / NN) * NN;
        const int on = face_node<NN>(axis, (dir & 1) ? 0 : NN - 1, b, cc);
        qo = sm.sQ[(((it >> 25) & 31u) * 3 + axis) * NP + on];
      }
      const CT qs = sm.sQ[(fle * 3 + axis) * NP + sn];
      sm.sFq[(fle * 6 + dir) * NF2 + fn] = (dir & 1) ? CT(-0.5) * (qs + qo) : CT(0.5) * (qo + qs);
    }
  }
  // the matrix-core contraction reads only phase 1's LDS (sF, sA) and writes its own array, so
  // it runs beside the flux pairs, before the barrier that phase 3 needs anyway
  if constexpr (CONTRACT == 2 && !kLinesPerLane<NN>) contract_mfma_bf16<NN, EPB, NT>(sF, sA, sOut, tid);
  if constexpr (CONTRACT == 1) contract_lines_f32<NN, EPB, NT>(sF, sA, sOut, tid, ldesc);
  if constexpr (CONTRACT == 2 && kLinesPerLane<NN>) contract_lines_bf16<NN, EPB, NT>(sF, sm.sAp, sOut, tid, ldesc);
  __syncthreads();

  if (!io) return;

  // ---- phase 3: line contractions, lifting, sponge, RK update ---------------------------
  const int idx[3] = {node / NF2, (node / NN) % NN, node % NN};
  const CT* F = sF + le * 4 * NP;
  const CT* Fl = sFl + le * 6 * NF2 * 2;
  CT drho = CT(0);
  CT du[3] = {CT(0), CT(0), CT(0)};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const unsigned am = (mask >> (2 * a)) & 3u;
    if (am == 0u || !valid) continue;
    const int ix = idx[a];
    const int ls = (a == 0) ? NF2 : (a == 1 ? NN : 1);
    const int base = node - ix * ls;
    const int fn = (a == 0) ? idx[1] * NN + idx[2] : (a == 1 ? idx[0] * NN + idx[2] : idx[0] * NN + idx[1]);
    const int dlo = (a == 0) ? 4 : (a == 1 ? 2 : 0);
    const CT* Fa = F + (1 + a) * NP;
    CT dp = CT(0), dm = CT(0);
    if constexpr (CONTRACT != 0) {
      dp = sOut[(le * 6 + 2 * a) * NP + node];
      dm = sOut[(le * 6 + 2 * a + 1) * NP + node];
    } else {
      const CT c0 = kShiftPressureLines ? F[base + (NN / 2) * ls] : CT(0);
#pragma unroll
      for (int mm = 0; mm < NN; ++mm) {
        const CT av = sA[ix * NN + mm];
        if constexpr (kShiftPressureLines) dp += av * (F[base + mm * ls] - c0);
        else dp += av * F[base + mm * ls];
        dm += av * Fa[base + mm * ls];
      }
    }
    const CT iw = sIw[le * 3 + a];
    const CT l0 = sL0[ix], ln = sLN[ix];
    const CT* flo = Fl + (dlo * NF2 + fn) * 2;
    const CT* fhi = Fl + ((dlo + 1) * NF2 + fn) * 2;
    if (am & 1u) du[a] = iw * (dp - (flo[0] * l0 + fhi[0] * ln));
    if (am & 2u) drho += iw * (dm - (flo[1] * l0 + fhi[1] * ln));
  }
  if constexpr (LDG) {
    if (valid) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int ix = idx[a];
        const int ls = (a == 0) ? NF2 : (a == 1 ? NN : 1);
        const int base = node - ix * ls;
        const int fn = (a == 0) ? idx[1] * NN + idx[2] : (a == 1 ? idx[0] * NN + idx[2] : idx[0] * NN + idx[1]);
        const int dlo = (a == 0) ? 4 : (a == 1 ? 2 : 0);
        const CT* Qa = sm.sQ + (le * 3 + a) * NP;
        CT sq = CT(0);
#pragma unroll
        for (int mm = 0; mm < NN; ++mm) sq += sA[ix * NN + mm] * Qa[base + mm * ls];
        sq += sm.sFq[(le * 6 + dlo) * NF2 + fn] * sL0[ix] + sm.sFq[(le * 6 + dlo + 1) * NF2 + fn] * sLN[ix];
        drho += -dcoef * (sq * sIw[le * 3 + a]);
      }
    }
  }
  if (mask & kTermSponge) {
#pragma unroll
    for (int a = 0; a < 3; ++a) du[a] -= sg * (CT)q[1 + a];
    drho -= sg * (CT)q[0];
  }
  const CT rhs[4] = {drho, du[0], du[1], du[2]};

  if constexpr (EPI == EPI_ACCUM) {
    bool touch[4];
    touch[0] = (mask & (kTermSponge | 0x2Au)) != 0u;
#pragma unroll
    for (int a = 0; a < 3; ++a) touch[1 + a] = (mask & (kTermSponge | (1u << (2 * a)))) != 0u;
    const Rsrc ro = make_rsrc(ep_out + wofs, wb64);
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      if (touch[f]) bstore<double>(bload<double>(ro, gi + f * P) + (double)rhs[f], ro, gi + f * P);
    }
  } else if constexpr (EPI == EPI_LSRK) {
    const Rsrc rqo = make_rsrc(ep_qout + wofs, wb64);
    if constexpr (std::is_same<DUT, double>::value) {
      double dn[4];
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        dn[f] = ep_a * pre[f] + ep_dt * (double)rhs[f];
        if constexpr (!kDuPairMajor) bstore<double, kPolicyNT>(dn[f], rdu, gi + f * P);
        bstore<double, kPolicyNT>(q[f] + ep_b * dn[f], rqo, gi + f * P);
      }
      if constexpr (kDuPairMajor) {
        bstore2d<kPolicyNT>(dn[0], dn[1], rdu, du_off);
        bstore2d<kPolicyNT>(dn[2], dn[3], rdu, du_off + (uint32_t)P * 16u);
      }
    } else {
      const float af = (float)ep_a, dtf = (float)ep_dt;
      float dn[4];
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        dn[f] = af * (float)pre[f] + dtf * (float)rhs[f];
        if constexpr (!kDuNodeMajor) bstore<float, kPolicyNT>(dn[f], rdu, gi + f * P);
        bstore<double, kPolicyNT>(q[f] + ep_b * (double)dn[f], rqo, gi + f * P);
      }
      if constexpr (kDuNodeMajor) bstore4f<kPolicyNT>(dn, rdu, du_off);
    }
  } else {
    // Grid.cpp:705, 765-766, 826-827: out = (rhs*(dt*beta) + in*beta) + old*alpha. Nontemporal
    // stores, as in the LSRK epilogue: cached ones ran the stages 1.4-2.4 % slower (N = 2, 4, 5, 7)
    const double sb = ep_dt * ep_b;
    const Rsrc ro = make_rsrc(ep_out + wofs, wb64);
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      double v = (double)rhs[f] * sb + q[f] * ep_b;
      if (ep_a != 0.0) v += pre[f] * ep_a;
      bstore<double, kPolicyNT>(v, ro, gi + f * P);
    }
  }
}

// Waves per SIMD the register allocation must allow. The FP64 stage needs more than the 80
// registers a target of 6 leaves when the RK register is loaded early, and with late loads at
// every order but 2 and 5 nodes per line, and would spill to scratch (8-20 bytes per lane); those
// shapes target 5 (96 registers), and so do the matrix-core stages at 2 and 3 nodes per line
// (split-BF16: 20-28 bytes at 6; FP32 with the RK33 epilogue: 20 bytes). No kernel of the
// library uses scratch.
template <int NN, typename CT, int MINW, bool LATE, int CONTRACT = 0>
constexpr int waves_target() {
  constexpr bool fp64_heavy = sizeof(CT) == 8 && (!LATE || (NN != 2 && NN != 5));
  constexpr bool mfma_heavy = CONTRACT != 0 && NN <= 3;
  return (MINW > 5 && (fp64_heavy || mfma_heavy)) ? 5 : MINW;
}

// The waves per SIMD a shape's LDS leaves room for (160 KiB per CU, 4 SIMDs). The register target
// asks for no more: at 2 nodes per line the FP64 flux coefficients and pairs take 160 bytes of LDS
// per node, which holds those shapes to 3-4 waves per SIMD whatever their registers.
template <int NN, int EPB, typename CT, int CONTRACT, bool LDG>
constexpr int lds_waves() {
  constexpr int wgs = (160 * 1024) / (int)sizeof(StageSmem<NN, EPB, CT, CONTRACT, LDG>);
  constexpr int w = wgs * (Geo<NN, EPB>::NT / 64) / 4;
  return w < 1 ? 1 : w;
}
template <int NN, int EPB, int MINW, typename CT, bool LATE, int CONTRACT, bool LDG>
constexpr int stage_waves() {
  constexpr int r = waves_target<NN, CT, MINW, LATE, CONTRACT>();
  constexpr int l = lds_waves<NN, EPB, CT, CONTRACT, LDG>();
  return r < l ? r : l;
}

template <int NN, int EPB, int MINW, typename CT, int EPI, typename DUT, int CONTRACT = 0, bool LATE = false,
          bool LDG = false>
__global__ void __launch_bounds__((Geo<NN, EPB>::NT), (stage_waves<NN, EPB, MINW, CT, LATE, CONTRACT, LDG>())) k_rhs(
    const RhsLaunch L, const GroupMap gm) {
  __shared__ StageSmem<NN, EPB, CT, CONTRACT, LDG> sm;
  const uint32_t grp = map_group(blockIdx.x, gm);
  stage_body<NN, EPB, Geo<NN, EPB>::NT, CT, EPI, DUT, CONTRACT, LATE, LDG>(L, L.mesh, L.ext.trace, (int)(grp * EPB),
                                                                           (int)threadIdx.x, sm);
}

// ---- batched persistent stage kernels ---------------------------------------------------
// One kernel per order over the element groups (tasks) of every solve of that order in the batch.
// Each solve's stage parameters are a launch record prepared on the host (one per stage and
// arena parity), read through the constant address space. PERSIST: workgroup w takes tasks w,
// w + grid, ... (a static schedule: every workgroup leaves the loop after the last task);
// otherwise one task per workgroup.
using ConstLaunchPtr = const __attribute__((address_space(4))) RhsLaunch*;
using ConstTaskPtr = const __attribute__((address_space(4))) BatchTask*;

// The tasks of a launch in per-XCD slabs: the hardware sends workgroup w to XCD w mod 8, and XCD x
// takes the x-th eighth of the tasks, in order, so that a solve's neighbouring element groups meet
// in one XCD's L2 (in the hardware order they spread over all eight). One task per workgroup:
// workgroup w's task. Persistent (grid >= 8): the workgroups of XCD x take its slab's tasks
// round-robin. With the slabs, the 256-solve batch stage went from 1.084 to 1.001 ms mixed and from
// 1.34 to 1.27-1.29 ms FP64 (alternating runs on one box).
__device__ __forceinline__ void xcd_slab(uint32_t ntasks, uint32_t x, uint32_t* first, uint32_t* end) {
  const uint32_t per = ntasks >> 3, rem = ntasks & 7u;
  *first = x * per + (x < rem ? x : rem);
  *end = *first + per + (x < rem ? 1u : 0u);
}

template <int NN, int EPB, typename CT, typename DUT, bool PERSIST, bool LATE>
__global__ void __launch_bounds__((Geo<NN, EPB>::NT)) k_rhs_batch(const BatchLaunch B) {
  __shared__ StageSmem<NN, EPB, CT, 0> sm;
  const ConstTaskPtr tasks = (ConstTaskPtr)B.tasks;
  const uint32_t w = blockIdx.x, x = w & 7u;
  if constexpr (PERSIST) {
    // fewer than 8 workgroups leave some XCD without one: then the plain order
    uint32_t t0 = w, t1 = (uint32_t)B.ntasks, step = gridDim.x;
    if (gridDim.x >= 8) {
      xcd_slab((uint32_t)B.ntasks, x, &t0, &t1);
      step = (gridDim.x - x + 7) >> 3;  // the workgroups on XCD x
      t0 += w >> 3;
    }
    for (uint32_t t = t0; t < t1; t += step) {
      __syncthreads();  // the previous task's phase 3 still reads LDS
      // the lane index is re-read inside the loop: without it the compiler hoists the
      // task-independent lane arithmetic out of the loop and holds it in registers across tasks
      int tid;
      asm volatile("v_mov_b32 %0, %1" : "=v"(tid) : "v"((int)threadIdx.x));
      const int solve = tasks[t].solve, e0 = tasks[t].e0;
      const RhsLaunch& L = *(const RhsLaunch*)((ConstLaunchPtr)(B.records + (int64_t)solve * kBatchRecords + B.record));
      stage_body<NN, EPB, Geo<NN, EPB>::NT, CT, EPI_LSRK, DUT, 0, LATE>(L, L.mesh, L.ext.trace, e0, tid, sm);
    }
  } else {
    // the grid is one workgroup per task
    uint32_t first, end;
    xcd_slab(gridDim.x, x, &first, &end);
    const uint32_t t = first + (w >> 3);
    if (t >= (uint32_t)B.ntasks) return;
    const int solve = tasks[t].solve, e0 = tasks[t].e0;
    const RhsLaunch& L = *(const RhsLaunch*)((ConstLaunchPtr)(B.records + (int64_t)solve * kBatchRecords + B.record));
    stage_body<NN, EPB, Geo<NN, EPB>::NT, CT, EPI_LSRK, DUT, 0, LATE>(L, L.mesh, L.ext.trace, e0, (int)threadIdx.x,
                                                                     sm);
  }
}

#ifndef DG3D_TU_ORDER
// Every order of a batch in one persistent launch (batch schedule 4): 512-lane workgroups take the
// element groups of all solves from one task list (largest order first, per-XCD slabs) and run the
// stage body of the task's order, with that order's elements per workgroup (a 512-lane shape
// where one exists: 16, 8, 4 elements at 3, 4, 5 nodes per line, 32 at 2 and 2 at 6 leave half and
// an eighth of the lanes idle, 1 at 7 and 8). The LDS is the union of the orders' layouts and the
// registers the most any order needs, so every order runs at the occupancy of the heaviest.
template <int NN>
constexpr int all_epb() {
  return NN == 2 ? 32 : NN == 3 ? 16 : NN == 4 ? 8 : NN == 5 ? 4 : NN == 6 ? 2 : 1;
}
constexpr int kAllNT = 512;
template <typename CT>
union AllSmem {
  StageSmem<2, all_epb<2>(), CT, 0> s2;
  StageSmem<3, all_epb<3>(), CT, 0> s3;
  StageSmem<4, all_epb<4>(), CT, 0> s4;
  StageSmem<5, all_epb<5>(), CT, 0> s5;
  StageSmem<6, all_epb<6>(), CT, 0> s6;
  StageSmem<7, all_epb<7>(), CT, 0> s7;
  StageSmem<8, all_epb<8>(), CT, 0> s8;
};

// Two 512-lane workgroups per CU (4 waves per SIMD, 128 registers): without the bound the FP64
// kernel took 152 registers, so one workgroup fitted per CU and the stage ran 1.5x slower; at 6
// waves per SIMD both precisions spill 200-320 bytes.
template <typename CT, typename DUT>
__global__ void __launch_bounds__(kAllNT, 4) k_rhs_batch_all(const BatchLaunch B) {
  __shared__ AllSmem<CT> sm;
  const ConstTaskPtr tasks = (ConstTaskPtr)B.tasks;
  const uint32_t w = blockIdx.x, x = w & 7u;
  // fewer than 8 workgroups leave some XCD without one: then the plain order
  uint32_t t0 = w, t1 = (uint32_t)B.ntasks, step = gridDim.x;
  if (gridDim.x >= 8) {
    xcd_slab((uint32_t)B.ntasks, x, &t0, &t1);
    step = (gridDim.x - x + 7) >> 3;
    t0 += w >> 3;
  }
  for (uint32_t t = t0; t < t1; t += step) {
    __syncthreads();  // the previous task's phase 3 still reads LDS
    int tid;
    asm volatile("v_mov_b32 %0, %1" : "=v"(tid) : "v"((int)threadIdx.x));
    const int solve = tasks[t].solve, e0 = tasks[t].e0;
    const RhsLaunch& L = *(const RhsLaunch*)((ConstLaunchPtr)(B.records + (int64_t)solve * kBatchRecords + B.record));
    switch (L.n) {  // the same for every lane of the workgroup
      case 2: stage_body<2, all_epb<2>(), kAllNT, CT, EPI_LSRK, DUT, 0>(L, L.mesh, L.ext.trace, e0, tid, sm.s2); break;
      case 3: stage_body<3, all_epb<3>(), kAllNT, CT, EPI_LSRK, DUT, 0>(L, L.mesh, L.ext.trace, e0, tid, sm.s3); break;
      case 4: stage_body<4, all_epb<4>(), kAllNT, CT, EPI_LSRK, DUT, 0>(L, L.mesh, L.ext.trace, e0, tid, sm.s4); break;
      case 5: stage_body<5, all_epb<5>(), kAllNT, CT, EPI_LSRK, DUT, 0>(L, L.mesh, L.ext.trace, e0, tid, sm.s5); break;
      case 6: stage_body<6, all_epb<6>(), kAllNT, CT, EPI_LSRK, DUT, 0>(L, L.mesh, L.ext.trace, e0, tid, sm.s6); break;
      case 7: stage_body<7, all_epb<7>(), kAllNT, CT, EPI_LSRK, DUT, 0>(L, L.mesh, L.ext.trace, e0, tid, sm.s7); break;
      case 8: stage_body<8, all_epb<8>(), kAllNT, CT, EPI_LSRK, DUT, 0>(L, L.mesh, L.ext.trace, e0, tid, sm.s8); break;
      default: break;
    }
  }
}

// Exterior trace of the source face at one stage time: the face's halo trace (or zero) plus the
// boundary source, p = s(t + lead), rho' = s/c^2, u_axis = +-s/(rho0 c) entering the block.
// Node i of the source face's trace at stage time ts.
__device__ __forceinline__ void source_trace_node(const ExtArgs& x, const MeshArgs& m, int n, const double* halo,
                                                  double* out, double ts, int64_t i) {
  const int64_t nf2 = (int64_t)n * n;
  const int face = x.src_face;
  const int axis = 2 - face / 2;
  const int nb = (axis == 0) ? m.nE[1] : m.nE[0];
  const int nc = (axis == 2) ? m.nE[1] : m.nE[2];
  const int64_t count = (int64_t)nb * nc * nf2;
  if (i >= count) return;
  const int64_t fidx = i / nf2;
  const int eb = (int)(fidx / nc), ec = (int)(fidx % nc);
  int e3[3];
  e3[axis] = (face & 1) ? m.nE[axis] - 1 : 0;
  e3[axis == 0 ? 1 : 0] = eb;
  e3[axis == 2 ? 1 : 2] = ec;
  const int64_t e = ((int64_t)e3[0] * m.nE[1] + e3[1]) * m.nE[2] + e3[2];
  const double t = ts + (x.lead ? x.lead[i] : 0.0);
  const double sv = source_value(x, t);
  const double cd = m.mat[e * kMatStride + MAT_C], rd = m.mat[e * kMatStride + MAT_RHO0];
  const double h0 = halo ? halo[i * 3] : 0.0, h1 = halo ? halo[i * 3 + 1] : 0.0, h2 = halo ? halo[i * 3 + 2] : 0.0;
  out[i * 3 + 0] = h0 + sv / (cd * cd);
  out[i * 3 + 1] = h1 + ((face & 1) ? -1.0 : 1.0) * sv / (rd * cd);
  out[i * 3 + 2] = h2;
}

__global__ void k_source_trace(ExtArgs x, MeshArgs m, int n, const double* halo, double* out) {
  const double ts = x.clock ? x.clock[0] + x.t : x.t;
  source_trace_node(x, m, n, halo, out, ts, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// The source traces of every stage of one step in one launch (blockIdx.y = stage s): stage s at
// the step's start on the device clock (or 0 without one) plus offs.v[s], into out + s count 3.
// Only stage 0 adds `halo0`, the halo traces the first stage of a call reads.
__global__ void k_source_trace_step(ExtArgs x, MeshArgs m, int n, const double* halo0, double* out,
                                    StageTimes offs, int64_t count) {
  const int s = (int)blockIdx.y;
  const double ts = (x.clock ? x.clock[0] : 0.0) + offs.v[s];
  source_trace_node(x, m, n, s == 0 ? halo0 : nullptr, out + (int64_t)s * count * 3, ts,
                    (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Every batched solve's source trace for one stage: blockIdx.y picks the entry.
__global__ void k_batch_source(const BatchSource* list, const double* clocks, const double* offsets, int stage,
                               int first) {
  const BatchSource& b = list[blockIdx.y];
  const double ts = clocks[b.solve] + offsets[(int64_t)b.solve * 5 + stage];
  source_trace_node(b.x, b.m, b.n, first ? b.halo : nullptr, b.out, ts,
                    (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_batch_clocks(double* clocks, const double* dts, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) clocks[i] += dts[i];
}
#endif  // DG3D_TU_ORDER

// LDG operators for one axis (Element.cpp:623-758), FP64, one lane per node.
template <int NN>
__global__ void __launch_bounds__(Shape<NN>::NT) k_ldg(const LdgLaunch L) {
  using S = Shape<NN>;
  constexpr int NP = S::NP, NF2 = S::NF2, EPB = S::EPB;
  __shared__ double sV[EPB * NP];
  __shared__ double sFl[EPB * 2 * NF2];
  const MeshArgs& m = L.mesh;
  const int tid = threadIdx.x;
  const int e0 = (int)blockIdx.x * EPB;
  const int le = tid / NP;
  const int node = tid - le * NP;
  const int e = e0 + le;
  const bool valid = (le < EPB) && (e < (int)m.E);
  const int axis = L.axis;
  const int fin = (L.which == 0) ? 0 : 4 + axis;  // rho' or Q_axis
  if (valid) sV[le * NP + node] = L.in[reg_slot(m.fs, S::ES, group_elems(NN), fin, e, node)];
  __syncthreads();
  for (int f = tid; f < EPB * 2 * NF2; f += S::NT) {
    const int fle = f / (2 * NF2);
    const int r = f - fle * 2 * NF2;
    const int side = r / NF2;
    const int fn = r - side * NF2;
    const int fe = e0 + fle;
    if (fe >= (int)m.E) continue;
    const int b = fn / NN, cc = fn % NN;
    const int sn = face_node<NN>(axis, side ? NN - 1 : 0, b, cc);
    const int on = face_node<NN>(axis, side ? 0 : NN - 1, b, cc);
    const double vs = sV[fle * NP + sn];
    double vo = 0.0;
    int ec[3];
    elem_coords(m, (uint32_t)fe, ec);
    const int eo = neighbour(m, fe, ec, axis, side);
    const int dir = (axis == 0) ? 4 + side : (axis == 1 ? 2 + side : side);
    if (eo >= 0) {
      vo = L.in[reg_slot(m.fs, S::ES, group_elems(NN), fin, eo, on)];
    } else {
      const int64_t fidx = face_elem(m, ec, axis);
      const double* tr = L.ext.trace[dir];
      if (tr) vo = tr[(fidx * NF2 + fn) * 3 + (L.which == 0 ? 0 : 2)];
      if (L.which == 0 && dir == L.ext.src_face) {
        const double t = L.ext.t + (L.ext.lead ? L.ext.lead[fidx * NF2 + fn] : 0.0);
        const double cd = m.mat[(int64_t)fe * kMatStride + MAT_C];
        vo += source_value(L.ext, t) / (cd * cd);
      }
    }
    // central flux: +(v_o + v_s)/2 on the previous face, -(v_s + v_o)/2 on the next face
    sFl[(fle * 2 + side) * NF2 + fn] = side ? -0.5 * (vs + vo) : 0.5 * (vo + vs);
  }
  __syncthreads();
  if (!valid) return;
  const int idx[3] = {node / NF2, (node / NN) % NN, node % NN};
  const int ix = idx[axis];
  const int ls = (axis == 0) ? NF2 : (axis == 1 ? NN : 1);
  const int base = node - ix * ls;
  const int fn = (axis == 0) ? idx[1] * NN + idx[2] : (axis == 1 ? idx[0] * NN + idx[2] : idx[0] * NN + idx[1]);
  double s = 0.0;
  for (int mm = 0; mm < NN; ++mm) s += L.opd[0][ix * NN + mm] * sV[le * NP + base + mm * ls];
  s += sFl[(le * 2 + 0) * NF2 + fn] * L.opd[1][ix] + sFl[(le * 2 + 1) * NF2 + fn] * L.opd[2][ix];
  int ec[3];
  elem_coords(m, (uint32_t)e, ec);
  const double x = s * m.inv_w[axis][ec[axis]];
  if (L.which == 0) {
    L.out[reg_slot(m.fs, S::ES, group_elems(NN), 4 + axis, e, node)] = -x;
  } else {
    L.out[reg_slot(m.fs, S::ES, group_elems(NN), 0, e, node)] += -L.diff[e] * x;
  }
}

// Element::OperateQFromDRho (Element.cpp:623-690) for the three axes in one pass, in place:
// Q_a = -(1/w_a) (A rho' + L0 F_prev + LN F_next), with the central traces F_prev =
// (rho'_o + rho'_s)/2 and F_next = -(rho'_s + rho'_o)/2 (the exterior trace on block faces,
// the source included). FP64, one lane per node; writes fields 4..6, reads field 0.
template <int NN>
__global__ void __launch_bounds__((Geo<NN, Shape<NN>::EPB>::NT)) k_ldg_q(const LdgLaunch L) {
  constexpr int EPB = Shape<NN>::EPB;
  using S = Geo<NN, EPB>;
  constexpr int NP = S::NP, NF2 = S::NF2, ES = S::ES, NT = S::NT;
  constexpr int NITEMS = EPB * 6 * NF2;
  constexpr int R = (NITEMS + NT - 1) / NT;
  __shared__ double sV[EPB * NP];
  __shared__ double sFl[EPB * 6 * NF2];
  __shared__ double sA[NN * NN], sL0[NN], sLN[NN], sIw[EPB * 3];
  const MeshArgs& m = L.mesh;
  const int E = (int)m.E;
  const int tid = threadIdx.x;
  const int e0 = (int)blockIdx.x * EPB;
  const int le = tid / ES;
  const int node = tid - le * ES;
  const int e = e0 + le;
  const bool valid = (le < EPB) && (e < E) && node < NP;
  const int64_t ve = valid ? e : 0;

  // ---- phase 0: every global load of the workgroup (as in the stage kernel) ---------------
  const double v = valid ? L.in[state_slot<NN>((uint32_t)ve, node)] : 0.0;
  const double* trace_ptr[6];
#pragma unroll
  for (int d = 0; d < 6; ++d) trace_ptr[d] = L.ext.trace[d];
  double vo[R];
  unsigned item[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int f = tid + r * NT;
    vo[r] = 0.0;
    item[r] = ~0u;
    if (f < NITEMS) {
      const int fle = f / (6 * NF2);
      const int rr = f - fle * (6 * NF2);
      const int dir = rr / NF2;
      const int fn = rr - dir * NF2;
      const int fe = e0 + fle;
      if (fe < E) {
        const int axis = 2 - (dir >> 1);
        const int b = fn / NN, cc = fn - (fn / NN) * NN;
        const int sn = face_node<NN>(axis, (dir & 1) ? NN - 1 : 0, b, cc);
        const int on = face_node<NN>(axis, (dir & 1) ? 0 : NN - 1, b, cc);
        int ec[3];
        elem_coords(m, (uint32_t)fe, ec);
        const int nb = face_neighbour(m, fe, ec, dir);
        const double* tr = trace_ptr[0];
#pragma unroll
        for (int d = 1; d < 6; ++d)
          if (dir == d) tr = trace_ptr[d];
        // the neighbour's rho', or the exterior trace's; an exterior face without a trace reads
        // the lane's own node and is zeroed below (kItemExterior)
        const bool inner = nb >= 0;
        const bool state = inner || tr == nullptr;
        const int64_t oi = state_slot<NN>((uint32_t)(inner ? nb : fe), inner ? on : sn);
        const double* pv = state ? L.in + oi : tr + ((int64_t)(-1 - nb) * NF2 + fn) * 3;
        vo[r] = *pv;
        item[r] = (unsigned)fle | ((unsigned)dir << 6) | ((unsigned)fn << 9) | ((unsigned)sn << 16) |
                  ((!inner && tr == nullptr) ? kItemExterior : 0u);
      }
    }
  }
  // operators on the second wave and 1/width on the last, when they fit there, so that wave 0
  // issues only its own loads, as in the stage kernel
  constexpr int HB = ((NT / 64 - 1) * 64 + EPB * 3 <= NT) ? (NT / 64 - 1) * 64 : 0;
  constexpr int OB = (NT >= 128 && 64 + NN * NN <= NT) ? 64 : 0;
  const int ht = tid - HB, ot = tid - OB;
  double av = 0.0, l0v = 0.0, lnv = 0.0, iwv = 0.0;
  if (ot >= 0 && ot < NN * NN) av = L.opd[0][ot];
  if (ot >= 0 && ot < NN) {
    l0v = L.opd[1][ot];
    lnv = L.opd[2][ot];
  }
  if (ht >= 0 && ht < EPB * 3) {
    const int he = e0 + ht / 3, a = ht - (ht / 3) * 3;
    if (he < E) {
      int ec[3];
      elem_coords(m, (uint32_t)he, ec);
      iwv = sel3(a, m.inv_w[0], m.inv_w[1], m.inv_w[2])[sel3(a, ec[0], ec[1], ec[2])];
    }
  }

  // ---- phase 1: the workgroup's rho', operators and 1/width into LDS ----------------------
  if (valid) sV[le * NP + node] = v;
  if (ot >= 0 && ot < NN * NN) sA[ot] = av;
  if (ot >= 0 && ot < NN) {
    sL0[ot] = l0v;
    sLN[ot] = lnv;
  }
  if (ht >= 0 && ht < EPB * 3) sIw[ht] = iwv;
  __syncthreads();

  // ---- phase 2: central fluxes of the face nodes -----------------------------------------
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned it = item[r];
    if (it == ~0u) continue;
    const int fle = it & 63u, dir = (it >> 6) & 7u, fn = (it >> 9) & 127u;
    const int sn = (it >> 16) & 0x1FFu;
    const double vs = sV[fle * NP + sn];
    const double o = (it & kItemExterior) ? 0.0 : vo[r];
    sFl[(fle * 6 + dir) * NF2 + fn] = (dir & 1) ? -0.5 * (vs + o) : 0.5 * (o + vs);
  }
  __syncthreads();
  if (!valid) return;

  // ---- phase 3: Q_a = -(1/w_a) (A rho' + L0 F_prev + LN F_next) ---------------------------
  const int idx[3] = {node / NF2, (node / NN) % NN, node % NN};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int ix = idx[a];
    const int ls = (a == 0) ? NF2 : (a == 1 ? NN : 1);
    const int base = node - ix * ls;
    const int fn = (a == 0) ? idx[1] * NN + idx[2] : (a == 1 ? idx[0] * NN + idx[2] : idx[0] * NN + idx[1]);
    const int dlo = (a == 0) ? 4 : (a == 1 ? 2 : 0);
    double sv = 0.0;
#pragma unroll
    for (int mm = 0; mm < NN; ++mm) sv += sA[ix * NN + mm] * sV[le * NP + base + mm * ls];
    sv += sFl[(le * 6 + dlo) * NF2 + fn] * sL0[ix] + sFl[(le * 6 + dlo + 1) * NF2 + fn] * sLN[ix];
    // nontemporal: the stage that reads Q next ran 3-4 % faster than after cached stores
    __builtin_nontemporal_store(-(sv * sIw[le * 3 + a]), L.out + q_slot<NN>(m.fs, (uint32_t)ve, a, node));
  }
}

#ifndef DG3D_TU_ORDER
__global__ void k_axpby(double* out, double a, const double* in, double b, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = out[i] * a + in[i] * b;
}

// Pseudo-random values in [-amp, amp), all normal numbers (a 64-bit mix of the index).
__global__ void k_fill_pattern(double* out, double amp, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull;
    h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
    h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
    h ^= h >> 31;
    const double u = (double)(h >> 11) * 0x1.0p-53;  // [0, 1)
    out[i] = amp * (0.0625 + 0.9375 * u) * ((h & 1) ? 1.0 : -1.0);
  }
}

constexpr int kRedThreads = 256;
constexpr int kRedMaxBlocks = 1024;

__device__ __forceinline__ double nan_max(double a, double b) {
  if (isnan(a) || isnan(b)) return __builtin_nan("");
  return a > b ? a : b;
}

template <typename Op>
__device__ double block_reduce(double v, Op op, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kRedThreads / 64; ++w) v = op(v, sh[w]);
  }
  __syncthreads();
  return v;
}

// The state fields' region of a register, slots 0 .. 4 fs - 1: slot i belongs to field
// (i / kRunSlots) mod 4 (gaps, padding and the slots past the last element hold zeros).
__global__ void __launch_bounds__(kRedThreads) k_maxabs(const double* reg, int64_t fs, double* part) {
  __shared__ double sh[kRedThreads / 64];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 4 * fs; i += (int64_t)gridDim.x * blockDim.x) {
    const int f = (int)((uint32_t)i / (uint32_t)kRunSlots) & 3;  // 4 fs < 2^31
    const double x = fabs(reg[i]);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k == f) v[k] = nan_max(v[k], x);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double r = block_reduce(v[k], nan_max, sh);
    if (threadIdx.x == 0) part[blockIdx.x * 4 + k] = r;
  }
}

__global__ void __launch_bounds__(kRedThreads) k_l2diff(const double* a, const double* b, int64_t fs, double* part) {
  __shared__ double sh[kRedThreads / 64];
  auto add = [](double x, double y) { return x + y; };
  double d[4] = {0.0, 0.0, 0.0, 0.0}, r[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 4 * fs; i += (int64_t)gridDim.x * blockDim.x) {
    const int f = (int)((uint32_t)i / (uint32_t)kRunSlots) & 3;  // 4 fs < 2^31
    const double x = a[i], y = b[i];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k == f) {
        d[k] += (x - y) * (x - y);
        r[k] += y * y;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double dd = block_reduce(d[k], add, sh);
    const double rr = block_reduce(r[k], add, sh);
    if (threadIdx.x == 0) {
      part[(blockIdx.x * 4 + k) * 2 + 0] = dd;
      part[(blockIdx.x * 4 + k) * 2 + 1] = rr;
    }
  }
}

__global__ void __launch_bounds__(kRedThreads) k_energy(const double* reg, MeshArgs m, int n, const double* wq, double* part) {
  __shared__ double sh[kRedThreads / 64];
  auto add = [](double x, double y) { return x + y; };
  const int np = n * n * n;
  double en = 0.0, ms = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m.fs; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = i / m.es;  // node slot i of the block, element-major
    const int nd = (int)(i - e * m.es);
    if (e >= m.E || nd >= np) continue;  // padding slots
    const int a = nd / (n * n), b = (nd / n) % n, c = nd % n;
    int ec[3];
    elem_coords(m, (uint32_t)e, ec);
    const double jac = 0.125 / (m.inv_w[0][ec[0]] * m.inv_w[1][ec[1]] * m.inv_w[2][ec[2]]);
    const double w = wq[a] * wq[b] * wq[c] * jac;
    const double c0 = m.mat[e * kMatStride + MAT_C], r0 = m.mat[e * kMatStride + MAT_RHO0];
    const double* blk = reg + reg_slot(m.fs, m.es, m.grp, 0, e, nd);
    const double r = blk[0];
    const double u = blk[kRunSlots], v = blk[2 * kRunSlots], x = blk[3 * kRunSlots];
    en += w * 0.5 * (c0 * c0 * r * r / r0 + r0 * (u * u + v * v + x * x));
    ms += w * r;
  }
  en = block_reduce(en, add, sh);
  ms = block_reduce(ms, add, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x * 2 + 0] = en;
    part[blockIdx.x * 2 + 1] = ms;
  }
}

__global__ void __launch_bounds__(kRedThreads) k_cfl(MeshArgs m, double dxi_half, double* part) {
  __shared__ double sh[kRedThreads / 64];
  auto mn = [](double x, double y) { return x < y ? x : y; };
  double v = 1e300;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m.E; e += (int64_t)gridDim.x * blockDim.x) {
    int ec[3];
    elem_coords(m, (uint32_t)e, ec);
    // d-dimensional CFL: c dt sum_a 1/dx_a <= CFL with dx_a = w_a dxi_min / 2
    const double iw = m.inv_w[0][ec[0]] + m.inv_w[1][ec[1]] + m.inv_w[2][ec[2]];
    v = mn(v, dxi_half / (iw * m.mat[e * kMatStride + MAT_C]));
  }
  v = block_reduce(v, mn, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void k_probe(const double* reg, MeshArgs m, int n, const int64_t* nodes, int count, double* buf,
                        const int* step, int cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = step[0];
  if (i >= count || row >= cap) return;
  double* out = buf + (int64_t)row * count;
  const int64_t g = nodes[i];  // host node index
  const int64_t np = (int64_t)n * n * n;
  const int64_t e = g / np;
  const double c = m.mat[e * kMatStride + MAT_C], r0 = m.mat[e * kMatStride + MAT_RHO0];
  const double ba = m.mat[e * kMatStride + MAT_BA];
  const double r = reg[reg_slot(m.fs, m.es, m.grp, 0, e, (int)(g - e * np))];
  out[i] = c * c * (r + ba * 0.5 / r0 * r * r);
}

__global__ void k_clock_set(double* clock, double t) {
  if (threadIdx.x == 0) clock[0] = t;
}

__global__ void k_step_end(double* clock, double dt, int* probe_step, int cap) {
  if (threadIdx.x != 0) return;
  clock[0] += dt;
  if (probe_step[0] < cap) probe_step[0] += 1;
}
#endif  // DG3D_TU_ORDER

// The fused stage with the LDG diffusion term, VALU contraction.
template <int NN, int EPB>
hipError_t launch_rhs_ldg_e(const RhsLaunch& L, hipStream_t s) {
  using S = Geo<NN, EPB>;
  const int64_t blocks = (L.mesh.E + EPB - 1) / EPB;
  if (blocks <= 0) return hipSuccess;
  dim3 grid((unsigned)blocks), block(S::NT);
  const GroupMap gm = group_map_for(L, blocks, EPB);
  if (L.epilogue == EPI_LSRK) {
    if (!L.mixed) hipLaunchKernelGGL((k_rhs<NN, EPB, 1, double, EPI_LSRK, double, 0, false, true>), grid, block, 0, s, L, gm);
    else hipLaunchKernelGGL((k_rhs<NN, EPB, 1, float, EPI_LSRK, float, 0, false, true>), grid, block, 0, s, L, gm);
  } else if (L.epilogue == EPI_RK33) {
    if (!L.mixed) hipLaunchKernelGGL((k_rhs<NN, EPB, 1, double, EPI_RK33, double, 0, false, true>), grid, block, 0, s, L, gm);
    else hipLaunchKernelGGL((k_rhs<NN, EPB, 1, float, EPI_RK33, double, 0, false, true>), grid, block, 0, s, L, gm);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// The stage with the LDG term: the default elements-per-workgroup shape, VALU contraction, no
// occupancy target, early RK-register loads. (The alternative shape, which the tuner often picks
// for the stage without diffusion, measured 2 % slower with it at 64^3 N=4.)
template <int NN>
hipError_t launch_rhs_ldg(const RhsLaunch& L, hipStream_t s) {
  return launch_rhs_ldg_e<NN, Variants<NN>::EPB0>(L, s);
}

template <int NN, int EPB, int MINW, bool LATE>
hipError_t launch_rhs_v(const RhsLaunch& L, hipStream_t s) {
  using S = Geo<NN, EPB>;
  const int64_t blocks = (L.mesh.E + EPB - 1) / EPB;
  if (blocks <= 0) return hipSuccess;
  dim3 grid((unsigned)blocks), block(S::NT);
  const GroupMap gm = group_map_for(L, blocks, EPB);
  if (L.epilogue == EPI_ACCUM) {
    if (L.mixed) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_rhs<NN, EPB, MINW, double, EPI_ACCUM, double>), grid, block, 0, s, L, gm);
  } else if (L.epilogue == EPI_LSRK) {
    if (!L.mixed) hipLaunchKernelGGL((k_rhs<NN, EPB, MINW, double, EPI_LSRK, double, 0, LATE>), grid, block, 0, s, L, gm);
    else if (L.contraction == 1) hipLaunchKernelGGL((k_rhs<NN, EPB, MINW, float, EPI_LSRK, float, 1>), grid, block, 0, s, L, gm);
    else if (L.contraction == 2) hipLaunchKernelGGL((k_rhs<NN, EPB, MINW, float, EPI_LSRK, float, 2>), grid, block, 0, s, L, gm);
    else hipLaunchKernelGGL((k_rhs<NN, EPB, MINW, float, EPI_LSRK, float, 0, LATE>), grid, block, 0, s, L, gm);
  } else {
    if (!L.mixed) hipLaunchKernelGGL((k_rhs<NN, EPB, MINW, double, EPI_RK33, double, 0, LATE>), grid, block, 0, s, L, gm);
    else if (L.contraction == 1) hipLaunchKernelGGL((k_rhs<NN, EPB, MINW, float, EPI_RK33, double, 1>), grid, block, 0, s, L, gm);
    else if (L.contraction == 2) hipLaunchKernelGGL((k_rhs<NN, EPB, MINW, float, EPI_RK33, double, 2>), grid, block, 0, s, L, gm);
    else hipLaunchKernelGGL((k_rhs<NN, EPB, MINW, float, EPI_RK33, double, 0, LATE>), grid, block, 0, s, L, gm);
  }
  return hipGetLastError();
}

// Variant bits: 1 alternative elements per workgroup, 2 occupancy target, 4 late RK-register
// loads (VALU contraction only; the matrix-core contractions ignore the bit).
template <int NN>
hipError_t launch_rhs_n(const RhsLaunch& L, hipStream_t s) {
  using V = Variants<NN>;
  if (L.ldg) return launch_rhs_ldg<NN>(L, s);
  // the accumulate epilogue (Block::Operate*) only uses the default shape
  const int v = (L.epilogue == EPI_ACCUM) ? 0 : L.variant;
  switch (v) {
    case 0: return launch_rhs_v<NN, V::EPB0, 1, false>(L, s);
    case 1: return launch_rhs_v<NN, V::EPB1, 1, false>(L, s);
    case 2: return launch_rhs_v<NN, V::EPB0, V::MINW1, false>(L, s);
    case 3: return launch_rhs_v<NN, V::EPB1, V::MINW1, false>(L, s);
    case 4: return launch_rhs_v<NN, V::EPB0, 1, true>(L, s);
    case 5: return launch_rhs_v<NN, V::EPB1, 1, true>(L, s);
    case 6: return launch_rhs_v<NN, V::EPB0, V::MINW1, true>(L, s);
    case 7: return launch_rhs_v<NN, V::EPB1, V::MINW1, true>(L, s);
    default: return hipErrorInvalidValue;
  }
}

template <int NN>
hipError_t launch_ldg_q_n(const LdgLaunch& L, hipStream_t s) {
  constexpr int EPB = Shape<NN>::EPB;
  const int64_t blocks = (L.mesh.E + EPB - 1) / EPB;
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_ldg_q<NN>), dim3((unsigned)blocks), dim3(Geo<NN, EPB>::NT), 0, s, L);
  return hipGetLastError();
}

template <int NN>
hipError_t launch_ldg_n(const LdgLaunch& L, hipStream_t s) {
  using S = Shape<NN>;
  const int64_t blocks = (L.mesh.E + S::EPB - 1) / S::EPB;
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_ldg<NN>), dim3((unsigned)blocks), dim3(S::NT), 0, s, L);
  return hipGetLastError();
}

// Per-order launchers. The stage kernels of one order are instantiated in a translation unit of
// their own (this file built with DG3D_TU_ORDER = 2..8), so the orders compile in parallel; the
// file built without it holds the dispatch on the order and the order-independent kernels.
template <int NN, int EPB, bool PERSIST, bool LATE>
hipError_t launch_batch_shape(const BatchLaunch& B, int grid, hipStream_t s) {
  constexpr int NT = Geo<NN, EPB>::NT;
  if (!PERSIST) grid = B.ntasks;
  if (B.mixed) hipLaunchKernelGGL((k_rhs_batch<NN, EPB, float, float, PERSIST, LATE>), dim3(grid), dim3(NT), 0, s, B);
  else hipLaunchKernelGGL((k_rhs_batch<NN, EPB, double, double, PERSIST, LATE>), dim3(grid), dim3(NT), 0, s, B);
  return hipGetLastError();
}
// variant bit 0: elements-per-workgroup shape (EPB0 / EPB1); bit 1: one task per workgroup;
// bit 2: late RK-register loads
template <int NN>
hipError_t launch_batch_n(const BatchLaunch& B, int variant, int grid, hipStream_t s) {
  using V = Variants<NN>;
  switch (variant) {
    case 0: return launch_batch_shape<NN, V::EPB0, true, false>(B, grid, s);
    case 1: return launch_batch_shape<NN, V::EPB1, true, false>(B, grid, s);
    case 2: return launch_batch_shape<NN, V::EPB0, false, false>(B, grid, s);
    case 3: return launch_batch_shape<NN, V::EPB1, false, false>(B, grid, s);
    case 4: return launch_batch_shape<NN, V::EPB0, true, true>(B, grid, s);
    case 5: return launch_batch_shape<NN, V::EPB1, true, true>(B, grid, s);
    case 6: return launch_batch_shape<NN, V::EPB0, false, true>(B, grid, s);
    default: return launch_batch_shape<NN, V::EPB1, false, true>(B, grid, s);
  }
}

// resident workgroups of the persistent form of a variant (the grid of its launches)
template <int NN, int EPB, bool LATE>
int resident_shape(int mixed) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  constexpr int NT = Geo<NN, EPB>::NT;
  hipError_t e =
      mixed ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_rhs_batch<NN, EPB, float, float, true, LATE>, NT, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_rhs_batch<NN, EPB, double, double, true, LATE>, NT, 0);
  if (e != hipSuccess || per < 1) per = 1;
  return cus * per;
}
template <int NN>
int resident_n(int variant, int mixed) {
  using V = Variants<NN>;
  const bool alt = (variant & 1) != 0, late = (variant & 4) != 0;
  if (late) return alt ? resident_shape<NN, V::EPB1, true>(mixed) : resident_shape<NN, V::EPB0, true>(mixed);
  return alt ? resident_shape<NN, V::EPB1, false>(mixed) : resident_shape<NN, V::EPB0, false>(mixed);
}

}  // namespace

#define DG3D_ORDER_API(n)                                                                     \
  hipError_t launch_rhs_o##n(const RhsLaunch& L, hipStream_t s);                             \
  hipError_t launch_ldg_o##n(const LdgLaunch& L, hipStream_t s);                             \
  hipError_t launch_ldg_q_o##n(const LdgLaunch& L, hipStream_t s);                           \
  hipError_t launch_batch_o##n(const BatchLaunch& B, int variant, int grid, hipStream_t s);  \
  int resident_o##n(int variant, int mixed);
DG3D_ORDER_API(2)
DG3D_ORDER_API(3)
DG3D_ORDER_API(4)
DG3D_ORDER_API(5)
DG3D_ORDER_API(6)
DG3D_ORDER_API(7)
DG3D_ORDER_API(8)
#undef DG3D_ORDER_API

#ifdef DG3D_TU_ORDER
#define DG3D_CAT2(a, b) a##b
#define DG3D_CAT(a, b) DG3D_CAT2(a, b)
hipError_t DG3D_CAT(launch_rhs_o, DG3D_TU_ORDER)(const RhsLaunch& L, hipStream_t s) {
  return launch_rhs_n<DG3D_TU_ORDER>(L, s);
}
hipError_t DG3D_CAT(launch_ldg_o, DG3D_TU_ORDER)(const LdgLaunch& L, hipStream_t s) {
  return launch_ldg_n<DG3D_TU_ORDER>(L, s);
}
hipError_t DG3D_CAT(launch_ldg_q_o, DG3D_TU_ORDER)(const LdgLaunch& L, hipStream_t s) {
  return launch_ldg_q_n<DG3D_TU_ORDER>(L, s);
}
hipError_t DG3D_CAT(launch_batch_o, DG3D_TU_ORDER)(const BatchLaunch& B, int variant, int grid, hipStream_t s) {
  return launch_batch_n<DG3D_TU_ORDER>(B, variant, grid, s);
}
int DG3D_CAT(resident_o, DG3D_TU_ORDER)(int variant, int mixed) { return resident_n<DG3D_TU_ORDER>(variant, mixed); }
#else

hipError_t launch_rhs(const RhsLaunch& L, hipStream_t s) {
  switch (L.n) {
    case 2: return launch_rhs_o2(L, s);
    case 3: return launch_rhs_o3(L, s);
    case 4: return launch_rhs_o4(L, s);
    case 5: return launch_rhs_o5(L, s);
    case 6: return launch_rhs_o6(L, s);
    case 7: return launch_rhs_o7(L, s);
    case 8: return launch_rhs_o8(L, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_ldg(const LdgLaunch& L, hipStream_t s) {
  switch (L.n) {
    case 2: return launch_ldg_o2(L, s);
    case 3: return launch_ldg_o3(L, s);
    case 4: return launch_ldg_o4(L, s);
    case 5: return launch_ldg_o5(L, s);
    case 6: return launch_ldg_o6(L, s);
    case 7: return launch_ldg_o7(L, s);
    case 8: return launch_ldg_o8(L, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_ldg_q(const LdgLaunch& L, hipStream_t s) {
  switch (L.n) {
    case 2: return launch_ldg_q_o2(L, s);
    case 3: return launch_ldg_q_o3(L, s);
    case 4: return launch_ldg_q_o4(L, s);
    case 5: return launch_ldg_q_o5(L, s);
    case 6: return launch_ldg_q_o6(L, s);
    case 7: return launch_ldg_q_o7(L, s);
    case 8: return launch_ldg_q_o8(L, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_rhs_batch(const BatchLaunch& B, int n, int variant, int grid, hipStream_t s) {
  if (B.ntasks <= 0) return hipSuccess;
  if (grid < 1) grid = 1;
  if (variant < 0 || variant >= kNumBatchVariants) return hipErrorInvalidValue;
  switch (n) {
    case 2: return launch_batch_o2(B, variant, grid, s);
    case 3: return launch_batch_o3(B, variant, grid, s);
    case 4: return launch_batch_o4(B, variant, grid, s);
    case 5: return launch_batch_o5(B, variant, grid, s);
    case 6: return launch_batch_o6(B, variant, grid, s);
    case 7: return launch_batch_o7(B, variant, grid, s);
    case 8: return launch_batch_o8(B, variant, grid, s);
    default: return hipErrorInvalidValue;
  }
}

int batch_max_grid(int n, int variant, int mixed) {
  switch (n) {
    case 2: return resident_o2(variant, mixed);
    case 3: return resident_o3(variant, mixed);
    case 4: return resident_o4(variant, mixed);
    case 5: return resident_o5(variant, mixed);
    case 6: return resident_o6(variant, mixed);
    case 7: return resident_o7(variant, mixed);
    case 8: return resident_o8(variant, mixed);
    default: return 0;
  }
}

int batch_all_epb(int n) {
  switch (n) {
    case 2: return all_epb<2>();
    case 3: return all_epb<3>();
    case 4: return all_epb<4>();
    case 5: return all_epb<5>();
    case 6: return all_epb<6>();
    case 7: return all_epb<7>();
    case 8: return all_epb<8>();
    default: return 0;
  }
}

hipError_t launch_rhs_batch_all(const BatchLaunch& B, int grid, hipStream_t s) {
  if (B.ntasks <= 0) return hipSuccess;
  if (grid < 1) grid = 1;
  if (B.mixed) hipLaunchKernelGGL((k_rhs_batch_all<float, float>), dim3(grid), dim3(kAllNT), 0, s, B);
  else hipLaunchKernelGGL((k_rhs_batch_all<double, double>), dim3(grid), dim3(kAllNT), 0, s, B);
  return hipGetLastError();
}

int batch_all_max_grid(int mixed) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const hipError_t e =
      mixed ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_rhs_batch_all<float, float>, kAllNT, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_rhs_batch_all<double, double>, kAllNT, 0);
  if (e != hipSuccess || per < 1) per = 1;
  return cus * per;
}

int batch_epb(int variant, int n) {
  const bool alt = (variant & 1) != 0;
  switch (n) {
    case 2: return alt ? Variants<2>::EPB1 : Variants<2>::EPB0;
    case 3: return alt ? Variants<3>::EPB1 : Variants<3>::EPB0;
    case 4: return alt ? Variants<4>::EPB1 : Variants<4>::EPB0;
    case 5: return alt ? Variants<5>::EPB1 : Variants<5>::EPB0;
    case 6: return alt ? Variants<6>::EPB1 : Variants<6>::EPB0;
    case 7: return alt ? Variants<7>::EPB1 : Variants<7>::EPB0;
    case 8: return alt ? Variants<8>::EPB1 : Variants<8>::EPB0;
    default: return 0;
  }
}

hipError_t launch_source_trace(const ExtArgs& x, const MeshArgs& m, int n, const double* halo, double* out,
                               hipStream_t s) {
  if (x.src_face < 0) return hipSuccess;
  const int axis = 2 - x.src_face / 2;
  const int64_t nb = (axis == 0) ? m.nE[1] : m.nE[0];
  const int64_t nc = (axis == 2) ? m.nE[1] : m.nE[2];
  const int64_t count = nb * nc * n * n;
  hipLaunchKernelGGL(k_source_trace, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, x, m, n, halo, out);
  return hipGetLastError();
}

hipError_t launch_source_trace_step(const ExtArgs& x, const MeshArgs& m, int n, const double* halo0, double* out,
                                    const StageTimes& offs, int stages, hipStream_t s) {
  if (x.src_face < 0) return hipSuccess;
  if (stages < 1 || stages > kMaxStages) return hipErrorInvalidValue;
  const int64_t count = source_trace_count(x.src_face, m, n);
  hipLaunchKernelGGL(k_source_trace_step, dim3((unsigned)((count + 255) / 256), (unsigned)stages), dim3(256), 0, s, x,
                     m, n, halo0, out, offs, count);
  return hipGetLastError();
}

hipError_t launch_batch_sources(const BatchSource* list, int count, int64_t max_items, const double* clocks,
                                const double* offsets, int stage, int first, hipStream_t s) {
  if (count <= 0 || max_items <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_source, dim3((unsigned)((max_items + 255) / 256), (unsigned)count), dim3(256), 0, s,
                     list, clocks, offsets, stage, first);
  return hipGetLastError();
}

hipError_t launch_batch_clocks(double* clocks, const double* dts, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_clocks, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, clocks, dts, count);
  return hipGetLastError();
}

hipError_t launch_axpby(double* out, double a, const double* in, double b, int64_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)blocks), dim3(256), 0, s, out, a, in, b, count);
  return hipGetLastError();
}

hipError_t launch_fill_pattern(double* out, double amp, int64_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill_pattern, dim3((unsigned)blocks), dim3(256), 0, s, out, amp, count);
  return hipGetLastError();
}

int reduce_blocks(int64_t count) {
  int64_t b = (count + kRedThreads - 1) / kRedThreads;
  if (b > kRedMaxBlocks) b = kRedMaxBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

hipError_t launch_maxabs(const double* reg, int64_t fs, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_maxabs, dim3(reduce_blocks(fs)), dim3(kRedThreads), 0, s, reg, fs, part);
  return hipGetLastError();
}
hipError_t launch_l2diff(const double* a, const double* b, int64_t fs, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_l2diff, dim3(reduce_blocks(fs)), dim3(kRedThreads), 0, s, a, b, fs, part);
  return hipGetLastError();
}
hipError_t launch_energy(const double* reg, const MeshArgs& m, int n, const double* wq, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_energy, dim3(reduce_blocks(m.fs)), dim3(kRedThreads), 0, s, reg, m, n, wq, part);
  return hipGetLastError();
}
hipError_t launch_cfl(const MeshArgs& m, double dxi_half, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_cfl, dim3(reduce_blocks(m.E)), dim3(kRedThreads), 0, s, m, dxi_half, part);
  return hipGetLastError();
}

// ---- the allocation-free reductions ---------------------------------------------------------
// Per-block partials of a reduction and its combined values live in this array, one copy per
// device (each device holds its own until the process ends). Host code never passes the array's
// name to a kernel: it resolves the device address once per device (scratch_address) and passes
// that. Every reduction launch writes every partial the combine then reads, and the combine
// writes every value the host then reads, so nothing has to start at zero; calls that overlap on
// one device race on the array.
static_assert(kScratchPartials >= kRedMaxBlocks * 8, "8 partials per block at the most blocks");
static __device__ __align__(256) double scratch[kScratchPartials + kScratchValues];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch[kMaxDevices] = {};
}  // namespace

hipError_t scratch_address(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[device]) {
    // the address is resolved with `device` current, as the caller's context was created on it
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return hipSuccess;
}

// One lane combines the partials in block order, the order dg3d_reduce sums them in on the
// host: NaN-sticky max of 4 fields, or sums of per_block values.
__global__ void k_reduce_combine(const double* part, int blocks, int kind, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (kind == 0) {
    for (int f = 0; f < 4; ++f) {
      double m = 0.0;
      for (int b = 0; b < blocks; ++b) {
        const double v = part[b * 4 + f];
        m = (isnan(v) || isnan(m)) ? __builtin_nan("") : (m > v ? m : v);
      }
      out[f] = m;
    }
  } else if (kind == 1) {
    double e = 0.0, r = 0.0;
    for (int b = 0; b < blocks; ++b) {
      e += part[b * 2];
      r += part[b * 2 + 1];
    }
    out[0] = e;
    out[1] = r;
  } else {
    for (int f = 0; f < 4; ++f) {
      double d = 0.0, r = 0.0;
      for (int b = 0; b < blocks; ++b) {
        d += part[(b * 4 + f) * 2];
        r += part[(b * 4 + f) * 2 + 1];
      }
      out[f] = d;
      out[4 + f] = r;
    }
  }
}

hipError_t launch_reduce_combine(const double* part, int blocks, int kind, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_combine, dim3(1), dim3(64), 0, s, part, blocks, kind, out);
  return hipGetLastError();
}
hipError_t launch_probe(const double* reg, const MeshArgs& m, int n, const int64_t* nodes, int count, double* out,
                        const int* step, int cap, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_probe, dim3((count + 63) / 64), dim3(64), 0, s, reg, m, n, nodes, count, out, step, cap);
  return hipGetLastError();
}

hipError_t launch_clock_set(double* clock, double t, hipStream_t s) {
  hipLaunchKernelGGL(k_clock_set, dim3(1), dim3(64), 0, s, clock, t);
  return hipGetLastError();
}

hipError_t launch_step_end(double* clock, double dt, int* probe_step, int cap, hipStream_t s) {
  hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, s, clock, dt, probe_step, cap);
  return hipGetLastError();
}

#endif  // DG3D_TU_ORDER

}  // namespace dg3d
