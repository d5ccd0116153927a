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
#include <type_traits>

#include "dg3d_device.hpp"

namespace dg3d {
namespace {

template <int NN>
struct Shape {
  static constexpr int NP = NN * NN * NN;
  static constexpr int NF2 = NN * NN;
  // elements per workgroup: whole elements, lanes padded to whole waves with < 10% idle
  static constexpr int EPB = (NN == 2) ? 32 : (NN == 3) ? 9 : (NN == 4) ? 4 : (NN == 5) ? 2
                           : (NN == 6) ? 2 : (NN == 7) ? 2 : 1;
  static constexpr int NT = ((EPB * NP + 63) / 64) * 64;
};

// Workgroups are dispatched round-robin over the 8 XCDs; renumber them so that XCD x
// processes one contiguous range of elements and its L2 sees the neighbour traces.
__device__ __forceinline__ int64_t remap_block(int64_t w, int64_t G, int on) {
  if (!on) return w;
  const int64_t X = 8;
  const int64_t per = G / X, rem = G % X, x = w % X, i = w / X;
  return x * per + (x < rem ? x : rem) + i;
}

// kernelFuncs.cpp:17-20: P/rho0 = (rho' + rho'^2 * (B/A) * 0.5 / rho0) * c^2 / rho0
template <typename CT>
__device__ __forceinline__ CT momentum_f(CT rho, CT c, CT ba, CT ir0) {
  return (rho + rho * rho * ba * CT(0.5) * ir0) * (c * c * ir0);
}
// kernelFuncs.cpp:26-30: rho0 u if B/A == 0 exactly, else u (rho' + rho0)
template <typename CT>
__device__ __forceinline__ CT mass_f(CT u, CT rho, CT r0, CT ba) {
  return (ba == CT(0)) ? u * r0 : u * (rho + r0);
}

// Upwind acoustic flux pair of one line end (Element.cpp:540-600) with z = rho0 c and the
// reference's coefficients simplified by z/rho0 = c (division-free).
template <typename CT>
__device__ __forceinline__ void flux_pair(int side, CT mom_s, CT mass_s, CT c1, CT r1, CT mom_o,
                                          CT mass_o, CT c2, CT r2, CT& F0, CT& F1) {
  const CT z1 = r1 * c1, z2 = r2 * c2;
  const CT iz = CT(1) / (z1 + z2);
  if (side == 0) {
    F0 = -iz * (z1 * r2 * mom_o + z2 * r1 * mom_s + z1 * c2 * mass_o - z2 * c1 * mass_s);
    F1 = -iz * (c2 * mass_o + c1 * mass_s + r2 * mom_o - r1 * mom_s);
  } else {
    F0 = iz * (z2 * r1 * mom_s + z1 * r2 * mom_o + z2 * c1 * mass_s - z1 * c2 * mass_o);
    F1 = iz * (c1 * mass_s + c2 * mass_o + r1 * mom_s - r2 * mom_o);
  }
}

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
__device__ __forceinline__ int neighbour(const MeshArgs& m, int e, const int* ec, int axis, int side) {
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
__device__ __forceinline__ int face_elem(const MeshArgs& m, const int* ec, int axis) {
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

template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void st_stream(T v, T* p) { __builtin_nontemporal_store(v, p); }

template <int NN, typename CT, int EPI, typename DUT>
__global__ void __launch_bounds__(Shape<NN>::NT) k_rhs(const RhsLaunch L) {
  using S = Shape<NN>;
  constexpr int NP = S::NP, NF2 = S::NF2, EPB = S::EPB, NT = S::NT;
  constexpr int NITEMS = EPB * 6 * NF2;            // face work items of the workgroup
  constexpr int R = (NITEMS + NT - 1) / NT;        // face items per lane
  __shared__ CT sF[EPB * 4 * NP];                  // P/rho0, m_x, m_y, m_z per node
  __shared__ CT sFl[EPB * 6 * NF2 * 2];            // flux pair per face node
  __shared__ CT sNbM[EPB * 6 * 4];                 // neighbour c, rho0, B/A, 1/rho0 per face
  __shared__ int sNbE[EPB * 6];                    // neighbour element or -1
  __shared__ int sFe[EPB * 3];                     // face-element index per axis
  __shared__ CT sA[NN * NN];
  __shared__ CT sL0[NN];
  __shared__ CT sLN[NN];

  const MeshArgs& m = L.mesh;
  const unsigned mask = L.mask;
  const int tid = threadIdx.x;
  const int e0 = (int)(remap_block(blockIdx.x, gridDim.x, L.xcd_remap) * EPB);
  const int E = (int)m.E;
  const int64_t fs = m.fs;

  const CT* opA;
  const CT* opL0;
  const CT* opLN;
  if constexpr (std::is_same<CT, double>::value) {
    opA = L.opd[0]; opL0 = L.opd[1]; opLN = L.opd[2];
  } else {
    opA = L.opf[0]; opL0 = L.opf[1]; opLN = L.opf[2];
  }
  if (tid < NN * NN) sA[tid] = opA[tid];
  if (tid < NN) {
    sL0[tid] = opL0[tid];
    sLN[tid] = opLN[tid];
  }

  // ---- phase 0: issue every global load of the stage before the first barrier ----------
  const int le = tid / NP;
  const int node = tid - le * NP;
  const int e = e0 + le;
  const bool valid = (le < EPB) && (e < E);
  const int es = valid ? e : 0;
  const int64_t gi = (int64_t)es * NP + node;

  double q[4], pre[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int f = 0; f < 4; ++f) q[f] = L.q_in[f * fs + gi];
  if (EPI == EPI_LSRK && L.a != 0.0) {
    const DUT* dup = static_cast<const DUT*>(L.du);
#pragma unroll
    for (int f = 0; f < 4; ++f) pre[f] = (double)ld_stream(dup + f * fs + gi);
  }
  if (EPI == EPI_RK33 && L.a != 0.0) {
#pragma unroll
    for (int f = 0; f < 4; ++f) pre[f] = ld_stream(L.old + f * fs + gi);
  }
  const double* mp = m.mat + (int64_t)es * kMatStride;
  const CT c = (CT)mp[MAT_C], r0 = (CT)mp[MAT_RHO0], ba = (CT)mp[MAT_BA], sg = (CT)mp[MAT_SIGMA];
  const CT ir0 = (CT)mp[MAT_IRHO0];

  // neighbour traces of this lane's face items (rho', u_axis), exterior items read the halo
  double nrho[R], nu[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    nrho[r] = 0.0;
    nu[r] = 0.0;
    const int f = tid + r * NT;
    const int fle = f / (6 * NF2);
    const int rr = f - fle * 6 * NF2;
    const int dir = rr / NF2;
    const int fn = rr - dir * NF2;
    const int axis = 2 - (dir >> 1);
    const int side = dir & 1;
    const int fe = e0 + fle;
    const bool live = (f < NITEMS) && (fe < E) && (((mask >> (2 * axis)) & 3u) != 0u);
    int ec[3];
    elem_coords(m, live ? (uint32_t)fe : 0u, ec);
    const int eo = live ? neighbour(m, fe, ec, axis, side) : -1;
    const int b = fn / NN, cc = fn - b * NN;
    const int on = face_node<NN>(axis, side ? 0 : NN - 1, b, cc);
    const int esrc = (eo >= 0) ? eo : 0;
    const double vr = L.q_in[(int64_t)esrc * NP + on];
    const double vu = L.q_in[(1 + axis) * fs + (int64_t)esrc * NP + on];
    if (eo >= 0) {
      nrho[r] = vr;
      nu[r] = vu;
    } else if (live) {
      const double* tr = L.ext.trace[dir];
      if (tr) {
        const int64_t ti = ((int64_t)face_elem(m, ec, axis) * NF2 + fn) * 3;
        nrho[r] = tr[ti];
        nu[r] = tr[ti + 1];
      }
    }
  }
  // per-element helper lanes: neighbour element and material of each face, face indices
  if (tid < EPB * 6) {
    const int hle = tid / 6, dir = tid - hle * 6;
    const int he = e0 + hle;
    const int axis = 2 - (dir >> 1);
    int nbe = -1;
    CT nm[4] = {CT(1), CT(1), CT(0), CT(1)};
    if (he < E) {
      int ec[3];
      elem_coords(m, (uint32_t)he, ec);
      nbe = neighbour(m, he, ec, axis, dir & 1);
      // the halo element has the adjacent element's material (Block.cpp:167-248)
      const double* nmp = m.mat + (int64_t)(nbe >= 0 ? nbe : he) * kMatStride;
      nm[0] = (CT)nmp[MAT_C];
      nm[1] = (CT)nmp[MAT_RHO0];
      nm[2] = (CT)nmp[MAT_BA];
      nm[3] = (CT)nmp[MAT_IRHO0];
      if (dir < 3) sFe[hle * 3 + (2 - dir)] = face_elem(m, ec, 2 - dir);
    }
    sNbE[tid] = nbe;
#pragma unroll
    for (int k = 0; k < 4; ++k) sNbM[tid * 4 + k] = nm[k];
  }
  if (valid) {
    const CT rho = (CT)q[0];
    CT* F = sF + le * 4 * NP;
    F[node] = momentum_f(rho, c, ba, ir0);
#pragma unroll
    for (int a = 0; a < 3; ++a) F[(1 + a) * NP + node] = mass_f((CT)q[1 + a], rho, r0, ba);
  }
  __syncthreads();

  // ---- phase 1: upwind flux pair of every face node -------------------------------------
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int f = tid + r * NT;
    if (f >= NITEMS) continue;
    const int fle = f / (6 * NF2);
    const int rr = f - fle * 6 * NF2;
    const int dir = rr / NF2;
    const int fn = rr - dir * NF2;
    const int axis = 2 - (dir >> 1);
    const int side = dir & 1;
    const int fe = e0 + fle;
    if (fe >= E || ((mask >> (2 * axis)) & 3u) == 0u) continue;
    const int b = fn / NN, cc = fn - b * NN;
    const int sn = face_node<NN>(axis, side ? NN - 1 : 0, b, cc);
    const CT* F = sF + fle * 4 * NP;
    const CT mom_s = F[sn];
    const CT mass_s = F[(1 + axis) * NP + sn];
    const double* mps = m.mat + (int64_t)fe * kMatStride;
    const CT cs = (CT)mps[MAT_C], r0s = (CT)mps[MAT_RHO0];
    const CT* nm = sNbM + (fle * 6 + dir) * 4;
    const CT co = nm[0], r0o = nm[1], bao = nm[2], ir0o = nm[3];
    double rho_o = nrho[r], u_o = nu[r];
    if (sNbE[fle * 6 + dir] < 0 && dir == L.ext.src_face) {
      const int fidx = sFe[fle * 3 + axis];
      const double t = L.ext.t + (L.ext.lead ? L.ext.lead[(int64_t)fidx * NF2 + fn] : 0.0);
      const double sv = source_value(L.ext, t);
      const double cd = mps[MAT_C], rd = mps[MAT_RHO0];
      rho_o += sv / (cd * cd);
      u_o += (side ? -1.0 : 1.0) * sv / (rd * cd);
    }
    const CT mom_o = momentum_f((CT)rho_o, co, bao, ir0o);
    const CT mass_o = mass_f((CT)u_o, (CT)rho_o, r0o, bao);
    CT F0, F1;
    flux_pair<CT>(side, mom_s, mass_s, cs, r0s, mom_o, mass_o, co, r0o, F0, F1);
    CT* Fl = sFl + ((fle * 6 + dir) * NF2 + fn) * 2;
    Fl[0] = F0;
    Fl[1] = F1;
  }
  __syncthreads();
  if (!valid) return;

  // ---- phase 2: line contractions, lifting, sponge, RK update ---------------------------
  const int idx[3] = {node / NF2, (node / NN) % NN, node % NN};
  int ecs[3];
  elem_coords(m, (uint32_t)e, ecs);
  const CT* F = sF + le * 4 * NP;
  const CT* Fl = sFl + le * 6 * NF2 * 2;
  CT drho = CT(0);
  CT du[3] = {CT(0), CT(0), CT(0)};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const unsigned am = (mask >> (2 * a)) & 3u;
    if (am == 0u) continue;
    const int ix = idx[a];
    const int ls = (a == 0) ? NF2 : (a == 1 ? NN : 1);
    const int base = node - ix * ls;
    const int fn = (a == 0) ? idx[1] * NN + idx[2] : (a == 1 ? idx[0] * NN + idx[2] : idx[0] * NN + idx[1]);
    const int dlo = (a == 0) ? 4 : (a == 1 ? 2 : 0);
    const CT* Fa = F + (1 + a) * NP;
    CT dp = CT(0), dm = CT(0);
#pragma unroll
    for (int mm = 0; mm < NN; ++mm) {
      const CT av = sA[ix * NN + mm];
      dp += av * F[base + mm * ls];
      dm += av * Fa[base + mm * ls];
    }
    const CT iw = (CT)m.inv_w[a][ecs[a]];
    const CT l0 = sL0[ix], ln = sLN[ix];
    const CT* flo = Fl + (dlo * NF2 + fn) * 2;
    const CT* fhi = Fl + ((dlo + 1) * NF2 + fn) * 2;
    if (am & 1u) du[a] = iw * (dp - (flo[0] * l0 + fhi[0] * ln) * ir0);
    if (am & 2u) drho += iw * (dm - r0 * (flo[1] * l0 + fhi[1] * ln));
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
#pragma unroll
    for (int f = 0; f < 4; ++f)
      if (touch[f]) L.out[f * fs + gi] += (double)rhs[f];
  } else if constexpr (EPI == EPI_LSRK) {
    DUT* dup = static_cast<DUT*>(L.du);
    if constexpr (std::is_same<DUT, double>::value) {
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const double dn = L.a * pre[f] + L.dt * (double)rhs[f];
        st_stream(dn, dup + f * fs + gi);
        st_stream(q[f] + L.b * dn, L.q_out + f * fs + gi);
      }
    } else {
      const float af = (float)L.a, dtf = (float)L.dt;
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        const float dn = af * (float)pre[f] + dtf * (float)rhs[f];
        st_stream(dn, dup + f * fs + gi);
        st_stream(q[f] + L.b * (double)dn, L.q_out + f * fs + gi);
      }
    }
  } else {
    // Grid.cpp:705, 765-766, 826-827: out = (rhs*(dt*beta) + in*beta) + old*alpha
    const double sb = L.dt * L.b;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      double v = (double)rhs[f] * sb + q[f] * L.b;
      if (L.a != 0.0) v += pre[f] * L.a;
      L.out[f * fs + gi] = v;
    }
  }
}

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
  const int64_t gi = (int64_t)e * NP + node;
  const int axis = L.axis;
  const int fin = (L.which == 0) ? 0 : 4 + axis;  // rho' or Q_axis
  if (valid) sV[le * NP + node] = L.in[fin * m.fs + gi];
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
      vo = L.in[fin * m.fs + (int64_t)eo * NP + on];
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
    L.out[(4 + axis) * m.fs + gi] = -x;
  } else {
    L.out[gi] += -L.diff[e] * x;
  }
}

__global__ void k_axpby(double* out, double a, const double* in, double b, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = out[i] * a + in[i] * b;
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

__global__ void __launch_bounds__(kRedThreads) k_maxabs(const double* reg, int64_t fs, int nfields, double* part) {
  __shared__ double sh[kRedThreads / 64];
  for (int f = 0; f < nfields; ++f) {
    double v = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < fs; i += (int64_t)gridDim.x * blockDim.x)
      v = nan_max(v, fabs(reg[f * fs + i]));
    v = block_reduce(v, nan_max, sh);
    if (threadIdx.x == 0) part[blockIdx.x * nfields + f] = v;
  }
}

__global__ void __launch_bounds__(kRedThreads) k_l2diff(const double* a, const double* b, int64_t fs, int nfields, double* part) {
  __shared__ double sh[kRedThreads / 64];
  auto add = [](double x, double y) { return x + y; };
  for (int f = 0; f < nfields; ++f) {
    double d = 0.0, r = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < fs; i += (int64_t)gridDim.x * blockDim.x) {
      const double x = a[f * fs + i], y = b[f * fs + i];
      d += (x - y) * (x - y);
      r += y * y;
    }
    d = block_reduce(d, add, sh);
    r = block_reduce(r, add, sh);
    if (threadIdx.x == 0) {
      part[(blockIdx.x * nfields + f) * 2 + 0] = d;
      part[(blockIdx.x * nfields + f) * 2 + 1] = r;
    }
  }
}

__global__ void __launch_bounds__(kRedThreads) k_energy(const double* reg, MeshArgs m, int n, const double* wq, double* part) {
  __shared__ double sh[kRedThreads / 64];
  auto add = [](double x, double y) { return x + y; };
  const int64_t np = (int64_t)n * n * n;
  double en = 0.0, ms = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m.fs; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = i / np;
    const int nd = (int)(i - e * np);
    const int a = nd / (n * n), b = (nd / n) % n, c = nd % n;
    int ec[3];
    elem_coords(m, (uint32_t)e, ec);
    const double jac = 0.125 / (m.inv_w[0][ec[0]] * m.inv_w[1][ec[1]] * m.inv_w[2][ec[2]]);
    const double w = wq[a] * wq[b] * wq[c] * jac;
    const double c0 = m.mat[e * kMatStride + MAT_C], r0 = m.mat[e * kMatStride + MAT_RHO0];
    const double r = reg[i];
    const double u = reg[m.fs + i], v = reg[2 * m.fs + i], x = reg[3 * m.fs + i];
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

__global__ void k_probe(const double* reg, MeshArgs m, int n, const int64_t* nodes, int count, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t g = nodes[i];
  const int64_t np = (int64_t)n * n * n;
  const int64_t e = g / np;
  const double c = m.mat[e * kMatStride + MAT_C], r0 = m.mat[e * kMatStride + MAT_RHO0];
  const double ba = m.mat[e * kMatStride + MAT_BA];
  const double r = reg[g];
  out[i] = c * c * (r + ba * 0.5 / r0 * r * r);
}

template <int NN>
hipError_t launch_rhs_n(const RhsLaunch& L, hipStream_t s) {
  using S = Shape<NN>;
  const int64_t blocks = (L.mesh.E + S::EPB - 1) / S::EPB;
  if (blocks <= 0) return hipSuccess;
  dim3 grid((unsigned)blocks), block(S::NT);
  if (L.epilogue == EPI_ACCUM) {
    if (L.mixed) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_rhs<NN, double, EPI_ACCUM, double>), grid, block, 0, s, L);
  } else if (L.epilogue == EPI_LSRK) {
    if (L.mixed) hipLaunchKernelGGL((k_rhs<NN, float, EPI_LSRK, float>), grid, block, 0, s, L);
    else hipLaunchKernelGGL((k_rhs<NN, double, EPI_LSRK, double>), grid, block, 0, s, L);
  } else {
    if (L.mixed) hipLaunchKernelGGL((k_rhs<NN, float, EPI_RK33, double>), grid, block, 0, s, L);
    else hipLaunchKernelGGL((k_rhs<NN, double, EPI_RK33, double>), grid, block, 0, s, L);
  }
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

}  // namespace

hipError_t launch_rhs(const RhsLaunch& L, hipStream_t s) {
  switch (L.n) {
    case 2: return launch_rhs_n<2>(L, s);
    case 3: return launch_rhs_n<3>(L, s);
    case 4: return launch_rhs_n<4>(L, s);
    case 5: return launch_rhs_n<5>(L, s);
    case 6: return launch_rhs_n<6>(L, s);
    case 7: return launch_rhs_n<7>(L, s);
    case 8: return launch_rhs_n<8>(L, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_ldg(const LdgLaunch& L, hipStream_t s) {
  switch (L.n) {
    case 2: return launch_ldg_n<2>(L, s);
    case 3: return launch_ldg_n<3>(L, s);
    case 4: return launch_ldg_n<4>(L, s);
    case 5: return launch_ldg_n<5>(L, s);
    case 6: return launch_ldg_n<6>(L, s);
    case 7: return launch_ldg_n<7>(L, s);
    case 8: return launch_ldg_n<8>(L, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_axpby(double* out, double a, const double* in, double b, int64_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_axpby, dim3((unsigned)blocks), dim3(256), 0, s, out, a, in, b, count);
  return hipGetLastError();
}

int reduce_blocks(int64_t count) {
  int64_t b = (count + kRedThreads - 1) / kRedThreads;
  if (b > kRedMaxBlocks) b = kRedMaxBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

hipError_t launch_maxabs(const double* reg, int64_t fs, int nfields, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_maxabs, dim3(reduce_blocks(fs)), dim3(kRedThreads), 0, s, reg, fs, nfields, part);
  return hipGetLastError();
}
hipError_t launch_l2diff(const double* a, const double* b, int64_t fs, int nfields, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_l2diff, dim3(reduce_blocks(fs)), dim3(kRedThreads), 0, s, a, b, fs, nfields, part);
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
hipError_t launch_probe(const double* reg, const MeshArgs& m, int n, const int64_t* nodes, int count, double* out,
                        hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_probe, dim3((count + 63) / 64), dim3(64), 0, s, reg, m, n, nodes, count, out);
  return hipGetLastError();
}

}  // namespace dg3d
