// oracle.cpp — CPU restatement of 3DDG's hot path. TEST INFRASTRUCTURE ONLY (the checker):
// loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
// product. Each function cites the reference lines it restates.
//
// Parity pins (tests/test_oracle_pins.py, tests/test_basis.py), each a closed form, an invariant,
// brute force or the reference's own code, never this file's formulas retyped:
//   LGL data and M, S: numpy.polynomial and closed forms;
//   per-line solve: the reference's own Matrix.h QR solve (oracle/_ref);
//   line operator: polynomial exactness through the element loops (B/A = 0);
//   kernel functions with B/A != 0 (kernelFuncs.cpp:13-30): exact derivatives of the quadratic
//     equation of state and mass flux for linear rho', u, and the B/A == 0 switch;
//   heterogeneous upwind flux (z1 != z2): a plane pulse crossing an impedance step reflects and
//     transmits with R = (Z2 - Z1)/(Z1 + Z2), T = 2 Z2/(Z1 + Z2) (three pairs, both orientations);
//   upwind flux at the zero-state halo: energy non-increase in an absorbing box;
//   boundary source: a zero column driven by s(t) on the -z face carries s(t - z/c), with the
//     error falling spectrally with the order;
//   sponge sigma: tables by brute force; SSP-RK3 / LSRK54: orders 3 and 4 on y' = -sigma y;
//   LDG (QFromDRho / DRhoFromQ): exactness for a cubic, heat-kernel decay of a Fourier mode;
//   forwardEuler: order 1, one step (1 - sigma dt) exactly, equal to RK33's first stage, LDG
//     on the x axis only;
//   literal (FP32) mode: agreement with FP64 at FP32 level. Its bitwise FP32 rounding pattern
//     is parity unpinned: it would take a run of the whole reference (MPI, HDF5), which this
//     checker does not build.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int NF = 7;  // rho', u, v, w, Qx, Qy, Qz (Node.h:77-83 minus coords)

// LSRK(5,4) 2N-storage coefficients, Carpenter & Kennedy (1994), NASA TM-109112.
const double LS_A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                        -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double LS_B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                        1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                        2277821191437.0 / 14882151754819.0};
const double LS_C[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                        2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

struct Mat {  // per-element material (Element.h:146-151)
  double c0 = 1500, rho0 = 1000, ba = 0, sigma = 0, diff = 0;
};

// Gaussian elimination with partial pivoting; a is n*n (destroyed), b is n (solution).
void solve_dense(int n, double* a, double* b) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[i * n + k]) > std::fabs(a[p * n + k])) p = i;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[p * n + j]);
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < n; ++i) {
      double f = a[i * n + k] / a[k * n + k];
      for (int j = k; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
      b[i] -= f * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < n; ++j) s -= a[k * n + j] * b[j];
    b[k] = s / a[k * n + k];
  }
}

// (M q w) x = (S q) f - flux with the reference's operation order (Element.cpp:602,
// Matrix.h:590-622): scalar products first, then the matrix-vector product, then the solve.
void line_solve(int n, const double* M, const double* S, double q, double w, const double* f,
                const double* flux, double* x) {
  double a[81];
  for (int i = 0; i < n * n; ++i) a[i] = (q * M[i]) * w;
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += (q * S[i * n + k]) * f[k];
    x[i] = s - flux[i];
  }
  solve_dense(n, a, x);
}

}  // namespace

struct oracle {
  int N = 0, n = 0, np = 0;
  int nE[3] = {0, 0, 0};
  int H[3] = {0, 0, 0};  // nE + 2
  int periodic[3] = {0, 0, 0};
  int literal = 0;
  std::vector<double> M, S, xi, wq;
  double corner[3] = {0, 0, 0};
  std::vector<double> width[3];   // per interior element along each axis
  std::vector<double> ecorner[3]; // lower corner per halo-inclusive index (size H)
  std::vector<double> ewidth[3];  // width per halo-inclusive index (halo = nearest interior)
  std::vector<Mat> mat;           // per halo-inclusive element
  std::vector<std::vector<double>> regs;  // [elem][field][node]
  // source
  int src_face = -1;
  double src_amp = 0, src_f = 0, src_phase = 0, src_t0 = 0, src_tau = 1;
  std::vector<double> src_lead;
  double time = 0.0;

  size_t eidx(int i, int j, int k) const {  // i,j,k in -1..nE
    return ((size_t)(i + 1) * H[1] + (j + 1)) * H[2] + (k + 1);
  }
  size_t nelem() const { return (size_t)H[0] * H[1] * H[2]; }
  double* node(int reg, size_t e, int f) { return &regs[reg][(e * NF + f) * np]; }
  const double* node(int reg, size_t e, int f) const { return &regs[reg][(e * NF + f) * np]; }
  bool is_halo(int i, int j, int k) const {
    return i < 0 || j < 0 || k < 0 || i >= nE[0] || j >= nE[1] || k >= nE[2];
  }
  // value as stored by the policy (FP32 state rounds every stored value)
  double st(double v) const { return literal ? (double)(float)v : v; }
};

namespace {

// Scalar policy: the reference evaluates flux constants and node scalars in float.
template <typename T>
struct Pol {
  // kernelFuncs.cpp:17-20 (scalar momentum = P/rho0)
  static T momentum(T u, T rho, T c, T r0, T ba) {
    (void)u;
    return (rho + rho * rho * ba * T(0.5) * (T(1.0) / r0)) * (c * c / r0);
  }
  // kernelFuncs.cpp:26-30 (scalar mass flux), exact switch at B/A == 0
  static T mass(T u, T rho, T c, T r0, T ba) {
    (void)c;
    if (ba == T(0.0)) return u * r0;
    return u * (rho + r0);
  }
};

// Matrix overloads (kernelFuncs.cpp:13-16, 21-25) act on FP64 line vectors with the scalar
// constants rounded to T.
template <typename T>
void momentum_line(int n, const double* rho, T c, T r0, T ba, double* out) {
  const double k1 = (double)ba, k2 = 0.5, k3 = (double)(T(1.0) / r0), k4 = (double)(c * c / r0);
  for (int i = 0; i < n; ++i) {
    double sq = rho[i] * rho[i];
    out[i] = (rho[i] + ((sq * k1) * k2) * k3) * k4;
  }
}
template <typename T>
void mass_line(int n, const double* u, const double* rho, T r0, T ba, double* out) {
  for (int i = 0; i < n; ++i) {
    if (ba == T(0.0)) out[i] = u[i] * (double)r0;
    else out[i] = u[i] * (rho[i] + (double)r0);
  }
}

// Face flux pair (mode 0 momentum, mode 1 mass) of one line end, Element.cpp:540-600.
// side 0: face with neighbors[prev] (node 0 of self, node N of other);
// side 1: face with neighbors[next] (node N of self, node 0 of other).
template <typename T>
void face_flux(int side, int mode, T us, T rs, const Mat& ms, T uo, T ro, const Mat& mo, double* F) {
  const T cs = (T)ms.c0, r0s = (T)ms.rho0, bas = (T)ms.ba;
  const T co = (T)mo.c0, r0o = (T)mo.rho0, bao = (T)mo.ba;
  const T z1 = r0s * cs, z2 = r0o * co, r1 = r0s, r2 = r0o;
  T c1, c2, c3, c4;
  // kernel functions of self/other: mode 0 uses momentum first, mass second; mode 1 swaps
  auto kf = [&](int which, bool other) -> T {
    const bool mom = (mode == 0) == (which == 0);
    if (other) return mom ? Pol<T>::momentum(uo, ro, co, r0o, bao) : Pol<T>::mass(uo, ro, co, r0o, bao);
    return mom ? Pol<T>::momentum(us, rs, cs, r0s, bas) : Pol<T>::mass(us, rs, cs, r0s, bas);
  };
  double acc;
  if (side == 0) {
    if (mode == 0) {
      c1 = z1 * r2 / (z1 + z2);
      c2 = z2 * r1 / (z1 + z2);
      c3 = z1 * z2 / r2 / (z1 + z2);
      c4 = T(-1.0) * z1 * z2 / r1 / (z1 + z2);
    } else {
      c1 = z2 / r2 / (z1 + z2);
      c2 = z1 / r1 / (z1 + z2);
      c3 = r2 / (z1 + z2);
      c4 = T(-1.0) * r1 / (z1 + z2);
    }
    acc = (double)(T)(kf(0, true) * c1);
    acc += (double)(T)(kf(0, false) * c2);
    acc += (double)(T)(kf(1, true) * c3);
    acc += (double)(T)(kf(1, false) * c4);
    acc *= -1.0;
  } else {
    if (mode == 0) {
      c1 = z2 * r1 / (z1 + z2);
      c2 = z1 * r2 / (z1 + z2);
      c3 = z1 * z2 / r1 / (z1 + z2);
      c4 = T(-1.0) * z1 * z2 / r2 / (z1 + z2);
    } else {
      c1 = z1 / r1 / (z1 + z2);
      c2 = z2 / r2 / (z1 + z2);
      c3 = r1 / (z1 + z2);
      c4 = T(-1.0) * r2 / (z1 + z2);
    }
    acc = (double)(T)(kf(0, false) * c1);
    acc += (double)(T)(kf(0, true) * c2);
    acc += (double)(T)(kf(1, false) * c3);
    acc += (double)(T)(kf(1, true) * c4);
  }
  *F = acc;
}

inline int dir_axis(int dir) { return 2 - dir / 2; }  // 0,1 -> z; 2,3 -> y; 4,5 -> x

// source value s(t) (no reference counterpart)
double source_value(const oracle* o, double t) {
  const double x = (t - o->src_t0) / o->src_tau;
  return o->src_amp * std::sin(2.0 * M_PI * o->src_f * t + o->src_phase) * std::exp(-x * x);
}

// State of the other element's face node as read through neighbors[dir] (Element.cpp:557-600)
// plus the exterior source where the other element is the halo on the source face.
void other_trace(const oracle* o, int reg, int dir, int ei, int ej, int ek, size_t eo, int onode,
                 int fnode, int axis, double* rho, double* u, double* Q) {
  *rho = o->node(reg, eo, 0)[onode];
  *u = o->node(reg, eo, 1 + axis)[onode];
  *Q = o->node(reg, eo, 4 + axis)[onode];
  if (o->src_face == dir) {
    int oi = ei, oj = ej, ok = ek;
    if (axis == 0) oi += (dir % 2) ? 1 : -1;
    if (axis == 1) oj += (dir % 2) ? 1 : -1;
    if (axis == 2) ok += (dir % 2) ? 1 : -1;
    if (o->is_halo(oi, oj, ok)) {
      const Mat& m = o->mat[eo];
      size_t fe;
      if (axis == 0) fe = (size_t)ej * o->nE[2] + ek;
      else if (axis == 1) fe = (size_t)ei * o->nE[2] + ek;
      else fe = (size_t)ei * o->nE[1] + ej;
      double lead = o->src_lead.empty() ? 0.0 : o->src_lead[fe * o->n * o->n + fnode];
      double s = source_value(o, o->time + lead);
      double sign = (dir % 2) ? -1.0 : 1.0;
      *rho += s / (m.c0 * m.c0);
      *u += sign * s / (m.rho0 * m.c0);
    }
  }
}

// Element::OperateUFromV for one interior element (Element.cpp:475-620).
template <typename T>
void operate_ufromv(oracle* o, int axis, int mode, int in, int out, int ei, int ej, int ek) {
  const int n = o->n, N = o->N;
  const size_t es = o->eidx(ei, ej, ek);
  const Mat& ms = o->mat[es];
  const int prevDir = (axis == 0) ? 4 : (axis == 1 ? 2 : 0);
  const int nextDir = prevDir + 1;
  int pi = ei, pj = ej, pk = ek, ni = ei, nj = ej, nk = ek;
  if (axis == 0) { pi--; ni++; }
  if (axis == 1) { pj--; nj++; }
  if (axis == 2) { pk--; nk++; }
  const size_t ep = o->eidx(pi, pj, pk), en = o->eidx(ni, nj, nk);
  const Mat& mp = o->mat[ep];
  const Mat& mn = o->mat[en];
  const T q = (mode == 0) ? (T)ms.rho0 : T(1.0F) / (T)ms.rho0;  // Element.cpp:504-505
  const double w = (double)(T)o->ewidth[axis][axis == 0 ? ei + 1 : (axis == 1 ? ej + 1 : ek + 1)];
  const double* rho = o->node(in, es, 0);
  const double* vel = o->node(in, es, 1 + axis);
  double* dst = o->node(out, es, mode == 0 ? 1 + axis : 0);
  double frho[9], fu[9], f[9], flux[9], x[9];
  const int s_ax[3] = {n * n, n, 1};
  for (int b = 0; b < n; ++b) {
    for (int c = 0; c < n; ++c) {
      int base;  // node index with the axis index zero
      if (axis == 0) base = b * n + c;
      else if (axis == 1) base = b * n * n + c;
      else base = b * n * n + c * n;
      const int fnode = b * n + c;
      for (int m = 0; m < n; ++m) {
        frho[m] = rho[base + m * s_ax[axis]];
        fu[m] = vel[base + m * s_ax[axis]];
      }
      for (int m = 0; m < n; ++m) flux[m] = 0.0;
      double ro, uo, Qo;
      // previous face: self node 0, other node N (Element.cpp:540-561)
      other_trace(o, in, prevDir, ei, ej, ek, ep, base + N * s_ax[axis], fnode, axis, &ro, &uo, &Qo);
      face_flux<T>(0, mode, (T)fu[0], (T)frho[0], ms, (T)uo, (T)ro, mp, &flux[0]);
      // next face: self node N, other node 0 (Element.cpp:571-600)
      other_trace(o, in, nextDir, ei, ej, ek, en, base, fnode, axis, &ro, &uo, &Qo);
      face_flux<T>(1, mode, (T)fu[N], (T)frho[N], ms, (T)uo, (T)ro, mn, &flux[N]);
      if (mode == 0) momentum_line<T>(n, frho, (T)ms.c0, (T)ms.rho0, (T)ms.ba, f);
      else mass_line<T>(n, fu, frho, (T)ms.rho0, (T)ms.ba, f);
      line_solve(n, o->M.data(), o->S.data(), (double)q, w, f, flux, x);
      for (int m = 0; m < n; ++m) {  // Element.cpp:605-616: (float)result added to state
        double& d = dst[base + m * s_ax[axis]];
        d = o->st(d + (double)(T)x[m]);
      }
    }
  }
}

// Element::OperateQFromDRho (Element.cpp:623-690): Q_axis of `out` = -(M w)^{-1}(S rho + F).
template <typename T>
void operate_qfromdrho(oracle* o, int axis, int in, int out, int ei, int ej, int ek) {
  const int n = o->n, N = o->N;
  const size_t es = o->eidx(ei, ej, ek);
  const int prevDir = (axis == 0) ? 4 : (axis == 1 ? 2 : 0);
  int pi = ei, pj = ej, pk = ek, ni = ei, nj = ej, nk = ek;
  if (axis == 0) { pi--; ni++; }
  if (axis == 1) { pj--; nj++; }
  if (axis == 2) { pk--; nk++; }
  const size_t ep = o->eidx(pi, pj, pk), en = o->eidx(ni, nj, nk);
  const double w = (double)(T)o->ewidth[axis][axis == 0 ? ei + 1 : (axis == 1 ? ej + 1 : ek + 1)];
  const double* rho = o->node(in, es, 0);
  double* dst = o->node(out, es, 4 + axis);
  const int s_ax[3] = {n * n, n, 1};
  double f[9], flux[9], x[9];
  for (int b = 0; b < n; ++b) {
    for (int c = 0; c < n; ++c) {
      int base;
      if (axis == 0) base = b * n + c;
      else if (axis == 1) base = b * n * n + c;
      else base = b * n * n + c * n;
      const int fnode = b * n + c;
      for (int m = 0; m < n; ++m) { f[m] = rho[base + m * s_ax[axis]]; flux[m] = 0.0; }
      double ro, uo, Qo;
      other_trace(o, in, prevDir, ei, ej, ek, ep, base + N * s_ax[axis], fnode, axis, &ro, &uo, &Qo);
      flux[0] = ro;
      flux[0] += f[0];
      flux[0] *= 0.5;
      other_trace(o, in, prevDir + 1, ei, ej, ek, en, base, fnode, axis, &ro, &uo, &Qo);
      flux[N] = f[N];
      flux[N] += ro;
      flux[N] *= -0.5;
      // (M w) x = S rho + Flux  (note +Flux, Element.cpp:679)
      for (int m = 0; m < n; ++m) flux[m] = -flux[m];
      line_solve(n, o->M.data(), o->S.data(), 1.0, w, f, flux, x);
      for (int m = 0; m < n; ++m) dst[base + m * s_ax[axis]] = o->st(-1.0 * (double)(T)x[m]);
    }
  }
}

// Element::OperateDRhoFromQ (Element.cpp:691-758): rho' of `out` += (M w)^{-1}(-D (S Q + F)).
template <typename T>
void operate_drhofromq(oracle* o, int axis, int in, int out, int ei, int ej, int ek) {
  const int n = o->n, N = o->N;
  const size_t es = o->eidx(ei, ej, ek);
  const int prevDir = (axis == 0) ? 4 : (axis == 1 ? 2 : 0);
  int pi = ei, pj = ej, pk = ek, ni = ei, nj = ej, nk = ek;
  if (axis == 0) { pi--; ni++; }
  if (axis == 1) { pj--; nj++; }
  if (axis == 2) { pk--; nk++; }
  const size_t ep = o->eidx(pi, pj, pk), en = o->eidx(ni, nj, nk);
  const double w = (double)(T)o->ewidth[axis][axis == 0 ? ei + 1 : (axis == 1 ? ej + 1 : ek + 1)];
  const double D = (double)(T)o->mat[es].diff;
  const double* Qin = o->node(in, es, 4 + axis);
  double* dst = o->node(out, es, 0);
  const int s_ax[3] = {n * n, n, 1};
  double q[9], flux[9], x[9], sq[9];
  for (int b = 0; b < n; ++b) {
    for (int c = 0; c < n; ++c) {
      int base;
      if (axis == 0) base = b * n + c;
      else if (axis == 1) base = b * n * n + c;
      else base = b * n * n + c * n;
      const int fnode = b * n + c;
      for (int m = 0; m < n; ++m) { q[m] = Qin[base + m * s_ax[axis]]; flux[m] = 0.0; }
      double ro, uo, Qo;
      other_trace(o, in, prevDir, ei, ej, ek, ep, base + N * s_ax[axis], fnode, axis, &ro, &uo, &Qo);
      flux[0] = Qo;
      flux[0] += q[0];
      flux[0] *= 0.5;
      other_trace(o, in, prevDir + 1, ei, ej, ek, en, base, fnode, axis, &ro, &uo, &Qo);
      flux[N] = q[N];
      flux[N] += Qo;
      flux[N] *= -0.5;
      // rhs = ((S Q) + Flux) * (-1) * D, solve (M w) x = rhs (Element.cpp:747)
      for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += o->S[i * n + k] * q[k];
        sq[i] = ((s + flux[i]) * -1.0) * D;
      }
      double a[81];
      for (int i = 0; i < n * n; ++i) a[i] = o->M[i] * w;
      solve_dense(n, a, sq);
      for (int m = 0; m < n; ++m) {
        double& d = dst[base + m * s_ax[axis]];
        d = o->st(d + (double)(T)sq[m]);
      }
      (void)x;
    }
  }
}

template <typename T>
void operate_sponge(oracle* o, int in, int out, int ei, int ej, int ek) {
  // Element.cpp:760-781: u_a += -1 * old.u_a * sigma; rho' += -1 * old.rho' * sigma
  const size_t es = o->eidx(ei, ej, ek);
  const T sg = (T)o->mat[es].sigma;
  for (int nd = 0; nd < o->np; ++nd) {
    for (int a = 0; a < 3; ++a) {
      double& d = o->node(out, es, 1 + a)[nd];
      d = o->st(d + (double)(T(-1.0F) * (T)o->node(in, es, 1 + a)[nd] * sg));
    }
    double& d = o->node(out, es, 0)[nd];
    d = o->st(d + (double)(T(-1.0F) * (T)o->node(in, es, 0)[nd] * sg));
  }
}

template <typename T>
void operate_all(oracle* o, int op, int axis, int mode, int in, int out) {
  // Block::Operate* loops over interior elements i,j,k in [1, nE] (Block.cpp:343-403).
  const int n0 = o->nE[0], n1 = o->nE[1], n2 = o->nE[2];
  const long total = (long)n0 * n1 * n2;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < total; ++t) {
    const int ei = (int)(t / ((long)n1 * n2));
    const int ej = (int)((t / n2) % n1);
    const int ek = (int)(t % n2);
    if (op == 0) operate_ufromv<T>(o, axis, mode, in, out, ei, ej, ek);
    else if (op == 1) operate_sponge<T>(o, in, out, ei, ej, ek);
    else if (op == 2) operate_qfromdrho<T>(o, axis, in, out, ei, ej, ek);
    else operate_drhofromq<T>(o, axis, in, out, ei, ej, ek);
  }
}

}  // namespace

extern "C" {

void oracle_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

void oracle_line_solve(int n, const double* mass, const double* stiff, double q, double w,
                       const double* f, const double* flux, double* x) {
  line_solve(n, mass, stiff, q, w, f, flux, x);
}

oracle* oracle_create(const oracle_desc* d) {
  if (!d || d->order < 1 || d->order > 8) return nullptr;
  oracle* o = new oracle();
  o->N = d->order;
  o->n = d->order + 1;
  o->np = o->n * o->n * o->n;
  for (int a = 0; a < 3; ++a) {
    o->nE[a] = d->num_elements[a];
    o->H[a] = d->num_elements[a] + 2;
    o->periodic[a] = d->periodic[a];
    o->corner[a] = d->corner[a];
    if (o->nE[a] < 1) { delete o; return nullptr; }
  }
  o->literal = d->literal;
  const int n = o->n;
  o->M.assign(d->mass, d->mass + n * n);
  o->S.assign(d->stiff, d->stiff + n * n);
  o->xi.assign(d->lgl_nodes, d->lgl_nodes + n);
  o->wq.assign(d->lgl_weights, d->lgl_weights + n);
  for (int a = 0; a < 3; ++a) {
    o->width[a].assign(d->width[a], d->width[a] + o->nE[a]);
    o->ewidth[a].resize(o->H[a]);
    o->ecorner[a].resize(o->H[a]);
    for (int h = 0; h < o->H[a]; ++h) {
      int idx = std::min(std::max(h - 1, 0), o->nE[a] - 1);
      o->ewidth[a][h] = o->width[a][idx];
    }
    // Element corners: STARTING_CORNER + sum of widths (Block.cpp:172,190,208 for one width)
    double acc = o->corner[a];
    for (int h = 1; h <= o->nE[a]; ++h) {
      o->ecorner[a][h] = acc;
      acc += o->width[a][h - 1];
    }
    o->ecorner[a][o->H[a] - 1] = acc;
    o->ecorner[a][0] = o->corner[a] - o->width[a][0];
  }
  o->mat.assign(o->nelem(), Mat());
  o->regs.assign(d->num_registers, std::vector<double>(o->nelem() * NF * o->np, 0.0));
  return o;
}

void oracle_destroy(oracle* o) { delete o; }

void oracle_set_material(oracle* o, const double* c0, const double* rho0, const double* ba,
                         const double* sigma, const double* diff) {
  for (int i = -1; i <= o->nE[0]; ++i)
    for (int j = -1; j <= o->nE[1]; ++j)
      for (int k = -1; k <= o->nE[2]; ++k) {
        const int ci = std::min(std::max(i, 0), o->nE[0] - 1);
        const int cj = std::min(std::max(j, 0), o->nE[1] - 1);
        const int ck = std::min(std::max(k, 0), o->nE[2] - 1);
        const size_t src = ((size_t)ci * o->nE[1] + cj) * o->nE[2] + ck;
        Mat& m = o->mat[o->eidx(i, j, k)];
        if (c0) m.c0 = c0[src];
        if (rho0) m.rho0 = rho0[src];
        if (ba) m.ba = ba[src];
        if (diff) m.diff = diff[src];
        // halo sponge: only interior elements apply the sponge (Block.cpp:389-402)
        if (sigma) m.sigma = o->is_halo(i, j, k) ? 0.0 : sigma[src];
      }
}

void oracle_node_coord(const oracle* o, int ei, int ej, int ek, int node, double* xyz) {
  const int n = o->n;
  const int idx[3] = {node / (n * n), (node / n) % n, node % n};
  const int e[3] = {ei, ej, ek};
  for (int a = 0; a < 3; ++a) {
    const double w = o->ewidth[a][e[a] + 1];
    const double c = o->ecorner[a][e[a] + 1];
    if (o->literal) {  // Element.cpp:334-336 in float
      const float cf = (float)c, wf = (float)w;
      xyz[a] = (double)(wf * ((float)o->xi[idx[a]] / 2.F + .5F) + cf);
    } else {
      xyz[a] = w * (o->xi[idx[a]] / 2.0 + 0.5) + c;
    }
  }
}

void oracle_fill_gaussian(oracle* o, int reg, const int* gauss, double amp_rho, const double* amp_vel,
                          double ic_sigma, const double* center) {
  // Node::InitializeRho / InitializeVelocities (Node.cpp:76-101), every node incl. halos.
  const bool any = gauss[0] || gauss[1] || gauss[2];
  for (int i = -1; i <= o->nE[0]; ++i)
    for (int j = -1; j <= o->nE[1]; ++j)
      for (int k = -1; k <= o->nE[2]; ++k) {
        const size_t e = o->eidx(i, j, k);
        for (int nd = 0; nd < o->np; ++nd) {
          double xyz[3];
          oracle_node_coord(o, i, j, k, nd, xyz);
          for (int f = 0; f < NF; ++f) o->node(reg, e, f)[nd] = 0.0;
          if (!any) continue;
          if (o->literal) {
            float dist = 0;
            for (int a = 0; a < 3; ++a)
              if (gauss[a]) dist += ((float)xyz[a] - (float)center[a]) * ((float)xyz[a] - (float)center[a]);
            dist = (float)std::sqrt((double)dist);
            const double g = std::exp((double)(-(dist) * (dist) / (1.F / (float)ic_sigma)));
            o->node(reg, e, 0)[nd] = (double)(float)((double)(float)amp_rho * g);
            for (int a = 0; a < 3; ++a) o->node(reg, e, 1 + a)[nd] = (double)(float)((double)(float)amp_vel[a] * g);
          } else {
            double d2 = 0;
            for (int a = 0; a < 3; ++a)
              if (gauss[a]) d2 += (xyz[a] - center[a]) * (xyz[a] - center[a]);
            const double g = std::exp(-d2 * ic_sigma);
            o->node(reg, e, 0)[nd] = amp_rho * g;
            for (int a = 0; a < 3; ++a) o->node(reg, e, 1 + a)[nd] = amp_vel[a] * g;
          }
        }
      }
}

void oracle_set_state(oracle* o, int reg, const double* host, int nfields) {
  const size_t E = (size_t)o->nE[0] * o->nE[1] * o->nE[2];
  for (int f = 0; f < nfields; ++f)
    for (size_t e = 0; e < E; ++e) {
      const int i = (int)(e / ((size_t)o->nE[1] * o->nE[2]));
      const int j = (int)((e / o->nE[2]) % o->nE[1]);
      const int k = (int)(e % o->nE[2]);
      double* d = o->node(reg, o->eidx(i, j, k), f);
      for (int nd = 0; nd < o->np; ++nd) d[nd] = o->st(host[(f * E + e) * o->np + nd]);
    }
}

void oracle_get_state(const oracle* o, int reg, double* host, int nfields) {
  const size_t E = (size_t)o->nE[0] * o->nE[1] * o->nE[2];
  for (int f = 0; f < nfields; ++f)
    for (size_t e = 0; e < E; ++e) {
      const int i = (int)(e / ((size_t)o->nE[1] * o->nE[2]));
      const int j = (int)((e / o->nE[2]) % o->nE[1]);
      const int k = (int)(e % o->nE[2]);
      const double* s = o->node(reg, o->eidx(i, j, k), f);
      for (int nd = 0; nd < o->np; ++nd) host[(f * E + e) * o->np + nd] = s[nd];
    }
}

// Face element/node enumeration shared with dg3d_set_halo.
static void face_geom(const oracle* o, int face, int* axis, int* nb, int* nc) {
  *axis = dir_axis(face);
  const int a = *axis;
  const int b = (a == 0) ? 1 : 0;
  const int c = (a == 2) ? 1 : 2;
  *nb = o->nE[b];
  *nc = o->nE[c];
}

static void halo_access(oracle* o, int reg, int face, double* trace, bool write) {
  int axis, nb, nc;
  face_geom(o, face, &axis, &nb, &nc);
  const int n = o->n;
  const bool high = face % 2;
  for (int fb = 0; fb < nb; ++fb)
    for (int fc = 0; fc < nc; ++fc) {
      int e3[3];
      e3[axis] = high ? o->nE[axis] : -1;
      e3[axis == 0 ? 1 : 0] = fb;
      e3[axis == 2 ? 1 : 2] = fc;
      const size_t e = o->eidx(e3[0], e3[1], e3[2]);
      for (int b = 0; b < n; ++b)
        for (int c = 0; c < n; ++c) {
          int idx[3];
          idx[axis] = high ? 0 : n - 1;  // the halo's node layer that touches the block
          idx[axis == 0 ? 1 : 0] = b;
          idx[axis == 2 ? 1 : 2] = c;
          const int nd = (idx[0] * n + idx[1]) * n + idx[2];
          double* t = &trace[(((size_t)fb * nc + fc) * n * n + b * n + c) * 3];
          if (write) {
            o->node(reg, e, 0)[nd] = o->st(trace ? t[0] : 0.0);
            o->node(reg, e, 1 + axis)[nd] = o->st(trace ? t[1] : 0.0);
            o->node(reg, e, 4 + axis)[nd] = o->st(trace ? t[2] : 0.0);
          } else {
            t[0] = o->node(reg, e, 0)[nd];
            t[1] = o->node(reg, e, 1 + axis)[nd];
            t[2] = o->node(reg, e, 4 + axis)[nd];
          }
        }
    }
}

void oracle_set_halo(oracle* o, int reg, int face, const double* trace) {
  if (!trace) {
    int axis, nb, nc;
    face_geom(o, face, &axis, &nb, &nc);
    std::vector<double> z((size_t)nb * nc * o->n * o->n * 3, 0.0);
    halo_access(o, reg, face, z.data(), true);
    return;
  }
  halo_access(o, reg, face, const_cast<double*>(trace), true);
}

void oracle_get_halo(const oracle* o, int reg, int face, double* trace) {
  halo_access(const_cast<oracle*>(o), reg, face, trace, false);
}

void oracle_zero(oracle* o, int reg) { std::fill(o->regs[reg].begin(), o->regs[reg].end(), 0.0); }

void oracle_copy(oracle* o, int dst, int src) { o->regs[dst] = o->regs[src]; }

void oracle_axpby(oracle* o, int out, double self_scalar, int in, double other_scalar) {
  // Element::selfWeightedAdd (Element.cpp:288-307) over interior elements, all 7 fields.
  for (int i = 0; i < o->nE[0]; ++i)
    for (int j = 0; j < o->nE[1]; ++j)
      for (int k = 0; k < o->nE[2]; ++k) {
        const size_t e = o->eidx(i, j, k);
        for (int f = 0; f < NF; ++f) {
          double* d = o->node(out, e, f);
          const double* s = o->node(in, e, f);
          for (int nd = 0; nd < o->np; ++nd) {
            if (o->literal) {
              const float a = (float)self_scalar, b = (float)other_scalar;
              d[nd] = (double)((float)d[nd] * a + (float)s[nd] * b);
            } else {
              d[nd] = d[nd] * self_scalar + s[nd] * other_scalar;
            }
          }
        }
      }
}

void oracle_operate(oracle* o, int op, int axis, int mode, int in_reg, int out_reg) {
  if (o->literal) operate_all<float>(o, op, axis, mode, in_reg, out_reg);
  else operate_all<double>(o, op, axis, mode, in_reg, out_reg);
}

void oracle_communicate(oracle* o, int reg) {
  // Grid::periodicSelf (Grid.cpp:385-440): halo elements copy state and material from the
  // opposite interior element on periodic axes; the first matching rule (i, then j, then k)
  // decides the source element.
  for (int i = -1; i <= o->nE[0]; ++i)
    for (int j = -1; j <= o->nE[1]; ++j)
      for (int k = -1; k <= o->nE[2]; ++k) {
        int fi = i, fj = j, fk = k;
        bool send = false;
        if (i == -1) { fi = o->nE[0] - 1; send = o->periodic[0]; }
        else if (i == o->nE[0]) { fi = 0; send = o->periodic[0]; }
        else if (j == -1) { fj = o->nE[1] - 1; send = o->periodic[1]; }
        else if (j == o->nE[1]) { fj = 0; send = o->periodic[1]; }
        else if (k == -1) { fk = o->nE[2] - 1; send = o->periodic[2]; }
        else if (k == o->nE[2]) { fk = 0; send = o->periodic[2]; }
        if (!send) continue;
        const size_t d = o->eidx(i, j, k), s = o->eidx(fi, fj, fk);
        for (int f = 0; f < NF; ++f) std::memcpy(o->node(reg, d, f), o->node(reg, s, f), sizeof(double) * o->np);
        Mat& md = o->mat[d];
        const Mat& msrc = o->mat[s];
        md.c0 = msrc.c0; md.rho0 = msrc.rho0; md.ba = msrc.ba; md.diff = msrc.diff;
      }
}

void oracle_set_source(oracle* o, int face, double amplitude, double frequency, double phase,
                       double t0, double tau, const double* lead) {
  o->src_face = face;
  o->src_amp = amplitude;
  o->src_f = frequency;
  o->src_phase = phase;
  o->src_t0 = t0;
  o->src_tau = tau;
  o->src_lead.clear();
  if (lead && face >= 0) {
    int axis, nb, nc;
    face_geom(o, face, &axis, &nb, &nc);
    o->src_lead.assign(lead, lead + (size_t)nb * nc * o->n * o->n);
  }
}

void oracle_set_time(oracle* o, double t) { o->time = t; }

static void eval_rhs(oracle* o, int in, int out, int adv, int diff, int sponge) {
  // the L(in) accumulation order of Grid::RK33 (Grid.cpp:655-705)
  if (diff) {
    for (int a = 0; a < 3; ++a) oracle_operate(o, 2, a, 0, in, in);
    oracle_communicate(o, in);
    for (int a = 0; a < 3; ++a) oracle_operate(o, 3, a, 0, in, out);
  }
  if (adv)
    for (int a = 0; a < 3; ++a) {
      oracle_operate(o, 0, a, 0, in, out);
      oracle_operate(o, 0, a, 1, in, out);
    }
  if (sponge) oracle_operate(o, 1, 0, 0, in, out);
}

void oracle_rk33(oracle* o, double dt, int nw, int old, int k1, int k2, int adv, int diff, int sponge) {
  // Grid::RK33 single-block branch (Grid.cpp:642-835)
  const double t = o->time;
  oracle_zero(o, k1);
  oracle_zero(o, k2);
  oracle_zero(o, nw);
  const bool lit = o->literal;
  const double dtf = lit ? (double)(float)dt : dt;
  oracle_communicate(o, old);
  o->time = t;
  eval_rhs(o, old, k1, adv, diff, sponge);
  oracle_axpby(o, k1, dtf, old, 1.0);
  oracle_communicate(o, k1);
  o->time = t + dt;
  eval_rhs(o, k1, k2, adv, diff, sponge);
  if (lit) {
    oracle_axpby(o, k2, (double)((float)dt * .25F), k1, (double).25F);
    oracle_axpby(o, k2, 1.0, old, (double).75F);
  } else {
    oracle_axpby(o, k2, dt * 0.25, k1, 0.25);
    oracle_axpby(o, k2, 1.0, old, 0.75);
  }
  oracle_communicate(o, k2);
  o->time = t + 0.5 * dt;
  eval_rhs(o, k2, nw, adv, diff, sponge);
  if (lit) {
    oracle_axpby(o, nw, (double)((float)dt * 2.0F / 3.0F), k2, (double)(2.0F / 3.0F));
    oracle_axpby(o, nw, 1.0, old, (double)(1.0F / 3.0F));
  } else {
    oracle_axpby(o, nw, dt * (2.0 / 3.0), k2, 2.0 / 3.0);
    oracle_axpby(o, nw, 1.0, old, 1.0 / 3.0);
  }
  o->time = t + dt;
}

void oracle_rk33_blocks(oracle* const* blocks, int count, double dt, int steps) {
  // The reference's parallel model (driver.cpp:70-90, Grid.cpp:176-375): one block per MPI rank,
  // each rank single-threaded. Here one std::thread per block, each running `steps` of
  // Grid::RK33 on its block with registers 0 (state), 1, 2 (k1, k2) and 3 (new), then copying
  // new into 0 as the driver's pointer swap does. The blocks exchange no halos: the timing
  // leaves the reference's MPI exchange out.
  std::vector<std::thread> pool;
  pool.reserve((size_t)count);
  for (int b = 0; b < count; ++b)
    pool.emplace_back([=]() {
#ifdef _OPENMP
      omp_set_num_threads(1);  // this thread's own parallel regions run on itself only
#endif
      for (int s = 0; s < steps; ++s) {
        oracle_rk33(blocks[b], dt, 3, 0, 1, 2, 1, 0, 1);
        oracle_copy(blocks[b], 0, 3);
      }
    });
  for (auto& t : pool) t.join();
}

void oracle_forward_euler(oracle* o, double dt, int nw, int old) {
  // Grid::forwardEuler single-block branch (Grid.cpp:567-640): no evaluate flags; the LDG pair
  // runs for the x axis only (Grid.cpp:581,597), then advection for the three axes, the sponge
  // and new = L(old) dt + old.
  const double t = o->time;
  oracle_zero(o, nw);
  oracle_communicate(o, old);
  oracle_operate(o, 2, 0, 0, old, old);
  oracle_communicate(o, old);
  oracle_operate(o, 3, 0, 0, old, nw);
  for (int a = 0; a < 3; ++a) {
    oracle_operate(o, 0, a, 0, old, nw);
    oracle_operate(o, 0, a, 1, old, nw);
  }
  oracle_operate(o, 1, 0, 0, old, nw);
  oracle_axpby(o, nw, o->literal ? (double)(float)dt : dt, old, 1.0);
  o->time = t + dt;
}

void oracle_lsrk54(oracle* o, double dt, int q, int du, int rhs, int adv, int diff, int sponge) {
  // 2N-storage: du = A_s du + dt L(q, t + C_s dt); q = q + B_s du (interior only).
  const double t = o->time;
  for (int s = 0; s < 5; ++s) {
    oracle_zero(o, rhs);
    oracle_communicate(o, q);
    o->time = t + LS_C[s] * dt;
    eval_rhs(o, q, rhs, adv, diff, sponge);
    oracle_axpby(o, du, LS_A[s], rhs, dt);
    oracle_axpby(o, q, 1.0, du, LS_B[s]);
    if (s == 0) {
      // after the first stage q is a new state whose halo is zero (as a zeroed register)
      for (int f = 0; f < 6; ++f)
        if (!o->periodic[dir_axis(f)]) oracle_set_halo(o, q, f, nullptr);
    }
  }
  o->time = t + dt;
}

void oracle_maxabs(const oracle* o, int reg, double* out) {
  for (int f = 0; f < 4; ++f) out[f] = 0.0;
  for (int i = 0; i < o->nE[0]; ++i)
    for (int j = 0; j < o->nE[1]; ++j)
      for (int k = 0; k < o->nE[2]; ++k) {
        const size_t e = o->eidx(i, j, k);
        for (int f = 0; f < 4; ++f) {
          const double* s = o->node(reg, e, f);
          for (int nd = 0; nd < o->np; ++nd) {
            const double v = s[nd];
            if (std::fabs(v) > out[f] || std::isnan(v)) out[f] = std::fabs(v);
          }
        }
      }
}

void oracle_energy(const oracle* o, int reg, double* out) {
  double en = 0.0, massint = 0.0;
  const int n = o->n;
  for (int i = 0; i < o->nE[0]; ++i)
    for (int j = 0; j < o->nE[1]; ++j)
      for (int k = 0; k < o->nE[2]; ++k) {
        const size_t e = o->eidx(i, j, k);
        const Mat& m = o->mat[e];
        const double jac = o->width[0][i] * o->width[1][j] * o->width[2][k] / 8.0;
        for (int nd = 0; nd < o->np; ++nd) {
          const int a = nd / (n * n), b = (nd / n) % n, c = nd % n;
          const double wgt = o->wq[a] * o->wq[b] * o->wq[c] * jac;
          const double r = o->node(reg, e, 0)[nd];
          double u2 = 0.0;
          for (int d = 1; d <= 3; ++d) u2 += o->node(reg, e, d)[nd] * o->node(reg, e, d)[nd];
          en += wgt * 0.5 * (m.c0 * m.c0 * r * r / m.rho0 + m.rho0 * u2);
          massint += wgt * r;
        }
      }
  out[0] = en;
  out[1] = massint;
}

void oracle_sponge_sigma(const int* nE, const int* nB, const int* bpos, const double* ew,
                         const int* sw, double sigma_max, int ramp, double* sigma) {
  // Block::Block (Block.cpp:163-233) in float: distance to each sponge face from the
  // element's global integer index, the nearest face normalised by its width, then the ramp.
  const float w[3] = {(float)ew[0], (float)ew[1], (float)ew[2]};
  float dist[6] = {1000000000, 1000000000, 1000000000, 1000000000, 1000000000, 1000000000};
  size_t out = 0;
  for (int i = 0; i < nE[0] + 2; ++i) {
    const int gi = bpos[0] * nE[0] + i - 1;
    if (sw[0] > 0) dist[0] = std::abs(gi) * w[0];
    if (sw[1] > 0) dist[1] = std::abs(nE[0] * nB[0] - std::abs(gi) - 1) * w[0];
    for (int j = 0; j < nE[1] + 2; ++j) {
      const int gj = bpos[1] * nE[1] + j - 1;
      if (sw[2] > 0) dist[2] = std::abs(gj) * w[1];
      if (sw[3] > 0) dist[3] = std::abs(nE[1] * nB[1] - std::abs(gj) - 1) * w[1];
      for (int k = 0; k < nE[2] + 2; ++k) {
        const int gk = bpos[2] * nE[2] + k - 1;
        if (sw[4] > 0) dist[4] = std::abs(gk) * w[2];
        if (sw[5] > 0) dist[5] = std::abs(nE[2] * nB[2] - std::abs(gk) - 1) * w[2];
        int edge = 0;
        float shortest = dist[0];
        for (int t = 1; t < 6; ++t)
          if (dist[t] < shortest) { edge = t; shortest = dist[t]; }
        shortest /= float(sw[edge]) * w[int(edge * .5)];
        float s;
        if (sw[edge] == 0) s = 0.F;
        else if (ramp == 0) s = std::fmax(0.0F, 1.0F - shortest) * (float)sigma_max;
        else {
          const float r = std::fmax(0.0F, 1.0F - shortest);
          s = r * r * (float)sigma_max;
        }
        if (i >= 1 && i <= nE[0] && j >= 1 && j <= nE[1] && k >= 1 && k <= nE[2]) sigma[out++] = (double)s;
      }
    }
  }
}

}  // extern "C"
