// LGL basis generation and the reference's hex-double file formats.
// Formats follow Basis.cpp:135-171 (coeff: "<2 chars><count>" line, zeros line, weights
// line; entry k holds k+1 values so that getZeros(order) = entry[order]), Basis.cpp:178-236
// (dpsi: a line starting with 'p' then rows) and Grid.cpp:920-989 ("Order=<N>", n rows of
// M, n rows of S). Hex decoding matches Grid.cpp:63-75 / Basis.cpp:52-65 (the 64 IEEE bits).
#include "basis.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace dg3d {

namespace {

// Legendre P_0..P_m at x by the three-term recurrence; returns P_m and P_{m-1}.
void legendre(int m, double x, double* pm, double* pm1) {
  double p0 = 1.0, p1 = x;
  if (m == 0) { *pm = 1.0; *pm1 = 0.0; return; }
  for (int k = 2; k <= m; ++k) {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *pm = p1;
  *pm1 = p0;
}

// Gauss-Legendre nodes/weights (m points) by Newton on P_m, for exact mass integration.
void gauss_legendre(int m, std::vector<double>* x, std::vector<double>* w) {
  x->assign(m, 0.0);
  w->assign(m, 0.0);
  for (int i = 0; i < m; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (m + 0.5));
    for (int it = 0; it < 100; ++it) {
      double pm, pm1;
      legendre(m, z, &pm, &pm1);
      double dp = m * (z * pm - pm1) / (z * z - 1.0);
      double dz = pm / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double pm, pm1;
    legendre(m, z, &pm, &pm1);
    double dp = m * (z * pm - pm1) / (z * z - 1.0);
    (*x)[m - 1 - i] = z;
    (*w)[m - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// Lagrange basis l_j(x) through nodes, in barycentric form.
void lagrange_values(const std::vector<double>& nodes, const std::vector<double>& bary, double x,
                     std::vector<double>* out) {
  const int n = (int)nodes.size();
  out->assign(n, 0.0);
  for (int j = 0; j < n; ++j) {
    if (x == nodes[j]) { (*out)[j] = 1.0; return; }
  }
  double den = 0.0;
  for (int j = 0; j < n; ++j) den += bary[j] / (x - nodes[j]);
  for (int j = 0; j < n; ++j) (*out)[j] = bary[j] / (x - nodes[j]) / den;
}

}  // namespace

LineBasis make_line_basis(int order) {
  LineBasis b;
  b.order = order;
  const int n = order + 1;
  b.nodes.assign(n, 0.0);
  b.weights.assign(n, 0.0);
  if (order == 0) {
    b.nodes[0] = 0.0;
    b.weights[0] = 2.0;
  } else {
    // LGL nodes: endpoints and the roots of P_N'. Newton on (x P_N - P_{N-1}) = 0 from
    // Chebyshev-Gauss-Lobatto starting points; weights 2 / (N (N+1) P_N(x)^2).
    for (int i = 0; i <= order; ++i) {
      double x = -std::cos(M_PI * i / order);
      for (int it = 0; it < 200; ++it) {
        double pn, pn1;
        legendre(order, x, &pn, &pn1);
        double dx = (x * pn - pn1) / ((order + 1.0) * pn);
        x -= dx;
        if (std::fabs(dx) < 1e-17) break;
      }
      if (i == 0) x = -1.0;
      if (i == order) x = 1.0;
      b.nodes[i] = x;
    }
    // Symmetrise to remove the last-bit asymmetry of the iteration.
    for (int i = 0; i < n / 2; ++i) {
      double s = 0.5 * (b.nodes[n - 1 - i] - b.nodes[i]);
      b.nodes[i] = -s;
      b.nodes[n - 1 - i] = s;
    }
    if (n % 2 == 1) b.nodes[n / 2] = 0.0;
    for (int i = 0; i < n; ++i) {
      double pn, pn1;
      legendre(order, b.nodes[i], &pn, &pn1);
      b.weights[i] = 2.0 / (order * (order + 1.0) * pn * pn);
    }
  }
  // Barycentric weights and the derivative matrix D_ij = l_j'(x_i).
  std::vector<double> bary(n, 1.0);
  for (int j = 0; j < n; ++j)
    for (int m = 0; m < n; ++m)
      if (m != j) bary[j] /= (b.nodes[j] - b.nodes[m]);
  b.dpsi.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    double diag = 0.0;
    for (int j = 0; j < n; ++j) {
      if (i == j) continue;
      double d = (bary[j] / bary[i]) / (b.nodes[i] - b.nodes[j]);
      b.dpsi[i * n + j] = d;
      diag -= d;
    }
    b.dpsi[i * n + i] = diag;
  }
  // Exact mass on [-1,1] by Gauss-Legendre with n+1 points (exact to degree 2n+1 >= 2N),
  // then halved for the unit interval (the reference scales by the full element width).
  std::vector<double> gx, gw, lv;
  gauss_legendre(n + 1, &gx, &gw);
  b.mass.assign(n * n, 0.0);
  for (size_t q = 0; q < gx.size(); ++q) {
    lagrange_values(b.nodes, bary, gx[q], &lv);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) b.mass[i * n + j] += 0.5 * gw[q] * lv[i] * lv[j];
  }
  // Weak stiffness S_ij = int l_i' l_j = w_j D_ji (LGL quadrature is exact to degree 2N-1).
  b.stiff.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) b.stiff[i * n + j] = b.weights[j] * b.dpsi[j * n + i];
  return b;
}

bool make_line_operator(int n, const double* mass, const double* stiff, LineOperator* out) {
  // Solve M X = [S | e_0 | e_N] by Gaussian elimination with partial pivoting.
  const int nr = n + 2;
  std::vector<double> a(mass, mass + n * n);
  std::vector<double> rhs(n * nr, 0.0);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) rhs[i * nr + j] = stiff[i * n + j];
    rhs[i * nr + n] = (i == 0) ? 1.0 : 0.0;
    rhs[i * nr + n + 1] = (i == n - 1) ? 1.0 : 0.0;
  }
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[i * n + k]) > std::fabs(a[p * n + k])) p = i;
    if (std::fabs(a[p * n + k]) < 1e-300) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[p * n + j]);
      for (int j = 0; j < nr; ++j) std::swap(rhs[k * nr + j], rhs[p * nr + j]);
    }
    for (int i = k + 1; i < n; ++i) {
      double f = a[i * n + k] / a[k * n + k];
      for (int j = k; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
      for (int j = 0; j < nr; ++j) rhs[i * nr + j] -= f * rhs[k * nr + j];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    for (int j = 0; j < nr; ++j) {
      double s = rhs[k * nr + j];
      for (int m = k + 1; m < n; ++m) s -= a[k * n + m] * rhs[m * nr + j];
      rhs[k * nr + j] = s / a[k * n + k];
    }
  }
  out->n = n;
  out->A.assign(n * n, 0.0);
  out->L0.assign(n, 0.0);
  out->LN.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) out->A[i * n + j] = rhs[i * nr + j];
    out->L0[i] = rhs[i * nr + n];
    out->LN[i] = rhs[i * nr + n + 1];
  }
  return true;
}

std::string double_to_hex(double v) {
  uint64_t bits;
  std::memcpy(&bits, &v, 8);
  char buf[32];
  std::snprintf(buf, sizeof(buf), "%016llx", (unsigned long long)bits);
  return buf;
}

bool hex_to_double(const std::string& s, double* out) {
  if (s.empty() || s.size() > 16) return false;
  uint64_t bits = 0;
  for (char ch : s) {
    int v;
    if (ch >= '0' && ch <= '9') v = ch - '0';
    else if (ch >= 'a' && ch <= 'f') v = ch - 'a' + 10;
    else if (ch >= 'A' && ch <= 'F') v = ch - 'A' + 10;
    else return false;
    bits = (bits << 4) | (uint64_t)v;
  }
  std::memcpy(out, &bits, 8);
  return true;
}

bool write_coeff_file(const std::string& path, int max_order) {
  std::ofstream f(path);
  if (!f) return false;
  for (int N = 0; N <= max_order; ++N) {
    LineBasis b = make_line_basis(N);
    f << "n=" << (N + 1) << "\n";
    for (int i = 0; i <= N; ++i) f << (i ? " " : "") << double_to_hex(b.nodes[i]);
    f << "\n";
    for (int i = 0; i <= N; ++i) f << (i ? " " : "") << double_to_hex(b.weights[i]);
    f << "\n";
  }
  return (bool)f;
}

bool write_dpsi_file(const std::string& path, int max_order) {
  std::ofstream f(path);
  if (!f) return false;
  for (int N = 0; N <= max_order; ++N) {
    LineBasis b = make_line_basis(N);
    const int n = N + 1;
    f << "p=" << n << "\n";
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) f << (j ? " " : "") << double_to_hex(b.dpsi[i * n + j]);
      f << "\n";
    }
  }
  return (bool)f;
}

bool write_matrices_file(const std::string& path, int max_order) {
  std::ofstream f(path);
  if (!f) return false;
  for (int N = 1; N <= max_order; ++N) {
    LineBasis b = make_line_basis(N);
    const int n = N + 1;
    f << "Order=" << N << "\n";
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) f << (j ? " " : "") << double_to_hex(b.mass[i * n + j]);
      f << "\n";
    }
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) f << (j ? " " : "") << double_to_hex(b.stiff[i * n + j]);
      f << "\n";
    }
  }
  return (bool)f;
}

bool read_coeff_file(const std::string& path, int order, std::vector<double>* zeros,
                     std::vector<double>* weights, std::string* err) {
  std::ifstream f(path);
  if (!f) {
    if (err) *err = "Opening file '" + path + "' failed, it either doesn't exist or is not accessible.";
    return false;
  }
  // Entries are (count line, zeros line, weights line); the entry with index `order`
  // (0-based, in file order) is the one the reference's getZeros(order) returns.
  std::string line;
  int phase = 0, entry = -1, count = 0;
  while (std::getline(f, line)) {
    if (phase == 0) {
      if (line.size() < 3) { if (err) *err = "bad count line in " + path; return false; }
      count = std::atoi(line.c_str() + 2);
      ++entry;
    } else {
      std::vector<double>* dst = (phase == 1) ? zeros : weights;
      if (entry == order) {
        std::istringstream iss(line);
        dst->clear();
        std::string tok;
        for (int i = 0; i < count; ++i) {
          double v;
          if (!(iss >> tok) || !hex_to_double(tok, &v)) {
            if (err) *err = "bad hex value in " + path;
            return false;
          }
          dst->push_back(v);
        }
      }
    }
    phase = (phase + 1) % 3;
    if (entry > order && phase == 0) break;
  }
  if (entry < order || (int)zeros->size() != order + 1 || (int)weights->size() != order + 1) {
    if (err) *err = "order " + std::to_string(order) + " not found in " + path;
    return false;
  }
  return true;
}

bool read_matrices_file(const std::string& path, int order, std::vector<double>* mass,
                        std::vector<double>* stiff, std::string* err) {
  std::ifstream f(path);
  if (!f) {
    if (err) *err = "Opening file '" + path + "' failed, it either doesn't exist or is not accessible.";
    return false;
  }
  const int n = order + 1;
  mass->assign(n * n, 0.0);
  stiff->assign(n * n, 0.0);
  std::string line;
  int cur = -1, row = 0;
  bool found = false;
  while (std::getline(f, line)) {
    if (line.compare(0, 6, "Order=") == 0) {
      cur = std::atoi(line.c_str() + 6);
      if (cur > order) break;
      continue;
    }
    if (cur != order) continue;
    found = true;
    std::istringstream iss(line);
    std::string tok;
    for (int c = 0; c < n; ++c) {
      double v;
      if (!(iss >> tok) || !hex_to_double(tok, &v)) {
        if (err) *err = "bad hex value in " + path;
        return false;
      }
      if (row < n) (*mass)[row * n + c] = v;
      else if (row < 2 * n) (*stiff)[(row - n) * n + c] = v;
    }
    ++row;
  }
  if (!found || row < 2 * n) {
    if (err) *err = "order " + std::to_string(order) + " not found in " + path;
    return false;
  }
  return true;
}

}  // namespace dg3d
