// One-dimensional Legendre-Gauss-Lobatto (LGL) basis data for the tensor-product
// hexahedra: nodes, weights, unit-interval mass matrix M, weak stiffness S, and the
// reference's three hex-double text formats that carry them.
//
// The reference reads these numbers from files it does not ship:
//   hardcoded_coeff.txt  (zeros, weights)  Basis.cpp:74-171
//   hardcoded_coeff2.txt (dpsi rows)        Basis.cpp:178-236
//   matrices.txt         (M rows, S rows)   Grid.cpp:920-989
// This module generates them from the mathematics and reads/writes the same formats.
#pragma once
#include <string>
#include <vector>

namespace dg3d {

struct LineBasis {
  int order = 0;                 // N; n = N+1 nodes
  std::vector<double> nodes;     // LGL nodes on [-1,1], ascending
  std::vector<double> weights;   // LGL quadrature weights on [-1,1]
  std::vector<double> dpsi;      // D[i*n+j] = l_j'(xi_i) on [-1,1]
  std::vector<double> mass;      // unit-interval mass M[i*n+j] = (1/2) int_{-1}^{1} l_i l_j
  std::vector<double> stiff;     // S[i*n+j] = int l_i' l_j  (scale free)
};

// Generate the basis for order N >= 0 (N <= 16) in FP64.
LineBasis make_line_basis(int order);

// Line operator used by every kernel: with q the per-mode scale and w the element width,
// the reference's per-line solve (M q w) x = S q f - Flux (Element.cpp:602) equals
//   x = (1/w) * A f - (1/(q w)) * (Flux_0 * L0 + Flux_N * LN),
// A = M^{-1} S, L0 = M^{-1} e_0, LN = M^{-1} e_N.
struct LineOperator {
  int n = 0;
  std::vector<double> A;   // n*n row-major
  std::vector<double> L0;  // n
  std::vector<double> LN;  // n
};
// Returns false if M is singular (pivot below 1e-300).
bool make_line_operator(int n, const double* mass, const double* stiff, LineOperator* out);

// Hex-double encoding used by all three files: 16 hex digits of the IEEE-754 bits.
std::string double_to_hex(double v);
bool hex_to_double(const std::string& s, double* out);

// Writers for the reference formats (orders 0..max_order; matrices.txt starts at order 1).
bool write_coeff_file(const std::string& path, int max_order);    // hardcoded_coeff.txt
bool write_dpsi_file(const std::string& path, int max_order);     // hardcoded_coeff2.txt
bool write_matrices_file(const std::string& path, int max_order); // matrices.txt

// Readers with the reference's parsing rules. Return false with a message on failure.
bool read_coeff_file(const std::string& path, int order, std::vector<double>* zeros,
                     std::vector<double>* weights, std::string* err);
bool read_matrices_file(const std::string& path, int order, std::vector<double>* mass,
                        std::vector<double>* stiff, std::string* err);

}  // namespace dg3d
