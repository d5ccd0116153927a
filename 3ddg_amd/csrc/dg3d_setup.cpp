// Host-side setup entry points of the C ABI: the LGL basis files the reference reads but does
// not ship, and the per-element sponge coefficient of Block::Block.
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "basis.hpp"
#include "dg3d.h"

extern "C" {

int dg3d_basis_generate(int order, double* nodes, double* weights, double* mass, double* stiff, double* dpsi) {
  if (order < 0 || order > 16) return DG3D_E_INVALID;
  dg3d::LineBasis b = dg3d::make_line_basis(order);
  const int n = order + 1;
  for (int i = 0; i < n; ++i) {
    if (nodes) nodes[i] = b.nodes[i];
    if (weights) weights[i] = b.weights[i];
  }
  for (int i = 0; i < n * n; ++i) {
    if (mass) mass[i] = b.mass[i];
    if (stiff) stiff[i] = b.stiff[i];
    if (dpsi) dpsi[i] = b.dpsi[i];
  }
  return DG3D_OK;
}

int dg3d_basis_write_files(const char* dir, int max_order) {
  if (!dir || max_order < 1 || max_order > 16) return DG3D_E_INVALID;
  const std::string d(dir);
  if (!dg3d::write_coeff_file(d + "/hardcoded_coeff.txt", max_order)) return DG3D_E_IO;
  if (!dg3d::write_dpsi_file(d + "/hardcoded_coeff2.txt", max_order)) return DG3D_E_IO;
  if (!dg3d::write_matrices_file(d + "/matrices.txt", max_order)) return DG3D_E_IO;
  return DG3D_OK;
}

int dg3d_basis_read_files(const char* coeff_path, const char* matrices_path, int order, double* nodes,
                          double* weights, double* mass, double* stiff) {
  if (!coeff_path || !matrices_path || order < 1) return DG3D_E_INVALID;
  std::vector<double> z, w, m, s;
  std::string err;
  if (!dg3d::read_coeff_file(coeff_path, order, &z, &w, &err)) return DG3D_E_IO;
  if (!dg3d::read_matrices_file(matrices_path, order, &m, &s, &err)) return DG3D_E_IO;
  const int n = order + 1;
  for (int i = 0; i < n; ++i) {
    if (nodes) nodes[i] = z[i];
    if (weights) weights[i] = w[i];
  }
  for (int i = 0; i < n * n; ++i) {
    if (mass) mass[i] = m[i];
    if (stiff) stiff[i] = s[i];
  }
  return DG3D_OK;
}

// Block::Block (Block.cpp:163-233). Distances to the six block-union faces are taken from the
// element's global integer position times the axis width, in float; the nearest sponge face
// (first minimum in face order -x,+x,-y,+y,-z,+z) normalises by its sponge width in elements.
// The distance slots keep their last value when a face has no sponge, as in the reference.
int dg3d_sponge_sigma(const int* nE, const int* nB, const int* bpos, const double* ew, const int* sw,
                      double sigma_max, int ramp, double* sigma) {
  if (!nE || !nB || !bpos || !ew || !sw || !sigma || (ramp != 0 && ramp != 1)) return DG3D_E_INVALID;
  const float w[3] = {(float)ew[0], (float)ew[1], (float)ew[2]};
  const float smax = (float)sigma_max;
  float dist[6];
  for (float& d : dist) d = 1000000000.0f;
  size_t k_out = 0;
  for (int i = 0; i <= nE[0] + 1; ++i) {
    const int g0 = bpos[0] * nE[0] + i - 1;
    if (sw[0] > 0) dist[0] = (float)std::abs(g0) * w[0];
    if (sw[1] > 0) dist[1] = (float)std::abs(nE[0] * nB[0] - std::abs(g0) - 1) * w[0];
    for (int j = 0; j <= nE[1] + 1; ++j) {
      const int g1 = bpos[1] * nE[1] + j - 1;
      if (sw[2] > 0) dist[2] = (float)std::abs(g1) * w[1];
      if (sw[3] > 0) dist[3] = (float)std::abs(nE[1] * nB[1] - std::abs(g1) - 1) * w[1];
      for (int k = 0; k <= nE[2] + 1; ++k) {
        const int g2 = bpos[2] * nE[2] + k - 1;
        if (sw[4] > 0) dist[4] = (float)std::abs(g2) * w[2];
        if (sw[5] > 0) dist[5] = (float)std::abs(nE[2] * nB[2] - std::abs(g2) - 1) * w[2];
        const bool interior = i >= 1 && i <= nE[0] && j >= 1 && j <= nE[1] && k >= 1 && k <= nE[2];
        if (!interior) continue;
        int face = 0;
        float best = dist[0];
        for (int f = 1; f < 6; ++f)
          if (dist[f] < best) {
            face = f;
            best = dist[f];
          }
        best /= (float)sw[face] * w[face / 2];
        float s = 0.0f;
        if (sw[face] != 0) {
          const float r = std::fmax(0.0f, 1.0f - best);
          s = (ramp == 0) ? r * smax : r * r * smax;
        }
        sigma[k_out++] = (double)s;
      }
    }
  }
  return DG3D_OK;
}

}  // extern "C"
