// C harness around the reference's own dense algebra (cppsrc/Matrix.h, included from the
// reference tree at build time, never copied). It evaluates the line solve of
// Element::OperateUFromV, result = (M q w) / (S q f - Flux) (Element.cpp:602), with the
// reference's QR decomposition and its rank cut-off (Matrix.h:289-445, 625-628).
// Test infrastructure: used only to pin the oracle's per-line solve.
#include "Matrix.h"

extern "C" {

// q and w are passed as float, as the reference holds them (Element.cpp:504-505, Element.h:134).
// Returns the numerical rank found by decomposeQRQ for M q w.
int ref_line_solve(int n, const double* mass, const double* stiff, float q, float w, const double* f,
                   const double* flux, double* x) {
  Matrix M(n), S(n), fv(n, 1), Fl(n, 1);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      M[i][j] = mass[i * n + j];
      S[i][j] = stiff[i * n + j];
    }
    fv[i][0] = f[i];
    Fl[i][0] = flux[i];
  }
  Matrix Mqw = M * q * w;
  Matrix rhs = S * q * fv - Fl;
  Matrix r = Mqw / rhs;
  for (int i = 0; i < n; ++i) x[i] = r[0][i];
  Matrix Mcopy = Mqw;
  QResult qr = decomposeQRQ(Mcopy);
  return (int)qr.Rank;
}

}
