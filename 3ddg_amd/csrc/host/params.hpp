// Run parameters in the reference's text format (driver.cpp:47-65 defaults, driver.cpp:129-196
// parser, BCParam.h:64-86 / BCParam.cpp:115 boundary and initial-condition defaults).
#pragma once
#include <ostream>
#include <string>

namespace dg3d_host {

struct Params {
  // driver.cpp:48-65
  unsigned int NUM_TIME_SUBSTEPS = 1;
  unsigned int NUM_OUTSTEPS = 1;
  unsigned int FREQ_TERMINAL = 1;
  bool BOOL_FILE = true;
  unsigned int NUM_ELEMENTS[3] = {4, 3, 2};
  unsigned int NUM_BLOCKS[3] = {5, 4, 6};
  unsigned short ORDER = 4;
  float ELEMENT_WIDTH[3] = {1.F, 2.F, 3.F};
  float STARTING_CORNER[3] = {0.25F, 0.5F, 0.75F};
  float DT = .001F;
  float C0 = 1500.0F;
  float RHO0 = 1000.0F;
  float B_OVER_A = 0.0F;
  float DIFF_CONST = 0.0F;
  bool eval_advect = false;
  bool eval_diff = false;
  bool eval_sponge = false;
  // BCParamDefault (BCParam.cpp:115)
  bool IS_PERIODIC[3] = {false, false, false};
  unsigned int spongeWidth[6] = {10, 10, 10, 10, 10, 10};
  float sigmaMax = -1.0F;
  bool IC_X_GAUSS = false, IC_Y_GAUSS = false, IC_Z_GAUSS = false;
  float IC_drho_amplitude = -1.0F;
  float IC_vx_amplitude = -1.0F, IC_vy_amplitude = -1.0F, IC_vz_amplitude = -1.0F;
  float IC_sigma = -1.0F;
  float IC_center_x = -1.0F, IC_center_y = -1.0F, IC_center_z = -1.0F;

  // Keys this implementation adds (the reference warns on unknown keys and ignores them, so
  // files using them stay readable by both).
  int SCHEME = 0;          // 0 = RK33 (reference), 1 = LSRK54, 2 = EULER (Grid::forwardEuler)
  float CFL = 0.0F;        // > 0: dt from the d-dimensional CFL condition instead of DT
  int SPONGE_RAMP = 0;     // 0 = linear (reference), 1 = quadratic
  int PRECISION = 0;       // 0 = FP64, 1 = mixed (FP64 state, FP32 right-hand side)
  int FUSED = 1;           // 1 = fused stage kernels, 0 = Grid::RK33 through Block operators
  int SOURCE_FACE = -1;    // boundary tone burst: face, amplitude, frequency, phase, t0, tau
  double SOURCE[5] = {0, 0, 0, 0, 1};
};

// Parses `path` with the reference's rules: lines starting with '#' or shorter than 2 characters
// are skipped; "NAME = values" otherwise; unknown names print a warning to `warn` and are
// ignored. Every line is echoed into `log` inside "Parameters used:\n{ ... }". Returns false if
// the file cannot be opened.
bool parse_parameters(const std::string& path, Params* p, std::ostream& log, std::ostream& warn);

}  // namespace dg3d_host
