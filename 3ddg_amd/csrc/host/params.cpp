#include "params.hpp"

#include <fstream>
#include <sstream>

namespace dg3d_host {

bool parse_parameters(const std::string& path, Params* P, std::ostream& log, std::ostream& warn) {
  std::ifstream f(path);
  if (!f.is_open()) return false;
  Params& p = *P;
  std::string line;
  log << "Parameters used:\n{\n";
  while (std::getline(f, line)) {
    log << "  " << line << "\n";
    if (line.empty() || line[0] == '#' || line.length() < 2) continue;
    // name is everything before " = " (the whole line when absent, which then matches nothing)
    const size_t pos = line.find(" = ");
    const std::string name = line.substr(0, pos);
    const std::string values = (pos == std::string::npos) ? std::string() : line.substr(pos + 3);
    std::istringstream iss(values);
    if (name == "NUM_TIME_SUBSTEPS") iss >> p.NUM_TIME_SUBSTEPS;
    else if (name == "NUM_OUTSTEPS") iss >> p.NUM_OUTSTEPS;
    else if (name == "FREQ_TERMINAL") iss >> p.FREQ_TERMINAL;
    else if (name == "BOOL_FILE") iss >> p.BOOL_FILE;
    else if (name == "C0") iss >> p.C0;
    else if (name == "RHO0") iss >> p.RHO0;
    else if (name == "B_OVER_A") iss >> p.B_OVER_A;
    else if (name == "DIFF_CONST") iss >> p.DIFF_CONST;
    else if (name == "NUM_ELEMENTS") iss >> p.NUM_ELEMENTS[0] >> p.NUM_ELEMENTS[1] >> p.NUM_ELEMENTS[2];
    else if (name == "NUM_BLOCKS") iss >> p.NUM_BLOCKS[0] >> p.NUM_BLOCKS[1] >> p.NUM_BLOCKS[2];
    else if (name == "ORDER") iss >> p.ORDER;
    else if (name == "ELEMENT_WIDTH") iss >> p.ELEMENT_WIDTH[0] >> p.ELEMENT_WIDTH[1] >> p.ELEMENT_WIDTH[2];
    else if (name == "STARTING_CORNER") iss >> p.STARTING_CORNER[0] >> p.STARTING_CORNER[1] >> p.STARTING_CORNER[2];
    else if (name == "DT") iss >> p.DT;
    else if (name == "EVALUATE_DIFFUSION") iss >> p.eval_diff;
    else if (name == "EVALUATE_ADVECTION") iss >> p.eval_advect;
    else if (name == "EVALUATE_SPONGE") iss >> p.eval_sponge;
    else if (name == "BC_IS_PERIODIC") iss >> p.IS_PERIODIC[0] >> p.IS_PERIODIC[1] >> p.IS_PERIODIC[2];
    else if (name == "BC_SIGMA_MAX") iss >> p.sigmaMax;
    else if (name == "BC_SPONGE_WIDTH")
      iss >> p.spongeWidth[0] >> p.spongeWidth[1] >> p.spongeWidth[2] >> p.spongeWidth[3] >> p.spongeWidth[4] >>
          p.spongeWidth[5];
    else if (name == "IC_X_GAUSS") iss >> p.IC_X_GAUSS;
    else if (name == "IC_Y_GAUSS") iss >> p.IC_Y_GAUSS;
    else if (name == "IC_Z_GAUSS") iss >> p.IC_Z_GAUSS;
    else if (name == "IC_DRHO_AMPLITUDE") iss >> p.IC_drho_amplitude;
    else if (name == "IC_VX_AMPLITUDE") iss >> p.IC_vx_amplitude;
    else if (name == "IC_VY_AMPLITUDE") iss >> p.IC_vy_amplitude;
    else if (name == "IC_VZ_AMPLITUDE") iss >> p.IC_vz_amplitude;
    else if (name == "IC_SIGMA") iss >> p.IC_sigma;
    else if (name == "IC_CENTER_X") iss >> p.IC_center_x;
    else if (name == "IC_CENTER_Y") iss >> p.IC_center_y;
    else if (name == "IC_CENTER_Z") iss >> p.IC_center_z;
    else if (name == "SCHEME") {
      std::string v;
      iss >> v;
      p.SCHEME = (v == "LSRK54" || v == "1") ? 1 : ((v == "EULER" || v == "2") ? 2 : 0);
    } else if (name == "CFL") iss >> p.CFL;
    else if (name == "SPONGE_RAMP") iss >> p.SPONGE_RAMP;
    else if (name == "PRECISION") {
      std::string v;
      iss >> v;
      p.PRECISION = (v == "MIXED" || v == "1") ? 1 : 0;
    } else if (name == "FUSED") iss >> p.FUSED;
    else if (name == "SOURCE") iss >> p.SOURCE_FACE >> p.SOURCE[0] >> p.SOURCE[1] >> p.SOURCE[2] >> p.SOURCE[3] >> p.SOURCE[4];
    else warn << "WARNING: IGNORING INVALID PARAMETER LINE: \"" << line << "\"\n";
  }
  log << "}\n";
  return true;
}

}  // namespace dg3d_host
