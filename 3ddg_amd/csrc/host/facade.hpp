// C++ host facade with the reference's class names and method shapes (Grid.h:29-76,
// Block.h:34-106), forwarding to the C ABI of include/dg3d.h. One Solver owns the device
// context of the block; each Grid is one state register of it.
#pragma once
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dg3d.h"
#include "params.hpp"

namespace dg3d_host {

struct Mesh {
  int order = 0;
  int nE[3] = {0, 0, 0};      // elements per axis of the block held on the device
  int nB[3] = {1, 1, 1};      // reference blocks (MPI ranks) the block is the union of
  double corner[3] = {0, 0, 0};
  std::vector<double> width[3];
  std::vector<double> nodes, weights, mass, stiff;
  int64_t E() const { return (int64_t)nE[0] * nE[1] * nE[2]; }
  int n() const { return order + 1; }
  int64_t dof() const { return E() * n() * n() * n(); }
};

class Solver;

class Block {
 public:
  Block(Solver* s, int reg) : solver_(s), reg_(reg) {}
  void OperateUFromV(unsigned short axis, unsigned short mode, Block* oldBlock);
  void OperateSponge(unsigned short axis, unsigned short mode, Block* oldBlock);
  void OperateQFromDRho(unsigned short axis, unsigned short mode, Block* oldBlock);
  void OperateDRhoFromQ(unsigned short axis, unsigned short mode, Block* oldBlock);
  // The FP64 drop-in takes the scalars in double (the reference passes float).
  void selfWeightedAdd(Block* other, double selfScalar, double otherScalar);
  void setZero();
  int reg() const { return reg_; }

 private:
  Solver* solver_;
  int reg_;
};

class Grid {
 public:
  Grid(Solver* s, int reg, bool evalAdvect, bool evalDiff, bool evalSponge);
  // Grid::RK33 (Grid.cpp:642-902) through the Block operators.
  void RK33(double dt, Grid* oldGrid, Grid* k1, Grid* k2, std::ofstream& logFile);
  // Grid::forwardEuler (Grid.cpp:567-640): ignores the evaluate flags, as the reference does;
  // the LDG pair runs for the x axis only. Needs a Solver with Q fields.
  void forwardEuler(double dt, Grid* oldGrid, std::ofstream& logFile);
  void communicate();
  void setZero();
  void copyFrom(const Grid& g);
  Block mBlock;
  double mTime = 0.0;

 private:
  Solver* solver_;
  bool mEvalAdvect, mEvalDiff, mEvalSponge;
};

class Solver {
 public:
  Solver(const Mesh& mesh, const int* periodic, int precision, int registers, bool with_q);
  ~Solver();
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;
  dg3d_ctx* ctx() const { return ctx_; }
  bool with_q() const { return with_q_; }
  void check(int rc, const char* what) const;
  const Mesh& mesh() const { return mesh_; }

 private:
  dg3d_ctx* ctx_ = nullptr;
  Mesh mesh_;
  bool with_q_ = false;
};

// Problem setup from reference parameters (global mesh of NUM_BLOCKS x NUM_ELEMENTS elements).
bool load_basis(int order, const std::vector<std::string>& dirs, Mesh* m, std::string* err);
void build_mesh(const Params& p, Mesh* m);
void material_from_params(const Params& p, const Mesh& m, std::vector<double>* c0, std::vector<double>* rho0,
                          std::vector<double>* ba, std::vector<double>* sigma, std::vector<double>* diff);
// Gaussian IC (Node.cpp:76-101) on the block's nodes and on the halo face layers.
void gaussian_ic(const Params& p, const Mesh& m, std::vector<double>* state, std::vector<double> traces[6]);
double cfl_dt(const Mesh& m, double c0, double cfl);

}  // namespace dg3d_host
