#include "facade.hpp"

#include <algorithm>
#include <cmath>

namespace dg3d_host {

Solver::Solver(const Mesh& mesh, const int* periodic, int precision, int registers, bool with_q)
    : mesh_(mesh), with_q_(with_q) {
  dg3d_desc d{};
  d.order = mesh.order;
  for (int a = 0; a < 3; ++a) {
    d.num_elements[a] = mesh.nE[a];
    d.periodic[a] = periodic ? periodic[a] : 0;
  }
  d.precision = precision;
  d.contraction = DG3D_CONTRACT_VALU;
  d.num_registers = registers;
  d.with_q = with_q ? 1 : 0;
  d.xcd_remap = 2;
  d.mass = mesh.mass.data();
  d.stiff = mesh.stiff.data();
  d.lgl_nodes = mesh.nodes.data();
  d.lgl_weights = mesh.weights.data();
  const int rc = dg3d_create(&d, &ctx_);
  if (rc != DG3D_OK) throw std::runtime_error(std::string("dg3d_create: ") + dg3d_create_error());
  check(dg3d_set_geometry(ctx_, mesh.corner, mesh.width[0].data(), mesh.width[1].data(), mesh.width[2].data()),
        "set_geometry");
}

Solver::~Solver() { dg3d_destroy(ctx_); }

void Solver::check(int rc, const char* what) const {
  if (rc != DG3D_OK) throw std::runtime_error(std::string(what) + ": " + dg3d_last_error(ctx_));
}

void Block::OperateUFromV(unsigned short axis, unsigned short mode, Block* old) {
  solver_->check(dg3d_operate(solver_->ctx(), DG3D_OP_UFROMV, axis, mode, old->reg_, reg_), "OperateUFromV");
}
void Block::OperateSponge(unsigned short axis, unsigned short mode, Block* old) {
  solver_->check(dg3d_operate(solver_->ctx(), DG3D_OP_SPONGE, axis, mode, old->reg_, reg_), "OperateSponge");
}
void Block::OperateQFromDRho(unsigned short axis, unsigned short mode, Block* old) {
  solver_->check(dg3d_operate(solver_->ctx(), DG3D_OP_QFROMDRHO, axis, mode, old->reg_, reg_), "OperateQFromDRho");
}
void Block::OperateDRhoFromQ(unsigned short axis, unsigned short mode, Block* old) {
  solver_->check(dg3d_operate(solver_->ctx(), DG3D_OP_DRHOFROMQ, axis, mode, old->reg_, reg_), "OperateDRhoFromQ");
}
void Block::selfWeightedAdd(Block* other, double a, double b) {
  solver_->check(dg3d_axpby(solver_->ctx(), reg_, a, other->reg_, b), "selfWeightedAdd");
}
void Block::setZero() { solver_->check(dg3d_zero(solver_->ctx(), reg_), "setZero"); }

Grid::Grid(Solver* s, int reg, bool adv, bool diff, bool sponge)
    : mBlock(s, reg), solver_(s), mEvalAdvect(adv), mEvalDiff(diff), mEvalSponge(sponge) {}

void Grid::communicate() { solver_->check(dg3d_halo(solver_->ctx(), mBlock.reg()), "communicate"); }
void Grid::setZero() { mBlock.setZero(); }
void Grid::copyFrom(const Grid& g) {
  solver_->check(dg3d_copy(solver_->ctx(), mBlock.reg(), g.mBlock.reg()), "copy");
  mTime = g.mTime;
}

void Grid::RK33(double dt, Grid* oldGrid, Grid* k1, Grid* k2, std::ofstream& logFile) {
  (void)logFile;
  mTime += dt;
  oldGrid->mTime += dt;
  k1->setZero();
  k2->setZero();
  setZero();
  Grid* src[3] = {oldGrid, k1, k2};
  Grid* dst[3] = {k1, k2, this};
  for (int s = 0; s < 3; ++s) {
    src[s]->communicate();
    if (mEvalDiff) {
      for (unsigned short a = 0; a < 3; ++a) src[s]->mBlock.OperateQFromDRho(a, 0, &src[s]->mBlock);
      src[s]->communicate();
      for (unsigned short a = 0; a < 3; ++a) dst[s]->mBlock.OperateDRhoFromQ(a, 0, &src[s]->mBlock);
    }
    if (mEvalAdvect)
      for (unsigned short a = 0; a < 3; ++a) {
        dst[s]->mBlock.OperateUFromV(a, 0, &src[s]->mBlock);
        dst[s]->mBlock.OperateUFromV(a, 1, &src[s]->mBlock);
      }
    if (mEvalSponge) dst[s]->mBlock.OperateSponge(0, 0, &src[s]->mBlock);
    if (s == 0) {
      k1->mBlock.selfWeightedAdd(&oldGrid->mBlock, dt, 1.0);
    } else if (s == 1) {
      k2->mBlock.selfWeightedAdd(&k1->mBlock, dt * 0.25, 0.25);
      k2->mBlock.selfWeightedAdd(&oldGrid->mBlock, 1.0, 0.75);
    } else {
      mBlock.selfWeightedAdd(&k2->mBlock, dt * (2.0 / 3.0), 2.0 / 3.0);
      mBlock.selfWeightedAdd(&oldGrid->mBlock, 1.0, 1.0 / 3.0);
    }
  }
}

void Grid::forwardEuler(double dt, Grid* oldGrid, std::ofstream& logFile) {
  (void)logFile;
  if (!solver_->with_q())
    throw std::runtime_error("Grid::forwardEuler runs the x-axis LDG pair: the Solver needs Q fields");
  mTime += dt;
  oldGrid->mTime += dt;
  setZero();
  oldGrid->communicate();
  oldGrid->mBlock.OperateQFromDRho(0, 0, &oldGrid->mBlock);
  oldGrid->communicate();
  mBlock.OperateDRhoFromQ(0, 0, &oldGrid->mBlock);
  for (unsigned short a = 0; a < 3; ++a) {
    mBlock.OperateUFromV(a, 0, &oldGrid->mBlock);
    mBlock.OperateUFromV(a, 1, &oldGrid->mBlock);
  }
  mBlock.OperateSponge(0, 0, &oldGrid->mBlock);
  mBlock.selfWeightedAdd(&oldGrid->mBlock, dt, 1.0);
}

bool load_basis(int order, const std::vector<std::string>& dirs, Mesh* m, std::string* err) {
  const int n = order + 1;
  m->nodes.assign(n, 0.0);
  m->weights.assign(n, 0.0);
  m->mass.assign(n * n, 0.0);
  m->stiff.assign(n * n, 0.0);
  for (const std::string& d : dirs) {
    const std::string c = d + "/hardcoded_coeff.txt", mt = d + "/matrices.txt";
    if (dg3d_basis_read_files(c.c_str(), mt.c_str(), order, m->nodes.data(), m->weights.data(), m->mass.data(),
                              m->stiff.data()) == DG3D_OK)
      return true;
  }
  if (err) *err = "Opening file 'hardcoded_coeff.txt' or 'matrices.txt' failed, they either don't exist or are not accessible.";
  return false;
}

void build_mesh(const Params& p, Mesh* m) {
  m->order = p.ORDER;
  for (int a = 0; a < 3; ++a) {
    m->nB[a] = (int)p.NUM_BLOCKS[a];
    m->nE[a] = (int)(p.NUM_ELEMENTS[a] * p.NUM_BLOCKS[a]);
    m->corner[a] = (double)p.STARTING_CORNER[a];
    m->width[a].assign(m->nE[a], (double)p.ELEMENT_WIDTH[a]);
  }
}

void material_from_params(const Params& p, const Mesh& m, std::vector<double>* c0, std::vector<double>* rho0,
                          std::vector<double>* ba, std::vector<double>* sigma, std::vector<double>* diff) {
  const int64_t E = m.E();
  c0->assign(E, (double)p.C0);
  rho0->assign(E, (double)p.RHO0);
  ba->assign(E, (double)p.B_OVER_A);
  diff->assign(E, (double)p.DIFF_CONST);
  sigma->assign(E, 0.0);
  if (p.eval_sponge) {
    const int one[3] = {1, 1, 1}, zero[3] = {0, 0, 0};
    const double ew[3] = {(double)p.ELEMENT_WIDTH[0], (double)p.ELEMENT_WIDTH[1], (double)p.ELEMENT_WIDTH[2]};
    int sw[6];
    for (int f = 0; f < 6; ++f) sw[f] = (int)p.spongeWidth[f];
    dg3d_sponge_sigma(m.nE, one, zero, ew, sw, (double)p.sigmaMax, p.SPONGE_RAMP, sigma->data());
  }
}

static double gauss_at(const Params& p, double x, double y, double z) {
  double d2 = 0.0;
  if (p.IC_X_GAUSS) d2 += (x - p.IC_center_x) * (x - p.IC_center_x);
  if (p.IC_Y_GAUSS) d2 += (y - p.IC_center_y) * (y - p.IC_center_y);
  if (p.IC_Z_GAUSS) d2 += (z - p.IC_center_z) * (z - p.IC_center_z);
  return std::exp(-d2 * (double)p.IC_sigma);
}

void gaussian_ic(const Params& p, const Mesh& m, std::vector<double>* st, std::vector<double> traces[6]) {
  const int n = m.n();
  const int64_t np = (int64_t)n * n * n, fs = m.E() * np;
  st->assign(4 * fs, 0.0);
  for (int f = 0; f < 6; ++f) traces[f].clear();
  if (!(p.IC_X_GAUSS || p.IC_Y_GAUSS || p.IC_Z_GAUSS)) return;
  std::vector<double> ec[3];  // element lower corners
  for (int a = 0; a < 3; ++a) {
    ec[a].resize(m.nE[a]);
    double acc = m.corner[a];
    for (int i = 0; i < m.nE[a]; ++i) {
      ec[a][i] = acc;
      acc += m.width[a][i];
    }
  }
  auto coord = [&](int a, int e, int node) { return m.width[a][e] * (m.nodes[node] / 2.0 + 0.5) + ec[a][e]; };
  const double amp[4] = {(double)p.IC_drho_amplitude, (double)p.IC_vx_amplitude, (double)p.IC_vy_amplitude,
                         (double)p.IC_vz_amplitude};
  for (int i = 0; i < m.nE[0]; ++i)
    for (int j = 0; j < m.nE[1]; ++j)
      for (int k = 0; k < m.nE[2]; ++k) {
        const int64_t e = ((int64_t)i * m.nE[1] + j) * m.nE[2] + k;
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b)
            for (int c = 0; c < n; ++c) {
              const double g = gauss_at(p, coord(0, i, a), coord(1, j, b), coord(2, k, c));
              const int64_t idx = e * np + (a * n + b) * n + c;
              for (int f = 0; f < 4; ++f) (*st)[f * fs + idx] = amp[f] * g;
            }
      }
  // halo face layers: the node layer of the halo element that touches the block
  for (int face = 0; face < 6; ++face) {
    const int ax = 2 - face / 2;
    const bool high = face % 2;
    const int bx = (ax == 0) ? 1 : 0, cx = (ax == 2) ? 1 : 2;
    const double w = high ? m.width[ax][m.nE[ax] - 1] : m.width[ax][0];
    const double hc = high ? ec[ax][m.nE[ax] - 1] + m.width[ax][m.nE[ax] - 1] : m.corner[ax] - m.width[ax][0];
    const double xa = w * (m.nodes[high ? 0 : n - 1] / 2.0 + 0.5) + hc;
    traces[face].assign((size_t)m.nE[bx] * m.nE[cx] * n * n * 3, 0.0);
    for (int eb = 0; eb < m.nE[bx]; ++eb)
      for (int ecc = 0; ecc < m.nE[cx]; ++ecc)
        for (int b = 0; b < n; ++b)
          for (int c = 0; c < n; ++c) {
            double xyz[3];
            xyz[ax] = xa;
            xyz[bx] = coord(bx, eb, b);
            xyz[cx] = coord(cx, ecc, c);
            const double g = gauss_at(p, xyz[0], xyz[1], xyz[2]);
            double* t = &traces[face][(((size_t)eb * m.nE[cx] + ecc) * n * n + b * n + c) * 3];
            t[0] = amp[0] * g;
            t[1] = amp[1 + ax] * g;
            t[2] = 0.0;
          }
  }
}

double cfl_dt(const Mesh& m, double c0, double cfl) {
  double dxi = 2.0;
  for (int i = 0; i + 1 < m.n(); ++i) dxi = std::min(dxi, m.nodes[i + 1] - m.nodes[i]);
  double inv = 0.0;
  for (int a = 0; a < 3; ++a) inv += 1.0 / *std::min_element(m.width[a].begin(), m.width[a].end());
  return cfl * 0.5 * dxi / (inv * c0);
}

}  // namespace dg3d_host
