/*
 * oracle.h — CPU restatement of drewalan/3DDG's per-RK-stage right-hand side and update.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP implementation: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The product
 * (3ddg_amd/) never links, calls or falls back to it.
 *
 * It restates the reference's data structures and loops: elements with a one-element halo
 * layer (Block.cpp:139-267), node arrays per element (Node.h:77-83), the per-line dense solve
 * of Element::OperateUFromV (Element.cpp:475-620), the LDG operators (Element.cpp:623-758),
 * the sponge (Element.cpp:760-781), the RK algebra (Element.cpp:267-307, Block.cpp:297-340)
 * and Grid::RK33's single-block branch (Grid.cpp:642-835). Two arithmetic policies:
 *   literal = FP32 state and FP32 scalar constants with FP64 line algebra (the reference);
 *   fp64    = FP64 everywhere (the FP64 drop-in's definition).
 * Extensions the reference lacks (LSRK54, boundary source, quadratic ramp, per-element
 * material and widths) are defined in the same structures.
 */
#ifndef DG3D_ORACLE_H
#define DG3D_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle;

typedef struct oracle_desc {
  int order;
  int num_elements[3];
  int periodic[3];
  int literal;             /* 1 = FP32 state/constants (reference), 0 = FP64 */
  int num_registers;
  const double* mass;      /* n*n unit-interval M */
  const double* stiff;     /* n*n S */
  const double* lgl_nodes; /* n */
  const double* lgl_weights; /* n */
  double corner[3];        /* STARTING_CORNER */
  const double* width[3];  /* per-axis element widths (nE[a] each) */
} oracle_desc;

oracle* oracle_create(const oracle_desc* d);
void oracle_destroy(oracle* o);
void oracle_set_threads(int nthreads);

/* Per-element material over interior elements (E values, element order); halo elements take
 * the material of the nearest interior element. NULL keeps the current value. */
void oracle_set_material(oracle* o, const double* c0, const double* rho0, const double* ba,
                         const double* sigma, const double* diff);

/* Node coordinates of every node (interior and halo) as Element.cpp:325-348 computes them. */
void oracle_node_coord(const oracle* o, int ei, int ej, int ek, int node, double* xyz);

/* Gaussian initial condition (Node.cpp:76-101) on every node including halos. */
void oracle_fill_gaussian(oracle* o, int reg, const int* gauss_xyz, double amp_rho,
                          const double* amp_vel, double ic_sigma, const double* center);

void oracle_set_state(oracle* o, int reg, const double* host, int nfields);
void oracle_get_state(const oracle* o, int reg, double* host, int nfields);
/* trace layout as dg3d_set_halo: [faceElem][n*n][3] = (rho', u_axis, Q_axis); NULL = zero */
void oracle_set_halo(oracle* o, int reg, int face, const double* trace);
void oracle_get_halo(const oracle* o, int reg, int face, double* trace);

void oracle_zero(oracle* o, int reg);
void oracle_copy(oracle* o, int dst, int src);
void oracle_axpby(oracle* o, int out, double self_scalar, int in, double other_scalar);
/* op: 0 UFromV, 1 Sponge, 2 QFromDRho, 3 DRhoFromQ (as dg3d_operate) */
void oracle_operate(oracle* o, int op, int axis, int mode, int in_reg, int out_reg);
void oracle_communicate(oracle* o, int reg);

/* Exterior source (as dg3d_source); lead may be NULL. Applied while evaluating at time t. */
void oracle_set_source(oracle* o, int face, double amplitude, double frequency, double phase,
                       double t0, double tau, const double* lead);
void oracle_set_time(oracle* o, double t);

/* Grid::RK33 (Grid.cpp:642-835): new_reg = RK33(old_reg) using k1, k2 registers. */
void oracle_rk33(oracle* o, double dt, int new_reg, int old_reg, int k1, int k2, int eval_advect,
                 int eval_diff, int eval_sponge);
/* `steps` of Grid::RK33 (advection and sponge, no diffusion) on each of `count` independent
 * blocks, one std::thread per block with one OpenMP thread each: the reference's one block per
 * single-threaded MPI rank, without the halo exchange. Registers 0..3 of every block are used. */
void oracle_rk33_blocks(oracle* const* blocks, int count, double dt, int steps);
/* Grid::forwardEuler (Grid.cpp:567-640): new_reg = old_reg + dt L(old_reg), with the LDG pair on
 * the x axis only, advection and sponge always on. Writes Q of old_reg. */
void oracle_forward_euler(oracle* o, double dt, int new_reg, int old_reg);
/* One LSRK(5,4) step on q_reg with du_reg and a scratch register rhs_reg. */
void oracle_lsrk54(oracle* o, double dt, int q_reg, int du_reg, int rhs_reg, int eval_advect,
                   int eval_diff, int eval_sponge);

/* max|rho'| NaN-sticky (HDF5Functions.cpp:201), max|u|,|v|,|w|; energy; int rho' dV */
void oracle_maxabs(const oracle* o, int reg, double* out4);
void oracle_energy(const oracle* o, int reg, double* out2);

/* Sponge sigma restated from Block.cpp:163-233 (ramp 0 linear, 1 quadratic). */
void oracle_sponge_sigma(const int* num_elements, const int* num_blocks, const int* block_pos,
                         const double* elem_width, const int* sponge_width, double sigma_max,
                         int ramp, double* sigma);

/* Per-line solve used by OperateUFromV, exposed for the Matrix.h cross-check:
 * x = (M q w)^{-1} (S q f - flux) in FP64 with partial-pivoting elimination. */
void oracle_line_solve(int n, const double* mass, const double* stiff, double q, double w,
                       const double* f, const double* flux, double* x);

#ifdef __cplusplus
}
#endif
#endif
