/*
 * dg3d.h — C ABI of the MI355X (gfx950) implementation of 3DDG's per-RK-stage
 * right-hand side and update for the first-order Westervelt-like acoustic system
 * (rho', u, v, w) on tensor-product LGL hexahedra.
 *
 * The reference has no FFI layer: its call surface is C++ methods. Each entry point
 * below names the reference method it replaces (file:line in drewalan/3DDG cppsrc/).
 * A C++ host facade (3ddg_amd/csrc/host/) with the reference's Grid/Block/Element
 * names forwards to these functions; Python binds them with ctypes.
 *
 * Conventions
 *  - Every function returns int status: DG3D_OK (0) or a negative DG3D_E_* code; the
 *    message is available from dg3d_last_error(ctx). No C++ exception crosses the ABI.
 *  - Host arrays are borrowed for the duration of the call only.
 *  - A context is used from one host thread at a time. Calls are ordered on the
 *    context's HIP stream; dg3d_get_state, dg3d_reduce and dg3d_sync synchronise.
 *  - Host state layout (the reference's mElements[i][j][k].mNodes[a][b][c] order):
 *      field-major: rho', u, v, w [, Qx, Qy, Qz];
 *      within a field: element (i,j,k) with k fastest, then node (a,b,c) with c fastest.
 *    i/a run along x, j/b along y, k/c along z (Element.cpp:325-348).
 *  - Face/direction numbering is the reference's neighbour order (Element.h:144):
 *      0 = k-1 (-z), 1 = k+1 (+z), 2 = j-1 (-y), 3 = j+1 (+y), 4 = i-1 (-x), 5 = i+1 (+x).
 */
#ifndef DG3D_H
#define DG3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Incremented whenever a structure layout or an entry point's signature changes; a caller compares
 * it with dg3d_abi_version() of the library it loads. */
#define DG3D_ABI_VERSION 1

/* status codes */
#define DG3D_OK 0
#define DG3D_E_INVALID (-1)     /* bad argument or call order */
#define DG3D_E_HIP (-2)         /* HIP runtime error */
#define DG3D_E_NOMEM (-3)       /* device allocation failed */
#define DG3D_E_NAN (-4)         /* NaN/Inf detected by a checked reduction */
#define DG3D_E_UNSUPPORTED (-5) /* order/precision/option not built */
#define DG3D_E_IO (-6)          /* file missing, unreadable or malformed */

/* precision of the fused step (dg3d_step) */
#define DG3D_FP64 0  /* FP64 state, FP64 RHS: the drop-in */
#define DG3D_MIXED 1 /* FP64 state, FP32 RHS and FP32 LSRK register */

/* derivative contraction used by the MIXED step */
#define DG3D_CONTRACT_VALU 0      /* FP32 FMA on the vector ALUs */
#define DG3D_CONTRACT_MFMA_F32 1  /* v_mfma_f32_4x4x1_16b_f32, one line per lane */
#define DG3D_CONTRACT_MFMA_BF16 2 /* split-BF16 (3 pieces per operand), FP32 accumulate */

/* time schemes */
#define DG3D_RK33 0   /* Shu-Osher SSP-RK3 as Grid::RK33 (Grid.cpp:642-902) */
#define DG3D_LSRK54 1 /* Carpenter-Kennedy 5-stage 4th-order 2N-storage */

/* operators of dg3d_operate (Block.h:49-52) */
#define DG3D_OP_UFROMV 0    /* Block::OperateUFromV(axis, mode, old) */
#define DG3D_OP_SPONGE 1    /* Block::OperateSponge(0, 0, old) */
#define DG3D_OP_QFROMDRHO 2 /* Block::OperateQFromDRho(axis, 0, old) */
#define DG3D_OP_DRHOFROMQ 3 /* Block::OperateDRhoFromQ(axis, 0, old) */

/* reductions of dg3d_reduce */
#define DG3D_REDUCE_MAXABS 0 /* out[0..3] = max|rho'|,|u|,|v|,|w|; NaN-sticky like HDF5Functions.cpp:201 */
#define DG3D_REDUCE_ENERGY 1 /* out[0] = acoustic energy (LGL quadrature), out[1] = int rho' dV */
#define DG3D_REDUCE_RELL2 2  /* out[f] = ||reg - reg2||_2 / ||reg2||_2 per field f=0..3 */
#define DG3D_REDUCE_DIFFL2 3 /* out[f] = ||reg - reg2||_2, out[4+f] = ||reg2||_2 */

typedef struct dg3d_ctx dg3d_ctx;

typedef struct dg3d_desc {
  int order;               /* N (1..7): n = N+1 LGL nodes per axis */
  int num_elements[3];     /* interior elements per axis of the block (x, y, z); a field of the
                              block may take at most 2^29 - 1 node slots: the elements x (order+1)^3
                              nodes, each element's rounded up to a multiple of 32 at 3, 5, 6 and 7
                              nodes per line, and the gap at the end of each group of elements (up
                              to 14 %) */
  int periodic[3];         /* BC_IS_PERIODIC (driver.cpp:170) */
  int precision;           /* DG3D_FP64 or DG3D_MIXED (for dg3d_step) */
  int contraction;         /* DG3D_CONTRACT_* (MIXED only) */
  int num_registers;       /* state registers (Grid copies); >= 1 */
  int with_q;              /* allocate the LDG fields Qx, Qy, Qz (EVALUATE_DIFFUSION) */
  int xcd_remap;           /* workgroup order before autotuning: 0 hardware round-robin over the
                              8 XCDs, 1 one contiguous slab per XCD, 2 runs of consecutive
                              element groups per XCD sized to the element rows, 3 slabs of a
                              traversal tiled over 8 rows x 8 groups (L2 reuse of the traces) */
  const double* mass;      /* n*n row-major unit-interval mass M (matrices.txt, Grid.cpp:979) */
  const double* stiff;     /* n*n row-major weak stiffness S (matrices.txt, Grid.cpp:980) */
  const double* lgl_nodes; /* n LGL nodes on [-1,1] (hardcoded_coeff.txt, Basis.cpp:142-153) */
  const double* lgl_weights; /* n LGL weights (hardcoded_coeff.txt, Basis.cpp:154-165) */
} dg3d_desc;

/* Exterior forcing on one block face: the halo state of that face becomes
 * p = s(t + lead), u_normal = s/(rho0 c) (entering the block), rho' = s/c^2, with
 * s(t) = amplitude * sin(2 pi f t + phase) * exp(-((t - t0)/tau)^2).
 * No counterpart in the reference (it has initial conditions only). */
typedef struct dg3d_source {
  int face;            /* direction 0..5, or -1 for none */
  double amplitude;    /* Pa */
  double frequency;    /* Hz */
  double phase;        /* rad */
  double t0;           /* envelope centre, s */
  double tau;          /* envelope width, s */
  const double* lead;  /* per face node time lead in s [faceElems][n*n], or NULL */
} dg3d_source;

/* ---- lifecycle ------------------------------------------------------------ */

/* Replaces Grid::Grid + Grid::readMatrices (Grid.cpp:176-375, 920-989): allocates the
 * block's device registers and precomputes the line operator from M and S. */
int dg3d_create(const dg3d_desc* desc, dg3d_ctx** out);
void dg3d_destroy(dg3d_ctx* ctx);
const char* dg3d_last_error(const dg3d_ctx* ctx);
/* Library-level message for a failed dg3d_create (ctx not available). */
const char* dg3d_create_error(void);
int dg3d_abi_version(void);

/* ---- setup ---------------------------------------------------------------- */

/* Element geometry (Element.cpp:278-336, Block.cpp:172-210). corner = STARTING_CORNER;
 * widths per axis (nE[a] values each): the reference uses one ELEMENT_WIDTH per axis
 * (driver.cpp:161); per-element values along an axis give stretched meshes. */
int dg3d_set_geometry(dg3d_ctx* ctx, const double* corner, const double* wx, const double* wy,
                      const double* wz);

/* Per-element material and sponge (Element.h:146-151; Block.cpp:175-233 for sigma).
 * Arrays of E values in element order; NULL leaves the current values. */
int dg3d_set_material(dg3d_ctx* ctx, const double* c0, const double* rho0, const double* b_over_a,
                      const double* sigma, const double* diffusion);

/* ---- registers (Grid copies) ---------------------------------------------- */

/* Write / read a register's interior state (host layout above; nfields 4 or 7). */
int dg3d_set_state(dg3d_ctx* ctx, int reg, const double* host, int nfields);
int dg3d_get_state(dg3d_ctx* ctx, int reg, double* host, int nfields);

/* Halo of a register on one face: trace[faceElem][n*n][3] = (rho', u_normal, Q_normal) of
 * the face node layer of the halo element (what Element::OperateUFromV reads through
 * neighbors[dir], Element.cpp:531-600). faceElem enumerates the face's elements in (i,j,k)
 * order with the face axis left out; face nodes likewise. NULL = zero state. */
int dg3d_set_halo(dg3d_ctx* ctx, int reg, int face, const double* trace);

/* Block::setZero (Block.cpp:297-310): zero state and halo. */
int dg3d_zero(dg3d_ctx* ctx, int reg);
/* Grid copy constructor (Grid.cpp:78-174): dst = src, halo included. */
int dg3d_copy(dg3d_ctx* ctx, int dst, int src);
/* Block::selfWeightedAdd (Block.cpp:327-340): out = out*self_scalar + in*other_scalar
 * over the interior, all allocated fields. */
int dg3d_axpby(dg3d_ctx* ctx, int out, double self_scalar, int in, double other_scalar);
/* Block::Operate* (Block.cpp:343-403): out_reg accumulates op(in_reg) for (axis, mode). */
int dg3d_operate(dg3d_ctx* ctx, int op, int axis, int mode, int in_reg, int out_reg);
/* Grid::communicate (Grid.cpp:487-565). One block on one GPU: periodic faces read the
 * opposite interior directly, so no data moves; non-periodic halos keep their state. */
int dg3d_halo(dg3d_ctx* ctx, int reg);

/* ---- fused path ----------------------------------------------------------- */

int dg3d_set_source(dg3d_ctx* ctx, const dg3d_source* src);

/* Advance register 0 by nsteps steps of the given scheme (one fused kernel per stage).
 * Replaces the driver's step loop over Grid::RK33 (driver.cpp:256-277). *t is the time of
 * register 0 on entry and is advanced by nsteps*dt. The halo of register 0 (set by
 * dg3d_set_halo) is read by the first stage only, as the reference reads the initial
 * grid's halo on step 1 stage 1 and zeroed registers afterwards (Grid.cpp:648-650). */
int dg3d_step(dg3d_ctx* ctx, int scheme, double dt, int nsteps, double* t);

/* One stage of a scheme on register 0 (stage = 0..stages-1), for stage-level tests. */
int dg3d_stage(dg3d_ctx* ctx, int scheme, int stage, double dt, double t);

/* Autotuner of the fused stage kernel (no counterpart in the reference): times every kernel
 * variant (0..7: elements per workgroup, occupancy target, RK-register loads issued with the
 * node's own loads or after the first barrier) under every workgroup order
 * (0..3, as xcd_remap above) on stage launches that pass a pseudo-random state between two
 * work registers, through the scheme's stage coefficients in turn, and keeps the fastest pair
 * for `scheme`; register 0 and the solution are left unchanged. dg3d_step runs it on its first call
 * unless a pair was pinned with dg3d_set_variant. With the VALU contraction the results do not
 * depend on the pair (the per-node arithmetic is the same). */
int dg3d_autotune(dg3d_ctx* ctx, int scheme, int* variant, int* xcd_remap);
int dg3d_set_variant(dg3d_ctx* ctx, int scheme, int variant, int xcd_remap);

/* The workgroup order of the stage kernels (no counterpart in the reference): out[w] = element
 * group of workgroup w, for a launch of `groups` groups of elems_per_group elements over a block
 * of num_elements[3], under workgroup order `order` (0..3, as xcd_remap). A host function, for
 * checking that every order visits every group once. */
int dg3d_group_order(int order, const int* num_elements, int elems_per_group, int64_t groups, int64_t* out);

/* CFL time step in FP64 on the device: dt = cfl / max_e( c(e) * sum_a 2 / (w_a(e) * dxi_min) ),
 * dxi_min the smallest LGL node gap on [-1,1] (the d-dimensional node-spacing CFL). */
int dg3d_cfl_dt(dg3d_ctx* ctx, double cfl, double* dt);

/* Reductions over a register (see DG3D_REDUCE_*); out holds up to 8 doubles. */
int dg3d_reduce(dg3d_ctx* ctx, int kind, int reg, int reg2, double* out);

/* Allocation-free variant of dg3d_reduce. No counterpart in the reference: it is for callers
 * that may not allocate (serving loops, harnesses that forbid allocation). Same arguments and
 * results, bit for bit, as dg3d_reduce. The per-block partials and combined values live in a
 * file-scope device array of the library, 8200 doubles (64.1 KiB), one copy per device, held
 * until the process ends; its address is resolved when a context is created on that device. The
 * call allocates no device or host memory; its error messages go into space the context
 * reserved. A grid-stride reduction of at most 1024 blocks needs at most 8192 partials, so every
 * register fits; a reduction needing more is rejected with DG3D_E_INVALID. Calls that overlap on
 * one device (contexts on one GPU used from several host threads) race on the array; calls on
 * one context are ordered by its stream. Synchronises, as dg3d_reduce does. */
int dg3d_reduce_noalloc(dg3d_ctx* ctx, int kind, int reg, int reg2, double* out);

/* Probe nodes: after every step of dg3d_step the pressure P(rho') at each probe node
 * (global node index into a field) is recorded. */
int dg3d_set_probes(dg3d_ctx* ctx, int count, const int64_t* node_index, int capacity_steps);
int dg3d_get_probes(dg3d_ctx* ctx, double* out, int* steps_recorded);

int dg3d_sync(dg3d_ctx* ctx);
/* Launch mode of dg3d_step (no counterpart in the reference; default on): 1 replays each step
 * as a captured HIP graph, one per (scheme, dt, state buffer), captured on first use and
 * dropped by dg3d_set_variant, dg3d_autotune, dg3d_set_material, dg3d_set_source and
 * dg3d_set_probes; 0 launches every kernel directly. The results do not depend on the mode. */
int dg3d_set_graphs(dg3d_ctx* ctx, int enable);

/* Average device time of the last dg3d_step's stage kernels in ms (HIP events on the
 * context stream), and the number of stage launches it averaged. */
int dg3d_last_step_timing(dg3d_ctx* ctx, double* avg_stage_ms, int* launches);

/* ---- batched solves (no counterpart in the reference) -------------------------------------
 * Many independent blocks of different element counts and orders, each what one dg3d_ctx
 * would hold (the reference builds one such block per MPI rank, Grid.cpp:176-375), advanced
 * together by LSRK54: per stage, one persistent launch per order present covers all solves of
 * that order (the launches run as parallel graph branches), whole steps are replayed from HIP
 * graphs, and elements per workgroup are autotuned per order. The first stage of the first
 * dg3d_batch_step reads the halo traces given here (as dg3d_set_halo on register 0 before
 * dg3d_step); later stages see zero halos. A solve may carry a boundary source (as
 * dg3d_set_source), evaluated on the solve's own clock from t = 0. Each solve follows the
 * trajectory of dg3d_step on a context of its own with the same precision (same per-node
 * arithmetic). */
typedef struct dg3d_batch dg3d_batch;
typedef struct dg3d_solve_desc {
  int order;                 /* 1..7 */
  int num_elements[3];
  int periodic[3];
  const double* mass;        /* (order+1)^2, as dg3d_desc */
  const double* stiff;
  const double* widths[3];   /* element widths per axis (num_elements[a] each) */
  const double* c0;          /* per element, E = product of num_elements */
  const double* rho0;
  const double* b_over_a;    /* NULL = 0 */
  const double* sigma;       /* NULL = 0 */
  const double* state;       /* initial rho', u, v, w: [4][E][(order+1)^3], as dg3d_set_state */
  const double* halo[6];     /* first-stage halo traces per face (as dg3d_set_halo), or NULL */
  double dt;                 /* time step of this solve */
  /* boundary tone burst on one non-periodic face, as dg3d_set_source (source_face -1: none);
   * the solve's time starts at 0 */
  int source_face;
  double source_amplitude, source_frequency, source_phase, source_t0, source_tau;
  const double* source_lead; /* per face node delay [faceElem][n*n], or NULL */
} dg3d_solve_desc;

int dg3d_batch_create(int count, const dg3d_solve_desc* solves, int precision, dg3d_batch** out);
const char* dg3d_batch_create_error(void);
int dg3d_batch_destroy(dg3d_batch* batch);
const char* dg3d_batch_last_error(const dg3d_batch* batch);
/* Variant per order, 0..7: bit 0 the elements-per-workgroup shape (the stage kernel's two),
 * bit 1 one element group per workgroup instead of the persistent loop, bit 2 RK-register loads
 * after the first barrier. dg3d_batch_autotune writes one variant per order, in
 * dg3d_batch_info's order. */
int dg3d_batch_autotune(dg3d_batch* batch, int* variants);
int dg3d_batch_set_variant(dg3d_batch* batch, int variant);
/* How a stage runs the orders: 0 every order on a parallel branch of its own, 1 all orders one
 * after another on one stream (largest first, each with the device to itself), 2 two branches
 * with about the same number of nodes, 3 the largest order on one branch and the others one after
 * another on a second beside it, 4 every order in one persistent launch (one kernel whose
 * workgroups take element groups of any order). dg3d_batch_autotune (and the first
 * dg3d_batch_step) times the five on whole stages after choosing the variants; results do not
 * depend on it. */
int dg3d_batch_set_schedule(dg3d_batch* batch, int schedule);
int dg3d_batch_get_schedule(const dg3d_batch* batch, int* schedule);
/* nsteps LSRK54 steps of every solve, each with its own dt; synchronises. */
int dg3d_batch_step(dg3d_batch* batch, int nsteps);
/* State of one solve in the dg3d_get_state layout (4 fields). */
int dg3d_batch_get_state(dg3d_batch* batch, int solve, double* host);
/* Nodes over all solves; per order present (descending, up to 7 entries per array): the order,
 * its tuned variant (-1 before tuning), its element-group count and its persistent grid. */
int dg3d_batch_info(const dg3d_batch* batch, int64_t* total_nodes, int* num_orders, int* orders, int* variants,
                    int* tasks, int* grids);
/* Average stage time of the last dg3d_batch_step (HIP events around the whole run). */
int dg3d_batch_last_step_timing(const dg3d_batch* batch, double* avg_stage_ms, int* launches);

/* ---- host-side setup helpers (no device work) ------------------------------ */

/* LGL basis for order N generated in FP64 (replaces the missing hardcoded_coeff*.txt and
 * matrices.txt inputs of Basis.cpp:74-237 and Grid.cpp:920-989). Arrays of n or n*n. */
int dg3d_basis_generate(int order, double* nodes, double* weights, double* mass, double* stiff,
                        double* dpsi);
/* Write hardcoded_coeff.txt, hardcoded_coeff2.txt and matrices.txt into dir. */
int dg3d_basis_write_files(const char* dir, int max_order);
/* Read order N from the reference's files (Basis.cpp:130-171, Grid.cpp:940-989). */
int dg3d_basis_read_files(const char* coeff_path, const char* matrices_path, int order,
                          double* nodes, double* weights, double* mass, double* stiff);

/* Sponge sigma per element (Block.cpp:163-233): the block at block_pos among num_blocks,
 * each with nE elements per axis; widths are the element widths per axis; ramp 0 = linear
 * max(0,1-d)*sigma_max (the reference), 1 = quadratic max(0,1-d)^2*sigma_max. */
int dg3d_sponge_sigma(const int* num_elements, const int* num_blocks, const int* block_pos,
                      const double* elem_width, const int* sponge_width, double sigma_max,
                      int ramp, double* sigma);

#ifdef __cplusplus
}
#endif
#endif /* DG3D_H */
