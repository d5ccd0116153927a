"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same seeded inputs.

Bar: relative L2 <= 1e-12 per field for the FP64 drop-in (fine-grained operators, fused RK33 and
LSRK54 steps over whole runs of small configurations), <= 1e-5 for the mixed-precision step.
"""
import numpy as np
import pytest

from helpers import dg3d, heterogeneous, problem, random_halos, random_state, rel_l2, small_c1, small_c2, small_c5
from oracle_bind import Oracle, load_problem_into

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_MIXED = 1e-5


def _pair(p, nreg=4, with_q=False):
    ctx = problem.make_context(p, num_registers=nreg, with_q=with_q, halos=False)
    o = Oracle(p, literal=False, num_registers=nreg)
    return ctx, o


def _load_same(ctx, o, p, st, halos, reg=0):
    ctx.set_state(reg, st)
    o.set_state(reg, st)
    for f, tr in enumerate(halos):
        ctx.set_halo(reg, f, tr)
        o.set_halo(reg, f, tr)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("order", [1, 3, 4, 7])
def test_operate_ufromv(axis, mode, order):
    print(f"operate_ufromv order={order} axis={axis} mode={mode}", flush=True)
    p = heterogeneous(small_c2(n_e=3, order=order))
    p.num_elements = (3, 4, 2)
    p.widths = [np.full(3, 1e-3), np.linspace(0.8e-3, 1.2e-3, 4), np.full(2, 0.9e-3)]
    E = 24
    rng = np.random.default_rng(order)
    p.c0, p.rho0, p.b_over_a, p.sigma, p.diffusion = (1500 + 100 * rng.random(E), 1000 + 200 * rng.random(E),
                                                       np.where(rng.random(E) < 0.3, 0, 8 * rng.random(E)),
                                                       1e4 * rng.random(E), np.zeros(E))
    p.source = problem.Source()
    ctx, o = _pair(p, nreg=2)
    st = rng.standard_normal(4 * p.dof) * np.repeat([1e-3, 1e-4, 1e-4, 1e-4], p.dof)
    halos = random_halos(p)
    _load_same(ctx, o, p, st, halos)
    ctx.zero(1)
    o.zero(1)
    ctx.operate(dg3d.OP_UFROMV, axis, mode, 0, 1)
    o.operate(0, axis, mode, 0, 1)
    g = ctx.get_state(1)
    r = o.get_state(1)
    field = 1 + axis if mode == 0 else 0
    err = rel_l2(g, r)
    assert np.linalg.norm(r.reshape(4, -1)[field]) > 0
    assert err[field] < TOL_FP64, err
    others = [f for f in range(4) if f != field]
    assert np.abs(g.reshape(4, -1)[others]).max() == 0.0


def test_operate_sponge_and_axpby():
    print("operate_sponge_and_axpby", flush=True)
    p = heterogeneous(small_c1())
    p.sigma = np.random.default_rng(5).random(p.E) * 1e4
    ctx, o = _pair(p, nreg=3)
    st = random_state(p)
    _load_same(ctx, o, p, st, [None] * 6)
    st2 = random_state(p, seed=9)
    _load_same(ctx, o, p, st2, [None] * 6, reg=2)
    for x in (ctx, o):
        x.zero(1)
    ctx.operate(dg3d.OP_SPONGE, 0, 0, 0, 1)
    o.operate(1, 0, 0, 0, 1)
    assert rel_l2(ctx.get_state(1), o.get_state(1)).max() < 1e-15
    ctx.axpby(1, 0.37, 2, -1.25)
    o.axpby(1, 0.37, 2, -1.25)
    assert rel_l2(ctx.get_state(1), o.get_state(1)).max() < 1e-15


def _run_both(p, scheme, steps, precision=dg3d.FP64):
    """Run the fused GPU step and the oracle from the reference-built initial grid."""
    ctx = problem.make_context(p, precision=precision)
    o = Oracle(p, literal=False, num_registers=4)
    load_problem_into(o, p, 0)
    dt = problem.cfl_dt_host(p)
    t = ctx.step(scheme, dt, steps, 0.0)
    o.set_time(0.0)
    for _ in range(steps):
        if scheme == dg3d.RK33:
            o.rk33(dt, 3, 0, 1, 2)
            o.copy(0, 3)
        else:
            o.lsrk54(dt, 0, 1, 2)
    assert abs(t - steps * dt) < 1e-12 * steps * dt
    return ctx.get_state(0), o.get_state(0), ctx, o


def test_rk33_c1_fused_vs_oracle():
    print("rk33_c1_fused_vs_oracle", flush=True)
    p = small_c1()
    g, r, _, _ = _run_both(p, dg3d.RK33, 20)
    err = rel_l2(g, r)
    assert err.max() < TOL_FP64, err


def test_rk33_facade_vs_fused():
    """Grid::RK33 through fine-grained Block operators equals the fused stage kernels."""
    print("rk33_facade_vs_fused", flush=True)
    p = heterogeneous(small_c2(n_e=4, order=3))
    ctx = problem.make_context(p, num_registers=4)
    ctx2 = problem.make_context(p, num_registers=1)
    halos = random_halos(p, scale=1e-4)
    st = random_state(p, scale=1e-3)
    for c in (ctx, ctx2):
        c.set_state(0, st)
        for f, tr in enumerate(halos):
            c.set_halo(0, f, tr)
    dt = problem.cfl_dt_host(p)
    # the source is evaluated at the stage times by both paths only in the fused path; drop it
    for c in (ctx, ctx2):
        c.set_source(-1)
    grids = [pkg_grid(ctx, r) for r in range(4)]
    old, new, k1, k2 = grids
    for s in range(3):
        new.RK33(dt, old, k1, k2)
        old, new = new, old
    ctx2.step(dg3d.RK33, dt, 3, 0.0)
    err = rel_l2(ctx.get_state(old.reg), ctx2.get_state(0))
    assert err.max() < 1e-13, err


def pkg_grid(ctx, reg):
    import importlib
    return importlib.import_module("3ddg_amd").grid.Grid(ctx, reg)


def test_lsrk54_c1_full_run():
    """The reference's CPU-sized workload on the GPU: 8^3 elements, N=3, Gaussian pulse, 100
    LSRK54 steps vs the oracle."""
    print("lsrk54_c1_full_run", flush=True)
    p = small_c1(steps=100)
    g, r, ctx, o = _run_both(p, dg3d.LSRK54, p.steps)
    err = rel_l2(g, r)
    assert err.max() < TOL_FP64, err
    e_g = ctx.reduce(dg3d.REDUCE_ENERGY, 0)
    e_o = o.energy(0)
    assert abs(e_g[0] - e_o[0]) < 1e-12 * e_o[0]


@pytest.mark.parametrize("order", [2, 4, 5])
def test_lsrk54_burst_nonlinear(order):
    """Water, B/A=5, z=0 tone burst, sponge: small version of the 64^3 workload."""
    print(f"lsrk54_burst_nonlinear order={order}", flush=True)
    p = small_c2(n_e=4, order=order, steps=150)
    g, r, _, _ = _run_both(p, dg3d.LSRK54, p.steps)
    err = rel_l2(g, r)
    assert np.linalg.norm(r) > 0
    assert err.max() < TOL_FP64, err


def test_lsrk54_stretched_two_layer_focused():
    print("lsrk54_stretched_two_layer_focused", flush=True)
    p = small_c5(n_e=5, steps=150)
    g, r, _, _ = _run_both(p, dg3d.LSRK54, p.steps)
    err = rel_l2(g, r)
    assert err.max() < TOL_FP64, err


def test_periodic_box():
    print("periodic_box", flush=True)
    p = small_c1()
    p.periodic = (1, 1, 1)
    p.sigma[:] = 0.0
    g, r, ctx, o = _run_both(p, dg3d.LSRK54, 60)
    err = rel_l2(g, r)
    assert err.max() < TOL_FP64, err


def test_mixed_precision_c1():
    print("mixed_precision_c1", flush=True)
    p = small_c1(steps=100)
    ctx64 = problem.make_context(p, precision=dg3d.FP64)
    ctx32 = problem.make_context(p, precision=dg3d.MIXED)
    dt = problem.cfl_dt_host(p)
    ctx64.step(dg3d.LSRK54, dt, p.steps, 0.0)
    ctx32.step(dg3d.LSRK54, dt, p.steps, 0.0)
    err = rel_l2(ctx32.get_state(0), ctx64.get_state(0))
    assert err.max() < TOL_MIXED, err
    assert err.max() > 0.0  # it is a different arithmetic


def test_reductions_and_cfl():
    print("reductions_and_cfl", flush=True)
    p = heterogeneous(small_c1())
    ctx, o = _pair(p, nreg=2)
    st = random_state(p)
    _load_same(ctx, o, p, st, [None] * 6)
    m = ctx.reduce(dg3d.REDUCE_MAXABS, 0)
    assert np.array_equal(m[:4], o.maxabs(0))
    e = ctx.reduce(dg3d.REDUCE_ENERGY, 0)
    eo = o.energy(0)
    assert abs(e[0] - eo[0]) < 1e-12 * eo[0] and abs(e[1] - eo[1]) < 1e-12 * abs(eo[1]) + 1e-30
    assert abs(ctx.cfl_dt(0.3) - problem.cfl_dt_host(p)) < 1e-15 * problem.cfl_dt_host(p)
    st[12345] = np.nan  # NaN-sticky max (HDF5Functions.cpp:201)
    ctx.set_state(0, st)
    assert np.isnan(ctx.reduce(dg3d.REDUCE_MAXABS, 0)[0])
    ctx.set_state(1, st)
    d = ctx.reduce(dg3d.REDUCE_DIFFL2, 0, 1)
    assert np.isnan(d[0])


@pytest.mark.parametrize("precision", [dg3d.FP64, dg3d.MIXED])
@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
def test_kernel_variants_bitwise_identical(precision, scheme):
    """Every autotuner candidate (elements per workgroup, occupancy target, XCD numbering)
    computes the same per-node arithmetic, so the states agree bit for bit."""
    print(f"kernel_variants precision={precision} scheme={scheme}", flush=True)
    p = small_c2(n_e=5, order=4, steps=4)
    dt = problem.cfl_dt_host(p)
    ref = None
    for v in range(8):
        for x in range(4):
            ctx = problem.make_context(p, precision=precision)
            ctx.set_variant(scheme, v, x)
            ctx.step(scheme, dt, p.steps, 0.0)
            st = ctx.get_state(0)
            ctx.close()
            if ref is None:
                ref = st
            assert np.array_equal(st, ref), (v, x)
    ctx = problem.make_context(p, precision=precision)
    v, x = ctx.autotune(scheme)
    assert 0 <= v < 8 and 0 <= x < 4


@pytest.mark.parametrize("precision", [dg3d.FP64, dg3d.MIXED])
@pytest.mark.parametrize("order", [5, 6])
def test_variants_bitwise_identical_heterogeneous(precision, order):
    """At 6 and 7 nodes per line the two elements-per-workgroup shapes read the element material
    differently (at 6 nodes, one element per workgroup reads it once per wave through the scalar
    cache, two elements per workgroup per lane): with per-element c, rho0, B/A and sponge every
    variant still gives the same bits, and FP64 matches the oracle."""
    print(f"variants_heterogeneous precision={precision} order={order}", flush=True)
    p = heterogeneous(small_c2(n_e=5, order=order, steps=3), seed=11)
    st0 = random_state(p, seed=12)
    dt = problem.cfl_dt_host(p)
    ref = None
    for v in range(8):
        ctx = problem.make_context(p, precision=precision)
        ctx.set_state(0, st0)
        ctx.set_variant(dg3d.LSRK54, v, 1)
        ctx.step(dg3d.LSRK54, dt, p.steps, 0.0)
        st = ctx.get_state(0)
        ctx.close()
        if ref is None:
            ref = st
        assert np.array_equal(st, ref), v
    if precision == dg3d.FP64:
        o = Oracle(p, literal=False, num_registers=3)
        load_problem_into(o, p, 0)
        o.set_state(0, st0)
        o.set_time(0.0)
        for _ in range(p.steps):
            o.lsrk54(dt, 0, 1, 2)
        err = rel_l2(ref, o.get_state(0))
        assert err.max() < 1e-12, err


@pytest.mark.parametrize("contraction", [dg3d.CONTRACT_MFMA_F32, dg3d.CONTRACT_MFMA_BF16])
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7])
def test_mixed_matrix_core_contraction(contraction, order):
    """The mixed step with the derivative contractions on the matrix cores (FP32 operands, or
    three-piece split-BF16) stays within the mixed tolerance of the FP64 step, and within FP32
    rounding of the same step with VALU contractions."""
    print(f"mixed_matrix_core contraction={contraction} order={order}", flush=True)
    p = heterogeneous(small_c2(n_e=4, order=order, steps=60))
    dt = problem.cfl_dt_host(p)
    st = {}
    for name, prec, con in (("fp64", dg3d.FP64, dg3d.CONTRACT_VALU), ("valu", dg3d.MIXED, dg3d.CONTRACT_VALU),
                            ("mfma", dg3d.MIXED, contraction)):
        ctx = problem.make_context(p, precision=prec, contraction=con)
        ctx.step(dg3d.LSRK54, dt, p.steps, 0.0)
        st[name] = ctx.get_state(0)
        ctx.close()
    assert np.linalg.norm(st["fp64"]) > 0
    err64 = rel_l2(st["mfma"], st["fp64"])
    errv = rel_l2(st["mfma"], st["valu"])
    assert err64.max() < TOL_MIXED, err64
    assert errv.max() < 5e-6, errv


@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7])
def test_fp32_matrix_core_contraction_equals_valu(order, scheme):
    """The FP32 matrix-core contraction feeds each line's nodes to v_mfma_f32_4x4x1f32 in node
    order, so it accumulates the same chain of correctly rounded multiply-adds as the VALU
    contraction: the two mixed steps agree bit for bit, with either scheme."""
    print(f"fp32_matrix_core_equals_valu order={order} scheme={scheme}", flush=True)
    p = heterogeneous(small_c2(n_e=4, order=order, steps=20))
    dt = problem.cfl_dt_host(p)
    st = []
    for con in (dg3d.CONTRACT_VALU, dg3d.CONTRACT_MFMA_F32):
        ctx = problem.make_context(p, precision=dg3d.MIXED, contraction=con)
        ctx.step(scheme, dt, p.steps, 0.0)
        st.append(ctx.get_state(0))
        ctx.close()
    assert np.linalg.norm(st[0]) > 0
    assert np.array_equal(st[0], st[1]), np.abs(st[0] - st[1]).max()


@pytest.mark.parametrize("contraction", [dg3d.CONTRACT_MFMA_F32, dg3d.CONTRACT_MFMA_BF16])
@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
def test_matrix_core_variants(contraction, scheme):
    """Each variant is deterministic (two runs agree bit for bit), and the variants agree to FP32
    rounding (with one line per lane, as here at 5 nodes per line, a line's sums do not depend on
    the workgroup shape; the 16x16x16 split-BF16 tiles at 8 nodes per line place a line in one of
    two diagonal blocks by the shape, which changes the order of the matrix core's sums)."""
    print(f"matrix_core_variants contraction={contraction} scheme={scheme}", flush=True)
    p = small_c2(n_e=5, order=4, steps=4)
    dt = problem.cfl_dt_host(p)
    ref = None
    for v in range(8):
        for x in range(4):
            runs = []
            for _ in range(2 if x == 0 else 1):
                ctx = problem.make_context(p, precision=dg3d.MIXED, contraction=contraction)
                ctx.set_variant(scheme, v, x)
                ctx.step(scheme, dt, p.steps, 0.0)
                runs.append(ctx.get_state(0))
                ctx.close()
            if len(runs) == 2:
                assert np.array_equal(runs[0], runs[1]), (v, x)
            if ref is None:
                ref = runs[0]
            err = rel_l2(runs[0], ref)
            assert err.max() < 1e-6, (v, x, err)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("order", [1, 3, 4, 7])
def test_operate_ldg(axis, order):
    """Element::OperateQFromDRho (in place, Grid.cpp:667) then OperateDRhoFromQ into another
    register, on heterogeneous D, anisotropic widths and random halo traces (with Q)."""
    print(f"operate_ldg order={order} axis={axis}", flush=True)
    p = heterogeneous(small_c2(n_e=3, order=order))
    p.num_elements = (3, 4, 2)
    p.widths = [np.full(3, 1e-3), np.linspace(0.8e-3, 1.2e-3, 4), np.full(2, 0.9e-3)]
    E = 24
    rng = np.random.default_rng(10 + order)
    p.c0, p.rho0, p.b_over_a = 1500 + 100 * rng.random(E), 1000 + 200 * rng.random(E), 5.0 * rng.random(E)
    p.sigma, p.diffusion = np.zeros(E), 1.0 + 4.0 * rng.random(E)
    p.source = problem.Source()
    ctx, o = _pair(p, nreg=2, with_q=True)
    st = np.concatenate([rng.standard_normal(4 * p.dof) * np.repeat([1e-3, 1e-4, 1e-4, 1e-4], p.dof),
                         1e-2 * rng.standard_normal(3 * p.dof)])
    halos = random_halos(p)
    _load_same(ctx, o, p, st, halos)
    ctx.operate(dg3d.OP_QFROMDRHO, axis, 0, 0, 0)
    o.operate(2, axis, 0, 0, 0)
    g, r = ctx.get_state(0, nfields=7), o.get_state(0, nfields=7)
    err = rel_l2(g, r, nfields=7)
    assert np.linalg.norm(r.reshape(7, -1)[4 + axis]) > 0
    assert err[4 + axis] < TOL_FP64, err
    assert np.array_equal(g.reshape(7, -1)[:4], st.reshape(7, -1)[:4])  # rho', u, v, w untouched
    ctx.zero(1)
    o.zero(1)
    ctx.operate(dg3d.OP_DRHOFROMQ, axis, 0, 0, 1)
    o.operate(3, axis, 0, 0, 1)
    g1, r1 = ctx.get_state(1, nfields=7), o.get_state(1, nfields=7)
    err1 = rel_l2(g1, r1, nfields=7)
    assert np.linalg.norm(r1.reshape(7, -1)[0]) > 0
    assert err1[0] < TOL_FP64, err1


@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
@pytest.mark.parametrize("order", [2, 4])
def test_fused_stage_with_ldg_diffusion(scheme, order):
    """dg3d_step with diffusion on: per stage the Q kernel, then the fused stage with the
    OperateDRhoFromQ term, against the oracle's Grid::RK33 order (QFromDRho, DRhoFromQ,
    UFromV, sponge); heterogeneous D and material, B/A=5 burst source."""
    print(f"fused_ldg scheme={scheme} order={order}", flush=True)
    p = heterogeneous(small_c2(n_e=4, order=order, steps=40))
    rng = np.random.default_rng(order)
    p.diffusion = 1.0e-3 * (1.0 + rng.random(p.E))   # m^2/s at 0.75 mm elements: a visible term
    ctx = problem.make_context(p, with_q=True, num_registers=1)
    o = Oracle(p, literal=False, num_registers=4)
    load_problem_into(o, p, 0)
    dt = problem.cfl_dt_host(p)
    ctx.step(scheme, dt, p.steps, 0.0)
    o.set_time(0.0)
    for _ in range(p.steps):
        if scheme == dg3d.RK33:
            o.rk33(dt, 3, 0, 1, 2, adv=1, diff=1, sponge=1)
            o.copy(0, 3)
        else:
            o.lsrk54(dt, 0, 1, 2, adv=1, diff=1, sponge=1)
    g, r = ctx.get_state(0), o.get_state(0)
    err = rel_l2(g, r)
    assert np.linalg.norm(r) > 0
    assert err.max() < TOL_FP64, err
    # the diffusion term is not negligible in this run
    ctx2 = problem.make_context(p, with_q=True, num_registers=1)
    ctx2.set_material(p.c0, p.rho0, p.b_over_a, p.sigma, np.zeros(p.E))
    ctx2.step(scheme, dt, p.steps, 0.0)
    assert rel_l2(ctx2.get_state(0), r).max() > 1e3 * TOL_FP64


def test_fused_ldg_mixed_vs_fp64():
    print("fused_ldg_mixed", flush=True)
    p = heterogeneous(small_c2(n_e=4, order=4, steps=40))
    p.diffusion = np.full(p.E, 1.0e-3)
    dt = problem.cfl_dt_host(p)
    st = []
    for prec in (dg3d.FP64, dg3d.MIXED):
        ctx = problem.make_context(p, precision=prec, with_q=True, num_registers=1)
        ctx.step(dg3d.LSRK54, dt, p.steps, 0.0)
        st.append(ctx.get_state(0))
        ctx.close()
    err = rel_l2(st[1], st[0])
    assert 0.0 < err.max() < TOL_MIXED, err


def _smooth_random_state(p, seed, rho_scale=0.3, u_scale=0.3):
    """Random rho' (kg/m^3) and velocities (m/s) on every node: every flux term, including the
    nonlinear rho'^2 and rho' u parts, carries data."""
    rng = np.random.default_rng(seed)
    st = np.empty((4, p.dof))
    st[0] = rho_scale * (rng.random(p.dof) - 0.5)
    st[1:] = u_scale * (rng.random((3, p.dof)) - 0.5)
    return st.reshape(-1)


@pytest.mark.parametrize("precision", [dg3d.FP64, dg3d.MIXED])
@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
@pytest.mark.parametrize("order,n_e", [(7, 3), (1, 5)])
def test_fused_orders_1_and_7_vs_oracle(order, n_e, scheme, precision):
    """The fused step at 8 and 2 nodes per line, where the LSRK register (pair-major FP64,
    node-major FP32) and the group-blocked state have their largest and smallest groups of
    elements (8 and 512) and the stage its own launch
    bounds: heterogeneous c and rho0 with B/A = 5 everywhere, random state on every node, the tone
    burst at its peak (the clock starts 30 steps before the envelope's centre), 60 steps. Every
    variant of the stage kernel runs; 27 elements at N = 7 leave a partial last workgroup under
    two elements per workgroup, 125 elements at N = 1 under both 32 and 16. FP64 equals the
    oracle to 1e-12 and every FP64 variant gives the same bits; mixed is within 1e-5."""
    print(f"fused_orders_1_7 order={order} n_e={n_e} scheme={scheme} precision={precision}", flush=True)
    p = heterogeneous(small_c2(n_e=n_e, order=order, steps=60), seed=21)
    p.b_over_a = np.full(p.E, 5.0)
    st0 = _smooth_random_state(p, seed=22)
    dt = problem.cfl_dt_host(p)
    t0 = p.source.t0 - 30 * dt
    ref = None
    for v in range(8):
        ctx = problem.make_context(p, precision=precision)
        ctx.set_state(0, st0)
        ctx.set_variant(scheme, v, 1)
        t = ctx.step(scheme, dt, p.steps, t0)
        st = ctx.get_state(0)
        ctx.close()
        assert abs(t - (t0 + p.steps * dt)) < 1e-12 * t
        if ref is None:
            ref = st
        elif precision == dg3d.FP64:
            assert np.array_equal(st, ref), v
        else:
            assert rel_l2(st, ref).max() < 1e-6, v
    o = Oracle(p, literal=False, num_registers=4)
    load_problem_into(o, p, 0)
    o.set_state(0, st0)
    o.set_time(t0)
    for _ in range(p.steps):
        if scheme == dg3d.RK33:
            o.rk33(dt, 3, 0, 1, 2)
            o.copy(0, 3)
        else:
            o.lsrk54(dt, 0, 1, 2)
    r = o.get_state(0)
    err = rel_l2(ref, r)
    print(f"  rel L2 vs oracle {err}", flush=True)
    assert np.linalg.norm(r - st0) > 1e-3 * np.linalg.norm(st0)
    assert err.max() < (TOL_FP64 if precision == dg3d.FP64 else TOL_MIXED), err


# elements per workgroup of the stage kernel's two shapes, per order
_EPB = {1: (32, 16), 2: (8, 16), 3: (4, 8), 4: (2, 4), 5: (2, 1), 6: (2, 1), 7: (1, 2)}


@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7])
def test_interior_workgroups_vs_oracle(order, scheme):
    """Element columns away from every block face, whose face neighbours are all gathered from the
    state, beside columns on the periodic x faces, the exterior faces and the source face: a
    4 x 4 x 3*EPB block (EPB the larger of the two elements-per-workgroup shapes, so that both
    shapes have whole columns inside), heterogeneous material, random state, the burst near its
    peak, 4 steps. Every variant gives the same bits, and they equal the oracle to 1e-12."""
    print(f"interior_workgroups order={order} scheme={scheme}", flush=True)
    nE = (4, 4, 3 * max(_EPB[order]))
    p = heterogeneous(problem._burst_problem("interior", order, nE, 0.75e-3, 4), seed=31)
    p.periodic = (True, False, False)
    st0 = _smooth_random_state(p, seed=32)
    dt = problem.cfl_dt_host(p)
    t0 = p.source.t0 - 2 * dt
    ref = None
    for v in range(8):
        ctx = problem.make_context(p)
        ctx.set_state(0, st0)
        ctx.set_variant(scheme, v, 1)
        ctx.step(scheme, dt, p.steps, t0)
        st = ctx.get_state(0)
        ctx.close()
        if ref is None:
            ref = st
        assert np.array_equal(st, ref), v
    o = Oracle(p, literal=False, num_registers=4)
    load_problem_into(o, p, 0)
    o.set_state(0, st0)
    o.set_time(t0)
    for _ in range(p.steps):
        if scheme == dg3d.RK33:
            o.rk33(dt, 3, 0, 1, 2)
            o.copy(0, 3)
        else:
            o.lsrk54(dt, 0, 1, 2)
    err = rel_l2(ref, o.get_state(0))
    print(f"  rel L2 vs oracle {err}", flush=True)
    assert err.max() < TOL_FP64, err


# elements per workgroup of the LDG Q kernel, per order
_EPB_LDG = {1: 32, 2: 8, 3: 4, 4: 2, 5: 2, 6: 2, 7: 1}


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7])
def test_ldg_q_kernel_orders_vs_oracle(order):
    """LSRK54 with LDG diffusion on a 3 x 3 x 2*EPB block, EPB the Q kernel's elements per
    workgroup, so that a workgroup's column holds neighbours in z (the LDG stage reads their Q
    from LDS) beside gathered ones; periodic y, heterogeneous D and material, the burst near its
    peak, 5 steps, under each of the four workgroup orders: the same bits under every order, and
    equal to the oracle."""
    print(f"ldg_q_orders order={order}", flush=True)
    nE = (3, 3, 2 * _EPB_LDG[order])
    p = heterogeneous(problem._burst_problem("ldg", order, nE, 0.75e-3, 5), seed=41)
    p.periodic = (False, True, False)
    rng = np.random.default_rng(42)
    p.diffusion = 1.0e-3 * (1.0 + rng.random(p.E))
    st0 = _smooth_random_state(p, seed=43)
    dt = problem.cfl_dt_host(p)
    t0 = p.source.t0 - 2 * dt
    ref = None
    for x in range(4):
        ctx = problem.make_context(p, with_q=True, num_registers=1)
        ctx.set_state(0, st0)
        ctx.set_variant(dg3d.LSRK54, 0, x)
        ctx.step(dg3d.LSRK54, dt, p.steps, t0)
        st = ctx.get_state(0)
        ctx.close()
        if ref is None:
            ref = st
        assert np.array_equal(st, ref), x
    o = Oracle(p, literal=False, num_registers=4)
    load_problem_into(o, p, 0)
    o.set_state(0, st0)
    o.set_time(t0)
    for _ in range(p.steps):
        o.lsrk54(dt, 0, 1, 2, adv=1, diff=1, sponge=1)
    err = rel_l2(ref, o.get_state(0))
    print(f"  rel L2 vs oracle {err}", flush=True)
    assert err.max() < TOL_FP64, err


@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
@pytest.mark.parametrize("order", [1, 4])
def test_register_slots_past_the_elements_stay_zero(scheme, order):
    """The slots past a group's elements in each field run, and those of a last, partial group,
    are never written by a stage, yet the reductions cover them. A zero state with
    no source, autotuned (the tuner streams a pseudo-random state through the work registers)
    and stepped (LSRK54 hands the state over by pointer, so register 0 becomes a work register):
    the max |.| over every slot stays exactly zero."""
    print(f"slots_past_elements scheme={scheme} order={order}", flush=True)
    p = small_c2(n_e=3, order=order, steps=3)
    p.source = problem.Source()
    ctx = problem.make_context(p)
    ctx.set_state(0, np.zeros(4 * p.dof))
    ctx.autotune(scheme)
    ctx.step(scheme, problem.cfl_dt_host(p), p.steps, 0.0)
    m = ctx.reduce(dg3d.REDUCE_MAXABS, 0)
    ctx.close()
    assert np.all(np.asarray(m)[:4] == 0.0), m


@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
def test_autotune_mid_run_leaves_the_solution(scheme):
    """dg3d_autotune between steps times its candidates on work registers only (from a
    pseudo-random state, which it clears afterwards): 6 steps, a tuning, the same pair pinned
    again and 6 more steps give the bits of 12 steps in one go (heterogeneous material, B/A=5,
    the burst source on)."""
    print(f"autotune_mid_run scheme={scheme}", flush=True)
    p = heterogeneous(small_c2(n_e=4, order=3, steps=12), seed=51)
    dt = problem.cfl_dt_host(p)
    t0 = p.source.t0 - 6 * dt
    ctx = problem.make_context(p)
    ctx.set_variant(scheme, 2, 1)
    ctx.step(scheme, dt, 12, t0)
    ref = ctx.get_state(0)
    ctx.close()
    ctx = problem.make_context(p)
    ctx.set_variant(scheme, 2, 1)
    t = ctx.step(scheme, dt, 6, t0)
    ctx.autotune(scheme)
    ctx.set_variant(scheme, 2, 1)
    ctx.step(scheme, dt, 6, t)
    st = ctx.get_state(0)
    ctx.close()
    assert np.linalg.norm(ref) > 0
    assert np.array_equal(st, ref)
