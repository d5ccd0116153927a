"""Pins of the CPU oracle: the reference's own Matrix.h solve, closed forms, invariants and
brute force. These run on the CPU only."""
import numpy as np
import pytest

import oracle_bind as ob
from helpers import dg3d, problem, rel_l2, small_c1


# ---- per-line solve vs the reference's QR solve (Element.cpp:602, Matrix.h:289-445) ----

@pytest.mark.skipif(not ob.ref_available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7])
def test_line_solve_matches_reference_matrix_h(N):
    x, w, M, S, _ = dg3d.basis_generate(N)
    rng = np.random.default_rng(N)
    for trial in range(20):
        f = rng.standard_normal(N + 1)
        flux = np.zeros(N + 1)
        flux[0], flux[-1] = rng.standard_normal(2)
        q = float(np.float32(1000.0 if trial % 2 == 0 else 1.0 / 1000.0))
        width = float(np.float32(0.125 * (1 + rng.random())))
        xr, rank = ob.ref_line_solve(M, S, q, width, f, flux)
        if rank < N + 1:
            continue  # truncated solves are covered by the next test
        xo = ob.line_solve(M, S, q, width, f, flux)
        assert np.abs(xo - xr).max() <= 1e-12 * np.abs(xr).max()


@pytest.mark.skipif(not ob.ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_reference_rank_cutoff_truncates_small_elements():
    """Matrix.h:320 stops the QR when a row norm^2 of M q w falls below 1e-12: with q = 1/rho0
    (the mass equation) and 1 mm elements the reference's solve loses rank. The oracle and the
    HIP path solve exactly instead."""
    N = 4
    x, w, M, S, _ = dg3d.basis_generate(N)
    f = np.linspace(1, 2, N + 1)
    flux = np.zeros(N + 1)
    xr, rank = ob.ref_line_solve(M, S, 1.0 / 1000.0, 1e-3, f, flux)
    assert rank < N + 1
    xo = ob.line_solve(M, S, 1.0 / 1000.0, 1e-3, f, flux)
    exact = np.linalg.solve(M * (1e-3 * 1e-3), (S * 1e-3) @ f)
    assert np.abs(xo - exact).max() < 1e-9 * np.abs(exact).max()
    assert np.abs(xr - exact).max() > 1e-3 * np.abs(exact).max()


# ---- line operator: polynomial exactness through the 3-D element loops ----------------

def _poly_problem(N=3):
    nE = (3, 2, 2)
    p = problem._base("poly", N, nE, (0.1, -0.2, 0.3), [np.array([0.1, 0.12, 0.09]), np.array([0.1, 0.1]),
                                                         np.array([0.11, 0.1])], 1500.0, 1000.0, 0.0)
    return p


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_oracle_operator_exact_for_polynomials(axis):
    N = 3
    p = _poly_problem(N)
    o = ob.Oracle(p, num_registers=2)
    coords = [problem.axis_coords(p, a) for a in range(3)]
    n = p.n
    nE = p.num_elements
    poly = np.polynomial.Polynomial([0.3, -1.0, 2.0, 0.7])
    X = [None, None, None]
    X[0] = coords[0][:, None, None, :, None, None]
    X[1] = coords[1][None, :, None, None, :, None]
    X[2] = coords[2][None, None, :, None, None, :]
    shape = tuple(nE) + (n, n, n)
    xa = np.broadcast_to(X[axis], shape)
    st = np.zeros((4,) + shape)
    st[0] = 1e-4 * poly(xa)              # rho' varies along `axis`
    st[1 + axis] = 1e-5 * poly(xa)       # so does u_axis
    o.set_state(0, st.reshape(-1))
    # halo = the same polynomials at the halo face nodes (a continuous exterior state)
    ecs = problem.element_corners(p.corner, p.widths)
    for face in range(6):
        a = problem.FACE_AXIS[face]
        b_ax, c_ax = (1, 2) if a == 0 else ((0, 2) if a == 1 else (0, 1))
        tr = np.zeros((nE[b_ax], nE[c_ax], n, n, 3))
        if a == axis:
            high = face % 2
            xf = (ecs[a][-1] + p.widths[a][-1]) if high else ecs[a][0]
            tr[..., 0] = 1e-4 * poly(xf)
            tr[..., 1] = 1e-5 * poly(xf)
        else:
            cb = coords[axis]
            arr = cb[:, None, :, None] if axis == b_ax else cb[None, :, None, :]
            val = np.broadcast_to(arr, (nE[b_ax], nE[c_ax], n, n))
            tr[..., 0] = 1e-4 * poly(val)
            tr[..., 1] = 0.0
        o.set_halo(0, face, tr.reshape(-1))
    for mode in (0, 1):
        o.zero(1)
        o.operate(0, axis, mode, 0, 1)
        out = o.get_state(1).reshape((4,) + shape)
        dpoly = poly.deriv()
        if mode == 0:  # du_axis = -(1/rho0) dP/dx = -(c^2/rho0) drho'/dx
            expect = -(1500.0 ** 2 / 1000.0) * 1e-4 * dpoly(xa)
            got = out[1 + axis]
        else:          # drho' = -d(rho0 u_axis)/dx
            expect = -1000.0 * 1e-5 * dpoly(xa)
            got = out[0]
        assert np.abs(got - expect).max() < 1e-10 * np.abs(expect).max()


# ---- sponge sigma (Block.cpp:163-233) --------------------------------------------------

def _sigma_brute(nE, nB, bpos, sw, smax, ramp):
    """Integer-distance restatement for uniform widths: d = elements to the nearest sponge face."""
    out = []
    big = 10 ** 9
    for i in range(nE[0]):
        for j in range(nE[1]):
            for k in range(nE[2]):
                g = (bpos[0] * nE[0] + i, bpos[1] * nE[1] + j, bpos[2] * nE[2] + k)
                tot = (nE[0] * nB[0], nE[1] * nB[1], nE[2] * nB[2])
                d = [g[0] if sw[0] else big, tot[0] - g[0] - 1 if sw[1] else big,
                     g[1] if sw[2] else big, tot[1] - g[1] - 1 if sw[3] else big,
                     g[2] if sw[4] else big, tot[2] - g[2] - 1 if sw[5] else big]
                f = int(np.argmin(d))
                if d[f] >= big:
                    out.append(0.0)
                    continue
                r = max(0.0, 1.0 - d[f] / sw[f])
                out.append(smax * (r if ramp == 0 else r * r))
    return np.array(out)


@pytest.mark.parametrize("case", [
    ((8, 8, 8), (1, 1, 1), (0, 0, 0), (2, 2, 2, 2, 2, 2), 0.125, 0),
    ((8, 8, 8), (1, 1, 1), (0, 0, 0), (2, 2, 2, 2, 2, 2), 0.125, 1),
    ((8, 8, 16), (2, 2, 1), (1, 0, 0), (1, 1, 1, 1, 1, 1), 0.125, 0),
    ((10, 6, 12), (1, 1, 1), (0, 0, 0), (8, 8, 8, 8, 0, 8), 0.25, 1),
    ((4, 4, 4), (1, 1, 1), (0, 0, 0), (0, 0, 0, 0, 0, 0), 0.5, 0),
])
def test_sponge_sigma_product_oracle_bruteforce(case):
    nE, nB, bpos, sw, h, ramp = case
    smax = 15000.0
    prod = dg3d.sponge_sigma(nE, nB, bpos, (h, h, h), sw, smax, ramp)
    orc = ob.sponge_sigma(nE, nB, bpos, (h, h, h), sw, smax, ramp)
    assert np.array_equal(prod, orc)
    brute = _sigma_brute(nE, nB, bpos, sw, smax, ramp)
    assert np.abs(prod - brute).max() <= 1e-6 * smax


def test_sponge_two_element_table():
    """W=2, linear: the element-constant sigma takes the values {sigma_max, sigma_max/2, 0}."""
    s = dg3d.sponge_sigma((8, 8, 8), (1, 1, 1), (0, 0, 0), (0.125,) * 3, (2,) * 6, 1000.0, 0).reshape(8, 8, 8)
    assert s[0, 4, 4] == 1000.0 and s[1, 4, 4] == 500.0 and s[2, 4, 4] == 0.0 and s[7, 4, 4] == 1000.0
    assert set(np.unique(s)) == {0.0, 500.0, 1000.0}


# ---- time integration: orders on y' = -sigma y ---------------------------------------

@pytest.mark.parametrize("scheme,order", [("rk33", 3), ("lsrk54", 4)])
def test_rk_order_on_sponge_decay(scheme, order):
    p = _poly_problem(2)
    sig = 1000.0
    p.sigma = np.full(p.E, sig)
    o = ob.Oracle(p, num_registers=4)
    T = 1e-3
    errs = []
    for steps in (8, 16, 32):
        st = np.zeros(4 * p.dof)
        st[: p.dof] = 1.0
        o.set_state(0, st)
        for f in range(6):
            o.set_halo(0, f, None)
        o.zero(1)
        dt = T / steps
        for _ in range(steps):
            if scheme == "rk33":
                o.rk33(dt, 3, 0, 1, 2, adv=0, diff=0, sponge=1)
                o.copy(0, 3)
            else:
                o.lsrk54(dt, 0, 1, 2, adv=0, diff=0, sponge=1)
        errs.append(abs(o.get_state(0)[0] - np.exp(-sig * T)))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(np.abs(rates - order) < 0.25), rates


# ---- energy and literal/FP64 agreement -----------------------------------------------

def mass_norm_energy(p, st):
    """0.5 * (c^2/rho0 rho'^T M3 rho' + rho0 u^T M3 u) with M3 = M x M x M scaled by the element
    volume: the norm in which the upwind DG semi-discretisation is dissipative."""
    n = p.n
    shape = (4,) + tuple(p.num_elements) + (n, n, n)
    s = st.reshape(shape)
    vol = (p.widths[0][:, None, None] * p.widths[1][None, :, None] * p.widths[2][None, None, :])
    M = p.mass
    def quad(f):
        g = np.einsum("ia,...abc->...ibc", M, f)
        g = np.einsum("jb,...ibc->...ijc", M, g)
        g = np.einsum("kc,...ijc->...ijk", M, g)
        return np.einsum("...ijk,...ijk->...", f, g) * vol
    c2r = (p.c0 ** 2 / p.rho0).reshape(p.num_elements)
    r0 = p.rho0.reshape(p.num_elements)
    return 0.5 * float(np.sum(c2r * quad(s[0]) + r0 * (quad(s[1]) + quad(s[2]) + quad(s[3]))))


def test_energy_non_increasing_upwind():
    p = small_c1()
    p.sigma[:] = 0.0
    o = ob.Oracle(p)
    ob.load_problem_into(o, p)
    for f in range(6):
        o.set_halo(0, f, None)  # no inflow: a closed absorbing box
    dt = problem.cfl_dt_host(p)
    e = [mass_norm_energy(p, o.get_state(0))]
    for _ in range(40):
        o.lsrk54(dt, 0, 1, 2)
        e.append(mass_norm_energy(p, o.get_state(0)))
    e = np.array(e)
    assert np.all(np.diff(e) <= 1e-12 * e[0])
    assert e[-1] < e[0]


def test_literal_fp32_close_to_fp64():
    p = small_c1()
    dt = problem.cfl_dt_host(p)
    ol = ob.Oracle(p, literal=True)
    on = ob.Oracle(p, literal=False)
    for o in (ol, on):
        ob.load_problem_into(o, p)
        for _ in range(10):
            o.rk33(dt, 3, 0, 1, 2)
            o.copy(0, 3)
    err = rel_l2(ol.get_state(0), on.get_state(0))
    assert err.max() < 2e-5, err
    assert err[0] > 0


def test_stability_cfl_definition_long_run():
    """With dt = CFL / max(c sum 2/(w dxi)) and CFL 0.3, LSRK54 stays bounded on noise."""
    p = small_c1()
    p.sigma[:] = 0.0
    o = ob.Oracle(p)
    rng = np.random.default_rng(0)
    o.set_state(0, 1e-6 * rng.standard_normal(4 * p.dof))
    dt = problem.cfl_dt_host(p)
    e0 = o.energy(0)[0]
    for _ in range(150):
        o.lsrk54(dt, 0, 1, 2)
    assert o.energy(0)[0] <= e0


# ---- LDG diffusion (Element.cpp:623-758): exactness and the heat kernel ------------------

@pytest.mark.parametrize("axis", [0, 1, 2])
def test_oracle_ldg_exact_for_polynomials(axis):
    """Q_a = d rho'/dx_a and the DRhoFromQ increment D d^2 rho'/dx_a^2 (so that rho' follows the
    heat equation), exactly for a cubic at N=3 through the element loops, with the exterior
    continuing the polynomial."""
    N = 3
    p = _poly_problem(N)
    D = 3.0
    p.diffusion = np.full(p.E, D)
    o = ob.Oracle(p, num_registers=2)
    coords = [problem.axis_coords(p, a) for a in range(3)]
    n = p.n
    nE = p.num_elements
    poly = np.polynomial.Polynomial([0.3, -1.0, 2.0, 0.7])
    d1, d2 = poly.deriv(), poly.deriv(2)
    X = [coords[0][:, None, None, :, None, None], coords[1][None, :, None, None, :, None],
         coords[2][None, None, :, None, None, :]]
    shape = tuple(nE) + (n, n, n)
    xa = np.broadcast_to(X[axis], shape)
    st = np.zeros((7,) + shape)
    st[0] = 1e-4 * poly(xa)
    o.set_state(0, st.reshape(-1))
    ecs = problem.element_corners(p.corner, p.widths)
    for face in range(6):
        a = problem.FACE_AXIS[face]
        b_ax, c_ax = (1, 2) if a == 0 else ((0, 2) if a == 1 else (0, 1))
        tr = np.zeros((nE[b_ax], nE[c_ax], n, n, 3))
        if a == axis:
            xf = (ecs[a][-1] + p.widths[a][-1]) if face % 2 else ecs[a][0]
            tr[..., 0] = 1e-4 * poly(xf)
            tr[..., 2] = 1e-4 * d1(xf)    # the exterior Q_axis
        else:
            cb = coords[axis]
            arr = cb[:, None, :, None] if axis == b_ax else cb[None, :, None, :]
            tr[..., 0] = 1e-4 * poly(np.broadcast_to(arr, (nE[b_ax], nE[c_ax], n, n)))
        o.set_halo(0, face, tr.reshape(-1))
    o.operate(2, axis, 0, 0, 0)                      # QFromDRho, in place
    q = o.get_state(0, nfields=7).reshape((7,) + shape)
    expect_q = 1e-4 * d1(xa)
    assert np.abs(q[4 + axis] - expect_q).max() < 1e-10 * np.abs(expect_q).max()
    o.zero(1)
    o.operate(3, axis, 0, 0, 1)                      # DRhoFromQ into register 1
    out = o.get_state(1, nfields=7).reshape((7,) + shape)
    expect = D * 1e-4 * d2(xa)
    assert np.abs(out[0] - expect).max() < 1e-9 * np.abs(expect).max()


def _heat_decay_error(N, ne=4, D=2.0e-3, T=5.0, steps=200):
    nE = (ne, ne, ne)
    L = 1.0
    p = problem._base("heat", N, nE, (0, 0, 0), problem._uniform_widths(nE, (L / ne,) * 3), 1500.0, 1000.0, 0.0)
    p.periodic = (1, 1, 1)
    p.diffusion = np.full(p.E, D)
    o = ob.Oracle(p, num_registers=3)
    coords = problem.axis_coords(p, 1)
    n = p.n
    k = 2 * np.pi / L
    shape = nE + (n, n, n)
    y = np.broadcast_to(coords[None, :, None, None, :, None], shape)
    st = np.zeros((7,) + shape)
    st[0] = 1e-4 * np.sin(k * y)
    o.set_state(0, st.reshape(-1))
    dt = T / steps
    o.set_time(0.0)
    for _ in range(steps):
        o.lsrk54(dt, 0, 1, 2, adv=0, diff=1, sponge=0)
    got = o.get_state(0).reshape((4,) + shape)
    decay = np.exp(-D * k * k * T)
    expect = 1e-4 * np.sin(k * y) * decay
    assert np.abs(got[1:]).max() == 0.0   # diffusion acts on rho' only
    return np.abs(got[0] - expect).max() / 1e-4, decay


def test_oracle_ldg_heat_kernel_decay():
    """A Fourier mode of rho' in a periodic box under diffusion alone decays as
    exp(-D k^2 t) (the Gaussian heat kernel's mode-by-mode form), through LSRK54; the error
    falls spectrally with the order (about 10x per order here)."""
    e4, decay = _heat_decay_error(4)
    e6, _ = _heat_decay_error(6)
    assert decay < 0.7          # a decay that a sign or scale error of D would miss
    assert e6 < 5e-6, e6
    assert e4 / e6 > 50, (e4, e6)


# ---- Grid::forwardEuler (Grid.cpp:567-640) --------------------------------------------

def _periodic_box(N=3, ne=4, D=0.0):
    nE = (ne, ne, ne)
    p = problem._base("box", N, nE, (0, 0, 0), problem._uniform_widths(nE, (1.0 / ne,) * 3), 1500.0, 1000.0, 0.0)
    p.periodic = (1, 1, 1)
    p.diffusion = np.full(p.E, D)
    return p


def test_oracle_forward_euler_order_one_on_sponge_decay():
    """y' = -sigma y for a uniform rho' in a periodic box (advection and LDG vanish): one step
    gives (1 - sigma dt) y exactly, and the error at a fixed time halves with dt (order 1)."""
    p = _periodic_box(N=2, ne=2)
    sig = 1000.0
    p.sigma = np.full(p.E, sig)
    T = 1e-3
    errs = []
    for steps in (1, 16, 32, 64):
        o = ob.Oracle(p, num_registers=2)
        st = np.zeros(7 * p.dof)
        st[: p.dof] = 1.0
        o.set_state(0, st)
        dt = T / steps
        cur, nxt = 0, 1
        for _ in range(steps):
            o.forward_euler(dt, nxt, cur)
            cur, nxt = nxt, cur
        y = o.get_state(cur)[: p.dof]
        if steps == 1:
            assert np.abs(y - (1.0 - sig * dt)).max() < 1e-14
        else:
            errs.append(abs(y.mean() - np.exp(-sig * T)))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(np.abs(rates - 1.0) < 0.1), rates


def test_oracle_forward_euler_is_rk33_first_stage():
    """With D = 0, forwardEuler's step is Grid::RK33's first stage k1 = L(u0) dt + u0
    (Grid.cpp:705): a Gaussian pulse with the initial-condition halos, sponge, B/A = 5."""
    p = small_c1()
    p.b_over_a = np.full(p.E, 5.0)
    dt = problem.cfl_dt_host(p)
    a = ob.Oracle(p, num_registers=2)
    ob.load_problem_into(a, p, 0)
    a.set_time(0.0)
    a.forward_euler(dt, 1, 0)
    b = ob.Oracle(p, num_registers=4)
    ob.load_problem_into(b, p, 0)
    b.set_time(0.0)
    b.rk33(dt, 3, 0, 1, 2, adv=1, diff=0, sponge=1)
    got, k1 = a.get_state(1), b.get_state(1)
    assert np.linalg.norm(k1 - b.get_state(0)) > 0
    assert np.abs(got - k1).max() <= 1e-15 * np.abs(k1).max()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_oracle_forward_euler_ldg_on_x_only(axis):
    """forwardEuler runs the LDG pair for the x axis only (Grid.cpp:581,597). For a mode along x
    its step equals RK33's first stage with the full three-axis LDG; for a mode along y or z the
    diffusion leaves it unchanged, while the three-axis stage changes it visibly."""
    p = _periodic_box(N=4, ne=4, D=1.0)
    n = p.n
    coords = problem.axis_coords(p, axis)
    shape = tuple(p.num_elements) + (n, n, n)
    idx = [None] * 6
    idx[axis] = slice(None)
    idx[3 + axis] = slice(None)
    xa = np.broadcast_to(coords[tuple(idx[a] if a in (axis, 3 + axis) else None for a in range(6))], shape)
    st = np.zeros((7,) + shape)
    st[0] = 1e-4 * np.sin(2 * np.pi * xa)
    dt = problem.cfl_dt_host(p)
    a = ob.Oracle(p, num_registers=2)
    a.set_state(0, st.reshape(-1))
    a.forward_euler(dt, 1, 0)
    b = ob.Oracle(p, num_registers=4)
    b.set_state(0, st.reshape(-1))
    b.rk33(dt, 3, 0, 1, 2, adv=1, diff=1, sponge=1)
    e_rho = a.get_state(1)[: p.dof]
    k1_rho = b.get_state(1)[: p.dof]
    rho0 = st[0].reshape(-1)
    change = np.abs(k1_rho - rho0).max()
    assert change > 1e-4 * 1e-4          # the diffusion step is far above rounding
    if axis == 0:
        assert np.abs(e_rho - k1_rho).max() < 1e-9 * change
    else:
        assert np.abs(e_rho - rho0).max() < 1e-9 * change
