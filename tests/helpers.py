"""Shared helpers for the parity tests: problem variants and state comparison."""
import importlib

import numpy as np

pk = importlib.import_module("3ddg_amd")
problem = pk.problem
dg3d = pk.dg3d


def rel_l2(a, b, nfields=4):
    """Per-field ||a-b|| / ||b|| (absolute when ||b|| == 0)."""
    a = np.asarray(a).reshape(nfields, -1)
    b = np.asarray(b).reshape(nfields, -1)
    out = []
    for f in range(nfields):
        nb = np.linalg.norm(b[f])
        d = np.linalg.norm(a[f] - b[f])
        out.append(d / nb if nb > 0 else d)
    return np.array(out)


def heterogeneous(p, seed=1):
    """Per-element random material (c, rho0, B/A) around water: exercises the heterogeneous
    flux coefficients of Element.cpp:540-600."""
    rng = np.random.default_rng(seed)
    p.c0 = 1500.0 * (1.0 + 0.1 * rng.random(p.E))
    p.rho0 = 1000.0 * (1.0 + 0.2 * rng.random(p.E))
    p.b_over_a = np.where(rng.random(p.E) < 0.3, 0.0, 10.0 * rng.random(p.E))
    return p


def random_state(p, seed=2, scale=None):
    """Smooth IC plus random noise on all four fields (so every term is non-zero)."""
    rng = np.random.default_rng(seed)
    st = problem.initial_state(p).reshape(4, -1)
    amp = scale if scale is not None else max(float(np.abs(st).max()), 1e-3)
    st = st + amp * 0.1 * rng.standard_normal(st.shape)
    st[1:] += amp * 1e-3 * rng.standard_normal(st[1:].shape)
    return st.reshape(-1)


def random_halos(p, seed=3, scale=1e-3):
    rng = np.random.default_rng(seed)
    out = []
    for face in range(6):
        a = problem.FACE_AXIS[face]
        b, c = (1, 2) if a == 0 else ((0, 2) if a == 1 else (0, 1))
        size = p.num_elements[b] * p.num_elements[c] * p.n ** 2 * 3
        out.append(scale * rng.standard_normal(size))
    return out


def small_c1(steps=100):
    return problem.config_c1(steps=steps)


def small_c2(n_e=4, order=4, steps=60):
    p = problem.config_c2(num_elements=n_e, steps=steps, order=order)
    return p


def small_c5(n_e=4, steps=60):
    return problem.config_c5(num_elements=n_e, steps=steps)


def discrete_mass(p, state):
    """The discrete mass of rho': the sum over elements of the element volume times
    sum_abc w_a w_b w_c rho'_abc. With the exact mass matrix, 1^T M is the vector of LGL weights,
    so this is the integral of rho' that the weak form (S = D^T W, single-valued face fluxes)
    conserves. Returns it and the same sum of |rho'| (its scale)."""
    n = p.n
    w = np.asarray(p.weights)
    w3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1)
    rho = np.asarray(state).reshape(-1)[: p.E * n ** 3].reshape(p.E, n ** 3)
    wx, wy, wz = (np.asarray(p.widths[a], float) for a in range(3))
    vol = (wx[:, None, None] * wy[None, :, None] * wz[None, None, :]).reshape(-1)  # k fastest
    return float(np.sum(vol * (rho @ w3))), float(np.sum(vol * (np.abs(rho) @ w3)))


def periodic_mass_box(steps, uniform_rho0=True):
    """8^3 periodic box without sponge around the c1 Gaussian pulse, with per-element c and B/A
    (B/A = 0 on some elements); rho0 uniform unless asked otherwise."""
    p = heterogeneous(small_c1(steps=steps))
    if uniform_rho0:
        p.rho0 = np.full(p.E, 1000.0)
    p.periodic = (1, 1, 1)
    p.sigma = np.zeros(p.E)
    return p


def plane_burst_column():
    """The 64^3 N=4 workload's 1 MHz, 1e6 Pa tone burst (B/A = 5) in a column of 2 x 2 x 64
    elements of 0.75 mm, periodic across (x, y) so that the wave is exactly plane, with the
    8-element sponge at the far (+z) end only."""
    h = 0.75e-3
    p = problem._burst_problem("plane_column", 4, (2, 2, 64), h, 1)
    p.periodic = (1, 1, 0)
    problem.set_sponge(p, (0, 0, 0, 0, 0, 8), 3 * 1500.0 / (8 * h), ramp=1)
    return p
