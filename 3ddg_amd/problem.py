"""Problem definitions for the block solver: mesh, material, sponge, initial state, halo traces
and boundary source, built on the host and handed to the C ABI.

The reference defines a run by parameters1.txt (driver.cpp:129-196) plus three data files it does
not ship. This module builds the same quantities:
  * node coordinates  = width * (xi/2 + 1/2) + element corner          (Element.cpp:334-336)
  * Gaussian IC        = A * exp(-IC_SIGMA * r^2) on every node, halos too (Node.cpp:76-101)
  * sponge sigma       = dg3d_sponge_sigma (Block.cpp:163-233)
and the workloads used for measurement, whose extensions (LSRK54, CFL dt, quadratic ramp,
boundary tone burst, stretched z, two-layer medium) the reference does not have:
  c1  8^3 unit cube, N=3, beta=0, Gaussian pulse, 2-element quadratic sponge, LSRK54 CFL 0.3, 100 steps
  c2  64^3, N=4, water beta=3.5, z=0 tone burst 1 MHz 1e6 Pa, 8-element sponge, LSRK54, 500 steps
  c3  c2 at N=7
  c5  48^3, N=5, geometric z (ratio 1.02), two layers, focused 1 MHz 5e6 Pa burst, 2000 steps
"""
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import dg3d

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

# reference direction numbering (Element.h:144): 0 -z, 1 +z, 2 -y, 3 +y, 4 -x, 5 +x
FACE_AXIS = [2, 2, 1, 1, 0, 0]


class MT19937_64:
    """64-bit Mersenne Twister (Matsumoto & Nishimura 2004), the engine std::mt19937_64 names."""

    def __init__(self, seed=5489):
        self.mt = [0] * 312
        self.mt[0] = seed & 0xFFFFFFFFFFFFFFFF
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & 0xFFFFFFFFFFFFFFFF
        self.idx = 312

    def _twist(self):
        mt = self.mt
        for i in range(312):
            x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self.idx = 0

    def __call__(self):
        if self.idx >= 312:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & 0xFFFFFFFFFFFFFFFF

    def uniform01(self):
        return self() / 2.0 ** 64


@dataclass
class Source:
    face: int = -1
    amplitude: float = 0.0
    frequency: float = 0.0
    phase: float = 0.0
    t0: float = 0.0
    tau: float = 1.0
    lead: np.ndarray = None  # [faceElem][n*n] seconds, or None


@dataclass
class Problem:
    name: str
    order: int
    num_elements: tuple
    corner: np.ndarray
    widths: list                      # 3 arrays (per axis)
    c0: np.ndarray                    # per element (reference element order)
    rho0: np.ndarray
    b_over_a: np.ndarray
    sigma: np.ndarray
    diffusion: np.ndarray
    nodes: np.ndarray                 # LGL nodes (n)
    weights: np.ndarray
    mass: np.ndarray
    stiff: np.ndarray
    periodic: tuple = (0, 0, 0)
    ic: dict = None                   # Gaussian IC parameters or None
    source: Source = field(default_factory=Source)
    scheme: int = dg3d.LSRK54
    cfl: float = 0.3
    dt: float = None                  # fixed dt (reference DT) or None for CFL
    steps: int = 100

    @property
    def n(self):
        return self.order + 1

    @property
    def E(self):
        return int(np.prod(self.num_elements))

    @property
    def dof(self):
        return self.E * self.n ** 3


def load_basis(order, data_dir=DATA_DIR):
    """LGL nodes/weights and M, S for `order` from the reference-format files in data_dir."""
    coeff = os.path.join(data_dir, "hardcoded_coeff.txt")
    mats = os.path.join(data_dir, "matrices.txt")
    if os.path.exists(coeff) and os.path.exists(mats):
        return dg3d.basis_read_files(coeff, mats, order)
    x, w, M, S, _ = dg3d.basis_generate(order)
    return x, w, M, S


def element_corners(corner, widths):
    """Lower corner of every element along each axis: STARTING_CORNER + cumulative widths."""
    return [corner[a] + np.concatenate([[0.0], np.cumsum(widths[a])[:-1]]) for a in range(3)]


def axis_coords(p, a):
    """Node coordinates along axis a: [element index][node index]."""
    ec = element_corners(p.corner, p.widths)[a]
    return p.widths[a][:, None] * (p.nodes[None, :] / 2.0 + 0.5) + ec[:, None]


def gaussian(p, xs, ys, zs):
    ic = p.ic
    d2 = 0.0
    for use, coord, cen in zip(ic["gauss"], (xs, ys, zs), ic["center"]):
        if use:
            d2 = d2 + (coord - cen) ** 2
    return np.exp(-d2 * ic["sigma"])


def initial_state(p):
    """Field-major state (rho', u, v, w), element (i,j,k) then node (a,b,c), k/c fastest."""
    n = p.n
    nE = p.num_elements
    st = np.zeros((4,) + tuple(nE) + (n, n, n))
    if p.ic is None or not any(p.ic["gauss"]):
        return st.reshape(-1)
    X = axis_coords(p, 0)
    Y = axis_coords(p, 1)
    Z = axis_coords(p, 2)
    xs = X[:, None, None, :, None, None]
    ys = Y[None, :, None, None, :, None]
    zs = Z[None, None, :, None, None, :]
    g = gaussian(p, xs, ys, zs)
    st[0] = p.ic["amp_rho"] * g
    for a in range(3):
        st[1 + a] = p.ic["amp_vel"][a] * g
    return st.reshape(-1)


def halo_traces(p):
    """Halo face traces of the initial grid: the Gaussian IC at the node layer of each halo
    element that touches the block, (rho', u_axis, Q_axis=0) per face node (dg3d_set_halo)."""
    out = []
    n = p.n
    nE = p.num_elements
    if p.ic is None or not any(p.ic["gauss"]):
        return [None] * 6
    coords = [axis_coords(p, a) for a in range(3)]
    ecs = element_corners(p.corner, p.widths)
    for face in range(6):
        a = FACE_AXIS[face]
        high = face % 2 == 1
        b_ax = 1 if a == 0 else 0
        c_ax = 1 if a == 2 else 2
        w = p.widths[a][-1] if high else p.widths[a][0]
        hc = (ecs[a][-1] + p.widths[a][-1]) if high else (p.corner[a] - p.widths[a][0])
        layer = 0 if high else n - 1
        xa = w * (p.nodes[layer] / 2.0 + 0.5) + hc
        cb = coords[b_ax]  # [nb][n]
        cc = coords[c_ax]
        full = [None, None, None]
        full[a] = np.full((nE[b_ax], nE[c_ax], n, n), xa)
        full[b_ax] = np.broadcast_to(cb[:, None, :, None], (nE[b_ax], nE[c_ax], n, n))
        full[c_ax] = np.broadcast_to(cc[None, :, None, :], (nE[b_ax], nE[c_ax], n, n))
        g = gaussian(p, full[0], full[1], full[2])
        tr = np.zeros((nE[b_ax], nE[c_ax], n, n, 3))
        tr[..., 0] = p.ic["amp_rho"] * g
        tr[..., 1] = p.ic["amp_vel"][a] * g
        out.append(tr.reshape(-1))
    return out


def focus_lead(p, focal_depth, center_xy):
    """Per face node time lead (s) of a burst focused at depth F on the z=0 face:
    (sqrt(F^2 + r^2) - F) / c, so outer nodes fire first."""
    n = p.n
    X = axis_coords(p, 0)
    Y = axis_coords(p, 1)
    nx, ny = p.num_elements[0], p.num_elements[1]
    xs = np.broadcast_to(X[:, None, :, None], (nx, ny, n, n))
    ys = np.broadcast_to(Y[None, :, None, :], (nx, ny, n, n))
    r2 = (xs - center_xy[0]) ** 2 + (ys - center_xy[1]) ** 2
    c = p.c0.reshape(p.num_elements)[:, :, 0]
    return ((np.sqrt(focal_depth ** 2 + r2) - focal_depth) / c[:, :, None, None]).reshape(-1)


def seeded_phase(seed=0):
    return 2.0 * math.pi * MT19937_64(seed).uniform01()


def _uniform_widths(nE, h):
    return [np.full(nE[a], float(h[a])) for a in range(3)]


def _base(name, order, nE, corner, widths, c0, rho0, ba):
    x, w, M, S = load_basis(order)
    E = int(np.prod(nE))
    return Problem(name=name, order=order, num_elements=tuple(nE), corner=np.asarray(corner, float),
                   widths=widths, c0=np.full(E, float(c0)), rho0=np.full(E, float(rho0)),
                   b_over_a=np.full(E, float(ba)), sigma=np.zeros(E), diffusion=np.zeros(E),
                   nodes=x, weights=w, mass=M, stiff=S)


def set_sponge(p, sponge_width, sigma_max, ramp):
    ew = [float(p.widths[a][0]) for a in range(3)]
    p.sigma = dg3d.sponge_sigma(p.num_elements, (1, 1, 1), (0, 0, 0), ew, sponge_width, sigma_max, ramp)


def config_c1(steps=100, scheme=dg3d.LSRK54):
    """8^3 unit cube, N=3, c=1500, rho0=1000, beta=0; Gaussian pressure pulse 1e3 Pa,
    sigma 0.08 at the centre; 2-element quadratic sponge with sigma_max = 3c/(2h)."""
    nE = (8, 8, 8)
    h = 1.0 / 8
    p = _base("c1", 3, nE, (0, 0, 0), _uniform_widths(nE, (h, h, h)), 1500.0, 1000.0, 0.0)
    p.ic = dict(gauss=(1, 1, 1), amp_rho=1e3 / 1500.0 ** 2, amp_vel=(0.0, 0.0, 0.0),
                sigma=1.0 / (2 * 0.08 ** 2), center=(0.5, 0.5, 0.5))
    set_sponge(p, (2, 2, 2, 2, 2, 2), 3 * 1500.0 / (2 * h), ramp=1)
    p.scheme, p.cfl, p.steps = scheme, 0.3, steps
    return p


def _burst_problem(name, order, nE, h, steps):
    p = _base(name, order, nE, (0, 0, 0), _uniform_widths(nE, (h, h, h)), 1500.0, 1000.0, 5.0)
    # 8-element sponge on every face but the source face (-z)
    set_sponge(p, (8, 8, 8, 8, 0, 8), 3 * 1500.0 / (8 * h), ramp=1)
    f = 1.0e6
    tau = 1.0 / f
    p.source = Source(face=0, amplitude=1.0e6, frequency=f, phase=seeded_phase(0), t0=3 * tau, tau=tau)
    p.scheme, p.cfl, p.steps = dg3d.LSRK54, 0.3, steps
    return p


def config_c2(num_elements=64, steps=500, order=4):
    """One GPU's block: 64^3 elements of 0.75 mm, N=4, water (beta=3.5 -> B/A=5); a 1 MHz,
    1e6 Pa Gaussian-modulated tone burst (envelope width 1/f, centre 3/f) enters through the
    z=0 face (phase from mt19937_64(0)); 8-element quadratic sponge elsewhere; LSRK54 at CFL 0.3."""
    nE = (num_elements,) * 3
    return _burst_problem("c2" if order == 4 else f"c2_N{order}", order, nE, 0.75e-3, steps)


def config_c3(num_elements=64, steps=500):
    """c2 at N=7 (512 nodes per element)."""
    p = config_c2(num_elements, steps, order=7)
    p.name = "c3"
    return p


def config_c5(num_elements=48, steps=2000):
    """48^3, N=5; x, y uniform, z geometric with ratio 1.02 over the same length; two layers
    split at the element face nearest z = L/2 ((1500, 1000, beta 3.5) below, (1540, 1050,
    beta 4.5) above); 1 MHz, 5e6 Pa burst on z=0 focused at depth L/2."""
    n_e = num_elements
    L = n_e * 0.75e-3
    nE = (n_e, n_e, n_e)
    r = 1.02
    h0 = L * (r - 1) / (r ** n_e - 1)
    wz = h0 * r ** np.arange(n_e)
    widths = [np.full(n_e, L / n_e), np.full(n_e, L / n_e), wz]
    p = _base("c5", 5, nE, (0, 0, 0), widths, 1500.0, 1000.0, 5.0)
    zf = np.concatenate([[0.0], np.cumsum(wz)])
    k_if = int(np.argmin(np.abs(zf - L / 2)))  # interface at face k_if
    kidx = np.broadcast_to(np.arange(n_e)[None, None, :], nE).reshape(-1)
    upper = kidx >= k_if
    p.c0 = np.where(upper, 1540.0, 1500.0)
    p.rho0 = np.where(upper, 1050.0, 1000.0)
    p.b_over_a = np.where(upper, 2 * (4.5 - 1), 2 * (3.5 - 1))
    # sponge in integer element units (element-constant, as the reference)
    p.sigma = dg3d.sponge_sigma(nE, (1, 1, 1), (0, 0, 0), (1.0, 1.0, 1.0), (8, 8, 8, 8, 0, 8),
                                3 * 1500.0 / (8 * L / n_e), 1)
    f = 1.0e6
    tau = 1.0 / f
    p.source = Source(face=0, amplitude=5.0e6, frequency=f, phase=seeded_phase(0), t0=3 * tau, tau=tau)
    p.source.lead = focus_lead(p, L / 2, (L / 2, L / 2))
    p.scheme, p.cfl, p.steps = dg3d.LSRK54, 0.3, steps
    return p


def make_context(p, precision=dg3d.FP64, contraction=dg3d.CONTRACT_VALU, num_registers=1, with_q=False,
                 xcd_remap=2, halos=True):
    """A device context loaded with the problem: geometry, material, IC (register 0), IC halo
    traces on register 0, and the source."""
    ctx = dg3d.Context(p.order, p.num_elements, p.mass, p.stiff, p.nodes, p.weights, periodic=p.periodic,
                       precision=precision, contraction=contraction, num_registers=num_registers,
                       with_q=with_q, xcd_remap=xcd_remap)
    ctx.set_geometry(p.corner, *p.widths)
    ctx.set_material(p.c0, p.rho0, p.b_over_a, p.sigma, p.diffusion)
    ctx.set_state(0, initial_state(p))
    if halos:
        for face, tr in enumerate(halo_traces(p)):
            if tr is not None and not p.periodic[FACE_AXIS[face]]:
                ctx.set_halo(0, face, tr)
    s = p.source
    if s.face >= 0:
        ctx.set_source(s.face, s.amplitude, s.frequency, s.phase, s.t0, s.tau, s.lead)
    return ctx


def cfl_dt_host(p):
    """CFL step on the host, as dg3d_cfl_dt: cfl / max_e(c_e * sum_a 2 / (w_a dxi_min))."""
    dxi = float(np.min(np.diff(p.nodes)))
    inv = (1.0 / p.widths[0][:, None, None] + 1.0 / p.widths[1][None, :, None]
           + 1.0 / p.widths[2][None, None, :]).reshape(-1)
    return p.cfl * float(np.min(0.5 * dxi / (inv * p.c0)))


def config_c4(num_solves=256, steps=100, seed=0):
    """Batch of independent unit-cube solves. Per solve, from mt19937_64(seed) in this order:
    elements per side uniform in 4..16, order N uniform in 2..6, pulse amplitude uniform in
    [1e4, 1e6] Pa, beta uniform in [0, 5] with B/A = max(0, 2 (beta - 1)). Each: water (c=1500,
    rho0=1000), Gaussian pressure pulse of width 0.08 at the centre, 2-element quadratic sponge
    (sigma_max = 3c/(2h)), LSRK54 at CFL 0.3."""
    rng = MT19937_64(seed)
    out = []
    for s in range(num_solves):
        ne = 4 + rng() % 13
        order = 2 + rng() % 5
        amp = 1.0e4 + (1.0e6 - 1.0e4) * rng.uniform01()
        beta = 5.0 * rng.uniform01()
        ba = max(0.0, 2.0 * (beta - 1.0))
        nE = (ne, ne, ne)
        h = 1.0 / ne
        p = _base(f"c4_{s}", order, nE, (0, 0, 0), _uniform_widths(nE, (h, h, h)), 1500.0, 1000.0, ba)
        p.ic = dict(gauss=(1, 1, 1), amp_rho=amp / 1500.0 ** 2, amp_vel=(0.0, 0.0, 0.0),
                    sigma=1.0 / (2 * 0.08 ** 2), center=(0.5, 0.5, 0.5))
        set_sponge(p, (2, 2, 2, 2, 2, 2), 3 * 1500.0 / (2 * h), ramp=1)
        p.scheme, p.cfl, p.steps = dg3d.LSRK54, 0.3, steps
        out.append(p)
    return out


def batch_solve(p):
    """The dg3d_batch_create description of one problem: the arrays make_context would upload,
    the initial halo traces, and the CFL time step."""
    halos = halo_traces(p)
    return dict(order=p.order, num_elements=p.num_elements, periodic=p.periodic, mass=p.mass, stiff=p.stiff,
                widths=p.widths, c0=p.c0, rho0=p.rho0, b_over_a=p.b_over_a, sigma=p.sigma,
                state=initial_state(p),
                halo=[None if (tr is None or p.periodic[FACE_AXIS[f]]) else tr for f, tr in enumerate(halos)],
                dt=cfl_dt_host(p),
                source=None if p.source.face < 0 else dict(face=p.source.face, amplitude=p.source.amplitude,
                                                           frequency=p.source.frequency, phase=p.source.phase,
                                                           t0=p.source.t0, tau=p.source.tau, lead=p.source.lead))


def make_batch(problems, precision=dg3d.MIXED):
    """A device batch of the problems."""
    return dg3d.Batch([batch_solve(p) for p in problems], precision=precision)
