"""System-level properties of the GPU path, which no formula restated from the reference supplies:
the mass the scheme conserves in a periodic box, and the second harmonic of a plane tone burst
against the Fubini solution (the closed form of weak plane-wave steepening)."""
import importlib

import numpy as np
import pytest

from helpers import dg3d, discrete_mass, periodic_mass_box, plane_burst_column, problem

bench = importlib.import_module("bench")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
def test_mass_conserved_in_periodic_box(scheme):
    """Periodic 8^3 box, no sponge, Gaussian pulse; c and B/A vary per element (B/A = 0 on some,
    so both forms of the mass flux meet at faces), rho0 uniform. Each face's mass flux is one
    value seen from both sides (Element.cpp:541-600 is symmetric under exchanging the sides), and
    the volume term sums to zero, so the discrete mass stays constant to rounding over 200 steps.
    A flux evaluated differently from the two sides, for instance with Z1 and Z2 exchanged on one
    of them, would change it at first order. (With rho0 varying the lifted flux is scaled by each
    side's own rho0, so the reference's scheme conserves mass only for uniform rho0.)"""
    print(f"mass_conserved_in_periodic_box scheme={scheme}", flush=True)
    p = periodic_mass_box(200)
    ctx = problem.make_context(p)
    dt = problem.cfl_dt_host(p)
    m0, a0 = discrete_mass(p, ctx.get_state(0))
    t = ctx.step(scheme, dt, p.steps, 0.0)
    st = ctx.get_state(0)
    ctx.close()
    m1, a1 = discrete_mass(p, st)
    assert t > 0 and a0 > 0
    assert np.abs(st.reshape(4, -1)[1:]).max() > 0          # the pulse has moved
    assert abs(m1 - m0) < 1e-12 * a0, (m0, m1, a0, a1)


def test_second_harmonic_of_a_plane_burst_matches_fubini():
    """The 1 MHz, 1e6 Pa tone burst of the 64^3 N=4 workload in a column of 2 x 2 x 64 elements,
    periodic across (x, y), so that the wave is exactly plane. At the on-axis probe 2.25 mm deep
    the fundamental and the second harmonic of the pressure trace (Hann window over the whole
    burst) match the first-order Fubini solution p1 + z beta/(2 rho0 c^3) d(p1^2)/dt, beta =
    1 + B/(2A), to 3 % and 5 % (measured: 0.984 and 1.022 of it; the fundamental's deficit is the
    scheme's dissipation over 3 elements at about 2 elements per wavelength). Through beta, a
    missing 1/2 in the B/A term of the pressure (kernelFuncs.cpp:17-20) would raise the harmonic by
    about 70 %, and dropping rho' from the mass flux (kernelFuncs.cpp:26-30) would lower it by
    about 30 %."""
    print("second_harmonic_of_a_plane_burst", flush=True)
    p = plane_burst_column()
    dt = problem.cfl_dt_host(p)
    pr = bench.probe_node(p)
    depth = bench.probe_depth(p, pr)
    assert abs(depth - 2.25e-3) < 1e-12
    steps = bench.burst_window_steps(p, dt, depth)
    ctx = problem.make_context(p)
    ctx.set_probes(np.array([pr]), steps + 1)
    ctx.step(dg3d.LSRK54, dt, steps, 0.0)
    tr = ctx.get_probes()[:, 0]
    ctx.close()
    f = p.source.frequency
    h1, h2 = bench.dft_bin(tr, dt, f), bench.dft_bin(tr, dt, 2 * f)
    e2, e1, sigma = bench.fubini_estimate(p, depth, dt, len(tr))
    assert 0.005 < sigma < 0.05                              # weak, pre-shock regime
    assert abs(h1 / e1 - 1.0) < 0.03, (h1, e1)
    assert abs(h2 / e2 - 1.0) < 0.05, (h2, e2, h2 / e2)
