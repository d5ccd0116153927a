"""dg3d_step replays each step as a captured HIP graph. The replays must equal direct launches bit
for bit, call after call, with the program's own work and synchronisations between the calls:
boundary source with per-node delays (device clock), probes (device row counter), initial-condition
halos (first step launched directly), LDG diffusion (Q kernel in the graph), both schemes, both
precisions, and graphs dropped by a variant change and by a new dt."""
import numpy as np
import pytest

from helpers import dg3d, heterogeneous, problem, random_state, small_c1, small_c2, small_c5

pytestmark = pytest.mark.gpu


def _probe_nodes(p, count=5):
    rng = np.random.default_rng(11)
    return rng.integers(0, p.dof, size=count).astype(np.int64)


def _calls(p, scheme, precision, graphs, with_q=False):
    """Several dg3d_step calls of uneven length; the state, time and probes after each."""
    ctx = problem.make_context(p, precision=precision, with_q=with_q)
    ctx.set_graphs(graphs)
    ctx.set_probes(_probe_nodes(p), 64)
    dt = problem.cfl_dt_host(p)
    t = 0.0
    out = []
    for i, n in enumerate([1, 4, 3, 2, 5]):
        if i == 3:
            ctx.set_variant(scheme, 1, 3)   # drops the captured graphs
        step_dt = dt if i < 4 else 0.5 * dt
        t = ctx.step(scheme, step_dt, n, t)
        energy = ctx.reduce(dg3d.REDUCE_ENERGY, 0)   # eager work between the replays
        out.append((ctx.get_state(0), t, energy[0]))
    probes = ctx.get_probes()
    ctx.close()
    return out, probes


def _check(p, scheme, precision, with_q=False):
    direct, pd = _calls(p, scheme, precision, False, with_q)
    graph, pg = _calls(p, scheme, precision, True, with_q)
    for i, ((sd, td, ed), (sg, tg, eg)) in enumerate(zip(direct, graph)):
        assert np.linalg.norm(sd) > 0
        assert td == tg, (i, td, tg)
        assert np.array_equal(sd, sg), (i, np.abs(sd - sg).max())
        assert ed == eg, (i, ed, eg)
    assert len(pd) == len(pg) == 15
    assert np.array_equal(np.asarray(pd), np.asarray(pg))


@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
@pytest.mark.parametrize("precision", [dg3d.FP64, dg3d.MIXED])
def test_graphs_burst_source(scheme, precision):
    print(f"graphs_burst_source scheme={scheme} precision={precision}", flush=True)
    _check(heterogeneous(small_c2(n_e=4, order=3)), scheme, precision)


def test_graphs_focused_source_with_delays():
    print("graphs_focused_source", flush=True)
    _check(small_c5(n_e=4), dg3d.LSRK54, dg3d.FP64)


@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
def test_graphs_ic_halos(scheme):
    """The Gaussian IC's halo traces are read by the first stage of the first call only."""
    print(f"graphs_ic_halos scheme={scheme}", flush=True)
    _check(small_c1(), scheme, dg3d.FP64)


@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
def test_graphs_ldg_diffusion(scheme):
    print(f"graphs_ldg scheme={scheme}", flush=True)
    p = heterogeneous(small_c2(n_e=3, order=2))
    p.diffusion = np.full(p.E, 1.0e-3)
    _check(p, scheme, dg3d.FP64, with_q=True)


@pytest.mark.parametrize("order,n_e", [(5, 5), (1, 9), (6, 5)])
def test_probe_rows_are_the_pressure_at_the_probe_nodes(order, n_e):
    """Each probe row holds P = c^2 (rho' + rho'^2 (B/A) / (2 rho0)) at the probe nodes (host
    numbering) after a step: checked against the state read back after every single step, on meshes
    whose last group of elements is partial (125 elements in groups of 18 and 10 at 6 and 7 nodes per
    line, 729 in groups of 512 at 2), with heterogeneous material and random state."""
    print(f"probe_rows order={order} n_e={n_e}", flush=True)
    p = heterogeneous(small_c2(n_e=n_e, order=order), seed=7)
    rng = np.random.default_rng(3)
    nodes = np.unique(np.concatenate([rng.integers(0, p.dof, size=40), [0, p.dof - 1]])).astype(np.int64)
    ctx = problem.make_context(p)
    ctx.set_state(0, random_state(p, seed=4))
    ctx.set_probes(nodes, 8)
    dt = problem.cfl_dt_host(p)
    t = 0.0
    expect = []
    np3 = p.n ** 3
    e = nodes // np3
    for _ in range(3):
        t = ctx.step(dg3d.LSRK54, dt, 1, t)
        rho = ctx.get_state(0)[: p.dof][nodes]
        c, r0, ba = p.c0[e], p.rho0[e], p.b_over_a[e]
        expect.append(c * c * (rho + ba * 0.5 / r0 * rho * rho))
    got = ctx.get_probes()
    ctx.close()
    assert got.shape == (3, nodes.size)
    expect = np.asarray(expect)
    assert np.abs(expect).max() > 0
    np.testing.assert_allclose(got, expect, rtol=1e-13, atol=1e-13 * np.abs(expect).max())


@pytest.mark.parametrize("scheme", [dg3d.LSRK54, dg3d.RK33])
def test_step_calls_equal_stage_calls_with_source(scheme):
    """dg3d_step writes the source face's traces of all stages of a step in one launch at the
    step's start, from the device clock; dg3d_stage writes one stage's trace just before that
    stage, at the host's stage time. The times are the same FP64 sums, so the states agree bit
    for bit, with graphs and with direct launches."""
    print(f"step_calls_equal_stage_calls scheme={scheme}", flush=True)
    p = heterogeneous(small_c2(n_e=4, order=3))
    dt = problem.cfl_dt_host(p)
    ns = 5 if scheme == dg3d.LSRK54 else 3
    states = []
    for graphs in (True, False):
        ctx = problem.make_context(p, precision=dg3d.FP64)
        ctx.set_variant(scheme, 2, 1)
        ctx.set_graphs(graphs)
        t = ctx.step(scheme, dt, 7, 0.0)
        states.append((ctx.get_state(0), t))
        ctx.close()
    ctx = problem.make_context(p, precision=dg3d.FP64)
    ctx.set_variant(scheme, 2, 1)
    t = 0.0
    for _ in range(7):
        for st in range(ns):
            ctx.stage(scheme, st, dt, t)
        t = t + dt
    ref = ctx.get_state(0)
    ctx.close()
    assert np.linalg.norm(ref) > 0
    for s, ts in states:
        assert ts == t
        assert np.array_equal(s, ref), np.abs(s - ref).max()
