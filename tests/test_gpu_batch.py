"""Batched solves (dg3d_batch_*): independent blocks of different sizes and orders advanced by
per-order persistent launches replayed from HIP graphs.

Each solve of a batch must follow the same trajectory as dg3d_step on a context of its own (the
per-node arithmetic is shared), and, in FP64, the oracle to the drop-in bar of 1e-12.
"""
import numpy as np
import pytest

from helpers import dg3d, heterogeneous, problem, rel_l2
from oracle_bind import Oracle, load_problem_into

pytestmark = pytest.mark.gpu


def pulse_problem(order, nE, periodic=(0, 0, 0), ba=0.0, amp=1e5, hetero_seed=None):
    nE = tuple(nE)
    h = tuple(1.0 / n for n in nE)
    p = problem._base(f"b{order}_{nE[0]}{nE[1]}{nE[2]}", order, nE, (0, 0, 0), problem._uniform_widths(nE, h),
                      1500.0, 1000.0, ba)
    p.ic = dict(gauss=(1, 1, 1), amp_rho=amp / 1500.0 ** 2, amp_vel=(0.0, 0.0, 0.0), sigma=1.0 / (2 * 0.08 ** 2),
                center=(0.5, 0.5, 0.5))
    p.periodic = periodic
    if not any(periodic):
        problem.set_sponge(p, (2, 2, 2, 2, 2, 2), 3 * 1500.0 / (2 * min(h)), ramp=1)
    if hetero_seed is not None:
        heterogeneous(p, seed=hetero_seed)
    p.scheme, p.cfl = dg3d.LSRK54, 0.3
    return p


def mixed_set():
    """Every order 1..7, uneven element counts, heterogeneous material, one periodic block."""
    return [
        pulse_problem(1, (6, 5, 4), ba=5.0),
        pulse_problem(2, (4, 4, 5), hetero_seed=3),
        pulse_problem(3, (4, 4, 4), ba=3.0),
        pulse_problem(4, (3, 4, 3), hetero_seed=4),
        pulse_problem(5, (3, 3, 3), periodic=(1, 1, 1)),
        pulse_problem(6, (2, 3, 3), ba=8.0),
        pulse_problem(7, (2, 2, 3), hetero_seed=5),
        pulse_problem(3, (5, 3, 4), ba=2.0),
    ]


def single_run(p, precision, steps):
    ctx = problem.make_context(p, precision=precision)
    ctx.step(dg3d.LSRK54, problem.cfl_dt_host(p), steps, 0.0)
    st = ctx.get_state(0)
    ctx.close()
    return st


@pytest.mark.parametrize("precision", [dg3d.FP64, dg3d.MIXED])
def test_batch_matches_single_contexts(precision):
    """Graph-replayed batch steps, in two calls, equal dg3d_step on one context per solve."""
    print(f"batch_matches_single precision={precision}", flush=True)
    ps = mixed_set()
    b = problem.make_batch(ps, precision=precision)
    b.step(3)   # the first step (halo traces) runs directly, the rest from graphs
    b.step(4)   # continues from the other arena parity
    info = b.info()
    assert info["orders"] == sorted({p.order for p in ps}, reverse=True)
    for i, p in enumerate(ps):
        ref = single_run(p, precision, 7)
        got = b.get_state(i)
        assert np.linalg.norm(ref) > 0
        err = rel_l2(got, ref)
        assert err.max() < 1e-14, (i, p.name, err)
    b.close()


def test_batch_fp64_vs_oracle():
    print("batch_fp64_vs_oracle", flush=True)
    ps = [pulse_problem(2, (4, 3, 5), hetero_seed=7), pulse_problem(4, (3, 3, 4), ba=5.0),
          pulse_problem(6, (2, 3, 2), periodic=(0, 1, 0))]
    steps = 12
    b = problem.make_batch(ps, precision=dg3d.FP64)
    b.step(steps)
    for i, p in enumerate(ps):
        o = Oracle(p, literal=False, num_registers=4)
        load_problem_into(o, p, 0)
        o.set_time(0.0)
        dt = problem.cfl_dt_host(p)
        for _ in range(steps):
            o.lsrk54(dt, 0, 1, 2)
        r = o.get_state(0)
        err = rel_l2(b.get_state(i), r)
        assert np.linalg.norm(r) > 0
        assert err.max() < 1e-12, (p.name, err)
    b.close()


def test_batch_mixed_vs_fp64():
    print("batch_mixed_vs_fp64", flush=True)
    ps = mixed_set()
    b64 = problem.make_batch(ps, precision=dg3d.FP64)
    b32 = problem.make_batch(ps, precision=dg3d.MIXED)
    for b in (b64, b32):
        b.step(20)
    for i, p in enumerate(ps):
        err = rel_l2(b32.get_state(i), b64.get_state(i))
        assert err.max() < 1e-5, (p.name, err)
        assert err.max() > 0.0
    b64.close()
    b32.close()


@pytest.mark.parametrize("precision", [dg3d.FP64, dg3d.MIXED])
def test_batch_shapes_bitwise_identical(precision):
    """Every variant (two elements-per-workgroup shapes, persistent or one group per workgroup,
    early or late RK-register loads) and every stage schedule (parallel branches per order, one
    stream, two balanced branches, the largest order beside the others, every order in one
    persistent launch) computes the same per-node arithmetic."""
    print(f"batch_shapes precision={precision}", flush=True)
    ps = mixed_set()
    states = []
    for v in range(8):
        b = problem.make_batch(ps, precision=precision)
        b.set_variant(v)
        b.step(5)
        states.append([b.get_state(i) for i in range(len(ps))])
        b.close()
    for v in range(1, 8):
        for i in range(len(ps)):
            assert np.array_equal(states[0][i], states[v][i]), (v, i)
    for sc in range(5):
        b = problem.make_batch(ps, precision=precision)
        b.set_variant(5)
        b.set_schedule(sc)
        assert b.schedule() == sc
        b.step(5)
        for i in range(len(ps)):
            assert np.array_equal(states[0][i], b.get_state(i)), ("schedule", sc, i)
        b.close()
    b = problem.make_batch(ps, precision=precision)
    v = b.autotune()
    assert len(v) == b.info()["num_orders"] and all(0 <= x < 8 for x in v)
    assert 0 <= b.schedule() < 5
    b.close()


def test_batch_rejects_bad_input():
    print("batch_rejects_bad_input", flush=True)
    p = pulse_problem(3, (3, 3, 3))
    d = problem.batch_solve(p)
    d["dt"] = -1.0
    with pytest.raises(dg3d.Dg3dError):
        dg3d.Batch([d])
    d = problem.batch_solve(p)
    d["order"] = 9
    with pytest.raises(dg3d.Dg3dError):
        dg3d.Batch([d])


def _with_source(p, face, lead=False):
    """The tone burst of the 64^3 workload on one face of a pulse problem (the sponge stays)."""
    f0 = 2.0e5
    tau = 2.0 / f0
    p.source = problem.Source(face=face, amplitude=2.0e5, frequency=f0, phase=problem.seeded_phase(face),
                              t0=0.2 * tau, tau=tau)
    if lead:
        a = problem.FACE_AXIS[face]
        b, c = (1, 2) if a == 0 else ((0, 2) if a == 1 else (0, 1))
        rng = np.random.default_rng(face)
        p.source.lead = 1e-6 * rng.random(p.num_elements[b] * p.num_elements[c] * p.n ** 2)
    return p


@pytest.mark.parametrize("precision", [dg3d.FP64, dg3d.MIXED])
def test_batch_sources_match_single_contexts(precision):
    """Solves with a boundary source (several faces, one with per-node delays, with and without
    initial halos) follow dg3d_step on contexts of their own bit for bit, over two calls."""
    print(f"batch_sources precision={precision}", flush=True)
    ps = [
        _with_source(pulse_problem(3, (4, 4, 5), ba=5.0), 0),
        _with_source(pulse_problem(2, (5, 3, 4), hetero_seed=8), 5, lead=True),
        _with_source(pulse_problem(4, (3, 4, 3)), 2),
        pulse_problem(3, (3, 3, 3)),
        _with_source(pulse_problem(5, (3, 3, 2), periodic=(1, 0, 0)), 1, lead=True),
    ]
    b = problem.make_batch(ps, precision=precision)
    b.step(3)
    b.step(4)
    for i, p in enumerate(ps):
        ref = single_run(p, precision, 7)
        got = b.get_state(i)
        assert np.linalg.norm(ref) > 0
        assert np.array_equal(got, ref), (i, p.name, rel_l2(got, ref))
    b.close()
    # the source is not negligible: without it the first solve ends elsewhere
    plain = single_run(pulse_problem(3, (4, 4, 5), ba=5.0), precision, 7)
    assert rel_l2(plain, single_run(ps[0], precision, 7)).max() > 1e-6


def test_batch_sources_fp64_vs_oracle():
    """Sourced solves of a batch against the oracle's LSRK54 with the same burst, 1e-12 in FP64."""
    print("batch_sources_fp64_vs_oracle", flush=True)
    ps = [_with_source(pulse_problem(2, (4, 3, 5), hetero_seed=7), 3, lead=True),
          _with_source(pulse_problem(4, (3, 3, 4), ba=5.0), 0),
          pulse_problem(3, (3, 3, 3))]
    steps = 12
    b = problem.make_batch(ps, precision=dg3d.FP64)
    b.step(steps)
    for i, p in enumerate(ps):
        o = Oracle(p, literal=False, num_registers=4)
        load_problem_into(o, p, 0)
        o.set_time(0.0)
        dt = problem.cfl_dt_host(p)
        for _ in range(steps):
            o.lsrk54(dt, 0, 1, 2)
        r = o.get_state(0)
        err = rel_l2(b.get_state(i), r)
        print(f"  solve {i} {p.name}: {err.max():.2e}", flush=True)
        assert np.linalg.norm(r) > 0
        assert err.max() < 1e-12, (p.name, err)
    b.close()
