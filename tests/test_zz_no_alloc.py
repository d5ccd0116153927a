"""dg3d_reduce_noalloc (the allocation-free variant of dg3d_reduce) through the C ABI against
dg3d_reduce: every reduction kind, bit for bit, on two calls in a row, the second on a smaller
block, so that the per-device scratch array still holds the first call's partials beyond the
second call's blocks."""
import numpy as np
import pytest

from helpers import dg3d, heterogeneous, problem, random_state, small_c1, small_c2

pytestmark = pytest.mark.gpu

KINDS = (dg3d.REDUCE_MAXABS, dg3d.REDUCE_ENERGY, dg3d.REDUCE_RELL2, dg3d.REDUCE_DIFFL2)


def _loaded(p, seed):
    ctx = problem.make_context(p, num_registers=2, halos=False)
    ctx.set_state(0, random_state(p, seed=seed))
    ctx.set_state(1, random_state(p, seed=seed + 1))
    return ctx


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def test_reduce_noalloc_matches_reduce_two_calls():
    print("reduce_noalloc two calls, 16^3 N=4 then 3^3 N=2", flush=True)
    big = heterogeneous(small_c2(n_e=16, order=4, steps=1))   # 512 000 nodes: 1024 blocks
    small = heterogeneous(small_c2(n_e=3, order=2, steps=1))  # 729 nodes: 3 blocks
    cb = _loaded(big, 31)
    cs = _loaded(small, 41)
    for kind in KINDS:
        first = cb.reduce_noalloc(kind, 0, 1)
        second = cs.reduce_noalloc(kind, 0, 1)          # stale partials of `first` remain beyond 3 blocks
        assert _same(first, cb.reduce(kind, 0, 1)), kind
        assert _same(second, cs.reduce(kind, 0, 1)), kind
        assert np.all(np.isfinite(second[:2])) and np.abs(second).max() > 0
    cb.close()
    cs.close()


def test_reduce_noalloc_nan_sticky_and_errors():
    print("reduce_noalloc NaN and bad arguments", flush=True)
    p = small_c1()
    ctx = _loaded(p, 51)
    st = random_state(p, seed=52)
    st[777] = np.nan
    ctx.set_state(0, st)
    got = ctx.reduce_noalloc(dg3d.REDUCE_MAXABS, 0)
    assert np.isnan(got[0]) and _same(got, ctx.reduce(dg3d.REDUCE_MAXABS, 0))
    with pytest.raises(RuntimeError):
        ctx.reduce_noalloc(dg3d.REDUCE_MAXABS, 5)      # no such register
    with pytest.raises(RuntimeError):
        ctx.reduce_noalloc(17, 0)                      # no such kind
    # the context stays usable after the rejected calls
    st[777] = 0.0
    ctx.set_state(0, st)
    assert _same(ctx.reduce_noalloc(dg3d.REDUCE_ENERGY, 0), ctx.reduce(dg3d.REDUCE_ENERGY, 0))
    ctx.close()
