"""The LSRK register's addressing in the stage kernel (stage_body in 3ddg_amd/csrc/dg3d_kernels.hip),
emulated on the CPU for every stage shape, both precisions and partial last workgroups.

The register is never read across a face, so the kernel lays it out per precision and order:
pair-major FP64 (two arrays of (rho', u) and (v, w) pairs) and node-major FP32 (the four fields of
a node together) at every order except 7 nodes per line, where both stay field-major. The 16-byte
forms are addressed from each workgroup's first element with a buffer resource of the
workgroup's own size. The emulation checks that every (node slot, field) of every element gets
its own location inside the 4 * fs values the context allocates, that each access stays inside
its buffer resource's range, and that the lanes past the workgroup's elements only touch their
own workgroup's range.
"""
import pytest

SHAPES = {2: (32, 16), 3: (8, 16), 4: (4, 8), 5: (2, 4), 6: (2, 1), 7: (2, 1), 8: (1, 2)}


def elem_stride(n):
    return n ** 3 if (n == 2 or n ** 3 % 32 == 0) else (n ** 3 + 31) // 32 * 32


def field_slots(E, n):
    return (E * elem_stride(n) + 2047) // 2048 * 2048 + 256


def layout(n, fp64):
    es = elem_stride(n)
    if fp64:
        return "pair" if n != 7 else "field"
    return "node" if n != 7 else "field"


def locations(n, epb, E, fp64):
    """(value index of every (element, slot, field), the accesses of idle lanes) over all
    workgroups, following the kernel's arithmetic."""
    es = elem_stride(n)
    nt = (epb * es + 63) // 64 * 64
    fs = field_slots(E, n)
    size = 8 if fp64 else 4
    kind = layout(n, fp64)
    seen = {}
    for e0 in range(0, E, epb):
        ne = min(E - e0, epb)
        for tid in range(nt):
            le, node = divmod(tid, es)
            e = e0 + le
            io = le < epb and e < E
            ge = e if io else min(e0 + epb - 1, E - 1)
            if kind == "field":
                idx = [f * fs + ge * es + node for f in range(4)]
            else:
                off = ((ge - e0) * es + node) * 16          # bytes from the workgroup's base
                assert off + 16 <= ne * es * 16               # inside the buffer resource
                if kind == "node":                            # 4 floats per node
                    base = e0 * es * 4
                    idx = [base + off // size + f for f in range(4)]
                else:                                         # 2 pair arrays of 2 doubles
                    base = e0 * es * 2
                    idx = [base + off // size + f if f < 2 else 2 * fs + base + off // size + f - 2
                           for f in range(4)]
            for i in idx:
                assert 0 <= i < 4 * fs
            if io:
                for f, i in enumerate(idx):
                    assert (e, node, f) not in seen
                    seen[(e, node, f)] = i
            elif kind == "node":
                lo = e0 * es * 4
                assert lo <= idx[0] and idx[3] < lo + ne * es * 4
            elif kind == "pair":
                lo = e0 * es * 2
                assert lo <= idx[0] and idx[1] < lo + ne * es * 2
                assert 2 * fs + lo <= idx[2] and idx[3] < 2 * fs + lo + ne * es * 2
    return seen


@pytest.mark.parametrize("n", sorted(SHAPES))
@pytest.mark.parametrize("fp64", [True, False])
def test_lsrk_register_one_location_per_value(n, fp64):
    for epb in SHAPES[n]:
        for E in (1, epb, 3 * epb + 1, 7):
            seen = locations(n, epb, E, fp64)
            es = elem_stride(n)
            assert len(seen) == 4 * E * es
            assert len(set(seen.values())) == len(seen), (n, epb, E, fp64)
