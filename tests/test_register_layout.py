"""The register addressing of the stage kernel (stage_body in 3ddg_amd/csrc/dg3d_kernels.hip),
emulated on the CPU for every stage shape, both precisions and partial last workgroups.

Registers are element-blocked: the four state fields of element e are one block of 4 ES slots
(field f at (4 e + f) ES + node), and the LDG fields Q_x, Q_y, Q_z of a 7-field register follow
from slot 4 fs in blocks of 3 ES slots. A workgroup reaches its elements' blocks through buffer
resources that start at its first element's block and span its own elements only.

The LSRK register is never read across a face, so its layout within an element's block is the
kernel's choice: pair-major FP64 (the element's (rho', u) pairs, then its (v, w) pairs) and
node-major FP32 (the four fields of a node together) at every order except 7 nodes per line,
where it is laid out like the state.

The emulation checks that every (element, node slot, field) gets its own location inside the
values the context allocates (4 fs for the state and the LSRK register, 7 fs with Q), that every
access stays inside its buffer resource's range, and that the lanes past the workgroup's elements
touch only their own workgroup's range.
"""
import pytest

SHAPES = {2: (32, 16), 3: (8, 16), 4: (4, 8), 5: (2, 4), 6: (2, 1), 7: (2, 1), 8: (1, 2)}


def elem_stride(n):
    return n ** 3 if (n == 2 or n ** 3 % 32 == 0) else (n ** 3 + 31) // 32 * 32


def field_slots(E, n):
    return (E * elem_stride(n) + 2047) // 2048 * 2048 + 256


def du_layout(n, fp64):
    if fp64:
        return "pair" if n != 7 else "field"
    return "node" if n != 7 else "field"


def lane_accesses(n, epb, E, e0, tid, fs):
    """The lane's element (or None past the workgroup's elements) and, per access, (absolute
    value index, resource start, resource end) for: the state fields, the Q fields, and the LSRK
    register in both precisions."""
    es = elem_stride(n)
    ne = min(E - e0, epb)
    le, node = divmod(tid, es)
    e = e0 + le
    io = le < epb and e < E
    ge = e if io else min(e0 + epb - 1, E - 1)
    out = {}
    # state: resource over the run of the workgroup's blocks; gi + f * es
    base, size = e0 * 4 * es, ne * 4 * es
    gi = (ge - e0) * 4 * es + node
    out["state"] = [(base + gi + f * es, base, base + size) for f in range(4)]
    # Q fields: blocks of 3 es from 4 fs
    qb, qs = 4 * fs + e0 * 3 * es, ne * 3 * es
    qi = (ge - e0) * 3 * es + node
    out["q"] = [(qb + qi + a * es, qb, qb + qs) for a in range(3)]
    for fp64 in (True, False):
        kind = du_layout(n, fp64)
        vsize = 8 if fp64 else 4
        lo, hi = e0 * 4 * es, e0 * 4 * es + ne * 4 * es           # in register values
        if kind == "field":
            idx = [lo + gi + f * es for f in range(4)]
        elif kind == "node":
            off = ((ge - e0) * es + node) * 16                      # bytes from the base
            assert off + 16 <= ne * 4 * es * vsize
            idx = [lo + off // vsize + f for f in range(4)]
        else:
            off = ((ge - e0) * 2 * es + node) * 16
            assert off + es * 16 + 16 <= ne * 4 * es * vsize
            idx = [lo + off // vsize + f if f < 2 else lo + (off + es * 16) // vsize + f - 2 for f in range(4)]
        out["du64" if fp64 else "du32"] = [(i, lo, hi) for i in idx]
    return (e if io else None), node, out


@pytest.mark.parametrize("n", sorted(SHAPES))
def test_registers_one_location_per_value(n):
    es = elem_stride(n)
    for epb in SHAPES[n]:
        for E in (1, epb, 3 * epb + 1, 7):
            fs = field_slots(E, n)
            nt = (epb * es + 63) // 64 * 64
            seen = {k: {} for k in ("state", "q", "du64", "du32")}
            for e0 in range(0, E, epb):
                for tid in range(nt):
                    e, node, acc = lane_accesses(n, epb, E, e0, tid, fs)
                    for kind, lst in acc.items():
                        total = 7 * fs if kind == "q" else 4 * fs
                        for f, (i, lo, hi) in enumerate(lst):
                            assert lo <= i < hi <= total, (kind, n, epb, E, e0, tid)
                            if e is not None:
                                assert (e, node, f) not in seen[kind]
                                seen[kind][(e, node, f)] = i
            for kind, m in seen.items():
                nf = 3 if kind == "q" else 4
                assert len(m) == nf * E * es, (kind, n, epb, E)
                assert len(set(m.values())) == len(m), (kind, n, epb, E)


def test_state_slots_match_the_element_blocked_rule():
    """The state and Q slots the lanes address are dg3d_device.hpp's reg_slot:
    (4 e + f) ES + node for f < 4, 4 fs + (3 e + f - 4) ES + node for the Q fields."""
    n, epb, E = 5, 2, 5
    es = elem_stride(n)
    fs = field_slots(E, n)
    for e0 in range(0, E, epb):
        for tid in range((epb * es + 63) // 64 * 64):
            e, node, acc = lane_accesses(n, epb, E, e0, tid, fs)
            if e is None:
                continue
            for f, (i, _, _) in enumerate(acc["state"]):
                assert i == (4 * e + f) * es + node
            for a, (i, _, _) in enumerate(acc["q"]):
                assert i == 4 * fs + (3 * e + a) * es + node
