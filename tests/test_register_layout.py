"""The register addressing of the stage kernel (stage_body in 3ddg_amd/csrc/dg3d_kernels.hip),
emulated on the CPU for every stage shape, both precisions and partial last workgroups.

Registers are group-blocked: G consecutive elements form a group whose state fields are 4 runs
of P = 4096 slots, field f of element e at (e / G) 4P + f P + (e % G) ES + node; the LDG fields
Q_x, Q_y, Q_z of a 7-field register follow from slot 4 fs in blocks of 3 runs. G ES <= P, and G is
a multiple of every elements-per-workgroup shape, so a workgroup's elements lie in one group. A
workgroup reaches them through buffer resources that start at its first element's slot.

The LSRK register is never read across a face, so its layout inside a group's block of 4P values
is the kernel's choice: pair-major FP64 (the group's (rho', u) pairs in the first half, its
(v, w) pairs in the second) and node-major FP32 (the four fields of a node together). The emulation
also covers the register laid out like the state ("field"), the kernel's earlier form.

The emulation checks that every (element, node slot, field) gets its own location inside the
values the context allocates (4 fs for the state and the LSRK register, 7 fs with Q), that every
access stays inside its buffer resource's range (the lanes past the workgroup's elements
included, whose resources span only the workgroup's own elements' pieces).
"""
import pytest

P = 4096
SHAPES = {2: (32, 16), 3: (8, 16), 4: (4, 8), 5: (2, 4), 6: (2, 1), 7: (2, 1), 8: (1, 2)}
GROUP = {2: 512, 3: 128, 4: 64, 5: 32, 6: 18, 7: 10, 8: 8}


def elem_stride(n):
    return n ** 3 if (n == 2 or n ** 3 % 32 == 0) else (n ** 3 + 31) // 32 * 32


def field_slots(E, n):
    return -(-E // GROUP[n]) * P


def du_layout(n, fp64):
    return "pair" if fp64 else "node"


def lane_accesses(n, epb, E, e0, tid, fs):
    """The lane's element (or None past the workgroup's elements) and, per register kind, the
    accesses (absolute value index, resource start, resource end) of: the state fields, the Q
    fields, and the LSRK register in both precisions."""
    es, G = elem_stride(n), GROUP[n]
    ne = min(E - e0, epb)
    le, node = divmod(tid, es)
    e = e0 + le
    io = le < epb and e < E
    ge = e if io else min(e0 + epb - 1, E - 1)
    g, eg = divmod(e0, G)
    gi = (ge - e0) * es + node
    out = {}
    base = g * 4 * P + eg * es
    out["state"] = [(base + gi + f * P, base, base + 3 * P + ne * es) for f in range(4)]
    qb = 4 * fs + g * 3 * P + eg * es
    out["q"] = [(qb + gi + a * P, qb, qb + 2 * P + ne * es) for a in range(3)]
    for fp64, kind in ((True, du_layout(n, True)), (False, du_layout(n, False)), (True, "field")):
        vsize = 8 if fp64 else 4
        d0 = g * 4 * P                                           # the group's block
        if kind == "field":
            lo = d0 + eg * es
            acc = [(lo + gi + f * P, lo, lo + 3 * P + ne * es) for f in range(4)]
        elif kind == "node":
            lo = d0 + eg * es * 4
            off = gi * 16                                        # bytes from the base
            acc = [(lo + off // vsize + f, lo, lo + ne * es * 16 // vsize) for f in range(4)]
        else:
            lo = d0 + eg * es * 2
            off = gi * 16
            hi = lo + (2 * P * 8 + ne * es * 16) // vsize
            acc = [(lo + off // vsize + f, lo, hi) if f < 2 else (lo + (off + P * 16) // vsize + f - 2, lo, hi)
                   for f in range(4)]
        out[("du64" if fp64 else "du32") + ("f" if kind == "field" else "")] = acc
    return (e if io else None), node, out


@pytest.mark.parametrize("n", sorted(SHAPES))
def test_registers_one_location_per_value(n):
    es = elem_stride(n)
    assert GROUP[n] * es <= P
    for epb in SHAPES[n]:
        assert GROUP[n] % epb == 0
        for E in (1, epb, 3 * epb + 1, 7, GROUP[n] + 1, 2 * GROUP[n] + epb):
            fs = field_slots(E, n)
            nt = (epb * es + 63) // 64 * 64
            seen = {k: {} for k in ("state", "q", "du64", "du32", "du64f")}
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


def test_state_slots_match_the_group_blocked_rule():
    """The state and Q slots the lanes address are dg3d_device.hpp's reg_slot:
    (e / G) 4P + f P + (e % G) ES + node for f < 4 and 4 fs + (e / G) 3P + (f - 4) P + (e % G) ES
    + node for the Q fields, at every order with a partial last group."""
    for n, (epb0, _) in SHAPES.items():
        es, G = elem_stride(n), GROUP[n]
        E = 2 * G + epb0
        fs = field_slots(E, n)
        for e0 in range(0, E, epb0):
            for tid in range((epb0 * es + 63) // 64 * 64):
                e, node, acc = lane_accesses(n, epb0, E, e0, tid, fs)
                if e is None:
                    continue
                for f, (i, _, _) in enumerate(acc["state"]):
                    assert i == (e // G) * 4 * P + f * P + (e % G) * es + node
                for a, (i, _, _) in enumerate(acc["q"]):
                    assert i == 4 * fs + (e // G) * 3 * P + a * P + (e % G) * es + node
