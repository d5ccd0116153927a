"""The C-ABI library loads and exports every entry point include/dg3d.h declares (no device calls)."""
import ctypes
import os
import re

from helpers import dg3d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dg3d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dg3d_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(dg3d.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_symbols()) == set(dg3d._SIGS)


def test_abi_version_and_create_error_without_device():
    L = dg3d.lib()
    assert L.dg3d_abi_version() == 1
    # invalid descriptors are rejected on the host before any device call
    d = dg3d.Desc()
    d.order = 0
    h = ctypes.c_void_p()
    assert L.dg3d_create(ctypes.byref(d), ctypes.byref(h)) == dg3d.E_UNSUPPORTED
    assert b"order" in L.dg3d_create_error()
    assert not h.value


def _orders_by_definition(order, nE, epb, G):
    """The workgroup orders written out with plain integer arithmetic (the stage kernels use
    host-prepared magic divisors instead)."""
    X = 8
    out = []
    for w in range(G):
        if order in (1, 3):
            per, rem, x, i = G // X, G % X, w % X, w // X
            s = x * per + min(x, rem) + i
            if order == 3 and nE[2] % epb == 0:
                kg = nE[2] // epb
                tk, tj = min(kg, 8), min(nE[1], 8)
                if kg % tk == 0 and nE[1] % tj == 0:
                    per_jb, per_kb = tj * nE[0] * kg, nE[0] * tj * tk
                    jb, r = divmod(s, per_jb)
                    kb, r = divmod(r, per_kb)
                    ii, r = divmod(r, tj * tk)
                    jj, kk = divmod(r, tk)
                    s = (ii * nE[1] + jb * tj + jj) * kg + kb * tk + kk
            out.append(s)
        elif order == 2:
            if nE[2] % epb != 0:
                C = 4
            else:
                sy = nE[2] // epb
                C = sy // 8 if sy % 8 == 0 else sy
            full = G // (X * C) * (X * C)
            if w >= full:
                out.append(w)
            else:
                x, i = w % X, w // X
                out.append((i // C) * (X * C) + x * C + i % C)
        else:
            out.append(w)
    return out


def test_workgroup_orders_are_permutations():
    """Every workgroup order visits each element group exactly once, on meshes that divide into
    the tiles and on meshes that do not, and equals its definition."""
    cases = [((64, 64, 64), 2), ((40, 40, 40), 1), ((13, 7, 5), 1), ((9, 16, 24), 2), ((8, 8, 8), 4),
             ((5, 3, 17), 4), ((1, 1, 1), 1), ((48, 48, 48), 4), ((3, 11, 16), 16), ((30, 9, 14), 2)]
    for nE, epb in cases:
        G = nE[0] * nE[1] * ((nE[2] + epb - 1) // epb) if nE[2] % epb == 0 else \
            (nE[0] * nE[1] * nE[2] + epb - 1) // epb
        for order in range(4):
            got = dg3d.group_order(order, nE, epb, G)
            assert sorted(got.tolist()) == list(range(G)), (nE, epb, order)
            if G <= 70000:
                assert got.tolist() == _orders_by_definition(order, nE, epb, G), (nE, epb, order)


def test_create_rejects_bad_descriptors_on_the_host():
    """Descriptors the library cannot run are rejected before any device call: 2^29 node slots or more
    in one block (32-bit offsets within a field), a bad contraction or precision, a missing
    operator, an empty axis."""
    import numpy as np
    L = dg3d.lib()
    n = 8
    arrs = [np.zeros(n * n), np.zeros(n * n), np.zeros(n), np.zeros(n)]

    def desc(order=7, nE=(8, 8, 8), precision=dg3d.FP64, contraction=0, regs=1, missing=False):
        d = dg3d.Desc()
        d.order = order
        for a in range(3):
            d.num_elements[a] = nE[a]
        d.precision = precision
        d.contraction = contraction
        d.num_registers = regs
        ptrs = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in arrs]
        d.mass, d.stiff, d.lgl_nodes, d.lgl_weights = ptrs
        if missing:
            d.stiff = ctypes.POINTER(ctypes.c_double)()
        return d

    def create(d):
        h = ctypes.c_void_p()
        rc = L.dg3d_create(ctypes.byref(d), ctypes.byref(h))
        assert not h.value
        return rc, L.dg3d_create_error()

    # 102^3 elements at N=7 take 512 slots each: 5.4e8 >= 2^29 node slots
    rc, msg = create(desc(nE=(102, 102, 102)))
    assert rc == dg3d.E_INVALID and b"2^29" in msg
    rc, msg = create(desc(contraction=7))
    assert rc == dg3d.E_INVALID and b"contraction" in msg
    rc, msg = create(desc(precision=5))
    assert rc == dg3d.E_INVALID and b"precision" in msg
    rc, msg = create(desc(missing=True))
    assert rc == dg3d.E_INVALID and b"required" in msg
    rc, msg = create(desc(nE=(4, 0, 4)))
    assert rc == dg3d.E_INVALID and b"num_elements" in msg
    rc, msg = create(desc(regs=0))
    assert rc == dg3d.E_INVALID and b"num_registers" in msg


def test_batch_rejects_bad_solves_on_the_host():
    """dg3d_batch_create checks every solve description before any device call: time step,
    order, and a boundary source's face, periodicity and envelope width."""
    from helpers import problem
    p = problem.config_c1()
    src = dict(face=0, amplitude=1e3, frequency=1e5, phase=0.0, t0=0.0, tau=1e-5)

    def rejected(**changes):
        d = problem.batch_solve(p)
        d.update(changes)
        try:
            dg3d.Batch([d])
        except dg3d.Dg3dError as e:
            return e.args
        raise AssertionError(f"accepted {changes}")

    assert "dt must be positive" in str(rejected(dt=-1.0))
    assert "order must be 1..7" in str(rejected(order=9))
    assert "source face must be 0..5" in str(rejected(source=dict(src, face=6)))
    assert "source on a periodic face" in str(rejected(source=dict(src, face=4), periodic=(1, 0, 0)))
    assert "source tau must be positive" in str(rejected(source=dict(src, tau=0.0)))
