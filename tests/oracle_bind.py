"""ctypes binding of the CPU oracle (oracle/liboracle.so). Test infrastructure: the checker."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_matrix.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OracleDesc(ctypes.Structure):
    _fields_ = [
        ("order", ctypes.c_int),
        ("num_elements", ctypes.c_int * 3),
        ("periodic", ctypes.c_int * 3),
        ("literal", ctypes.c_int),
        ("num_registers", ctypes.c_int),
        ("mass", _dp),
        ("stiff", _dp),
        ("lgl_nodes", _dp),
        ("lgl_weights", _dp),
        ("corner", ctypes.c_double * 3),
        ("width", _dp * 3),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_LIB)
        sig = {
            "oracle_create": (ctypes.c_void_p, [ctypes.POINTER(OracleDesc)]),
            "oracle_destroy": (None, [ctypes.c_void_p]),
            "oracle_set_threads": (None, [ctypes.c_int]),
            "oracle_set_material": (None, [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]),
            "oracle_node_coord": (None, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
            "oracle_fill_gaussian": (None, [ctypes.c_void_p, ctypes.c_int, _ip, ctypes.c_double, _dp, ctypes.c_double, _dp]),
            "oracle_set_state": (None, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int]),
            "oracle_get_state": (None, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int]),
            "oracle_set_halo": (None, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]),
            "oracle_get_halo": (None, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]),
            "oracle_zero": (None, [ctypes.c_void_p, ctypes.c_int]),
            "oracle_copy": (None, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
            "oracle_axpby": (None, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double]),
            "oracle_operate": (None, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "oracle_communicate": (None, [ctypes.c_void_p, ctypes.c_int]),
            "oracle_set_source": (None, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, _dp]),
            "oracle_set_time": (None, [ctypes.c_void_p, ctypes.c_double]),
            "oracle_rk33": (None, [ctypes.c_void_p, ctypes.c_double] + [ctypes.c_int] * 7),
            "oracle_lsrk54": (None, [ctypes.c_void_p, ctypes.c_double] + [ctypes.c_int] * 6),
            "oracle_forward_euler": (None, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
            "oracle_rk33_blocks": (None, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_double, ctypes.c_int]),
            "oracle_maxabs": (None, [ctypes.c_void_p, ctypes.c_int, _dp]),
            "oracle_energy": (None, [ctypes.c_void_p, ctypes.c_int, _dp]),
            "oracle_sponge_sigma": (None, [_ip, _ip, _ip, _dp, _ip, ctypes.c_double, ctypes.c_int, _dp]),
            "oracle_line_solve": (None, [ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


class Oracle:
    """CPU restatement of one block (registers = the reference's Grid copies)."""

    def __init__(self, p, literal=False, num_registers=4, threads=0):
        self.L = lib()
        self.p = p
        d = OracleDesc()
        d.order = p.order
        d.num_elements[:] = list(p.num_elements)
        d.periodic[:] = [int(bool(v)) for v in p.periodic]
        d.literal = int(literal)
        d.num_registers = num_registers
        self._keep = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in
                      (p.mass, p.stiff, p.nodes, p.weights, p.widths[0], p.widths[1], p.widths[2])]
        d.mass, d.stiff, d.lgl_nodes, d.lgl_weights = (k.ctypes.data_as(_dp) for k in self._keep[:4])
        d.corner[:] = [float(v) for v in p.corner]
        for a in range(3):
            d.width[a] = self._keep[4 + a].ctypes.data_as(_dp)
        if threads:
            self.L.oracle_set_threads(threads)
        self.h = self.L.oracle_create(ctypes.byref(d))
        if not self.h:
            raise RuntimeError("oracle_create failed")
        self.nf_size = p.E * p.n ** 3
        self.set_material(p.c0, p.rho0, p.b_over_a, p.sigma, p.diffusion)
        s = p.source
        if s.face >= 0:
            self.set_source(s.face, s.amplitude, s.frequency, s.phase, s.t0, s.tau, s.lead)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oracle_destroy(self.h)
            self.h = None

    def set_material(self, c0, rho0, ba, sigma, diff):
        self._mat = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (c0, rho0, ba, sigma, diff)]
        self.L.oracle_set_material(self.h, *[_p(a) for a in self._mat])

    def fill_ic(self, reg=0):
        ic = self.p.ic
        g = (ctypes.c_int * 3)(*ic["gauss"])
        av = np.asarray(ic["amp_vel"], float)
        cen = np.asarray(ic["center"], float)
        self.L.oracle_fill_gaussian(self.h, reg, g, float(ic["amp_rho"]), _p(av), float(ic["sigma"]), _p(cen))

    def set_state(self, reg, st):
        st = np.ascontiguousarray(st, dtype=np.float64)
        self.L.oracle_set_state(self.h, reg, _p(st), st.size // self.nf_size)

    def get_state(self, reg, nfields=4):
        out = np.zeros(nfields * self.nf_size)
        self.L.oracle_get_state(self.h, reg, _p(out), nfields)
        return out

    def face_size(self, face):
        a = [2, 2, 1, 1, 0, 0][face]
        nE = self.p.num_elements
        b, c = (1, 2) if a == 0 else ((0, 2) if a == 1 else (0, 1))
        return nE[b] * nE[c] * self.p.n ** 2 * 3

    def set_halo(self, reg, face, trace):
        self.L.oracle_set_halo(self.h, reg, face, _p(trace))

    def get_halo(self, reg, face):
        out = np.zeros(self.face_size(face))
        self.L.oracle_get_halo(self.h, reg, face, _p(out))
        return out

    def zero(self, reg):
        self.L.oracle_zero(self.h, reg)

    def copy(self, dst, src):
        self.L.oracle_copy(self.h, dst, src)

    def axpby(self, out, a, inp, b):
        self.L.oracle_axpby(self.h, out, float(a), inp, float(b))

    def operate(self, op, axis, mode, inp, out):
        self.L.oracle_operate(self.h, op, axis, mode, inp, out)

    def communicate(self, reg):
        self.L.oracle_communicate(self.h, reg)

    def set_source(self, face, amp, freq, phase, t0, tau, lead=None):
        self._lead = None if lead is None else np.ascontiguousarray(lead, dtype=np.float64)
        self.L.oracle_set_source(self.h, face, amp, freq, phase, t0, tau, _p(self._lead))

    def set_time(self, t):
        self.L.oracle_set_time(self.h, float(t))

    def rk33(self, dt, new, old, k1, k2, adv=1, diff=0, sponge=1):
        self.L.oracle_rk33(self.h, float(dt), new, old, k1, k2, adv, diff, sponge)

    def forward_euler(self, dt, new, old):
        self.L.oracle_forward_euler(self.h, float(dt), new, old)

    def lsrk54(self, dt, q, du, rhs, adv=1, diff=0, sponge=1):
        self.L.oracle_lsrk54(self.h, float(dt), q, du, rhs, adv, diff, sponge)

    def maxabs(self, reg):
        out = np.zeros(4)
        self.L.oracle_maxabs(self.h, reg, _p(out))
        return out

    def energy(self, reg):
        out = np.zeros(2)
        self.L.oracle_energy(self.h, reg, _p(out))
        return out


def rk33_blocks(oracles, dt, steps):
    """`steps` RK33 steps on each oracle block, one std::thread per block (registers 0..3)."""
    arr = (ctypes.c_void_p * len(oracles))(*[o.h for o in oracles])
    lib().oracle_rk33_blocks(arr, len(oracles), float(dt), int(steps))


def sponge_sigma(nE, nB, bpos, ew, sw, smax, ramp):
    L = lib()
    out = np.zeros(int(np.prod(nE)))
    L.oracle_sponge_sigma((ctypes.c_int * 3)(*nE), (ctypes.c_int * 3)(*nB), (ctypes.c_int * 3)(*bpos),
                          _p(np.asarray(ew, float)), (ctypes.c_int * 6)(*sw), float(smax), int(ramp), out.ctypes.data_as(_dp))
    return out


def line_solve(M, S, q, w, f, flux):
    n = len(f)
    x = np.zeros(n)
    lib().oracle_line_solve(n, _p(M), _p(S), float(q), float(w), _p(f), _p(flux), x.ctypes.data_as(_dp))
    return x


def ref_available():
    return os.path.exists(REF_LIB)


_ref = None


def ref_line_solve(M, S, q, w, f, flux):
    """The reference's own Matrix.h solve (oracle/_ref), returns (x, rank)."""
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_LIB)
        _ref.ref_line_solve.restype = ctypes.c_int
        _ref.ref_line_solve.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_float, ctypes.c_float, _dp, _dp, _dp]
    n = len(f)
    x = np.zeros(n)
    rank = _ref.ref_line_solve(n, _p(M), _p(S), float(q), float(w), _p(f), _p(flux), x.ctypes.data_as(_dp))
    return x, rank


def load_problem_into(o, p, reg=0):
    """Initial grid as the reference builds it: IC on every node incl. halos (Node.cpp:76-101)."""
    if p.ic is not None and any(p.ic["gauss"]):
        o.fill_ic(reg)
    else:
        o.zero(reg)
