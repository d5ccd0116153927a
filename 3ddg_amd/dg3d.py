"""ctypes binding of the C ABI in include/dg3d.h (libdg3d.so, built in-tree).

The library is required: there is no Python or CPU fallback for any device entry point.
Loading fails loudly when libdg3d.so is missing.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libdg3d.so")
ABI_VERSION = 1  # the structure layouts and signatures below (DG3D_ABI_VERSION of include/dg3d.h)

OK = 0
E_INVALID, E_HIP, E_NOMEM, E_NAN, E_UNSUPPORTED, E_IO = -1, -2, -3, -4, -5, -6
FP64, MIXED = 0, 1
CONTRACT_VALU, CONTRACT_MFMA_F32, CONTRACT_MFMA_BF16 = 0, 1, 2
RK33, LSRK54 = 0, 1
OP_UFROMV, OP_SPONGE, OP_QFROMDRHO, OP_DRHOFROMQ = 0, 1, 2, 3
REDUCE_MAXABS, REDUCE_ENERGY, REDUCE_RELL2, REDUCE_DIFFL2 = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Desc(ctypes.Structure):
    _fields_ = [
        ("order", ctypes.c_int),
        ("num_elements", ctypes.c_int * 3),
        ("periodic", ctypes.c_int * 3),
        ("precision", ctypes.c_int),
        ("contraction", ctypes.c_int),
        ("num_registers", ctypes.c_int),
        ("with_q", ctypes.c_int),
        ("xcd_remap", ctypes.c_int),
        ("mass", _dp),
        ("stiff", _dp),
        ("lgl_nodes", _dp),
        ("lgl_weights", _dp),
    ]


class Source(ctypes.Structure):
    _fields_ = [
        ("face", ctypes.c_int),
        ("amplitude", ctypes.c_double),
        ("frequency", ctypes.c_double),
        ("phase", ctypes.c_double),
        ("t0", ctypes.c_double),
        ("tau", ctypes.c_double),
        ("lead", _dp),
    ]


class SolveDesc(ctypes.Structure):
    _fields_ = [
        ("order", ctypes.c_int),
        ("num_elements", ctypes.c_int * 3),
        ("periodic", ctypes.c_int * 3),
        ("mass", _dp),
        ("stiff", _dp),
        ("widths", _dp * 3),
        ("c0", _dp),
        ("rho0", _dp),
        ("b_over_a", _dp),
        ("sigma", _dp),
        ("state", _dp),
        ("halo", _dp * 6),
        ("dt", ctypes.c_double),
        ("source_face", ctypes.c_int),
        ("source_amplitude", ctypes.c_double),
        ("source_frequency", ctypes.c_double),
        ("source_phase", ctypes.c_double),
        ("source_t0", ctypes.c_double),
        ("source_tau", ctypes.c_double),
        ("source_lead", _dp),
    ]


# every entry point declared in include/dg3d.h, with its ctypes signature
_SIGS = {
    "dg3d_abi_version": (ctypes.c_int, []),
    "dg3d_create_error": (ctypes.c_char_p, []),
    "dg3d_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "dg3d_create": (ctypes.c_int, [ctypes.POINTER(Desc), ctypes.POINTER(ctypes.c_void_p)]),
    "dg3d_destroy": (None, [ctypes.c_void_p]),
    "dg3d_set_geometry": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "dg3d_set_material": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    "dg3d_set_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int]),
    "dg3d_get_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int]),
    "dg3d_set_halo": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]),
    "dg3d_zero": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dg3d_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "dg3d_axpby": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double]),
    "dg3d_operate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int]),
    "dg3d_halo": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dg3d_set_source": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Source)]),
    "dg3d_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]),
    "dg3d_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "dg3d_cfl_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, _dp]),
    "dg3d_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _ip]),
    "dg3d_set_variant": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "dg3d_set_graphs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dg3d_group_order": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_int64)]),
    "dg3d_reduce": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    "dg3d_reduce_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    "dg3d_set_probes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                        ctypes.c_int]),
    "dg3d_get_probes": (ctypes.c_int, [ctypes.c_void_p, _dp, _ip]),
    "dg3d_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "dg3d_last_step_timing": (ctypes.c_int, [ctypes.c_void_p, _dp, _ip]),
    "dg3d_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(SolveDesc), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "dg3d_batch_create_error": (ctypes.c_char_p, []),
    "dg3d_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "dg3d_batch_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "dg3d_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, _ip]),
    "dg3d_batch_set_variant": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dg3d_batch_set_schedule": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dg3d_batch_get_schedule": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "dg3d_batch_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dg3d_batch_get_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "dg3d_batch_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), _ip, _ip, _ip, _ip, _ip]),
    "dg3d_batch_last_step_timing": (ctypes.c_int, [ctypes.c_void_p, _dp, _ip]),
    "dg3d_basis_generate": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]),
    "dg3d_basis_write_files": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "dg3d_basis_read_files": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, _dp, _dp, _dp, _dp]),
    "dg3d_sponge_sigma": (ctypes.c_int, [_ip, _ip, _ip, _dp, _ip, ctypes.c_double, ctypes.c_int, _dp]),
}

_lib = None


def lib():
    """Load libdg3d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python 3ddg_amd/build.py` "
                               "(there is no fallback implementation)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.dg3d_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI version {L.dg3d_abi_version()}, this binding "
                               f"describes version {ABI_VERSION}: rebuild the library")
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Dg3dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dg3d error {code}: {msg}")
        self.code = code


# ---- host-side setup helpers --------------------------------------------------

def basis_generate(order):
    n = order + 1
    x, w, M, S, D = (np.zeros(n), np.zeros(n), np.zeros(n * n), np.zeros(n * n), np.zeros(n * n))
    rc = lib().dg3d_basis_generate(order, _ptr(x), _ptr(w), _ptr(M), _ptr(S), _ptr(D))
    if rc:
        raise Dg3dError(rc, "basis_generate")
    return x, w, M.reshape(n, n), S.reshape(n, n), D.reshape(n, n)


def basis_write_files(directory, max_order):
    rc = lib().dg3d_basis_write_files(str(directory).encode(), max_order)
    if rc:
        raise Dg3dError(rc, "basis_write_files")


def basis_read_files(coeff_path, matrices_path, order):
    n = order + 1
    x, w, M, S = np.zeros(n), np.zeros(n), np.zeros(n * n), np.zeros(n * n)
    rc = lib().dg3d_basis_read_files(str(coeff_path).encode(), str(matrices_path).encode(), order,
                                     _ptr(x), _ptr(w), _ptr(M), _ptr(S))
    if rc:
        raise Dg3dError(rc, f"basis_read_files({coeff_path}, {matrices_path}, {order})")
    return x, w, M.reshape(n, n), S.reshape(n, n)


def group_order(order, num_elements, elems_per_group, groups):
    """Element group of every workgroup of a stage launch under a workgroup order (0..3)."""
    nE = (ctypes.c_int * 3)(*num_elements)
    out = np.zeros(int(groups), dtype=np.int64)
    rc = lib().dg3d_group_order(int(order), nE, int(elems_per_group), int(groups),
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    if rc:
        raise Dg3dError(rc, "group_order")
    return out


def sponge_sigma(num_elements, num_blocks, block_pos, elem_width, sponge_width, sigma_max, ramp):
    nE = (ctypes.c_int * 3)(*num_elements)
    nB = (ctypes.c_int * 3)(*num_blocks)
    bp = (ctypes.c_int * 3)(*block_pos)
    ew = _f64(elem_width)
    sw = (ctypes.c_int * 6)(*sponge_width)
    out = np.zeros(int(np.prod(num_elements)))
    rc = lib().dg3d_sponge_sigma(nE, nB, bp, _ptr(ew), sw, float(sigma_max), int(ramp), _ptr(out))
    if rc:
        raise Dg3dError(rc, "sponge_sigma")
    return out


# ---- device context -----------------------------------------------------------

class Context:
    """One block's device state (the reference's Grid/Block for one rank)."""

    def __init__(self, order, num_elements, mass, stiff, nodes, weights, periodic=(0, 0, 0), precision=FP64,
                 contraction=CONTRACT_VALU, num_registers=1, with_q=False, xcd_remap=2):
        self._L = lib()
        self.order = order
        self.n = order + 1
        self.num_elements = tuple(int(v) for v in num_elements)
        self.E = int(np.prod(self.num_elements))
        self.np_ = self.n ** 3
        self.nfields = 7 if with_q else 4
        self._keep = [_f64(mass).ravel(), _f64(stiff).ravel(), _f64(nodes), _f64(weights)]
        d = Desc()
        d.order = order
        d.num_elements[:] = list(self.num_elements)
        d.periodic[:] = [int(bool(p)) for p in periodic]
        d.precision = precision
        d.contraction = contraction
        d.num_registers = num_registers
        d.with_q = int(bool(with_q))
        d.xcd_remap = int(xcd_remap)
        d.mass, d.stiff, d.lgl_nodes, d.lgl_weights = (_ptr(a) for a in self._keep)
        h = ctypes.c_void_p()
        rc = self._L.dg3d_create(ctypes.byref(d), ctypes.byref(h))
        if rc:
            raise Dg3dError(rc, self._L.dg3d_create_error().decode())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.dg3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc:
            raise Dg3dError(rc, f"{what}: {self._L.dg3d_last_error(self._h).decode()}")

    @property
    def field_size(self):
        return self.E * self.np_

    def set_geometry(self, corner, wx, wy, wz):
        c = _f64(corner)
        ws = [_f64(w) for w in (wx, wy, wz)]
        self._check(self._L.dg3d_set_geometry(self._h, _ptr(c), *[_ptr(w) for w in ws]), "set_geometry")

    def set_material(self, c0=None, rho0=None, b_over_a=None, sigma=None, diffusion=None):
        arrs = [None if a is None else _f64(a) for a in (c0, rho0, b_over_a, sigma, diffusion)]
        for a in arrs:
            if a is not None and a.size != self.E:
                raise ValueError("material arrays need one value per element")
        self._check(self._L.dg3d_set_material(self._h, *[_ptr(a) for a in arrs]), "set_material")

    def set_state(self, reg, state):
        s = _f64(state)
        nf = s.size // self.field_size
        if nf * self.field_size != s.size:
            raise ValueError("state size is not a whole number of fields")
        self._check(self._L.dg3d_set_state(self._h, reg, _ptr(s), nf), "set_state")

    def get_state(self, reg, nfields=4):
        out = np.zeros(nfields * self.field_size)
        self._check(self._L.dg3d_get_state(self._h, reg, _ptr(out), nfields), "get_state")
        return out

    def set_halo(self, reg, face, trace):
        t = None if trace is None else _f64(trace)
        self._check(self._L.dg3d_set_halo(self._h, reg, face, _ptr(t)), "set_halo")

    def zero(self, reg):
        self._check(self._L.dg3d_zero(self._h, reg), "zero")

    def copy(self, dst, src):
        self._check(self._L.dg3d_copy(self._h, dst, src), "copy")

    def axpby(self, out, self_scalar, inp, other_scalar):
        self._check(self._L.dg3d_axpby(self._h, out, float(self_scalar), inp, float(other_scalar)), "axpby")

    def operate(self, op, axis, mode, in_reg, out_reg):
        self._check(self._L.dg3d_operate(self._h, op, axis, mode, in_reg, out_reg), "operate")

    def halo(self, reg):
        self._check(self._L.dg3d_halo(self._h, reg), "halo")

    def set_source(self, face=-1, amplitude=0.0, frequency=0.0, phase=0.0, t0=0.0, tau=1.0, lead=None):
        s = Source()
        s.face, s.amplitude, s.frequency, s.phase, s.t0, s.tau = face, amplitude, frequency, phase, t0, tau
        self._lead = None if lead is None else _f64(lead)
        s.lead = _ptr(self._lead)
        self._check(self._L.dg3d_set_source(self._h, ctypes.byref(s)), "set_source")

    def step(self, scheme, dt, nsteps, t):
        tt = ctypes.c_double(t)
        self._check(self._L.dg3d_step(self._h, scheme, float(dt), int(nsteps), ctypes.byref(tt)), "step")
        return tt.value

    def stage(self, scheme, stage, dt, t):
        self._check(self._L.dg3d_stage(self._h, scheme, stage, float(dt), float(t)), "stage")

    def autotune(self, scheme):
        v, x = ctypes.c_int(), ctypes.c_int()
        self._check(self._L.dg3d_autotune(self._h, scheme, ctypes.byref(v), ctypes.byref(x)), "autotune")
        return v.value, x.value

    def set_variant(self, scheme, variant, xcd_remap):
        self._check(self._L.dg3d_set_variant(self._h, scheme, int(variant), int(xcd_remap)), "set_variant")

    def set_graphs(self, enable):
        """dg3d_step replays captured HIP graphs (True, the default) or launches directly."""
        self._check(self._L.dg3d_set_graphs(self._h, 1 if enable else 0), "set_graphs")

    def cfl_dt(self, cfl):
        v = ctypes.c_double()
        self._check(self._L.dg3d_cfl_dt(self._h, float(cfl), ctypes.byref(v)), "cfl_dt")
        return v.value

    def reduce(self, kind, reg, reg2=0):
        out = np.zeros(8)
        self._check(self._L.dg3d_reduce(self._h, kind, reg, reg2, _ptr(out)), "reduce")
        return out

    def reduce_noalloc(self, kind, reg, reg2=0):
        """dg3d_reduce_noalloc: the same reduction with the library's per-device scratch array."""
        out = np.zeros(8)
        self._check(self._L.dg3d_reduce_noalloc(self._h, kind, reg, reg2, _ptr(out)), "reduce_noalloc")
        return out

    def set_probes(self, nodes, capacity):
        arr = np.ascontiguousarray(nodes, dtype=np.int64)
        self._probe_count = arr.size
        self._check(self._L.dg3d_set_probes(self._h, arr.size, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                            int(capacity)), "set_probes")

    def get_probes(self):
        steps = ctypes.c_int()
        self._check(self._L.dg3d_get_probes(self._h, None, ctypes.byref(steps)), "get_probes")
        out = np.zeros(max(steps.value, 1) * max(getattr(self, "_probe_count", 1), 1))
        self._check(self._L.dg3d_get_probes(self._h, _ptr(out), ctypes.byref(steps)), "get_probes")
        return out[: steps.value * self._probe_count].reshape(steps.value, self._probe_count)

    def sync(self):
        self._check(self._L.dg3d_sync(self._h), "sync")

    def last_step_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        self._check(self._L.dg3d_last_step_timing(self._h, ctypes.byref(ms), ctypes.byref(n)), "timing")
        return ms.value, n.value


# ---- batched solves -----------------------------------------------------------

class Batch:
    """Independent blocks of different sizes and orders advanced together by LSRK54 (the batched
    entry points of include/dg3d.h). `solves` is a sequence of dicts with the keys order,
    num_elements, mass, stiff, widths (3 arrays), c0, rho0, b_over_a, sigma, state (4 fields,
    dg3d_set_state layout), dt, and optionally periodic, halo (6 traces or None) and source (a
    dict with face, amplitude, frequency, phase, t0, tau and optionally lead, as set_source)."""

    def __init__(self, solves, precision=MIXED):
        self._L = lib()
        self._keep = []
        descs = (SolveDesc * len(solves))()
        self.sizes = []
        for d, s in zip(descs, solves):
            order = int(s["order"])
            d.order = order
            d.num_elements[:] = [int(x) for x in s["num_elements"]]
            d.periodic[:] = [int(x) for x in s.get("periodic", (0, 0, 0))]
            d.mass = self._arr(s["mass"])
            d.stiff = self._arr(s["stiff"])
            for a in range(3):
                d.widths[a] = self._arr(s["widths"][a])
            for name in ("c0", "rho0", "b_over_a", "sigma", "state"):
                setattr(d, name, self._arr(s.get(name)))
            halo = s.get("halo") or [None] * 6
            for f in range(6):
                d.halo[f] = self._arr(halo[f])
            d.dt = float(s["dt"])
            src = s.get("source")
            d.source_face = -1 if src is None else int(src["face"])
            if src is not None:
                d.source_amplitude = float(src["amplitude"])
                d.source_frequency = float(src["frequency"])
                d.source_phase = float(src["phase"])
                d.source_t0 = float(src["t0"])
                d.source_tau = float(src["tau"])
                d.source_lead = self._arr(src.get("lead"))
            E = int(np.prod(s["num_elements"]))
            self.sizes.append((E, (order + 1) ** 3))
        h = ctypes.c_void_p()
        rc = self._L.dg3d_batch_create(len(solves), descs, int(precision), ctypes.byref(h))
        if rc:
            raise Dg3dError(rc, self._L.dg3d_batch_create_error().decode())
        self._h = h
        self._keep = []

    def _arr(self, a):
        if a is None:
            return None
        a = _f64(a).ravel()
        self._keep.append(a)
        return _ptr(a)

    def _check(self, rc, what):
        if rc:
            raise Dg3dError(rc, f"{what}: {self._L.dg3d_batch_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._L.dg3d_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self):
        out = (ctypes.c_int * 8)()
        self._check(self._L.dg3d_batch_autotune(self._h, out), "batch_autotune")
        return list(out[: self.info()["num_orders"]])

    def set_variant(self, variant):
        self._check(self._L.dg3d_batch_set_variant(self._h, int(variant)), "batch_set_variant")

    def set_schedule(self, schedule):
        self._check(self._L.dg3d_batch_set_schedule(self._h, int(schedule)), "batch_set_schedule")

    def schedule(self):
        v = ctypes.c_int()
        self._check(self._L.dg3d_batch_get_schedule(self._h, ctypes.byref(v)), "batch_get_schedule")
        return v.value

    def step(self, nsteps):
        self._check(self._L.dg3d_batch_step(self._h, int(nsteps)), "batch_step")

    def get_state(self, solve):
        E, np_ = self.sizes[solve]
        out = np.zeros(4 * E * np_)
        self._check(self._L.dg3d_batch_get_state(self._h, int(solve), _ptr(out)), "batch_get_state")
        return out

    def info(self):
        nodes = ctypes.c_int64()
        k = ctypes.c_int()
        orders, variants, tasks, grids = [(ctypes.c_int * 8)() for _ in range(4)]
        self._check(self._L.dg3d_batch_info(self._h, ctypes.byref(nodes), ctypes.byref(k), orders, variants, tasks,
                                            grids), "batch_info")
        n = k.value
        return dict(total_nodes=nodes.value, num_orders=n, orders=list(orders[:n]), variants=list(variants[:n]),
                    tasks=list(tasks[:n]), grids=list(grids[:n]))

    def last_step_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        self._check(self._L.dg3d_batch_last_step_timing(self._h, ctypes.byref(ms), ctypes.byref(n)), "batch_timing")
        return ms.value, n.value
