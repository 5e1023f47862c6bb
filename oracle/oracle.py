"""ctypes binding of the CPU restatement (oracle/mcf_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and only as the checker / the timed CPU
baseline.  The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmcf_oracle.so")


class Params(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("np1x", "np2x", "np1y", "np2y", "np1z", "np2z")] + \
        [(n, C.c_double) for n in ("ql1x", "ql2x", "ql1y", "ql2y", "ql1z", "ql2z",
                                   "vinitx1", "vinitx2", "vinity1", "vinity2", "vinitz1", "vinitz2",
                                   "binitx1", "binitx2", "binity1", "binity2", "binitz1", "binitz2",
                                   "cfl", "facdep")] + \
        [("maxstp", C.c_int), ("nitera", C.c_int), ("tfinal", C.c_double)] + \
        [(n, C.c_int) for n in ("bcx1", "bcx2", "bcy1", "bcy2", "bcz1", "bcz2",
                                "ifchem", "ifthrm", "ifrad", "ifgrv")] + \
        [(n, C.c_double) for n in ("pinit_over_kb", "Hini", "pini", "H2ini", "Heini", "Hepini",
                                   "Cini", "COini", "Cpini")] + \
        [(n, C.c_int) for n in ("nsplitx", "nsplity", "nsplitz")] + \
        [("kicool", C.c_int)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        D = C.c_double
        pd = C.POINTER(C.c_double)
        L.mcfo_create.restype = P
        L.mcfo_create.argtypes = [C.POINTER(Params)]
        L.mcfo_destroy.argtypes = [P]
        L.mcfo_set_threads.argtypes = [P, C.c_int]
        L.mcfo_set_threads.restype = C.c_int
        L.mcfo_set_dtf.argtypes = [P, C.c_void_p]
        L.mcfo_set_dtf.restype = C.c_int
        L.mcfo_field.restype = pd
        L.mcfo_field.argtypes = [P, C.c_int]
        L.mcfo_grid.restype = pd
        L.mcfo_grid.argtypes = [P, C.c_int, C.c_int]
        L.mcfo_get_time.argtypes = [P, pd, C.POINTER(C.c_int)]
        L.mcfo_set_time.argtypes = [P, D, C.c_int]
        L.mcfo_bbrv.argtypes = [P, pd, pd, pd]
        L.mcfo_cc4.argtypes = [P]
        L.mcfo_couran.restype = D
        L.mcfo_couran.argtypes = [P]
        L.mcfo_stblty.restype = D
        L.mcfo_stblty.argtypes = [P]
        L.mcfo_timestep.restype = D
        L.mcfo_timestep.argtypes = [P, D, C.POINTER(C.c_int)]
        for nm in ("mcfo_source", "mcfo_ct", "mcfo_step"):
            getattr(L, nm).argtypes = [P, D]
        L.mcfo_sweep.argtypes = [P, C.c_int, D]
        L.mcfo_dissip.argtypes = [P]
        L.mcfo_run.restype = C.c_int
        L.mcfo_run.argtypes = [P, C.c_int, pd, C.POINTER(C.c_int)]
        L.mcfo_rieman.argtypes = [pd, D, pd, pd]
        L.mcfo_chareq.argtypes = [pd, pd, pd]
        L.mcfo_fcool_ki.restype = D
        L.mcfo_fcool_ki.argtypes = [D, D]
        L.mcfo_imc.restype = D
        L.mcfo_imc.argtypes = [D, D, D, D, D]
        L.mcfo_vlimit1d.argtypes = [C.c_int, pd, pd, pd, pd]
        L.mcfo_init_columns.argtypes = [P]
        L.mcfo_column.restype = pd
        L.mcfo_column.argtypes = [P, C.c_int, C.c_int]
        L.mcfo_grav_solve.restype = C.c_int
        L.mcfo_grav_solve.argtypes = [P]
        L.mcfo_grav_kick.argtypes = [P, D]
        L.mcfo_phi.restype = pd
        L.mcfo_phi.argtypes = [P]
        L.mcfo_fcool_cell.restype = D
        L.mcfo_fcool_cell.argtypes = [P, C.c_int, C.c_int, C.c_int, D]
        _lib = L
    return _lib


COOLING = {"ki": 1, "chem": 0}
COLUMNS = ("Ntot", "NH2", "NnC", "NCO", "tCII")


def to_params(p, nsplit=(2, 2, 2), cooling="ki") -> Params:
    out = Params()
    for name, _ in Params._fields_:
        if name.startswith("nsplit") or name == "kicool":
            continue
        setattr(out, name, getattr(p, name))
    out.nsplitx, out.nsplity, out.nsplitz = nsplit
    out.kicool = COOLING[cooling]
    return out


class Oracle:
    """One single-block CPU instance of the reference step.

    cooling: "ki" for the Koyama-Inutsuka Fcool the KI configurations
    substitute, "chem" for the reference's chemistry-network Fcool."""

    def __init__(self, p, state: np.ndarray | None = None, time: float = 0.0, ifevo: int = 1,
                 nsplit=(2, 2, 2), threads: int = 0, cooling: str = "ki"):
        self.p = p
        self.n = p.n
        self._params = to_params(p, nsplit, cooling)
        self.h = lib().mcfo_create(C.byref(self._params))
        if not self.h:
            raise ValueError("oracle: grid or block split not supported")
        # threads < 1: OpenMP's default (OMP_NUM_THREADS or all cores)
        self.threads = lib().mcfo_set_threads(self.h, threads)
        if self.threads < 1:
            raise MemoryError("oracle work arrays")
        nx, ny, nz = self.n
        self.shape = (nz + 4, ny + 4, nx + 4)
        if state is not None:
            self.set_state(state)
            lib().mcfo_cc4(self.h)
            lib().mcfo_init_columns(self.h)
            if p.ifgrv == 2:  # INITIA: GRAVTY(0,2) after the state is read (M3:481-484)
                lib().mcfo_grav_solve(self.h)
        lib().mcfo_set_time(self.h, time, ifevo)

    def __del__(self):
        if getattr(self, "h", None):
            lib().mcfo_destroy(self.h)
            self.h = None

    def set_dtf(self, dtf):
        """Inflow density field (float32 [k, j, i] over the interior), or
        None for the uniform inflow density."""
        if dtf is None:
            lib().mcfo_set_dtf(self.h, None)
            return
        self._dtf = np.ascontiguousarray(dtf, dtype=np.float32)
        assert self._dtf.shape == tuple(self.n[::-1])
        if lib().mcfo_set_dtf(self.h, self._dtf.ctypes.data) != 0:
            raise MemoryError("oracle DTF field")

    def field(self, fid: int) -> np.ndarray:
        ptr = lib().mcfo_field(self.h, fid)
        return np.ctypeslib.as_array(ptr, shape=self.shape)

    def set_state(self, st: np.ndarray):
        for f in range(16):
            self.field(f + 1)[...] = st[f]

    def state(self) -> np.ndarray:
        return np.stack([self.field(f + 1).copy() for f in range(16)])

    def columns(self) -> np.ndarray | None:
        """[q, dir, k, j, i]: the column densities of COLUMNS seen from the
        left (dir 0) and right (dir 1) x ends, or None when not carried."""
        if not lib().mcfo_column(self.h, 0, 0):
            return None
        return np.stack([np.stack([np.ctypeslib.as_array(lib().mcfo_column(self.h, q, d), shape=self.shape).copy()
                                   for d in range(2)]) for q in range(5)])

    def phi(self) -> np.ndarray | None:
        """Gravitational potential (ifgrv = 2), the layout of a field."""
        ptr = lib().mcfo_phi(self.h)
        if not ptr:
            return None
        return np.ctypeslib.as_array(ptr, shape=self.shape).copy()

    def grav_solve(self):
        if lib().mcfo_grav_solve(self.h) != 0:
            raise ValueError("oracle: no self-gravity (ifgrv != 2)")

    def grav_kick(self, dt):
        lib().mcfo_grav_kick(self.h, dt)

    def fcool_cell(self, i, j, k, T):
        return lib().mcfo_fcool_cell(self.h, i, j, k, T)

    def bcc(self) -> np.ndarray:
        return np.stack([self.field(17 + c).copy() for c in range(3)])

    @property
    def time(self):
        t = C.c_double()
        e = C.c_int()
        lib().mcfo_get_time(self.h, C.byref(t), C.byref(e))
        return t.value, e.value

    def run(self, nsteps: int):
        dts = np.zeros(nsteps)
        tags = np.zeros(nsteps, dtype=np.int32)
        n = lib().mcfo_run(self.h, nsteps, dts.ctypes.data_as(C.POINTER(C.c_double)),
                           tags.ctypes.data_as(C.POINTER(C.c_int)))
        return dts[:n], tags[:n]

    def step(self, dt: float):
        lib().mcfo_step(self.h, dt)

    def timestep(self, tsave: float = 1e30):
        st = C.c_int()
        dt = lib().mcfo_timestep(self.h, tsave, C.byref(st))
        return dt, st.value

    def couran(self):
        return lib().mcfo_couran(self.h)

    def stblty(self):
        return lib().mcfo_stblty(self.h)

    def source(self, dt):
        lib().mcfo_source(self.h, dt)

    def sweep(self, d, dt):
        lib().mcfo_sweep(self.h, d, dt)

    def ct(self, dt):
        lib().mcfo_ct(self.h, dt)

    def dissip(self):
        lib().mcfo_dissip(self.h)


def rieman(ql, qr, gamma):
    a = np.array(list(ql) + list(qr), dtype=np.float64)
    pa = C.c_double(); va = C.c_double()
    lib().mcfo_rieman(a.ctypes.data_as(C.POINTER(C.c_double)), gamma, C.byref(pa), C.byref(va))
    return pa.value, va.value


def chareq(ql, qr):
    a = np.array(list(ql) + list(qr), dtype=np.float64)
    b = C.c_double(); v = C.c_double()
    lib().mcfo_chareq(a.ctypes.data_as(C.POINTER(C.c_double)), C.byref(b), C.byref(v))
    return b.value, v.value


def fcool_ki(rho, T):
    return lib().mcfo_fcool_ki(rho, T)


def imc(P, n, dt, rho, gammi1):
    return lib().mcfo_imc(P, n, dt, rho, gammi1)


def vlimit1d(q, T, dxx):
    q = np.ascontiguousarray(q, dtype=np.float64)
    n = q.size - 4
    T = np.ascontiguousarray(T, dtype=np.float64)
    dxx = np.ascontiguousarray(dxx, dtype=np.float64)
    out = np.zeros(n + 2)
    pd = C.POINTER(C.c_double)
    lib().mcfo_vlimit1d(n, q.ctypes.data_as(pd), T.ctypes.data_as(pd), dxx.ctypes.data_as(pd),
                        out.ctypes.data_as(pd))
    return out
