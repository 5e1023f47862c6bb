"""On-disk formats of the MCF3DSG* family (host side, numpy).

* ``INPUT3D.DAT`` -- 16 list-directed lines read by INITIA
  (3DMHD-CfCA/MCF3DSGincfca3.f90:134-151).
* restart ``000{rank:03d}.dat`` -- Fortran unformatted sequential FP64, one
  record per (k, j) in [1, N+1]^2 holding (U1..U8, ndH, ndp, ndH2, ndHe,
  ndHep, ndC, ndCO, ndCp, Phi) interleaved for i = 1..N+1
  (written at MCF3DSGincfca3.f90:749-757, read at :454-465).
* snapshot ``{nunit:03d}{rank:03d}.dat`` -- the same layout in FP32 (:701-717).

Global state convention used across this package: a float64 array of shape
(16, nz+4, ny+4, nx+4), C-contiguous, i.e. the Fortran array
U(-1:nx+2, -1:ny+2, -1:nz+2, 16) with x fastest.  Fields 0..7 are U(:,:,:,1..8)
(rho, vx, vy, vz, p, face Bx, face By, face Bz) and 8..15 the species in the
restart order above.
"""
from __future__ import annotations

import dataclasses
import os
import struct
from typing import Iterable

import numpy as np

NFIELD = 16
FIELD_NAMES = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz",
               "ndH", "ndp", "ndH2", "ndHe", "ndHep", "ndC", "ndCO", "ndCp")


@dataclasses.dataclass
class Input3D:
    """Parameters of INPUT3D.DAT (MCF3DSGincfca3.f90:134-151) plus the WNM
    state that INITIA hard-wires (:155-157)."""
    np1x: int = 32
    np2x: int = 32
    np1y: int = 32
    np2y: int = 32
    np1z: int = 32
    np2z: int = 32
    ql1x: float = 10.0
    ql2x: float = 10.0
    ql1y: float = 10.0
    ql2y: float = 10.0
    ql1z: float = 10.0
    ql2z: float = 10.0
    vinitx1: float = 9.778
    vinitx2: float = -9.778
    vinity1: float = 0.0
    vinity2: float = 0.0
    vinitz1: float = 0.0
    vinitz2: float = 0.0
    binitx1: float = 0.0
    binitx2: float = 0.0
    binity1: float = 2.2307
    binity2: float = 2.2307
    binitz1: float = 0.0
    binitz2: float = 0.0
    cfl: float = 0.49
    facdep: float = 1.0
    maxstp: int = 4
    nitera: int = 5000
    tfinal: float = 100.0
    bcx1: int = 4
    bcx2: int = 4
    bcy1: int = 3
    bcy2: int = 3
    bcz1: int = 3
    bcz2: int = 3
    ifchem: int = 0
    ifthrm: int = 2
    ifrad: int = 0
    ifgrv: int = 0
    pinit_over_kb: float = 8.810807e3
    Hini: float = 0.9219098
    pini: float = 0.9503446e-2
    H2ini: float = 0.9465513e-8
    Heini: float = 0.9155226e-1
    Hepini: float = 0.5655353e-3
    Cini: float = 0.1565848e-8
    COini: float = 0.2202631e-20
    Cpini: float = 0.1433520e-3

    @property
    def n(self):
        return (self.np1x + self.np2x, self.np1y + self.np2y, self.np1z + self.np2z)

    def to_text(self) -> str:
        def f(v):
            return repr(float(v))
        lines = [
            f"{self.np1x},{self.np2x}", f"{self.np1y},{self.np2y}", f"{self.np1z},{self.np2z}",
            f"{f(self.ql1x)},{f(self.ql2x)}", f"{f(self.ql1y)},{f(self.ql2y)}", f"{f(self.ql1z)},{f(self.ql2z)}",
            f"{f(self.vinitx1)},{f(self.vinitx2)}", f"{f(self.vinity1)},{f(self.vinity2)}",
            f"{f(self.vinitz1)},{f(self.vinitz2)}",
            f"{f(self.binitx1)},{f(self.binitx2)}", f"{f(self.binity1)},{f(self.binity2)}",
            f"{f(self.binitz1)},{f(self.binitz2)}",
            f"{f(self.cfl)},{f(self.facdep)}",
            f"{self.maxstp},{self.nitera},{f(self.tfinal)}",
            f"{self.bcx1},{self.bcx2},{self.bcy1},{self.bcy2},{self.bcz1},{self.bcz2}",
            f"{self.ifchem},{self.ifthrm},{self.ifrad},{self.ifgrv}",
        ]
        return "\n".join(lines) + "\n"

    @classmethod
    def from_text(cls, text: str) -> "Input3D":
        rows = [ln.split("!")[0].strip() for ln in text.strip().splitlines()][:16]

        def nums(r):
            return [x.strip().lower().replace("d", "e") for x in r.split(",") if x.strip()]
        v = [nums(r) for r in rows]
        p = cls()
        p.np1x, p.np2x = map(int, v[0]); p.np1y, p.np2y = map(int, v[1]); p.np1z, p.np2z = map(int, v[2])
        p.ql1x, p.ql2x = map(float, v[3]); p.ql1y, p.ql2y = map(float, v[4]); p.ql1z, p.ql2z = map(float, v[5])
        p.vinitx1, p.vinitx2 = map(float, v[6]); p.vinity1, p.vinity2 = map(float, v[7])
        p.vinitz1, p.vinitz2 = map(float, v[8])
        p.binitx1, p.binitx2 = map(float, v[9]); p.binity1, p.binity2 = map(float, v[10])
        p.binitz1, p.binitz2 = map(float, v[11])
        p.cfl, p.facdep = map(float, v[12])
        p.maxstp, p.nitera = int(v[13][0]), int(v[13][1]); p.tfinal = float(v[13][2])
        p.bcx1, p.bcx2, p.bcy1, p.bcy2, p.bcz1, p.bcz2 = map(int, v[14])
        p.ifchem, p.ifthrm, p.ifrad, p.ifgrv = map(int, v[15])
        return p


def empty_state(n: Iterable[int]) -> np.ndarray:
    nx, ny, nz = n
    return np.zeros((NFIELD, nz + 4, ny + 4, nx + 4), dtype=np.float64)


def _write_record(f, payload: bytes):
    f.write(struct.pack("<i", len(payload)))
    f.write(payload)
    f.write(struct.pack("<i", len(payload)))


def write_restart_rank(path: str, block: np.ndarray, phi: np.ndarray | None = None):
    """block: (16, nk+1, nj+1, ni+1) values at local 1..N+1 (x fastest)."""
    nf, nk1, nj1, ni1 = block.shape
    assert nf == NFIELD
    if phi is None:
        phi = np.zeros((nk1, nj1, ni1))
    rec = np.empty((nk1, nj1, ni1, 17), dtype="<f8")
    rec[..., :16] = np.moveaxis(block, 0, -1)
    rec[..., 16] = phi
    with open(path, "wb") as f:
        for k in range(nk1):
            for j in range(nj1):
                _write_record(f, rec[k, j].tobytes())


def read_restart_rank(path: str, ni: int, nj: int, nk: int, dtype="<f8"):
    """Returns (block (16, nk+1, nj+1, ni+1), phi)."""
    item = np.dtype(dtype).itemsize
    reclen = 17 * (ni + 1) * item
    out = np.empty((nk + 1, nj + 1, ni + 1, 17), dtype=np.float64)
    with open(path, "rb") as f:
        for k in range(nk + 1):
            for j in range(nj + 1):
                (n0,) = struct.unpack("<i", f.read(4))
                if n0 != reclen:
                    raise ValueError(f"{path}: record length {n0}, expected {reclen}")
                out[k, j] = np.frombuffer(f.read(reclen), dtype=dtype).reshape(ni + 1, 17)
                f.read(4)
    return np.ascontiguousarray(np.moveaxis(out[..., :16], -1, 0)), out[..., 16]


def rank_coords(rank: int, nsplit):
    """IST/JST/KST of a rank (MCF3DSGincfca3.f90:75)."""
    sx, sy, _ = nsplit
    ist = rank % sx
    kst = rank // (sx * sy)
    jst = rank // sx - sy * kst
    return ist, jst, kst


NSPLIT_TABLE = {4: (2, 2, 1), 8: (2, 2, 2), 16: (4, 2, 2), 32: (4, 4, 2), 64: (4, 4, 4),
                128: (8, 4, 4), 256: (8, 8, 4), 512: (8, 8, 8), 1024: (8, 8, 16)}


def split_state(state: np.ndarray, nsplit):
    """Global state -> per-rank restart blocks (local 1..N+1 on each axis)."""
    _, nz4, ny4, nx4 = state.shape
    nx, ny, nz = nx4 - 4, ny4 - 4, nz4 - 4
    sx, sy, sz = nsplit
    cx, cy, cz = nx // sx, ny // sy, nz // sz
    blocks = []
    for r in range(sx * sy * sz):
        i0, j0, k0 = rank_coords(r, nsplit)
        gi, gj, gk = i0 * cx, j0 * cy, k0 * cz
        # global index g <-> array slot g+1
        blocks.append(state[:, gk + 2:gk + cz + 3, gj + 2:gj + cy + 3, gi + 2:gi + cx + 3].copy())
    return blocks


def assemble_state(blocks, n, nsplit) -> np.ndarray:
    """Per-rank local 1..N+1 blocks -> global state (interior 1..N plus the
    N+1 plane of the last rank along each axis; other ghosts zero)."""
    nx, ny, nz = n
    sx, sy, sz = nsplit
    cx, cy, cz = nx // sx, ny // sy, nz // sz
    st = empty_state(n)
    for r, b in enumerate(blocks):
        i0, j0, k0 = rank_coords(r, nsplit)
        ei = cx + (1 if i0 == sx - 1 else 0)
        ej = cy + (1 if j0 == sy - 1 else 0)
        ek = cz + (1 if k0 == sz - 1 else 0)
        gi, gj, gk = i0 * cx, j0 * cy, k0 * cz
        st[:, gk + 2:gk + 2 + ek, gj + 2:gj + 2 + ej, gi + 2:gi + 2 + ei] = b[:, :ek, :ej, :ei]
    return st


def write_counters(run_dir: str, time: float = 0.0, nunit: int = 2):
    """tsave.DAT / tsave2D.DAT / time.DAT as INITIA and EVOLVE read them
    (MCF3DSGincfca3.f90:309-312, 520-530).  nunit >= 2 makes INITIA read the
    restart files."""
    with open(os.path.join(run_dir, "tsave.DAT"), "w") as f:
        f.write(f"{fortran_d(time)}\n{nunit:8d}\n")
    with open(os.path.join(run_dir, "tsave2D.DAT"), "w") as f:
        f.write("  0\n")
    with open(os.path.join(run_dir, "time.DAT"), "w") as f:
        for _ in range(max(nunit, 1)):
            f.write(f"{fortran_d(time)}\n")


def fortran_d(v: float) -> str:
    """Format as Fortran (1p1d25.17)."""
    s = f"{v:.17E}"
    mant, ex = s.split("E")
    return f"{mant}D{int(ex):+03d}".rjust(25)
