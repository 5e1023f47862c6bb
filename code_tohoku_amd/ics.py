"""Seeded initial states for the per-timestep MHD update (host side, numpy).

All generators return the global state array described in ``restart.py``
(16 fields on the (-1..N+2)^3 Fortran index box, x fastest) with the
interior cells 1..N set, the N+1 planes set (face B on the far faces, periodic
images elsewhere) and the remaining ghosts left to the solver's own fills.

Face B is built from a staggered vector potential, so the discrete
divergence the constrained-transport update preserves,
  (Bx(i+1)-Bx(i))/dx(i) + (By(j+1)-By(j))/dy(j) + (Bz(k+1)-Bz(k))/dz(k),
vanishes to round-off.  Units are the code units of the reference
(1 pc, 1 Myr, m_p cm^-3; kb = 8.63359, see MCF3DSGincfca3.f90:11).
"""
from __future__ import annotations

import numpy as np

from .restart import Input3D, empty_state

KB = 8.63359
# composition of the reference's inflow boundary (MCF3DSGincfca3.f90:182-190)
_CM = None


def bbrv_cm(p: Input3D) -> np.ndarray:
    xc = 1.4e-4
    cm = np.zeros(8)
    cm[0] = 0.91 / 1.27
    cm[1] = 0.91 * p.pini / p.Hini / 1.27
    cm[2] = 0.91 * p.H2ini / p.Hini / 1.27
    cm[3] = 0.09 / 1.27
    cm[4] = 0.09 * p.Hepini / p.Heini / 1.27
    cm[7] = xc / 1.27
    cm[5] = xc * p.Cini / p.Cpini / 1.27
    cm[6] = xc * p.COini / p.Cpini / 1.27
    return cm


def grid(p: Input3D):
    """Cell edges x(i) and widths dx(i) for i = -1..N+2 (INITIA M3:255-298)."""
    out = []
    np1 = (p.np1x, p.np1y, p.np1z)
    ql1 = (p.ql1x, p.ql1y, p.ql1z)
    for d, n in enumerate(p.n):
        idx = np.arange(-1, n + 3)
        if d == 0:
            dx = np.where(idx <= p.np1x, p.ql1x / p.np1x, p.ql2x / p.np2x)
        else:
            dx = np.full(n + 4, ql1[d] / np1[d])
        x = np.empty(n + 4)
        x[0] = -dx[1]
        for i in range(1, n + 4):
            x[i] = x[i - 1] + dx[i]
        out.append((x, dx))
    return out


def _modes(rng, nmodes, kmax, periodic_x):
    ks = []
    while len(ks) < nmodes:
        k = rng.integers(-kmax, kmax + 1, size=3)
        if not periodic_x:
            k = k.astype(float)
            k[0] = rng.uniform(-kmax, kmax)
        if np.any(k != 0):
            ks.append(k)
    return np.array(ks, dtype=float), rng.uniform(0, 2 * np.pi, nmodes), rng.normal(size=nmodes)


def _field(mods, X, Y, Z, L):
    ks, ph, amp = mods
    f = np.zeros(np.broadcast_shapes(X.shape, Y.shape, Z.shape))
    for k, p0, a in zip(ks, ph, amp):
        f += a * np.cos(2 * np.pi * (k[0] * X / L[0] + k[1] * Y / L[1] + k[2] * Z / L[2]) + p0)
    return f / np.sqrt(0.5 * np.sum(amp ** 2))


def _wrap_periodic(st, n, periodic):
    """Set the N+1 plane of periodic axes to the image of plane 1."""
    nx, ny, nz = n
    if periodic[0]:
        st[:, :, :, nx + 2] = st[:, :, :, 2]
    if periodic[1]:
        st[:, :, ny + 2, :] = st[:, :, 2, :]
    if periodic[2]:
        st[:, nz + 2, :, :] = st[:, 2, :, :]


def random_state(p: Input3D, seed: int = 1, vamp: float = 8.0, bamp: float = 2.0,
                 b0=(4.0, 4.0, 1.0), rho0: float = 1.3, pt_range=(2.5, 12.0),
                 converging: bool | None = None) -> np.ndarray:
    """Smooth random state (density, velocity, pressure, composition and
    face B all perturbed) on the grid of ``p``.  Thermal pressure n*T (in
    10^3 K cm^-3) spans ``pt_range``, which puts cells on both sides of the
    limiter's T = 3 switch and in both the explicit and implicit cooling
    regimes.  With ``converging`` (default: when x has inflow boundaries) vx
    follows -v0*tanh((x-Lx/2)/w) like the colliding-flow problem."""
    rng = np.random.default_rng(seed)
    n = p.n
    nx, ny, nz = n
    periodic = (p.bcx1 == 3 and p.bcx2 == 3, True, True)
    if converging is None:
        converging = not periodic[0]
    (xe, dxe), (ye, dye), (ze, dze) = grid(p)
    L = (xe[nx + 1] - xe[1], ye[ny + 1] - ye[1], ze[nz + 1] - ze[1])
    # cell centres for slots 0..N+3 (index -1..N+2)
    xc = xe - 0.5 * dxe
    yc = ye - 0.5 * dye
    zc = ze - 0.5 * dze
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    st = empty_state(n)
    mk = lambda: _modes(rng, 6, 3, periodic[0])  # noqa: E731
    f_rho, f_pt, f_h2, f_vx, f_vy, f_vz = (_field(mk(), X, Y, Z, L) for _ in range(6))
    rho = rho0 * np.exp(0.35 * f_rho)
    cm = bbrv_cm(p)
    h2frac = 0.15 * (1.0 + np.tanh(f_h2))          # molecular fraction 0..0.3
    sp = np.empty((8,) + rho.shape)
    for s in range(8):
        sp[s] = cm[s] * rho
    sp[2] = h2frac * 0.91 / 1.27 * rho * 0.5
    sp[0] = cm[0] * rho * (1.0 - h2frac)
    lo, hi = np.log(pt_range[0]), np.log(pt_range[1])
    pt = np.exp(lo + (hi - lo) * 0.5 * (1.0 + np.tanh(f_pt)))
    ntot = sp[1] + sp[0] + sp[2] + sp[3] + sp[4]
    nnum = rho / 1.27
    pres = KB * ntot * (pt / nnum)                  # p = kb * sum(n) * T, T = (nT)/n
    vx = vamp * f_vx
    if converging:
        w = 0.1 * L[0]
        vx = -abs(p.vinitx1) * np.tanh((X - (xe[1] + 0.5 * L[0])) / w) + 0.3 * vamp * f_vx
    vy = vamp * f_vy
    vz = vamp * f_vz
    st[0], st[1], st[2], st[3], st[4] = rho, vx, vy, vz, pres
    st[8:16] = sp
    # face B from a staggered vector potential: A_d on the edges
    ex, ey, ez = xe, ye, ze              # edge coordinates x(i-1) sit at slot i
    xs = np.concatenate(([ex[0] - dxe[0]], ex[:-1]))  # x(i-1) for slot i
    ys = np.concatenate(([ey[0] - dye[0]], ey[:-1]))
    zs = np.concatenate(([ez[0] - dze[0]], ez[:-1]))
    Zc, Ys, Xs = np.meshgrid(zc, ys, xs, indexing="ij")   # Az: edge (x(i-1), y(j-1)), z centre
    Zs, Yc, Xs2 = np.meshgrid(zs, yc, xs, indexing="ij")  # Ay: edge (x(i-1), z(k-1))
    Zs2, Ys2, Xc = np.meshgrid(zs, ys, xc, indexing="ij")  # Ax: edge (y(j-1), z(k-1))
    scale = bamp * L[0] / (2 * np.pi * 2.0)
    Ax = scale * _field(mk(), Xc, Ys2, Zs2, L)
    Ay = scale * _field(mk(), Xs2, Yc, Zs, L)
    Az = scale * _field(mk(), Xs, Ys, Zc, L)
    dxa = dxe[None, None, :]
    dya = dye[None, :, None]
    dza = dze[:, None, None]
    Bx = np.zeros_like(rho); By = np.zeros_like(rho); Bz = np.zeros_like(rho)
    Bx[:, :-1, :] = (Az[:, 1:, :] - Az[:, :-1, :]) / dya[:, :-1, :]
    Bx[:-1, :, :] -= (Ay[1:, :, :] - Ay[:-1, :, :]) / dza[:-1, :, :]
    By[:-1, :, :] = (Ax[1:, :, :] - Ax[:-1, :, :]) / dza[:-1, :, :]
    By[:, :, :-1] -= (Az[:, :, 1:] - Az[:, :, :-1]) / dxa[:, :, :-1]
    Bz[:, :, :-1] = (Ay[:, :, 1:] - Ay[:, :, :-1]) / dxa[:, :, :-1]
    Bz[:, :-1, :] -= (Ax[:, 1:, :] - Ax[:, :-1, :]) / dya[:, :-1, :]
    st[5] = b0[0] + Bx
    st[6] = b0[1] + By
    st[7] = b0[2] + Bz
    _wrap_periodic(st, n, periodic)
    return st


def face_divergence(st: np.ndarray, p: Input3D) -> np.ndarray:
    """Discrete div B on interior cells, the quantity CC(3) preserves."""
    nx, ny, nz = p.n
    (_, dxe), (_, dye), (_, dze) = grid(p)
    I = slice(2, nx + 2); J = slice(2, ny + 2); K = slice(2, nz + 2)
    bx, by, bz = st[5], st[6], st[7]
    d = (bx[K, J, 3:nx + 3] - bx[K, J, I]) / dxe[None, None, I]
    d += (by[K, 3:ny + 3, I] - by[K, J, I]) / dye[None, J, None]
    d += (bz[3:nz + 3, J, I] - bz[K, J, I]) / dze[K, None, None]
    return d


def brio_wu_input(n: int = 64, ny: int | None = None, periodic_x: bool = True) -> Input3D:
    """Config 1 box: n cells on [0,1] per axis, CFL 0.4, no cooling.

    The reference's floors (pmin ~ 15.8, 1.27*rmin ~ 0.248; MCF3DSGincfca3.f90:
    173-176) and its species-derived gamma = 5/3 rule out the textbook
    Brio-Wu values, so the states are rescaled by the MHD similarity
    rho -> 2 rho, p -> 200 p, B -> sqrt(200) B (v = 0): rho 2 | 0.25,
    p 200 | 20, Bx = 10.607, By = +-14.142.  T = p/(kb n) >= 9 everywhere, so
    the limiter's minmod branch is taken uniformly, and the CFL limit binds
    dt rather than the cooling-time limit of Stblty."""
    ny = n if ny is None else ny
    h = 1.0 / n
    p = Input3D(np1x=n // 2, np2x=n - n // 2, np1y=ny // 2, np2y=ny - ny // 2, np1z=ny // 2, np2z=ny - ny // 2,
                ql1x=0.5, ql2x=0.5, ql1y=h * (ny // 2), ql2y=h * (ny - ny // 2),
                ql1z=h * (ny // 2), ql2z=h * (ny - ny // 2),
                vinitx1=0.0, vinitx2=0.0, binitx1=0.0, binitx2=0.0, binity1=0.0, binity2=0.0,
                cfl=0.4, facdep=1.0, tfinal=100.0, ifthrm=0)
    if periodic_x:
        p.bcx1 = p.bcx2 = 3
    return p


BRIO_WU = dict(rhoL=2.0, rhoR=0.25, pL=200.0, pR=20.0, bx=0.75 * np.sqrt(200.0), byL=np.sqrt(200.0),
               byR=-np.sqrt(200.0))


def brio_wu_state(p: Input3D) -> np.ndarray:
    """Planar shock tube along x: left state for x < 0.5, right state beyond
    (a double shock tube when x is periodic).  Only neutral H is present, so
    gamma = 5/3; the other species start at the reference's floors."""
    nx, ny, nz = p.n
    (xe, dxe), _, _ = grid(p)
    xc = xe - 0.5 * dxe
    left = xc < 0.5
    st = empty_state(p.n)
    bw = BRIO_WU
    rho = np.where(left, bw["rhoL"], bw["rhoR"])
    pres = np.where(left, bw["pL"], bw["pR"])
    by = np.where(left, bw["byL"], bw["byR"])
    floors = (0.0, 1e-20, 1e-20, 0.1949628 * 0.09, 1e-20, 1e-20, 1e-20, 1e-20)
    st[0] = rho[None, None, :]
    st[4] = pres[None, None, :]
    st[5] = bw["bx"]
    st[6] = by[None, None, :]
    st[8] = rho[None, None, :]
    for s in range(1, 8):
        st[8 + s] = floors[s]
    return st


# --------------------------------------------------------------------------
# Config 2: thermally bistable converging flow
# --------------------------------------------------------------------------
KMS = 1.0 / 0.97779        # 1 km/s in code velocity units
MICROGAUSS = 1.0 / 0.44828  # 1 muG in code field units


def park_miller(n: int, idum: int = 1) -> np.ndarray:
    """ran0 of GenDen/IGEN1.f90 (end of file): Park-Miller minimal standard,
    Schrage's factorisation, AM = 1./IM evaluated in single precision as the
    PARAMETER statement does."""
    IA, IM, IQ, IR = 16807, 2147483647, 127773, 2836
    am = float(np.float32(1.0) / np.float32(IM))
    out = np.empty(n)
    for t in range(n):
        k = idum // IQ
        idum = IA * (idum - k * IQ) - IR * k
        if idum < 0:
            idum += IM
        out[t] = am * idum
    return out


def kolmogorov_logfield(n: int, kmax: int | None = None, idum: int = 1) -> np.ndarray:
    """sum over 0 < |k| <= kmax of |k|^(-11/6) cos(2 pi k.x + phi_k) at the
    cell centres of an n^3 box (GenDen/IGEN1.f90:102-113), phases phi_k =
    2 pi ran0 drawn in the reference's loop order (l, ll, lll from -kmax to
    kmax).  Evaluated as one inverse FFT instead of a direct sum."""
    kmax = min(64, n // 4) if kmax is None else kmax
    ks = np.arange(-kmax, kmax + 1)
    L, LL, LLL = np.meshgrid(ks, ks, ks, indexing="ij")
    kw = np.sqrt((L ** 2 + LL ** 2 + LLL ** 2).astype(float))
    sel = (kw > 0) & (kw <= kmax)
    nmode = int(sel.sum())
    phase = 2.0 * np.pi * park_miller(nmode)
    amp = np.sqrt(kw[sel] ** (-11.0 / 3.0))
    l, ll, lll = L[sel], LL[sel], LLL[sel]
    # array index m = i-1 sits at the cell centre (m + 1/2)/n: fold the
    # half-cell shift into the coefficient
    coef = amp * np.exp(1j * (phase + np.pi * (l + ll + lll) / n))
    C = np.zeros((n, n, n), dtype=complex)
    np.add.at(C, (l % n, ll % n, lll % n), coef)
    u = np.real(np.fft.ifftn(C)) * n ** 3        # u[i_x, i_y, i_z], i = 0..n-1 <-> cells 1..n
    return np.ascontiguousarray(u.transpose(2, 1, 0))  # -> [k, j, i]


def converging_flow_input(n: int = 256, nsplit=(4, 4, 4)) -> Input3D:
    """Config 2 box: 20 pc cube, x inflow at +-10 km/s, periodic y/z, B = 5 muG
    at 45 deg in the x-y plane, CFL 0.49, KI cooling + conduction.  The WNM
    composition of MCF3DSGincfca3.f90:156-157 is scaled to n_tot = 1 cm^-3 and
    the inflow pressure is set to T = 6000 K."""
    h = n // 2
    v = 10.0 * KMS
    b = 5.0 * MICROGAUSS / np.sqrt(2.0)
    base = Input3D()
    ntot0 = base.Hini + base.pini + 2 * base.H2ini + base.Heini + base.Hepini
    s = 1.0 / ntot0
    p = Input3D(np1x=h, np2x=n - h, np1y=h, np2y=n - h, np1z=h, np2z=n - h,
                ql1x=10.0, ql2x=10.0, ql1y=10.0, ql2y=10.0, ql1z=10.0, ql2z=10.0,
                vinitx1=v, vinitx2=-v, binitx1=b, binitx2=b, binity1=b, binity2=b,
                cfl=0.49, facdep=1.0, tfinal=100.0, bcx1=4, bcx2=4, ifthrm=2,
                Hini=base.Hini * s, pini=base.pini * s, H2ini=base.H2ini * s, Heini=base.Heini * s,
                Hepini=base.Hepini * s, Cini=base.Cini * s, COini=base.COini * s, Cpini=base.Cpini * s)
    nsum = p.pini + p.Hini + p.H2ini + p.Heini + p.Hepini
    p.pinit_over_kb = nsum * 6000.0
    return p


def converging_flow_state(p: Input3D, drho: float = 0.1, idum: int = 1) -> np.ndarray:
    """Config 2 initial state, following the reference's DTF initialisation
    path (MCF3DSGincfca3.f90:434-447): log-normal Kolmogorov density
    rho = rho0 exp(a u) with a normalising u to rms drho (GenDen/IGEN1.f90:
    115-125), species proportional to rho, uniform pressure,
    vx = -v0 tanh((x(i) - ql1x)/2), uniform face field."""
    nx, ny, nz = p.n
    assert nx == ny == nz, "cubic grids only"
    u = kolmogorov_logfield(nx, idum=idum)
    a = drho / np.sqrt(np.mean(u ** 2))
    rho0 = p.Hini + p.pini + 2.0 * p.H2ini + 4.0 * p.Heini + 4.0 * p.Hepini
    st = empty_state(p.n)
    rho = np.full((nz + 4, ny + 4, nx + 4), rho0)
    rho[2:nz + 2, 2:ny + 2, 2:nx + 2] = rho0 * np.exp(a * u)
    comp = np.array([p.Hini, p.pini, p.H2ini, p.Heini, p.Hepini, p.Cini, p.COini, p.Cpini]) / rho0
    for s_ in range(8):
        st[8 + s_] = comp[s_] * rho
    st[0] = rho
    st[4] = p.pinit_over_kb * KB * 1.0e-3
    (xe, _), _, _ = grid(p)
    st[1] = (-p.vinitx1 * np.tanh(0.5 * (xe - p.ql1x)))[None, None, :]
    st[5] = p.binitx1
    st[6] = p.binity1
    st[7] = p.binitz1
    _wrap_periodic(st, p.n, (False, True, True))
    return st


def inflow_density_field(st: np.ndarray, p: Input3D) -> np.ndarray:
    """The DTF field of a state: its interior density as REAL*4, [k, j, i].
    INITIA's fresh-start path sets the density to this field
    (MCF3DSGincfca3.f90:437), and the x-inflow ghosts read it shifted by -v t
    (:818-826, 854-861), so the gas entering the box carries the initial
    density structure."""
    nx, ny, nz = p.n
    return np.ascontiguousarray(st[0, 2:nz + 2, 2:ny + 2, 2:nx + 2], dtype=np.float32)


def turbulence_input(n: int = 512) -> Input3D:
    """Decaying MHD turbulence box: 10 pc periodic cube, n_tot = 10 cm^-3
    (the WNM composition ratios of MCF3DSGincfca3.f90:156-157), T = 100 K,
    B = 3 muG along z, CFL 0.49, KI cooling + conduction.  The reference's
    pressure floor (pmin = 1.83e3 K cm^-3, MCF3DSGincfca3.f90:173) is above
    n T = 1e3 K cm^-3, so the first update raises the gas to ~183 K."""
    h = n // 2
    base = Input3D()
    ntot0 = base.Hini + base.pini + 2 * base.H2ini + base.Heini + base.Hepini
    s = 10.0 / ntot0
    bz = 3.0 * MICROGAUSS
    p = Input3D(np1x=h, np2x=n - h, np1y=h, np2y=n - h, np1z=h, np2z=n - h,
                ql1x=5.0, ql2x=5.0, ql1y=5.0, ql2y=5.0, ql1z=5.0, ql2z=5.0,
                vinitx1=0.0, vinitx2=0.0, binitx1=0.0, binitx2=0.0, binity1=0.0, binity2=0.0,
                binitz1=bz, binitz2=bz, cfl=0.49, facdep=1.0, tfinal=100.0, bcx1=3, bcx2=3, ifthrm=2,
                Hini=base.Hini * s, pini=base.pini * s, H2ini=base.H2ini * s, Heini=base.Heini * s,
                Hepini=base.Hepini * s, Cini=base.Cini * s, COini=base.COini * s, Cpini=base.Cpini * s)
    nsum = p.pini + p.Hini + p.H2ini + p.Heini + p.Hepini
    p.pinit_over_kb = nsum * 100.0
    return p


def solenoidal_velocity(n: int, kmin: float = 1.0, kmax: float = 4.0, seed: int = 1) -> np.ndarray:
    """Divergence-free random field on an n^3 periodic grid, (3, n, n, n) in
    (z, y, x) index order: Gaussian Fourier amplitudes with flat power on
    kmin <= |k| <= kmax (k in units of the box wavenumber), the component
    along k removed, zero mean, unit rms speed."""
    try:
        import scipy.fft as sfft
        irfftn = lambda a: sfft.irfftn(a, s=(n, n, n), workers=-1)
    except ImportError:  # pragma: no cover
        irfftn = lambda a: np.fft.irfftn(a, s=(n, n, n))
    rng = np.random.default_rng(seed)
    kz = np.fft.fftfreq(n, 1.0 / n)[:, None, None]
    ky = np.fft.fftfreq(n, 1.0 / n)[None, :, None]
    kx = np.fft.rfftfreq(n, 1.0 / n)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    shell = (k2 >= kmin * kmin) & (k2 <= kmax * kmax)
    idx = np.nonzero(shell)
    m = len(idx[0])
    a = rng.normal(size=(3, m)) + 1j * rng.normal(size=(3, m))
    kv = np.stack([np.broadcast_to(kx, shell.shape)[idx], np.broadcast_to(ky, shell.shape)[idx],
                   np.broadcast_to(kz, shell.shape)[idx]])
    a -= kv * (np.sum(kv * a, axis=0) / np.sum(kv * kv, axis=0))
    v = np.empty((3, n, n, n))
    for c in range(3):
        F = np.zeros(shell.shape, dtype=np.complex128)
        F[idx] = a[c]
        v[c] = irfftn(F)
    v -= v.mean(axis=(1, 2, 3), keepdims=True)
    v /= np.sqrt(np.mean(np.sum(v * v, axis=0)))
    return v


def turbulence_state(p: Input3D, mach: float = 5.0, seed: int = 1) -> np.ndarray:
    """Uniform density, pressure and B = Bz; solenoidal velocity with k in
    [1, 4] scaled to rms Mach ``mach`` with respect to the adiabatic sound
    speed."""
    nx, ny, nz = p.n
    assert nx == ny == nz, "cubic grids only"
    rho0 = p.Hini + p.pini + 2.0 * p.H2ini + 4.0 * p.Heini + 4.0 * p.Hepini
    s4 = p.Hini + p.pini + p.Heini + p.Hepini
    gamma = (5.0 * s4 + 7.0 * p.H2ini) / (3.0 * s4 + 5.0 * p.H2ini)
    p0 = p.pinit_over_kb * KB * 1.0e-3
    cs = np.sqrt(gamma * p0 / rho0)
    v = solenoidal_velocity(nx, seed=seed) * (mach * cs)
    st = empty_state(p.n)
    st[0] = rho0
    comp = np.array([p.Hini, p.pini, p.H2ini, p.Heini, p.Hepini, p.Cini, p.COini, p.Cpini])
    for s_ in range(8):
        st[8 + s_] = comp[s_]
    st[4] = p0
    I = slice(2, nx + 2)
    for c in range(3):
        st[1 + c][I, I, I] = v[c]
    st[5] = p.binitx1
    st[6] = p.binity1
    st[7] = p.binitz1
    _wrap_periodic(st, p.n, (True, True, True))
    return st
