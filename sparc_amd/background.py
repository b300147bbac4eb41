"""Host-side FP64 set-up of one solve: the work SPARC does once before its
time loop (Initialize_all, solar_data, mp_Initialize_RHS[_2d], INIT_DAMPING,
FORWARD_SOURCE).  The results are handed to the GPU engine through
``sparc_set_background`` / ``sparc_set_profile`` exactly as the reference's
module globals would be.

Single-precision literals of the reference (no -r8, makefile:27) are kept
through ``sp()`` so set-up values match the reference to the last bit where
the operation order allows.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np


def sp(x: float) -> float:
    """A Fortran default-real literal stored into a REAL*8."""
    return float(np.float32(x))


PI_R = sp(3.14159265358979)      # initialize.f90:38
RSUN_R = sp(69598946770.0)       # initialize.f90:38
DIMC = 10.0 ** 6                 # solar_data, initialize.f90:751
DIMRHO = sp(1e-3)                # 10.0**(-3.0), initialize.f90:755
DIML = RSUN_R                    # initialize.f90:756
DIMB = (4.0 * PI_R * DIMC ** 2.0 * DIMRHO) ** 0.5


# --------------------------------------------------------------------------
# DBYD2 (dbyd2.f90:8-237) + MTRIDAG ispecial=0 (mtridag.f90:94-111), vectorised
# over the first index.  Used for set-up gradients only.
# --------------------------------------------------------------------------
def dbyd2(B: np.ndarray, ibc: int, A: np.ndarray | None = None) -> np.ndarray:
    """Sixth-order compact derivative of B(n1, n2) along axis 1 (w.r.t.
    index/(n2-1)); A carries caller boundary values for ibc 2, 3, 4."""
    B = np.asarray(B, dtype=np.float64)
    n1, n2 = B.shape
    if not 0 <= ibc <= 4 or n2 < 5:
        raise ValueError("dbyd2: ibc must be 0..4 and n2 >= 5")
    A = np.zeros_like(B) if A is None else np.array(A, dtype=np.float64)
    d2i = float(n2 - 1)
    C1, C2 = d2i * 7.0 / 9.0, d2i / 36.0
    A[:, 2:n2 - 2] = C1 * (B[:, 3:n2 - 1] - B[:, 1:n2 - 3]) + C2 * (B[:, 4:n2] - B[:, 0:n2 - 4])
    c1, c2, c3 = sp(-43. / 96.) * d2i, sp(-5. / 6.) * d2i, 9. / 8. * d2i
    c4, c5 = sp(1. / 6.) * d2i, sp(-1. / 96.) * d2i
    A[:, 1] = c1 * B[:, 0] + c2 * B[:, 1] + c3 * B[:, 2] + c4 * B[:, 3] + c5 * B[:, 4]
    A[:, n2 - 2] = -(c1 * B[:, n2 - 1] + c2 * B[:, n2 - 2] + c3 * B[:, n2 - 3]
                     + c4 * B[:, n2 - 4] + c5 * B[:, n2 - 5])
    b1, b2, b3, b4 = sp(-10 / 3.) * d2i, -3.0 * d2i, 6. * d2i, sp(1. / 3.) * d2i
    if ibc == 0:
        A[:, 0] = 0.0
        A[:, n2 - 1] = 0.0
    if ibc in (1, 3):
        A[:, 0] = b1 * B[:, 0] + b2 * B[:, 1] + b3 * B[:, 2] + b4 * B[:, 3]
    if ibc in (1, 2):
        A[:, n2 - 1] = -(b1 * B[:, n2 - 1] + b2 * B[:, n2 - 2] + b3 * B[:, n2 - 3]
                         + b4 * B[:, n2 - 4])
    alpha = sp(1. / 3.)
    alpha2, gamma2, alpha1, beta1 = 3. / 4., 1. / 8., 6.0, 3.
    cst = 1. / (alpha2 - beta1 * gamma2)
    up1 = (alpha1 * alpha2 - beta1) * cst
    if ibc % 2:
        A[:, 0] = (alpha2 * A[:, 0] - beta1 * A[:, 1]) * cst
    if ibc in (1, 2):
        A[:, n2 - 1] = (alpha2 * A[:, n2 - 1] - beta1 * A[:, n2 - 2]) * cst
    uppr = np.zeros(n2)
    uppr[2:n2 - 2] = alpha
    uppr[0] = 0.0 if ibc % 2 == 0 else up1
    uppr[1] = alpha2
    uppr[n2 - 2] = gamma2
    lowr = uppr[::-1].copy()
    lowr[n2 - 1] = up1 if ibc in (1, 2) else 0.0
    # MTRIDAG ispecial=0
    betad = np.empty(n2)
    gam = np.zeros(n2)
    betad[0] = 1.0
    for j in range(1, n2):
        gam[j] = uppr[j - 1] / betad[j - 1]
        betad[j] = 1.0 - lowr[j] * gam[j]
    U = np.empty_like(A)
    U[:, 0] = A[:, 0] / betad[0]
    for j in range(1, n2):
        U[:, j] = (A[:, j] - lowr[j] * U[:, j - 1]) / betad[j]
    for j in range(n2 - 2, -1, -1):
        U[:, j] = U[:, j] - gam[j + 1] * U[:, j + 1]
    return U


def _f32(x: float) -> float:
    return float(np.float32(x))


def dbyd1(B: np.ndarray, ibc: int, A: np.ndarray | None = None) -> np.ndarray:
    """DBYD1 (dbyd1.f:8-242) + TRIDAG ispecial=0 (tridag.f): the sixth-order
    compact derivative of B(n1, n2) along axis 0 (w.r.t. index/(n1-1)); A
    carries caller boundary values for ibc 2, 3, 4.  dbyd1.f has no IMPLICIT
    NONE and declares only A, B and the matrix REAL*8, so d1i, the stencil
    coefficients, alpha and the preconditioning constants are single
    precision, each computed in single arithmetic (dbyd1.f:27-28)."""
    B = np.asarray(B, dtype=np.float64)
    n1, n2 = B.shape
    if not 0 <= ibc <= 4:
        raise ValueError("dbyd1: ibc must be 0..4 (5, periodic, is not used here)")
    A = np.zeros_like(B) if A is None else np.array(A, dtype=np.float64)
    d1i = _f32(n1 - 1)
    if n1 < 5:  # dbyd1.f:44-77
        if ibc == 0:
            A[:] = 0.0
        elif ibc == 1:
            A[:] = (B[n1 - 1] - B[0])[None, :]
        elif ibc == 2:
            A[:] = A[0][None, :]
        elif ibc == 3:
            A[:] = A[n1 - 1][None, :]
        else:  # in place, row by row: later rows see the rewritten first row
            for i in range(n1):
                A[i] = (float(i) * A[0] + float(n1 - 1 - i) * A[n1 - 1]) / d1i
        return A
    f = np.float32
    C1 = float(f(f(d1i) * f(7.0)) / f(9.0))
    C2 = float(f(d1i) / f(36.0))
    A[2:n1 - 2] = C1 * (B[3:n1 - 1] - B[1:n1 - 3]) + C2 * (B[4:n1] - B[0:n1 - 4])
    c1 = float(f(-43.) / f(96.) * f(d1i))
    c2 = float(f(-5.) / f(6.) * f(d1i))
    c3 = float(f(9.) / f(8.) * f(d1i))
    c4 = float(f(1.) / f(6.) * f(d1i))
    c5 = float(f(-1.) / f(96.) * f(d1i))
    A[1] = c1 * B[0] + c2 * B[1] + c3 * B[2] + c4 * B[3] + c5 * B[4]
    A[n1 - 2] = -(c1 * B[n1 - 1] + c2 * B[n1 - 2] + c3 * B[n1 - 3] + c4 * B[n1 - 4]
                  + c5 * B[n1 - 5])
    b1 = float(f(-10) / f(3.) * f(d1i))
    b2 = float(f(-3.0) * f(d1i))
    b3 = float(f(6.) * f(d1i))
    b4 = float(f(1.) / f(3.) * f(d1i))
    if ibc == 0:
        A[0] = 0.0
        A[n1 - 1] = 0.0
    if ibc in (1, 3):
        A[0] = b1 * B[0] + b2 * B[1] + b3 * B[2] + b4 * B[3]
    if ibc in (1, 2):
        A[n1 - 1] = -(b1 * B[n1 - 1] + b2 * B[n1 - 2] + b3 * B[n1 - 3] + b4 * B[n1 - 4])
    alpha = float(f(1.) / f(3.))
    alpha2, gamma2, alpha1, beta1 = 0.75, 0.125, 6.0, 3.0
    cst = float(f(1.) / f(f(alpha2) - f(f(beta1) * f(gamma2))))
    up1 = float(f(f(f(alpha1) * f(alpha2)) - f(beta1)) * f(cst))
    if ibc % 2:
        A[0] = (alpha2 * A[0] - beta1 * A[1]) * cst
    if ibc in (1, 2):
        A[n1 - 1] = (alpha2 * A[n1 - 1] - beta1 * A[n1 - 2]) * cst
    uppr = np.zeros(n1)
    uppr[2:n1 - 2] = alpha
    uppr[0] = 0.0 if ibc % 2 == 0 else up1
    uppr[1] = alpha2
    uppr[n1 - 2] = gamma2
    lowr = uppr[::-1].copy()
    lowr[n1 - 1] = up1 if ibc in (1, 2) else 0.0
    # TRIDAG ispecial=0 (tridag.f), diag = 1
    U = np.empty_like(A)
    bet = 1.0
    U[0] = A[0] / bet
    gam = np.zeros(n1)
    for j in range(1, n1):
        gam[j] = uppr[j - 1] / bet
        bet = 1.0 - lowr[j] * gam[j]
        U[j] = (A[j] - lowr[j] * U[j - 1]) / bet
    for j in range(n1 - 2, -1, -1):
        U[j] = U[j] - gam[j + 1] * U[j + 1]
    return U


def ddz(f: np.ndarray, stretch: np.ndarray, ibc: int = 1) -> np.ndarray:
    """ddz (derivatives.f90:10-24) of a (nx, ny, nz) field."""
    nx, ny, nz = f.shape
    d = dbyd2(f.reshape(nx * ny, nz, order="F"), ibc)
    return (d * stretch[None, :]).reshape((nx, ny, nz), order="F")


# --------------------------------------------------------------------------
# Background model
# --------------------------------------------------------------------------
M_CM = 1.0e8  # 1 Mm in cm


def polytrope_table(nz: int, z_top_mm: float, z_bot_mm: float, d_top_mm: float = 0.5,
                    m: float = 1.5, gamma1: float = 5.0 / 3.0, g: float = 2.775841e4,
                    rho_ref: float = 2.78e-7, d_ref_cm: float = 3.92e7) -> np.ndarray:
    """Adiabatic polytrope with temperature linear in depth, in the column
    format of the reference's ``polytrope`` file (solar_data,
    initialize.f90:735-769): z/Rsun, c, rho, p, g (cgs), Gamma_1; row 1 at the
    bottom.

    Depth d = z_s - z below the T=0 level z_s = z_top + d_top.  Hydrostatic
    balance gives rho = rho_ref (d/d_ref)^m, p = rho g d/(m+1),
    c^2 = Gamma_1 p / rho.  The grid is uniform in sqrt(d), i.e. uniform in
    acoustic travel time, as in the shipped ``polytrope`` (its c decreases by
    a constant step per row).  rho_ref and d_ref are the shipped file's top
    row (rho = 2.78e-7 g/cc where c^2 (m+1)/(Gamma_1 g) = 3.92e7 cm).
    """
    zs = (z_top_mm + d_top_mm) * M_CM
    d_top = d_top_mm * M_CM
    d_bot = zs - z_bot_mm * M_CM
    s = np.linspace(math.sqrt(d_bot), math.sqrt(d_top), nz)
    d = s * s
    z_cm = zs - d
    rho = rho_ref * (d / d_ref_cm) ** m
    p = rho * g * d / (m + 1.0)
    c = np.sqrt(gamma1 * p / rho)
    tab = np.empty((nz, 6))
    tab[:, 0] = 1.0 + z_cm / RSUN_R
    tab[:, 1] = c
    tab[:, 2] = rho
    tab[:, 3] = p
    tab[:, 4] = g
    tab[:, 5] = gamma1
    return tab


def solar_data(tab: np.ndarray) -> dict:
    """Nondimensionalisation of solar_data (initialize.f90:728-774)."""
    z = tab[:, 0].copy()
    c = tab[:, 1] / DIMC
    return {
        "z": z,
        "c2": (tab[:, 1] / DIMC) ** 2.0,
        "c": c,
        "rho0": tab[:, 2] / DIMRHO,
        "p0": tab[:, 3] / (DIMRHO * DIMC ** 2.0),
        "g": tab[:, 4] * DIML / DIMC ** 2.0,
        "gamma": tab[:, 5].copy(),
    }


def nearest_index(z: np.ndarray, target: float) -> int:
    """1-based index of the last grid point nearest ``target``
    (physics.f90:44-63: a scan with ``.LE.`` keeps the last tie)."""
    temp = 1.0
    best = 1
    for k in range(z.size):
        if abs(z[k] - target) <= temp:
            temp = abs(z[k] - target)
            best = k + 1
    return best


def init_rhs(z, c2, rho0, p0, g, npmlbot: int, npmltop: int, is2d: bool,
             tempin: np.ndarray | None = None) -> dict:
    """Derived background of mp_Initialize_RHS (physics.f90:204-303) or
    mp_Initialize_RHS_2d (physics2d.f90:235-342), PML profiles included.
    c2, rho0, p0 are (nx, ny, nz); z, g are (nz,); tempin is the PML speed
    (sqrt(c2) by default, the fast speed with a magnetic field)."""
    nx, ny, nz = c2.shape
    unstretch = dbyd2(np.asarray(z)[None, :], 1)[0]
    stretch = 1.0 / unstretch
    rhoinv = 1.0 / rho0
    c2rho0 = c2 * rho0
    gradrho0_z = ddz(rho0, stretch, 1)
    gradp0_z = ddz(p0, stretch, 1)
    g = np.array(g, dtype=np.float64)
    if tempin is None:
        tempin = c2 ** 0.5
    dampz = np.zeros_like(c2)
    pd = PI_R * sp(0.005) * DIML / DIMC
    alphaz = np.full(nz, pd)
    botfac = 3. * 3. if is2d else 2. * 3. * 3.
    ktop = nz - 6 if is2d else nz - 1
    for k1 in range(1, nz + 1):
        k = k1 - 1
        if k1 < npmlbot + 1:
            L = z[npmlbot - 1] - z[0]
            alphaz[k] = PI_R * sp(0.005) * DIML / DIMC * (z[k] - z[0]) / L
            dampz[:, :, k] = botfac * tempin[:, :, 0] / L * ((z[npmlbot - 1] - z[k]) / L) ** 2.0
            fd = alphaz[k] / (dampz[:, :, k] + alphaz[k])
            gradp0_z[:, :, k] = gradp0_z[:, :, k] * fd
            gradrho0_z[:, :, k] = gradrho0_z[:, :, k] * fd
            g[k] = g[k] * fd[0, 0]
        if k1 > nz - npmltop:
            L = z[nz - 1] - z[nz - npmltop]
            alphaz[k] = PI_R * sp(0.005) * DIML / DIMC * (z[nz - 1] - z[k]) / L
            dampz[:, :, k] = 3. * 3. * tempin[:, :, ktop] / L * ((z[nz - npmltop] - z[k]) / L) ** 2.0
            fd = alphaz[k] / (dampz[:, :, k] + alphaz[k])
            gradp0_z[:, :, k] = gradp0_z[:, :, k] * fd
            gradrho0_z[:, :, k] = gradrho0_z[:, :, k] * fd
            g[k] = g[k] * fd[0, 0]
    bzpml = -alphaz[None, None, :] - dampz
    az = -dampz
    return dict(unstretch=unstretch, stretch=stretch, rhoinv=rhoinv, c2rho0=c2rho0,
                gradrho0_z=gradrho0_z, gradp0_z=gradp0_z, g=g, az=az, bzpml=bzpml)


def wilson_planes(rho0_3d: np.ndarray, rho_q: np.ndarray, k_ref: int) -> np.ndarray:
    """erad_2d / orad_2d of mp_Initialize_RHS (physics.f90:66-90): per pixel the
    1-based plane in [k_ref-50, k_ref+10] whose rho0 is closest to the quiet
    profile's rho_q(k_ref), later planes winning ties (the <= of the search),
    starting from a tolerance of 1.  Pixels with no plane within it get 0."""
    nx, ny, nz = rho0_3d.shape
    target = float(rho_q[k_ref - 1])
    best = np.ones((nx, ny))
    out = np.zeros((nx, ny), dtype=np.int32)
    for k in range(max(1, k_ref - 50), min(nz, k_ref + 10) + 1):
        d = np.abs(target - rho0_3d[:, :, k - 1])
        take = d <= best
        out[take] = k
        best[take] = d[take]
    return out


def sponge_z(z: np.ndarray, npml_zeroed: int = 90) -> np.ndarray:
    """spongez(k) of mp_Initialize_RHS (physics.f90:103-107) with the planes
    1..90 zeroed as the PML branch does (physics.f90:293); option 10 applies
    it to v for k >= nz-35.  Literals are single-rounded as in the reference."""
    z = np.asarray(z, dtype=np.float64)
    with np.errstate(over="ignore"):
        s = (200.0 / (1.0 + np.exp((sp(1.002) - z) / sp(0.0001)))
             + 300.0 / (1.0 + np.exp((z - (z[0] + sp(0.005))) / sp(0.0005))))
    s[:min(npml_zeroed, z.size)] = 0.0
    return s


def source_dep(z: np.ndarray, e_rad: int) -> np.ndarray:
    """source_dep(k) = exp(-10**7 * 4 (z(k) - z(e_rad))^2) (physics.f90:109),
    the vertical shape of option 2's source; e_rad is 1-based."""
    z = np.asarray(z, dtype=np.float64)
    return np.exp(-sp(1e7) * 4.0 * (z - z[e_rad - 1]) ** 2.0)


def damping_rates(nx: int, ny: int, xlength: float, ylength: float, is2d: bool) -> np.ndarray:
    """INIT_DAMPING (damping.f90:10-65) / INIT_DAMPING_2D (71-120):
    rates((nx/2+1), ny) with normx*normy folded."""
    constx = 2. * PI_R / xlength * RSUN_R
    consty = 2. * PI_R / ylength * RSUN_R
    i = np.arange(nx // 2 + 1, dtype=np.float64)
    kay = np.zeros((nx // 2 + 1, ny))
    if is2d:
        kay[:, 0] = (constx ** 2. * i ** 2. + consty ** 2. * 0.0 ** 2.) ** 0.5
    else:
        for j in range(1, ny + 1):
            jj = (j - 1.0) if j <= ny // 2 + 1 else (ny - j + 1.0)
            kay[:, j - 1] = (constx ** 2. * i ** 2. + consty ** 2. * jj ** 2.) ** 0.5
    kayref = 902.0
    normx, normy = 1.0 / nx, 1.0 / ny
    return DIML / DIMC * sp(1e-6) * 150.0 * (kay / kayref) ** sp(2.2) * normx * normy


def decay_coeff(n: int) -> np.ndarray:
    """decay_coeff_x (initialize.f90:215-221): evaluated in single precision
    (``i - n/3.0`` and ``exp`` of a default real), times normx."""
    i = np.arange(1, n // 2 + 2, dtype=np.float32)
    t = np.float32(1.0) / (np.float32(1.0) + np.exp((i - np.float32(n) / np.float32(3.0))
                                                     * np.float32(1.0)))
    return t.astype(np.float64) * (1.0 / n)


def fwdsource_reference(nt: int, timestep: float, nupeak: float = sp(0.0017),
                        nuwidth: float = sp(0.0055)) -> np.ndarray:
    """FORWARD_SOURCE (driver.f90:1046-1071): 1e-5 cos(2 pi nu t)
    exp(-2 (t - 0.5/nu)^2 w^2), zeroed below 1e-5 before scaling."""
    t = (np.arange(1, nt + 1) - 1.0) * timestep
    f = np.cos(2. * PI_R * nupeak * t) * np.exp(-2. * (t - 0.5 / nupeak) ** 2. * nuwidth ** 2.)
    f[np.abs(f) < sp(0.00001)] = 0.0
    return f * 10. ** (-5.)


def ricker_source(nt: int, timestep: float, nu: float = 3.0e-3, t0: float | None = None,
                  amp: float = 1e-5) -> np.ndarray:
    """Ricker wavelet peaked at ``nu`` (Hz), sampled at t = (k-1)*timestep,
    with the reference's convention of zeroing |f| < 1e-5 before scaling by
    1e-5 (driver.f90:1066-1068)."""
    if t0 is None:
        t0 = 1.2 / nu
    t = (np.arange(1, nt + 1) - 1.0) * timestep
    a = (np.pi * nu * (t - t0)) ** 2
    f = (1.0 - 2.0 * a) * np.exp(-a)
    f[np.abs(f) < 1e-5] = 0.0
    return f * amp


def gaussian_mask(nx: int, ny: int, xlength: float, ylength: float, x0_mm: float, y0_mm: float,
                  sigma_mm: float) -> np.ndarray:
    """Horizontal Gaussian exp(-r^2 / (2 sigma^2)) on the periodic grid
    x = (i-1)/nx (initialize.f90:197-203), centred at (x0, y0) Mm measured from
    the box centre like ``(x-0.5)*xlength`` (driver.f90:612)."""
    x = ((np.arange(nx) / nx) - 0.5) * xlength / M_CM
    y = ((np.arange(ny) / ny) - 0.5) * ylength / M_CM if ny > 1 else np.zeros(1)
    dx = (x[:, None] - x0_mm)
    dy = (y[None, :] - y0_mm)
    return np.exp(-(dx ** 2 + dy ** 2) / (2.0 * sigma_mm ** 2))


def sinc_mask_2d(nx: int, xlength: float, xloc_mm: float) -> np.ndarray:
    """The 2-D forward source mask of DETERMINE_STATUS (driver.f90:608-624):
    exp(-d^2/(400 dx^2)) sinc(0.43 pi d)."""
    x = np.arange(nx) / nx
    dx = (x[1] - x[0]) * xlength * 10. ** (-8)
    xdist = (x - 0.5) * xlength * 10. ** (-8) - xloc_mm
    with np.errstate(divide="ignore", invalid="ignore"):
        sincx = np.sin(0.43 * PI_R * xdist) / (0.43 * PI_R * xdist)
    sincx[np.argmin(np.abs(xdist))] = 1.0
    return (np.exp(-xdist ** 2. / (400. * dx ** 2.)) * sincx)[:, None]


@dataclass
class Setup:
    """Everything one solve needs, in reference units (nondimensional)."""
    nx: int
    ny: int
    nz: int
    option: int
    xlength: float
    ylength: float
    timestep: float
    npmlbot: int
    npmltop: int
    z: np.ndarray
    c2: np.ndarray
    rho0: np.ndarray
    p0: np.ndarray
    g_raw: np.ndarray
    derived: dict = field(default_factory=dict)
    rates: np.ndarray | None = None
    decay_x: np.ndarray | None = None
    decay_y: np.ndarray | None = None
    mhd: dict | None = None           # magnetic backgrounds by abi id (3-D)
    rho_q: np.ndarray | None = None   # quiet density profile (Wilson depression search)

    @property
    def deltat(self) -> float:
        return self.timestep * DIMC / DIML  # initialize.f90:472

    @property
    def stretchx(self) -> float:
        return RSUN_R / self.xlength  # initialize.f90:223

    @property
    def stretchy(self) -> float:
        return RSUN_R / self.ylength


def build_setup(tab: np.ndarray, nx: int, ny: int, option: int, xlength: float, ylength: float,
                timestep: float, npmlbot: int, npmltop: int, c2_3d: np.ndarray | None = None,
                rho0_3d: np.ndarray | None = None) -> Setup:
    """solar_data + mp_Initialize_RHS[_2d] for a z-stratified model, with
    optional 3-D c^2 / rho0 replacing the profiles (e.g. model_c, the
    sound-speed perturbation of driver.f90:151-155)."""
    nz = tab.shape[0]
    sd = solar_data(tab)
    shape = (nx, ny, nz)
    c2 = np.broadcast_to(sd["c2"], shape).copy(order="F") if c2_3d is None else np.asfortranarray(c2_3d)
    rho0 = np.broadcast_to(sd["rho0"], shape).copy(order="F") if rho0_3d is None else np.asfortranarray(rho0_3d)
    p0 = np.broadcast_to(sd["p0"], shape).copy(order="F")
    is2d = option == 17
    der = init_rhs(sd["z"], c2, rho0, p0, sd["g"], npmlbot, npmltop, is2d)
    S = Setup(nx=nx, ny=ny, nz=nz, option=option, xlength=xlength, ylength=ylength,
              timestep=timestep, npmlbot=npmlbot, npmltop=npmltop, z=sd["z"], c2=c2,
              rho0=rho0, p0=p0, g_raw=sd["g"], derived=der)
    S.rates = damping_rates(nx, ny, xlength, ylength, is2d)
    S.decay_x = decay_coeff(nx)
    S.decay_y = decay_coeff(ny)
    return S


def perturb_density(S: Setup, dlnrho: np.ndarray, hydrostatic: bool = True) -> Setup:
    """The set-up of a density perturbation rho0 -> rho0 exp(dlnrho) at fixed
    sound speed and gravity, with the background kept hydrostatic:
    d(p0')/dz changes by -g d(rho0) (p0' = -rho0 g).  This is the
    perturbation the density kernel of kernels.f90:292-313 is the
    derivative for: c^2 rho0 (the bulk modulus), rho0 (inertia), rho0' (the
    buoyancy of rho') and p0' (the Lagrangian pressure) all move with it.
    dlnrho must vanish in the PML layers (their profile factors multiply the
    gradients there).  hydrostatic=False keeps p0' as it was (a background
    out of balance; the density kernel is not its derivative)."""
    if S.mhd is not None:
        raise ValueError("perturb_density: magnetic set-ups are not covered")
    d = np.asarray(dlnrho, dtype=np.float64)
    rho_new = np.asfortranarray(np.broadcast_to(S.rho0, S.c2.shape) * np.exp(d))
    drho = rho_new - S.rho0
    der = init_rhs(S.z, S.c2, rho_new, S.p0, S.g_raw, S.npmlbot, S.npmltop, S.option == 17)
    der["gradp0_z"] = S.derived["gradp0_z"] - (S.derived["g"][None, None, :] * drho
                                              if hydrostatic else 0.0)
    return replace(S, rho0=rho_new, derived=der)


def is_uniform(f: np.ndarray) -> bool:
    return bool(np.all(f == f[:1, :1, :]))


def background_arrays(S: Setup) -> dict:
    """Background id -> array (compact z-profile when horizontally uniform)."""
    from . import abi
    d = S.derived
    out = {}
    for bid, arr in ((abi.BG_RHO0, S.rho0), (abi.BG_RHOINV, d["rhoinv"]),
                     (abi.BG_C2RHO0, d["c2rho0"]), (abi.BG_GRADP0_Z, d["gradp0_z"]),
                     (abi.BG_GRADRHO0_Z, d["gradrho0_z"]), (abi.BG_AZ, d["az"]),
                     (abi.BG_BZPML, d["bzpml"]), (abi.BG_C2, S.c2)):
        out[bid] = np.ascontiguousarray(arr[0, 0, :]) if is_uniform(arr) else arr
    return out


# --------------------------------------------------------------------------
# Magnetic backgrounds: a flux tube from a vector potential, and the
# reference's magnetic set-up of mp_Initialize_RHS
# --------------------------------------------------------------------------
def ddh(f: np.ndarray, axis: int, stretch: float) -> np.ndarray:
    """ddx / ddy (derivatives.f90:28-170) on the host in FP64: periodic
    spectral derivative along axis 0 or 1 with the Nyquist bin dropped,
    w.r.t. x/Rsun (stretch = Rsun/length)."""
    n = f.shape[axis]
    k = np.fft.rfftfreq(n) * n
    mult = 2.0 * np.pi * k * stretch
    if n % 2 == 0:
        mult[-1] = 0.0
    shape = [1] * f.ndim
    shape[axis] = k.size
    return np.fft.irfft(np.fft.rfft(f, axis=axis) * (1j * mult).reshape(shape), n=n, axis=axis)


def _solve_h(div: np.ndarray, sx: float, sy: float) -> np.ndarray:
    """chi with lap_h(chi) = div per plane (spectral, periodic, chi mean 0),
    using the squared wavenumbers of ddh so that ddh(ddh(chi)) summed over x
    and y reproduces div exactly for every mode ddh keeps."""
    nx, ny = div.shape[:2]
    kx = np.fft.fftfreq(nx) * nx
    ky = np.fft.fftfreq(ny) * ny
    mx = 2.0 * np.pi * kx * sx
    my = 2.0 * np.pi * ky * sy
    if nx % 2 == 0:
        mx[nx // 2] = 0.0
    if ny % 2 == 0:
        my[ny // 2] = 0.0
    lap = -(mx[:, None] ** 2 + my[None, :] ** 2)
    F = np.fft.fft2(div, axes=(0, 1))
    with np.errstate(divide="ignore", invalid="ignore"):
        G = np.where(lap[:, :, None] != 0.0, F / lap[:, :, None], 0.0)
    return np.fft.ifft2(G, axes=(0, 1)).real


def _tube_profile(zc, zmm, p0, b_ph, r_ph_mm, decay_mm, beta_min):
    """Axis field B0(z) and radius R(z) (Mm) of the flux tube: B0 is the
    smooth minimum (4-norm) of b_ph exp(z/decay) and the field of plasma
    beta beta_min, normalised to b_ph at z = 0; R = r_ph sqrt(b_ph/B0)."""
    cap = np.sqrt(8.0 * np.pi * p0 / beta_min)

    def smin(a, b):
        return (a ** -4.0 + b ** -4.0) ** -0.25

    cap0 = float(np.interp(0.0, zc, cap))
    # amplitude of the exponential branch that puts B0(0) at b_ph
    a0 = (b_ph ** -4.0 - cap0 ** -4.0) ** -0.25
    b0 = smin(a0 * np.exp(zmm / decay_mm), cap)
    return b0, r_ph_mm * np.sqrt(b_ph / b0)


def flux_tube(tab: np.ndarray, nx: int, ny: int, xlength: float, ylength: float,
              b_ph: float = 3000.0, r_ph_mm: float = 4.0, decay_mm: float = 10.0,
              beta_min: float = 1.5):
    """Axisymmetric Gaussian flux tube in the polytrope of ``tab``, on the
    centred (nx, ny) grid, in cgs (B in gauss).

    Field: B_z = B0(z) exp(-r^2/R(z)^2) with R = r_ph sqrt(b_ph/B0(z)), so
    the flux pi R^2 B0 is the same at every height.  On the axis B0 is the
    smooth minimum (4-norm) of b_ph exp(z/decay) -- decaying with depth like
    code_python/tubex.py's exp(z/decay_z) -- and sqrt(8 pi p0/beta_min), the
    field of plasma beta beta_min, normalised to b_ph at the photosphere
    z = 0.  The horizontal field is the horizontal gradient that cancels
    d(B_z)/dz in the divergence with the package's own operators (spectral
    x, y; compact z), so div B = 0 on the grid up to each plane's mean (the
    flux the box truncates, ~1e-6 of the tube's dB/dz); it is the discrete
    form of the vector potential's B_r = -r (dB0/dz)/(2 B0) B_z.

    Gas: lateral (thin-tube) pressure balance p = p0 - B_z^2/(8 pi), which
    beta_min > 1 keeps positive, at the temperature outside (rho = rho0 p/p0,
    so c is unchanged and the tube is less dense: a Wilson depression).  The
    full magnetohydrostatic balance of tubex.py (lines 73-96, pressure from
    the horizontal Lorentz force and density from the vertical one) has no
    positive solution for a 3 kG tube in this polytrope's top 0.5 Mm: the
    field has to weaken upward there, and the horizontal field that
    weakening needs pulls the tube's pressure below zero.

    Returns dict(rho, p, c, bx, by, bz, z) of (nx, ny, nz) arrays (z in cm).
    """
    nz = tab.shape[0]
    zc = (tab[:, 0] - 1.0) * RSUN_R               # cm, photosphere at 0
    zmm = zc / M_CM
    x = (np.arange(nx) / nx - 0.5) * xlength / M_CM
    y = (np.arange(ny) / ny - 0.5) * ylength / M_CM
    r2 = x[:, None, None] ** 2 + y[None, :, None] ** 2
    p0, rho0 = tab[:, 3], tab[:, 2]
    b0, R = _tube_profile(zc, zmm, p0, b_ph, r_ph_mm, decay_mm, beta_min)
    bz = b0[None, None, :] * np.exp(-r2 / R[None, None, :] ** 2)
    sx, sy = RSUN_R / xlength, RSUN_R / ylength
    unstretch = dbyd2(np.asarray(tab[:, 0])[None, :], 1)[0]
    stretch = 1.0 / unstretch
    # lap_h chi = -dBz/dz in x/Rsun units; B_h = grad_h chi (per cm)
    chi = _solve_h(-ddz(bz, stretch, 1) * RSUN_R, sx, sy)
    bx = ddh(chi, 0, sx) / RSUN_R
    by = ddh(chi, 1, sy) / RSUN_R
    p = p0[None, None, :] - bz ** 2 / (8.0 * np.pi)
    if p.min() <= 0:
        raise ValueError("flux tube: negative pressure (beta_min must exceed 1)")
    rho = rho0[None, None, :] * p / p0[None, None, :]
    c = np.broadcast_to(tab[:, 1][None, None, :], p.shape)
    return dict(rho=np.asfortranarray(rho), p=np.asfortranarray(p), c=np.asfortranarray(c),
                bx=np.asfortranarray(bx), by=np.asfortranarray(by), bz=np.asfortranarray(bz),
                z=zc)


def flux_tube_vector_potential(tab: np.ndarray, nx: int, ny: int, xlength: float,
                               ylength: float, b_ph: float = 3000.0, r_ph_mm: float = 4.0,
                               decay_mm: float = 10.0, beta_min: float = 1.5):
    """The flux tube of ``flux_tube`` built from a vector potential, B =
    B_mean z + curl A, with the package's own derivative operators (spectral
    x and y, compact z), so that div B = 0 on the grid by construction: the x,
    y and z operators act on different indices and commute.

    The tube's flux function is psi(r, z) = r A_phi = (Phi/2 pi)(1 - exp(-r^2
    / R(z)^2)), Phi = pi B0 R^2 the same at every height (B0(z), R(z) as in
    flux_tube), whose B_z = B0 exp(-r^2/R^2).  A periodic box holds the tube's
    net flux as the uniform field B_mean = Phi/(Lx Ly) (divergence-free and
    constant in z); the rest is the periodic Coulomb-gauge potential
    A = (-d phi/dy, d phi/dx, 0) with lap_h phi = B_z - B_mean per plane,
    which is A_phi(r, z) phi_hat less the uniform field's (B_mean/2)(-y, x)
    near the tube.  Then B_x = -dA_y/dz, B_y = dA_x/dz and B_z = dA_y/dx -
    dA_x/dy + B_mean.  The gas is set as in flux_tube.

    Returns flux_tube's dict plus ax, ay (gauss cm) and b_mean (gauss)."""
    nz = tab.shape[0]
    zc = (tab[:, 0] - 1.0) * RSUN_R
    zmm = zc / M_CM
    x = (np.arange(nx) / nx - 0.5) * xlength / M_CM
    y = (np.arange(ny) / ny - 0.5) * ylength / M_CM
    r2 = x[:, None, None] ** 2 + y[None, :, None] ** 2
    p0, rho0 = tab[:, 3], tab[:, 2]
    b0, R = _tube_profile(zc, zmm, p0, b_ph, r_ph_mm, decay_mm, beta_min)
    bz_t = b0[None, None, :] * np.exp(-r2 / R[None, None, :] ** 2)
    phi_flux = np.pi * b0 * (R * M_CM) ** 2          # Phi(z), gauss cm^2
    b_mean = float(np.mean(phi_flux)) / (xlength * ylength)
    sx, sy = RSUN_R / xlength, RSUN_R / ylength
    # Rsun units: phi' in G Rsun^2, A' in G Rsun, derivatives w.r.t. x/Rsun
    phi = _solve_h(bz_t - b_mean, sx, sy)
    ax = -ddh(phi, 1, sy)
    ay = ddh(phi, 0, sx)
    unstretch = dbyd2(np.asarray(tab[:, 0])[None, :], 1)[0]
    stretch = 1.0 / unstretch
    bx = -ddz(ay, stretch, 1)
    by = ddz(ax, stretch, 1)
    bz = ddh(ay, 0, sx) - ddh(ax, 1, sy) + b_mean
    p = p0[None, None, :] - bz ** 2 / (8.0 * np.pi)
    if p.min() <= 0:
        raise ValueError("flux tube: negative pressure (beta_min must exceed 1)")
    rho = rho0[None, None, :] * p / p0[None, None, :]
    c = np.broadcast_to(tab[:, 1][None, None, :], p.shape)
    F = np.asfortranarray
    return dict(rho=F(rho), p=F(p), c=F(c), bx=F(bx), by=F(by), bz=F(bz), z=zc,
                ax=F(ax * RSUN_R), ay=F(ay * RSUN_R), b_mean=b_mean, flux=phi_flux)


def magnetic_from_vectorpsi_2d(psi: np.ndarray, stretch: np.ndarray, stretchx: float):
    """The 2-D field from the vector potential model_vectorpsi
    (physics2d.f90:124-140): box = -ddz(psi, ibc 1), boz = DBYD1(psi,
    ibc 0) * stretchx (the compact x derivative, not the FFT), and the curl
    curlboy = ddz(box, ibc 1) - ddx(boz) (physics2d.f90:160-162).  psi is
    (nx, 1, nz); returns box, boz, curlboy."""
    psi = np.asarray(psi, dtype=np.float64)
    nx, ny, nz = psi.shape
    box = -ddz(psi, stretch, 1)
    boz = np.empty_like(psi)
    for j in range(ny):
        boz[:, j, :] = dbyd1(psi[:, j, :], 0) * stretchx
    curlboy = ddz(box, stretch, 1) - ddh(boz, 0, stretchx)
    return box, boz, curlboy


def magnetic_setup(tab: np.ndarray, fields: dict, nx: int, ny: int, option: int,
                   xlength: float, ylength: float, timestep: float, npmlbot: int, npmltop: int,
                   cap2: float = 8.0) -> "Setup":
    """mp_Initialize_RHS with MAGNETIC (physics.f90:113-160, 204-316) from 3-D
    rho, c, p (cgs) and B (gauss) as READ_IN_FIELD reads them
    (all_modules.f90:486-513):

    * nondimensionalisation (B/dimb, p/(dimrho dimc^2), rho/dimrho, c/dimc);
    * the Alfven-speed reduction cap2/(cap2 + B^2/(rho0 cq^2)) with the quiet
      sound speed cq(z), times ((k-nz)/10)^4 on the top 11 planes, applied to
      B as its square root (cap2 = 8 in the reference, i.e. v_A <= sqrt(8) c);
    * curl B0 as the reference forms it: curlboz = dx boy - dy box,
      curlbox = dy boz and curlboy = -dx boz (its dz terms are computed and
      then overwritten with zero, physics.f90:133-157);
    * horizontal gradients of rho0 and p0 (physics.f90:310-316);
    * the PML profiles from the fast speed sqrt(c^2 + B^2/rho0)
      (physics.f90:258), the z gradients of rho0 and p0, rhoinv, c2rho0.
    Returns a Setup whose ``mhd`` holds the extra backgrounds by abi id."""
    from . import abi
    sd = solar_data(tab)
    nz = tab.shape[0]
    rho0 = fields["rho"] / DIMRHO
    p0 = fields["p"] / (DIMRHO * DIMC ** 2.0)
    c2 = (fields["c"] / DIMC) ** 2.0
    box, boy, boz = (fields[k] / DIMB for k in ("bx", "by", "bz"))
    cq = sd["c"]
    red = (box ** 2 + boy ** 2 + boz ** 2) / (rho0 * cq[None, None, :] ** 2)
    red = cap2 / (cap2 + red)
    for k1 in range(nz - 10, nz + 1):
        red[:, :, k1 - 1] *= ((k1 - nz) / 10.0) ** 4.0
    sq = red ** 0.5
    box, boy, boz = box * sq, boy * sq, boz * sq
    sx, sy = RSUN_R / xlength, RSUN_R / ylength
    curlbz = ddh(boy, 0, sx) - ddh(box, 1, sy)
    curlbx = ddh(boz, 1, sy)
    curlby = -ddh(boz, 0, sx)
    tempin = (c2 + (box ** 2 + boy ** 2 + boz ** 2) / rho0) ** 0.5
    der = init_rhs(sd["z"], c2, rho0, p0, sd["g"], npmlbot, npmltop, False, tempin=tempin)
    S = Setup(nx=nx, ny=ny, nz=nz, option=option, xlength=xlength, ylength=ylength,
              timestep=timestep, npmlbot=npmlbot, npmltop=npmltop, z=sd["z"],
              c2=np.asfortranarray(c2), rho0=np.asfortranarray(rho0),
              p0=np.asfortranarray(p0), g_raw=sd["g"], derived=der)
    S.rates = damping_rates(nx, ny, xlength, ylength, False)
    S.decay_x = decay_coeff(nx)
    S.decay_y = decay_coeff(ny)
    S.mhd = {
        abi.BG_BX: box, abi.BG_BY: boy, abi.BG_BZ: boz,
        abi.BG_CURLBX: curlbx, abi.BG_CURLBY: curlby, abi.BG_CURLBZ: curlbz,
        abi.BG_GRADP0_X: ddh(p0, 0, sx), abi.BG_GRADP0_Y: ddh(p0, 1, sy),
        abi.BG_GRADRHO0_X: ddh(rho0, 0, sx), abi.BG_GRADRHO0_Y: ddh(rho0, 1, sy),
    }
    S.rho_q = sd["rho0"]
    return S
