"""Synthetic block set-ups (host-side initialisation, numpy).

These play the role of the reference's init_domain (domain.f90:30-315):
grid coordinates, material, anelastic coefficients, PML flags, boundary
codes, face frames, sources, receivers and the time step.  Metric terms are
NOT computed here: pass ``metrics_fn`` (the device pipeline
:func:`waveqlab3d_q_gpu_amd.block.grid_metrics`, or any function with the
signature ``X -> (metricx, metricy, metricz, J)``).

Array layout: ``(ncomp, nz, ny, nx)`` C order == Fortran ``A(x, y, z, c)``.
Axis convention (grid.f90:1350-1357): r (y, 2nd index) is depth, the r-min
face (face 2, "Ly") is the free surface, the r-max face (face 3) the bottom.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .block import SourceSpec

FACE_NAMES = ("Lx", "Rx", "Ly", "Ry", "Lz", "Rz")


@dataclass
class Problem:
    n: tuple                       # (nq, nr, ns)
    mesh: int
    X: np.ndarray                  # (3, nz, ny, nx)
    M: np.ndarray                  # (3, nz, ny, nx) lambda, mu, rho
    bc: tuple
    dt: float
    nsteps: int
    pml: tuple = (0,) * 6
    npml: tuple = (0,) * 6
    nmech: int = 0
    tau: np.ndarray = None
    weight: np.ndarray = None
    Qs_inv: np.ndarray = None
    Qp_inv: np.ndarray = None
    F0: np.ndarray = None          # initial fields (None = zero)
    sources: list = field(default_factory=list)
    receivers: np.ndarray = None   # (nrec, 3) 1-based node indices
    name: str = ""
    sat_form: int = 0              # WQL_SAT_BC / WQL_SAT_ALPHA_THETA


# ---------------------------------------------------------------------------
# grids
# ---------------------------------------------------------------------------
def cartesian_grid(n, a, b):
    """cartesian_grid_3d (grid.f90:214-241): x_i = a + (i-1)*(b-a)/(n-1)."""
    nq, nr, ns = n
    bh = [(b[k] - a[k]) / float(n[k] - 1) for k in range(3)]
    x = a[0] + np.arange(nq) * bh[0]
    y = a[1] + np.arange(nr) * bh[1]
    z = a[2] + np.arange(ns) * bh[2]
    X = np.empty((3, ns, nr, nq))
    X[0] = x[None, None, :]
    X[1] = y[None, :, None]
    X[2] = z[:, None, None]
    return X


def gen_surf_mesh(XL, XR, XB, XT, YL, YR, YB, YT):
    """Gen_Surf_Mesh (grid.f90:2420-2457): the bilinear Coons patch of two
    coordinate functions from their four boundary curves; XB, XT have mx
    points, XL, XR my points; returns X, Y as (mx, my) arrays (Fortran
    X(i, j))."""
    mx, my = len(XB), len(XL)
    q = (np.arange(mx) * (1.0 / (mx - 1)))[:, None]
    r = (np.arange(my) * (1.0 / (my - 1)))[None, :]

    def patch(L, R, B, T):
        L, R, B, T = (np.asarray(v, dtype=np.float64) for v in (L, R, B, T))
        return ((1.0 - q) * L[None, :] + q * R[None, :] + (1.0 - r) * B[:, None] + r * T[:, None]
                - (1.0 - q) * (1.0 - r) * L[0] - q * (1.0 - r) * R[0] - r * (1.0 - q) * T[0]
                - (r * q) * T[mx - 1])

    return patch(XL, XR, XB, XT), patch(YL, YR, YB, YT)


def gen3_curve_mesh(left, right, top, bottom, front, back):
    """Gen3_Curve_Mesh (grid.f90:2292-2418): transfinite interpolation of one
    coordinate from its six boundary surfaces, left/right (nr, ns), top/bottom
    (nq, ns), front/back (nq, nr) (Fortran index order); returns (ns, nr, nq)."""
    nr, ns = left.shape
    nq = top.shape[0]
    q = (np.arange(nq) * (1.0 / (nq - 1)))[None, None, :]
    r = (np.arange(nr) * (1.0 / (nr - 1)))[None, :, None]
    s = (np.arange(ns) * (1.0 / (ns - 1)))[:, None, None]
    Lf, Rf = left.T[:, :, None], right.T[:, :, None]            # (ns, nr, 1)
    Tp, Bt = top.T[:, None, :], bottom.T[:, None, :]            # (ns, 1, nq)
    Fr, Bk = front.T[None, :, :], back.T[None, :, :]            # (1, nr, nq)
    U = (1.0 - q) * Lf + q * Rf
    V = (1.0 - r) * Tp + r * Bt
    W = (1.0 - s) * Fr + s * Bk
    UW = ((1.0 - q) * ((1.0 - r) * left[0, :][:, None, None] + r * left[nr - 1, :][:, None, None])
          + q * ((1.0 - r) * right[0, :][:, None, None] + r * right[nr - 1, :][:, None, None]))
    UV = ((1.0 - r) * ((1.0 - s) * top[:, 0][None, None, :] + s * top[:, ns - 1][None, None, :])
          + r * ((1.0 - s) * bottom[:, 0][None, None, :] + s * bottom[:, ns - 1][None, None, :]))
    VW = ((1.0 - s) * ((1.0 - q) * front[0, :][None, :, None] + q * front[nq - 1, :][None, :, None])
          + s * ((1.0 - q) * back[0, :][None, :, None] + q * back[nq - 1, :][None, :, None]))
    UVW1 = (1.0 - q) * ((1.0 - r) * ((1.0 - s) * top[0, 0] + s * top[0, ns - 1])
                        + r * ((1.0 - s) * bottom[0, 0] + s * bottom[0, ns - 1]))
    UVW2 = q * ((1.0 - r) * ((1.0 - s) * top[nq - 1, 0] + s * top[nq - 1, ns - 1])
                + r * ((1.0 - s) * bottom[nq - 1, 0] + s * bottom[nq - 1, ns - 1]))
    return U + V + W - UV - UW - VW + UVW1 + UVW2


def curve_grid_analytical_test_problem(n, a, b, lc, rc):
    """Curve_Grid_3D3 (grid.f90:473-1909) with profile 'analytical_test_problem'
    and no topography: the x faces carry a Gaussian bump (width^2 0.025) plus
    a sine pattern of amplitude 0.1, scaled by rc (x-low face) and lc (x-high
    face) (:777-833); the other four faces are Coons patches of the x faces'
    edges (:929-1180); the volume is the transfinite interpolation of the six
    faces (:1874-1876).  Returns X (3, ns, nr, nq)."""
    mx, my, mz = n
    ax, ay, az = a
    bx, by, bz = b
    pi = 3.141592653589793
    hy = (by - ay) / float(my - 1)
    hz = (bz - az) / float(mz - 1)
    yline = ay + np.arange(my) * hy
    zline = az + np.arange(mz) * hz
    # x faces: coordinates from their (straight) boundary curves
    Yl, Zl = gen_surf_mesh(np.full(mz, ay), np.full(mz, by), yline, yline,
                           zline, zline, np.full(my, az), np.full(my, bz))
    Yr, Zr = Yl.copy(), Zl.copy()

    def bump(Y, Z, c):
        return (0.1 * c * np.exp(-((1.0 * Y - 0.5 * (ay + by)) ** 2
                                   + 1.0 * (Z - 0.5 * (az + bz)) ** 2) / 0.025)
                + 0.1 * c * np.sin(2.0 / (by - ay) * pi * Y) * np.sin(2.0 / (bz - az) * pi * Z))

    Xl = ax + bump(Yl, Zl, rc)
    Xr = bx + bump(Yr, Zr, lc)
    # front (k = 1) and back (k = mz) faces
    def xface(Xf_l, Xf_r, Yf_l, Yf_r, jlo, jhi):
        hxb = (Xf_r[jlo] - Xf_l[jlo]) / float(mx - 1)
        hxt = (Xf_r[jhi] - Xf_l[jhi]) / float(mx - 1)
        hyb = (Yf_r[jlo] - Yf_l[jlo]) / float(mx - 1)
        hyt = (Yf_r[jhi] - Yf_l[jhi]) / float(mx - 1)
        i = np.arange(mx)
        XB = Xf_l[jlo] + i * hxb
        XT = Xf_l[jhi] + i * hxt
        YB = Yf_l[jlo] + i * hyb
        YT = Yf_l[jhi] + i * hyt
        return XB, XT, YB, YT

    XB, XT, YB, YT = xface(Xl[:, 0], Xr[:, 0], Yl[:, 0], Yr[:, 0], 0, my - 1)
    Xfront, Yfront = gen_surf_mesh(Xl[:, 0], Xr[:, 0], XB, XT, Yl[:, 0], Yr[:, 0], YB, YT)
    Zfront = np.full((mx, my), az)
    XB, XT, YB, YT = xface(Xl[:, mz - 1], Xr[:, mz - 1], Yl[:, mz - 1], Yr[:, mz - 1], 0, my - 1)
    Xback, Yback = gen_surf_mesh(Xl[:, mz - 1], Xr[:, mz - 1], XB, XT, Yl[:, mz - 1],
                                 Yr[:, mz - 1], YB, YT)
    Zback = np.full((mx, my), bz)
    # top (j = 1) and bottom (j = my) faces
    i = np.arange(mx)
    for jj, name in ((0, "top"), (my - 1, "bottom")):
        ZT = Zback[:, jj]
        ZB = Zfront[:, jj]
        XBy = Xl[jj, 0] + i * ((Xr[jj, 0] - Xl[jj, 0]) / float(mx - 1))
        XTy = Xl[jj, mz - 1] + i * ((Xr[jj, mz - 1] - Xl[jj, mz - 1]) / float(mx - 1))
        Xf, Zf = gen_surf_mesh(Xl[jj, :], Xr[jj, :], XBy, XTy, Zl[jj, :], Zr[jj, :], ZB, ZT)
        if name == "top":
            Xtop, Ztop = Xf, Zf
        else:
            Xbottom, Zbottom = Xf, Zf
    Ytop = np.full((mx, mz), ay)
    Ybottom = np.full((mx, mz), by)
    X = np.empty((3, mz, my, mx))
    X[0] = gen3_curve_mesh(Xl, Xr, Xtop, Xbottom, Xfront, Xback)
    X[1] = gen3_curve_mesh(Yl, Yr, Ytop, Ybottom, Yfront, Yback)
    X[2] = gen3_curve_mesh(Zl, Zr, Ztop, Zbottom, Zfront, Zback)
    return X


def gaussian_hills(xs, zs, seed, nhills=12, width=(500.0, 2000.0), height=(-300.0, 300.0)):
    """Surface elevation e(x, z) = sum of seeded Gaussian hills."""
    rng = np.random.default_rng(seed)
    Lx0, Lx1 = xs.min(), xs.max()
    Lz0, Lz1 = zs.min(), zs.max()
    cx = rng.uniform(Lx0, Lx1, nhills)
    cz = rng.uniform(Lz0, Lz1, nhills)
    w = rng.uniform(width[0], width[1], nhills)
    hgt = rng.uniform(height[0], height[1], nhills)
    e = np.zeros((zs.size, xs.size))
    for k in range(nhills):
        e += hgt[k] * np.exp(-(((xs[None, :] - cx[k]) ** 2 + (zs[:, None] - cz[k]) ** 2)
                               / (2.0 * w[k] ** 2)))
    return e


def topography_grid(n, h, seed=7, **hill_kw):
    """Curvilinear block: x, z uniform with spacing h; depth coordinate y runs
    from the surface y_top(x, z) = -e(x, z) (r = 1) down to the flat bottom
    y = (nr-1)*h (r = nr), linearly in r (transfinite blend between the two
    bounding surfaces)."""
    nq, nr, ns = n
    xs = np.arange(nq) * h
    zs = np.arange(ns) * h
    e = gaussian_hills(xs, zs, seed, **hill_kw)       # (ns, nq)
    ytop = -e
    ybot = (nr - 1) * h
    r = np.arange(nr) / (nr - 1)
    X = np.empty((3, ns, nr, nq))
    X[0] = xs[None, None, :]
    X[2] = zs[:, None, None]
    X[1] = ytop[:, None, :] + r[None, :, None] * (ybot - ytop[:, None, :])
    return X


def von_karman(shape, h, seed, std=0.05, corr=1000.0, hurst=0.3):
    """Seeded von Karman random field (spectrum (1 + k^2 a^2)^-(H + 3/2)),
    zero mean, standard deviation ``std``; shape (nz, ny, nx)."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape).astype(np.float32)
    spec = np.fft.rfftn(noise)
    kz = np.fft.fftfreq(shape[0], d=h) * 2 * np.pi
    ky = np.fft.fftfreq(shape[1], d=h) * 2 * np.pi
    kx = np.fft.rfftfreq(shape[2], d=h) * 2 * np.pi
    k2 = (kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2)
    amp = (1.0 + k2 * corr * corr) ** (-(hurst + 1.5) / 2.0)
    field_ = np.fft.irfftn(spec * amp.astype(np.float32), s=shape, axes=(0, 1, 2)).astype(np.float64)
    field_ -= field_.mean()
    sd = field_.std()
    return field_ * (std / sd) if sd > 0 else field_


# ---------------------------------------------------------------------------
# material and anelastic coefficients
# ---------------------------------------------------------------------------
def material_from_rho_s_p(n, rho, cs, cp):
    """init_material default problem (material.f90:696-702)."""
    nq, nr, ns = n
    M = np.empty((3, ns, nr, nq))
    M[0] = rho * (cp * cp - 2.0 * cs * cs)
    M[1] = rho * cs * cs
    M[2] = rho
    return M


def material_from_speeds(rho, cs, cp):
    M = np.empty((3,) + cs.shape)
    M[1] = rho * cs * cs
    M[0] = rho * cp * cp - 2.0 * M[1]
    M[2] = rho
    return M


def q_coefficients(nmech, fmin=0.05, fmax=20.0):
    """build_q{4,8}_fixed_coefficients (anelastic_q8_model.f90:100-116,
    anelastic_q4_model.f90:97-111)."""
    pi = 3.141592653589793
    taumin = 1.0 / (2.0 * pi * fmax)
    taumax = 1.0 / (2.0 * pi * fmin)
    tau = np.array([math.exp(math.log(taumin) + (2.0 * k - 1.0) / (2.0 * nmech)
                             * math.log(taumax / taumin)) for k in range(1, nmech + 1)])
    if nmech == 8:
        w = np.array([1.50589707, 0.0, 0.52793567, 0.53065494,
                      0.32862132, 0.64375916, 0.0, 1.32751442])
    elif nmech == 4:
        w = np.array([1.549360, 0.804277, 0.887718, 1.464160])
    else:
        raise ValueError("nmech must be 4 or 8")
    return tau, w


def modulus_correction(M, Qs_inv, Qp_inv, tau, weight, fref=1.0):
    """Unrelaxed-modulus correction at fref (material.f90:425-453 for Q8,
    258-280 for Q4), per node; returns the corrected M."""
    pi = 3.141592653589793
    wref = 2.0 * pi * fref
    vS = np.zeros_like(Qs_inv)
    vP = np.zeros_like(Qp_inv)
    for l in range(len(tau)):
        dS = (wref * wref * (tau[l] * tau[l]) + 1.0) * (1.0 / Qs_inv)
        dP = (wref * wref * (tau[l] * tau[l]) + 1.0) * (1.0 / Qp_inv)
        vS = vS + weight[l] / dS
        vP = vP + weight[l] / dP
    if np.any(vS >= 1.0) or np.any(vP >= 1.0):
        raise ValueError("invalid anelastic modulus correction: weight/Q sum >= 1")
    lam, mu, rho = M[0], M[1], M[2]
    vs = np.sqrt(mu / rho)
    vp = np.sqrt((lam + 2.0 * mu) / rho)
    muS = rho * (vs * vs) / (1.0 - vS)
    muP = rho * (vp * vp) / (1.0 - vP)
    vs = np.sqrt(muS / rho)
    vp = np.sqrt(muP / rho)
    out = np.empty_like(M)
    out[1] = (vs * vs) * rho
    out[0] = (vp * vp) * rho - 2.0 * out[1]
    out[2] = rho
    return out


def relaxation_dt_limit(tau):
    """q{4,8}_relaxation_dt_limit: 2*min(tau) (anelastic_q8_model.f90:148-156)."""
    return 2.0 * float(np.min(tau))


def block_time_step(h, cfl, cs, cp):
    """block_time_step (block.f90:495-510): cfl*h/sqrt(cs^2 + cp^2)."""
    return cfl * h / math.sqrt(cs * cs + cp * cp)


# ---------------------------------------------------------------------------
# face frames (unit_normals.f90:8-284), vectorised
# ---------------------------------------------------------------------------
def face_frames(metx, mety, metz):
    """local_orth_vectors_{q,r,s} for the six faces.

    Returns lists n_l, n_m, n_n of arrays (3, n2, n1) per face:
    q faces (3, nz, ny), r faces (3, nz, nx), s faces (3, ny, nx)."""
    nl, nm, nn = [], [], []
    for f in range(6):
        ax = f // 2
        idx = 0 if f % 2 == 0 else -1
        sl = [slice(None)] * 3
        sl[2 - ax] = idx                       # array dims are (z, y, x)
        sl = tuple(sl)
        l, m, n = face_frames_from_gradient(metx[ax][sl], mety[ax][sl], metz[ax][sl])
        nl.append(np.ascontiguousarray(l))
        nm.append(np.ascontiguousarray(m))
        nn.append(np.ascontiguousarray(n))
    return nl, nm, nn


# ---------------------------------------------------------------------------
# sources and receivers
# ---------------------------------------------------------------------------
def nearest_node(X, p):
    """Nearest node to point p (1-based), ties to the last node in Fortran
    loop order as in Find_Coordinates (seismogram.f90:767-876)."""
    d2 = (X[0] - p[0]) ** 2 + (X[1] - p[1]) ** 2 + (X[2] - p[2]) ** 2
    flat = d2.ravel()
    best = flat.min()
    k = np.flatnonzero(flat == best)[-1]
    z, y, x = np.unravel_index(k, d2.shape)
    return (int(x) + 1, int(y) + 1, int(z) + 1), tuple(float(X[c, z, y, x]) for c in range(3))


def point_source(X, p, kind, t_init, duration, order=4, **m):
    near, phys = nearest_node(X, p)
    return SourceSpec(kind=kind, t_init=t_init, duration=duration, near=near, near_phys=phys,
                      location=tuple(p), order=order, **m)


def surface_receivers(n, count, seed, margin):
    """Seeded receiver nodes on the r-min (free-surface) face."""
    rng = np.random.default_rng(seed)
    nq, nr, ns = n
    xi = rng.integers(1 + margin, nq - margin + 1, count)
    zi = rng.integers(1 + margin, ns - margin + 1, count)
    return np.stack([xi, np.ones(count, dtype=np.int64), zi], axis=1)


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------
def kat_cartesian_41():
    """inputfile/test_elastic_traditional_o6.in: 41^3 Cartesian [-2,2]^3 (km),
    rho_s_p = (2.7, 3.464, 6), characteristic SAT on all faces, Cos-Bell
    moment source (order 4) at the origin, CFL 0.2, t_final 0.02 (6 steps)."""
    n = (41, 41, 41)
    X = cartesian_grid(n, (-2.0, -2.0, -2.0), (2.0, 2.0, 2.0))
    M = material_from_rho_s_p(n, 2.7, 3.464, 6.0)
    spat = 4.0 / float(41 - 1)
    dt = block_time_step(spat, 0.2, 3.464, 6.0)
    nsteps = int(math.floor(0.02 / dt))
    src = point_source(X, (0.0, 0.0, 0.0), L.WQL_SRC_COSBELL, t_init=0.05, duration=0.0,
                       order=4, mxz=-10.0)
    return Problem(n=n, mesh=L.WQL_CARTESIAN, X=X, M=M, bc=(1,) * 6, dt=dt, nsteps=nsteps,
                   sources=[src], name="kat41")


@dataclass
class TwoBlockProblem:
    """Two blocks coupled across the x-high face of ``left`` and the x-low
    face of ``right`` (nblocks = 2, interface direction 'q'), with the
    interface coupling and the face coordinates (y, z) of every interface
    node as (nz, ny) arrays."""
    left: Problem
    right: Problem
    coupling: int
    yf: np.ndarray
    zf: np.ndarray
    dt: float
    nsteps: int
    name: str = ""
    s0: tuple = None               # TPV28 initial stress (None: the source's values)
    W0: np.ndarray = None          # rate-and-state initial state (None: 0)
    alpha: float = None            # linear friction coefficient (None: the reference's 1e14)


def tpv5_cartesian(n=21, sat_form=1, coupling=None):
    """test_problems/test_rup_cart_fric.in: TPV5 on two n^3 Cartesian blocks,
    block 1 [-3,0] x [6,9] x [-1.5,1.5], block 2 [0,3] x [6,9] x [-1.5,1.5]
    (km), rho_s_p = (2.67, 3.464, 6), lqrs/rqrs = (1,2,1)/(0,1,1) and
    (0,2,1)/(1,1,1), slip-weakening friction, CFL 0.5, t_final 0.5 (46 steps
    at n = 21).  sat_form 1 selects the SAT form of the archived golden files
    (test_problems/truth/1)."""
    if coupling is None:
        coupling = L.WQL_IFACE_SLIP_WEAKENING_TPV5
    nn = (n, n, n)
    XL = cartesian_grid(nn, (-3.0, 6.0, -1.5), (0.0, 9.0, 1.5))
    XR = cartesian_grid(nn, (0.0, 6.0, -1.5), (3.0, 9.0, 1.5))
    M = material_from_rho_s_p(nn, 2.67, 3.464, 6.0)
    spat = 3.0 / float(n - 1)
    dt = block_time_step(spat, 0.5, 3.464, 6.0)
    nsteps = int(math.floor(0.5 / dt))
    left = Problem(n=nn, mesh=L.WQL_CARTESIAN, X=XL, M=M.copy(), bc=(1, 0, 2, 1, 1, 1), dt=dt,
                   nsteps=nsteps, name="tpv5_block1", sat_form=sat_form)
    right = Problem(n=nn, mesh=L.WQL_CARTESIAN, X=XR, M=M.copy(), bc=(0, 1, 2, 1, 1, 1), dt=dt,
                    nsteps=nsteps, name="tpv5_block2", sat_form=sat_form)
    return TwoBlockProblem(left=left, right=right, coupling=coupling,
                           yf=np.ascontiguousarray(XL[1][:, :, -1]),
                           zf=np.ascontiguousarray(XL[2][:, :, -1]), dt=dt, nsteps=nsteps,
                           name=f"tpv5_cartesian_{n}")


def tpv28_curvilinear(n=21, sat_form=1, s0=None):
    """test_problems/test_rup_curv_fric.in: problem TPV28 on two n^3
    curvilinear blocks over the boxes of the Cartesian TPV5 test, meshed by
    Curve_Grid_3D3 with profile 'analytical_test_problem' (block 1: lc = 1,
    rc = 0; block 2: lc = 0, rc = 1, so both carry the same curved fault
    face), rho_s_p = (2.67, 3.464, 6), slip-weakening friction, CFL 0.5,
    t_final 0.5.  sat_form 1 selects the SAT form of the archived golden
    files (test_problems/truth/2); s0 overrides the TPV28 initial stress
    (xx, yy, zz, xy, xz, yz), whose source values are (-60, 0, 60, 0, 29.38,
    0) (initial_stress_condition.f90:78-84) while the golden files were made
    with zz = -60."""
    nn = (n, n, n)
    XL = curve_grid_analytical_test_problem(nn, (-3.0, 6.0, -1.5), (0.0, 9.0, 1.5), 1.0, 0.0)
    XR = curve_grid_analytical_test_problem(nn, (0.0, 6.0, -1.5), (3.0, 9.0, 1.5), 0.0, 1.0)
    M = material_from_rho_s_p(nn, 2.67, 3.464, 6.0)
    dt = block_time_step(3.0 / float(n - 1), 0.5, 3.464, 6.0)
    nsteps = int(math.floor(0.5 / dt))
    left = Problem(n=nn, mesh=L.WQL_CURVILINEAR, X=XL, M=M.copy(), bc=(1, 0, 2, 1, 1, 1), dt=dt,
                   nsteps=nsteps, name="tpv28_block1", sat_form=sat_form)
    right = Problem(n=nn, mesh=L.WQL_CURVILINEAR, X=XR, M=M.copy(), bc=(0, 1, 2, 1, 1, 1), dt=dt,
                    nsteps=nsteps, name="tpv28_block2", sat_form=sat_form)
    return TwoBlockProblem(left=left, right=right, coupling=L.WQL_IFACE_SLIP_WEAKENING_TPV28,
                           yf=np.ascontiguousarray(XL[1][:, :, -1]),
                           zf=np.ascontiguousarray(XL[2][:, :, -1]), dt=dt, nsteps=nsteps,
                           name=f"tpv28_curvilinear_{n}", s0=None if s0 is None else tuple(s0))


def rate_state_initial_state(yf, zf, T0, Vin=1.0e-12, L0=0.02, f0=0.6, b=0.012, V0=1.0e-6):
    """Init_Vel_State for TPV101 / TPV102 (Interface_Condition.f90:2757-2787):
    the state psi in which a fault loaded by the prestress T0 = (T0_n, T0_m,
    T0_l) (3, nz, ny) slides steadily at Vin,
    theta = L0/V0 exp((a ln(2 sinh(T0_l / (a sigma_n))) - f0 - a ln(Vin/V0)) / b),
    psi = f0 + b ln(V0/L0 theta), sigma_n = -T0_n.  (The reference defines
    this routine but does not call it: its interface starts at psi = 0.)"""
    def sbox(x, w, wr):
        ax = np.abs(x)
        out = np.where(ax <= w, 1.0, 0.0)
        mid = (ax > w) & (ax < w + wr)
        with np.errstate(divide="ignore", invalid="ignore"):
            val = 0.5 * (1.0 + np.tanh(wr / (ax - w - wr) + wr / (ax - w)))
        return np.where(mid, val, out)
    a = 0.008 + 0.008 * (1.0 - sbox(yf, 15.0, 3.0) * sbox(zf, 15.0, 3.0))
    sigma_n = -T0[0]
    x = T0[2] / (a * sigma_n)
    # ln(2 sinh x) without overflow: x + ln(1 - e^(-2x))
    log2sinh = x + np.log1p(-np.exp(-2.0 * x))
    log_theta = math.log(L0 / V0) + (a * log2sinh - f0 - a * math.log(Vin / V0)) / b
    return f0 + b * (math.log(V0 / L0) + log_theta)


def tpv101_cartesian(n=21, problem=101, nsteps=None, t_final=1.5):
    """Rate-and-state friction, problem TPV101 or TPV102 (inputfile/tpv101.in,
    tpv102.in; coupling 'rate-and-state_friction') on two n^3 Cartesian
    blocks over the boxes of the TPV5 test ([-3, 0] / [0, 3] x [6, 9] x
    [-1.5, 1.5] km), rho/cs/cp = 2.67/3.464/6, characteristic SAT outside and
    a free surface on y-min, CFL 0.5.  Initial state: the steady state of the
    prestress at Vin = 1e-12 m/s (rate_state_initial_state) and particle
    velocities -+ Vin/2 in z (Init_Vel_State); the nucleation stress ramps in
    over the first second."""
    nn = (n, n, n)
    XL = cartesian_grid(nn, (-3.0, 6.0, -1.5), (0.0, 9.0, 1.5))
    XR = cartesian_grid(nn, (0.0, 6.0, -1.5), (3.0, 9.0, 1.5))
    M = material_from_rho_s_p(nn, 2.67, 3.464, 6.0)
    dt = block_time_step(3.0 / float(n - 1), 0.5, 3.464, 6.0)
    if nsteps is None:
        nsteps = int(math.floor(t_final / dt))
    Vin = 1.0e-12
    F0L = np.zeros((9,) + nn)
    F0R = np.zeros((9,) + nn)
    F0L[2] = -0.5 * Vin
    F0R[2] = 0.5 * Vin
    yf = np.ascontiguousarray(XL[1][:, :, -1])
    zf = np.ascontiguousarray(XL[2][:, :, -1])
    # block L's x-high face: n = (1, 0, 0); prestress = traction of
    # (-120, 0, -120, 0, 75, 0) in its (n, m, l) frame
    one = np.ones_like(yf)
    l, m, nv = face_frames_from_gradient(one, 0.0 * one, 0.0 * one)
    s = (-120.0, 0.0, -120.0, 0.0, 75.0, 0.0)
    Tx = nv[0] * s[0] + nv[1] * s[3] + nv[2] * s[4]
    Ty = nv[0] * s[3] + nv[1] * s[1] + nv[2] * s[5]
    Tz = nv[0] * s[4] + nv[1] * s[5] + nv[2] * s[2]
    T0 = np.stack([nv[0] * Tx + nv[1] * Ty + nv[2] * Tz, m[0] * Tx + m[1] * Ty + m[2] * Tz,
                   l[0] * Tx + l[1] * Ty + l[2] * Tz])
    W0 = rate_state_initial_state(yf, zf, T0, Vin)
    coupling = L.WQL_IFACE_RATE_STATE_TPV101 if problem == 101 else L.WQL_IFACE_RATE_STATE_TPV102
    left = Problem(n=nn, mesh=L.WQL_CARTESIAN, X=XL, M=M.copy(), bc=(1, 0, 2, 1, 1, 1), dt=dt,
                   nsteps=nsteps, F0=F0L, name=f"tpv{problem}_block1")
    right = Problem(n=nn, mesh=L.WQL_CARTESIAN, X=XR, M=M.copy(), bc=(0, 1, 2, 1, 1, 1), dt=dt,
                    nsteps=nsteps, F0=F0R, name=f"tpv{problem}_block2")
    tp = TwoBlockProblem(left=left, right=right, coupling=coupling, yf=yf, zf=zf, dt=dt,
                         nsteps=nsteps, name=f"tpv{problem}_cartesian_{n}")
    tp.W0 = W0
    return tp


def rate_weakening_cartesian(n=21, nsteps=None, t_final=0.6):
    """Rate-and-state friction, problem 'rate-weakening' (Interface_Condition.
    f90:410-416, 438-444, 2387-2426), on two n^3 Cartesian blocks
    [-3, 0] / [0, 3] x [6.5, 9.5] x [-1.5, 1.5] km around the pulse centre
    (y, z) = (8, 0), rho/cs/cp = 2.67/3.464/6, characteristic SAT outside and
    a free surface on y-min, CFL 0.5.  Initial state psi = 0.4367 and zero
    particle velocity (Init_Vel_State, :2822-2845); the prestress pulse is on
    from t > 0."""
    nn = (n, n, n)
    XL = cartesian_grid(nn, (-3.0, 6.5, -1.5), (0.0, 9.5, 1.5))
    XR = cartesian_grid(nn, (0.0, 6.5, -1.5), (3.0, 9.5, 1.5))
    M = material_from_rho_s_p(nn, 2.67, 3.464, 6.0)
    dt = block_time_step(3.0 / float(n - 1), 0.5, 3.464, 6.0)
    if nsteps is None:
        nsteps = int(math.floor(t_final / dt))
    yf = np.ascontiguousarray(XL[1][:, :, -1])
    zf = np.ascontiguousarray(XL[2][:, :, -1])
    left = Problem(n=nn, mesh=L.WQL_CARTESIAN, X=XL, M=M.copy(), bc=(1, 0, 2, 1, 1, 1), dt=dt,
                   nsteps=nsteps, name="rw_block1")
    right = Problem(n=nn, mesh=L.WQL_CARTESIAN, X=XR, M=M.copy(), bc=(0, 1, 2, 1, 1, 1), dt=dt,
                    nsteps=nsteps, name="rw_block2")
    tp = TwoBlockProblem(left=left, right=right, coupling=L.WQL_IFACE_RATE_STATE_RW, yf=yf,
                         zf=zf, dt=dt, nsteps=nsteps, name=f"rate_weakening_cartesian_{n}")
    tp.W0 = np.full(yf.shape, 0.4367)
    return tp


def linear_friction_curvilinear(n=24, alpha=5.0, nmech=4, npml=6, nsteps=30):
    """Linear friction tau = alpha v (coupling 'linear_friction',
    Interface_Condition.f90:360-366, 639-742) across the curved interface of
    two_block_curvilinear: the Gaussian pulse of block L crosses it with slip
    (alpha in MPa s/m; the shear impedance harmonic mean of the two sides is
    about 4.7)."""
    tp = two_block_curvilinear(n=n, nmech=nmech, npml=npml, coupling=L.WQL_IFACE_LINEAR_FRICTION,
                               nsteps=nsteps)
    tp.alpha = alpha
    tp.name = f"linear_friction_curvilinear_{n}_alpha{alpha:g}"
    return tp


def fault_two_block(n=(201, 401, 201), h=0.1, nmech=8, npml=20, cfl=0.5, nsteps=2000,
                    curvilinear=True, seed=2025, qs=100.0, qp=200.0, nrec=32):
    """Two blocks of n = (nq, nr, ns) nodes each, block L over x in
    [-(nq-1)h, 0] and block R over [0, (nq-1)h], y (depth, r) from the
    surface to (nr-1)h, z in [-(ns-1)h/2, (ns-1)h/2] (km), coupled by a planar
    vertical TPV5 slip-weakening fault at x = 0 (mu_s 0.677, mu_d 0.525, Dc
    0.4 m, sigma_n -120 MPa, tau 70 MPa, 81.6 MPa in the 3 km nucleation
    patch at y = 7.5 km, z = 0); rho/cs/cp = 2.67/3.464/6, anelastic Q (Qs,
    Qp), free surface on y-min, PML of npml nodes on the other outer faces.

    curvilinear: both blocks are curvilinear, the surface carries 12 seeded
    Gaussian hills (widths 0.5-2 km, heights -0.3..0.3 km) over the whole
    domain and each column blends linearly from its surface point to the flat
    bottom (the transfinite grid of topography_grid); x depends on q only, so
    the fault stays the plane x = 0 and its nodes coincide on both sides.
    Otherwise both blocks are Cartesian with a flat surface.

    Receivers: nrec surface nodes, half in each block, 2-8 km from the fault
    (off-fault stations)."""
    nq, nr, ns = n
    Lz = (ns - 1) * h
    Ly = (nr - 1) * h
    mesh = L.WQL_CURVILINEAR if curvilinear else L.WQL_CARTESIAN

    def grid(x0, x1):
        X = cartesian_grid(n, (x0, 0.0, -0.5 * Lz), (x1, Ly, 0.5 * Lz))
        if curvilinear:
            xs_all = np.linspace(-(nq - 1) * h, (nq - 1) * h, 2 * nq - 1)
            zs = X[2][:, 0, 0]
            rng = np.random.default_rng(seed)
            cx = rng.uniform(xs_all[0], xs_all[-1], 12)
            cz = rng.uniform(zs[0], zs[-1], 12)
            w = rng.uniform(0.5, 2.0, 12)
            hgt = rng.uniform(-0.3, 0.3, 12)
            xs = X[0][0, 0, :]
            e = np.zeros((ns, nq))
            for k in range(12):
                e += hgt[k] * np.exp(-(((xs[None, :] - cx[k]) ** 2 + (zs[:, None] - cz[k]) ** 2)
                                       / (2.0 * w[k] ** 2)))
            r = np.arange(nr) / (nr - 1)
            ytop = -e
            X[1] = ytop[:, None, :] + r[None, :, None] * (Ly - ytop[:, None, :])
        return X

    XL = grid(-(nq - 1) * h, 0.0)
    XR = grid(0.0, (nq - 1) * h)
    M = material_from_rho_s_p(n, 2.67, 3.464, 6.0)
    tau = weight = Qs_inv = Qp_inv = None
    if nmech:
        tau, weight = q_coefficients(nmech)
        Qs_inv = np.full(XL.shape[1:], 1.0 / qs)
        Qp_inv = np.full(XL.shape[1:], 1.0 / qp)
        M = modulus_correction(M, Qs_inv, Qp_inv, tau, weight)
    dt = block_time_step(h, cfl, 3.464, 6.0)
    if nmech:
        dt = min(dt, relaxation_dt_limit(tau))
    pmlL = (1, 0, 0, 1, 1, 1) if npml else (0,) * 6
    pmlR = (0, 1, 0, 1, 1, 1) if npml else (0,) * 6
    rng = np.random.default_rng(seed + 1)
    half = nrec // 2
    margin = max(npml, 6) + 1

    def stations(side):
        # surface nodes 2-8 km from the fault, clear of the PML slabs
        dist = rng.uniform(2.0, 8.0, half)
        ix = np.rint(dist / h).astype(int)
        ix = np.clip(ix, 0, nq - 1 - margin)
        ix = (nq - 1 - ix) if side == 0 else ix
        iz = rng.integers(margin, ns - margin, half)
        return np.stack([ix + 1, np.ones(half, dtype=int), iz + 1], axis=1)

    kw = dict(mesh=mesh, dt=dt, nsteps=nsteps, nmech=nmech, tau=tau, weight=weight,
              Qs_inv=Qs_inv, Qp_inv=Qp_inv)
    left = Problem(n=n, X=XL, M=M, bc=(1, 0, 2, 1, 1, 1), pml=pmlL,
                   npml=tuple(npml * p for p in pmlL), receivers=stations(0),
                   name="fault_block1", **kw)
    right = Problem(n=n, X=XR, M=M.copy(), bc=(0, 1, 2, 1, 1, 1), pml=pmlR,
                    npml=tuple(npml * p for p in pmlR), receivers=stations(1),
                    name="fault_block2", **kw)
    kind = "curvilinear" if curvilinear else "cartesian"
    return TwoBlockProblem(left=left, right=right, coupling=L.WQL_IFACE_SLIP_WEAKENING_TPV5,
                           yf=np.ascontiguousarray(XL[1][:, :, -1]),
                           zf=np.ascontiguousarray(XL[2][:, :, -1]), dt=dt, nsteps=nsteps,
                           name=f"fault_two_block_{kind}_{nq}x{nr}x{ns}")


def locked_interface_41():
    """inputfile/test_elastic_locked_interface_o6.in: two 41^3 Cartesian blocks,
    [-2,0] x [-1,1]^2 and [0,2] x [-1,1]^2 (km), rho_s_p = (2.7, 3.464, 6),
    locked interface, characteristic SAT on the outer faces, CFL 0.2,
    t_final 0.15 (103 steps); block 1 holds an order-4 Cos-Bell moment
    source at the origin (on its interface face) with mXX = -0.25,
    mXZ = -10, t_init = 0.05."""
    nn = (41, 41, 41)
    XL = cartesian_grid(nn, (-2.0, -1.0, -1.0), (0.0, 1.0, 1.0))
    XR = cartesian_grid(nn, (0.0, -1.0, -1.0), (2.0, 1.0, 1.0))
    M = material_from_rho_s_p(nn, 2.7, 3.464, 6.0)
    dt = block_time_step(2.0 / 40.0, 0.2, 3.464, 6.0)
    nsteps = int(math.floor(0.15 / dt))
    src = point_source(XL, (0.0, 0.0, 0.0), L.WQL_SRC_COSBELL, t_init=0.05, duration=0.0,
                       order=4, mxx=-0.25, mxz=-10.0)
    left = Problem(n=nn, mesh=L.WQL_CARTESIAN, X=XL, M=M.copy(), bc=(1, 0, 1, 1, 1, 1), dt=dt,
                   nsteps=nsteps, sources=[src], name="locked41_block1")
    right = Problem(n=nn, mesh=L.WQL_CARTESIAN, X=XR, M=M.copy(), bc=(0, 1, 1, 1, 1, 1), dt=dt,
                    nsteps=nsteps, name="locked41_block2")
    return TwoBlockProblem(left=left, right=right, coupling=L.WQL_IFACE_LOCKED,
                           yf=np.ascontiguousarray(XL[1][:, :, -1]),
                           zf=np.ascontiguousarray(XL[2][:, :, -1]), dt=dt, nsteps=nsteps,
                           name="locked_interface_41")


def two_block_curvilinear(n=24, nmech=4, npml=6, coupling=None, cfl=0.4, nsteps=30):
    """Two curvilinear n^3 blocks, block L over x in [-3, 0] and block R over
    [0, 3] (y, z in [0, 3], km), both mapped by the same smooth map so the
    interface x = 0 becomes a curved surface whose nodes coincide; different
    materials on the two sides (2.7/3.464/6 and 2.4/2.9/5.1), anelastic Q
    (nmech mechanisms, Qs = 50, Qp = 100), PML of npml nodes on the outer x
    faces, free surface on y-min, characteristic SAT elsewhere; a Gaussian
    velocity pulse in block L next to the interface."""
    if coupling is None:
        coupling = L.WQL_IFACE_LOCKED
    nn = (n, n, n)

    def mapped(a, b):
        X = cartesian_grid(nn, a, b)
        x, y, z = X[0].copy(), X[1].copy(), X[2].copy()
        X[0] = x + 0.12 * np.sin(2.0 * np.pi * y / 3.0) * np.sin(2.0 * np.pi * z / 3.0)
        X[1] = y + 0.08 * np.sin(np.pi * x / 3.0) * np.cos(np.pi * z / 3.0)
        X[2] = z + 0.05 * np.sin(np.pi * y / 3.0) * np.sin(np.pi * (x + 3.0) / 6.0)
        return X

    XL = mapped((-3.0, 0.0, 0.0), (0.0, 3.0, 3.0))
    XR = mapped((0.0, 0.0, 0.0), (3.0, 3.0, 3.0))
    ML = material_from_rho_s_p(nn, 2.7, 3.464, 6.0)
    MR = material_from_rho_s_p(nn, 2.4, 2.9, 5.1)
    tau = weight = Qs_inv = Qp_inv = None
    if nmech:
        tau, weight = q_coefficients(nmech)
        Qs_inv = np.full(XL.shape[1:], 1.0 / 50.0)
        Qp_inv = np.full(XL.shape[1:], 1.0 / 100.0)
        ML = modulus_correction(ML, Qs_inv, Qp_inv, tau, weight)
        MR = modulus_correction(MR, Qs_inv, Qp_inv, tau, weight)
    dt = block_time_step(3.0 / (n - 1) * 0.8, cfl, 3.464, 6.0)
    if nmech:
        dt = min(dt, relaxation_dt_limit(tau))
    r2 = (XL[0] + 0.6) ** 2 + (XL[1] - 1.5) ** 2 + (XL[2] - 1.5) ** 2
    g = np.exp(-r2 / 0.15)
    F0 = np.zeros((9,) + XL.shape[1:])
    F0[0] = g
    F0[2] = 0.5 * g
    pmlL = (1, 0, 0, 0, 0, 0) if npml else (0,) * 6
    pmlR = (0, 1, 0, 0, 0, 0) if npml else (0,) * 6
    nL = tuple(npml * p for p in pmlL)
    nR = tuple(npml * p for p in pmlR)
    kw = dict(mesh=L.WQL_CURVILINEAR, dt=dt, nsteps=nsteps, nmech=nmech, tau=tau, weight=weight,
              Qs_inv=Qs_inv, Qp_inv=Qp_inv)
    left = Problem(n=nn, X=XL, M=ML, bc=(1, 0, 2, 1, 1, 1), pml=pmlL, npml=nL, F0=F0,
                   name="curv_block1", **kw)
    right = Problem(n=nn, X=XR, M=MR, bc=(0, 1, 2, 1, 1, 1), pml=pmlR, npml=nR,
                    name="curv_block2", **kw)
    return TwoBlockProblem(left=left, right=right, coupling=coupling,
                           yf=np.ascontiguousarray(XL[1][:, :, -1]),
                           zf=np.ascontiguousarray(XL[2][:, :, -1]), dt=dt, nsteps=nsteps,
                           name=f"two_block_curvilinear_{n}")


def cartesian_bump(n=65, h=100.0, nsteps=200, seed=1234, cfl=0.5, nrec=8):
    """Plain FP64 Cartesian block: n^3 nodes, spacing h (m); rho = 2670,
    cp = 6000, cs = 3464; free surface on r-min, characteristic SAT elsewhere;
    zero stress and a Gaussian vx bump (sigma 300 m, 1 m/s) at the centre plus
    seeded U[-1e-6, 1e-6] noise on all 9 fields."""
    nn = (n, n, n)
    L_ = (n - 1) * h
    X = cartesian_grid(nn, (0.0, 0.0, 0.0), (L_, L_, L_))
    rho, cp, cs = 2670.0, 6000.0, 3464.0
    M = material_from_rho_s_p(nn, rho, cs, cp)
    c = L_ / 2.0
    r2 = (X[0] - c) ** 2 + (X[1] - c) ** 2 + (X[2] - c) ** 2
    rng = np.random.default_rng(seed)
    F0 = rng.uniform(-1.0e-6, 1.0e-6, size=(9,) + X.shape[1:])
    F0[0] += np.exp(-r2 / (2.0 * 300.0 ** 2))
    dt = block_time_step(h, cfl, cs, cp)
    bc = (1, 1, 2, 1, 1, 1)
    rec = surface_receivers(nn, nrec, seed + 1, margin=2)
    return Problem(n=nn, mesh=L.WQL_CARTESIAN, X=X, M=M, bc=bc, dt=dt, nsteps=nsteps, F0=F0,
                   receivers=rec, name=f"cartesian_bump_{n}")


def curvilinear_q(n=257, h=50.0, nsteps=500, nmech=8, npml=20, seed=2024, cfl=0.5,
                  nrec=64, f0=2.0, vk=True, hills=True, layered=True, qs_scale=50.0):
    """Curvilinear block with topography (12 seeded Gaussian hills), layered
    cs from 2000 to 3500 m/s with depth, cp = sqrt(3)*cs, rho = 2700, seeded
    5 % von Karman perturbation (1 km, H 0.3), anelastic Q (nmech = 4 or 8)
    with Qs = qs_scale*cs[km/s], Qp = 2 Qs; PML of npml nodes on the four
    sides and the bottom; free surface on top (r-min); Ricker moment-tensor
    source (f0) and nrec surface receivers."""
    nn = (n, n, n)
    X = topography_grid(nn, h, seed=seed) if hills else cartesian_grid(
        nn, (0.0, 0.0, 0.0), ((n - 1) * h,) * 3)
    depth = X[1] - X[1][:, :1, :]                  # depth below local surface
    Ldep = (n - 1) * h
    cs = 2000.0 + 1500.0 * np.clip(depth / Ldep, 0.0, 1.0) if layered else np.full(X.shape[1:], 3000.0)
    if vk:
        cs = cs * (1.0 + von_karman(X.shape[1:], h, seed + 11))
    cp = math.sqrt(3.0) * cs
    rho = np.full(X.shape[1:], 2700.0)
    M = material_from_speeds(rho, cs, cp)
    tau = weight = Qs_inv = Qp_inv = None
    if nmech:
        tau, weight = q_coefficients(nmech)
        Qs = qs_scale * cs / 1000.0
        Qs_inv = 1.0 / Qs
        Qp_inv = 1.0 / (2.0 * Qs)
        M = modulus_correction(M, Qs_inv, Qp_inv, tau, weight)
    cs_ref, cp_ref = 3500.0, math.sqrt(3.0) * 3500.0
    dt = block_time_step(h, cfl, cs_ref * 1.2, cp_ref * 1.2)
    if nmech:
        dt = min(dt, relaxation_dt_limit(tau))
    if npml:
        pml = (1, 1, 0, 1, 1, 1)
        npm = (npml, npml, 0, npml, npml, npml)
    else:
        pml, npm = (0,) * 6, (0,) * 6
    bc = (1, 1, 2, 1, 1, 1)
    Lh = (n - 1) * h
    src_pt = (0.5 * Lh, 0.35 * Lh, 0.5 * Lh)
    src = point_source(X, src_pt, L.WQL_SRC_RICKER, t_init=1.2 / f0, duration=f0, order=4,
                       mxx=1.0e15, myy=-0.5e15, mzz=-0.5e15, mxy=0.3e15)
    rec = surface_receivers(nn, nrec, seed + 5, margin=max(npml, 6) + 1)
    return Problem(n=nn, mesh=L.WQL_CURVILINEAR, X=X, M=M, bc=bc, dt=dt, nsteps=nsteps,
                   pml=pml, npml=npm, nmech=nmech, tau=tau, weight=weight, Qs_inv=Qs_inv,
                   Qp_inv=Qp_inv, sources=[src], receivers=rec,
                   name=f"curvilinear_q{nmech}_{n}")


def differentiation_q4(n=129, h=100.0, nsteps=400, f0=1.5, seed=4):
    """The differentiation workload: n^3 curvilinear block, h = 100 m,
    seeded topography, 5 % von Karman perturbations of the wave speeds and an
    independent 5 % one of rho (lambda and mu follow rho, so the speeds keep
    theirs), Q4, characteristic SAT on the sides and bottom, free surface on
    top, Ricker f0 = 1.5 Hz, 64 surface receivers."""
    prob = curvilinear_q(n=n, h=h, nsteps=nsteps, nmech=4, npml=0, seed=seed, nrec=64, f0=f0)
    rho = prob.M[2] * (1.0 + von_karman(prob.M.shape[1:], h, seed + 23))
    ratio = rho / prob.M[2]
    prob.M = np.stack([prob.M[0] * ratio, prob.M[1] * ratio, rho])
    prob.name = f"differentiation_q4_{n}"
    return prob


# ---------------------------------------------------------------------------
def make_device_block(prob: Problem, precision=L.WQL_FP64, device="cuda", state_mirror=False):
    """Create a :class:`Block` whose set-up arrays are float64 torch tensors on
    the GPU: metric terms from the device pipeline (grid_metrics), face
    frames from the device metrics, material and Q from the problem.  The
    state starts at the init_* values (F = 0, DF = 1e40, memory = 0).  With
    state_mirror the block also gets device tensors ``F`` (and ``eta``,
    ``Deta`` when anelastic) that ``download`` fills and ``upload`` reads."""
    import torch

    from .block import Block, grid_metrics
    dev = torch.device(device)
    X = torch.from_numpy(np.ascontiguousarray(prob.X)).to(dev)
    metx, mety, metz, J = grid_metrics(X)
    del X
    # face frames from the device metrics (small face slices copied to host)
    fm = []
    for m in (metx, mety, metz):
        fm.append(m)
    nl, nm_, nn_ = [], [], []
    for f in range(6):
        ax = f // 2
        idx = 0 if f % 2 == 0 else -1
        sl = [slice(None)] * 3
        sl[2 - ax] = idx
        sl = tuple(sl)
        q = [m[ax][sl].cpu().numpy() for m in fm]
        parts = face_frames_from_gradient(*q)
        for lst, arr in zip((nl, nm_, nn_), parts):
            lst.append(torch.from_numpy(np.ascontiguousarray(arr)).to(dev))
    kw = {}
    if prob.nmech:
        kw.update(Qs_inv=torch.from_numpy(np.ascontiguousarray(prob.Qs_inv)).to(dev),
                  Qp_inv=torch.from_numpy(np.ascontiguousarray(prob.Qp_inv)).to(dev))
    nq, nr, ns = prob.n
    if state_mirror:
        kw["F"] = torch.zeros((9, ns, nr, nq), dtype=torch.float64, device=dev)
        if prob.nmech:
            for name in ("eta", "Deta"):
                kw[name] = torch.zeros((6, prob.nmech, ns, nr, nq), dtype=torch.float64, device=dev)
    blk = Block(nq=nq, nr=nr, ns=ns, mesh=prob.mesh, precision=precision, nmech=prob.nmech,
                bc=prob.bc, pml=prob.pml, npml=prob.npml,
                tau=tuple(prob.tau) if prob.tau is not None else (),
                weight=tuple(prob.weight) if prob.weight is not None else (),
                metricx=metx, metricy=mety, metricz=metz, J=J,
                M=torch.from_numpy(np.ascontiguousarray(prob.M)).to(dev),
                n_l=nl, n_m=nm_, n_n=nn_, on_device=True, sat_form=prob.sat_form, **kw)
    blk.upload(L.WQL_GEOMETRY | L.WQL_MATERIAL | L.WQL_QINV | L.WQL_FACES)
    if any(src.order == 0 for src in prob.sources):
        blk.set_coordinates(prob.X)
    blk.set_sources(prob.sources)
    return blk


def face_frames_from_gradient(qx, qy, qz):
    """gen_orth_vectors applied to n = grad(xi)/|grad(xi)| on one face
    (unit_normals.f90:8-284); returns (l, m, n), each (3,) + face shape."""
    q = np.stack([qx, qy, qz])
    nrm = np.sqrt(q[0] ** 2 + q[1] ** 2 + q[2] ** 2)
    n = q / nrm
    tol = 1.0e-12
    m = np.zeros_like(n)
    d1 = np.sqrt(n[0] ** 2 + (n[1] - 1.0) ** 2 + n[2] ** 2)
    d2 = np.sqrt(n[0] ** 2 + (n[1] + 1.0) ** 2 + n[2] ** 2)
    alt = ~((d1 >= tol) & (d2 >= tol))
    m[1] = np.where(alt, 0.0, 1.0)
    m[2] = np.where(alt, 1.0, 0.0)
    a = n[0] * m[0] + n[1] * m[1] + n[2] * m[2]
    m = m - a * n
    m = (1.0 / np.sqrt(m[0] ** 2 + m[1] ** 2 + m[2] ** 2)) * m
    l = np.stack([n[1] * m[2] - n[2] * m[1],
                  n[2] * m[0] - n[0] * m[2],
                  n[0] * m[1] - n[1] * m[0]])
    return l, m, n


def make_block(prob: Problem, metrics, precision=L.WQL_FP64, frames=None, upload=True):
    """Create a host-array :class:`Block` from a problem and its metrics and
    upload it (F from prob.F0 if given, otherwise the init_* state)."""
    from .block import Block
    metx, mety, metz, J = metrics
    if frames is None:
        frames = face_frames(metx, mety, metz)
    nl, nm, nn = frames
    nq, nr, ns = prob.n
    kw = {}
    if prob.nmech:
        shp = (6, prob.nmech, ns, nr, nq)
        kw.update(eta=np.zeros(shp), Deta=np.zeros(shp),
                  Qs_inv=np.ascontiguousarray(prob.Qs_inv), Qp_inv=np.ascontiguousarray(prob.Qp_inv))
    pq, pdq = [None] * 6, [None] * 6
    from .block import pml_slab_box
    for f in range(6):
        if prob.pml[f]:
            lo, hi = pml_slab_box(f, prob.n, prob.npml[f])
            shp = (9, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
            pq[f] = np.zeros(shp)
            pdq[f] = np.full(shp, 1.0e40)
    # a copy: downloads write the block's F, which must not alias the problem's F0
    F = (np.array(prob.F0, dtype=np.float64, order="C", copy=True) if prob.F0 is not None
         else np.zeros((9, ns, nr, nq)))
    blk = Block(nq=nq, nr=nr, ns=ns, mesh=prob.mesh, precision=precision, nmech=prob.nmech,
                bc=prob.bc, pml=prob.pml, npml=prob.npml,
                tau=tuple(prob.tau) if prob.tau is not None else (),
                weight=tuple(prob.weight) if prob.weight is not None else (),
                metricx=np.ascontiguousarray(metx), metricy=np.ascontiguousarray(mety),
                metricz=np.ascontiguousarray(metz), J=np.ascontiguousarray(J),
                M=np.ascontiguousarray(prob.M), F=F, DF=np.full((9, ns, nr, nq), 1.0e40),
                pml_Q=pq, pml_DQ=pdq, n_l=nl, n_m=nm, n_n=nn, sat_form=prob.sat_form, **kw)
    if upload:
        blk.upload(L.WQL_GEOMETRY | L.WQL_MATERIAL | L.WQL_QINV | L.WQL_FACES | L.WQL_F)
        if any(src.order == 0 for src in prob.sources):
            blk.set_coordinates(prob.X)
        blk.set_sources(prob.sources)
    return blk
