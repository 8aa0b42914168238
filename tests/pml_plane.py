"""Plane-wave pulses into PML slabs (shared by the PML physics tests).

A block long along one axis (100 nodes, h = 50 m) and 12 nodes across it,
with a 20-node PML slab at both ends of the long axis (PMLB of those two
faces, pml.f90:11-168) and a free surface (bc 2) on those two faces, so that
without the PML every pulse returns whole; the four faces across carry no
penalty (bc 0), so a field that depends on the long axis only stays one
(the SBP derivative of a constant is zero).  The initial state is a Gaussian
velocity pulse (sigma 150 m) at the centre, which splits into two pulses
running into the two slabs: P for velocity along the axis, S across it.
Measured: the SBP-norm energy of the whole block at every step, and the
energy left outside the slabs once the incident pulses have left that region
(the reflected energy)."""
import math

import numpy as np

from waveqlab3d_q_gpu_amd import problems as PB

H_PHYS, NPML, N_AX, N_T = 50.0, 20, 100, 12
RHO, CS, CP = 2670.0, 3464.0, 6000.0
SIGMA = 150.0


def plane_problem(axis, curvilinear=False, pml=True, outer_bc=2):
    n = [N_T] * 3
    n[axis] = N_AX
    nn = tuple(n)
    X = PB.cartesian_grid(nn, (0.0,) * 3, tuple((k - 1) * H_PHYS for k in nn))
    if curvilinear:
        # the coordinates across the axis bend with position along it (a
        # curvilinear mesh with non-diagonal metrics); the axis coordinate
        # itself depends on its own index only, so plane waves stay planes
        others = [a for a in range(3) if a != axis]
        X[others[0]] = X[others[0]] + 0.3 * H_PHYS * np.sin(
            2.0 * np.pi * X[axis] / (N_AX * H_PHYS)) * np.cos(
            np.pi * X[others[1]] / ((N_T - 1) * H_PHYS))
    M = PB.material_from_rho_s_p(nn, RHO, CS, CP)
    bc, pm, npm = [0] * 6, [0] * 6, [0] * 6
    for f in (2 * axis, 2 * axis + 1):
        bc[f] = outer_bc
        if pml:
            pm[f], npm[f] = 1, NPML
    dt = PB.block_time_step(H_PHYS, 0.5, CS, CP)
    return PB.Problem(n=nn, mesh=1 if curvilinear else 0, X=X, M=M, bc=tuple(bc), dt=dt,
                      nsteps=0, pml=tuple(pm), npml=tuple(npm),
                      name=f"plane_axis{axis}{'_curv' if curvilinear else ''}")


def initial_pulse(prob, axis, comp):
    F = np.zeros((9,) + prob.X.shape[1:])
    xa = prob.X[axis]
    F[comp] = np.exp(-(xa - 0.5 * xa.max()) ** 2 / (2.0 * SIGMA ** 2))
    return F


def schedule(prob, axis, comp):
    """(nsteps, t_clear): 1.3 times the time for the pulse to reach a slab,
    cross it, return and reach the centre again; t_clear = the time the
    incident pulse's tail (4 sigma) has left the region outside the slabs."""
    c = CP if comp == axis else CS
    dist = (N_AX // 2 - NPML) * H_PHYS
    T = 1.3 * (2.0 * dist + 2.0 * NPML * H_PHYS) / c
    return int(math.ceil(T / prob.dt)), (dist + 4.0 * SIGMA) / c


def norm_weights(n):
    """Diagonal SBP-6 norm H (exact closure weights 13649/43200, ...)."""
    h = np.ones(n)
    w = [13649 / 43200, 12013 / 8640, 2711 / 4320, 5359 / 4320, 7877 / 8640, 43801 / 43200]
    h[:6] = w
    h[-6:] = w[::-1]
    return h / (n - 1)


class Energy:
    """SBP-norm energy 1/2 sum H J (rho |v|^2 + sigma : C^-1 : sigma) of the
    block, and of the nodes outside the slabs."""

    def __init__(self, prob, J, axis):
        nq, nr, ns = prob.n
        H = [norm_weights(nq), norm_weights(nr), norm_weights(ns)]
        self.w = H[2][:, None, None] * H[1][None, :, None] * H[0][None, None, :] * J
        lam, mu, rho = prob.M
        self.lam, self.mu, self.rho = lam, mu, rho
        idx = np.arange(prob.n[axis])
        inner = (idx >= NPML) & (idx < prob.n[axis] - NPML)
        shp = [1, 1, 1]
        shp[2 - axis] = -1
        self.inner = np.broadcast_to(inner.reshape(shp), J.shape)

    def density(self, F):
        v2 = F[0] ** 2 + F[1] ** 2 + F[2] ** 2
        s = F[3:]
        tr = s[0] + s[1] + s[2]
        ss = s[0] ** 2 + s[1] ** 2 + s[2] ** 2 + 2.0 * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2)
        comp = (ss - self.lam / (3.0 * self.lam + 2.0 * self.mu) * tr ** 2) / (2.0 * self.mu)
        return 0.5 * self.w * (self.rho * v2 + comp)

    def __call__(self, F):
        e = self.density(F)
        return float(e.sum()), float(e[self.inner].sum())


def run(step, prob, axis, comp, J):
    """Drive ``step()`` (one RK step, returns F) over the schedule; returns
    (E0, energies of the block per step, reflected energy fraction)."""
    nsteps, t_clear = schedule(prob, axis, comp)
    en = Energy(prob, J, axis)
    E0, _ = en(initial_pulse(prob, axis, comp))
    tot, inner = [], []
    for k in range(nsteps):
        F = step()
        a, b = en(F)
        tot.append(a); inner.append(b)
    t = prob.dt * np.arange(1, nsteps + 1)
    refl = max(v for v, tt in zip(inner, t) if tt > t_clear) / E0
    return E0, np.array(tot), refl
