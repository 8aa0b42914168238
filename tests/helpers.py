"""Shared set-up for the parity tests: one Problem drives both the oracle
block and the GPU block with identical inputs (the oracle's metrics)."""
import numpy as np

from waveqlab3d_q_gpu_amd import _lib as L
from waveqlab3d_q_gpu_amd import problems as PB


def oracle_block(O, prob, metrics):
    nq, nr, ns = prob.n
    b = O.OracleBlock(nq, nr, ns, curvilinear=prob.mesh)
    b.metx, b.mety, b.metz, b.J = metrics
    b.M[...] = prob.M
    b.bc = list(prob.bc)
    b.sat_form = prob.sat_form
    for f in range(6):
        if prob.pml[f]:
            b.set_pml(f, prob.npml[f])
    if prob.nmech:
        b.set_anelastic(prob.nmech, prob.Qs_inv, prob.Qp_inv, correct_modulus=False)
    b.normals()
    for s in prob.sources:
        b.add_source(kind=s.kind, t_init=s.t_init, duration=s.duration, mxx=s.mxx, myy=s.myy,
                     mzz=s.mzz, mxy=s.mxy, mxz=s.mxz, myz=s.myz, loc=s.location, near=s.near,
                     near_phys=s.near_phys, order=s.order, alpha=s.alpha)
    if any(s.order == 0 for s in prob.sources):
        b.X = np.ascontiguousarray(prob.X, dtype=np.float64)
    if prob.F0 is not None:
        b.F[...] = prob.F0
    return b


def gpu_block(prob, metrics, precision=L.WQL_FP64, frames=None):
    return PB.make_block(prob, metrics, precision=precision, frames=frames)


def rel_l2(a, b):
    nb = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0)


def field_errors(gpu_F, orc_F):
    return [rel_l2(gpu_F[c], orc_F[c]) for c in range(gpu_F.shape[0])]


def receiver_flat(prob):
    nq, nr, ns = prob.n
    r = np.asarray(prob.receivers)
    return (r[:, 0] - 1) + nq * ((r[:, 1] - 1) + nr * (r[:, 2] - 1))


def small_curvilinear(n=30, nmech=8, npml=8, nsteps=12, seed=5):
    """A reduced copy of the production curvilinear set-up that the oracle
    finishes in seconds (same physics, smaller block, h scaled so the
    topography and source still fit)."""
    prob = PB.curvilinear_q(n=n, h=50.0 * 256 / (n - 1) / 4, nsteps=nsteps, nmech=nmech,
                            npml=npml, seed=seed, nrec=6, f0=4.0, qs_scale=50.0)
    return prob


def oracle_two_block(O, tp):
    """Oracle blocks L, R and interface of a TwoBlockProblem (metrics from
    the oracle's own pipeline)."""
    L_ = oracle_block(O, tp.left, O.metrics(tp.left.X))
    R_ = oracle_block(O, tp.right, O.metrics(tp.right.X))
    nz, ny = tp.yf.shape
    alpha = getattr(tp, "alpha", None)
    I = O.OracleIface(ny, nz, tp.yf, tp.zf, coupling=tp.coupling, s0=tp.s0,
                      W0=getattr(tp, "W0", None), alpha=1.0e14 if alpha is None else alpha)
    return L_, R_, I


TPV5_KEYS = (("S", 4), ("Svel", 3), ("Uface", 9), ("Vface", 9), ("Uhat", 6), ("Vhat", 6),
             ("trup", 1))


def tpv5_truth(name="tpv5_cart_truth.npz"):
    """The reference's archived two-block outputs (tests/golden/make_tpv5_truth.py):
    tpv5_cart_truth.npz (truth/1) or tpv28_curv_truth.npz (truth/2)."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k].astype(np.float64) for k, _ in TPV5_KEYS}
        for label in ("21_11_11", "1_11_11"):
            out["station_" + label] = z["station_" + label].astype(np.float64)
        for key in ("Hslice", "Vslice"):
            out[key] = z[key].astype(np.float64)
        return out
