"""Realized quality factor of the standard-linear-solid memory variables
(shared by the realized-Q tests).

With strain rate e'(t), a memory variable obeys tau_i eta_i' = w_i Q^-1 M e'
- eta_i and the stress sigma' = M e' - sum_i eta_i (RHS_Interior.f90:
2748-2906, the relaxation equations of the anelastic Q4/Q8 point kernels).
For e = e0 exp(i w t) the steady state is sigma = M(w) e with
M(w)/M = 1 - sum_i (w_i/Q) / (1 + i w tau_i), whose ratio of real to
imaginary part is the realized Q (the frequency-domain modulus of
anelastic_q8_model.f90:119-145)."""
import numpy as np


def complex_modulus(f, tau, weight, Q):
    """M(w)/M for frequencies f (Hz)."""
    w = 2.0 * np.pi * np.asarray(f, dtype=np.float64)[..., None]
    x = w * np.asarray(tau)[None, :]
    return 1.0 - np.sum((np.asarray(weight) / Q) / (1.0 + 1j * x), axis=-1)


def analytic_q(f, tau, weight, Q):
    m = complex_modulus(f, tau, weight, Q)
    return np.abs(m.real / m.imag)


def fit_harmonic(t, y, f):
    """Least-squares a + b t + c cos(wt) + d sin(wt) over samples (t, y):
    returns the complex amplitude c - i d (y ~ Re[(c - i d) e^{iwt}])."""
    w = 2.0 * np.pi * f
    A = np.stack([np.ones_like(t), t, np.cos(w * t), np.sin(w * t)], axis=1)
    coef, *_ = np.linalg.lstsq(A, y, rcond=None)
    return coef[2] - 1j * coef[3]


def uniform_problem(nmech, n=12, qs=50.0, qp=100.0):
    """A uniform block (rho 2700, cs 3000, cp = sqrt(3) cs), anelastic Q4/Q8
    with constant Qs, Qp and no boundary penalty on any face (bc 0, no
    interface): a velocity field linear in x has the same gradient at every
    node and uniform stress, so the velocity keeps the value it is given and
    every node is the constitutive point model driven by that gradient."""
    import math

    from waveqlab3d_q_gpu_amd import problems as PB
    nn = (n, n, n)
    # 1000 km across: the CFL limit of the stencil (about 7 s) stays far
    # above the drive's time steps, so round-off never grows
    X = PB.cartesian_grid(nn, (0.0,) * 3, (1.0e6,) * 3)
    M = PB.material_from_rho_s_p(nn, 2700.0, 3000.0, math.sqrt(3.0) * 3000.0)
    tau, w = PB.q_coefficients(nmech)
    shp = X.shape[1:]
    return PB.Problem(n=nn, mesh=0, X=X, M=M, bc=(0,) * 6, dt=0.0, nsteps=0, nmech=nmech,
                      tau=tau, weight=w, Qs_inv=np.full(shp, 1.0 / qs),
                      Qp_inv=np.full(shp, 1.0 / qp), name=f"uniform_q{nmech}_{n}")


def schedule(f, tau):
    """(dt, nsteps) for two periods at frequency f: at most T/400 and the
    shortest relaxation time."""
    T = 1.0 / f
    dt = min(float(np.min(tau)), T / 400.0)
    return dt, int(np.ceil(2.0 * T / dt))


def drive(step, f, tau, e0=1.0e-6):
    """Harmonic drive at f Hz for two periods: before each step the velocity
    is set to v_x = v_y = a_n x with a_n = e0 w cos(w (t_n + dt/2)), i.e. the
    strain rates e_xx = g_xy = a_n are held over the step.  ``step(a, dt,
    t)`` sets the velocity, advances one RK step and returns the stress
    (sxx, syy, sxy) of one node.  Returns t, strain, stresses at the end of
    each step."""
    dt, nsteps = schedule(f, tau)
    w = 2.0 * np.pi * f
    t = 0.0
    e = 0.0
    ts, es, ss = [], [], []
    for n in range(nsteps):
        a = e0 * w * np.cos(w * (t + 0.5 * dt))
        s = step(a, dt, t)
        t = (n + 1) * dt
        e += a * dt
        ts.append(t); es.append(e); ss.append(s)
    return np.array(ts), np.array(es), np.array(ss)


def measured_moduli(ts, es, ss, f, tau):
    """Complex moduli M_xx, M_yy, M_xy = sigma / strain at f from the driven
    run: sigma and the strain are fitted by a constant, a trend, the harmonic
    at f and the free relaxations exp(-t/tau_i) of the memory variables
    (the general solution of the forced relaxation equations), and the
    harmonic amplitudes divided."""
    w = 2.0 * np.pi * f
    basis = [np.ones_like(ts), ts, np.cos(w * ts), np.sin(w * ts)]
    basis += [np.exp(-ts / ti) for ti in tau]
    A = np.stack(basis, axis=1)
    scale = np.linalg.norm(A, axis=0)
    An = A / scale

    def amp(y):
        coef, *_ = np.linalg.lstsq(An, y, rcond=None)
        coef = coef / scale
        return coef[2] - 1j * coef[3]

    ce = amp(es)
    return [amp(ss[:, k]) / ce for k in range(3)]


def analytic_moduli(f, prob):
    """M_xx, M_yy, M_xy at f for the uniaxial strain e_xx with shear g_xy
    (the frequency-domain form of the relaxation equations)."""
    lam, mu = prob.M[0].flat[0], prob.M[1].flat[0]
    Qs, Qp = prob.Qs_inv.flat[0], prob.Qp_inv.flat[0]
    x = 2.0 * np.pi * f * np.asarray(prob.tau)
    r = np.asarray(prob.weight) / (1.0 + 1j * x)
    mxx = (lam + 2.0 * mu) * (1.0 - Qp * r.sum())
    myy = lam - ((lam + 2.0 * mu) * Qp - 2.0 * mu * Qs) * r.sum()
    mxy = mu * (1.0 - Qs * r.sum())
    return mxx, myy, mxy


FREQS = 0.08 * (15.0 / 0.08) ** (np.arange(12) / 11.0)


def oracle_stepper(O, prob):
    """step() of :func:`drive` on the oracle (node (1, 1, 1) sampled)."""
    from tests.helpers import oracle_block
    b = oracle_block(O, prob, O.metrics(prob.X))
    X = prob.X

    def step(a, dt, t):
        b.F[0] = a * X[0]
        b.F[1] = a * X[0]
        b.time_step(dt, 1, t)
        return b.F[3, 0, 0, 0], b.F[4, 0, 0, 0], b.F[6, 0, 0, 0]
    return step
