"""The oracle's two-block interface (wql_oracle_iface.c) against the
reference's archived TPV5 golden files and against what the interface
conditions fix.

Golden files: test_problems/truth/1/test_rup_cart_fric_interface.* (stored
as tests/golden/tpv5_cart_truth.npz), checked by python/read_binary.py:9-37
with max |difference| <= 1e-3 over 46 frames of S, Svel, Uface, Vface, Uhat,
Vhat; and the volume slices test_rup_cart_fric.{Hslice,Vslice}.  The files predate two changes of the current source: they carry the
alpha/theta SAT form (commented out at BoundaryConditions.f90:833-841 and
905-913) and one slip-rate increment per stage (each MPI rank owns its copy of
the interface; on one rank both block evaluations of
enforce_bound_iface_conditions, domain.f90:615-644, add to the shared DS).
The oracle reproduces them with sat_form = 1 and one increment per stage.
"""
import numpy as np
import pytest

from tests.helpers import TPV5_KEYS, oracle_two_block, tpv5_truth
from waveqlab3d_q_gpu_amd import _lib as L
from waveqlab3d_q_gpu_amd import problems as PB


def check_slices(fr, truth):
    """Block 1's volume slices (the planes r = 1 and s = 1, all nine fields,
    every frame) against the archived .Hslice / .Vslice to float32 rounding
    (|values| <= 40; the nearest other planes differ by O(10))."""
    for key in ("Hslice", "Vslice"):
        ref = truth[key]
        assert fr[key].shape == ref.shape, key
        d = float(np.abs(fr[key] - ref).max())
        assert d <= 6e-8 * max(1.0, float(np.abs(ref).max())), (key, d)


def test_tpv5_cartesian_golden_files(oracle):
    tp = PB.tpv5_cartesian(21, sat_form=1)
    assert tp.nsteps == 46
    Lb, Rb, I = oracle_two_block(oracle, tp)
    _, fr = oracle.two_block_time_step(Lb, Rb, I, tp.dt, tp.nsteps, 0.0, slices=True)
    truth = tpv5_truth()
    check_slices(fr, truth)
    for key, nc in TPV5_KEYS:
        mine, ref = fr[key], truth[key]
        assert mine.shape == ref.shape, key
        d = float(np.abs(mine - ref).max())
        # the reference's own tolerance (read_binary.py:37) ...
        assert d <= 1.0e-3, (key, d)
        # ... and float32 rounding of the stored values (|values| <= 120)
        scale = max(1.0, float(np.abs(ref[ref < 1e8]).max()))
        assert d <= 4e-7 * scale + 1e-7 or key == "trup", (key, d)
    # the rupture front reached the free surface and the nucleation patch
    assert np.isfinite(fr["trup"]).all() and (fr["trup"][-1] < 1e9).any()


def test_tpv5_current_sat_form_differs(oracle):
    """The golden files pin the alpha/theta form: the current BCm/BCp form
    differs from them by far more than the tolerance (so the comparison above
    would catch the wrong form)."""
    tp = PB.tpv5_cartesian(21, sat_form=0)
    Lb, Rb, I = oracle_two_block(oracle, tp)
    _, fr = oracle.two_block_time_step(Lb, Rb, I, tp.dt, 10, 0.0)
    truth = tpv5_truth()
    assert float(np.abs(fr["Uface"] - truth["Uface"][:10]).max()) > 1e-2


def _uniform_pair(oracle, coupling, jump=0.0, seed=3):
    """L and R with different materials, fields constant in space: velocity
    and traction n.sigma (n = x) continuous across the interface; the other
    stress components differ."""
    rng = np.random.default_rng(seed)
    tp = PB.tpv5_cartesian(16, sat_form=0, coupling=coupling)
    tp.right.M = PB.material_from_rho_s_p(tp.right.n, 2.2, 2.9, 5.1)
    Lb, Rb, I = oracle_two_block(oracle, tp)
    v = rng.normal(size=3)
    sxx, sxy, sxz = rng.normal(size=3)
    for b, other in ((Lb, rng.normal(size=3)), (Rb, rng.normal(size=3))):
        vals = [v[0], v[1], v[2], sxx, other[0], other[1], sxy, sxz, other[2]]
        for c in range(9):
            b.F[c] = vals[c]
    Rb.F[1] += jump
    for b in (Lb, Rb):
        b.DF[...] = 0.0
    I.DS[...] = 0.0
    return Lb, Rb, I


def test_locked_sat_vanishes_when_interface_conditions_hold(oracle):
    Lb, Rb, I = _uniform_pair(oracle, L.WQL_IFACE_LOCKED)
    I.enforce(Lb, Rb, 0.0, 1, 2)
    I.enforce(Lb, Rb, 0.0, 1, 1)
    scale = float(np.abs(Lb.F).max()) * Lb.tau0 / Lb.hq * 10.0
    assert float(np.abs(Lb.DF).max()) <= 1e-13 * scale
    assert float(np.abs(Rb.DF).max()) <= 1e-13 * scale
    # no slip rate either
    assert float(np.abs(I.DS[:3]).max()) <= 1e-14


def test_locked_sat_opposes_a_velocity_jump(oracle):
    """A tangential velocity jump [v_y] = V_R - V_L > 0: the penalty on the
    velocity rate pushes v_y of L up and v_y of R down (DF = -tau0/h * F
    with F of opposite signs), and the hats see slip rate |jump|."""
    Lb, Rb, I = _uniform_pair(oracle, L.WQL_IFACE_LOCKED, jump=0.5)
    I.enforce(Lb, Rb, 0.0, 1, 2)
    I.enforce(Lb, Rb, 0.0, 1, 1)
    fL = Lb.DF[1][:, :, -1]
    fR = Rb.DF[1][:, :, 0]
    assert (fL > 0).all() and (fR < 0).all()
    # welded hats are continuous: no slip rate even with a jump in the grid values
    assert float(np.abs(I.DS[:3]).max()) <= 1e-14


def test_locked_interface_kat_41(oracle):
    """inputfile/test_elastic_locked_interface_o6.in: the reference prints
    'block 1 max field: 2.0699E+03', 'block 2 max field: 1.6389E+03'
    (PHASE2_SCRATCH_MEMORY_AUDIT.md:222-223; close_domain, domain.f90:583-586)
    after 103 steps of two 41^3 blocks with a locked interface and a Cos-Bell
    source on block 1's interface face."""
    tp = PB.locked_interface_41()
    assert tp.nsteps == 103 and tuple(tp.left.sources[0].near) == (41, 21, 21)
    Lb, Rb, I = oracle_two_block(oracle, tp)
    oracle.two_block_time_step(Lb, Rb, I, tp.dt, tp.nsteps, 0.0)
    assert "%.4E" % np.abs(Lb.F).max() == "2.0699E+03"
    assert "%.4E" % np.abs(Rb.F).max() == "1.6389E+03"


def test_tpv5_station_files(oracle):
    """test_problems/truth/1/test_rup_cart_fric_{21_11_11,1_11_11}.dat (t, vx,
    vy, vz with 6 decimals) are the particle velocities of the fault nodes
    (21, 11, 11) of block 1 and (1, 11, 11) of block 2: the oracle's face
    fields at y = z = 11 reproduce them to the printed digits."""
    tp = PB.tpv5_cartesian(21, sat_form=1)
    Lb, Rb, I = oracle_two_block(oracle, tp)
    _, fr = oracle.two_block_time_step(Lb, Rb, I, tp.dt, tp.nsteps, 0.0)
    truth = tpv5_truth()
    t = np.arange(tp.nsteps) * tp.dt
    for key, label in (("Uface", "21_11_11"), ("Vface", "1_11_11")):
        st = truth["station_" + label]
        assert np.abs(st[:, 0] - t).max() <= 6e-7
        v = fr[key][:, 0:3, 10, 10]
        assert np.abs(v - st[:, 1:4]).max() <= 6e-7, (label, np.abs(v - st[:, 1:4]).max())


# TPV28 initial stress of the archived golden files test_problems/truth/2
# (xx, yy, zz, xy, xz, yz): zz = -60 where the current source has +60
# (initial_stress_condition.f90:78-84)
TPV28_TRUTH_STRESS = (-60.0, 0.0, -60.0, 0.0, 29.38, 0.0)


def test_tpv28_curvilinear_golden_files(oracle):
    """test_problems/truth/2 (test_rup_curv_fric.in): TPV28 on two curvilinear
    21^3 blocks meshed by Curve_Grid_3D3 ('analytical_test_problem'), with the
    same historical choices as truth/1 (alpha/theta SAT, one slip-rate
    increment per stage) and the initial stress with zz = -60.  All seven
    outputs match to float32 rounding; the station files to their digits."""
    tp = PB.tpv28_curvilinear(21, sat_form=1, s0=TPV28_TRUTH_STRESS)
    assert tp.nsteps == 46
    Lb, Rb, I = oracle_two_block(oracle, tp)
    _, fr = oracle.two_block_time_step(Lb, Rb, I, tp.dt, tp.nsteps, 0.0, slices=True)
    truth = tpv5_truth("tpv28_curv_truth.npz")
    check_slices(fr, truth)
    for key, _ in TPV5_KEYS:
        d = float(np.abs(fr[key] - truth[key]).max())
        assert d <= 1.0e-3, (key, d)
        scale = max(1.0, float(np.abs(truth[key][truth[key] < 1e8]).max()))
        assert d <= 4e-7 * scale + 1e-7 or key == "trup", (key, d)
    for key, label in (("Uface", "21_11_11"), ("Vface", "1_11_11")):
        st = truth["station_" + label]
        assert np.abs(fr[key][:, 0:3, 10, 10] - st[:, 1:4]).max() <= 6e-7


def test_tpv28_source_stress_differs_from_golden_files(oracle):
    """With the source's initial stress (zz = +60) the prestress at the first
    frame already differs from the files by far more than their tolerance."""
    tp = PB.tpv28_curvilinear(21, sat_form=1)
    Lb, Rb, I = oracle_two_block(oracle, tp)
    _, fr = oracle.two_block_time_step(Lb, Rb, I, tp.dt, 1, 0.0)
    truth = tpv5_truth("tpv28_curv_truth.npz")
    assert float(np.abs(fr["Uhat"][0] - truth["Uhat"][0]).max()) > 1.0


def test_transfinite_grid_reduces_to_cartesian():
    """Curve_Grid_3D3 with flat x faces (lc = rc = 0) is the uniform grid, and
    the two TPV28 blocks share their fault nodes (to rounding)."""
    a, b = (0.0, 6.0, -1.5), (3.0, 9.0, 1.5)
    flat = PB.curve_grid_analytical_test_problem((21, 17, 13), a, b, 0.0, 0.0)
    assert np.abs(flat - PB.cartesian_grid((21, 17, 13), a, b)).max() <= 1e-13
    tp = PB.tpv28_curvilinear(21)
    assert np.abs(tp.left.X[:, :, :, -1] - tp.right.X[:, :, :, 0]).max() <= 1e-14
    assert np.abs(tp.left.X[0][:, :, -1]).max() > 0.05      # the fault face is curved


# ---------------------------------------------------------------------------
# rate-and-state friction (TPV101 / TPV102)
# ---------------------------------------------------------------------------
def _rs_root(Phi, eta, sigma, psi, V0, a):
    """The slip rate of V + a sigma/eta asinh(V/(2 V0) e^(psi/a)) = Phi/eta by
    bisection on [0, Phi/eta] (the function increases in V), 200 halvings."""
    def f(V):
        return V + a * sigma / eta * np.arcsinh(0.5 * V / V0 * np.exp(psi / a)) - Phi / eta
    lo, hi = 0.0, Phi / eta
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def test_regula_falsi_slip_rate_is_the_root(oracle):
    """Regula_Falsi (Interface_Condition.f90:1170-1239) returns the root of
    f (:1241-1265) within the accuracy its stopping rule allows, from any
    guess, for slip rates from 1e-12 to 10 m/s: it stops when consecutive
    iterates differ by less than 1e-12 m/s, which on its one-sided
    convergence leaves up to ~1e-10 m/s (measured 9e-11) and, above 1e-6 m/s,
    at most 1e-4 relative (measured 8e-5)."""
    rng = np.random.default_rng(101)
    worst = 0.0
    for _ in range(200):
        sigma = rng.uniform(20.0, 150.0)
        a = rng.uniform(0.006, 0.02)
        eta = rng.uniform(2.0, 6.0)
        psi = rng.uniform(0.3, 0.9)
        Vt = 10.0 ** rng.uniform(-12, 1)
        # a stress transfer that makes Vt the exact root
        Phi = eta * Vt + sigma * a * np.arcsinh(0.5 * Vt / 1e-6 * np.exp(psi / a))
        V = oracle.rs_regula_falsi(rng.uniform(0.0, Phi / eta), Phi, eta, sigma, psi, 1e-6, a)
        ref = _rs_root(Phi, eta, sigma, psi, 1e-6, a)
        assert 0.0 <= V <= Phi / eta
        assert abs(ref - Vt) <= 1e-9 * Vt + 1e-20
        worst = max(worst, abs(V - Vt) / (1e-4 * Vt + 1e-10))
    assert worst <= 1.0, worst


def test_rate_state_steady_state(oracle):
    """The state rate of TPV101 / TPV102 (:431-436) vanishes at psi_ss = f0 -
    b ln(V/V0), is negative above it and positive below, and at psi_ss the
    friction a asinh(V/(2 V0) e^(psi/a)) is the steady-state law f0 + (a - b)
    ln(V/V0) (velocity weakening for a < b); the slip rate solved at the
    steady-state stress is V itself."""
    L0, f0, b, V0 = 0.02, 0.6, 0.012, 1e-6
    for a in (0.008, 0.016):
        for V in (1e-9, 1e-6, 1e-3, 0.1, 1.0, 5.0):
            psi = f0 - b * np.log(V / V0)
            rate = oracle.rs_state_rate(psi, V, L0, f0, b, V0)
            assert abs(rate) <= 1e-12 * b * V / L0 + 1e-300
            assert oracle.rs_state_rate(psi + 0.01, V, L0, f0, b, V0) < 0
            assert oracle.rs_state_rate(psi - 0.01, V, L0, f0, b, V0) > 0
            mu = a * np.arcsinh(0.5 * V / V0 * np.exp(psi / a))
            assert abs(mu - (f0 + (a - b) * np.log(V / V0))) <= 1e-9
            sigma, eta = 120.0, 5.0
            Phi = eta * V + sigma * mu
            got = oracle.rs_regula_falsi(0.5 * Phi / eta, Phi, eta, sigma, psi, V0, a)
            assert abs(got - V) <= 1e-4 * V + 1e-10, (a, V, got)   # the solver's floor, above


def test_rate_state_parameter_a():
    """a(y, z) of TPV101 / TPV102: 0.008 inside the 15 km box, 0.016 beyond
    the 3 km smooth_boxcar layer, monotone across it."""
    from oracle import oracle as O
    assert O.rs_a(7.5, 0.0) == 0.008 and O.rs_a(-15.0, 15.0) == 0.008
    assert O.rs_a(18.0, 0.0) == 0.016 and O.rs_a(0.0, -18.5) == 0.016
    vals = [O.rs_a(15.0 + 3.0 * k / 50.0, 0.0) for k in range(1, 50)]
    assert all(0.008 <= v <= 0.016 for v in vals)
    assert sum(0.009 < v < 0.015 for v in vals) >= 10
    assert all(v2 >= v1 for v1, v2 in zip(vals, vals[1:]))


def test_rate_state_initial_state_is_steady_sliding(oracle):
    """The initial state of the TPV101 set-up (Init_Vel_State,
    Interface_Condition.f90:2757-2787) makes the prestress the steady-state
    friction at Vin = 1e-12 m/s: sigma_n a asinh(Vin/(2 V0) e^(psi/a)) = T0_l."""
    tp = PB.tpv101_cartesian(21)
    a = np.vectorize(oracle.rs_a)(tp.yf, tp.zf)
    sigma, T0l = 120.0, 75.0
    fric = sigma * a * np.arcsinh(0.5 * 1e-12 / 1e-6 * np.exp(tp.W0 / a))
    assert np.allclose(fric, T0l, rtol=1e-10)


def test_tpv101_nucleates_and_ruptures(oracle):
    """TPV101 on the two 21^3 blocks for 1.5 s: the slip rate starts at Vin
    (steady sliding), the nucleation stress ramps in, the rupture grows from
    the nucleation zone (slip rate > 1 m/s, rupture times increasing with
    distance from (7.5, 0))."""
    from helpers import oracle_two_block
    tp = PB.tpv101_cartesian(21)
    Lo, Ro, Io = oracle_two_block(oracle, tp)
    _, fr = oracle.two_block_time_step(Lo, Ro, Io, tp.dt, tp.nsteps, 0.0)
    sv = np.sqrt((fr["Svel"][:, 1:] ** 2).sum(axis=1)).max(axis=(1, 2))
    assert sv[0] < 1e-10 and sv[-1] > 1.0
    # the .state output holds psi at the start of each step: the steady
    # state first, then lower where the fault slips fast (state evolution
    # towards f0 - b ln(V/V0))
    st = fr["state"][:, 0]
    assert np.array_equal(st[0], tp.W0.reshape(st.shape[1:]))
    assert st[-1].min() < st[0].min() - 0.05
    tr = fr["trup"][-1, 0]
    ok = tr < 1e8
    assert ok.sum() > 100
    r = np.sqrt((tp.yf - 7.5) ** 2 + tp.zf ** 2)
    # later rupture farther out: rank correlation of rupture time with distance
    from scipy.stats import spearmanr
    rho = spearmanr(r[ok], tr[ok]).correlation
    assert rho > 0.8, rho


def test_linear_friction_limits(oracle):
    """Linear_Friction (Interface_Condition.f90:639-742): with tau = alpha v
    the hats tend to the welded interface's as alpha grows (the reference's
    alpha = 1e14 is its locked limit, :360-366), and alpha = 0 is a
    frictionless interface (zero shear traction hats) with the normal
    components still welded.  For any alpha the hats keep both outgoing
    characteristics (p = Z1 u - T on the minus side, q = Z2 v + T on the
    plus side) and the shear traction is alpha times the slip rate."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        U, V = rng.normal(size=6), rng.normal(size=6)
        Z1s, Z2s = rng.uniform(3, 10, 2)
        Z1p, Z2p = Z1s * 1.7, Z2s * 1.8
        Pw, Qw = oracle.welded_hats(U, V, Z1s, Z2s, Z1p, Z2p)
        Pl, Ql = oracle.linear_friction_hats(U, V, Z1s, Z2s, Z1p, Z2p, 1.0e14)
        assert np.allclose(Pl, Pw, rtol=0, atol=1e-12) and np.allclose(Ql, Qw, rtol=0, atol=1e-12)
        P0, Q0 = oracle.linear_friction_hats(U, V, Z1s, Z2s, Z1p, Z2p, 0.0)
        assert np.all(P0[4:] == 0.0) and np.all(Q0[4:] == 0.0)
        assert P0[0] == Pw[0] and Q0[0] == Qw[0] and P0[3] == Pw[3]
        for alpha in (0.0, 0.7, 5.0, 300.0):
            P, Q = oracle.linear_friction_hats(U, V, Z1s, Z2s, Z1p, Z2p, alpha)
            Zs1, Zs2 = (Z1p, Z1s, Z1s), (Z2p, Z2s, Z2s)
            for c in range(3):
                assert abs((Zs1[c] * P[c] - P[3 + c]) - (Zs1[c] * U[c] - U[3 + c])) <= 1e-12 * 20
                assert abs((Zs2[c] * Q[c] + Q[3 + c]) - (Zs2[c] * V[c] + V[3 + c])) <= 1e-12 * 20
            for c in (1, 2):     # tau = alpha (v_plus - v_minus), same traction both sides
                assert abs(P[3 + c] - Q[3 + c]) <= 1e-14
                assert abs(P[3 + c] - alpha * (Q[c] - P[c])) <= 1e-10 * max(1.0, alpha)


def test_rate_weakening_state_law(oracle):
    """The state rate of problem 'rate-weakening' (:438-444) drives psi to
    the steady state of its flash-weakened friction: it vanishes where
    a asinh(V/(2 V0) e^(psi/a)) equals Fss(V), is negative above and positive
    below; Fss is the classical f0 - (b - a) ln(V/V0) far below V_w (at V = V0
    it is f0 to ~1e-43) and weakens towards 0.13 far above V_w."""
    f0, b, V0 = 0.7, 0.02, 1e-6
    for a, Vw in ((0.016, 0.17), (0.032, 5.17)):
        for V, Fss_expect in ((V0, f0), (1e-9, f0 - (b - a) * np.log(1e-9 / V0)),
                              (1e4 * Vw, None)):
            if Fss_expect is None:
                # V = 1e4 V_w: Fss = 0.13 + (f - 0.13) / (1 + 1e32)^(1/8) = 0.13 + ~1e-4 (f - 0.13)
                Fss_expect = 0.13 + (f0 - (b - a) * np.log(V / V0) - 0.13) * 1e-4
            psi = a * np.log(2.0 * V0 / V * np.sinh(Fss_expect / a))
            r = oracle.rw_state_rate(psi, V, a, Vw)
            scale = V / (2.0 * 0.2572) * a
            assert abs(r) <= 1e-6 * scale, (a, V, r)
            assert oracle.rw_state_rate(psi + 0.01, V, a, Vw) < 0
            assert oracle.rw_state_rate(psi - 0.01, V, a, Vw) > 0


def test_rate_weakening_parameters_and_prestress(oracle):
    """a(y, z) 0.016 / 0.032 and V_w 0.17 / 5.17 m/s inside / outside the
    20 km box (smooth_boxcar width 2.5); the prestress is the traction of a
    stress whose principal axis lies at psi = 50 degrees from x in the x-z
    plane (no shear traction on that plane), -126 MPa normal and 42 MPa
    shear on the fault plane x = const, plus the pulse 2 (0.2429 * 126) F(r)
    on T0_l from t > 0, F = 1 at (8, 0), exp(-1/2) at r = 1.35 km."""
    assert oracle.rw_a(8.0, 0.0) == 0.016 and oracle.rw_a(23.0, 0.0) == 0.032
    assert oracle.rw_vw(8.0, 0.0) == 0.17 and oracle.rw_vw(0.0, -23.0) == 5.17
    n, m, l = (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)
    T0 = oracle.rw_prestress(n, m, l, 2.0, 0.0, 0.0)      # t = 0: no pulse
    assert T0[0] == -126.0 and T0[1] == 0.0 and abs(T0[2] - 42.0) <= 1e-13
    psi = 50.0 / 180.0 * 3.14159265359
    nn = (np.cos(psi), 0.0, np.sin(psi))
    mm = (0.0, 1.0, 0.0)
    ll = (-np.sin(psi), 0.0, np.cos(psi))
    Tp = oracle.rw_prestress(nn, mm, ll, 2.0, 0.0, 0.0)
    assert abs(Tp[1]) <= 1e-12 and abs(Tp[2]) <= 1e-12
    pulse = oracle.rw_prestress(n, m, l, 8.0, 0.0, 1e-3)[2] - T0[2]
    assert abs(pulse - 2.0 * 0.2429 * 126.0) <= 1e-12
    pulse_r0 = oracle.rw_prestress(n, m, l, 8.0, 1.35, 1e-3)[2] - T0[2]
    assert abs(pulse_r0 - 2.0 * 0.2429 * 126.0 * np.exp(-0.5)) <= 1e-12


def test_rate_weakening_ruptures(oracle):
    """Problem 'rate-weakening' on two 21^3 blocks: from psi = 0.4367 (far
    below the steady state of its prestress) the fault slides at once; the
    slip rate peaks at several m/s, every fault node ruptures and the state
    evolves downwards."""
    from helpers import oracle_two_block
    tp = PB.rate_weakening_cartesian(21)
    Lo, Ro, Io = oracle_two_block(oracle, tp)
    _, fr = oracle.two_block_time_step(Lo, Ro, Io, tp.dt, tp.nsteps, 0.0)
    assert np.isfinite(Lo.F).all() and np.isfinite(Ro.F).all()
    sv = np.sqrt((fr["Svel"][:, 1:] ** 2).sum(axis=1)).max()
    assert sv > 1.0
    assert (fr["trup"][-1, 0] < 1e8).all()
    assert Io.W.max() < 0.4367
