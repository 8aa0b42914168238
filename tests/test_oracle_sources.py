"""Source-time functions of the oracle (source_time, moment_tensor.f90:813-872,
plus the Ricker wavelet, which the reference does not have), pinned by the
properties that define them rather than by their formulas:

* Gaussian: a normalised pulse, unit integral, peak w/sqrt(2 pi) at t0;
* cosine bell: unit integral over [0, t0], zero from t0 on;
* Ricker with peak frequency f0 and delay t0: equal to -g''/(2 pi^2 f0^2) for
  the Gaussian g = exp(-(pi f0 (t - t0))^2) (second difference of g), value 1
  at t0, zero mean, and an amplitude spectrum that peaks at f0.
"""
import math

import numpy as np

SRC = {"gaussian": 0, "cosbell": 1, "ricker": 2}


def _sample(O, kind, t, t0, w):
    return np.array([O.source_time(SRC[kind], float(x), t0, w) for x in t])


def test_gaussian_is_a_unit_pulse(oracle):
    t0, w = 0.7, 25.0
    t = np.linspace(t0 - 1.5, t0 + 1.5, 30001)
    g = _sample(oracle, "gaussian", t, t0, w)
    assert abs(np.trapezoid(g, t) - 1.0) < 1e-10
    assert abs(oracle.source_time(0, t0, t0, w) - w / math.sqrt(2 * math.pi)) < 1e-12
    assert np.argmax(g) == len(t) // 2


def test_cosine_bell_has_unit_integral_and_ends_at_t0(oracle):
    t0 = 0.4
    t = np.linspace(0.0, t0, 40001)
    b = _sample(oracle, "cosbell", t[:-1], t0, 0.0)
    assert abs(np.trapezoid(b, t[:-1]) + 0.5 * (t[1] - t[0]) * b[-1] - 1.0) < 1e-8
    assert all(oracle.source_time(1, x, t0, 0.0) == 0.0 for x in (t0, t0 + 0.1, 3.0))


def test_ricker_is_the_scaled_second_derivative_of_a_gaussian(oracle):
    f0, t0 = 2.0, 0.6
    h = 1e-4
    t = np.linspace(0.0, 1.2, 241)
    a = math.pi * f0
    g = lambda x: np.exp(-(a * (x - t0)) ** 2)
    d2 = (g(t + h) - 2 * g(t) + g(t - h)) / h ** 2
    r = _sample(oracle, "ricker", t, t0, f0)
    assert np.abs(r + d2 / (2 * a ** 2)).max() < 1e-6
    assert oracle.source_time(2, t0, t0, f0) == 1.0


def test_ricker_zero_mean_and_peak_frequency(oracle):
    f0, t0 = 2.0, 1.5
    dt = 1e-3
    t = np.arange(0.0, 3.0, dt)
    r = _sample(oracle, "ricker", t, t0, f0)
    assert abs(np.sum(r) * dt) < 1e-9
    n = 1 << 18                                   # zero-padded: 0.0038 Hz bins
    spec = np.abs(np.fft.rfft(r, n))
    freq = np.fft.rfftfreq(n, dt)
    assert abs(freq[np.argmax(spec)] - f0) < 0.01


def _mollified_weights(O, X, loc, near):
    """delta of the oracle's mollified source (set_moment_tensor_smooth) at
    every node of coordinates X (3, nz, ny, nx): DF(sxx) after one call at the
    Gaussian source-time function's peak with mXX = -1/A(peak)."""
    nz, ny, nx = X.shape[1:]
    b = O.OracleBlock(nx, ny, nz, curvilinear=1)
    b.X = np.ascontiguousarray(X)
    b.DF[...] = 0.0
    t0 = 0.5
    A = O.source_time(0, t0, t0, 10.0)            # Gaussian source-time at its peak
    b.add_source(kind=0, t_init=t0, duration=10.0, mxx=-1.0 / A, loc=loc, near=near,
                 near_phys=loc, order=0)
    b.set_moment_tensor(t0)
    return b.DF[3].copy()


def test_mollified_source_is_a_unit_gaussian_of_width_2h(oracle):
    """set_moment_tensor_smooth (moment_tensor.f90:1255-1330) on a uniform
    grid: the weights are a sampled 3-D Gaussian of standard deviation 2h, so
    their quadrature sum is 1 (to the reference's truncated pi, 1e-13), their
    first moments vanish about the source location and their second
    moments are (2h)^2 per axis; the other moment-tensor components are
    untouched."""
    n, h = 45, 50.0
    ax = np.arange(n) * h
    Z, Y, Xc = np.meshgrid(ax, ax, ax, indexing="ij")
    X = np.stack([Xc, Y, Z])
    loc = (22.3 * h, 21.6 * h, 22.9 * h)
    d = _mollified_weights(oracle, X, loc, (23, 23, 24))
    w = h ** 3
    assert abs(d.sum() * w - 1.0) < 1e-11
    for k, c in enumerate(loc):
        r = X[k] - c
        assert abs((d * r).sum() * w) < 1e-9 * h
        assert abs((d * r * r).sum() * w / (2 * h) ** 2 - 1.0) < 1e-9
    assert d.max() > 0


def test_mollified_source_takes_the_local_spacing(oracle):
    """On a grid whose x spacing doubles from one half to the other, the
    weight at each node is the uniform-grid weight of that node's own backward
    spacing (the forward one at the first node): a source in the fine half
    and one in the coarse half each integrate to 1 against their local cell."""
    n = 61
    hx = np.where(np.arange(n - 1) < 30, 40.0, 80.0)
    x = np.concatenate([[0.0], np.cumsum(hx)])
    ay = np.arange(n) * 40.0
    Z, Y, Xc = np.meshgrid(ay, ay, x, indexing="ij")
    X = np.stack([Xc, Y, Z])
    for i0, hloc in ((12, 40.0), (48, 80.0)):
        loc = (x[i0] + 0.3 * hloc, 30.2 * 40.0, 29.7 * 40.0)
        d = _mollified_weights(oracle, X, loc, (i0 + 1, 31, 30))
        assert abs(d.sum() * hloc * 40.0 * 40.0 - 1.0) < 1e-6
