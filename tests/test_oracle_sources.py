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
