"""The KH growth-rate fit of cueva.diagnostics on amplitude series whose
rate is known (no GPU): a pure exponential, an exponential after a flat
start, and a series that never grows."""
import numpy as np

from cueva import diagnostics


def test_rate_of_a_pure_exponential():
    t = np.linspace(0.0, 2.0, 41)
    a = 0.01 * np.exp(1.7 * t)
    win = diagnostics.growth_window(a)
    assert win is not None
    assert abs(diagnostics.growth_rate(t, a, win) - 1.7) < 1e-10
    # the window starts at the first sample above 1.5 A(0)
    assert a[win[0]] >= 1.5 * a[0] > a[win[0] - 1]


def test_window_skips_the_transient():
    t = np.linspace(0.0, 1.0, 101)
    a = np.where(t < 0.5, 0.01, 0.01 * np.exp(4.0 * (t - 0.5)))
    win = diagnostics.growth_window(a)
    assert t[win[0]] > 0.5
    assert abs(diagnostics.growth_rate(t, a, win) - 4.0) < 1e-10


def test_no_growth_gives_no_window():
    a = np.full(30, 0.01) * (1.0 + 0.01 * np.sin(np.arange(30)))
    assert diagnostics.growth_window(a) is None


def test_rate_from_a_quadratic_integral():
    # vy = A(t) f(x, y) gives an integral of vy^2 growing as A^2: the
    # amplitude rate is half its log slope, as fit_growth.gnu halves its fit
    t = np.linspace(0.0, 2.0, 41)
    amp = 0.01 * np.exp(1.3 * t)
    x = np.linspace(-0.5, 0.5, 64, endpoint=False)
    shape = np.sin(2 * np.pi * x) * np.exp(-x * x / 0.01)
    ints = np.array([np.sum((a * shape) ** 2) / 64 for a in amp])
    win = diagnostics.growth_window(amp)
    assert abs(diagnostics.growth_rate_of_integral(t, ints, win) - 1.3) < 1e-10
    assert abs(diagnostics.growth_rate(t, ints, win) - 2.6) < 1e-10
