"""The Strouhal estimator (imexlb.accuracy.strouhal) on synthetic lift
signals of known frequency: interpolated zero crossings resolve the period
far below the sampling step, and the reported standard error covers the
truth when the period jitters.  (The shedding log extends Monitor,
lbm.f90:205-227, which logs only the maximum speed.)"""
import numpy as np
import pytest

from imexlb.accuracy import strouhal, upward_crossings

D, U = 25.6, 0.05  # diameter and inlet speed of the 400x100 cylinder deck


def _signal(period, steps, every, phase=0.3, offset=0.0, harmonic=0.0, noise=0.0, seed=0):
    rng = np.random.default_rng(seed)
    t = np.arange(0, steps, every, dtype=float)
    ph = 2 * np.pi * t / period + phase
    x = np.sin(ph) + harmonic * np.sin(2 * ph + 1.1) + offset
    return x + noise * rng.standard_normal(t.size)


def test_crossings_of_a_line_are_exact():
    x = np.array([-3.0, -1.0, 1.0, 3.0, 1.0, -1.0, -3.0, -1.0, 1.0])
    # mean is -0.444..; crossings of x - mean between samples 1-2 and 7-8
    t = upward_crossings(x, 10)
    m = x.mean()
    assert np.allclose(t, [(1 + (m + 1) / 2) * 10, (7 + (m + 1) / 2) * 10], rtol=0, atol=1e-12)


@pytest.mark.parametrize("every", [1, 10, 100])
def test_pure_sinusoid(every):
    period = 2437.3
    st_true = D / (U * period)
    st, err, n = strouhal(_signal(period, 50000, every), every, D, U)
    assert n >= 9
    # the old FFT-bin estimate had a resolution of ~0.1 here; this one is
    # exact to the interpolation error of a sine between two samples
    tol = {1: 1e-9, 10: 1e-7, 100: 2e-4}[every]
    assert abs(st - st_true) <= tol * st_true, (st, st_true)
    assert err <= tol * st_true


def test_offset_harmonic_and_noise():
    """A drag-like offset, a second harmonic and noise do not move the
    estimate beyond its error bar."""
    period = 1843.0
    st_true = D / (U * period)
    st, err, n = strouhal(_signal(period, 60000, 10, offset=0.7, harmonic=0.25, noise=0.01), 10, D, U)
    assert n >= 15
    assert abs(st - st_true) <= 4 * err + 1e-6 * st_true, (st, st_true, err)
    assert abs(st - st_true) <= 1e-3 * st_true


def _cycles(period, scatter, ncycle, every, seed):
    """A lift signal whose successive cycles last period*(1 + scatter*N(0,1))
    steps each (sampled every `every` steps), and the mean of those periods
    over the second half of the record."""
    rng = np.random.default_rng(seed)
    lengths = period * (1 + scatter * rng.standard_normal(ncycle))
    edges = np.concatenate([[0.0], np.cumsum(lengths)])
    t = np.arange(0, edges[-1], every, dtype=float)
    k = np.searchsorted(edges, t, side="right") - 1
    x = np.sin(2 * np.pi * (t - edges[k]) / lengths[k])
    return x, lengths


def test_period_scatter_is_covered_by_the_error():
    """Cycle-to-cycle scatter of 5 %: the estimate stays within three
    standard errors of the mean period of the distribution in at least 18
    of 20 records, and its error bar has the size of that scatter over
    the square root of the number of periods."""
    period = 2000.0
    st_true = D / (U * period)
    hits = 0
    for seed in range(20):
        x, _ = _cycles(period, 0.05, 60, 10, seed)
        st, err, n = strouhal(x, 10, D, U)
        assert 25 <= n <= 31
        assert 0.5 * 0.05 / np.sqrt(n) <= err / st <= 2.0 * 0.05 / np.sqrt(n)
        hits += abs(st - st_true) <= 3 * err
    assert hits >= 18


def test_too_short_or_diverged_gives_none():
    assert strouhal(_signal(5000.0, 12000, 10), 10, D, U)[0] is None  # one period in the second half
    x = _signal(500.0, 20000, 10)
    x[-5] = np.nan
    assert strouhal(x, 10, D, U) == (None, None, 0)


def test_noise_near_zero_adds_no_cycles():
    """Rounding-like noise of a tenth of the amplitude on every sample makes
    plain sign changes many times per period; the hysteresis keeps one
    crossing per period and the estimate within 0.2 %."""
    period = 2500.0
    st_true = D / (U * period)
    x = _signal(period, 60000, 10, noise=0.1, seed=3)
    xm = x - x.mean()
    plain = np.count_nonzero((xm[:-1] < 0) & (xm[1:] >= 0))
    assert plain > 3 * 60000 / period  # the noise alone crosses zero far more often
    t = upward_crossings(x, 10)
    assert abs(len(t) - 60000 / period) <= 1
    st, err, n = strouhal(x, 10, D, U)
    assert abs(st - st_true) <= 2e-3 * st_true and abs(st - st_true) <= 4 * err
