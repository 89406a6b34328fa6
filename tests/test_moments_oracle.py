"""Pins of the oracle's s2dw_stat_moments (oracle/s2dw_oracle.c,
orc_stat_moments; reference src/mod/s2dw_stat_mod.f90:81-275), CPU only.

- against the reference's own routine, compiled from its Fortran sources into
  oracle/_ref/ref_dump, through tests/golden/ref_moments.npz: plain, binary
  mask, fractional mask (nj = nint(sum mask) differs from the count, and a
  mask of 1 + 5e-13 counts as 1), a zero-variance scale, a large mean
  (1000 + exponential) and a masked skewed case;
- against scipy.stats' central moments (an independent implementation);
- against exact rational arithmetic on a discrete uniform sample.

Tolerance: 1e-12 relative to max(|value|, 1) (the mean: relative to
max(|mean|, standard deviation); the excess kurtosis and sixth moment:
relative to the standardised moment before the -3 and -15), times the
condition number max(1, |mean| / standard deviation) for the variance and
higher moments: a rounding of the mean shifts every centred value, so FP64
sums in different orders agree only to that.
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.stats

import oracle as O

RTOL = 1e-12
CASES = ["plain", "mask", "frac", "const", "offset", "gamma"]


def split(x, B, N, alpha, J):
    his, _ = O.band_limits(J, B, alpha)
    out, o = [], 0
    for h in his:
        n = (2 * h - 1) * 2 * h * N
        out.append(np.asfortranarray(x[o:o + n].reshape((2 * h - 1, 2 * h, N), order="F")))
        o += n
    assert o == len(x)
    return out


EXCESS = [0.0, 0.0, 0.0, 3.0, 0.0, 15.0]  # added back to kur and six before scaling


def assert_moments(got, want, label="", rtol=RTOL):
    mean, var = want[0], want[1]
    sd = np.sqrt(np.maximum(var, 0))
    scale_mean = np.maximum(np.abs(mean), sd)
    assert np.all(np.abs(got[0] - mean) <= rtol * np.maximum(scale_mean, 1e-300)), (label, got[0], mean)
    cond = np.maximum(1.0, np.abs(mean) / np.maximum(sd, 1e-300))
    for k in range(1, 6):
        tol = rtol * cond * np.maximum(np.abs(want[k] + EXCESS[k]), 1.0)
        assert np.all(np.abs(got[k] - want[k]) <= tol), (label, k, got[k], want[k], np.abs(got[k] - want[k]))


def fixture_case(golden, name):
    d = golden("ref_moments.npz")
    B, N, J = (int(v) for v in d[name + "_meta"])
    alpha = float(d[name + "_alpha"][0])
    x = d[name + "_x"].copy()
    m = d[name + "_mask"].copy() if name + "_mask" in d.files else None
    after = d[name + "_x_after"] if name + "_x_after" in d.files else None
    return B, N, J, alpha, x, m, d[name + "_moments"], after


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_routine(golden, name):
    B, N, J, alpha, x, m, want, after = fixture_case(golden, name)
    wavs = split(x, B, N, alpha, J)
    masks = split(m, B, N, alpha, J) if m is not None else None
    code, *got = O.stat_moments(wavs, J, B, N, alpha, masks)
    assert code == (6 if name == "const" else 0)
    assert_moments(got, want, name)
    flat = np.concatenate([w.ravel(order="F") for w in wavs])
    if after is not None:
        assert np.array_equal(flat, after)  # masked in place, as the reference does
    else:
        assert np.array_equal(flat, x)


@pytest.mark.parametrize("name", ["plain", "offset"])
def test_oracle_matches_scipy_central_moments(golden, name):
    B, N, J, alpha, x, _, _, _ = fixture_case(golden, name)
    wavs = split(x, B, N, alpha, J)
    code, *got = O.stat_moments(wavs, J, B, N, alpha)
    assert code == 0
    want = np.zeros((6, J + 1))
    for j, w in enumerate(wavs):
        v = w.ravel(order="F")
        n = v.size
        var = np.var(v, ddof=1)
        c = [scipy.stats.moment(v, k) for k in (3, 4, 5, 6)]  # (1/n) sum (x - mean)^k
        want[:, j] = [np.mean(v), var, c[0] / var ** 1.5, c[1] / var ** 2 - 3, c[2] / var ** 2.5,
                      c[3] / var ** 3 - 15]
        assert n == (2 * (w.shape[1] // 2) - 1) * w.shape[1] * N
    assert_moments(got, want, name)


def test_oracle_exact_discrete_uniform():
    # B=2, alpha=2: J=1, both scales hold 3*4*1 = 12 values; x = 0..11
    B, N, alpha = 2, 1, 2.0
    J = O.comp_Jmax(B, alpha)
    assert J == 1
    wavs = [np.asfortranarray(np.arange(12, dtype=np.float64).reshape((3, 4, 1), order="F")) for _ in range(J + 1)]
    code, *got = O.stat_moments(wavs, J, B, N, alpha)
    assert code == 0
    xs = [Fraction(k) for k in range(12)]
    mu = sum(xs) / 12
    cm = {p: sum((x - mu) ** p for x in xs) / 12 for p in (2, 3, 4, 5, 6)}
    var = cm[2] * 12 / 11
    exact = [float(mu), float(var), 0.0, float(cm[4] / var ** 2) - 3, 0.0, float(cm[6] / var ** 3) - 15]
    for j in range(J + 1):
        assert_moments([g[j:j + 1] for g in got], [np.array([e]) for e in exact], "uniform")
    assert exact[3] == pytest.approx(-1.5016025641025641, abs=1e-15)


def test_oracle_error_paths():
    B, N, alpha = 4, 1, 2.0
    J = O.comp_Jmax(B, alpha)
    his, _ = O.band_limits(J, B, alpha)
    rng = np.random.default_rng(5)
    wavs = [np.asfortranarray(rng.standard_normal((2 * h - 1, 2 * h, N))) for h in his]
    masks = [np.asfortranarray(np.ones_like(w)) for w in wavs]
    masks[1][:] = 0.0  # nj = 0 on scale 1: a halting error in the reference
    code, *_ = O.stat_moments([w.copy(order="F") for w in wavs], J, B, N, alpha, masks)
    assert code == 8
    wavs[0][:] = 3.0  # zero variance: warning, zero higher moments
    code, mean, var, skew, kur, five, six = O.stat_moments(wavs, J, B, N, alpha)
    assert code == 6
    assert mean[0] == pytest.approx(3.0, abs=1e-15) and abs(var[0]) < 1e-12
    assert skew[0] == kur[0] == five[0] == six[0] == 0.0
    assert var[1] > 0.1 and kur[1] != 0.0
