"""rgb_to_spectral_radiance (crates/gemray/src/renderer/env_map_spectrum.rs:60-82) in the
oracle, pinned lobe by lobe by closed forms, plus the reference's own four tests
(env_map_spectrum.rs:88-118).

Each channel alone is one asymmetric Gaussian: exactly 1 at its centre mu (exp(-0) = 1) and
exp(-1/2) at mu - sigma_lo and at mu + sigma_hi, where (x - mu) / sigma = -1 and +1 exactly.
Those three points tell the two widths apart wherever they differ, so a swapped sigma_lo /
sigma_hi, a wrong centre or a lobe attached to the wrong channel fails here.  CPU only."""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

# (channel, centre nm, sigma below the centre, sigma above) as the reference documents
# them (env_map_spectrum.rs:65-71).
LOBES = [(0, 615.0, 45.0, 65.0), (1, 545.0, 45.0, 45.0), (2, 465.0, 40.0, 45.0)]
EXP_M_HALF = np.float32(math.exp(-0.5))


def one_hot(c):
    rgb = [0.0, 0.0, 0.0]
    rgb[c] = 1.0
    return rgb


def ulps(a, b):
    return abs(int(np.float32(a).view(np.int32)) - int(np.float32(b).view(np.int32)))


@pytest.mark.parametrize("c,mu,lo,hi", LOBES)
def test_each_lobe_peaks_at_its_centre_and_falls_to_exp_minus_half_one_sigma_out(c, mu, lo, hi):
    rgb = one_hot(c)
    assert po.rgb_to_spectral_radiance(rgb, mu) == 1.0
    assert ulps(po.rgb_to_spectral_radiance(rgb, mu - lo), EXP_M_HALF) <= 1
    assert ulps(po.rgb_to_spectral_radiance(rgb, mu + hi), EXP_M_HALF) <= 1
    # Two sigma out on each side: exp(-2).
    assert ulps(po.rgb_to_spectral_radiance(rgb, mu - 2 * lo), np.float32(math.exp(-2.0))) <= 1
    assert ulps(po.rgb_to_spectral_radiance(rgb, mu + 2 * hi), np.float32(math.exp(-2.0))) <= 1
    # Strictly below 1 on both sides of the centre, falling monotonically away from it.
    left = [po.rgb_to_spectral_radiance(rgb, mu - d) for d in (1, 10, 30, 60)]
    right = [po.rgb_to_spectral_radiance(rgb, mu + d) for d in (1, 10, 30, 60)]
    assert all(1.0 > a > b for a, b in zip(left, left[1:]))
    assert all(1.0 > a > b for a, b in zip(right, right[1:]))


def test_asymmetric_widths_are_not_swapped():
    # Red and blue are asymmetric: one sigma below the centre on the wide side's width would
    # read exp(-1/2 (wide/narrow)^2), not exp(-1/2).
    r_lo = po.rgb_to_spectral_radiance([1, 0, 0], 615.0 - 65.0)
    assert r_lo == pytest.approx(math.exp(-0.5 * (65.0 / 45.0) ** 2), rel=1e-6)
    b_hi = po.rgb_to_spectral_radiance([0, 0, 1], 465.0 + 40.0)
    assert b_hi == pytest.approx(math.exp(-0.5 * (40.0 / 45.0) ** 2), rel=1e-6)


def test_channels_combine_linearly_with_the_fma_chain():
    # r * R(l) + (g * G(l) + b * B(l)), each product and sum rounded once (fma).
    rng = np.random.default_rng(0)
    for _ in range(200):
        rgb = rng.uniform(0, 3, 3).astype(np.float32)
        lam = np.float32(rng.uniform(380, 780))
        R, G, B = (np.float32(po.rgb_to_spectral_radiance(one_hot(c), lam)) for c in range(3))
        inner = np.float32(np.float64(rgb[1]) * G + np.float64(np.float32(rgb[2] * B)))
        expect = np.float32(np.float64(rgb[0]) * R + np.float64(inner))
        assert po.rgb_to_spectral_radiance(rgb, lam) == expect


# The reference's own tests (env_map_spectrum.rs:88-118).
def test_zero_rgb_gives_zero_spectrum():
    for lam in (380.0, 500.0, 560.0, 650.0, 780.0):
        assert po.rgb_to_spectral_radiance([0, 0, 0], lam) == 0.0


def test_negative_components_are_clamped_not_propagated():
    v = po.rgb_to_spectral_radiance([-1.0, -2.0, -3.0], 550.0)
    assert v == 0.0 and not math.isnan(v)


def test_result_is_always_finite_and_non_negative():
    for lam in (300.0, 380.0, 450.0, 550.0, 650.0, 780.0, 900.0):
        v = po.rgb_to_spectral_radiance([1.0, 2.5, 100.0], lam)
        assert math.isfinite(v) and v >= 0.0


def test_red_weight_dominates_the_red_end_of_the_spectrum():
    assert po.rgb_to_spectral_radiance([1, 0, 0], 615.0) > po.rgb_to_spectral_radiance([0, 0, 1], 615.0)


def test_analytic_furnace_target_against_a_float64_quadrature():
    """furnace.rs:294-302: target = sum over lambda = 380..780 nm (1 nm) of CMF(lambda) x
    rgb_to_spectral_radiance([l0; 3], lambda) / 106.856.  Recomputed here in float64 from
    the closed-form lobes and the oracle's scalar CMF, so a wrong step, span, lobe sum or
    normalization in the oracle's loop moves it by far more than f32 rounding."""
    l0 = 2.5
    target = np.zeros(3, np.float32)
    po.lib().or_analytic_furnace_target(l0, po.fptr(target))

    def lobe(x, mu, lo, hi):
        s = lo if x < mu else hi
        return math.exp(-0.5 * ((x - mu) / s) ** 2)

    acc = np.zeros(3)
    cmf = np.zeros(3, np.float32)
    for step in range(401):
        lam = 380.0 + step
        spec = l0 * sum(lobe(lam, mu, lo, hi) for _, mu, lo, hi in LOBES)
        po.lib().or_cie_1931_cmf(lam, po.fptr(cmf))
        acc += cmf.astype(np.float64) * spec
    ref = acc / 106.856
    assert np.all(ref > 0)
    assert np.max(np.abs(target - ref) / ref) < 2e-5
