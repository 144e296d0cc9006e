"""The biaxial indicatrix on the device (gr_probe) against the oracle's restatement of
optics/birefringence.rs:435-836 (needs an MI355X).

Away from a degenerate pair of principal indices the device functions use only IEEE
add/mul/fma/div/sqrt in the reference's order (no transcendental), so wave_indices,
eigen_polarizations and mode_poynting_dir must match the oracle bit for bit; the
near-degenerate uniaxial short-cut takes sin/cos of theta from its cosine (as the uniaxial
kernels do) and is held to a few ulp.  Inputs: the three biaxial built-ins (Alexandrite,
Topaz, Tanzanite) at several wavelengths, synthetic strongly biaxial indicatrices, and
directions drawn uniformly on the sphere plus a dense sweep through both optic axes.
"""
import numpy as np
import pytest

import pyoracle as po
from gemray_studio_amd import _lib
from gemray_studio_amd import backend as B
from gemray_studio_amd import materials as M

pytestmark = pytest.mark.gpu
F = np.float32


@pytest.fixture(scope="module")
def gpu():
    b = B.GpuBackend.acquire()
    assert b.available, "no gfx950 device: the HIP path is required for these tests"
    yield b
    b.close()


def indicatrices():
    """(indicatrix, the gamma axis it was built from) -- the probe builds the frame from
    the same unnormalized axis."""
    out = []
    for name in ("Alexandrite", "Topaz", "Tanzanite"):
        mat = M.by_name(name)
        m = M.to_struct(mat)
        for lam in (400.0, 520.0, 589.3, 690.0):
            out.append((po.Indicatrix.of_material(m, lam), mat.c_axis))
    for n, g in (((1.60, 1.65, 1.75), [0.3, -0.5, 0.8]), ((1.4, 2.4, 2.49), [0.0, 1.0, 0.0]),
                 ((2.2, 2.21, 2.35), [0.95, 0.1, 0.3])):
        out.append((po.Indicatrix.from_gamma_axis(*n, g), g))
    return out


def directions(ind, rng, n):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    # The optic axes lie in the alpha-gamma plane at +-V from gamma, tan^2 V =
    # (b_a - b_b) / (b_b - b_g); sweep across both.
    b = 1.0 / ind.n.astype(np.float64) ** 2
    vz = np.arctan(np.sqrt(max((b[0] - b[1]) / (b[1] - b[2]), 0.0)))
    ax = ind.axes.astype(np.float64)
    t = np.linspace(-3e-3, 3e-3, 64)
    sweep = []
    for s in (1.0, -1.0):
        for dt in t:
            a = s * vz + dt
            sweep.append(np.cos(a) * ax[:, 2] + np.sin(a) * ax[:, 0] + 0.3 * dt * ax[:, 1])
    v = np.concatenate([v, np.array(sweep)])
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(F)


def probe_inputs(ind, gamma, w):
    n = len(w)
    return np.concatenate([np.tile(ind.n, (n, 1)), w, np.tile(gamma, (n, 1))], 1).astype(F)


def test_wave_indices_eigenvectors_and_walk_off_match_the_oracle(gpu):
    rng = np.random.default_rng(21)
    n_vec = n_vec_exact = 0
    worst = 0.0
    for ind, gamma in indicatrices():
        w = directions(ind, rng, 1500)
        x = probe_inputs(ind, gamma, w)
        got_n = gpu.probe(_lib.GR_PROBE_BX_WAVE_INDICES, x)[:, :2]
        want_n = np.array([ind.wave_indices(d) for d in w], F)
        assert np.array_equal(got_n.view(np.uint32), want_n.view(np.uint32)), ind.n
        for fn, ref in ((_lib.GR_PROBE_BX_EIGEN_SLOW, lambda d: ind.eigen_polarizations(d)[0]),
                        (_lib.GR_PROBE_BX_EIGEN_FAST, lambda d: ind.eigen_polarizations(d)[1]),
                        (_lib.GR_PROBE_BX_POYNTING_SLOW, lambda d: ind.mode_poynting_dir(d, True)),
                        (_lib.GR_PROBE_BX_POYNTING_FAST, lambda d: ind.mode_poynting_dir(d, False))):
            got = gpu.probe(fn, x)
            want = np.array([ref(d) for d in w], F)
            same = np.all(got.view(np.uint32) == want.view(np.uint32), axis=1)
            n_vec += len(w)
            n_vec_exact += int(same.sum())
            worst = max(worst, float(np.abs(got - want).max()))
    print(f"[biaxial probes] vectors bit-exact {n_vec_exact}/{n_vec}, worst abs diff {worst:.3e}", flush=True)
    assert n_vec_exact == n_vec, (n_vec_exact, n_vec, worst)


def test_degenerate_pairs_take_the_uniaxial_short_cut(gpu):
    # n_alpha == n_beta (positive uniaxial about gamma) and n_beta == n_gamma (negative
    # uniaxial about alpha): the device forms sin theta from cos theta where the reference
    # takes acos then sin/cos; within 4 ulp of the index.
    rng = np.random.default_rng(4)
    gamma = [0.35, 0.82, -0.45]
    for n in ((1.65, 1.65, 1.72), (1.76, 1.77, 1.77), (1.9, 1.9, 1.9)):
        ind = po.Indicatrix.from_gamma_axis(*n, gamma)
        w = directions(ind, rng, 500) if n[0] != n[2] else rng.normal(size=(200, 3)).astype(F)
        w = (w / np.linalg.norm(w, axis=1, keepdims=True)).astype(F)
        got = gpu.probe(_lib.GR_PROBE_BX_WAVE_INDICES, probe_inputs(ind, gamma, w))[:, :2]
        want = np.array([ind.wave_indices(d) for d in w], F)
        ulp = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
        assert ulp.max() <= 4, (n, ulp.max())
