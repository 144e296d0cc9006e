"""Per-function ULP budgets of the transport kernel's own device functions (gr_probe)
against the oracle's statement of the reference's route (needs an MI355X).

The kernel replaces angle round trips (acos -> sin/cos, atan -> sin/cos, atan2 -> sin/cos)
with algebraic identities; each such function is evaluated on the device on randomized
physical inputs and compared with the reference's route through the transcendentals:

  function               kernel form                              budget (abs, on O(1) values)
  n_eff(theta_c)         sin = sqrt(1 - cos^2)                     4 ulp of n_eff
  Poynting direction     cos/sin delta from tan rho                8e-7 per component
  TIR retardation        rational cos/sin delta from tan(delta/2)  8e-7
  frame rotation         double angle of (cos psi, sin psi)        8e-7
  field quadratic form   cos/sin 2psi from (Q, U) / |(Q, U)|       8e-7 x max(ao, ae)
                         (hardware reciprocal square root on (Q, U)
                         scaled to a unit exponent; |Q|, |U| down to
                         denormals)
  degree of polarization hardware square root and reciprocal        5e-7 x the value
  companion Fresnel      r_s, r_p, t_s, t_p and the transmission     3e-7 (r_s, r_p);
                         factor on the hardware reciprocal           1e-6 x m00
  lobe exponential       native exp (as the reference's shader)    2.4e-7 + |x| 1.2e-7 rel.
  CIE matching functions lobes on the native exp                   2e-6 of the largest
                                                                   component (the reference
                                                                   CPU uses a Cephes exp)

Budgets are a few ulp of the outputs' unit scale: the identities are exact in real
arithmetic, and the reference's own route rounds at every transcendental.
"""
import numpy as np
import pytest

import pyoracle as po
from gemray_studio_amd import _lib
from gemray_studio_amd import backend as B

pytestmark = pytest.mark.gpu
L = po.lib()
F = np.float32
N = 20000


@pytest.fixture(scope="module")
def gpu():
    b = B.GpuBackend.acquire()
    assert b.available, "no gfx950 device: the HIP path is required for these tests"
    yield b
    b.close()


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(F)


def ulp_dist(a, b):
    ia = a.astype(F).view(np.int32).astype(np.int64)
    ib = b.astype(F).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def test_effective_index_from_cosine(gpu):
    rng = np.random.default_rng(1)
    n_o = rng.uniform(1.4, 2.7, N).astype(F)
    dn = rng.uniform(-0.2, 0.2, N).astype(F)
    n_e = (n_o + dn).astype(F)
    c = rng.uniform(-1.0, 1.0, N).astype(F)
    c[:200] = F(1.0)  # along the optic axis
    c[200:400] = F(0.0)
    out = gpu.probe(_lib.GR_PROBE_NEFF, np.stack([n_o, n_e, c], 1))[:, 0]
    ref = np.array([L.or_neff_from_cos(float(a), float(b), float(x)) for a, b, x in zip(n_o, n_e, c)], F)
    assert ulp_dist(out, ref).max() <= 4


def test_poynting_direction(gpu):
    rng = np.random.default_rng(2)
    w, c = unit(rng, N), unit(rng, N)
    n_o = rng.uniform(1.4, 2.7, N).astype(F)
    n_e = (n_o + rng.uniform(-0.2, 0.2, N)).astype(F)
    out = gpu.probe(_lib.GR_PROBE_POYNTING, np.concatenate([w, c, n_o[:, None], n_e[:, None]], 1))
    ref = np.zeros((N, 3), F)
    for i in range(N):
        L.or_extraordinary_poynting_dir(po.fptr(w[i]), po.fptr(c[i]), float(n_o[i]), float(n_e[i]), po.fptr(ref[i]))
    assert np.abs(out - ref).max() <= 8e-7
    # The walk-off is real: most directions moved off the wave normal.
    assert np.mean(np.abs(ref - w).max(axis=1) > 1e-4) > 0.5


def test_tir_retardation(gpu):
    rng = np.random.default_rng(3)
    n1 = rng.uniform(1.4, 2.7, N).astype(F)
    sin_c = 1.0 / n1
    sin_i = rng.uniform(sin_c, 1.0).astype(F)  # past the critical angle
    cos_i = np.sqrt(np.maximum(1.0 - sin_i.astype(np.float64) ** 2, 0.0)).astype(F)
    out = gpu.probe(_lib.GR_PROBE_TIR_RETARDATION, np.stack([n1, cos_i, sin_i], 1))
    delta = np.array([L.or_tir_phase_delta(float(a), float(b), float(s)) for a, b, s in zip(n1, cos_i, sin_i)], F)
    ref = np.stack([np.cos(delta.astype(np.float64)), np.sin(delta.astype(np.float64))], 1)
    assert np.abs(out[:, :2] - ref).max() <= 8e-7


def test_frame_rotation(gpu):
    rng = np.random.default_rng(4)
    prev, cur, axis = unit(rng, N), unit(rng, N), unit(rng, N)
    out = gpu.probe(_lib.GR_PROBE_FRAME_ROTATION, np.concatenate([prev, cur, axis], 1))
    ref = np.zeros((N, 2), F)
    for i in range(N):
        L.or_frame_rotation_cs(po.fptr(prev[i]), po.fptr(cur[i]), po.fptr(axis[i]), po.fptr(ref[i]))
    assert np.abs(out[:, :2] - ref).max() <= 8e-7


def test_field_quadratic_form(gpu):
    rng = np.random.default_rng(5)
    s = unit(rng, N)
    p = np.cross(s, unit(rng, N))
    p = (p / np.linalg.norm(p, axis=1, keepdims=True)).astype(F)
    mag = 10.0 ** rng.uniform(-6, 0, N)
    mag[100:1100] = 10.0 ** rng.uniform(-44, -18, 1000)  # squares underflow; denormal Q, U
    ang = rng.uniform(0, 2 * np.pi, N)
    q = (mag * np.cos(ang)).astype(F)
    u = (mag * np.sin(ang)).astype(F)
    q[:100] = F(0.0)
    u[:100] = F(0.0)  # unpolarized in Q/U: psi = 0
    u[1100:1200] = F(0.0)  # linear along +-Q only
    ao = rng.uniform(0.0, 1.0, N).astype(F)
    out = gpu.probe(_lib.GR_PROBE_FIELD_QFORM, np.concatenate([q[:, None], u[:, None], s, p, ao[:, None]], 1))[:, 0]
    ae = (F(1.0) - ao).astype(F)
    ref = np.array([L.or_field_qform(float(q[i]), float(u[i]), po.fptr(s[i]), po.fptr(p[i]), float(ao[i]),
                                     float(ae[i])) for i in range(N)], F)
    scale = np.maximum(ao, ae)
    assert (np.abs(out - ref) / scale).max() <= 8e-7


def test_degree_of_polarization(gpu):
    """sqrt(Q^2 + U^2 + V^2) / I clamped to [0, 1] (polarization.rs:38-48), as the absorption
    step evaluates it, against the oracle's correctly rounded route: within 5e-7 of the value
    (4 ulp), over intensities from 1e-9 (below the 1e-7 cut, where both return 0) to 1e3,
    degrees from 0 to 1.2 (clamped) and components down to denormals."""
    rng = np.random.default_rng(12)
    i = 10.0 ** rng.uniform(-9, 3, N)
    dop = rng.uniform(0.0, 1.2, N)
    dop[:500] = 10.0 ** rng.uniform(-40, -10, 500)
    d = rng.normal(size=(N, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    quv = d * (i * dop)[:, None]
    quv[500:600, 1:] = 0.0
    x = np.concatenate([i[:, None], quv], 1).astype(F)
    out = gpu.probe(_lib.GR_PROBE_DOP, x)[:, 0]
    ref = np.array([L.or_degree_of_polarization(*map(float, x[k])) for k in range(N)], F)
    assert np.all((out == 0) == (ref == 0))
    assert np.all(np.abs(out.astype(np.float64) - ref) <= 5e-7 * ref.astype(np.float64) + 1e-38)
    assert np.count_nonzero(ref == 1.0) > 1000 and np.count_nonzero((ref > 0) & (ref < 1)) > 10000


def test_companion_fresnel(gpu):
    """A companion channel's Fresnel amplitudes and transmission Mueller element m00 on the
    hardware reciprocal, against the reference's divisions (refraction.rs:1049-1056,
    polarization.rs:119-137), entering and leaving stones of index 1.3-2.8, from normal to
    grazing incidence."""
    rng = np.random.default_rng(13)
    n = rng.uniform(1.3, 2.8, N)
    entering = rng.random(N) < 0.5
    n1 = np.where(entering, 1.0, n)
    n2 = np.where(entering, n, 1.0)
    cos_i = np.concatenate([rng.uniform(1e-3, 1.0, N - 200), np.full(200, 1.0)])
    s2 = (n1 / n2) ** 2 * (1.0 - cos_i ** 2)
    keep = s2 < 1.0
    x = np.stack([n1, n2, cos_i, np.sqrt(np.maximum(1.0 - s2, 0.0))], 1)[keep].astype(F)
    out = gpu.probe(_lib.GR_PROBE_FRESNEL_COMPANION, x).astype(np.float64)
    ref = np.zeros((len(x), 3), F)
    for k in range(len(x)):
        L.or_fresnel_companion(*map(float, x[k]), po.fptr(ref[k]))
    ref = ref.astype(np.float64)
    assert len(x) > 0.5 * N
    assert np.abs(out[:, :2] - ref[:, :2]).max() <= 3e-7
    assert np.all(np.abs(out[:, 2] - ref[:, 2]) <= 1e-6 * np.abs(ref[:, 2]))


def test_lobe_exponential(gpu):
    rng = np.random.default_rng(6)
    x = np.concatenate([-(10.0 ** rng.uniform(-6, 1.7, N - 8)), np.zeros(8)]).astype(F)
    out = gpu.probe(_lib.GR_PROBE_EXP, x[:, None])[:, 0].astype(np.float64)
    exact = np.exp(x.astype(np.float64))
    rel = np.abs(out - exact) / exact
    assert np.all(rel <= 2.4e-7 + np.abs(x) * 1.2e-7)
    # Against the reference CPU's Cephes polynomial (the oracle's exp_f32x8).
    ref = np.zeros(len(x), F)
    for i in range(0, len(x) - len(x) % 8, 8):
        L.or_exp_f32x8(po.fptr(x[i:i + 8].copy()), po.fptr(ref[i:i + 8]))
    m = (x > -5.0)[: len(x) - len(x) % 8]
    assert ulp_dist(out[: len(m)][m].astype(F), ref[: len(m)][m]).max() <= 8


def test_cie_matching_functions(gpu):
    # Every comb wavelength the kernel can see lies in [380, 780).
    rng = np.random.default_rng(7)
    lam = rng.uniform(380.0, 780.0, N).astype(F)
    out = gpu.probe(_lib.GR_PROBE_CMF, lam[:, None])
    ref = np.zeros((N, 3), F)
    for i in range(0, N, 8):
        buf = np.zeros(24, F)
        L.or_cie_1931_cmf_x8(po.fptr(lam[i:i + 8].copy()), po.fptr(buf))
        ref[i:i + 8] = buf.reshape(8, 3)
    err = np.abs(out.astype(np.float64) - ref) / np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-3)
    assert err.max() <= 2e-6
