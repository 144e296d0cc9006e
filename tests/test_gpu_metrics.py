"""Gemological metrics on the device (gr_optical_metrics) against the oracle's
evaluate_gem_optical_metrics (needs an MI355X).

Every metric must be bit-identical to the oracle's: the golden pose of the reference's
test (raytracer_tests.rs:1817-1841), and a sweep of cuts (the reference's round brilliant
and emerald cut, the parametric brilliant, the oval, random convex stones of 23 and 181
planes, the 233-plane step cut) x materials (Sellmeier and Cauchy dispersion, low to high
index) x poses (including pitch 0 and the pole, where the camera basis switches its up
vector), plus the reference's angular profile at one azimuth.  The device's arc cosine
(the only transcendental evaluated on the device here) is checked against the correctly
rounded value and against the oracle's libm on the range fire angles come from.
"""
import json
import os

import numpy as np
import pytest

import pyoracle as po
from gemray_studio_amd import _lib
from gemray_studio_amd import backend as B
from gemray_studio_amd import materials as M
from gemray_studio_amd import scenes
from helpers import bits

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")))
HALF_PI32 = float(np.float32(np.pi / 2))


@pytest.fixture(scope="module")
def gpu():
    b = B.GpuBackend.acquire()
    assert b.available, "no gfx950 device: the HIP path is required for these tests"
    yield b
    b.close()


def oracle(planes, material, pose):
    return po.gem_optical_metrics(planes, M.to_struct(material), *pose)


def device(gpu, planes, material, poses):
    out = gpu.optical_metrics(planes, material, poses)
    return np.array([[o[n] for n in po.METRIC_NAMES] for o in out], np.float32)


def test_golden_pose_bits(gpu):
    g = GOLD["metrics_golden_diamond"]
    d = device(gpu, po.standard_round_brilliant(), M.by_name("Diamond"), [(0.0, 0.45, 0.85, 0.95)])[0]
    assert [bits(v) for v in d] == [g[n] for n in po.METRIC_NAMES]


STONES = [
    ("srb-74", lambda: po.standard_round_brilliant()),
    ("emerald-34", lambda: po.emerald_cut()),
    ("brilliant-73", lambda: scenes.round_brilliant()),
    ("oval-69", lambda: scenes.oval_brilliant()),
    ("random-23", lambda: scenes.random_convex_stone(11, 23)),
    ("random-181", lambda: scenes.random_convex_stone(12, 181)),
    ("step-233", lambda: scenes.step_cut()),
]
MATERIALS = ["Diamond", "Sapphire", "Cubic Zirconia", "Quartz", "Synthetic Moissanite", "Emerald", "Tanzanite"]
POSES = [(0.0, 0.45, 0.85, 0.95), (0.6, 0.45, 0.85, 0.95), (0.0, 1.4, 0.85, 0.95), (2.1, 0.8, 0.3, 0.6),
         (-1.0, 0.0, 1.7, 1.2), (0.35, HALF_PI32, 0.85, 0.95), (3.0, 0.2, -2.0, 0.25), (1.2, 1.1, 0.9, 1.4)]


@pytest.mark.parametrize("stone", [s[0] for s in STONES])
def test_sweep_is_bit_identical_to_the_oracle(gpu, stone):
    planes = dict(STONES)[stone]()
    for name in MATERIALS:
        m = M.by_name(name)
        d = device(gpu, planes, m, POSES)
        for pose, got in zip(POSES, d):
            want = oracle(planes, m, pose)
            assert [bits(v) for v in got] == [bits(v) for v in want], (stone, name, pose, got, want)


def test_angular_profile_at_azimuth(gpu):
    planes = po.standard_round_brilliant()
    m = M.by_name("Diamond")
    az = float(np.float32(45.0) * (np.float32(np.pi) / np.float32(180.0)))
    b, e, w = gpu.angular_profile_at_azimuth(planes, m, az, 0.85, 0.95)
    assert len(b) == len(e) == len(w) == 19
    for i, pitch in enumerate(B.PROFILE_PITCHES_RAD):
        want = oracle(planes, m, (az, float(pitch), 0.85, 0.95))
        assert bits(b[i]) == bits(want[0]) and bits(e[i]) == bits(want[4]) and bits(w[i]) == bits(want[3]), i


def test_no_facets_and_invalid_inputs(gpu):
    m = M.by_name("Diamond")
    d = device(gpu, np.zeros((0, 4), np.float32), m, [(0.0, 0.45, 0.85, 0.95)] * 3)
    assert d.tolist() == [[85.0, 25.0, 75.0, 5.0, 5.0]] * 3
    with pytest.raises(_lib.GrError) as e:
        gpu.optical_metrics(po.standard_round_brilliant(), m, [])
    assert e.value.status == _lib.GR_ERR_INVALID
    with pytest.raises(_lib.GrError):
        gpu.optical_metrics(po.standard_round_brilliant(), m, [(0.0, float("nan"), 0.85, 0.95)])
    bad = po.standard_round_brilliant().copy()
    bad[3, :3] = 0.0
    with pytest.raises(_lib.GrError):
        gpu.optical_metrics(bad, m, [(0.0, 0.45, 0.85, 0.95)])


def test_arc_cosine_is_correctly_rounded(gpu):
    rng = np.random.default_rng(8)
    x = np.concatenate([rng.uniform(-1.0, 1.0, 40000), 1.0 - 10.0 ** rng.uniform(-7.5, -1, 40000),
                        [-1.0, 0.0, 1.0]]).astype(np.float32)
    out = gpu.probe(_lib.GR_PROBE_ACOS, x[:, None])[:, 0]
    cr = np.arccos(x.astype(np.float64)).astype(np.float32)
    assert np.array_equal(out.view(np.uint32), cr.view(np.uint32))
