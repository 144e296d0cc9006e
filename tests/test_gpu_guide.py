"""The denoiser's primary-hit guide buffers on the device (gr_guide_buffers) against the
oracle's generate_guide_buffers (apps/diagram-gui/src/bridge/guide_pass.rs:37-110) (needs an
MI355X).  The pass is IEEE add/mul/div/sqrt in the reference's order (the camera ray, the
convex clip), so depth, normal and facet must match bit for bit.
"""
import numpy as np
import pytest

import pyoracle as po
from gemray_studio_amd import _lib
from gemray_studio_amd import backend as B
from gemray_studio_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    b = B.GpuBackend.acquire()
    assert b.available, "no gfx950 device: the HIP path is required for these tests"
    yield b
    b.close()


CASES = [
    ("srb", lambda: po.standard_round_brilliant(), (0.35, 0.28, 5.0, 18.0), 96, 64),
    ("oval", lambda: scenes.oval_brilliant(), (1.1, 0.9, 4.0, 24.0), 120, 90),
    ("step-233", lambda: scenes.step_cut(), (-0.4, 1.5, 3.5, 30.0), 77, 131),
    ("random-181", lambda: scenes.random_convex_stone(12, 181), (2.0, -0.3, 4.5, 20.0), 64, 64),
    ("pole", lambda: po.standard_round_brilliant(), (0.0, float(np.float32(np.pi / 2)), 5.0, 18.0), 50, 50),
    ("emerald-34", lambda: po.emerald_cut(), (0.6, 0.45, 4.0, 40.0), 200, 100),
]


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_guide_buffers_are_bit_identical_to_the_oracle(gpu, name):
    _, make, (yaw, pitch, dist, fov), w, h = next(c for c in CASES if c[0] == name)
    planes = make()
    d_g, n_g, f_g = gpu.guide_buffers(B.camera_new(yaw, pitch, dist, fov), w, h, planes)
    d_o, n_o, f_o = po.guide_buffers(po.camera_new(yaw, pitch, dist, fov), w, h, planes)
    assert np.array_equal(f_g, f_o)
    assert np.array_equal(d_g.view(np.uint32), d_o.view(np.uint32))
    assert np.array_equal(n_g.view(np.uint32), n_o.view(np.uint32))
    assert (f_g >= 0).mean() > 0.1 and (f_g == -1).any()


def test_guide_buffers_edge_cases(gpu):
    cam = B.camera_new(0.35, 0.28, 5.0, 18.0)
    d, n, f = gpu.guide_buffers(cam, 0, 7, po.standard_round_brilliant())
    assert d.size == 0 and f.size == 0
    d, n, f = gpu.guide_buffers(cam, 3, 2, np.zeros((0, 4), np.float32))
    assert np.all(d == np.float32(1e30)) and np.all(f == 0) and np.all(n == 0)
    bad = po.standard_round_brilliant().copy()
    bad[5, 3] = np.nan
    with pytest.raises(_lib.GrError) as e:
        gpu.guide_buffers(cam, 8, 8, bad)
    assert e.value.status == _lib.GR_ERR_INVALID
