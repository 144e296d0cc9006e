"""CPU+GPU hybrid rendering (gemray_studio_amd.hybrid, crates/gemray/src/renderer/gpu/
hybrid.rs:82-336) with the device leg on the MI355X and the oracle as the host's CPU leg.

* Merge by sum: accum[p] += gpu[p] + cpu[p], the device leg being samples [0, g) and the
  CPU leg [g, g + c) -- bit for bit against the two legs run separately.
* Determinism: two renders with the same explicit split are byte-identical (hybrid.rs:64-74).
* A scene the device declines renders all-CPU (forced_cpu_only).
* The hybrid image estimates the same radiance as an all-device render of the same total
  (per-pixel Welch z, as the statistical image comparison).
"""
import dataclasses

import numpy as np
import pytest

import pyoracle as po
from gemray_studio_amd import backend as B
from gemray_studio_amd import materials as M
from gemray_studio_amd.hybrid import HybridSplit, render_hybrid, supports
from hybrid_util import oracle_cpu_trace

pytestmark = pytest.mark.gpu
CAM = (0.35, 0.28, 5.0, 18.0)


@pytest.fixture(scope="module")
def gpu():
    b = B.GpuBackend.acquire()
    assert b.available, "no gfx950 device: the HIP path is required for these tests"
    yield b
    b.close()


def scene(material=None, w=24, h=20):
    return B.SceneRef(camera=B.camera_new(*CAM), width=w, height=h, planes=po.standard_round_brilliant(),
                      material=material or M.by_name("Sapphire"), max_bounces=12,
                      environment=B.Studio("RingLights", 1.0, 0.85, 0.95))


def test_merge_is_the_sum_of_the_two_legs_and_deterministic(gpu):
    sc = scene()
    cpu = oracle_cpu_trace(CAM)
    n = sc.width * sc.height
    out = []
    for _ in range(2):
        accum = np.full((n, 3), 0.5, np.float32)
        st = render_hybrid(gpu, sc, HybridSplit(6, 4), accum, cpu)
        assert not st.forced_cpu_only and (st.gpu_samples, st.cpu_samples) == (6, 4)
        assert st.gpu_samples_per_sec > 0 and st.cpu_samples_per_sec > 0
        out.append(accum)
    g = np.zeros((n, 3), np.float32)
    gpu.accumulate(sc, 0, 6, g)
    expect = np.full((n, 3), 0.5, np.float32) + (g + cpu(sc, 6, 4))
    assert np.array_equal(out[0].view(np.uint32), expect.view(np.uint32))
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))
    # all-device and all-CPU splits
    a = np.zeros((n, 3), np.float32)
    render_hybrid(gpu, sc, HybridSplit(10, 0), a, cpu)
    g10 = np.zeros((n, 3), np.float32)
    gpu.accumulate(sc, 0, 10, g10)
    assert np.array_equal(a.view(np.uint32), (np.zeros((n, 3), np.float32) + (g10 + 0.0)).view(np.uint32))


def test_declined_scene_renders_all_cpu(gpu):
    declined = scene(dataclasses.replace(M.by_name("Diamond"), biaxial_delta_beta_alpha=0.002), 8, 6)
    assert not supports(gpu, declined) and supports(gpu, scene())
    cpu = oracle_cpu_trace(CAM)
    accum = np.zeros((48, 3), np.float32)
    st = render_hybrid(gpu, declined, HybridSplit(3, 2), accum, cpu)
    assert st.forced_cpu_only and (st.gpu_samples, st.cpu_samples) == (0, 5)
    assert np.array_equal(accum.view(np.uint32), (np.zeros((48, 3), np.float32) + cpu(declined, 0, 5)).view(np.uint32))
    assert HybridSplit.calibrated(gpu, declined, 16, 2, cpu) == HybridSplit(0, 16)


def test_calibrated_split_and_unbiased_merge(gpu):
    sc = scene(w=16, h=16)
    cpu = oracle_cpu_trace(CAM)
    split = HybridSplit.calibrated(gpu, sc, 64, 4, cpu)
    assert split.total_spp == 64 and split.gpu_samples > split.cpu_samples
    # The hybrid image (half the samples each way) and an all-device image of disjoint
    # samples estimate the same radiance: per-pixel Welch z on luminance.
    n, spp = 256, 128
    xyz_g, _ = gpu.trace_samples(sc, 0, n, 1000, spp, with_info=False)
    hyb = np.zeros((n, 3), np.float32)
    render_hybrid(gpu, sc, HybridSplit(spp // 2, spp // 2), hyb, cpu)
    # Per-sample variance from the device's own samples (both estimators share it).
    g = xyz_g.astype(np.float64).reshape(n, spp, 3)[:, :, 1]
    se = np.sqrt(2 * g.var(1, ddof=1) / spp)
    z = np.where(se > 0, (hyb[:, 1] / spp - g.mean(1)) / np.maximum(se, 1e-30), 0.0)
    assert int((np.abs(z) > 3.0).sum()) <= max(5 * n * 0.0027, 10)
