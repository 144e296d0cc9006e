"""The hybrid split's arithmetic and its all-CPU route (gemray_studio_amd.hybrid,
crates/gemray/src/renderer/gpu/hybrid.rs:82-336) (CPU only: without a device every scene is
rendered by the CPU leg, HybridStats::forced_cpu_only)."""
import numpy as np
import pytest

from gemray_studio_amd import _lib
from gemray_studio_amd import backend as B
from gemray_studio_amd import materials as M
from gemray_studio_amd.hybrid import HybridSplit, render_hybrid, supports
from hybrid_util import oracle_cpu_trace
import pyoracle as po


def test_split_from_rates():
    # hybrid.rs:180-205: floor(total * gpu_share + 0.5), capped at the total.
    assert HybridSplit.from_rates(10, 3.0, 1.0) == HybridSplit(8, 2)
    assert HybridSplit.from_rates(10, 1.0, 3.0) == HybridSplit(3, 7)
    assert HybridSplit.from_rates(9, 5.0, 0.0) == HybridSplit(9, 0)
    assert HybridSplit.from_rates(9, 0.0, 2.0) == HybridSplit(0, 9)
    assert HybridSplit.from_rates(7, 0.0, 0.0) == HybridSplit(3, 4)
    assert HybridSplit.from_rates(0, 1.0, 1.0) == HybridSplit(0, 0)
    assert HybridSplit.cpu_only(12) == HybridSplit(0, 12) and HybridSplit(5, 7).total_spp == 12
    for total in range(0, 300, 7):
        for g, c in ((1.0, 2.0), (2.5e9, 3.1e7), (1e-3, 1.0)):
            s = HybridSplit.from_rates(total, g, c)
            assert s.total_spp == total and 0 <= s.gpu_samples <= total
            assert abs(s.gpu_samples - total * g / (g + c)) <= 0.5 + 1e-9


def test_without_a_device_the_cpu_leg_renders_everything():
    gpu = B.GpuBackend.disabled()
    cam = (0.35, 0.28, 5.0, 18.0)
    scene = B.SceneRef(camera=B.camera_new(*cam), width=10, height=8, planes=po.standard_round_brilliant(),
                       material=M.by_name("Sapphire"), max_bounces=12, environment=B.Studio())
    cpu = oracle_cpu_trace(cam)
    accum = np.full((80, 3), 0.25, np.float32)
    st = render_hybrid(gpu, scene, HybridSplit(5, 3), accum, cpu)
    assert st.forced_cpu_only and st.gpu_samples == 0 and st.cpu_samples == 8
    expect = np.full((80, 3), 0.25, np.float32) + (np.zeros((80, 3), np.float32) + cpu(scene, 0, 8))
    assert np.array_equal(accum.view(np.uint32), expect.view(np.uint32))


class _FakeDevice:
    """Stands in for GpuBackend in supports(): accumulate raises the given status."""

    available = True

    def __init__(self, status):
        self.status = status

    def accumulate(self, scene, sample_offset, spp, accum):
        if self.status:
            raise _lib.GrError(self.status, "gr_accumulate")


def test_supports_declines_only_on_decline_codes():
    scene = B.SceneRef(camera=B.camera_new(0.35, 0.28, 5.0, 18.0), width=4, height=3,
                       planes=po.standard_round_brilliant(), material=M.by_name("Sapphire"), max_bounces=12,
                       environment=B.Studio())
    assert supports(_FakeDevice(0), scene)
    for st in (_lib.GR_DECLINE_NO_DEVICE, _lib.GR_DECLINE_ENV, _lib.GR_DECLINE_MATERIAL, _lib.GR_DECLINE_BUDGET):
        assert not supports(_FakeDevice(st), scene)
    for st in (5, 6):  # GR_ERR_INVALID, GR_ERR_DEVICE: raised, never a silent CPU fallback
        with pytest.raises(_lib.GrError):
            supports(_FakeDevice(st), scene)
