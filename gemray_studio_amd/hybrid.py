"""CPU+GPU hybrid frame rendering (crates/gemray/src/renderer/gpu/hybrid.rs:82-336).

The device and the host's CPU tracer trace disjoint sample ranges of one frame at the same
time -- the device samples [0, gpu_samples), the CPU [gpu_samples, gpu_samples +
cpu_samples) -- and their per-pixel sums are added into the caller's film (sum, never
average: the caller divides by the total once, at display time).  Every (pixel, absolute
sample) result is a pure function of its indices, so a render with an explicit split is
deterministic whatever thread traced what.

The CPU leg is the caller's: ``cpu_trace(scene, sample_offset, spp)`` returns the per-pixel
sums of samples [sample_offset, sample_offset + spp) as an (n_pixels, 3) float32 array (the
reference's ``cpu_trace_range``, hybrid.rs:337-470, over its CPU ``trace_spectral_ray``).
This package ships no CPU tracer; the device leg is ``GpuBackend.accumulate``.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass
from fractions import Fraction
from typing import Callable

import numpy as np

from .backend import GpuBackend, SceneRef
from ._lib import (GR_DECLINE_BUDGET, GR_DECLINE_ENV, GR_DECLINE_MATERIAL, GR_DECLINE_NO_DEVICE,
                   GrError)

CpuTrace = Callable[[SceneRef, int, int], np.ndarray]
_DECLINES = (GR_DECLINE_NO_DEVICE, GR_DECLINE_ENV, GR_DECLINE_MATERIAL, GR_DECLINE_BUDGET)


@dataclass(frozen=True)
class HybridSplit:
    """``HybridSplit`` (hybrid.rs:86-210): samples per engine."""

    gpu_samples: int
    cpu_samples: int

    @property
    def total_spp(self) -> int:
        return self.gpu_samples + self.cpu_samples

    @classmethod
    def cpu_only(cls, total_spp: int) -> "HybridSplit":
        return cls(0, total_spp)

    @classmethod
    def from_rates(cls, total_spp: int, gpu_rate: float, cpu_rate: float) -> "HybridSplit":
        """``HybridSplit::from_rates`` (hybrid.rs:180-205): the GPU's share of the total
        rate, rounded to nearest sample as floor(fma(total, share, 0.5)) in f64, capped at
        the total; both rates unmeasurable (zero) splits down the middle."""
        total_rate = gpu_rate + cpu_rate
        if total_rate <= 0.0:
            g = total_spp // 2
            return cls(g, total_spp - g)
        share = gpu_rate / total_rate
        # f64::mul_add: the exact product-sum, rounded once to f64.
        fused = float(Fraction(float(total_spp)) * Fraction(share) + Fraction(0.5))
        g = int(min(math.floor(fused), float(total_spp)))
        return cls(g, total_spp - g)

    @classmethod
    def calibrated(cls, gpu: GpuBackend, scene: SceneRef, total_spp: int, warmup_spp: int,
                   cpu_trace: CpuTrace) -> "HybridSplit":
        """``HybridSplit::calibrated`` (hybrid.rs:120-178): time a warm device dispatch (after a
        throwaway one) and a CPU dispatch of ``warmup_spp`` samples and split ``total_spp``
        by their rates; all-CPU when the device declines the scene.  Not reproducible run to
        run (it measures wall-clock rates); a render given the resulting split is."""
        if not supports(gpu, scene):
            return cls.cpu_only(total_spp)
        if total_spp == 0 or warmup_spp == 0:
            return cls(0, total_spp)
        n = scene.width * scene.height
        gpu.accumulate(scene, 0, warmup_spp, np.zeros((n, 3), np.float32))
        t0 = time.perf_counter()
        gpu.accumulate(scene, 0, warmup_spp, np.zeros((n, 3), np.float32))
        gpu_rate = _rate(n, warmup_spp, time.perf_counter() - t0)
        t0 = time.perf_counter()
        cpu_trace(scene, 0, warmup_spp)
        cpu_rate = _rate(n, warmup_spp, time.perf_counter() - t0)
        return cls.from_rates(total_spp, gpu_rate, cpu_rate)


@dataclass
class HybridStats:
    """``HybridStats`` (hybrid.rs:212-230)."""

    wall_time_s: float
    gpu_samples: int
    cpu_samples: int
    gpu_samples_per_sec: float
    cpu_samples_per_sec: float
    forced_cpu_only: bool


def supports(gpu: GpuBackend, scene: SceneRef) -> bool:
    """Whether the device traces this scene (``GemMaterial::gpu_supported`` and the
    boundary's environment check): a zero-sample accumulate validates and declines exactly
    as a real one would, and touches nothing.  Only a decline (GR_DECLINE_*) means "render
    on the CPU"; an invalid scene or a device failure is raised."""
    if not gpu.available:
        return False
    try:
        gpu.accumulate(scene, 0, 0, np.zeros((scene.width * scene.height, 3), np.float32))
    except GrError as e:
        if e.status in _DECLINES:
            return False
        raise
    return True


def render_hybrid(gpu: GpuBackend, scene: SceneRef, requested_split: HybridSplit, accum: np.ndarray,
                  cpu_trace: CpuTrace) -> HybridStats:
    """``render_hybrid`` (hybrid.rs:232-327): the device leg runs on its own thread for the
    whole CPU leg; then accum[p] += gpu[p] + cpu[p].  A scene the device declines is
    rendered all-CPU (``forced_cpu_only``)."""
    n = scene.width * scene.height
    if accum.shape != (n, 3) or accum.dtype != np.float32:
        raise AssertionError("accumulation buffer must have one entry per pixel")
    forced = not supports(gpu, scene)
    split = HybridSplit.cpu_only(requested_split.total_spp) if forced else requested_split
    wall0 = time.perf_counter()
    gpu_buf = np.zeros((n, 3), np.float32)
    result: dict = {}

    def gpu_leg():
        try:
            if split.gpu_samples:
                t0 = time.perf_counter()
                gpu.accumulate(scene, 0, split.gpu_samples, gpu_buf)
                result["elapsed"] = time.perf_counter() - t0
        except BaseException as e:  # re-raised on the caller's thread
            result["error"] = e

    worker = threading.Thread(target=gpu_leg, name="hybrid-gpu")
    worker.start()
    t0 = time.perf_counter()
    if split.cpu_samples:
        cpu_buf = np.ascontiguousarray(cpu_trace(scene, split.gpu_samples, split.cpu_samples), np.float32)
    else:
        cpu_buf = np.zeros((n, 3), np.float32)
    cpu_elapsed = time.perf_counter() - t0
    worker.join()
    if "error" in result:
        raise result["error"]
    if cpu_buf.shape != (n, 3):
        raise AssertionError("the CPU leg must return one sum per pixel")
    accum += gpu_buf + cpu_buf
    return HybridStats(time.perf_counter() - wall0, split.gpu_samples, split.cpu_samples,
                       _rate(n, split.gpu_samples, result.get("elapsed", 0.0)),
                       _rate(n, split.cpu_samples, cpu_elapsed), forced)


def _rate(n_pixels: int, samples: int, elapsed: float) -> float:
    """samples_per_sec (hybrid.rs:330-336)."""
    if samples == 0:
        return 0.0
    return n_pixels * float(samples) / max(elapsed, 2.220446049250313e-16)
