"""The render-backend boundary: gemray's ``GpuBackend`` over the HIP C ABI.

Mirrors ``crates/gemray/src/renderer/gpu_backend.rs:85-238``:

* ``GpuBackend.acquire()`` opens the gfx950 device or returns a backend that declines
  every call (``gpu_backend.rs:114-126``);
* ``GpuBackend.disabled()`` never touches a device (``:135``);
* ``adapter_label()`` is ``None`` when no device is held (``:146-153``);
* ``try_accumulate(scene, sample_offset, spp, accum)`` ADDS the samples into ``accum`` and
  returns ``True``, or returns ``False`` and leaves ``accum`` untouched so the caller can
  fall back to its CPU tracer (``:160-193``).  A length mismatch between ``accum`` and the
  frame is a programmer error and raises, as the reference asserts (``frame.rs:673-677``).

The package itself has no CPU tracer: a declined call is the caller's to handle.
"""
from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass
from typing import Optional, Sequence, Union

import numpy as np

from . import _lib
from ._lib import Camera, Env, GrError, JobDesc, JobStats, Options, Scene, ServeMetrics, Stats, check
from .materials import GemMaterial, MaterialStruct, to_struct

# LightingPreset::params (optics/raytracer/environment.rs:81-100): (temp_k, spot_mult).
PRESETS = {
    "Daylight": (6500.0, 1.0),
    "Incandescent": (3200.0, 1.2),
    "RingLights": (5000.0, 1.6),
    "DarkSpotlight": (6000.0, 2.4),
}


@dataclass(frozen=True)
class Studio:
    """``EnvironmentSource::Studio`` (environment.rs:189-194)."""

    preset: str = "Daylight"
    exposure: float = 1.0
    light_yaw: float = 0.0
    light_pitch: float = 0.0


@dataclass(frozen=True)
class HdrMap:
    """``EnvironmentSource::HdrMap`` -- an ``EnvironmentMap`` of shape (H, W, 3) f32.  With
    ``device`` (GpuBackend.upload_env_map) the calls reference that HBM copy instead of
    uploading ``rgb``."""

    rgb: np.ndarray
    device: Optional["EnvMapDevice"] = None

    @staticmethod
    def uniform(l0: float) -> "HdrMap":
        """``EnvironmentMap::uniform(1, 1, [l0; 3])`` (env_map.rs:163-168): the furnace."""
        return HdrMap(np.full((1, 1, 3), l0, dtype=np.float32))


Environment = Union[Studio, HdrMap]


def camera_new(yaw: float, pitch: float, distance: float, fov_deg: float) -> Camera:
    """``Camera::new`` (optics/raytracer/camera.rs:63-93), evaluated by the library."""
    cam = Camera()
    check(_lib.load().gr_camera_new(yaw, pitch, distance, fov_deg, ctypes.byref(cam)), "gr_camera_new")
    return cam


@dataclass
class SceneRef:
    """``GpuSceneRef`` (gpu_backend.rs:85-99)."""

    camera: Camera
    width: int
    height: int
    planes: np.ndarray  # (P, 4) f32: normal xyz, d
    material: GemMaterial
    max_bounces: int
    environment: Environment
    facet_finishes: Optional[Sequence[int]] = None


class _Packed:
    """A ``gr_scene`` plus the buffers its pointers reference (kept alive together)."""

    def __init__(self, scene: SceneRef):
        self.planes = np.ascontiguousarray(scene.planes, dtype=np.float32).reshape(-1, 4)
        self.material = to_struct(scene.material)
        self.camera = scene.camera
        self.env = Env()
        self.map = None
        env = scene.environment
        if isinstance(env, Studio):
            t, s = PRESETS[env.preset]
            self.env.kind = _lib.GR_ENV_STUDIO
            self.env.temp_k, self.env.spot_mult = t, s
            self.env.exposure, self.env.light_yaw, self.env.light_pitch = (
                env.exposure, env.light_yaw, env.light_pitch)
        elif isinstance(env, HdrMap):
            self.map = np.ascontiguousarray(env.rgb, dtype=np.float32)
            if self.map.ndim != 3 or self.map.shape[2] != 3:
                raise ValueError("HdrMap.rgb must have shape (H, W, 3)")
            self.env.kind = _lib.GR_ENV_MAP
            self.env.map_h, self.env.map_w = self.map.shape[0], self.map.shape[1]
            self.env.map_rgb = self.map.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            if env.device is not None:
                self.env.map = env.device.handle
        else:
            raise TypeError(f"unknown environment {env!r}")
        self.finishes = None
        if scene.facet_finishes is not None:
            self.finishes = np.ascontiguousarray(scene.facet_finishes, dtype=np.uint32)
        self.scene = Scene()
        self.scene.camera = ctypes.pointer(self.camera)
        self.scene.width, self.scene.height = scene.width, scene.height
        self.scene.planes = self.planes.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        self.scene.n_planes = self.planes.shape[0]
        if self.finishes is not None:
            self.scene.finishes = self.finishes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self.scene.material = ctypes.pointer(self.material)
        self.scene.max_bounces = scene.max_bounces
        self.scene.env = ctypes.pointer(self.env)


class DeviceScene:
    """A scene uploaded once to HBM (``gr_scene_upload``)."""

    def __init__(self, backend: "GpuBackend", scene: SceneRef):
        self._backend = backend
        self.width, self.height = scene.width, scene.height
        packed = _Packed(scene)
        h = ctypes.c_void_p()
        check(backend._L.gr_scene_upload(backend._ctx, ctypes.byref(packed.scene), ctypes.byref(h)),
              "gr_scene_upload")
        self.handle = h

    def free(self):
        if self.handle:
            self._backend._L.gr_scene_free(self._backend._ctx, self.handle)
            self.handle = None


class EnvMapDevice:
    """A device-resident environment map (gr_env_map_upload); release() drops the caller's
    reference (scenes and jobs that use it hold their own)."""

    def __init__(self, backend: "GpuBackend", rgb: np.ndarray):
        self._backend = backend
        self.rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        h = ctypes.c_void_p()
        check(backend._L.gr_env_map_upload(backend._ctx, self.rgb.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                           self.rgb.shape[1], self.rgb.shape[0], ctypes.byref(h)),
              "gr_env_map_upload")
        self.handle = h

    def release(self):
        if self.handle:
            self._backend._L.gr_env_map_release(self._backend._ctx, self.handle)
            self.handle = None


class Film:
    """A device-resident XYZ film (``gr_film_create``); accumulates across renders."""

    def __init__(self, backend: "GpuBackend", width: int, height: int):
        self._backend = backend
        self.width, self.height = width, height
        h = ctypes.c_void_p()
        check(backend._L.gr_film_create(backend._ctx, width, height, ctypes.byref(h)), "gr_film_create")
        self.handle = h

    def clear(self):
        check(self._backend._L.gr_film_clear(self._backend._ctx, self.handle), "gr_film_clear")

    def download(self) -> np.ndarray:
        out = np.zeros((self.height * self.width, 3), dtype=np.float32)
        check(self._backend._L.gr_film_download(
            self._backend._ctx, self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.size),
            "gr_film_download")
        return out

    def free(self):
        if self.handle:
            self._backend._L.gr_film_free(self._backend._ctx, self.handle)
            self.handle = None


# PROFILE_ANGLES_DEG (color/metrics.rs:42-45) and f32::to_radians of each, in f32.
PROFILE_ANGLES_DEG = np.arange(0.0, 91.0, 5.0).astype(np.float32)
PROFILE_AZIMUTHS_DEG = np.array([0.0, 45.0, 90.0, 135.0], np.float32)
PROFILE_PITCHES_RAD = PROFILE_ANGLES_DEG * (np.float32(np.pi) / np.float32(180.0))


class GpuBackend:
    """``renderer::gpu_backend::GpuBackend`` over ``libgemray_hip.so``."""

    def __init__(self, ctx=None):
        self._ctx = ctx
        self._L = _lib.load() if ctx is not None else None
        self._mu = threading.Lock()

    @classmethod
    def acquire(cls, device: int = 0, chunk_tuples: int = 0, hbm_budget_bytes: int = 0) -> "GpuBackend":
        L = _lib.load()
        opts = Options(device, chunk_tuples, hbm_budget_bytes)
        h = ctypes.c_void_p()
        st = L.gr_open(ctypes.byref(h), ctypes.byref(opts))
        if st != _lib.GR_OK:
            return cls(None)
        return cls(h)

    @classmethod
    def disabled(cls) -> "GpuBackend":
        return cls(None)

    @property
    def available(self) -> bool:
        return self._ctx is not None

    def close(self):
        if self._ctx is not None:
            self._L.gr_close(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def adapter_label(self) -> Optional[str]:
        if self._ctx is None:
            return None
        buf = ctypes.create_string_buffer(256)
        self._L.gr_adapter_label(self._ctx, buf, 256)
        return buf.value.decode()

    def _require(self):
        if self._ctx is None:
            raise GrError(_lib.GR_DECLINE_NO_DEVICE, "no gfx950 device held")

    def accumulate(self, scene: SceneRef, sample_offset: int, spp: int, accum: np.ndarray) -> None:
        """Like ``try_accumulate`` but raises ``GrError`` instead of returning False."""
        self._require()
        if accum.dtype != np.float32 or not accum.flags["C_CONTIGUOUS"]:
            raise TypeError("accum must be a C-contiguous float32 array")
        n_pixels = scene.width * scene.height
        if accum.size != 3 * n_pixels:
            raise AssertionError("accumulation buffer must have one entry per pixel")
        packed = _Packed(scene)
        with self._mu:
            st = self._L.gr_accumulate(self._ctx, ctypes.byref(packed.scene), sample_offset, spp,
                                       accum.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n_pixels)
        check(st, "gr_accumulate")

    def try_accumulate(self, scene: SceneRef, sample_offset: int, spp: int, accum: np.ndarray) -> bool:
        if self._ctx is None:
            return False
        try:
            self.accumulate(scene, sample_offset, spp, accum)
        except GrError:
            return False
        return True

    def trace_samples(self, scene: SceneRef, pixel_begin: int, pixel_count: int, sample_offset: int,
                      spp: int, with_info: bool = True):
        """Per-tuple XYZ (and segments / termination / facet hash) for a pixel range."""
        self._require()
        n = pixel_count * spp
        xyz = np.zeros((n, 3), dtype=np.float32)
        info = np.zeros((n, 4), dtype=np.uint32) if with_info else None
        packed = _Packed(scene)
        with self._mu:
            st = self._L.gr_trace_samples(
                self._ctx, ctypes.byref(packed.scene), pixel_begin, pixel_count, sample_offset, spp,
                xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                info.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if info is not None else None)
        check(st, "gr_trace_samples")
        return xyz, info

    def upload(self, scene: SceneRef) -> DeviceScene:
        self._require()
        return DeviceScene(self, scene)

    def autotune(self, scene: DeviceScene, spp: int) -> dict:
        """gr_autotune: times probe renders of the scene over the trace configurations and
        caches the fastest for its class and size; returns it."""
        self._require()
        c = _lib.TuneChoice()
        with self._mu:
            st = self._L.gr_autotune(self._ctx, scene.handle, spp, ctypes.byref(c))
        check(st, "gr_autotune")
        return {name: getattr(c, name) for name, _ in _lib.TuneChoice._fields_}

    def set_autotune(self, enabled: bool) -> None:
        self._require()
        check(self._L.gr_set_autotune(self._ctx, 1 if enabled else 0), "gr_set_autotune")

    def render_tuned(self, scene: DeviceScene, sample_offset: int, spp: int, film: Film, choice: dict) -> None:
        """gr_render_tuned: gr_render with the given configuration (keys of gr_tune_choice)."""
        self._require()
        c = _lib.TuneChoice()
        for name, value in choice.items():
            if name not in ("probe_ms", "probes"):
                setattr(c, name, value)
        with self._mu:
            st = self._L.gr_render_tuned(self._ctx, scene.handle, sample_offset, spp, film.handle, ctypes.byref(c))
        check(st, "gr_render_tuned")

    def upload_env_map(self, rgb: np.ndarray) -> "EnvMapDevice":
        """Uploads an environment map once; HdrMap(rgb, device=...) scenes reference it."""
        self._require()
        return EnvMapDevice(self, rgb)

    def film(self, width: int, height: int) -> Film:
        self._require()
        return Film(self, width, height)

    def render(self, scene: DeviceScene, sample_offset: int, spp: int, film: Film) -> None:
        self._require()
        with self._mu:
            st = self._L.gr_render(self._ctx, scene.handle, sample_offset, spp, film.handle)
        check(st, "gr_render")

    # ---- batched / served jobs (gr_job_*) -------------------------------------------

    def submit(self, scene: SceneRef, sample_offset: int, spp: int, subbatch_ms: int = 0,
               subbatch_samples: int = 0, subbatch_fixed: bool = False) -> int:
        """Queues a render job (a RenderRequest: scene, first_sample, samples); returns its id.
        With subbatch_ms > 0 the job is traced in contiguous sample sub-batches sized toward
        that many milliseconds each (from subbatch_samples, or always subbatch_samples when
        subbatch_fixed), whose sums ``take_delta`` drains as the worker's tracer does.
        Raises GrError on a decline (GR_DECLINE_*) or invalid input."""
        self._require()
        packed = _Packed(scene)
        desc = JobDesc(ctypes.pointer(packed.scene), sample_offset, spp, subbatch_ms, subbatch_samples,
                       1 if subbatch_fixed else 0)
        jid = ctypes.c_uint64()
        check(self._L.gr_job_submit(self._ctx, ctypes.byref(desc), ctypes.byref(jid)), "gr_job_submit")
        return int(jid.value)

    def poll(self, job: int) -> int:
        """GR_OK (done), GR_PENDING, GR_CANCELLED, GR_ERR_NOT_FOUND or the job's error."""
        self._require()
        return int(self._L.gr_job_poll(self._ctx, job))

    def wait(self, job: int, timeout_ms: int = -1) -> int:
        self._require()
        return int(self._L.gr_job_wait(self._ctx, job, timeout_ms))

    def cancel(self, job: int) -> None:
        self._require()
        check(self._L.gr_job_cancel(self._ctx, job), "gr_job_cancel")

    def job_film(self, job: int, accum: np.ndarray) -> None:
        """ADDS the finished job's film into accum (float32, width*height*3)."""
        self._require()
        if accum.dtype != np.float32 or not accum.flags["C_CONTIGUOUS"]:
            raise TypeError("accum must be a C-contiguous float32 array")
        check(self._L.gr_job_film(self._ctx, job, accum.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                  accum.size // 3), "gr_job_film")

    def job_info(self, job: int) -> dict:
        self._require()
        s = JobStats()
        check(self._L.gr_job_info(self._ctx, job, ctypes.byref(s)), "gr_job_info")
        return {
            "state": _lib.JOB_STATES[s.state], "width": s.width, "height": s.height,
            "sample_offset": s.sample_offset, "spp": s.spp, "samples_done": s.samples_done,
            "subbatches": s.subbatches, "footprint_bytes": int(s.footprint_bytes),
            "paths": int(s.paths), "events": {n: int(s.events[i]) for i, n in enumerate(_lib.EV_NAMES)},
            "queue_ms": s.queue_ms, "run_ms": s.run_ms, "latency_ms": s.latency_ms,
        }

    def take_delta(self, job: int, n_pixels: int):
        """PendingDelta::take: (first_sample, samples, xyz[n_pixels, 3]) summed over the
        sub-batches finished since the last take, or None when nothing is pending."""
        self._require()
        out = np.zeros((n_pixels, 3), np.float32)
        first, n = ctypes.c_uint32(), ctypes.c_uint32()
        check(self._L.gr_job_take_delta(self._ctx, job, ctypes.byref(first), ctypes.byref(n),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n_pixels),
              "gr_job_take_delta")
        return None if n.value == 0 else (int(first.value), int(n.value), out)

    def release(self, job: int) -> None:
        self._require()
        check(self._L.gr_job_release(self._ctx, job), "gr_job_release")

    def metrics(self) -> dict:
        self._require()
        m = ServeMetrics()
        check(self._L.gr_metrics(self._ctx, ctypes.byref(m)), "gr_metrics")
        return {name: (float(getattr(m, name)) if isinstance(getattr(m, name), float) else int(getattr(m, name)))
                for name, _ in ServeMetrics._fields_}

    def metrics_reset(self) -> None:
        self._require()
        check(self._L.gr_metrics_reset(self._ctx), "gr_metrics_reset")

    def set_tick_tuples(self, tuples: int) -> None:
        self._require()
        check(self._L.gr_serve_set_tick_tuples(self._ctx, tuples), "gr_serve_set_tick_tuples")

    def set_policy(self, policy: int, slack: float = 0.25) -> None:
        """How ticks are shared among running jobs: _lib.GR_SERVE_DEADLINE (deadline =
        submit + slack x estimated service) or _lib.GR_SERVE_EQUAL_SHARE."""
        self._require()
        check(self._L.gr_serve_set_policy(self._ctx, policy, slack), "gr_serve_set_policy")

    def probe(self, fn: int, inputs: np.ndarray) -> np.ndarray:
        """Evaluates device function fn (_lib.GR_PROBE_*) on rows of up to 9 float32 inputs;
        returns (n, 3) float32 outputs (a test probe of the kernel's own functions)."""
        self._require()
        # Rows travel as raw 32-bit words: uint32 inputs (the integer probes) keep their bits.
        inputs = np.asarray(inputs)
        bits = inputs.astype(np.uint32) if inputs.dtype.kind in "ui" else inputs.astype(np.float32).view(np.uint32)
        x = np.zeros((len(inputs), 9), np.uint32)
        x[:, :inputs.shape[1]] = bits
        out = np.zeros((len(inputs), 3), np.float32)
        check(self._L.gr_probe(self._ctx, fn, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(x),
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "gr_probe")
        return out

    def guide_buffers(self, camera: Camera, width: int, height: int, planes: np.ndarray):
        """The denoiser's primary-hit guide buffers (generate_guide_buffers,
        apps/diagram-gui/src/bridge/guide_pass.rs:37-110): per pixel the hit distance of the
        pixel-corner camera ray, the hit facet's normal and its index (1e6, zeros and -1 on a
        miss).  Returns (depth (h*w,), normal (h*w, 3), facet (h*w,) int32)."""
        self._require()
        p = np.ascontiguousarray(planes, dtype=np.float32).reshape(-1, 4)
        n = width * height
        depth = np.zeros(n, np.float32)
        normal = np.zeros((n, 3), np.float32)
        facet = np.zeros(n, np.int32)
        check(self._L.gr_guide_buffers(self._ctx, ctypes.byref(camera), width, height,
                                       p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.shape[0],
                                       depth.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       normal.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       facet.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "gr_guide_buffers")
        return depth, normal, facet

    def optical_metrics(self, planes: np.ndarray, material: GemMaterial, poses) -> list:
        """evaluate_gem_optical_metrics (color/metrics.rs:1082-1248) for each pose
        (cam_yaw, cam_pitch, light_yaw, light_pitch) of one stone, in one device call.
        Returns one dict per pose with brilliance_pct, fire_index, scintillation_pct,
        windowing_pct and extinction_pct."""
        self._require()
        p = np.ascontiguousarray(planes, dtype=np.float32).reshape(-1, 4)
        poses = list(poses)
        arr = (_lib.MetricsPose * len(poses))()
        for i, q in enumerate(poses):
            arr[i].cam_yaw, arr[i].cam_pitch, arr[i].light_yaw, arr[i].light_pitch = (float(v) for v in q)
        out = (_lib.GemMetrics * len(poses))()
        m = to_struct(material)
        check(self._L.gr_optical_metrics(self._ctx, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.shape[0],
                                         ctypes.byref(m), arr, len(poses), out), "gr_optical_metrics")
        return [{n: float(getattr(o, n)) for n, _ in _lib.GemMetrics._fields_} for o in out]

    def optical_metrics_batch(self, stones) -> list:
        """gr_optical_metrics_batch: stones is a sequence of (planes, material, poses); every
        pose of every stone in one launch.  Returns one list of per-pose dicts per stone."""
        self._require()
        keep, arr = [], (_lib.MetricsStone * len(stones))()
        total = 0
        for s, (planes, material, poses) in enumerate(stones):
            p = np.ascontiguousarray(planes, dtype=np.float32).reshape(-1, 4)
            poses = list(poses)
            pa = (_lib.MetricsPose * len(poses))()
            for i, q in enumerate(poses):
                pa[i].cam_yaw, pa[i].cam_pitch, pa[i].light_yaw, pa[i].light_pitch = (float(v) for v in q)
            m = to_struct(material)
            keep += [p, pa, m]
            arr[s].planes = p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            arr[s].n_planes = p.shape[0]
            arr[s].material = ctypes.cast(ctypes.byref(m), ctypes.c_void_p)
            arr[s].poses = pa
            arr[s].n_poses = len(poses)
            total += len(poses)
        out = (_lib.GemMetrics * total)()
        with self._mu:
            st = self._L.gr_optical_metrics_batch(self._ctx, arr, len(stones), out)
        check(st, "gr_optical_metrics_batch")
        res, k = [], 0
        for _, _, poses in stones:
            n = len(list(poses))
            res.append([{f: float(getattr(o, f)) for f, _ in _lib.GemMetrics._fields_} for o in out[k:k + n]])
            k += n
        return res

    def gem_optical_metrics(self, planes, material, cam_yaw, cam_pitch, light_yaw, light_pitch) -> dict:
        """One pose of ``optical_metrics`` (the reference's evaluate_gem_optical_metrics)."""
        return self.optical_metrics(planes, material, [(cam_yaw, cam_pitch, light_yaw, light_pitch)])[0]

    def angular_profile_at_azimuth(self, planes, material, cam_yaw, light_yaw, light_pitch):
        """evaluate_angular_profile_at_azimuth (color/metrics.rs:1262-1290): brilliance,
        extinction and windowing at camera pitches 0, 5, ..., 90 degrees (one device call)."""
        ms = self.optical_metrics(planes, material,
                                  [(cam_yaw, float(p), light_yaw, light_pitch) for p in PROFILE_PITCHES_RAD])
        return ([m["brilliance_pct"] for m in ms], [m["extinction_pct"] for m in ms],
                [m["windowing_pct"] for m in ms])

    def angular_profile(self, planes, material, light_yaw, light_pitch):
        """evaluate_angular_profile (color/metrics.rs:1298-1306): the canonical azimuth 0."""
        return self.angular_profile_at_azimuth(planes, material, 0.0, light_yaw, light_pitch)

    def set_trace_mode(self, mode: str) -> None:
        """"plain" (one lane per tuple) or "wavefront" (persistent lanes + finish pass)."""
        self._require()
        code = {"plain": _lib.GR_TRACE_PLAIN, "wavefront": _lib.GR_TRACE_WAVEFRONT}[mode]
        check(self._L.gr_set_trace_mode(self._ctx, code), "gr_set_trace_mode")

    def set_split_bounce(self, split_bounce: int) -> None:
        """Plain mode: park paths running after this many segments for a second launch (0 = never)."""
        self._require()
        check(self._L.gr_set_split_bounce(self._ctx, split_bounce), "gr_set_split_bounce")

    def set_hero_tables(self, mode: int) -> None:
        """0: evaluate every path's spectral functions; 1 (default): read the hero tables,
        building a scene's material table for traces of >= 2^26 tuples; 2: also for smaller
        traces.  Results are bit-identical in every mode."""
        self._require()
        check(self._L.gr_set_hero_tables(self._ctx, int(mode)), "gr_set_hero_tables")

    def set_chunk_tuples(self, tuples: int) -> None:
        self._require()
        check(self._L.gr_set_chunk_tuples(self._ctx, tuples), "gr_set_chunk_tuples")

    def last_stats(self) -> dict:
        self._require()
        s = Stats()
        check(self._L.gr_last_stats(self._ctx, ctypes.byref(s)), "gr_last_stats")
        return {
            "events": {name: int(s.events[i]) for i, name in enumerate(_lib.EV_NAMES)},
            "segments": int(s.events[0]),
            "paths": int(s.paths),
            "launches": int(s.launches),
            "trace_kernel_ms": float(s.trace_kernel_ms),
            "finish_kernel_ms": float(s.finish_kernel_ms),
            "total_ms": float(s.total_ms),
            "n_planes": int(s.n_planes),
            "spectral_table": bool(s.hero_tables & 1),
            "material_table": bool(s.hero_tables & 2),
        }
