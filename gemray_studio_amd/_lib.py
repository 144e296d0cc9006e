"""ctypes binding of libgemray_hip.so (include/gemray_hip.h).

The library is the only compute path: if it is missing or fails to load, every entry
point raises; there is no CPU fallback in this package.
"""
from __future__ import annotations

import ctypes
import os

from .materials import MaterialStruct

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgemray_hip.so")

GR_OK = 0
GR_DECLINE_NO_DEVICE = 1
GR_DECLINE_ENV = 2
GR_DECLINE_MATERIAL = 3
GR_DECLINE_BUDGET = 4
GR_ERR_INVALID = 5
GR_ERR_DEVICE = 6
GR_ERR_NOT_FOUND = 7
GR_PENDING = 8
GR_CANCELLED = 9
STATUS_NAMES = {
    0: "GR_OK", 1: "GR_DECLINE_NO_DEVICE", 2: "GR_DECLINE_ENV", 3: "GR_DECLINE_MATERIAL",
    4: "GR_DECLINE_BUDGET", 5: "GR_ERR_INVALID", 6: "GR_ERR_DEVICE", 7: "GR_ERR_NOT_FOUND",
    8: "GR_PENDING", 9: "GR_CANCELLED",
}
JOB_STATES = ["queued", "running", "done", "cancelled", "failed"]

GR_ENV_STUDIO = 0
GR_ENV_MAP = 1
GR_TRACE_PLAIN = 0
GR_TRACE_WAVEFRONT = 1
EV_NAMES = ["segment", "miss", "hit", "rotate", "absorb", "tir", "reflect", "refract",
            "entry_aniso", "walkoff", "rr", "scatter", "cert_miss", "exterior_miss", "first_miss",
            "absorption_fill"]
GR_EV_COUNT = len(EV_NAMES)


class GrError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)}")
        self.status = status


class Options(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("chunk_tuples", ctypes.c_uint64),
                ("hbm_budget_bytes", ctypes.c_uint64)]


class Camera(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("forward", ctypes.c_float * 3),
                ("right", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("fov_tan", ctypes.c_float)]


class Env(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("temp_k", ctypes.c_float), ("spot_mult", ctypes.c_float),
                ("exposure", ctypes.c_float), ("light_yaw", ctypes.c_float),
                ("light_pitch", ctypes.c_float), ("map_w", ctypes.c_int32),
                ("map_h", ctypes.c_int32), ("map_rgb", ctypes.POINTER(ctypes.c_float)),
                ("map", ctypes.c_void_p)]


class Scene(ctypes.Structure):
    _fields_ = [("camera", ctypes.POINTER(Camera)), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("planes", ctypes.POINTER(ctypes.c_float)),
                ("n_planes", ctypes.c_uint32), ("finishes", ctypes.POINTER(ctypes.c_uint32)),
                ("material", ctypes.POINTER(MaterialStruct)), ("max_bounces", ctypes.c_uint32),
                ("env", ctypes.POINTER(Env))]


class Stats(ctypes.Structure):
    _fields_ = [("events", ctypes.c_uint64 * GR_EV_COUNT), ("paths", ctypes.c_uint64),
                ("launches", ctypes.c_uint32), ("trace_kernel_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("n_planes", ctypes.c_uint32),
                ("finish_kernel_ms", ctypes.c_double), ("hero_tables", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class JobDesc(ctypes.Structure):
    _fields_ = [("scene", ctypes.POINTER(Scene)), ("sample_offset", ctypes.c_uint32),
                ("spp", ctypes.c_uint32), ("subbatch_ms", ctypes.c_uint32), ("subbatch_samples", ctypes.c_uint32),
                ("subbatch_fixed", ctypes.c_uint32)]


class JobStats(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("sample_offset", ctypes.c_uint32), ("spp", ctypes.c_uint32),
                ("samples_done", ctypes.c_uint32), ("subbatches", ctypes.c_uint32),
                ("footprint_bytes", ctypes.c_uint64), ("paths", ctypes.c_uint64),
                ("events", ctypes.c_uint64 * GR_EV_COUNT), ("queue_ms", ctypes.c_double),
                ("run_ms", ctypes.c_double), ("latency_ms", ctypes.c_double)]


class MetricsPose(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("cam_yaw", "cam_pitch", "light_yaw", "light_pitch")]


class GemMetrics(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in (
        "brilliance_pct", "fire_index", "scintillation_pct", "windowing_pct", "extinction_pct")]


class MetricsStone(ctypes.Structure):
    _fields_ = [("planes", ctypes.POINTER(ctypes.c_float)), ("n_planes", ctypes.c_uint32),
                ("material", ctypes.c_void_p), ("poses", ctypes.POINTER(MetricsPose)), ("n_poses", ctypes.c_uint32)]


class ServeMetrics(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "submitted", "admitted", "completed", "cancelled", "failed", "declined", "queued", "running",
        "paths", "segments", "ticks", "launches")] + [
        (n, ctypes.c_double) for n in (
            "busy_ms", "wall_ms", "latency_p50_ms", "latency_p99_ms", "latency_mean_ms",
            "latency_max_ms")] + [
        (n, ctypes.c_uint64) for n in ("hbm_in_use", "hbm_budget", "tick_tuples")] + [
        ("policy", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("slack", ctypes.c_double),
        ("rate_tuples_per_ms", ctypes.c_double)]


class TuneChoice(ctypes.Structure):
    _fields_ = [("waves_per_eu", ctypes.c_uint32), ("split_bounce", ctypes.c_uint32), ("cont_div", ctypes.c_uint32),
                ("trace_mode", ctypes.c_uint32), ("chunk_tuples", ctypes.c_uint64), ("probe_ms", ctypes.c_double),
                ("probes", ctypes.c_uint32), ("park_lanes", ctypes.c_uint32)]


# Every symbol include/gemray_hip.h declares.
EXPORTS = [
    "gr_open", "gr_close", "gr_adapter_label", "gr_set_chunk_tuples", "gr_accumulate",
    "gr_trace_samples", "gr_last_stats", "gr_scene_upload", "gr_scene_free", "gr_film_create",
    "gr_film_clear", "gr_film_free", "gr_film_download", "gr_render", "gr_render_noalloc", "gr_camera_new", "gr_env_map_upload", "gr_env_map_release",
    "gr_autotune", "gr_set_autotune", "gr_render_tuned", "gr_optical_metrics_batch",
    "gr_set_trace_mode", "gr_job_submit", "gr_job_poll", "gr_job_wait", "gr_job_cancel",
    "gr_job_film", "gr_job_info", "gr_job_release", "gr_metrics", "gr_metrics_reset",
    "gr_serve_set_tick_tuples", "gr_probe", "gr_set_split_bounce", "gr_set_hero_tables", "gr_serve_set_policy",
    "gr_optical_metrics", "gr_job_take_delta", "gr_guide_buffers",
]
GR_SERVE_EQUAL_SHARE, GR_SERVE_DEADLINE = 0, 1
(GR_PROBE_NEFF, GR_PROBE_POYNTING, GR_PROBE_TIR_RETARDATION, GR_PROBE_FRAME_ROTATION, GR_PROBE_FIELD_QFORM,
 GR_PROBE_EXP, GR_PROBE_CMF, GR_PROBE_ACOS, GR_PROBE_BX_WAVE_INDICES, GR_PROBE_BX_EIGEN_SLOW, GR_PROBE_BX_EIGEN_FAST,
 GR_PROBE_BX_POYNTING_SLOW, GR_PROBE_BX_POYNTING_FAST, GR_PROBE_SAMPLE_DRAWS, GR_PROBE_SEED, GR_PROBE_STREAM_DRAW,
 GR_PROBE_RGB_SPECTRAL, GR_PROBE_DOP, GR_PROBE_FRESNEL_COMPANION) = range(19)

_lib = None


def load():
    """Loads the HIP library; raises if it was not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m gemray_studio_amd.build_lib` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    u32 = ctypes.c_uint32
    fp = ctypes.POINTER(ctypes.c_float)
    L.gr_open.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Options)]
    L.gr_open.restype = ctypes.c_int
    L.gr_close.argtypes = [vp]
    L.gr_close.restype = None
    L.gr_adapter_label.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    L.gr_adapter_label.restype = ctypes.c_int
    L.gr_set_chunk_tuples.argtypes = [vp, ctypes.c_uint64]
    L.gr_set_chunk_tuples.restype = ctypes.c_int
    L.gr_accumulate.argtypes = [vp, ctypes.POINTER(Scene), u32, u32, fp, ctypes.c_size_t]
    L.gr_accumulate.restype = ctypes.c_int
    L.gr_trace_samples.argtypes = [vp, ctypes.POINTER(Scene), u32, u32, u32, u32, fp,
                                   ctypes.POINTER(ctypes.c_uint32)]
    L.gr_trace_samples.restype = ctypes.c_int
    L.gr_last_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.gr_last_stats.restype = ctypes.c_int
    L.gr_scene_upload.argtypes = [vp, ctypes.POINTER(Scene), ctypes.POINTER(vp)]
    L.gr_scene_upload.restype = ctypes.c_int
    L.gr_scene_free.argtypes = [vp, vp]
    L.gr_scene_free.restype = None
    L.gr_film_create.argtypes = [vp, u32, u32, ctypes.POINTER(vp)]
    L.gr_film_create.restype = ctypes.c_int
    L.gr_film_clear.argtypes = [vp, vp]
    L.gr_film_clear.restype = ctypes.c_int
    L.gr_film_free.argtypes = [vp, vp]
    L.gr_film_free.restype = None
    L.gr_film_download.argtypes = [vp, vp, fp, ctypes.c_size_t]
    L.gr_film_download.restype = ctypes.c_int
    L.gr_render.argtypes = [vp, vp, u32, u32, vp]
    L.gr_render.restype = ctypes.c_int
    L.gr_optical_metrics_batch.argtypes = [vp, ctypes.POINTER(MetricsStone), u32, ctypes.POINTER(GemMetrics)]
    L.gr_optical_metrics_batch.restype = ctypes.c_int
    L.gr_autotune.argtypes = [vp, vp, u32, ctypes.POINTER(TuneChoice)]
    L.gr_autotune.restype = ctypes.c_int
    L.gr_set_autotune.argtypes = [vp, ctypes.c_int]
    L.gr_set_autotune.restype = ctypes.c_int
    L.gr_render_tuned.argtypes = [vp, vp, u32, u32, vp, ctypes.POINTER(TuneChoice)]
    L.gr_render_tuned.restype = ctypes.c_int
    L.gr_env_map_upload.argtypes = [vp, fp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(vp)]
    L.gr_env_map_upload.restype = ctypes.c_int
    L.gr_env_map_release.argtypes = [vp, vp]
    L.gr_env_map_release.restype = None
    L.gr_render_noalloc.argtypes = [vp, vp, u32, u32, vp]
    L.gr_render_noalloc.restype = ctypes.c_int
    L.gr_camera_new.argtypes = [ctypes.c_float] * 4 + [ctypes.POINTER(Camera)]
    L.gr_camera_new.restype = ctypes.c_int
    L.gr_set_trace_mode.argtypes = [vp, ctypes.c_int]
    L.gr_set_trace_mode.restype = ctypes.c_int
    u64 = ctypes.c_uint64
    L.gr_job_submit.argtypes = [vp, ctypes.POINTER(JobDesc), ctypes.POINTER(u64)]
    L.gr_job_poll.argtypes = [vp, u64]
    L.gr_job_wait.argtypes = [vp, u64, ctypes.c_int64]
    L.gr_job_cancel.argtypes = [vp, u64]
    L.gr_job_film.argtypes = [vp, u64, fp, ctypes.c_size_t]
    L.gr_job_info.argtypes = [vp, u64, ctypes.POINTER(JobStats)]
    L.gr_job_release.argtypes = [vp, u64]
    L.gr_job_take_delta.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32), fp,
                                    ctypes.c_size_t]
    L.gr_job_take_delta.restype = ctypes.c_int
    L.gr_metrics.argtypes = [vp, ctypes.POINTER(ServeMetrics)]
    L.gr_metrics_reset.argtypes = [vp]
    L.gr_serve_set_tick_tuples.argtypes = [vp, u64]
    L.gr_serve_set_policy.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    L.gr_set_split_bounce.argtypes = [vp, ctypes.c_uint32]
    L.gr_set_split_bounce.restype = ctypes.c_int
    L.gr_set_hero_tables.argtypes = [vp, ctypes.c_int]
    L.gr_set_hero_tables.restype = ctypes.c_int
    L.gr_optical_metrics.argtypes = [vp, fp, ctypes.c_uint32, ctypes.POINTER(MaterialStruct),
                                     ctypes.POINTER(MetricsPose), ctypes.c_uint32, ctypes.POINTER(GemMetrics)]
    L.gr_optical_metrics.restype = ctypes.c_int
    L.gr_probe.argtypes = [vp, ctypes.c_int, fp, ctypes.c_size_t, fp]
    L.gr_probe.restype = ctypes.c_int
    L.gr_guide_buffers.argtypes = [vp, ctypes.POINTER(Camera), ctypes.c_uint32, ctypes.c_uint32, fp, ctypes.c_uint32,
                                   fp, fp, ctypes.POINTER(ctypes.c_int32)]
    L.gr_guide_buffers.restype = ctypes.c_int
    for name in ("gr_job_submit", "gr_job_poll", "gr_job_wait", "gr_job_cancel", "gr_job_film",
                 "gr_job_info", "gr_job_release", "gr_metrics", "gr_metrics_reset",
                 "gr_serve_set_tick_tuples", "gr_serve_set_policy"):
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(status: int, what: str) -> None:
    if status != GR_OK:
        raise GrError(status, what)
