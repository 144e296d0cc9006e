"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

OR_EV_NAMES = [
    "segment", "miss", "hit", "rotate", "absorb", "tir", "reflect", "refract",
    "entry_aniso", "walkoff", "rr", "scatter",
]
OR_EV_COUNT = len(OR_EV_NAMES)


class OrEnv(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("temp_k", ctypes.c_float),
        ("spot_mult", ctypes.c_float),
        ("exposure", ctypes.c_float),
        ("light_yaw", ctypes.c_float),
        ("light_pitch", ctypes.c_float),
        ("map_w", ctypes.c_int32),
        ("map_h", ctypes.c_int32),
        ("map_rgb", ctypes.POINTER(ctypes.c_float)),
    ]


class OrCamera(ctypes.Structure):
    _fields_ = [
        ("origin", ctypes.c_float * 3),
        ("forward", ctypes.c_float * 3),
        ("right", ctypes.c_float * 3),
        ("up", ctypes.c_float * 3),
        ("fov_tan", ctypes.c_float),
    ]


class OrPathInfo(ctypes.Structure):
    _fields_ = [
        ("segments", ctypes.c_uint32),
        ("termination", ctypes.c_uint32),
        ("facet_hash", ctypes.c_uint32),
        ("fate_hash", ctypes.c_uint32),
    ]


PATH_INFO_DTYPE = np.dtype([("segments", "<u4"), ("termination", "<u4"), ("facet_hash", "<u4"),
                            ("fate_hash", "<u4")])

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class use_library:
    """Context manager: route pyoracle calls to another oracle build (e.g. liboracle_cr.so,
    the restatement with correctly rounded transcendentals)."""

    def __init__(self, name: str):
        self.path = os.path.join(HERE, name)

    def __enter__(self):
        global _lib
        lib()
        self.saved = _lib
        alt = ctypes.CDLL(self.path)
        _declare(alt)
        _lib = alt
        return alt

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def _declare(L):
    fp = ctypes.POINTER(ctypes.c_float)
    u32 = ctypes.c_uint32
    L.or_hash_u32.argtypes = [u32]
    L.or_hash_u32.restype = u32
    L.or_stream_draw.argtypes = [u32, u32, u32]
    L.or_stream_draw.restype = ctypes.c_float
    L.or_low_discrepancy_base2.argtypes = [u32]
    L.or_low_discrepancy_base2.restype = ctypes.c_float
    L.or_cranley_patterson_rotate.argtypes = [ctypes.c_float, ctypes.c_float]
    L.or_cranley_patterson_rotate.restype = ctypes.c_float
    L.or_rgb_to_spectral_radiance.argtypes = [fp, ctypes.c_float]
    L.or_rgb_to_spectral_radiance.restype = ctypes.c_float
    u32p = ctypes.POINTER(u32)
    L.or_rng_records.argtypes = [ctypes.c_size_t, u32p, u32p, u32p, u32p, ctypes.c_int, u32p]
    L.or_rng_records.restype = None
    L.or_sample_draws.argtypes = [u32, u32, ctypes.POINTER(u32), fp]
    L.or_radical_inverse.argtypes = [u32, u32]
    L.or_radical_inverse.restype = ctypes.c_float
    L.or_wrapped_hero_wavelengths.argtypes = [ctypes.c_float, fp]
    L.or_exp_f32x8.argtypes = [fp, fp]
    L.or_cie_1931_cmf.argtypes = [ctypes.c_float, fp]
    L.or_cie_1931_cmf_x8.argtypes = [fp, fp]
    L.or_dispersion_evaluate.argtypes = [ctypes.c_void_p, ctypes.c_float]
    L.or_dispersion_evaluate.restype = ctypes.c_float
    L.or_blackbody.argtypes = [ctypes.c_float, ctypes.c_float]
    L.or_blackbody.restype = ctypes.c_float
    L.or_white_balance.argtypes = [ctypes.c_float, fp]
    L.or_sample_studio_environment.argtypes = [fp] + [ctypes.c_float] * 6
    L.or_sample_studio_environment.restype = ctypes.c_float
    L.or_env_map_radiance.argtypes = [ctypes.POINTER(OrEnv), fp, ctypes.c_float]
    L.or_env_map_radiance.restype = ctypes.c_float
    L.or_analytic_furnace_target.argtypes = [ctypes.c_float, fp]
    L.or_standard_round_brilliant.argtypes = [fp]
    L.or_standard_round_brilliant.restype = ctypes.c_int
    L.or_plane_new.argtypes = [fp, ctypes.c_float, fp]
    L.or_camera_new.argtypes = [ctypes.c_float] * 4 + [ctypes.POINTER(OrCamera)]
    L.or_generate_ray.argtypes = [ctypes.POINTER(OrCamera)] + [ctypes.c_float] * 6 + [fp, fp]
    L.or_intersect.argtypes = [fp, ctypes.c_int, fp, fp, fp, fp, ctypes.POINTER(ctypes.c_int)]
    L.or_intersect.restype = ctypes.c_int
    L.or_intersect_scalar.argtypes = L.or_intersect.argtypes
    L.or_intersect_scalar.restype = ctypes.c_int
    L.or_effective_extraordinary_index.argtypes = [ctypes.c_float] * 3
    L.or_effective_extraordinary_index.restype = ctypes.c_float
    L.or_walk_off_angle.argtypes = [ctypes.c_float] * 3
    L.or_walk_off_angle.restype = ctypes.c_float
    L.or_neff_from_cos.argtypes = [ctypes.c_float] * 3
    L.or_neff_from_cos.restype = ctypes.c_float
    L.or_frame_rotation_cs.argtypes = [fp, fp, fp, fp]
    L.or_frame_rotation_cs.restype = None
    L.or_field_qform.argtypes = [ctypes.c_float, ctypes.c_float, fp, fp, ctypes.c_float, ctypes.c_float]
    L.or_field_qform.restype = ctypes.c_float
    L.or_degree_of_polarization.argtypes = [ctypes.c_float] * 4
    L.or_degree_of_polarization.restype = ctypes.c_float
    L.or_fresnel_companion.argtypes = [ctypes.c_float] * 4 + [fp]
    L.or_fresnel_companion.restype = None
    L.or_extraordinary_poynting_dir.argtypes = [fp, fp, ctypes.c_float, ctypes.c_float, fp]
    L.or_tir_phase_delta.argtypes = [ctypes.c_float] * 3
    L.or_tir_phase_delta.restype = ctypes.c_float
    L.or_signed_frame_rotation_psi.argtypes = [fp, fp, fp]
    L.or_signed_frame_rotation_psi.restype = ctypes.c_float
    L.or_material_biaxial_indicatrix.argtypes = [ctypes.c_void_p, ctypes.c_float, fp]
    L.or_material_biaxial_indicatrix.restype = ctypes.c_int
    L.or_biaxial_from_gamma_axis.argtypes = [ctypes.c_float] * 3 + [fp, fp]
    L.or_biaxial_wave_indices.argtypes = [fp, fp, fp]
    L.or_biaxial_eigen_polarizations.argtypes = [fp, fp, fp]
    L.or_biaxial_d_to_e.argtypes = [fp, fp, fp]
    L.or_biaxial_mode_poynting_dir.argtypes = [fp, fp, ctypes.c_int, fp]
    L.or_biaxial_resolve_entry_mode.argtypes = [fp, fp, fp, ctypes.c_float, ctypes.c_float, ctypes.c_int, fp]
    L.or_biaxial_resolve_entry_mode.restype = ctypes.c_float
    L.or_poynting_direction.argtypes = [fp, fp, fp]
    L.or_uniaxial_eigen_polarizations.argtypes = [fp, fp, fp]
    L.or_effective_pleochroic_alpha.argtypes = [fp, ctypes.c_int, fp, fp, fp, fp, ctypes.c_float]
    L.or_effective_pleochroic_alpha.restype = ctypes.c_float
    L.or_guide_buffers.argtypes = [ctypes.POINTER(OrCamera), u32, u32, fp, ctypes.c_int, fp, fp,
                                   ctypes.POINTER(ctypes.c_int32)]
    L.or_sample_hg_batch.argtypes = [ctypes.c_int, fp, fp, ctypes.c_float, fp, fp]
    L.or_channel_absorption_alphas.argtypes = [ctypes.c_void_p, fp, fp, fp, fp, fp]
    L.or_maybe_scatter_or_extinguish.argtypes = [fp, ctypes.c_float, ctypes.c_float, fp, ctypes.c_float,
                                                 ctypes.c_float, u32, u32, fp, fp, fp, fp]
    L.or_maybe_scatter_or_extinguish.restype = ctypes.c_int
    L.or_trace_spectral_ray.argtypes = [
        fp, fp, fp, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, u32, ctypes.POINTER(OrEnv), u32,
        ctypes.c_float, fp, ctypes.POINTER(OrPathInfo), ctypes.POINTER(ctypes.c_uint64),
    ]
    L.or_render.argtypes = [
        ctypes.POINTER(OrCamera), u32, u32, fp, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, u32,
        ctypes.POINTER(OrEnv), u32, u32, u32, u32, ctypes.c_int, fp, fp, ctypes.c_void_p,
        ctypes.POINTER(ctypes.c_uint64),
    ]
    L.or_render.restype = ctypes.c_uint64
    L.or_shading_normal_near_edge.argtypes = [fp, ctypes.c_int, fp, ctypes.c_int, fp, ctypes.c_float, fp]
    L.or_emerald_cut.argtypes = [fp]
    L.or_emerald_cut.restype = ctypes.c_int
    L.or_gem_optical_metrics.argtypes = [fp, ctypes.c_int, ctypes.c_void_p] + [ctypes.c_float] * 4 + [fp]
    L.or_metrics_items.argtypes = [fp, ctypes.c_int, ctypes.c_void_p] + [ctypes.c_float] * 4 + [
        ctypes.c_void_p, fp, fp, ctypes.c_void_p]
    L.or_metrics_reduce.argtypes = [ctypes.c_void_p, fp, fp, ctypes.c_void_p, fp]


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def f32arr(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32))


def bits(v) -> int:
    return int(np.float32(v).view(np.uint32))


# Studio presets (optics/raytracer/environment.rs:81-100): (temp_k, spot_mult).
PRESETS = {
    "Daylight": (6500.0, 1.0),
    "Incandescent": (3200.0, 1.2),
    "RingLights": (5000.0, 1.6),
    "DarkSpotlight": (6000.0, 2.4),
}


class Env:
    """Keeps the map buffer alive alongside the OrEnv struct."""

    def __init__(self, struct: OrEnv, keep=None):
        self.struct = struct
        self.keep = keep


def studio_env(preset: str, exposure: float, light_yaw: float, light_pitch: float) -> Env:
    t, s = PRESETS[preset]
    e = OrEnv()
    e.kind = 0
    e.temp_k, e.spot_mult, e.exposure, e.light_yaw, e.light_pitch = t, s, exposure, light_yaw, light_pitch
    return Env(e)


def map_env(rgb: np.ndarray) -> Env:
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w, _ = rgb.shape
    e = OrEnv()
    e.kind = 1
    e.map_w, e.map_h = w, h
    e.map_rgb = fptr(rgb)
    return Env(e, rgb)


def uniform_env(l0: float) -> Env:
    return map_env(np.full((1, 1, 3), l0, dtype=np.float32))


def standard_round_brilliant() -> np.ndarray:
    out = np.zeros((74, 4), dtype=np.float32)
    n = lib().or_standard_round_brilliant(fptr(out))
    assert n == 74
    return out


def emerald_cut() -> np.ndarray:
    out = np.zeros((34, 4), dtype=np.float32)
    n = lib().or_emerald_cut(fptr(out))
    assert n == 34
    return out


METRIC_NAMES = ("brilliance_pct", "fire_index", "scintillation_pct", "windowing_pct", "extinction_pct")
MX_ITEMS = 5 * 18 * 18


def gem_optical_metrics(planes, material_struct, cam_yaw, cam_pitch, light_yaw, light_pitch) -> np.ndarray:
    """evaluate_gem_optical_metrics restated: float32[5] in METRIC_NAMES order."""
    p = np.ascontiguousarray(planes, dtype=np.float32).reshape(-1, 4)
    out = np.zeros(5, dtype=np.float32)
    lib().or_gem_optical_metrics(fptr(p), p.shape[0], ctypes.byref(material_struct), cam_yaw, cam_pitch,
                                 light_yaw, light_pitch, fptr(out))
    return out


def metrics_items(planes, material_struct, cam_yaw, cam_pitch, light_yaw, light_pitch):
    """Per-ray records (cls int32, angle, weight, temporal uint8), MX_ITEMS each."""
    p = np.ascontiguousarray(planes, dtype=np.float32).reshape(-1, 4)
    cls = np.zeros(MX_ITEMS, np.int32)
    angle = np.zeros(MX_ITEMS, np.float32)
    weight = np.zeros(MX_ITEMS, np.float32)
    temporal = np.zeros(MX_ITEMS, np.uint8)
    lib().or_metrics_items(fptr(p), p.shape[0], ctypes.byref(material_struct), cam_yaw, cam_pitch, light_yaw,
                           light_pitch, cls.ctypes.data, fptr(angle), fptr(weight), temporal.ctypes.data)
    return cls, angle, weight, temporal


def metrics_reduce(cls, angle, weight, temporal) -> np.ndarray:
    out = np.zeros(5, dtype=np.float32)
    cls = np.ascontiguousarray(cls, np.int32)
    temporal = np.ascontiguousarray(temporal, np.uint8)
    lib().or_metrics_reduce(cls.ctypes.data, fptr(np.ascontiguousarray(angle, np.float32)),
                            fptr(np.ascontiguousarray(weight, np.float32)), temporal.ctypes.data, fptr(out))
    return out


class Indicatrix:
    """A BiaxialIndicatrix (optics/birefringence.rs:349-870) as the oracle's 12 floats."""

    def __init__(self, p):
        self.p = f32arr(p)
        assert self.p.shape == (12,)

    @property
    def n(self):
        return self.p[:3]

    @property
    def axes(self):
        """3 x 3, columns = the alpha, beta, gamma principal axes."""
        return self.p[3:].reshape(3, 3).T.copy()

    @classmethod
    def from_gamma_axis(cls, n_alpha, n_beta, n_gamma, gamma_axis):
        out = np.zeros(12, np.float32)
        lib().or_biaxial_from_gamma_axis(n_alpha, n_beta, n_gamma, fptr(f32arr(gamma_axis)), fptr(out))
        return cls(out)

    @classmethod
    def of_material(cls, material_struct, lambda_nm):
        out = np.zeros(12, np.float32)
        if not lib().or_material_biaxial_indicatrix(ctypes.byref(material_struct), lambda_nm, fptr(out)):
            return None
        return cls(out)

    def wave_indices(self, w):
        out = np.zeros(2, np.float32)
        lib().or_biaxial_wave_indices(fptr(self.p), fptr(f32arr(w)), fptr(out))
        return out

    def eigen_polarizations(self, w):
        out = np.zeros(6, np.float32)
        lib().or_biaxial_eigen_polarizations(fptr(self.p), fptr(f32arr(w)), fptr(out))
        return out[:3], out[3:]

    def d_to_e(self, d):
        out = np.zeros(3, np.float32)
        lib().or_biaxial_d_to_e(fptr(self.p), fptr(f32arr(d)), fptr(out))
        return out

    def mode_poynting_dir(self, w, want_slow):
        out = np.zeros(3, np.float32)
        lib().or_biaxial_mode_poynting_dir(fptr(self.p), fptr(f32arr(w)), int(want_slow), fptr(out))
        return out

    def resolve_entry_mode(self, incident, normal, cos_i, n_seed, want_slow):
        wave = np.zeros(3, np.float32)
        n = lib().or_biaxial_resolve_entry_mode(fptr(self.p), fptr(f32arr(incident)), fptr(f32arr(normal)),
                                                cos_i, n_seed, int(want_slow), fptr(wave))
        return np.float32(n), wave


def guide_buffers(camera: OrCamera, width, height, planes):
    """generate_guide_buffers (apps/diagram-gui/src/bridge/guide_pass.rs:37-110): (depth,
    normal (n, 3), facet int32) per pixel."""
    p = np.ascontiguousarray(planes, dtype=np.float32)
    n = width * height
    depth = np.zeros(n, np.float32)
    normal = np.zeros((n, 3), np.float32)
    facet = np.zeros(n, np.int32)
    lib().or_guide_buffers(ctypes.byref(camera), width, height, fptr(p), p.shape[0], fptr(depth), fptr(normal),
                           facet.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return depth, normal, facet


def poynting_direction(k, e) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().or_poynting_direction(fptr(f32arr(k)), fptr(f32arr(e)), fptr(out))
    return out


def uniaxial_eigen_polarizations(w, c):
    out = np.zeros(6, np.float32)
    lib().or_uniaxial_eigen_polarizations(fptr(f32arr(w)), fptr(f32arr(c)), fptr(out))
    return out[:3], out[3:]


def effective_pleochroic_alpha(alpha, trichroic, axis, e, ea, eb, dop) -> np.float32:
    return np.float32(lib().or_effective_pleochroic_alpha(
        fptr(f32arr(alpha)), int(trichroic), fptr(f32arr(axis)), fptr(f32arr(e)), fptr(f32arr(ea)),
        fptr(f32arr(eb)), dop))


def camera_new(yaw, pitch, distance, fov_deg) -> OrCamera:
    c = OrCamera()
    lib().or_camera_new(yaw, pitch, distance, fov_deg, ctypes.byref(c))
    return c


def _finishes(finishes, n):
    if finishes is None:
        return None
    f = np.ascontiguousarray(finishes, dtype=np.uint32)
    assert f.shape == (n,)
    return f


def trace(origin, direction, planes, material_struct, max_bounces, env: Env, seed, hero_rand,
          events: Optional[np.ndarray] = None, finishes=None):
    o = f32arr(origin)
    d = f32arr(direction)
    p = np.ascontiguousarray(planes, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    info = OrPathInfo()
    ev = None
    if events is not None:
        ev = events.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    fin = _finishes(finishes, p.shape[0])
    lib().or_trace_spectral_ray(
        fptr(o), fptr(d), fptr(p), p.shape[0], fin.ctypes.data if fin is not None else None,
        ctypes.byref(material_struct), max_bounces,
        ctypes.byref(env.struct), seed, hero_rand, fptr(out), ctypes.byref(info), ev,
    )
    return out, info


def render(camera: OrCamera, width, height, planes, material_struct, max_bounces, env: Env,
           first_sample, spp, accum: Optional[np.ndarray] = None, pixel_begin=0,
           pixel_end=None, threads=None, per_sample=False, events: Optional[np.ndarray] = None,
           finishes=None):
    """render_core::trace_into restated; returns (accum, per_sample_xyz, per_sample_info,
    total_segments)."""
    if pixel_end is None:
        pixel_end = width * height
    if threads is None:
        threads = os.cpu_count() or 1
    p = np.ascontiguousarray(planes, dtype=np.float32)
    if accum is None:
        accum = np.zeros((height * width, 3), dtype=np.float32)
    n = (pixel_end - pixel_begin) * spp
    xyz = np.zeros((n, 3), dtype=np.float32) if per_sample else None
    info = np.zeros(n, dtype=PATH_INFO_DTYPE) if per_sample else None
    ev = None
    if events is not None:
        ev = events.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    fin = _finishes(finishes, p.shape[0])
    segs = lib().or_render(
        ctypes.byref(camera), width, height, fptr(p), p.shape[0], fin.ctypes.data if fin is not None else None,
        ctypes.byref(material_struct),
        max_bounces, ctypes.byref(env.struct), first_sample, spp, pixel_begin, pixel_end, threads,
        fptr(accum), fptr(xyz) if xyz is not None else None,
        info.ctypes.data if info is not None else None, ev,
    )
    return accum, xyz, info, int(segs)


def rng_records(pixel, sample, bounce, salts) -> np.ndarray:
    """or_rng_records: (n, 8 + len(salts)) uint32 -- seed, the three pixel-rotation hashes,
    the sample's reversed bits, the bit patterns of jx, jy, hero and of each salt's
    per-bounce draw (renderer/gpu/rng_check.rs:1-35)."""
    p = np.ascontiguousarray(pixel, np.uint32)
    s = np.ascontiguousarray(sample, np.uint32)
    b = np.ascontiguousarray(bounce, np.uint32)
    sl = np.ascontiguousarray(salts, np.uint32)
    out = np.zeros((len(p), 8 + len(sl)), np.uint32)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib().or_rng_records(len(p), p.ctypes.data_as(u32p), s.ctypes.data_as(u32p), b.ctypes.data_as(u32p),
                         sl.ctypes.data_as(u32p), len(sl), out.ctypes.data_as(u32p))
    return out


def rgb_to_spectral_radiance(rgb, lambda_nm) -> float:
    """rgb_to_spectral_radiance (renderer/env_map_spectrum.rs:60-82)."""
    return lib().or_rgb_to_spectral_radiance(fptr(f32arr(rgb)), float(lambda_nm))
