"""Seeded scene generators for the benchmark and parity workloads.

The reference ships fixed cuts (``crates/gemray/src/geometry/cuts.rs``) and built-in
materials (``optics/materials.rs``); the workloads here need parametric and seeded ones
(a parametric round brilliant, an oval brilliant, random convex stones, a 233-facet step
cut, a pleochroic uniaxial material and a procedural HDR environment).  Every generator is
deterministic in its seed.  Geometry is computed in float64 and stored as float32 planes
``(normal.x, normal.y, normal.z, d)`` with unit normals and the stone on the ``n.x + d <= 0``
side, the convention of ``geometry/plane.rs:3-20``.
"""
from __future__ import annotations

import math
from dataclasses import replace
from typing import List, Sequence, Tuple

import numpy as np

from .materials import DISP_SELLMEIER3, GemMaterial, by_name

Plane = Tuple[float, float, float, float]


def _plane(n, point) -> Plane:
    n = np.asarray(n, dtype=np.float64)
    n = n / np.linalg.norm(n)
    d = -float(np.dot(n, np.asarray(point, dtype=np.float64)))
    return (float(n[0]), float(n[1]), float(n[2]), d)


def _to_array(planes: Sequence[Plane]) -> np.ndarray:
    a = np.asarray(planes, dtype=np.float64)
    n = a[:, :3] / np.linalg.norm(a[:, :3], axis=1, keepdims=True)
    out = np.empty((a.shape[0], 4), dtype=np.float32)
    out[:, :3] = n
    out[:, 3] = a[:, 3]
    return out


def _tier(angle_deg: float, count: int, phase: float, point_r: float, point_y: float,
          upper: bool) -> List[Plane]:
    """``count`` facets at polar ``angle_deg`` from the vertical, azimuths
    ``phase + 2 pi i / count``, each passing through (point_r, point_y) in its azimuth."""
    a = math.radians(angle_deg)
    out = []
    for i in range(count):
        phi = phase + 2.0 * math.pi * i / count
        ny = math.cos(a) if upper else -math.cos(a)
        n = (math.sin(a) * math.cos(phi), ny, math.sin(a) * math.sin(phi))
        p = (point_r * math.cos(phi), point_y, point_r * math.sin(phi))
        out.append(_plane(n, p))
    return out


def round_brilliant(table: float = 0.57, crown_deg: float = 34.5, pavilion_deg: float = 40.75,
                    girdle: float = 0.03, girdle_planes: int = 16, star_deg: float = 21.5,
                    upper_girdle_deg: float = 41.0, lower_girdle_deg: float = 42.0) -> np.ndarray:
    """Parametric round brilliant of girdle radius 1: table, 8 stars, 8 crown mains,
    16 upper girdles, 8 pavilion mains, 16 lower girdles (57 facets) plus
    ``girdle_planes`` vertical girdle planes.  ``table`` is the table diameter and
    ``girdle`` the girdle thickness as fractions of the girdle diameter."""
    half_g = girdle  # girdle thickness 2 * half_g in radius-1 units = girdle * diameter
    crown_h = (1.0 - table) * math.tan(math.radians(crown_deg))
    table_y = half_g + crown_h
    planes: List[Plane] = [(0.0, 1.0, 0.0, -table_y)]
    # stars meet the table edge between the mains
    planes += _tier(star_deg, 8, math.pi / 8, table * 1.005, table_y, True)
    planes += _tier(crown_deg, 8, 0.0, 1.0, half_g, True)
    planes += _tier(upper_girdle_deg, 16, math.pi / 16, 1.0, half_g, True)
    for i in range(girdle_planes):
        phi = 2.0 * math.pi * i / girdle_planes
        planes.append((math.cos(phi), 0.0, math.sin(phi), -1.0))
    planes += _tier(pavilion_deg, 8, 0.0, 1.0, -half_g, False)
    planes += _tier(lower_girdle_deg, 16, math.pi / 16, 1.0, -half_g, False)
    return _to_array(planes)


def _scale_planes(planes: np.ndarray, sx: float, sy: float, sz: float) -> np.ndarray:
    """Planes of the stone mapped by x' = diag(sx, sy, sz) x."""
    n = planes[:, :3].astype(np.float64) / np.array([sx, sy, sz])
    d = planes[:, 3].astype(np.float64)
    norm = np.linalg.norm(n, axis=1)
    out = np.empty_like(planes)
    out[:, :3] = n / norm[:, None]
    out[:, 3] = d / norm
    return out


def oval_brilliant(aspect: float = 1.35) -> np.ndarray:
    """Oval brilliant with 69 planes: table, 8 stars, 8 crown mains, 16 upper girdles,
    12 girdle facets, 8 pavilion mains, 16 lower girdles; elongated along x."""
    rb = round_brilliant(girdle_planes=12)
    assert rb.shape[0] == 69
    return _scale_planes(rb, aspect, 1.0, 1.0)


def random_convex_stone(seed: int, n_planes: int, jitter_deg: float = 3.0) -> np.ndarray:
    """A seeded brilliant-style convex stone with exactly ``n_planes`` planes (17..1024):
    table + crown tiers + girdle ring + pavilion tiers, every cut angle jittered by up to
    +-``jitter_deg`` and every azimuth by up to +-``jitter_deg`` / 2."""
    if n_planes < 17:
        raise ValueError("need at least 17 planes")
    rng = np.random.default_rng(seed)
    rest = n_planes - 1  # table
    girdle_n = max(8, int(round(rest * 0.15)))
    crown_n = (rest - girdle_n) // 2
    pav_n = rest - girdle_n - crown_n
    table = float(rng.uniform(0.5, 0.62))
    crown_deg = float(rng.uniform(30.0, 37.0))
    pav_deg = float(rng.uniform(39.0, 42.5))
    half_g = float(rng.uniform(0.015, 0.04))
    crown_h = (1.0 - table) * math.tan(math.radians(crown_deg))
    table_y = half_g + crown_h
    planes: List[Plane] = [(0.0, 1.0, 0.0, -table_y)]

    def jit():
        return float(rng.uniform(-jitter_deg, jitter_deg))

    def tiers(total, base_deg, upper, y0):
        out = []
        k = max(1, min(4, total // 8))
        sizes = [total // k + (1 if i < total % k else 0) for i in range(k)]
        for t, cnt in enumerate(sizes):
            ang = base_deg + (t - (k - 1) / 2.0) * 4.0
            phase = float(rng.uniform(0, 2 * math.pi))
            for i in range(cnt):
                a = math.radians(ang + jit())
                phi = phase + 2 * math.pi * i / cnt + math.radians(jit() / 2)
                ny = math.cos(a) if upper else -math.cos(a)
                r = 1.0 - 0.05 * t
                out.append(_plane((math.sin(a) * math.cos(phi), ny, math.sin(a) * math.sin(phi)),
                                  (r * math.cos(phi), y0, r * math.sin(phi))))
        return out

    planes += tiers(crown_n, crown_deg, True, half_g)
    for i in range(girdle_n):
        phi = 2 * math.pi * i / girdle_n
        planes.append((math.cos(phi), 0.0, math.sin(phi), -1.0))
    planes += tiers(pav_n, pav_deg, False, -half_g)
    arr = _to_array(planes)
    assert arr.shape[0] == n_planes
    return arr


def step_cut(seed: int = 233, n_planes: int = 233) -> np.ndarray:
    """A seeded mixed step cut (step crown, brilliant pavilion) with ``n_planes`` planes:
    table, crown steps of decreasing angle, a faceted girdle and pavilion steps."""
    rng = np.random.default_rng(seed)
    planes: List[Plane] = [(0.0, 1.0, 0.0, -0.34)]
    remaining = n_planes - 1
    girdle_n = 32
    crown_steps = [(28.0, 24), (36.0, 24), (44.0, 24)]
    pav_n = remaining - girdle_n - sum(c for _, c in crown_steps)
    pav_steps = []
    base = [38.0, 41.0, 44.0, 48.0, 52.0]
    per = pav_n // len(base)
    for i, ang in enumerate(base):
        pav_steps.append((ang, per + (1 if i < pav_n - per * len(base) else 0)))
    for t, (ang, cnt) in enumerate(crown_steps):
        planes += _tier(ang + float(rng.uniform(-1, 1)), cnt, float(rng.uniform(0, 0.3)),
                        1.0 - 0.06 * t, 0.03 + 0.04 * t, True)
    for i in range(girdle_n):
        phi = 2 * math.pi * i / girdle_n
        planes.append((math.cos(phi), 0.0, math.sin(phi), -1.0))
    for t, (ang, cnt) in enumerate(pav_steps):
        planes += _tier(ang + float(rng.uniform(-1, 1)), cnt, float(rng.uniform(0, 0.3)),
                        1.0 - 0.1 * t, -0.03 - 0.08 * t, False)
    arr = _scale_planes(_to_array(planes), 1.2, 1.0, 1.0)
    assert arr.shape[0] == n_planes
    return arr


def seeded_unit_vector(seed: int) -> Tuple[float, float, float]:
    v = np.random.default_rng(seed).normal(size=3)
    v /= np.linalg.norm(v)
    return tuple(float(x) for x in np.float32(v))


def pleochroic_sapphire(seed: int, stone_cm: float = 0.8) -> GemMaterial:
    """The single-job workload's stone: Sapphire's Sellmeier fit (n_o ~ 1.768,
    n_e = n_o - 0.0081), a seeded random optic axis, and a pleochroic absorption tensor of
    8 Gaussian bands per axis centred on the 8 comb slots with peaks in [0.01, 2] cm^-1.
    ``absorption_path_scale`` converts model units (girdle radius 1) to cm."""
    rng = np.random.default_rng(seed)
    centres = [380.0 + 25.0 + 50.0 * k for k in range(8)]

    def bands():
        return tuple((c, 30.0, float(rng.uniform(0.01, 2.0))) for c in centres)

    o, e, beta = bands(), bands(), bands()
    m = by_name("Sapphire")
    return replace(m, name="Pleochroic sapphire (seeded)", o_bands=o, e_bands=e, beta_bands=beta,
                   c_axis=seeded_unit_vector(seed + 1), absorption_path_scale=stone_cm / 2.0)


def sellmeier_for(n_d: float, abbe: float) -> Tuple[float, ...]:
    """Single-term Sellmeier (plus a fixed IR term) hitting n_d at 589.3 nm and the Abbe
    number V = (n_d - 1) / (n_F - n_C); solved by bisection on the UV resonance."""
    ld, lf, lc = 0.5893, 0.4861, 0.6563
    c2, b2 = 100.0, 1.0

    def n_of(b1, c1, l):
        l2 = l * l
        return math.sqrt(1.0 + b1 * l2 / (l2 - c1) + b2 * l2 / (l2 - c2))

    def b1_for(c1):
        l2 = ld * ld
        target = n_d * n_d - 1.0 - b2 * l2 / (l2 - c2)
        return target * (l2 - c1) / l2

    lo, hi = 1e-5, 0.2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        b1 = b1_for(mid)
        v = (n_of(b1, mid, ld) - 1.0) / (n_of(b1, mid, lf) - n_of(b1, mid, lc))
        if v > abbe:
            lo = mid
        else:
            hi = mid
    c1 = 0.5 * (lo + hi)
    return (b1_for(c1), b2, 0.0, c1, c2, 1.0)


def stress_material() -> GemMaterial:
    """High-dispersion (Abbe ~ 20, n_d ~ 2.2) uniaxial stone with n_e - n_o = 0.17."""
    coeffs = sellmeier_for(2.2, 20.0)
    return GemMaterial(name="Stress uniaxial (Abbe 20, dn 0.17)", crystal_system="Tetragonal",
                       disp_kind=DISP_SELLMEIER3, disp=coeffs, birefringence_delta=0.17,
                       c_axis=seeded_unit_vector(5))


def procedural_env(seed: int, width: int = 2048, height: int = 1024, n_boxes: int = 6) -> np.ndarray:
    """Lat-long HDR environment (H, W, 3) f32: a dark vertical gradient plus seeded soft
    boxes (smooth angular falloff, radiance 5-40).  Texel (x, y) is the direction of
    ``EnvironmentMap::uv_to_direction`` (env_map.rs:216-224) at the texel centre."""
    rng = np.random.default_rng(seed)
    v = (np.arange(height, dtype=np.float64) + 0.5) / height
    u = (np.arange(width, dtype=np.float64) + 0.5) / width
    theta = v * math.pi
    phi = u * 2 * math.pi
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    sp, cp = np.sin(phi)[None, :], np.cos(phi)[None, :]
    dx, dy, dz = st * sp, np.broadcast_to(ct, (height, width)), st * cp
    img = np.empty((height, width, 3), dtype=np.float64)
    grad = 0.02 + 0.06 * np.clip(dy, 0, 1)
    img[..., 0] = grad * 0.9
    img[..., 1] = grad * 0.95
    img[..., 2] = grad * 1.1
    for _ in range(n_boxes):
        d = rng.normal(size=3)
        d[1] = abs(d[1]) + 0.2
        d /= np.linalg.norm(d)
        cosang = dx * d[0] + dy * d[1] + dz * d[2]
        width_cos = float(rng.uniform(0.97, 0.995))
        w = np.clip((cosang - width_cos) / (1 - width_cos), 0, 1) ** 2
        color = rng.uniform(0.7, 1.0, size=3)
        power = float(rng.uniform(5.0, 40.0))
        img += (w * power)[..., None] * color[None, None, :]
    return img.astype(np.float32)


def furnace_material() -> GemMaterial:
    from .materials import furnace_material as fm
    return fm()
