"""Gem materials: the host-side mirror of gemray's ``GemMaterial``.

Mirrors ``crates/gemray/src/optics/materials.rs:32-120`` (the struct) and the built-in
table at ``materials.rs:136-1222``.  Only the fields the transport estimator reads are
kept.  Coefficients are copied as data because they are the materials' published
dispersion fits; the reference's bit-exact golden vectors depend on them.

A material is plain data.  ``to_struct`` packs it into the ``gr_material`` C layout the
HIP library consumes (``include/gemray_hip.h``).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field, replace
from typing import List, Optional, Tuple

MAX_BANDS = 8

DISP_SELLMEIER1 = 0
DISP_SELLMEIER3 = 1
DISP_CAUCHY = 2

Band = Tuple[float, float, float]  # (center_nm, width_nm, peak)


@dataclass(frozen=True)
class GemMaterial:
    """The estimator-visible part of ``GemMaterial`` (materials.rs:32-120)."""

    name: str
    crystal_system: str  # "Cubic", "Trigonal", ...
    disp_kind: int
    disp: Tuple[float, ...]  # S1: (b1, c1); S3: (b0, b1, b2, c0, c1, c2); Cauchy: (a, b, c)
    birefringence_delta: float = 0.0
    o_bands: Tuple[Band, ...] = ()
    e_bands: Tuple[Band, ...] = ()
    beta_bands: Optional[Tuple[Band, ...]] = None
    c_axis: Tuple[float, float, float] = (0.0, 1.0, 0.0)
    biaxial_delta_beta_alpha: Optional[float] = None
    scattering_sigma_s: float = 0.0
    scattering_g: float = 0.0
    edge_rounding_radius: float = 0.0
    absorption_path_scale: float = 1.0

    @property
    def is_cubic(self) -> bool:
        return self.crystal_system == "Cubic"

    @property
    def is_anisotropic(self) -> bool:
        # transport.rs:548-550 (build_ray_material_context)
        return (not self.is_cubic) and abs(self.birefringence_delta) > 1e-4

    def with_absorption_path_scale(self, scale: float) -> "GemMaterial":
        return replace(self, absorption_path_scale=scale)

    def with_c_axis(self, c_axis) -> "GemMaterial":
        return replace(self, c_axis=tuple(float(v) for v in c_axis))

    def dispersion_evaluate(self, lambda_nm: float) -> float:
        """float64 reading of ``DispersionModel::evaluate`` (dispersion.rs:12-37)."""
        l2 = (lambda_nm * 1e-3) ** 2
        if self.disp_kind == DISP_SELLMEIER1:
            b1, c1 = self.disp[:2]
            return math.sqrt(max(1.0 + b1 * l2 / (l2 - c1), 1.0))
        if self.disp_kind == DISP_SELLMEIER3:
            b = self.disp[:3]
            c = self.disp[3:6]
            n2 = 1.0 + sum(b[i] * l2 / (l2 - c[i]) for i in range(3))
            return math.sqrt(max(n2, 1.0))
        a, b, c = self.disp[:3]
        return a + b / l2 + c / (l2 * l2)


def _sellmeier3(b, c):
    return DISP_SELLMEIER3, tuple(b) + tuple(c)


def _cauchy(a, b, c=0.0):
    return DISP_CAUCHY, (a, b, c)


def legacy_rgb_bands(rgb) -> Tuple[Band, ...]:
    """optics/absorption.rs:216-223."""
    return ((620.0, 55.0, rgb[0]), (540.0, 45.0, rgb[1]), (450.0, 45.0, rgb[2]))


def _m(name, system, disp, dn=0.0, o=(), e=None, beta=None, c_axis=(0.0, 1.0, 0.0), biax=None):
    kind, coeffs = disp
    return GemMaterial(
        name=name,
        crystal_system=system,
        disp_kind=kind,
        disp=coeffs,
        birefringence_delta=dn,
        o_bands=tuple(o),
        e_bands=tuple(o if e is None else e),
        beta_bands=None if beta is None else tuple(beta),
        c_axis=c_axis,
        biaxial_delta_beta_alpha=biax,
    )


# materials.rs:136-1222, estimator-visible fields only.
BUILTIN: List[GemMaterial] = [
    _m("Diamond", "Cubic", _sellmeier3([4.3356, 0.3306, 0.0], [0.011236, 0.030625, 1.0])),
    _m(
        "Sapphire",
        "Trigonal",
        _sellmeier3([1.431349, 0.650547, 5.341402], [0.00528, 0.01424, 325.015]),
        -0.0081,
        o=[(580.0, 90.0, 3.0)],
        e=[(700.0, 95.0, 2.1)],
    ),
    _m(
        "Ruby",
        "Trigonal",
        _sellmeier3([1.431349, 0.650547, 5.341402], [0.00528, 0.01424, 325.015]),
        -0.0081,
        o=[(410.0, 30.0, 3.0), (556.0, 45.0, 2.5)],
        e=[(400.0, 30.0, 2.9), (550.0, 45.0, 1.4)],
    ),
    _m(
        "Emerald",
        "Hexagonal",
        _cauchy(1.566794, 0.004273),
        -0.0060,
        o=[(430.0, 30.0, 2.8), (600.0, 40.0, 2.6)],
    ),
    _m("Zircon", "Tetragonal", _cauchy(1.890963, 0.011820), 0.0590, o=legacy_rgb_bands([0.2, 0.6, 1.8])),
    _m(
        "Alexandrite",
        "Orthorhombic",
        _sellmeier3([0.78522, 1.21202, 16.81], [0.0, 0.01262, 1000.0]),
        0.0076,
        o=[(565.0, 40.0, 1.9), (420.0, 32.0, 1.8)],
        e=[(595.0, 40.0, 3.1), (430.0, 32.0, 1.5)],
        beta=[(560.0, 40.0, 0.65), (422.0, 32.0, 3.5)],
        biax=0.001951,
    ),
    _m("Topaz", "Orthorhombic", _cauchy(1.614872, 0.004273), 0.0080, o=legacy_rgb_bands([1.2, 0.4, 0.1]), biax=0.0024),
    _m("Spinel", "Cubic", _sellmeier3([1.8938, 3.0755, 0.0], [0.009884, 250.462, 1.0]), o=legacy_rgb_bands([0.4, 2.2, 1.6])),
    _m(
        "Quartz",
        "Trigonal",
        _sellmeier3([0.2860414, 1.0704408, 1.1020224], [0.0, 0.0100586, 100.0]),
        0.0091,
        o=legacy_rgb_bands([0.8, 1.8, 0.6]),
    ),
    _m(
        "Tourmaline",
        "Trigonal",
        _cauchy(1.624481, 0.005183),
        -0.0210,
        o=[(720.0, 55.0, 3.0), (430.0, 45.0, 1.8)],
        e=[(720.0, 55.0, 1.0), (430.0, 45.0, 0.6)],
        c_axis=(1.0, 0.0, 0.0),
    ),
    _m(
        "Tanzanite",
        "Orthorhombic",
        _cauchy(1.674589, 0.009123),
        0.0130,
        o=[(595.0, 45.0, 1.3), (528.0, 35.0, 1.8)],
        e=[(595.0, 45.0, 1.6), (528.0, 35.0, 1.0), (455.0, 40.0, 2.6)],
        beta=[(595.0, 45.0, 3.2), (528.0, 35.0, 2.8)],
        biax=0.0070,
    ),
    _m("Synthetic Moissanite", "Hexagonal", _sellmeier3([1.163887, 4.408433, 21.53], [0.0, 0.03178, 1000.0]), 0.0415),
    _m("Cubic Zirconia", "Cubic", _sellmeier3([1.347091, 2.117788, 9.452943], [0.003912, 0.027802, 591.489])),
]


def by_name(name: str) -> GemMaterial:
    """``GemMaterial::by_name`` (materials.rs:1443-1452): exact (case-insensitive) match,
    else the longest built-in name contained in ``name``."""
    for m in BUILTIN:
        if m.name.lower() == name.lower():
            return m
    needle = name.lower()
    hits = [m for m in BUILTIN if m.name.lower() in needle]
    if not hits:
        raise KeyError(name)
    return max(hits, key=lambda m: len(m.name))


def furnace_material() -> GemMaterial:
    """Colourless n=1.5 cubic anchor (renderer/gpu/estimator_check/mod.rs:90-109)."""
    return _m("Furnace anchor (colourless n=1.5 cubic)", "Cubic", _cauchy(1.5, 0.0, 0.0))


def new_custom(name, mean_ri, dispersion_delta, birefringence_delta, absorption_rgb) -> GemMaterial:
    """``GemMaterial::new_custom`` (materials.rs:1257-1316), evaluated in float64."""
    ld, lf, lc = 0.5893, 0.4861, 0.6563
    k_fc = 1.0 / (1.0 / (lf * lf) - 1.0 / (lc * lc))
    b = max(dispersion_delta * k_fc, 0.0)
    a = max(mean_ri - b / (ld * ld), 1.0)
    system = "Trigonal" if abs(birefringence_delta) > 1e-4 else "Cubic"
    return _m(name, system, _cauchy(a, b, 0.0), birefringence_delta, o=legacy_rgb_bands(absorption_rgb))


class MaterialStruct(ctypes.Structure):
    """``gr_material`` (include/gemray_hip.h).  The oracle's ``or_material`` is a prefix of
    this layout, so one struct feeds both."""

    _fields_ = [
        ("disp_kind", ctypes.c_int32),
        ("disp", ctypes.c_float * 6),
        ("birefringence_delta", ctypes.c_float),
        ("crystal_cubic", ctypes.c_int32),
        ("c_axis", ctypes.c_float * 3),
        ("n_o_bands", ctypes.c_int32),
        ("n_e_bands", ctypes.c_int32),
        ("o_bands", (ctypes.c_float * 3) * MAX_BANDS),
        ("e_bands", (ctypes.c_float * 3) * MAX_BANDS),
        ("absorption_path_scale", ctypes.c_float),
        ("biaxial_delta_beta_alpha", ctypes.c_float),
        ("scattering_sigma_s", ctypes.c_float),
        ("edge_rounding_radius", ctypes.c_float),
        ("has_beta_bands", ctypes.c_int32),
        ("n_beta_bands", ctypes.c_int32),
        ("beta_bands", (ctypes.c_float * 3) * MAX_BANDS),
        ("scattering_g", ctypes.c_float),
    ]


def to_struct(m: GemMaterial) -> MaterialStruct:
    if len(m.o_bands) > MAX_BANDS or len(m.e_bands) > MAX_BANDS:
        raise ValueError(f"{m.name}: more than {MAX_BANDS} absorption bands per axis")
    s = MaterialStruct()
    s.disp_kind = m.disp_kind
    for i, v in enumerate(m.disp):
        s.disp[i] = v
    s.birefringence_delta = m.birefringence_delta
    s.crystal_cubic = 1 if m.is_cubic else 0
    for i in range(3):
        s.c_axis[i] = m.c_axis[i]
    s.n_o_bands = len(m.o_bands)
    s.n_e_bands = len(m.e_bands)
    for i, b in enumerate(m.o_bands):
        for j in range(3):
            s.o_bands[i][j] = b[j]
    for i, b in enumerate(m.e_bands):
        for j in range(3):
            s.e_bands[i][j] = b[j]
    s.absorption_path_scale = m.absorption_path_scale
    s.biaxial_delta_beta_alpha = (
        float("nan") if m.biaxial_delta_beta_alpha is None else m.biaxial_delta_beta_alpha)
    s.scattering_sigma_s = m.scattering_sigma_s
    s.scattering_g = m.scattering_g
    s.edge_rounding_radius = m.edge_rounding_radius
    if m.beta_bands is not None:
        if len(m.beta_bands) > MAX_BANDS:
            raise ValueError(f"{m.name}: more than {MAX_BANDS} beta absorption bands")
        s.has_beta_bands = 1
        s.n_beta_bands = len(m.beta_bands)
        for i, b in enumerate(m.beta_bands):
            for j in range(3):
                s.beta_bands[i][j] = b[j]
    return s
