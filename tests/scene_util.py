"""Builds one scene for both the HIP product path and the oracle."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

import pyoracle as po
from gemray_studio_amd import backend as B
from gemray_studio_amd import materials as M


@dataclass
class Case:
    name: str
    material: M.GemMaterial
    planes: np.ndarray
    env: object  # B.Studio | B.HdrMap
    width: int
    height: int
    max_bounces: int
    camera: tuple = (0.35, 0.28, 5.0, 18.0)  # test_camera() (estimator_check/mod.rs:132-138)
    finishes: object = None  # per-plane FacetFinish (0 polished, 1 frosted) or None

    def product_scene(self) -> B.SceneRef:
        cam = B.camera_new(*self.camera)
        return B.SceneRef(camera=cam, width=self.width, height=self.height, planes=self.planes,
                          material=self.material, max_bounces=self.max_bounces,
                          environment=self.env, facet_finishes=self.finishes)

    def oracle_env(self):
        if isinstance(self.env, B.Studio):
            return po.studio_env(self.env.preset, self.env.exposure, self.env.light_yaw,
                                 self.env.light_pitch)
        return po.map_env(self.env.rgb)

    def oracle_render(self, first_sample, spp, per_sample=True, events=None, threads=None,
                      pixel_begin=0, pixel_end=None):
        cam = po.camera_new(*self.camera)
        return po.render(cam, self.width, self.height, self.planes, M.to_struct(self.material),
                         self.max_bounces, self.oracle_env(), first_sample, spp,
                         per_sample=per_sample, events=events, threads=threads,
                         pixel_begin=pixel_begin, pixel_end=pixel_end, finishes=self.finishes)


def frosted_girdle(n_planes: int, girdle=range(33, 49)):
    """Finishes with the standard round brilliant's girdle band frosted
    (STANDARD_ROUND_BRILLIANT_GIRDLE_FACETS, geometry/cuts.rs:37), the rest polished."""
    f = np.zeros(n_planes, np.uint32)
    f[list(girdle)] = 1
    return f


def camera_bits_equal(a, b) -> bool:
    fa = np.ctypeslib.as_array(ctypes.cast(ctypes.byref(a), ctypes.POINTER(ctypes.c_float)), (13,))
    fb = np.ctypeslib.as_array(ctypes.cast(ctypes.byref(b), ctypes.POINTER(ctypes.c_float)), (13,))
    return bool(np.array_equal(fa.view(np.uint32), fb.view(np.uint32)))
