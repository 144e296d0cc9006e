"""The CPU leg the tests hand to gemray_studio_amd.hybrid: the oracle's render_core sums
(the reference's cpu_trace_range, hybrid.rs:337-470, restated by the oracle)."""
import numpy as np

import pyoracle as po
from gemray_studio_amd import backend as B
from gemray_studio_amd import materials as M


def oracle_cpu_trace(camera_args):
    def trace(scene: B.SceneRef, sample_offset: int, spp: int) -> np.ndarray:
        env = scene.environment
        oenv = (po.studio_env(env.preset, env.exposure, env.light_yaw, env.light_pitch)
                if isinstance(env, B.Studio) else po.map_env(env.rgb))
        acc, _, _, _ = po.render(po.camera_new(*camera_args), scene.width, scene.height, scene.planes,
                                 M.to_struct(scene.material), scene.max_bounces, oenv, sample_offset, spp,
                                 finishes=scene.facet_finishes)
        return acc
    return trace
