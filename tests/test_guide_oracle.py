"""The oracle's primary-hit guide buffers (generate_guide_buffers,
apps/diagram-gui/src/bridge/guide_pass.rs:37-110) against a float64 ray / convex
polyhedron intersection written independently here, and known answers on the unit cube
(intersect.rs:259-305) (CPU only).
"""
import numpy as np

import pyoracle as po
from gemray_studio_amd import scenes


def exact_first_hit(o, d, planes):
    """float64 slab clip: (t, facet) of the first boundary crossing with t > 1e-4, or None."""
    n = planes[:, :3].astype(np.float64)
    side = planes[:, 3].astype(np.float64) + n @ o
    den = n @ d
    with np.errstate(divide="ignore", invalid="ignore"):
        t = -side / den
    enter = den < -1e-7
    leave = den > 1e-7
    if np.any((~enter) & (~leave) & (side > 0)):
        return None
    t_near = t[enter].max() if enter.any() else -1e30
    t_far = t[leave].min() if leave.any() else 1e30
    if t_near > t_far:
        return None
    if t_near > 1e-4:
        return t_near, int(np.flatnonzero(enter)[np.argmax(t[enter])])
    if t_far > 1e-4:
        return t_far, int(np.flatnonzero(leave)[np.argmin(t[leave])])
    return None


def test_guide_buffers_match_an_exact_intersection():
    for planes, (yaw, pitch, dist, fov) in ((po.standard_round_brilliant(), (0.35, 0.28, 5.0, 18.0)),
                                            (scenes.oval_brilliant(), (1.1, 0.9, 4.0, 24.0)),
                                            (scenes.step_cut(), (-0.4, 1.5, 3.5, 30.0))):
        cam = po.camera_new(yaw, pitch, dist, fov)
        w, h = 40, 28
        depth, normal, facet = po.guide_buffers(cam, w, h, planes)
        hits = 0
        for p in range(w * h):
            o = np.zeros(3, np.float32)
            d = np.zeros(3, np.float32)
            po.lib().or_generate_ray(po.ctypes.byref(cam), float(p % w), float(p // w), float(w), float(h), 0.0, 0.0,
                                     po.fptr(o), po.fptr(d))
            ref = exact_first_hit(o.astype(np.float64), d.astype(np.float64), planes)
            if ref is None:
                # a miss, unless the ray grazes the stone within rounding
                if facet[p] != -1:
                    assert abs(depth[p] - 1e6) > 0 and facet[p] >= 0
                    continue
                assert depth[p] == np.float32(1e6) and np.all(normal[p] == 0)
                continue
            hits += 1
            assert abs(depth[p] - ref[0]) <= 1e-5 * max(1.0, ref[0]), (p, depth[p], ref)
            if facet[p] != ref[1]:
                # a different facet only on an edge: both planes pass through the hit point
                hp = o + depth[p] * d
                assert abs(planes[facet[p], :3] @ hp + planes[facet[p], 3]) < 1e-5
            assert np.array_equal(normal[p], planes[facet[p], :3])
        assert hits > 0.2 * w * h


def test_unit_cube_and_empty_planes():
    cube = np.array([[1, 0, 0, -1], [-1, 0, 0, -1], [0, 1, 0, -1], [0, -1, 0, -1], [0, 0, 1, -1], [0, 0, -1, -1]],
                    np.float32)
    cam = po.camera_new(0.0, 0.0, 5.5, 60.0)  # on +z, looking at the cube's +z face
    w = h = 9
    depth, normal, facet = po.guide_buffers(cam, w, h, cube)
    centre = (h // 2) * w + w // 2
    assert facet[centre] == 4 and normal[centre].tolist() == [0.0, 0.0, 1.0]
    assert abs(depth[centre] - 4.5) < 0.02
    # the corners of a 60 degree frame at distance 5.5 miss the unit cube
    assert facet[0] == -1 and depth[0] == np.float32(1e6) and np.all(normal[0] == 0)
    # no planes: the unbounded sentinel hit (intersect.rs:287-305)
    depth, normal, facet = po.guide_buffers(cam, 3, 2, np.zeros((0, 4), np.float32))
    assert np.all(depth == np.float32(1e30)) and np.all(facet == 0) and np.all(normal == 0)
