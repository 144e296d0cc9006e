"""The trace kernel's clip shortcuts, modelled in float32 numpy and checked against the
oracle's intersect_polyhedron restatement (CPU only).

The kernels (gemray_studio_amd/csrc/clip.h, clip_scan / exterior_miss) depart from
the reference's scan in two ways that must not change any result:

* the screen: a plane takes the exact division only if side + t_ext * denom > 0 (one fma,
  exact sign), i.e. its exact quotient is beyond the running extreme t_ext;
* the exterior shortcut: a ray that starts outside the facet it just met and moves away
  from it is a miss without the scan.

The model below reproduces the kernel's float32 operations (the same association order)
and the tests require the oracle's exact hit, t bits and facet on randomized rays,
including planes duplicated with d moved by 0-2 ulp (ties and near-ties).
"""
import ctypes

import numpy as np

import pyoracle as po
from gemray_studio_amd import scenes

L = po.lib()
F = np.float32


def oracle_intersect(planes, o, d):
    t = np.zeros(1, F)
    n = np.zeros(3, F)
    f = ctypes.c_int()
    hit = L.or_intersect(po.fptr(planes), len(planes), po.fptr(po.f32arr(o)), po.fptr(po.f32arr(d)),
                         po.fptr(t), po.fptr(n), ctypes.byref(f))
    return bool(hit), t[0], f.value


def dot_rows(p, v):
    """(p.x*v.x + p.y*v.y) + p.z*v.z in float32, rows of p against rows of v."""
    return ((p[..., 0] * v[..., 0] + p[..., 1] * v[..., 1]) + p[..., 2] * v[..., 2]).astype(F)


def model_scan(planes, ro, rd, rng=None, near_start=-1e30):
    """clip_scan<false> for a batch of rays (ro, rd: (R, 3) float32), the entering extreme
    starting at near_start (-1e30 as the reference; the transport kernel starts it at the
    1e-4 hit threshold).  The screen's fma sign is evaluated in float64, where
    t_ext * denom is exact and adding side keeps the exact sum's sign."""
    R = ro.shape[0]
    t_near = np.full(R, F(near_start))
    t_far = np.full(R, F(1e30))
    near_idx = np.full(R, -1)
    far_idx = np.full(R, -1)
    outside = np.zeros(R, bool)
    divisions = 0
    for i, p in enumerate(planes):
        denom = dot_rows(np.broadcast_to(p, (R, 4)), rd)
        side = (p[3] + dot_rows(np.broadcast_to(p, (R, 4)), ro)).astype(F)
        steep = np.abs(denom) > F(1e-7)
        entering = denom < 0
        t_ext = np.where(entering, t_near, t_far)
        g = t_ext.astype(np.float64) * denom.astype(np.float64) + side.astype(np.float64)
        cand = np.where(steep, g > 0, side > 0)
        outside |= cand & ~steep
        with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
            t = ((-side) / denom).astype(F)
        work = cand & steep
        divisions += int(np.count_nonzero(work))
        upd_n = work & entering & (t > t_near)
        upd_f = work & ~entering & (t < t_far)
        t_near = np.where(upd_n, t, t_near)
        near_idx = np.where(upd_n, i, near_idx)
        t_far = np.where(upd_f, t, t_far)
        far_idx = np.where(upd_f, i, far_idx)
    hit = (~outside) & ~(t_near > t_far) & ((t_near > F(1e-4)) | (t_far > F(1e-4)))
    use_near = t_near > F(1e-4)
    t_hit = np.where(use_near, t_near, t_far)
    facet = np.where(use_near, near_idx, far_idx)
    facet = np.where(facet < 0, 0, facet)
    return hit, t_hit, facet, divisions


def exterior_miss(p, ro, rd):
    denom = dot_rows(np.broadcast_to(p, ro.shape[:1] + (4,)), rd)
    side = (p[..., 3] + dot_rows(np.broadcast_to(p, ro.shape[:1] + (4,)), ro)).astype(F)
    return (np.abs(denom) > F(1e-7)) & (denom > 0) & (side > 0)


def random_rays(rng, n, spread=2.5):
    o = rng.normal(size=(n, 3)).astype(F)
    o = (o / np.linalg.norm(o, axis=1, keepdims=True) * spread).astype(F)
    # Aim at jittered points near the stone so that hits, grazes and misses all occur.
    target = (rng.normal(size=(n, 3)) * 0.6).astype(F)
    d = (target - o).astype(F)
    d = (d / np.sqrt((d * d).sum(axis=1, keepdims=True))).astype(F)
    return o, d


STONES = [
    ("round-brilliant-73", scenes.round_brilliant()),
    ("oval-69", scenes.oval_brilliant()),
    ("random-23", scenes.random_convex_stone(11, 23)),
    ("random-181", scenes.random_convex_stone(12, 181)),
    ("step-233", scenes.step_cut()),
]


def with_near_duplicates(planes, rng, n=8):
    """Appends copies of n random planes with d moved by 0, 1 or 2 ulp: their t values tie
    or differ by an ulp or two, which only the exact division and the first-index rule
    can order."""
    extra = planes[rng.integers(0, len(planes), size=n)].copy()
    for k in range(n):
        for _ in range(int(rng.integers(0, 3))):
            extra[k, 3] = np.nextafter(extra[k, 3], F(np.inf) if rng.random() < 0.5 else F(-np.inf))
    out = np.concatenate([planes, extra]).astype(F)
    return np.ascontiguousarray(out[rng.permutation(len(out))])


def test_screened_scan_selects_the_oracle_planes():
    rng = np.random.default_rng(1)
    total, divs, planes_seen = 0, 0, 0
    for name, planes in STONES:
        planes = with_near_duplicates(np.ascontiguousarray(planes, F), rng)
        # Exterior rays, then interior rays from points inside the stone.
        o, d = random_rays(rng, 1500)
        inner = (rng.normal(size=(1500, 3)) * 0.15).astype(F)
        o = np.concatenate([o, inner]).astype(F)
        d = np.concatenate([d, random_rays(rng, 1500)[1]]).astype(F)
        hit, t, facet, nd = model_scan(planes, o, d, rng)
        divs += nd
        planes_seen += len(planes) * len(o)
        for k in range(len(o)):
            h_ref, t_ref, f_ref = oracle_intersect(planes, o[k], d[k])
            assert hit[k] == h_ref, (name, k)
            if h_ref:
                assert t[k].view(np.uint32) == t_ref.view(np.uint32), (name, k)
                assert facet[k] == f_ref, (name, k)
        total += len(o)
    assert total == 15000
    # The screen is worth having: most planes skip the division.
    assert divs < 0.25 * planes_seen


def test_entering_extreme_floored_at_the_hit_threshold_selects_the_oracle_planes():
    """The transport kernel starts t_near at 1e-4: entering planes whose exact t is at most
    the hit threshold are screened out.  Same hit, t bits and facet as the oracle on exterior
    rays, interior rays and rays leaving a facet (origin = hit point + 1e-4 along the new
    direction, as the tracer continues a path), and interior scans take fewer divisions."""
    rng = np.random.default_rng(5)
    total = 0
    for name, planes in STONES:
        planes = with_near_duplicates(np.ascontiguousarray(planes, F), rng)
        o, d = random_rays(rng, 1000)
        hit0, t0, f0, _ = model_scan(planes, o, d)
        # Continue the exterior hits from their facet: reflected or straight on, inside.
        ok = hit0
        p_hit = (o[ok] + d[ok] * t0[ok][:, None]).astype(F)
        d2 = random_rays(rng, int(ok.sum()))[1]
        o2 = (p_hit + d2 * F(1e-4)).astype(F)
        inner = (rng.normal(size=(1000, 3)) * 0.15).astype(F)
        o = np.concatenate([o, o2, inner]).astype(F)
        d = np.concatenate([d, d2, random_rays(rng, 1000)[1]]).astype(F)
        hit, t, facet, nd = model_scan(planes, o, d, near_start=1e-4)
        _, _, _, nd_ref = model_scan(planes, o, d)
        assert nd < nd_ref, name
        for k in range(len(o)):
            h_ref, t_ref, f_ref = oracle_intersect(planes, o[k], d[k])
            assert hit[k] == h_ref, (name, k)
            if h_ref:
                assert t[k].view(np.uint32) == t_ref.view(np.uint32), (name, k)
                assert facet[k] == f_ref, (name, k)
        total += len(o)
    assert total > 10000


def test_exterior_shortcut_only_fires_on_misses():
    rng = np.random.default_rng(2)
    fired = 0
    checked = 0
    for name, planes in STONES:
        planes = np.ascontiguousarray(planes, F)
        # Surface points: the oracle's hits of rays from inside.
        inner = (rng.normal(size=(800, 3)) * 0.15).astype(F)
        d_in = random_rays(rng, 800)[1]
        for k in range(len(inner)):
            h, t, f = oracle_intersect(planes, inner[k], d_in[k])
            if not h:
                continue
            p = planes[f]
            hit_point = (inner[k] + d_in[k] * t).astype(F)
            n = p[:3]
            # Outgoing directions: refraction out (any direction on the outer side) and
            # grazing ones, plus reflections off the outside.
            for _ in range(3):
                v = rng.normal(size=3).astype(F)
                v = (v / np.sqrt(np.dot(v, v))).astype(F)
                if rng.random() < 0.3:
                    v = (v - n * np.dot(v, n) + n * F(rng.uniform(1e-6, 1e-3))).astype(F)
                    v = (v / np.sqrt(np.dot(v, v))).astype(F)
                elif np.dot(v, n) < 0:
                    v = (-v).astype(F)
                ro = (hit_point + v * F(1e-4)).astype(F)
                if exterior_miss(p[None, :], ro[None, :], v[None, :])[0]:
                    fired += 1
                    assert not oracle_intersect(planes, ro, v)[0], (name, k)
                checked += 1
    assert checked > 3000
    assert fired > 0.6 * checked


def certificate_of(planes, ro, rd, rng):
    """cert_kernel: the extreme planes of a missing central ray, or None."""
    t_near, near_idx, t_far, far_idx = F(-1e30), -1, F(1e30), -1
    outside = False
    for i, p in enumerate(planes):
        denom = dot_rows(p[None, :], rd[None, :])[0]
        side = F(p[3] + dot_rows(p[None, :], ro[None, :])[0])
        if not abs(denom) > F(1e-7):
            outside |= bool(side > 0)
            continue
        with np.errstate(divide="ignore"):
            t = F(-side / denom)
        if denom < 0 and t > t_near:
            t_near, near_idx = t, i
        if denom > 0 and t < t_far:
            t_far, far_idx = t, i
    if outside or not t_near > t_far or near_idx < 0 or far_idx < 0:
        return None
    return near_idx, far_idx


def certified(planes, cert, ro, rd):
    """camera_miss_certified for one ray."""
    i, j = cert
    di = dot_rows(planes[i][None, :], rd[None, :])[0]
    dj = dot_rows(planes[j][None, :], rd[None, :])[0]
    if not (abs(di) > F(1e-7) and di < 0 and abs(dj) > F(1e-7) and dj > 0):
        return False
    si = F(planes[i][3] + dot_rows(planes[i][None, :], ro[None, :])[0])
    sj = F(planes[j][3] + dot_rows(planes[j][None, :], ro[None, :])[0])
    return bool(F(-si / di) > F(-sj / dj))


def test_camera_miss_certificates_only_prove_misses():
    rng = np.random.default_rng(3)
    proved = checked = 0
    for name, planes in STONES:
        planes = np.ascontiguousarray(planes, F)
        o, d = random_rays(rng, 200)
        for k in range(len(o)):
            cert = certificate_of(planes, o[k], d[k], rng)
            if cert is None:
                continue
            # The pixel's other samples: the central direction jittered by up to a few
            # milliradians, some of them onto the stone.
            for _ in range(10):
                dj = (d[k] + rng.normal(size=3).astype(F) * F(10.0 ** rng.uniform(-4, -1.5))).astype(F)
                dj = (dj / np.sqrt(np.dot(dj, dj))).astype(F)
                checked += 1
                if certified(planes, cert, o[k], dj):
                    proved += 1
                    assert not oracle_intersect(planes, o[k], dj)[0], (name, k)
    assert checked > 1000
    assert proved > 0.5 * checked


def test_padding_planes_are_inert():
    # stage_planes pads the plane array to a multiple of four with (0, 0, 0, -1).
    rng = np.random.default_rng(4)
    for name, planes in STONES:
        planes = np.ascontiguousarray(planes, F)
        pad = (-len(planes)) % 4 or 4
        padded = np.concatenate([planes, np.tile(np.array([[0, 0, 0, -1]], F), (pad, 1))]).astype(F)
        o, d = random_rays(rng, 500)
        a = model_scan(planes, o, d, np.random.default_rng(9))
        b = model_scan(padded, o, d, np.random.default_rng(9))
        for x, y in zip(a[:3], b[:3]):
            assert np.array_equal(x, y), name


def test_vector_scan_of_the_oracle_matches_its_scalar_scan():
    """The oracle's tracer scans eight planes at a time (AVX2, as the reference's CPU tracer
    does, simd/slab.rs:186-270); it must select exactly what the one-plane-at-a-time scan
    selects, ties and near-parallel planes included."""
    rng = np.random.default_rng(9)
    for name, planes in STONES:
        planes = with_near_duplicates(np.ascontiguousarray(planes, F), rng, n=16)
        # A near-parallel plane (|denom| <= 1e-7 for rays along +x) in some stones.
        o, d = random_rays(rng, 1500)
        inner = (rng.normal(size=(600, 3)) * 0.15).astype(F)
        o = np.concatenate([o, inner]).astype(F)
        d = np.concatenate([d, random_rays(rng, 600)[1]]).astype(F)
        for k in range(len(o)):
            t1, n1, f1 = np.zeros(1, F), np.zeros(3, F), ctypes.c_int()
            t2, n2, f2 = np.zeros(1, F), np.zeros(3, F), ctypes.c_int()
            args = (po.fptr(planes), len(planes), po.fptr(po.f32arr(o[k])), po.fptr(po.f32arr(d[k])))
            h1 = L.or_intersect(*args, po.fptr(t1), po.fptr(n1), ctypes.byref(f1))
            h2 = L.or_intersect_scalar(*args, po.fptr(t2), po.fptr(n2), ctypes.byref(f2))
            assert h1 == h2, (name, k)
            if h1:
                assert t1.view(np.uint32)[0] == t2.view(np.uint32)[0] and f1.value == f2.value, (name, k)
