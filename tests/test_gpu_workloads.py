"""The batched and served workloads of bench.py, whole, through the job loop (needs an
MI355X).

* Batched: the 64 seeded convex stones (17-181 planes, +-3 deg, isotropic / uniaxial mix,
  128^2-1024^2 px, 64-1024 spp, studio presets, depth 16), all submitted at once.  Every
  job film is bit-identical to gr_accumulate of the same job (the reference's partition
  contract, gemray-net/tests/partition_correctness.rs:126-231: a result depends only on
  (pixel, absolute sample)); row windows of stones across the plane-count range and both
  material classes hold the per-path parity tiers against the oracle; every stone's
  gemological metrics at its job pose are bit-identical to the oracle's
  (color/metrics.rs:1082-1248).
* Served: the 40-job request mix (Poisson arrivals at 20 jobs/s, the single-job material /
  environment family, 10 % at 1920x1080 x 512 spp) submitted at its arrival times while
  earlier jobs run: every film is bit-identical to the same job traced alone, two stones'
  row windows hold the per-path tiers against the oracle, and the loop reports p50 / p99
  latency.
"""
import time

import numpy as np
import pytest

import pyoracle as po
from gemray_studio_amd import _lib
from gemray_studio_amd import backend as B
from gemray_studio_amd import materials as M
from gemray_studio_amd import scenes
from gemray_studio_amd import workloads
from scene_util import Case

pytestmark = pytest.mark.gpu

SEQ_MATCH_MIN = 0.995
FATE_MATCH_MIN_ANISO = 0.97
XYZ_RTOL = 2e-4


@pytest.fixture(scope="module")
def gpu():
    b = B.GpuBackend.acquire()
    assert b.available, "no gfx950 device: the HIP path is required for these tests"
    yield b
    b.close()


def accumulate(gpu, scene, first, spp):
    film = np.zeros((scene.width * scene.height, 3), np.float32)
    assert gpu.try_accumulate(scene, first, spp, film)
    return film


def collect(gpu, ids, jobs):
    films = []
    for jid, (sc, _, _) in zip(ids, jobs):
        assert gpu.wait(jid, 600000) == _lib.GR_OK
        film = np.zeros((sc.width * sc.height, 3), np.float32)
        gpu.job_film(jid, film)
        films.append(film)
    return films


def window_tiers(gpu, case, row, spp):
    """The per-path tiers of test_gpu_parity on one row of the frame (every tuple of the
    window against the oracle)."""
    w = case.width
    begin = row * w
    xyz_g, info_g = gpu.trace_samples(case.product_scene(), begin, w, 0, spp)
    _, xyz_o, info_o, _ = case.oracle_render(0, spp, per_sample=True, pixel_begin=begin, pixel_end=begin + w)
    seq = (info_g[:, 0] == info_o["segments"]) & (info_g[:, 1] == info_o["termination"]) & (
        info_g[:, 2] == info_o["facet_hash"])
    assert seq.mean() >= SEQ_MATCH_MIN, (case.name, seq.mean())
    same = seq & (info_g[:, 3] == info_o["fate_hash"])
    fate_min = FATE_MATCH_MIN_ANISO if case.material.is_anisotropic else SEQ_MATCH_MIN
    assert same.sum() / max(seq.sum(), 1) >= fate_min, case.name
    # T2 as on the full-size frames: matched paths within XYZ_RTOL of their own brightest
    # component at the 99.9th percentile (a rare dim path with a companion channel just
    # short of its critical angle may reach 1e-2 of itself) and every one within XYZ_RTOL of
    # the window's brightest path.
    scale = np.maximum(np.abs(xyz_o).max(axis=1), 1e-30)
    err = (np.abs(xyz_g - xyz_o).max(axis=1) / scale)[same]
    assert np.quantile(err, 0.999) <= XYZ_RTOL and np.median(err) <= 2e-6, (case.name, err.max())
    assert err.max() <= 1e-2, (case.name, err.max())
    assert np.abs(xyz_g - xyz_o).max(axis=1)[same].max() <= XYZ_RTOL * np.abs(xyz_o).max(), case.name
    assert np.mean(info_g[:, 0] > 1) > 0.05, case.name  # the window crosses the stone
    print(f"[workloads] {case.name} row {row}: seq={seq.mean():.4f} worst={err.max():.2e}", flush=True)


def test_batched_workload(gpu):
    poses, cams = [], []
    jobs = workloads.batched_jobs(poses=poses, cams=cams)
    assert len(jobs) == 64
    planes = [sc.planes.shape[0] for sc, _, _ in jobs]
    assert min(planes) >= 17 and max(planes) <= 181
    print(f"[workloads] batched: {len(jobs)} jobs, "
          f"{sum(sc.width * sc.height * spp for sc, _, spp in jobs)} paths", flush=True)
    want = [accumulate(gpu, *j) for j in jobs]
    gpu.metrics_reset()
    ids = [gpu.submit(*j) for j in jobs]
    films = collect(gpu, ids, jobs)
    for k, (w, f) in enumerate(zip(want, films)):
        assert np.array_equal(w.view(np.uint32), f.view(np.uint32)), f"job {k}"
        assert np.all(np.isfinite(f)) and float(f.sum()) > 0.0
    m = gpu.metrics()
    assert m["completed"] >= len(jobs)
    for jid in ids:
        gpu.release(jid)

    # Oracle windows: the fewest- and most-faceted stones and one in between, covering both
    # material classes.
    order = np.argsort(planes)
    picks = {int(order[0]), int(order[-1]), int(order[len(order) // 2])}
    aniso = [k for k in range(len(jobs)) if jobs[k][0].material.is_anisotropic]
    iso = [k for k in range(len(jobs)) if not jobs[k][0].material.is_anisotropic]
    if not any(k in aniso for k in picks):
        picks.add(aniso[0])
    if not any(k in iso for k in picks):
        picks.add(iso[0])
    for k in sorted(picks):
        sc = jobs[k][0]
        case = Case(f"job{k}-P{planes[k]}-{sc.material.name}", sc.material, sc.planes, sc.environment, sc.width,
                    sc.height, sc.max_bounces, camera=cams[k])
        window_tiers(gpu, case, sc.height // 2, 4)

    # Gemological metrics of every stone at its job's camera pose, bit for bit, all 64 stones
    # reduced on the device in one launch (and equal to one call per stone).
    batch = gpu.optical_metrics_batch([(sc.planes, sc.material, [pose]) for (sc, _, _), pose in zip(jobs, poses)])
    for k, ((sc, _, _), pose) in enumerate(zip(jobs, poses)):
        dev = gpu.optical_metrics(sc.planes, sc.material, [pose])[0]
        assert batch[k][0] == dev, k
        ref = po.gem_optical_metrics(sc.planes, M.to_struct(sc.material), *pose)
        for i, name in enumerate(po.METRIC_NAMES):
            assert np.float32(dev[name]) == ref[i], (k, name, dev[name], ref[i])


def test_served_workload(gpu):
    env = scenes.procedural_env(seed=2, width=2048, height=1024)
    cams = []
    jobs = workloads.served_jobs(env, cams=cams)
    arrivals = workloads.served_arrivals(len(jobs))
    print(f"[workloads] served: {len(jobs)} jobs over {arrivals[-1]:.2f} s of arrivals", flush=True)
    want = [accumulate(gpu, *j) for j in jobs]
    gpu.metrics_reset()
    ids = []
    t0 = time.perf_counter()
    for k, j in enumerate(jobs):
        delay = t0 + arrivals[k] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(gpu.submit(*j))
    films = collect(gpu, ids, jobs)
    for k, (w, f) in enumerate(zip(want, films)):
        assert np.array_equal(w.view(np.uint32), f.view(np.uint32)), f"served job {k}"
    m = gpu.metrics()
    assert m["completed"] >= len(jobs)
    assert 0.0 < m["latency_p50_ms"] <= m["latency_p99_ms"] <= m["latency_max_ms"]
    for jid in ids:
        gpu.release(jid)
    # Two stones of the mix against the oracle: a small random stone and a full-frame job.
    small = next(k for k, (sc, _, _) in enumerate(jobs) if sc.width < 512)
    full = next(k for k, (sc, _, _) in enumerate(jobs) if sc.width == 1920)
    for k in (small, full):
        sc = jobs[k][0]
        case = Case(f"served{k}", sc.material, sc.planes, sc.environment, sc.width, sc.height, sc.max_bounces,
                    camera=cams[k])
        window_tiers(gpu, case, sc.height // 2, 4)
