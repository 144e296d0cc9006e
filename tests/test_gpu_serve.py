"""The batched / served job loop (gr_job_*) through the C ABI (needs an MI355X).

A job's film must be bit-identical to gr_accumulate of the same scene and sample range,
alone and in any mix of jobs, whatever the tick size; admission follows the HBM budget;
cancellation, declines and metrics behave as the header states."""
import dataclasses
import time

import numpy as np
import pytest

import pyoracle as po
from gemray_studio_amd import _lib
from gemray_studio_amd import backend as B
from gemray_studio_amd import materials as M
from gemray_studio_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    b = B.GpuBackend.acquire()
    assert b.available, "no gfx950 device: the HIP path is required for these tests"
    yield b
    b.close()


def _mix():
    """Jobs of every material/environment class the batched kernel specialises on, with
    uneven plane counts, resolutions and sample counts."""
    srb = po.standard_round_brilliant()
    env = scenes.procedural_env(5, 128, 64)
    cam = B.camera_new(0.35, 0.28, 5.0, 18.0)
    cam2 = B.camera_new(0.9, 0.4, 4.0, 24.0)
    return [
        (B.SceneRef(cam, 40, 24, srb, M.by_name("Diamond"), 12, B.Studio("RingLights", 1.0, 0.85, 0.95)), 3, 8),
        (B.SceneRef(cam2, 17, 33, scenes.random_convex_stone(7, 40), M.by_name("Sapphire"), 16,
                    B.Studio("Daylight", 1.0, 0.4, 0.35)), 0, 12),
        (B.SceneRef(cam, 32, 32, scenes.oval_brilliant(), scenes.pleochroic_sapphire(3), 32, B.HdrMap(env)), 5, 6),
        (B.SceneRef(cam2, 24, 20, scenes.round_brilliant(), M.by_name("Ruby"), 16, B.HdrMap(env)), 1, 10),
        (B.SceneRef(cam, 16, 16, scenes.step_cut(), scenes.stress_material(), 64, B.HdrMap(env)), 2, 4),
        (B.SceneRef(cam, 20, 20, srb, M.furnace_material(), 12, B.HdrMap.uniform(2.5)), 0, 16),
        (B.SceneRef(cam2, 30, 18, scenes.random_convex_stone(9, 120), M.by_name("Tourmaline"), 16,
                    B.Studio("Incandescent", 1.3, 0.0, 1.2)), 7, 9),
    ]


def _accumulate(gpu, scene, first, spp):
    film = np.zeros((scene.width * scene.height, 3), np.float32)
    assert gpu.try_accumulate(scene, first, spp, film)
    return film


def _run_jobs(gpu, jobs):
    ids = [gpu.submit(sc, first, spp) for sc, first, spp in jobs]
    films = []
    for jid, (sc, _, _) in zip(ids, jobs):
        assert gpu.wait(jid, 60000) == _lib.GR_OK
        film = np.zeros((sc.width * sc.height, 3), np.float32)
        gpu.job_film(jid, film)
        films.append(film)
        gpu.release(jid)
    return films


def test_job_film_equals_accumulate_alone_and_in_the_mix(gpu):
    jobs = _mix()
    want = [_accumulate(gpu, *j) for j in jobs]
    alone = [_run_jobs(gpu, [j])[0] for j in jobs]
    for w, a in zip(want, alone):
        assert np.array_equal(w.view(np.uint32), a.view(np.uint32))
    # All jobs at once, with ticks small enough that every job is cut into many pieces.
    gpu.set_tick_tuples(1 << 16)
    try:
        mixed = _run_jobs(gpu, jobs)
    finally:
        gpu.set_tick_tuples(1 << 24)
    for w, m in zip(want, mixed):
        assert np.array_equal(w.view(np.uint32), m.view(np.uint32))
    # Reversed submission order and the default tick size.
    rev = _run_jobs(gpu, jobs[::-1])[::-1]
    for w, r in zip(want, rev):
        assert np.array_equal(w.view(np.uint32), r.view(np.uint32))
    # The other scheduling policies cut the frames differently; the films stay the same.
    gpu.set_tick_tuples(1 << 16)
    try:
        for policy, slack in ((_lib.GR_SERVE_EQUAL_SHARE, 0.0), (_lib.GR_SERVE_DEADLINE, 0.0),
                              (_lib.GR_SERVE_DEADLINE, 50.0)):
            gpu.set_policy(policy, slack)
            for w, m in zip(want, _run_jobs(gpu, jobs)):
                assert np.array_equal(w.view(np.uint32), m.view(np.uint32)), (policy, slack)
    finally:
        gpu.set_tick_tuples(1 << 24)
        gpu.set_policy(_lib.GR_SERVE_DEADLINE)


def test_job_counters_and_metrics(gpu):
    gpu.metrics_reset()
    jobs = _mix()[:3]
    ids = [gpu.submit(*j) for j in jobs]
    for jid in ids:
        assert gpu.wait(jid, 60000) == _lib.GR_OK
    infos = [gpu.job_info(jid) for jid in ids]
    for (sc, first, spp), info in zip(jobs, infos):
        assert info["state"] == "done"
        assert info["paths"] == sc.width * sc.height * spp
        assert info["events"]["segment"] >= info["paths"]
        assert info["latency_ms"] >= info["run_ms"] >= 0.0
        # The job's counters equal a plain render's.
        _accumulate(gpu, sc, first, spp)
        assert gpu.last_stats()["events"] == info["events"]
    m = gpu.metrics()
    assert m["completed"] == 3 and m["submitted"] == 3
    assert m["segments"] == sum(i["events"]["segment"] for i in infos)
    assert 0.0 < m["latency_p50_ms"] <= m["latency_p99_ms"] <= m["latency_max_ms"]
    assert m["running"] == 0 and m["queued"] == 0 and m["hbm_in_use"] == 0
    for jid in ids:
        gpu.release(jid)
    assert gpu.poll(ids[0]) == _lib.GR_ERR_NOT_FOUND


def test_late_small_job_overtakes_a_running_large_job(gpu):
    # About a second of work for the large job; the small one arrives 10 ms in.
    big = B.SceneRef(B.camera_new(0.35, 0.28, 5.0, 18.0), 1024, 1024, scenes.oval_brilliant(),
                     scenes.pleochroic_sapphire(3), 32, B.HdrMap(scenes.procedural_env(5, 128, 64)))
    small = _mix()[0]
    a = gpu.submit(big, 0, 1024)
    time.sleep(0.01)
    b = gpu.submit(*small)
    assert gpu.wait(b, 60000) == _lib.GR_OK
    assert gpu.poll(a) == _lib.GR_PENDING, "the small job should join while the large one runs"
    assert gpu.wait(a, 120000) == _lib.GR_OK
    assert gpu.job_info(b)["latency_ms"] < gpu.job_info(a)["latency_ms"]
    gpu.release(a)
    gpu.release(b)


def test_deadline_slack_zero_is_first_come_first_served(gpu):
    big = B.SceneRef(B.camera_new(0.35, 0.28, 5.0, 18.0), 512, 512, scenes.oval_brilliant(),
                     scenes.pleochroic_sapphire(3), 32, B.HdrMap(scenes.procedural_env(5, 128, 64)))
    small = _mix()[0]
    gpu.set_policy(_lib.GR_SERVE_DEADLINE, 0.0)
    try:
        a = gpu.submit(big, 0, 1024)
        time.sleep(0.01)
        b = gpu.submit(*small)
        assert gpu.wait(b, 60000) == _lib.GR_OK
        # The earlier job has the earlier deadline: it was done before the small one.
        assert gpu.poll(a) == _lib.GR_OK
        m = gpu.metrics()
        assert m["policy"] == _lib.GR_SERVE_DEADLINE and m["slack"] == 0.0
        assert m["rate_tuples_per_ms"] > 0.0
    finally:
        gpu.set_policy(_lib.GR_SERVE_DEADLINE)
    gpu.release(a)
    gpu.release(b)
    with pytest.raises(_lib.GrError):
        gpu.set_policy(7, 1.0)
    with pytest.raises(_lib.GrError):
        gpu.set_policy(_lib.GR_SERVE_DEADLINE, -1.0)


def test_cancel_frees_the_job(gpu):
    big = B.SceneRef(B.camera_new(0.35, 0.28, 5.0, 18.0), 1024, 1024, scenes.oval_brilliant(),
                     scenes.pleochroic_sapphire(3), 32, B.Studio())
    a = gpu.submit(big, 0, 1024)
    gpu.cancel(a)
    assert gpu.wait(a, 60000) == _lib.GR_CANCELLED
    film = np.zeros((1024 * 1024, 3), np.float32)
    with pytest.raises(_lib.GrError):
        gpu.job_film(a, film)
    gpu.release(a)
    assert gpu.metrics()["hbm_in_use"] == 0


def test_declines_at_submit(gpu):
    sc, first, spp = _mix()[0]
    biax = dataclasses.replace(M.by_name("Diamond"), biaxial_delta_beta_alpha=0.01)
    with pytest.raises(_lib.GrError) as e:
        gpu.submit(B.SceneRef(sc.camera, sc.width, sc.height, sc.planes, biax, sc.max_bounces, sc.environment),
                   first, spp)
    assert e.value.status == _lib.GR_DECLINE_MATERIAL
    with pytest.raises(_lib.GrError) as e:
        gpu.submit(sc, first, 0)
    assert e.value.status == _lib.GR_ERR_INVALID


def test_admission_follows_the_hbm_budget():
    jobs = _mix()[:4]
    # A 1 MiB budget: these jobs (films of at most 32x32 pixels, a 128x64 map) fit, a
    # 1024x1024 film (12 MiB) does not.
    b = B.GpuBackend.acquire(hbm_budget_bytes=1 << 20)
    try:
        huge = B.SceneRef(jobs[0][0].camera, 1024, 1024, jobs[0][0].planes, jobs[0][0].material, 12,
                          jobs[0][0].environment)
        with pytest.raises(_lib.GrError) as e:
            b.submit(huge, 0, 4)
        assert e.value.status == _lib.GR_DECLINE_BUDGET
        films = _run_jobs(b, jobs)
        m = b.metrics()
        assert m["completed"] == 4 and m["declined"] == 1
        assert m["hbm_budget"] == 1 << 20 and m["hbm_in_use"] == 0
        for (sc, first, spp), f in zip(jobs, films):
            assert np.array_equal(_accumulate(b, sc, first, spp).view(np.uint32), f.view(np.uint32))
    finally:
        b.close()
