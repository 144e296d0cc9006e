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
from scene_util import frosted_girdle

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
        (B.SceneRef(cam, 22, 26, srb, M.by_name("Zircon"), 16, B.HdrMap(env), facet_finishes=frosted_girdle(74)),
         4, 7),
        (B.SceneRef(cam2, 26, 22, scenes.oval_brilliant(), M.by_name("Tanzanite"), 24, B.HdrMap(env)), 3, 9),
        (B.SceneRef(cam, 18, 24, srb, M.by_name("Topaz"), 12, B.Studio("Daylight", 1.0, 0.4, 0.35)), 6, 5),
        (B.SceneRef(cam2, 20, 16, srb, dataclasses.replace(M.by_name("Ruby"), scattering_sigma_s=0.8,
                                                           scattering_g=-0.3), 16, B.HdrMap(env)), 2, 7),
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


def test_largest_stone_shares_a_launch_with_small_ones(gpu):
    """A 1024-plane job (the runtime's limit) in one class launch with 17- and 74-plane
    jobs: the launch stages planes in dynamic LDS sized to its largest scene, and every
    job's film equals its own accumulate bit for bit, alone and cut into small ticks."""
    srb = po.standard_round_brilliant()
    env = scenes.procedural_env(5, 128, 64)
    cam = B.camera_new(0.35, 0.28, 5.0, 18.0)
    ruby = M.by_name("Ruby")
    jobs = [
        (B.SceneRef(cam, 28, 20, scenes.random_convex_stone(29, 1024), ruby, 16, B.HdrMap(env)), 1, 8),
        (B.SceneRef(cam, 20, 24, scenes.random_convex_stone(3, 17), ruby, 16, B.HdrMap(env)), 0, 12),
        (B.SceneRef(cam, 24, 24, srb, ruby, 16, B.HdrMap(env)), 2, 6),
        (B.SceneRef(cam, 16, 20, scenes.random_convex_stone(31, 1024), M.by_name("Alexandrite"), 16,
                    B.HdrMap(env)), 0, 8),
    ]
    want = [_accumulate(gpu, *j) for j in jobs]
    for w, a in zip(want, [_run_jobs(gpu, [j])[0] for j in jobs]):
        assert np.array_equal(w.view(np.uint32), a.view(np.uint32))
    gpu.set_tick_tuples(1 << 16)
    try:
        mixed = _run_jobs(gpu, jobs)
    finally:
        gpu.set_tick_tuples(1 << 24)
    for w, m in zip(want, mixed):
        assert np.array_equal(w.view(np.uint32), m.view(np.uint32))
    assert all(np.isfinite(w).all() and w.max() > 0 for w in want)


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


def _drain(gpu, jid, n_pixels, timeout_s=60.0):
    """take_delta until the job settles (the worker's emitter loop), then once more."""
    deltas = []
    t0 = time.time()
    while gpu.poll(jid) == _lib.GR_PENDING:
        d = gpu.take_delta(jid, n_pixels)
        if d is not None:
            deltas.append(d)
        time.sleep(0.002)
        assert time.time() - t0 < timeout_s
    d = gpu.take_delta(jid, n_pixels)
    if d is not None:
        deltas.append(d)
    return deltas


def _batch_films(gpu, scene, first, spp, k):
    """gr_accumulate of each fixed sub-batch [first + i k, ...) of the range."""
    out = []
    s = first
    while s < first + spp:
        n = min(k, first + spp - s)
        out.append((s, n, _accumulate(gpu, scene, s, n)))
        s += n
    return out


def _check_fixed_deltas(deltas, batches, first, spp):
    # The taken deltas tile the range in order, and each is its sub-batches' films summed
    # in sample order (PendingDelta::add), bit for bit.
    assert deltas[0][0] == first and sum(d[1] for d in deltas) == sum(b[1] for b in batches)
    bi = 0
    for d_first, d_n, d in deltas:
        assert d_first == batches[bi][0]
        acc, covered = None, 0
        while covered < d_n:
            _, n, f = batches[bi]
            acc = f.copy() if acc is None else (acc + f).astype(np.float32)
            covered += n
            bi += 1
        assert covered == d_n
        assert np.array_equal(acc.view(np.uint32), d.view(np.uint32))
    return bi


def test_progressive_fixed_subbatches_are_bit_exact(gpu):
    sc, first, spp = _mix()[2][0], 5, 10
    n = sc.width * sc.height
    batches = _batch_films(gpu, sc, first, spp, 3)
    jid = gpu.submit(sc, first, spp, subbatch_ms=100, subbatch_samples=3, subbatch_fixed=True)
    deltas = _drain(gpu, jid, n)
    assert gpu.wait(jid, 60000) == _lib.GR_OK
    assert _check_fixed_deltas(deltas, batches, first, spp) == len(batches)
    # The running total: the sub-batch films added in order (tracer.rs:98-100).
    total = np.zeros((n, 3), np.float32)
    for _, _, f in batches:
        total = (total + f).astype(np.float32)
    film = np.zeros((n, 3), np.float32)
    gpu.job_film(jid, film)
    assert np.array_equal(film.view(np.uint32), total.view(np.uint32))
    info = gpu.job_info(jid)
    assert info["samples_done"] == spp and info["subbatches"] == len(batches) == 4
    assert gpu.take_delta(jid, n) is None
    gpu.release(jid)


def test_progressive_subbatches_in_a_mix_equal_alone(gpu):
    jobs = _mix()
    sc, first, spp = jobs[4][0], 2, 9
    n = sc.width * sc.height
    batches = _batch_films(gpu, sc, first, spp, 2)
    gpu.set_tick_tuples(1 << 16)
    try:
        others = [gpu.submit(*j) for j in jobs[:4]]
        jid = gpu.submit(sc, first, spp, subbatch_ms=100, subbatch_samples=2, subbatch_fixed=True)
        deltas = _drain(gpu, jid, n)
        for o in others:
            assert gpu.wait(o, 60000) == _lib.GR_OK
            gpu.release(o)
    finally:
        gpu.set_tick_tuples(1 << 24)
    assert _check_fixed_deltas(deltas, batches, first, spp) == len(batches)
    gpu.release(jid)


def test_progressive_adaptive_subbatches_cover_the_range(gpu):
    sc = B.SceneRef(B.camera_new(0.35, 0.28, 5.0, 18.0), 192, 128, scenes.oval_brilliant(),
                    scenes.pleochroic_sapphire(3), 32, B.HdrMap(scenes.procedural_env(5, 128, 64)))
    first, spp = 11, 96
    n = sc.width * sc.height
    jid = gpu.submit(sc, first, spp, subbatch_ms=4)
    deltas = _drain(gpu, jid, n)
    assert gpu.wait(jid, 60000) == _lib.GR_OK
    # Contiguous ranges from the first sample to the end.
    at = first
    for d_first, d_n, _ in deltas:
        assert d_first == at and d_n > 0
        at += d_n
    assert at == first + spp
    info = gpu.job_info(jid)
    assert info["samples_done"] == spp and info["subbatches"] >= 3
    # Sub-batch sums add up to the whole range (coalesced_deltas_sum_identically_to_un_coalesced_ones,
    # stream_emit/tests.rs:34-55: within 1e-3 relative).
    direct = _accumulate(gpu, sc, first, spp)
    total = np.zeros((n, 3), np.float32)
    for _, _, d in deltas:
        total = (total + d).astype(np.float32)
    scale = np.maximum(np.maximum(np.abs(total), np.abs(direct)), 1e-6)
    assert np.max(np.abs(total - direct) / scale) < 1e-3
    film = np.zeros((n, 3), np.float32)
    gpu.job_film(jid, film)
    assert np.max(np.abs(film - direct) / scale) < 1e-3
    gpu.release(jid)


def test_progressive_cancel_keeps_finished_subbatches(gpu):
    sc = B.SceneRef(B.camera_new(0.35, 0.28, 5.0, 18.0), 512, 512, scenes.oval_brilliant(),
                    scenes.pleochroic_sapphire(3), 32, B.HdrMap(scenes.procedural_env(5, 128, 64)))
    n = sc.width * sc.height
    jid = gpu.submit(sc, 0, 1024, subbatch_ms=100, subbatch_samples=4, subbatch_fixed=True)
    got = []
    t0 = time.time()
    while not got:
        d = gpu.take_delta(jid, n)
        if d is not None:
            got.append(d)
        time.sleep(0.002)
        assert time.time() - t0 < 60
    gpu.cancel(jid)
    assert gpu.wait(jid, 60000) == _lib.GR_CANCELLED
    d = gpu.take_delta(jid, n)
    if d is not None:
        got.append(d)
    info = gpu.job_info(jid)
    assert 4 <= info["samples_done"] < 1024 and info["samples_done"] % 4 == 0
    assert sum(x[1] for x in got) == info["samples_done"]
    # What was delivered is exact: the first delivered range equals its sub-batches' films.
    f0, n0, d0 = got[0]
    assert f0 == 0
    acc = None
    for s in range(0, n0, 4):
        f = _accumulate(gpu, sc, s, 4)
        acc = f if acc is None else (acc + f).astype(np.float32)
    assert np.array_equal(acc.view(np.uint32), d0.view(np.uint32))
    gpu.release(jid)
    assert gpu.metrics()["hbm_in_use"] == 0


def test_device_env_map_shared_by_scenes_and_jobs(gpu):
    """A map uploaded once (gr_env_map_upload) gives the same films as the host texels, in
    plain calls and in the job loop, and stays alive while jobs reference it after the
    caller's release."""
    env_rgb = scenes.procedural_env(6, 256, 128)
    dev = gpu.upload_env_map(env_rgb)
    cam = B.camera_new(0.35, 0.28, 5.0, 18.0)
    jobs = []
    for k, (w, h, spp) in enumerate([(40, 32, 8), (24, 24, 16), (33, 17, 5)]):
        planes = scenes.random_convex_stone(30 + k, 20 + 40 * k)
        mat = scenes.pleochroic_sapphire(40 + k)
        jobs.append((B.SceneRef(cam, w, h, planes, mat, 32, B.HdrMap(env_rgb)),
                     B.SceneRef(cam, w, h, planes, mat, 32, B.HdrMap(env_rgb, device=dev)), 2 * k, spp))
    for host, shared, first, spp in jobs:
        a, b = _accumulate(gpu, host, first, spp), _accumulate(gpu, shared, first, spp)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    ids = [gpu.submit(shared, first, spp) for _, shared, first, spp in jobs]
    dev.release()  # the jobs keep their own references
    for jid, (host, shared, first, spp) in zip(ids, jobs):
        assert gpu.wait(jid, 60000) == _lib.GR_OK
        film = np.zeros((shared.width * shared.height, 3), np.float32)
        gpu.job_film(jid, film)
        gpu.release(jid)
        assert np.array_equal(film.view(np.uint32), _accumulate(gpu, host, first, spp).view(np.uint32))
