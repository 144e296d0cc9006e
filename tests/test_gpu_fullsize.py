"""The HIP path at the full sizes of the single-job and stress configurations (needs an
MI355X).

* Single job: the pleochroic uniaxial sapphire on the 69-plane oval brilliant in the
  2048 x 1024 procedural HDR environment, 1920 x 1080, depth 32 (bench.py's workload).
* Stress: the Abbe-20, n_e - n_o = 0.17 stone on the 233-plane step cut in the same
  environment, 4096 x 4096, depth 64.

At these sizes the oracle cannot trace whole frames in a test, so each frame is checked
through
  - the per-path tiers of test_gpu_parity on row windows spread over the frame (every
    tuple of a window against the oracle: hero sequence, channel fates, XYZ);
  - size-independent identities of the whole frame: sample-range additivity (the
    reference's partition contract, partition_correctness.rs:126-231 / hybrid.rs:20-44:
    rendering [0, 8) then [8, 16) into one film equals the two films added) and chunk
    invariance (any chunk size gives the same bits), both bit for bit.
"""
import numpy as np
import pytest

from gemray_studio_amd import backend as B
from gemray_studio_amd import workloads
from gemray_studio_amd import scenes
from scene_util import Case

pytestmark = pytest.mark.gpu

SEQ_MATCH_MIN = 0.995
FATE_MATCH_MIN_ANISO = 0.97
AGREE_MIN_ANISO = 0.99
XYZ_RTOL = 2e-4
XYZ_RTOL_LOOSE = 1e-3
XYZ_RTOL_WORST = 1e-2


@pytest.fixture(scope="module")
def gpu():
    b = B.GpuBackend.acquire()
    assert b.available, "no gfx950 device: the HIP path is required for these tests"
    yield b
    b.close()


def single_job_case():
    scene, cam = workloads.single_job_scene()
    return Case("single-job", scene.material, scene.planes, scene.environment, scene.width, scene.height,
                scene.max_bounces, camera=cam)


def stress_case():
    scene, cam = workloads.stress_scene()
    return Case("stress", scene.material, scene.planes, scene.environment, scene.width, scene.height,
                scene.max_bounces, camera=cam)


def window_parity(gpu, case, rows, spp):
    w = case.width
    for row in rows:
        begin = row * w
        print(f"[fullsize] {case.name} row {row}: {w} px x {spp} spp", flush=True)
        xyz_g, info_g = gpu.trace_samples(case.product_scene(), begin, w, 0, spp)
        _, xyz_o, info_o, _ = case.oracle_render(0, spp, per_sample=True, pixel_begin=begin, pixel_end=begin + w)
        seq = (info_g[:, 0] == info_o["segments"]) & (info_g[:, 1] == info_o["termination"]) & (
            info_g[:, 2] == info_o["facet_hash"])
        assert seq.mean() >= SEQ_MATCH_MIN, (case.name, row, seq.mean())
        same = seq & (info_g[:, 3] == info_o["fate_hash"])
        assert same.sum() / max(seq.sum(), 1) >= FATE_MATCH_MIN_ANISO, (case.name, row)
        scale = np.maximum(np.abs(xyz_o).max(axis=1), 1e-30)
        err = np.abs(xyz_g - xyz_o).max(axis=1) / scale
        assert (seq & (err <= XYZ_RTOL_LOOSE)).mean() >= AGREE_MIN_ANISO, (case.name, row)
        e = err[same]
        k = int(np.flatnonzero(same)[np.argmax(e)])
        print(f"[fullsize]   matched paths: max={e.max():.3e} p99.9={np.quantile(e, 0.999):.3e} "
              f"median={np.median(e):.3e}; worst: pixel {begin + k // spp} sample {k % spp} segments "
              f"{info_g[k, 0]} gpu {xyz_g[k].tolist()} oracle {xyz_o[k].tolist()}", flush=True)
        # T2 on matched paths.  Nearly all are within XYZ_RTOL of their own brightest
        # component; a rare dim path keeps a companion channel just short of its critical
        # angle, where cos(theta_t) = sqrt(1 - sin^2(theta_t)) turns the few-ulp n_eff of the
        # kernel's identities into a larger relative change of that channel's transmission
        # (seen once in ~8000 paths of the stress stone: 4e-3 of a path 1e-5 times dimmer
        # than the frame's bright paths, while both oracle builds -- glibc and correctly
        # rounded libm -- agree on it bit for bit).  Measured against the window's
        # brightest path, every matched path is within XYZ_RTOL.
        assert np.quantile(e, 0.999) <= XYZ_RTOL and np.median(e) <= 2e-6, (case.name, row)
        assert e.max() <= XYZ_RTOL_WORST, (case.name, row, e.max())
        abs_err = np.abs(xyz_g - xyz_o).max(axis=1)[same]
        assert abs_err.max() <= XYZ_RTOL * np.abs(xyz_o).max(), (case.name, row)
        # The window is not all background: paths enter the stone.
        assert np.mean(info_g[:, 0] > 1) > 0.05


def frame_identities(gpu, case, spp):
    n = case.width * case.height
    sc = case.product_scene()
    f1 = np.zeros((n, 3), np.float32)
    f2 = np.zeros((n, 3), np.float32)
    both = np.zeros((n, 3), np.float32)
    gpu.accumulate(sc, 0, spp, f1)
    gpu.accumulate(sc, spp, spp, f2)
    gpu.accumulate(sc, 0, spp, both)
    gpu.accumulate(sc, spp, spp, both)
    assert np.array_equal(both.view(np.uint32), (f1 + f2).astype(np.float32).view(np.uint32))
    gpu.set_chunk_tuples(1 << 22)
    try:
        c = np.zeros((n, 3), np.float32)
        gpu.accumulate(sc, 0, spp, c)
    finally:
        gpu.set_chunk_tuples(1 << 25)
    assert np.array_equal(c.view(np.uint32), f1.view(np.uint32))
    assert np.all(np.isfinite(f1)) and float(f1.sum()) > 0.0


@pytest.mark.parametrize("make,rows,spp,frame_spp", [
    (single_job_case, (360, 580, 760), 8, 8),
    (stress_case, (1400, 2048, 2700), 2, 2),
], ids=["single-job-1920x1080", "stress-4096x4096"])
def test_full_size_configuration(gpu, make, rows, spp, frame_spp):
    case = make()
    window_parity(gpu, case, rows, spp)
    frame_identities(gpu, case, frame_spp)
