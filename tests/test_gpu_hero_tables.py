"""Hero tables (gr_set_hero_tables) against per-path evaluation, through the C ABI (needs an
MI355X).

The tables hold, per hero offset, values the kernels otherwise evaluate per path with the
same device functions (the finish's colour-matching values and map lobes; an absorbing
material's dispersion indices and absorption sums), so every traced tuple, film, counter and
job result must be bit-identical with the tables read and with every path evaluating."""
import numpy as np
import pytest

from gemray_studio_amd import _lib
from gemray_studio_amd import backend as B
from gemray_studio_amd.workloads import single_job_scene
from test_gpu_parity import CASE_IDS, _cases
from test_gpu_serve import _mix

import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    b = B.GpuBackend.acquire()
    assert b.available, "no gfx950 device: the HIP path is required for these tests"
    yield b
    b.set_hero_tables(1)
    b.close()


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("idx", range(len(CASE_IDS)), ids=CASE_IDS)
def test_every_tuple_identical_with_and_without_tables(gpu, idx):
    """Every material and environment class of the per-path parity suite: mode 2 builds the
    material table even for this small frame, so the tables are read by every fill and
    finish (the scattering stones' fills go through the scatter step's cache fill)."""
    case = _cases(po.standard_round_brilliant())[idx]
    scene = case.product_scene()
    n_pix = case.width * case.height
    print(f"[hero] {case.name} {case.width}x{case.height}x4", flush=True)
    out = {}
    for mode in (0, 2):
        gpu.set_hero_tables(mode)
        xyz, info = gpu.trace_samples(scene, 0, n_pix, 9, 4)
        out[mode] = (xyz, info, gpu.last_stats()["events"])
    gpu.set_hero_tables(1)
    assert np.array_equal(_bits(out[0][0]), _bits(out[2][0])), f"{case.name}: XYZ differs"
    assert np.array_equal(out[0][1], out[2][1]), f"{case.name}: per-path records differ"
    assert out[0][2] == out[2][2]


def test_single_job_frame_identical_with_the_default_threshold(gpu):
    """The single-job workload's frame at 33 spp (68 M tuples, over the default mode's 2^26
    threshold, so its material table is built by the render itself) against every path
    evaluating."""
    sc, _ = single_job_scene()
    films, stats = {}, {}
    for mode in (0, 1):
        gpu.set_hero_tables(mode)
        ds = gpu.upload(sc)
        film = gpu.film(sc.width, sc.height)
        print(f"[hero] single job mode {mode}", flush=True)
        gpu.render(ds, 17, 33, film)
        stats[mode] = gpu.last_stats()["events"]
        films[mode] = film.download()
        film.free()
        ds.free()
    gpu.set_hero_tables(1)
    assert np.array_equal(_bits(films[0]), _bits(films[1]))
    assert stats[0] == stats[1]
    assert films[1].sum() > 0.0


def test_job_loop_identical_with_and_without_tables(gpu):
    """Mode 1 reads the spectral table; mode 2 also builds every absorbing job's material
    table at admission (these jobs are under the default threshold)."""
    films = {}
    for mode in (0, 1, 2):
        gpu.set_hero_tables(mode)
        ids = [gpu.submit(sc, first, spp) for sc, first, spp in _mix()]
        out = []
        for jid, (sc, _, _) in zip(ids, _mix()):
            assert gpu.wait(jid, 60000) == _lib.GR_OK
            f = np.zeros((sc.width * sc.height, 3), np.float32)
            gpu.job_film(jid, f)
            gpu.release(jid)
            out.append(f)
        films[mode] = out
    gpu.set_hero_tables(1)
    for mode in (1, 2):
        for a, b in zip(films[0], films[mode]):
            assert np.array_equal(_bits(a), _bits(b)), f"mode {mode}"
