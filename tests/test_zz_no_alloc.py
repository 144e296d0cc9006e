"""gr_render_noalloc, the entry point that allocates nothing (its workspace is the library's
fixed per-GPU device array), against gr_render through the C ABI (needs an MI355X): two
calls in a row, the second on a smaller frame with fewer samples, each film bit-identical
to gr_render's of the same scene and range, the same event counts, and an oversized
sample count rejected."""
import numpy as np
import pytest

from gemray_studio_amd import _lib
from gemray_studio_amd import backend as B
from gemray_studio_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    b = B.GpuBackend.acquire()
    assert b.available, "no gfx950 device: the HIP path is required for these tests"
    yield b
    b.close()


def film_bits(gpu, scene, first, spp, noalloc):
    ds = gpu.upload(scene)
    film = gpu.film(scene.width, scene.height)
    try:
        if noalloc:
            st = gpu._L.gr_render_noalloc(gpu._ctx, ds.handle, first, spp, film.handle)
            assert st == _lib.GR_OK, st
        else:
            gpu.render(ds, first, spp, film)
        stats = gpu.last_stats()
        return film.download().view(np.uint32).copy(), stats
    finally:
        film.free()
        ds.free()


def test_noalloc_render_matches_render_on_two_calls(gpu):
    env = B.HdrMap(scenes.procedural_env(5, 256, 128))
    cam = B.camera_new(0.35, 0.28, 5.0, 18.0)
    big = B.SceneRef(cam, 160, 120, scenes.oval_brilliant(), scenes.pleochroic_sapphire(3), 32, env)
    small = B.SceneRef(B.camera_new(0.9, 0.4, 4.0, 24.0), 48, 40, scenes.random_convex_stone(7, 40),
                       scenes.pleochroic_sapphire(4), 16, env)
    # 160 x 120 x 64 = 1.2 M tuples: three launches of the 2^19-tuple workspace; then a
    # smaller frame whose single launch reuses the array the first call left behind.
    for scene, first, spp in ((big, 5, 64), (small, 3, 8)):
        print(f"[noalloc] {scene.width}x{scene.height} x {spp} spp from {first}", flush=True)
        want, st_want = film_bits(gpu, scene, first, spp, noalloc=False)
        got, st_got = film_bits(gpu, scene, first, spp, noalloc=True)
        assert np.array_equal(want, got)
        assert st_got["events"] == st_want["events"]
        assert st_got["segments"] > 0
    big_spp = 65536 + 1  # GR_MAX_SAMPLES_PER_CALL + 1
    ds = gpu.upload(small)
    film = gpu.film(small.width, small.height)
    try:
        assert gpu._L.gr_render_noalloc(gpu._ctx, ds.handle, 0, big_spp, film.handle) == _lib.GR_ERR_INVALID
        assert gpu._L.gr_render_noalloc(gpu._ctx, ds.handle, 0, 0, film.handle) == _lib.GR_OK
        assert not film.download().any()
    finally:
        film.free()
        ds.free()
