"""The autotuner on the device (needs an MI355X): every configuration it can choose -- each
occupancy variant, split bounce, early-parking threshold, continuation capacity, launch size
and trace mode --
renders a film bit-identical to the default configuration's, and so do the tuned choice and
a render with autotuning on."""
import numpy as np
import pytest

from gemray_studio_amd import backend as B
from gemray_studio_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    b = B.GpuBackend.acquire()
    assert b.available, "no gfx950 device: the HIP path is required for these tests"
    yield b
    b.close()


BASE = {"waves_per_eu": 3, "split_bounce": 6, "cont_div": 4, "trace_mode": 0, "chunk_tuples": 1 << 25,
        "park_lanes": 16}
VARIANTS = ([dict(BASE, waves_per_eu=w) for w in (2, 4)] + [dict(BASE, split_bounce=s) for s in (0, 3, 4, 5, 7, 8, 10)]
            + [dict(BASE, park_lanes=p) for p in (0, 1, 8, 24, 32, 64)]
            + [dict(BASE, cont_div=d) for d in (1, 2, 8, 64)] + [dict(BASE, chunk_tuples=c) for c in (1 << 14, 1 << 17)]
            + [dict(BASE, split_bounce=3, cont_div=1), dict(BASE, split_bounce=4, park_lanes=64, cont_div=64)]
            + [dict(BASE, trace_mode=1), dict(BASE, waves_per_eu=4, split_bounce=3, cont_div=8, chunk_tuples=1 << 15)])


def film_bits(gpu, ds, scene, first, spp, choice=None):
    film = gpu.film(scene.width, scene.height)
    try:
        if choice is None:
            gpu.render(ds, first, spp, film)
        else:
            gpu.render_tuned(ds, first, spp, film, choice)
        return film.download().view(np.uint32).copy()
    finally:
        film.free()


@pytest.mark.parametrize("name", ["pleochroic-oval-map", "diamond-studio", "alexandrite-map"])
def test_every_configuration_is_bit_identical(gpu, name):
    env = B.HdrMap(scenes.procedural_env(5, 256, 128))
    cam = B.camera_new(0.35, 0.28, 5.0, 18.0)
    from gemray_studio_amd import materials as M

    scene = {
        "pleochroic-oval-map": B.SceneRef(cam, 96, 64, scenes.oval_brilliant(), scenes.pleochroic_sapphire(3), 32, env),
        "diamond-studio": B.SceneRef(cam, 80, 64, scenes.round_brilliant(), M.by_name("Diamond"), 16,
                                     B.Studio("RingLights", 1.0, 0.85, 0.95)),
        "alexandrite-map": B.SceneRef(cam, 64, 64, scenes.random_convex_stone(5, 90), M.by_name("Alexandrite"), 24,
                                      env),
    }[name]
    ds = gpu.upload(scene)
    try:
        want = film_bits(gpu, ds, scene, 3, 16)
        for v in VARIANTS:
            assert np.array_equal(film_bits(gpu, ds, scene, 3, 16, v), want), (name, v)
        choice = gpu.autotune(ds, 16)
        print(f"[autotune] {name}: {choice}", flush=True)
        assert choice["probes"] >= 10 and choice["probe_ms"] > 0
        assert np.array_equal(film_bits(gpu, ds, scene, 3, 16, choice), want)
        gpu.set_autotune(True)
        try:
            assert np.array_equal(film_bits(gpu, ds, scene, 3, 16), want)
        finally:
            gpu.set_autotune(False)
    finally:
        ds.free()
