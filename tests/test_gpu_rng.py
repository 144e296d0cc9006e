"""Tier 1 RNG bit-exactness on the device (renderer/gpu/rng_check.rs:1-35, :105-109): over about
10^6 (pixel, sample, bounce) tuples, the kernel's own sampling functions (gr_probe) against
the oracle integer-exactly -- the path seed, the three pixel-rotation hashes, the sample's
reversed bits, the per-bounce draw of every stream -- and, one step further than the
reference's Tier 2 budget, the stratified draws jx, jy and hero bit for bit too (the kernel
evaluates them with the same operations in the same order).  Needs an MI355X."""
import numpy as np
import pytest

import pyoracle as po
from gemray_studio_amd import _lib
from gemray_studio_amd import backend as B

from test_rng_constants import REFERENCE_SALTS

pytestmark = pytest.mark.gpu

# The per-bounce streams (sampling.rs:12-41); the pixel rotations are covered by the hashes.
BOUNCE_SALTS = [REFERENCE_SALTS[k] for k in (
    "FRESNEL_BRANCH_STREAM", "RUSSIAN_ROULETTE_STREAM", "BIREFRINGENT_SPLIT_STREAM", "MODE_COUPLING_STREAM",
    "FROSTED_DIR_U_STREAM", "FROSTED_DIR_V_STREAM", "DISTANCE_SAMPLE_STREAM", "PHASE_DIR_U_STREAM",
    "PHASE_DIR_V_STREAM")]


@pytest.fixture(scope="module")
def gpu():
    b = B.GpuBackend.acquire()
    assert b.available, "no gfx950 device: the HIP path is required for these tests"
    yield b
    b.close()


def tuples(n_pairs=262144, n_bounces=4, seed=7):
    """n_pairs (pixel, sample) pairs x n_bounces bounces: pixels over a 7680x4320 frame and
    the full u32 range, samples over the first 65536 and arbitrary absolute offsets."""
    rng = np.random.default_rng(seed)
    half = n_pairs // 2
    pixel = np.concatenate([rng.integers(0, 7680 * 4320, half), rng.integers(0, 2 ** 32, n_pairs - half)])
    sample = np.concatenate([rng.integers(0, 65536, half), rng.integers(0, 2 ** 32, n_pairs - half)])
    pixel[:4], sample[:4] = [0, 1, 0, 2 ** 32 - 1], [0, 0, 1, 2 ** 32 - 1]
    pixel = np.repeat(pixel.astype(np.uint32), n_bounces)
    sample = np.repeat(sample.astype(np.uint32), n_bounces)
    bounce = np.tile(np.arange(n_bounces, dtype=np.uint32), n_pairs)
    bounce[-n_bounces:] = [5, 16, 31, 63]
    return pixel, sample, bounce


def test_rng_streams_are_bit_exact(gpu):
    pixel, sample, bounce = tuples()
    n = len(pixel)
    assert n >= 10 ** 6
    ref = po.rng_records(pixel, sample, bounce, BOUNCE_SALTS)
    u32 = lambda a: a.view(np.uint32)  # noqa: E731

    # Seed and the rotation hashes, one probe per rotation stream (x = pixel ^ salt), and the
    # sample's reversed bits (x = sample).
    for col, name in ((1, "PIXEL_JITTER_X_ROTATION_STREAM"), (2, "PIXEL_JITTER_Y_ROTATION_STREAM"),
                      (3, "HERO_WAVELENGTH_ROTATION_STREAM")):
        x = pixel ^ np.uint32(REFERENCE_SALTS[name])
        out = u32(gpu.probe(_lib.GR_PROBE_SEED, np.stack([pixel, sample, x], axis=1)))
        assert np.array_equal(out[:, 0], ref[:, 0]), "path seed"
        assert np.array_equal(out[:, 1], ref[:, col]), name
    out = u32(gpu.probe(_lib.GR_PROBE_SEED, np.stack([pixel, sample, sample], axis=1)))
    assert np.array_equal(out[:, 2], ref[:, 4]), "reversed sample bits"

    # The stratified draws (radical inverses of bases 3 and 5 from the host's sample table).
    out = u32(gpu.probe(_lib.GR_PROBE_SAMPLE_DRAWS, np.stack([pixel, sample], axis=1)))
    assert np.array_equal(out, ref[:, 5:8]), "jx, jy, hero"

    # Every per-bounce stream draw, three salts per probe call.
    seed = ref[:, 0]
    for k in range(0, len(BOUNCE_SALTS), 3):
        salts = np.broadcast_to(np.array(BOUNCE_SALTS[k:k + 3], np.uint32), (n, 3))
        out = u32(gpu.probe(_lib.GR_PROBE_STREAM_DRAW, np.concatenate([np.stack([seed, bounce], axis=1), salts], 1)))
        assert np.array_equal(out, ref[:, 8 + k:11 + k]), f"stream draws {k}..{k + 2}"


def test_rgb_to_spectral_closed_forms_on_the_device(gpu):
    """The kernel's rgb_to_spectral_radiance: each lobe is 1 at its centre and exp(-1/2) one
    sigma out on either side (the widths of env_map_spectrum.rs:65-71), within the native
    exponential's budget; and against the oracle on random inputs over the visible range."""
    lobes = [(0, 615.0, 45.0, 65.0), (1, 545.0, 45.0, 45.0), (2, 465.0, 40.0, 45.0)]
    rows = []
    for c, mu, lo, hi in lobes:
        rgb = [0.0, 0.0, 0.0]
        rgb[c] = 1.0
        rows += [rgb + [mu], rgb + [mu - lo], rgb + [mu + hi]]
    out = gpu.probe(_lib.GR_PROBE_RGB_SPECTRAL, np.array(rows, np.float32))[:, 0].reshape(3, 3)
    assert np.all(out[:, 0] == 1.0)
    assert np.all(np.abs(out[:, 1:] - np.exp(-0.5)) <= 3e-7)
    rng = np.random.default_rng(3)
    n = 20000
    x = np.concatenate([rng.uniform(0, 4, (n, 3)), rng.uniform(380, 780, (n, 1))], 1).astype(np.float32)
    x[:100, :3] *= -1.0  # negative components clamp to zero
    dev = gpu.probe(_lib.GR_PROBE_RGB_SPECTRAL, x)[:, 0]
    ref = np.array([po.rgb_to_spectral_radiance(r[:3], r[3]) for r in x], np.float32)
    scale = np.maximum(np.abs(x[:, :3]).sum(1), 1e-6)
    assert np.max(np.abs(dev - ref) / scale) < 1e-6
    assert np.all(dev[:100] == 0.0)
