"""The worker's progressive-delivery rules as the job loop implements them
(gemray_studio_amd/csrc/progressive.h: next_batch_size, PendingDelta), held to the
reference's own tests, apps/gemray-worker/src/stream_emit/tests.rs:34-92.  The header is
host C++ only; each test compiles a small driver against it with g++.  CPU only."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gemray_studio_amd", "csrc")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "progressive.h"
using gr::next_batch_size;
int main(int argc, char **argv) {
    const unsigned target = 100;  // TARGET_SUBBATCH (sizing.rs:12)
    if (argc == 1) {
        // next_batch_size_grows_when_well_under_budget / _shrinks_when_over_budget /
        // _never_returns_zero (tests.rs:74-92), plus the x4 growth cap and the zero-elapsed rule.
        std::printf("%u %u %u %u %u %u %u\n", next_batch_size(4, 10.0, target), next_batch_size(100, 400.0, target),
                    next_batch_size(0, 1.0, target), next_batch_size(1, 10000.0, target),
                    next_batch_size(4, 1.0, target), next_batch_size(7, 0.0, target),
                    next_batch_size(0xFFFFFFFFu, 1.0, target));
        // pending_delta_take_returns_none_when_empty / _is_empty_again_immediately_after_take.
        gr::PendingDelta d(3 * 16);
        unsigned first = 77, samples = 77;
        std::vector<float> out(3 * 16, -1.0f), a(3 * 16, 0.5f);
        const bool t0 = d.take(&first, &samples, out.data());
        d.add(0, 2, a.data());
        const bool t1 = d.take(&first, &samples, out.data());
        const unsigned f1 = first, s1 = samples;
        const bool t2 = d.take(&first, &samples, out.data());
        std::printf("%d %d %d %u %u %g\n", t0, t1, t2, f1, s1, out[5]);
        return 0;
    }
    // Coalescing: argv = n_floats, a.bin (samples [0, 3)), b.bin (samples [3, 8)), out.bin.
    const size_t n = std::strtoul(argv[1], nullptr, 10);
    std::vector<float> a(n), b(n), out(n);
    FILE *fa = std::fopen(argv[2], "rb"), *fb = std::fopen(argv[3], "rb");
    if (!fa || !fb || std::fread(a.data(), 4, n, fa) != n || std::fread(b.data(), 4, n, fb) != n) return 2;
    std::fclose(fa);
    std::fclose(fb);
    gr::PendingDelta d(n);
    d.add(0, 3, a.data());
    d.add(3, 5, b.data());
    unsigned first = 0, samples = 0;
    if (!d.take(&first, &samples, out.data())) return 3;
    std::printf("%u %u\n", first, samples);
    FILE *fo = std::fopen(argv[4], "wb");
    std::fwrite(out.data(), 4, n, fo);
    std::fclose(fo);
    return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("progressive")
    src = d / "driver.cpp"
    src.write_text(DRIVER)
    exe = d / "driver"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", f"-I{CSRC}", "-o", str(exe), str(src)],
                   check=True)
    return exe


def test_next_batch_size_and_pending_delta_rules(driver):
    lines = subprocess.run([str(driver)], check=True, capture_output=True, text=True).stdout.split("\n")
    grow, shrink, zero_prev, slow_one, capped, zero_elapsed, saturated = map(int, lines[0].split())
    assert grow > 4                      # next_batch_size_grows_when_well_under_budget
    assert 1 <= shrink < 100             # next_batch_size_shrinks_when_over_budget
    assert zero_prev >= 1 and slow_one >= 1  # next_batch_size_never_returns_zero
    assert capped == 16                  # growth is capped at x4 (sizing.rs:28)
    assert zero_elapsed == 28            # zero elapsed: prev x 4 (sizing.rs:23-25)
    assert saturated == 0xFFFFFFFF       # saturating_mul
    assert shrink == 25 and grow == 16   # 100 x 100/400; 4 x 10 capped at 16
    t0, t1, t2, f1, s1, v = lines[1].split()
    assert (t0, t1, t2) == ("0", "1", "0")  # none when empty; empty again right after a take
    assert (int(f1), int(s1), float(v)) == (0, 2, 0.5)


def test_coalesced_deltas_sum_identically_to_un_coalesced_ones(driver, tmp_path):
    # coalesced_deltas_sum_identically_to_un_coalesced_ones (tests.rs:34-55): the tiny scene
    # (4x4, Diamond on the standard round brilliant, yaw 0.4, pitch 0.3, distance 3, the viewer's
    # 42 degree field of view (render_core/mod.rs:65), Daylight,
    # light 0.85 / 0.95, depth 4) traced on the oracle for samples [0, 3), [3, 8) and [0, 8).
    srb = po.standard_round_brilliant()
    from gemray_studio_amd import materials as M

    m = M.to_struct(M.by_name("Diamond"))
    env = po.studio_env("Daylight", 1.0, 0.85, 0.95)
    cam = po.camera_new(0.4, 0.3, 3.0, 42.0)
    a, _, _, _ = po.render(cam, 4, 4, srb, m, 4, env, 0, 3)
    b, _, _, _ = po.render(cam, 4, 4, srb, m, 4, env, 3, 5)
    direct, _, _, _ = po.render(cam, 4, 4, srb, m, 4, env, 0, 8)
    fa, fb, fo = tmp_path / "a.bin", tmp_path / "b.bin", tmp_path / "o.bin"
    a.astype(np.float32).tofile(fa)
    b.astype(np.float32).tofile(fb)
    res = subprocess.run([str(driver), str(a.size), str(fa), str(fb), str(fo)], check=True, capture_output=True,
                         text=True)
    assert res.stdout.split() == ["0", "8"]
    coalesced = np.fromfile(fo, np.float32).reshape(direct.shape)
    assert np.array_equal(coalesced, a + b)
    scale = np.maximum(np.maximum(np.abs(coalesced), np.abs(direct)), 1e-6)
    assert np.max(np.abs(coalesced - direct) / scale) < 1e-3
    assert np.any(direct > 0)
