"""The interior clip's first-pass screen (clip_scan_screened, gemray_studio_amd/csrc/clip.h),
compiled for the host from the kernel header itself and run against the plain scan
(clip_scan<false>, whose selections test_clip_model.py ties to the oracle's
intersect_polyhedron, simd/slab.rs:94-152).  CPU only (hipcc host pass, no GPU).

The screen rules a plane out from FMA-form dot products and a rounding-error margin, so it
must leave every output of the scan unchanged: t_near, t_far (bits), their facet indices and
the outside flag.  Checked on rays aimed at the stone, rays from points inside it, rays
leaving a hit point in a random direction (as a path continues) and rays leaving it nearly
parallel to the facet it hit, over stones with near-duplicate planes (d moved by 0-2 ulp),
scaled by 1e-3 and 1e3, moved 900 units from the origin, and with normals scaled by 0.01
and 37.  Also checked: padding with copies of plane 0 (what the screened kernels stage)
leaves the plain scan unchanged, and -- so that the check is known to see a wrong margin --
the same screen with its margins set to zero does change outcomes on these rays."""
import os
import struct
import subprocess

import numpy as np

from gemray_studio_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gemray_studio_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
F = np.float32

CHECK_CPP = r"""
#include "clip.h"
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
using namespace gr;
struct Out { float tn, tf; int ni, fi; bool out; };
static bool same(const Out &a, const Out &b) {
    uint32_t x[2], y[2];
    std::memcpy(&x[0], &a.tn, 4); std::memcpy(&x[1], &a.tf, 4);
    std::memcpy(&y[0], &b.tn, 4); std::memcpy(&y[1], &b.tf, 4);
    return x[0] == y[0] && x[1] == y[1] && a.ni == b.ni && a.fi == b.fi && a.out == b.out;
}
static Out init() { return Out{1e-4f, 1e30f, -1, -1, false}; }
static Out plain(const std::vector<float4> &p, uint32_t n, V3 o, V3 d) {
    Out r = init();
    clip_scan<false>(p.data(), nullptr, n, o, d, r.tn, r.ni, r.tf, r.fi, r.out);
    return r;
}
static Out screened(const std::vector<float4> &p, uint32_t n, float w, float b, V3 o, V3 d) {
    Out r = init();
    clip_scan_screened(p.data(), n, w, b, o, d, r.tn, r.ni, r.tf, r.fi, r.out);
    return r;
}
static Out zero_margin(const std::vector<float4> &p, uint32_t n, V3 o, V3 d) {
    Out r = init();
    const ClipMargins m{0.0f, 0.0f};
    float mn = 0.0f, mf = 0.0f;
    for (uint32_t i = 0; i < ((n + 3u) & ~3u); ++i)
        clip_plane_screened(p[i], (int)i, o, d, m, mn, mf, r.tn, r.ni, r.tf, r.fi, r.out);
    return r;
}
int main(int argc, char **argv) {
    FILE *f = std::fopen(argv[1], "rb");
    long rays = 0, bad_screen = 0, bad_copy = 0, zero_diff = 0, continued = 0;
    int32_t hdr[2];
    while (std::fread(hdr, 4, 2, f) == 2) {
        const uint32_t n = (uint32_t)hdr[0], R = (uint32_t)hdr[1], n4 = (n + 3u) & ~3u;
        std::vector<float> pl(4 * n), ro(3 * R), rd(3 * R), rd2(3 * R), eps(R);
        if (std::fread(pl.data(), 4, pl.size(), f) != pl.size() || std::fread(ro.data(), 4, ro.size(), f) != ro.size() ||
            std::fread(rd.data(), 4, rd.size(), f) != rd.size() || std::fread(rd2.data(), 4, rd2.size(), f) != rd2.size() ||
            std::fread(eps.data(), 4, eps.size(), f) != eps.size())
            return 2;
        std::vector<float4> zpad(n4), cpad(n4);
        double w = 1e-20, b = 0.0;  // the bounds the runtime computes for the scene
        for (uint32_t i = 0; i < n4; ++i) {
            const uint32_t j = i < n ? i : 0u;
            const float4 p = make_float4(pl[4 * j], pl[4 * j + 1], pl[4 * j + 2], pl[4 * j + 3]);
            zpad[i] = i < n ? p : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
            cpad[i] = p;
            if (i < n) {
                w = std::fmax(w, std::fabs((double)p.w));
                b = std::fmax(b, std::fabs((double)p.x) + std::fabs((double)p.y) + std::fabs((double)p.z));
            }
        }
        const float wf = std::nextafter((float)w, INFINITY), bf = std::nextafter((float)b, INFINITY);
        auto check = [&](V3 o, V3 d) {
            const Out r = plain(zpad, n, o, d);
            if (!same(r, screened(cpad, n, wf, bf, o, d))) ++bad_screen;
            if (!same(r, plain(cpad, n, o, d))) ++bad_copy;
            if (!same(r, zero_margin(cpad, n, o, d))) ++zero_diff;
            ++rays;
            return r;
        };
        for (uint32_t k = 0; k < R; ++k) {
            const V3 o{ro[3 * k], ro[3 * k + 1], ro[3 * k + 2]}, d{rd[3 * k], rd[3 * k + 1], rd[3 * k + 2]};
            const Out r = check(o, d);
            const bool hit = !r.out && !(r.tn > r.tf) && (r.tn > 1e-4f || r.tf > 1e-4f);
            if (!hit) continue;
            const bool near = r.tn > 1e-4f;
            const float t = near ? r.tn : r.tf;
            const int fi = near ? r.ni : r.fi;
            if (fi < 0) continue;
            const V3 h{o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
            // Leave the hit point in a random direction, 1e-4 along it.
            const V3 d2{rd2[3 * k], rd2[3 * k + 1], rd2[3 * k + 2]};
            check(V3{h.x + d2.x * 1e-4f, h.y + d2.y * 1e-4f, h.z + d2.z * 1e-4f}, d2);
            // And nearly along the facet it hit: d2 with its normal component replaced by eps.
            const float4 p = cpad[fi];
            const float nn = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
            const V3 nu{p.x / nn, p.y / nn, p.z / nn};
            const float c = d2.x * nu.x + d2.y * nu.y + d2.z * nu.z;
            V3 g{d2.x - c * nu.x + eps[k] * nu.x, d2.y - c * nu.y + eps[k] * nu.y, d2.z - c * nu.z + eps[k] * nu.z};
            const float gl = std::sqrt(g.x * g.x + g.y * g.y + g.z * g.z);
            if (!(gl > 0.0f)) continue;
            g = V3{g.x / gl, g.y / gl, g.z / gl};
            check(V3{h.x + g.x * 1e-4f, h.y + g.y * 1e-4f, h.z + g.z * 1e-4f}, g);
            check(h, g);
            ++continued;
        }
    }
    std::printf("%ld %ld %ld %ld %ld\n", rays, continued, bad_screen, bad_copy, zero_diff);
    return 0;
}
"""


def near_duplicates(planes, rng, n=10):
    """Appends copies of n random planes with d moved by 0, 1 or 2 ulp (ties and
    near-ties only the exact division and the first-index rule can order)."""
    extra = planes[rng.integers(0, len(planes), size=n)].copy()
    for k in range(n):
        for _ in range(int(rng.integers(0, 3))):
            extra[k, 3] = np.nextafter(extra[k, 3], F(np.inf) if rng.random() < 0.5 else F(-np.inf))
    out = np.concatenate([planes, extra]).astype(F)
    return np.ascontiguousarray(out[rng.permutation(len(out))])


def unit_rows(v):
    return (v / np.sqrt((v * v).sum(axis=1, keepdims=True))).astype(F)


def transformed(planes, o, s, c, k):
    """The stone y = s x + c (planes (n, s d - n . c)), its normals and d scaled by k."""
    p = planes.astype(np.float64)
    out = np.empty_like(p)
    out[:, :3] = p[:, :3] * k
    out[:, 3] = (s * p[:, 3] - p[:, :3] @ np.asarray(c, np.float64)) * k
    return out.astype(F), (o.astype(np.float64) * s + np.asarray(c, np.float64)).astype(F)


def write_cases(path):
    rng = np.random.default_rng(20)
    stones = [scenes.round_brilliant(), scenes.oval_brilliant(), scenes.step_cut(),
              scenes.random_convex_stone(11, 23), scenes.random_convex_stone(12, 181)]
    variants = [(1.0, (0, 0, 0), 1.0), (1e-3, (0, 0, 0), 1.0), (1e3, (0, 0, 0), 1.0),
                (1.0, (700.0, -300.0, 450.0), 1.0), (0.05, (3.0, 1.0, -2.0), 1.0),
                (1.0, (0, 0, 0), 0.01), (1.0, (0, 0, 0), 37.0)]
    with open(path, "wb") as f:
        for stone in stones:
            planes = near_duplicates(np.ascontiguousarray(stone, F), rng)
            for s, c, k in variants:
                n_out, n_in = 1500, 700
                o = rng.normal(size=(n_out, 3))
                o = o / np.linalg.norm(o, axis=1, keepdims=True) * 2.5
                d = unit_rows(rng.normal(size=(n_out, 3)) * 0.6 - o)
                o = np.concatenate([o, rng.normal(size=(n_in, 3)) * 0.15]).astype(F)
                d = np.concatenate([d, unit_rows(rng.normal(size=(n_in, 3)))]).astype(F)
                pl, ot = transformed(planes, o, s, c, k)
                d2 = unit_rows(rng.normal(size=(len(o), 3)))
                eps = (10.0 ** rng.uniform(-9, -4, size=len(o)) * rng.choice([-1, 1], size=len(o))).astype(F)
                f.write(struct.pack("<ii", len(pl), len(o)))
                for a in (pl, ot, d, d2, eps):
                    f.write(np.ascontiguousarray(a, F).tobytes())


def test_screened_interior_scan_matches_the_plain_scan(tmp_path):
    src = tmp_path / "clip_check.hip"
    src.write_text(CHECK_CPP)
    exe = tmp_path / "clip_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "--offload-host-only", f"-I{CSRC}",
                    "-Wno-unused-command-line-argument", "-o", str(exe), str(src)], check=True, capture_output=True)
    cases = tmp_path / "cases.bin"
    write_cases(cases)
    out = subprocess.run([str(exe), str(cases)], check=True, capture_output=True, text=True).stdout
    rays, continued, bad_screen, bad_copy, zero_diff = map(int, out.split())
    assert rays > 150_000 and continued > 20_000, out
    assert bad_screen == 0, out
    assert bad_copy == 0, out
    # The check sees a wrong margin: without one the screen changes outcomes.
    assert zero_diff > 0, out
