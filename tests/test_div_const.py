"""The kernels divide by the Gaussian lobes' constant widths with div_by_const (one
multiply by the rounded reciprocal and one fma-corrected step, device_physics.h).  This
checks, exhaustively over every float32 comb wavelength in [380, 780] nm, that the
sequence returns the correctly rounded quotient the reference's division gives, for every
(centre, width below, width above) triple the kernels use -- the triples are read from the
kernel sources, so a new lobe is checked automatically.  CPU only (gcc)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gemray_studio_amd", "csrc")

CHECK_C = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float seq(float x, float c, float rc) {
    float q0 = x * rc;
#if CORRECT
    float r = fmaf(-q0, c, x);
    return fmaf(r, rc, q0);
#else
    (void)c;
    return q0;
#endif
}
static const float lobes[][3] = { LOBES };
int main(void) {
    long bad = 0, n = 0;
    for (unsigned k = 0; k < sizeof(lobes) / sizeof(lobes[0]); ++k) {
        const float mu = lobes[k][0], lo = lobes[k][1], hi = lobes[k][2];
        const float rlo = 1.0f / lo, rhi = 1.0f / hi;
        for (uint32_t u = u_of(380.0f); u <= u_of(780.0f); ++u) {
            const float l = f_of(u);
            const int below = l < mu;
            const float x = l - mu;
            volatile float s = below ? lo : hi;
            const float q = x / s;
            const float q2 = seq(x, below ? lo : hi, below ? rlo : rhi);
            ++n;
            if (u_of(q) != u_of(q2)) ++bad;
        }
    }
    printf("%ld %ld\n", bad, n);
    return 0;
}
"""


def lobe_triples():
    triples = set()
    for name in ("device_physics.h", "transport_kernels.h"):
        text = open(os.path.join(CSRC, name)).read()
        for fn in ("cmf_lobe", "asym_gauss"):
            for m in re.finditer(fn + r"\(l,\s*([0-9.]+)f,\s*([0-9.]+)f,\s*([0-9.]+)f\)", text):
                triples.add(tuple(m.groups()))
    return sorted(triples)


def run_check(tmp_path, correct):
    triples = lobe_triples()
    src = CHECK_C.replace("LOBES", ", ".join("{%sf, %sf, %sf}" % t for t in triples))
    c = tmp_path / "divc.c"
    c.write_text(src)
    exe = tmp_path / ("divc_ok" if correct else "divc_bad")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", f"-DCORRECT={int(correct)}", "-o", str(exe), str(c), "-lm"],
                   check=True)
    bad, n = map(int, subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split())
    return bad, n, len(triples)


def test_lobe_divisions_are_correctly_rounded_for_every_comb_wavelength(tmp_path):
    bad, n, k = run_check(tmp_path, True)
    assert k == 10  # 7 CIE lobes + 3 RGB-to-spectrum lobes
    assert n > 8e7
    assert bad == 0


SPARK_C = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
int main(void) {
    long bad = 0, n = 0;
    const float c = 0.04f, rc = 1.0f / 0.04f;
    for (uint32_t u = u_of(0.96f) + 1; u <= u_of(1.0f); ++u) {
        const float x = f_of(u) - 0.96f;
        volatile float cv = c;
        const float q = x / cv;
        const float q0 = x * rc;
        const float q2 = fmaf(fmaf(-q0, c, x), rc, q0);
        ++n;
        if (u_of(q) != u_of(q2)) ++bad;
    }
    printf("%ld %ld\n", bad, n);
    return 0;
}
"""


def test_ring_light_spark_division_is_correctly_rounded_where_used(tmp_path):
    # finish_escaped: spark = (rd - 0.96) / 0.04 is used only for rd in (0.96, 1].
    text = open(os.path.join(CSRC, "transport_kernels.h")).read()
    assert "div_by_const(rd - 0.96f, 0.04f, 1.0f / 0.04f)" in text
    c = tmp_path / "spark.c"
    c.write_text(SPARK_C)
    exe = tmp_path / "spark"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(c), "-lm"], check=True)
    bad, n = map(int, subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split())
    assert n > 600000 and bad == 0


def test_the_check_sees_an_uncorrected_quotient(tmp_path):
    # Without the fma step the product alone is off by an ulp for many wavelengths.
    bad, _, _ = run_check(tmp_path, False)
    assert bad > 1000
