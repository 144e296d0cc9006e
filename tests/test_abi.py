"""The C-ABI library loads, exports every declared symbol, and its host-side pieces
match the oracle.  No kernel is launched here (these run without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

import pyoracle as po
from gemray_studio_amd import _lib
from gemray_studio_amd import backend as B
from gemray_studio_amd import materials as M
from scene_util import camera_bits_equal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gemray_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gr_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == syms


def test_struct_sizes_match_header_layout():
    # gr_material: 1 + 6 + 1 + 1 + 3 + 2 + 24 + 24 + 1 + 3 + 2 + 24 + 1 four-byte fields.
    assert ctypes.sizeof(_lib.MaterialStruct) == 4 * 93
    assert ctypes.sizeof(_lib.Camera) == 4 * 13
    assert ctypes.sizeof(_lib.Env) == 4 * 8 + 8 + 8
    assert ctypes.sizeof(_lib.Stats) == 8 * 16 + 8 + 8 + 8 + 8 + 8 + 8 + 8


# ctypes mirror of each C struct of the header, checked field by field against the C
# compiler's own layout of include/gemray_hip.h.
MIRRORS = {
    "gr_options": _lib.Options, "gr_camera": _lib.Camera, "gr_material": _lib.MaterialStruct,
    "gr_env": _lib.Env, "gr_scene": _lib.Scene, "gr_stats": _lib.Stats,
    "gr_job_desc": _lib.JobDesc, "gr_job_stats": _lib.JobStats, "gr_serve_metrics": _lib.ServeMetrics,
    "gr_metrics_pose": _lib.MetricsPose, "gr_gem_metrics": _lib.GemMetrics, "gr_tune_choice": _lib.TuneChoice,
}


def test_ctypes_mirrors_match_the_c_layout(tmp_path):
    import subprocess
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {']
    for cname, py in MIRRORS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in MIRRORS.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


@pytest.mark.parametrize("args", [(0.35, 0.28, 5.0, 18.0), (0.4, 0.3, 3.0, 30.0),
                                  (2.1, -0.7, 7.5, 12.0), (0.0, 1.5707963, 4.0, 25.0)])
def test_camera_new_matches_oracle_bit_for_bit(args):
    assert camera_bits_equal(B.camera_new(*args), po.camera_new(*args))


def test_camera_new_rejects_invalid():
    cam = _lib.Camera()
    L = _lib.load()
    assert L.gr_camera_new(0.0, 0.0, -1.0, 18.0, ctypes.byref(cam)) == _lib.GR_ERR_INVALID
    assert L.gr_camera_new(float("nan"), 0.0, 1.0, 18.0, ctypes.byref(cam)) == _lib.GR_ERR_INVALID


def test_disabled_backend_declines_and_leaves_film_untouched():
    b = B.GpuBackend.disabled()
    assert b.adapter_label() is None
    accum = np.full((4, 3), 7.0, dtype=np.float32)
    scene = B.SceneRef(camera=B.camera_new(0.3, 0.2, 5.0, 18.0), width=2, height=2,
                       planes=np.array([[0, 1, 0, -1]], np.float32), material=M.by_name("Diamond"),
                       max_bounces=4, environment=B.Studio())
    assert b.try_accumulate(scene, 0, 4, accum) is False
    assert np.all(accum == 7.0)


def test_open_without_device_declines():
    import torch  # device enumeration only; no kernel runs
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present; the no-device decline is checked on CPU hosts")
    h = ctypes.c_void_p()
    st = _lib.load().gr_open(ctypes.byref(h), None)
    assert st == _lib.GR_DECLINE_NO_DEVICE
    assert not h.value
    assert B.GpuBackend.acquire().available is False
