"""The sampling streams' salts as the kernels and the oracle spell them, against the values
the reference defines (crates/gemray/src/optics/raytracer/sampling.rs:12-41 and the three
pixel-rotation streams, sampling.rs:111-113), read from the sources as text.  The per-path
parity tests only see a salt through the paths it moves; a mistyped salt of a rare stream
(frosted or scattering draws) would otherwise go unnoticed on most stones.  CPU only."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

REFERENCE_SALTS = {
    "FRESNEL_BRANCH_STREAM": 0x9E3779B1,
    "RUSSIAN_ROULETTE_STREAM": 0x517CC1B7,
    "BIREFRINGENT_SPLIT_STREAM": 0x2545F491,
    "MODE_COUPLING_STREAM": 0xCC9E2D51,
    "FROSTED_DIR_U_STREAM": 0x27D4EB2F,
    "FROSTED_DIR_V_STREAM": 0x165667B1,
    "DISTANCE_SAMPLE_STREAM": 0xA24BAED4,
    "PHASE_DIR_U_STREAM": 0x9FB21C65,
    "PHASE_DIR_V_STREAM": 0x1CE4E5B9,
    "PIXEL_JITTER_X_ROTATION_STREAM": 0xA511E9B3,
    "PIXEL_JITTER_Y_ROTATION_STREAM": 0x63D81B23,
    "HERO_WAVELENGTH_ROTATION_STREAM": 0x1B873593,
}


def salts_in(path):
    text = open(os.path.join(ROOT, path)).read()
    return {m.group(1): int(m.group(2), 16)
            for m in re.finditer(r"constexpr uint32_t (\w+_STREAM) = 0x([0-9A-Fa-f]+)u;", text)}


def test_device_salts_match_the_reference():
    assert salts_in("gemray_studio_amd/csrc/device_physics.h") == REFERENCE_SALTS


def test_oracle_salts_match_the_reference():
    assert salts_in("oracle/gemray_oracle.cpp") == REFERENCE_SALTS
