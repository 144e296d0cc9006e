"""The oracle reproduces every golden f32 bit pattern the reference's own tests record.

Fixture: tests/golden/reference_goldens.json, transcribed from
crates/gemray/tests/raytracer_tests.rs:1479-1490, :1673-1757 and :1794-1804.
"""
import json
import os

import numpy as np
import pytest

import pyoracle as po
from gemray_studio_amd import materials as M
from helpers import bits, f32_normalize, unit_draw

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")))
PI32 = float(np.float32(np.pi))


@pytest.fixture(scope="module")
def srb():
    return po.standard_round_brilliant()


@pytest.mark.parametrize("name,expected", GOLD["material_goldens"]["cases"])
def test_trace_spectral_ray_material_goldens(srb, name, expected):
    g = GOLD["material_goldens"]
    d = f32_normalize(g["ray"]["dir_unnormalized"])
    env = po.studio_env(g["env"]["preset"], g["env"]["exposure"], g["env"]["light_yaw"], g["env"]["light_pitch"])
    seed = g["seed"]
    xyz, info = po.trace(g["ray"]["origin"], d, srb, M.to_struct(M.by_name(name)), g["max_bounces"], env,
                         seed, unit_draw(po.lib().or_hash_u32(seed)))
    assert [bits(v) for v in xyz] == expected


@pytest.mark.parametrize("case", GOLD["studio_rig_cases"], ids=lambda c: str(c["seed"]))
@pytest.mark.parametrize("material", ["diamond", "sapphire"])
def test_trace_spectral_ray_studio_rig_goldens(srb, case, material):
    d = f32_normalize(case["dir"]) if case["normalize_dir"] else np.asarray(case["dir"], np.float32)
    yaw = PI32 if case["light_yaw"] == "PI" else case["light_yaw"]
    env = po.studio_env(case["preset"], 1.0, yaw, case["light_pitch"])
    m = M.to_struct(M.by_name(material))
    seed = case["seed"]
    xyz, _ = po.trace(case["origin"], d, srb, m, 12, env, seed, unit_draw(po.lib().or_hash_u32(seed)))
    assert [bits(v) for v in xyz] == case[material]


def test_sample_studio_environment_goldens():
    g = GOLD["studio_env"]
    t, s = po.PRESETS[g["preset"]]
    for dv, do_norm, exp in zip(g["dirs_unnormalized"], g["normalize"], g["expected"]):
        d = f32_normalize(dv) if do_norm else np.asarray(dv, np.float32)
        for lam, e in zip(g["lambdas"], exp):
            v = po.lib().or_sample_studio_environment(po.fptr(po.f32arr(d)), lam, t, s, g["exposure"],
                                                      g["light_yaw"], g["light_pitch"])
            assert bits(v) == e, (dv, lam)


def test_srb_has_74_planes_with_unit_normals(srb):
    assert srb.shape == (74, 4)
    n = np.linalg.norm(srb[:, :3].astype(np.float64), axis=1)
    assert np.all(np.abs(n - 1.0) < 1e-6)
    # geometry/cuts.rs:50 table, :126 culet
    assert srb[0].tolist() == [0.0, 1.0, 0.0, np.float32(-0.32)]
    assert srb[73].tolist() == [0.0, -1.0, 0.0, np.float32(-0.88)]
