"""The benchmark and test workloads, as seeded scene generators shared by bench.py and the
GPU tests (so a test runs exactly the jobs the benchmark times):

  single_job_scene  -- the single-job configuration: a pleochroic uniaxial sapphire on the
                       69-plane oval brilliant in the 2048x1024 procedural HDR map,
                       1920x1080, depth 32, framed with the reference's test camera
                       (renderer/gpu/estimator_check/mod.rs:132-138: yaw 0.35, pitch 0.28,
                       distance 5, 18 degrees, chosen so that most rays meet the stone);
  batched_jobs      -- 64 seeded convex stones for the batched launch;
  served_jobs       -- the served request mix;
  cpu_config_scene  -- the reference CPU tracer's configuration (Diamond, parametric
                       brilliant, uniform furnace environment, 256x256, depth 16);
  stress_scene      -- the stress configuration (233-plane step cut, Abbe 20, n_e - n_o
                       0.17, 4096x4096, depth 64);
  preview_jobs      -- interactive previews: many small frames at few samples sharing one
                       device-resident map (the served mix at the sizes a viewer requests).
"""
import numpy as np

SINGLE_JOB_CAMERA = (0.35, 0.28, 5.0, 18.0)


def single_job_scene(width=1920, height=1080):
    from gemray_studio_amd import backend as B
    from gemray_studio_amd import scenes

    material = scenes.pleochroic_sapphire(seed=2)
    planes = scenes.oval_brilliant()
    env = scenes.procedural_env(seed=2, width=2048, height=1024)
    cam_args = SINGLE_JOB_CAMERA
    scene = B.SceneRef(camera=B.camera_new(*cam_args), width=width, height=height, planes=planes,
                       material=material, max_bounces=32, environment=B.HdrMap(env))
    return scene, cam_args


ISOTROPIC = ["Diamond", "Cubic Zirconia", "Spinel"]
UNIAXIAL = ["Sapphire", "Ruby", "Emerald", "Zircon", "Quartz", "Tourmaline", "Synthetic Moissanite"]
PRESETS = ["Daylight", "Incandescent", "RingLights", "DarkSpotlight"]


def batched_jobs(seed=3, n_jobs=64, poses=None, cams=None):
    """Batched workload: n_jobs seeded convex stones with 17-181 planes (cut angles jittered
    +-3 deg), an isotropic/uniaxial material mix (uniaxial ones on a seeded optic axis),
    square frames of log-uniform 128^2..1024^2 pixels, log-uniform 64..1024 spp, studio
    lighting with a random preset, depth 16.  poses / cams, when given, receive each job's
    (camera yaw, pitch, light yaw, light pitch) and camera_new arguments."""
    from dataclasses import replace

    from gemray_studio_amd import backend as B
    from gemray_studio_amd import materials as M
    from gemray_studio_amd import scenes

    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n_jobs):
        n_planes = int(rng.integers(17, 182))
        planes = scenes.random_convex_stone(seed * 1000 + i, n_planes, 3.0)
        if rng.random() < 0.5:
            mat = M.by_name(ISOTROPIC[int(rng.integers(len(ISOTROPIC)))])
        else:
            mat = M.by_name(UNIAXIAL[int(rng.integers(len(UNIAXIAL)))])
            mat = replace(mat, c_axis=scenes.seeded_unit_vector(seed * 1000 + i))
        side = int(round(float(np.exp(rng.uniform(np.log(128), np.log(1024))))))
        spp = int(round(float(np.exp(rng.uniform(np.log(64), np.log(1024))))))
        cam_yaw, cam_pitch = float(rng.uniform(0, 2 * np.pi)), float(rng.uniform(0.2, 0.6))
        cam_args = (cam_yaw, cam_pitch, 5.0, float(rng.uniform(16.0, 28.0)))
        cam = B.camera_new(*cam_args)
        env = B.Studio(PRESETS[int(rng.integers(len(PRESETS)))], 1.0, float(rng.uniform(0, 2 * np.pi)),
                       float(rng.uniform(0.2, 1.0)))
        jobs.append((B.SceneRef(cam, side, side, planes, mat, 16, env), 0, spp))
        if poses is not None:
            poses.append((cam_yaw, cam_pitch, env.light_yaw, env.light_pitch))
        if cams is not None:
            cams.append(cam_args)
    return jobs


def served_jobs(env_map, seed=4, n_jobs=40, cams=None):
    """Served workload: the batched workload's per-job distribution with the single-job
    workload's material and environment family (a seeded pleochroic sapphire, the 2048x1024
    procedural HDR map, depth 32), plus 10 % of jobs at 1920x1080 x 512 spp on the 69-plane
    oval brilliant.  cams, when given, receives each job's camera_new arguments."""
    from gemray_studio_amd import backend as B
    from gemray_studio_amd import scenes

    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n_jobs):
        mat = scenes.pleochroic_sapphire(seed * 1000 + i)
        cam_args = (float(rng.uniform(0, 2 * np.pi)), float(rng.uniform(0.2, 0.6)), 5.0,
                    float(rng.uniform(16.0, 28.0)))
        cam = B.camera_new(*cam_args)
        if cams is not None:
            cams.append(cam_args)
        if rng.random() < 0.1:
            jobs.append((B.SceneRef(cam, 1920, 1080, scenes.oval_brilliant(), mat, 32, B.HdrMap(env_map)), 0, 512))
            continue
        planes = scenes.random_convex_stone(seed * 1000 + i, int(rng.integers(17, 182)), 3.0)
        side = int(round(float(np.exp(rng.uniform(np.log(128), np.log(1024))))))
        spp = int(round(float(np.exp(rng.uniform(np.log(64), np.log(1024))))))
        jobs.append((B.SceneRef(cam, side, side, planes, mat, 32, B.HdrMap(env_map)), 0, spp))
    return jobs


def cpu_config_scene(width=256, height=256):
    """The CPU tracer's configuration: Diamond (Sellmeier3, n_d 2.417, Abbe 55) on the
    parametric round brilliant (table 57 %, crown 34.5 deg, pavilion 40.75 deg, girdle 3 %),
    no absorption, the uniform white furnace environment, depth 16."""
    from gemray_studio_amd import backend as B
    from gemray_studio_amd import materials as M
    from gemray_studio_amd import scenes

    cam_args = SINGLE_JOB_CAMERA
    scene = B.SceneRef(camera=B.camera_new(*cam_args), width=width, height=height,
                       planes=scenes.round_brilliant(), material=M.by_name("Diamond"), max_bounces=16,
                       environment=B.HdrMap.uniform(1.0))
    return scene, cam_args


def stress_scene(width=4096, height=4096):
    """The stress configuration: the Abbe-20, n_e - n_o = 0.17 stone on the 233-plane step cut
    in the 2048x1024 procedural HDR map, depth 64."""
    from gemray_studio_amd import backend as B
    from gemray_studio_amd import scenes

    cam_args = (0.35, 0.28, 4.0, 24.0)
    scene = B.SceneRef(camera=B.camera_new(*cam_args), width=width, height=height, planes=scenes.step_cut(),
                       material=scenes.stress_material(), max_bounces=64,
                       environment=B.HdrMap(scenes.procedural_env(seed=2, width=2048, height=1024)))
    return scene, cam_args


def served_arrivals(n_jobs, rate_per_s=20.0, seed=5):
    """Arrival times (s, the first at 0) of the served mix: a Poisson process at rate_per_s."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_per_s, size=n_jobs))
    return arrivals - arrivals[0]


def preview_jobs(env, seed=6, n_jobs=400, cams=None):
    """Interactive previews: n_jobs square frames of log-uniform 96^2..256^2 pixels at
    log-uniform 4..32 spp, seeded convex stones of 17-181 planes (+-3 deg) in the single-job
    material family (a seeded pleochroic sapphire), lit by one shared environment (env: an
    HdrMap, normally device-resident), depth 32 -- the requests a viewer sends while the user
    turns the stone."""
    from gemray_studio_amd import backend as B
    from gemray_studio_amd import scenes

    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n_jobs):
        mat = scenes.pleochroic_sapphire(seed * 1000 + i)
        cam_args = (float(rng.uniform(0, 2 * np.pi)), float(rng.uniform(0.2, 0.6)), 5.0,
                    float(rng.uniform(16.0, 28.0)))
        planes = scenes.random_convex_stone(seed * 1000 + i, int(rng.integers(17, 182)), 3.0)
        side = int(round(float(np.exp(rng.uniform(np.log(96), np.log(256))))))
        spp = int(round(float(np.exp(rng.uniform(np.log(4), np.log(32))))))
        jobs.append((B.SceneRef(B.camera_new(*cam_args), side, side, planes, mat, 32, env), 0, spp))
        if cams is not None:
            cams.append(cam_args)
    return jobs
