"""Algorithmic FP32 operation counts per event class of the transport estimator.

Convention: add/sub/mul/min/max = 1, FMA = 2, div/sqrt/rcp = 1, each transcendental
call = 1; integer RNG work and compares = 0; channel-invariant work counted once.  The
counts are taken from the reference lines each class executes (one "segment" is one
bounce-loop iteration, transport.rs:781-929):

  clip       12 per plane: two 3-term dots (5 + 5), the add of d, one divide
             (simd/slab.rs:140-152)
  hit        24: hit point (6), plane-of-incidence normal cross + normalize (18)
  rotate     80: psi (32, raytracer/absorption.rs:44-54) + 8 x 6 Stokes rotation
  absorb    742: eigen-polarizations, field directions, quadratic forms and the exp
             polynomial for 8 channels (raytracer/absorption.rs:115-216)
  geometry   16: cos_i, sin_i, medium indices (refraction.rs:428-547), per hit
  aniso      80: per hit of an anisotropic stone, 8 extraordinary indices
  fresnel    22: hero Fresnel amplitudes and r_unpol (refraction.rs:1049-1056), per
             non-TIR hit
  tir       190: 8 channels of retardation or reflection (refraction.rs:599-637)
  reflect   394: 8 x 47 + 18 (refraction.rs:650-690)
  refract   790: 30 + 8 x 95 (refraction.rs:748-1011)
  entry      60: theta_c fixed point at an anisotropic entry (refraction.rs:248-269)
  walkoff    45: extraordinary Poynting direction (birefringence.rs:74-91)
  rr         43: Russian roulette (transport.rs:204-220)
  miss_map  245: 45 + 8 x 25 (env_map.rs:231-319, env_map_spectrum.rs:60-82)
  miss_studio 280: 120 + 8 x 20 (environment.rs:280-321)
  path      640: per-path setup and finish (comb, 8 dispersion evaluations, CMF x 8, MIS,
             XYZ, white balance), plus 8 x 6 per absorption band and axis
"""
OPS = {
    "clip_per_plane": 12,
    "hit": 24,
    "rotate": 80,
    "absorb": 742,
    "geometry": 16,
    "aniso_per_hit": 80,
    "fresnel": 22,
    "tir": 190,
    "reflect": 394,
    "refract": 790,
    "entry": 60,
    "walkoff": 45,
    "rr": 43,
    "miss_map": 245,
    "miss_studio": 280,
    "path": 640,
    "path_per_band": 48,
}


def algorithmic_flops(events: dict, n_planes: int, paths: int, anisotropic: bool,
                      absorbing: bool, env_studio: bool, n_bands: int = 0) -> float:
    """FP32 operations of a launch from its exact per-class event counters."""
    e = events
    hits = e["hit"]
    flops = e["segment"] * OPS["clip_per_plane"] * n_planes
    flops += hits * (OPS["hit"] + OPS["geometry"])
    flops += e["rotate"] * OPS["rotate"]
    if absorbing:
        flops += e["absorb"] * OPS["absorb"]
    if anisotropic:
        flops += hits * OPS["aniso_per_hit"]
    flops += (e["reflect"] + e["refract"]) * OPS["fresnel"]
    flops += e["tir"] * OPS["tir"] + e["reflect"] * OPS["reflect"] + e["refract"] * OPS["refract"]
    flops += e["entry_aniso"] * OPS["entry"] + e["walkoff"] * OPS["walkoff"]
    flops += e["rr"] * OPS["rr"]
    flops += e["miss"] * (OPS["miss_studio"] if env_studio else OPS["miss_map"])
    flops += paths * (OPS["path"] + OPS["path_per_band"] * n_bands)
    return float(flops)
