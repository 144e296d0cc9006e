"""Algorithmic FP32 operation counts per event class of the transport estimator.

Convention: add/sub/mul/min/max = 1, FMA = 2, div/sqrt/rcp = 1, each transcendental
call = 1; integer RNG work and compares = 0; channel-invariant work counted once.  The
counts are taken from the reference lines each class executes (one "segment" is one
bounce-loop iteration, transport.rs:781-929):

  clip       12 per plane: two 3-term dots (5 + 5), the add of d, one divide
             (simd/slab.rs:140-152).  A camera ray's origin is shared by every ray of the
             frame, so its per-plane sides (6 of the 12) are counted once, not per ray: 6
             per plane at the first segment
  cert       12: a camera ray the pixel's miss certificate settles -- two planes' exact t
             (the full clip is not run)
  exterior   12: a ray leaving the stone outwards, settled by the facet it left (one plane)
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
  scatter   104: free-path draw, 8 channels of extinction and the Henyey-Greenstein
             direction (scattering.rs:132-336)
  miss_map  245: 45 + 8 x 25 (env_map.rs:231-319, env_map_spectrum.rs:60-82)
  miss_studio 280: 120 + 8 x 20 (environment.rs:280-321)
  path_setup  60: draws, comb and camera ray (sampling.rs:232-271, camera.rs:96-114)
  dispersion 120: 8 Sellmeier evaluations (dispersion.rs:12-37), once per path that meets
             the stone (a path whose first segment misses never evaluates them)
  finish    460: CMF x 8, spectral MIS, XYZ and white balance (color.rs:121-330), once per
             escaping path (a path that never escapes carries zero radiance)
  band_fill  48 per absorption band (8 channels x 6), once per path whose cache is filled

With the hero tables (gr_set_hero_tables) a path looks up, instead of evaluating, the CMF
values (344 of the finish's 460), the map's spectral lobes (8 x 20 of miss_map's 8 x 25; the
weighted sum, 5 per channel, still runs), the dispersion indices and the band sums; a lookup
is charged nothing, and the tables' one-off fills are not part of a launch.
"""
OPS = {
    "clip_per_plane": 12,
    "clip_per_plane_camera": 6,
    "cert": 12,
    "exterior": 12,
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
    "scatter": 104,
    "miss_map": 245,
    "miss_studio": 280,
    "path_setup": 60,
    "dispersion": 120,
    "finish": 460,
    "band_fill": 48,
    "finish_cmf": 344,
    "miss_map_lobes": 160,
}


def algorithmic_flops(events: dict, n_planes: int, paths: int, anisotropic: bool,
                      absorbing: bool, env_studio: bool, n_bands: int = 0, spectral_table: bool = False,
                      material_table: bool = False) -> float:
    """FP32 operations of a launch from its exact per-class event counters.  Every path has
    exactly one first segment (its camera ray); the certificate and exterior shortcuts, the
    lazily filled caches and the hero-table lookups are charged at the work that runs."""
    e = events
    hits = e["hit"]
    later = e["segment"] - paths
    scanned_camera = paths - e["cert_miss"]
    flops = scanned_camera * OPS["clip_per_plane_camera"] * n_planes + e["cert_miss"] * OPS["cert"]
    flops += (later - e["exterior_miss"]) * OPS["clip_per_plane"] * n_planes + e["exterior_miss"] * OPS["exterior"]
    flops += hits * (OPS["hit"] + OPS["geometry"])
    flops += e["rotate"] * OPS["rotate"]
    if absorbing:
        flops += e["absorb"] * OPS["absorb"]
    if anisotropic:
        flops += hits * OPS["aniso_per_hit"]
    flops += (e["reflect"] + e["refract"]) * OPS["fresnel"]
    flops += e["tir"] * OPS["tir"] + e["reflect"] * OPS["reflect"] + e["refract"] * OPS["refract"]
    flops += e["entry_aniso"] * OPS["entry"] + e["walkoff"] * OPS["walkoff"]
    flops += e["rr"] * OPS["rr"] + e["scatter"] * OPS["scatter"]
    env_ops = OPS["miss_studio"] if env_studio else OPS["miss_map"]
    finish = OPS["finish"]
    if spectral_table:
        finish -= OPS["finish_cmf"]
        if not env_studio:
            env_ops -= OPS["miss_map_lobes"]
    flops += e["miss"] * (env_ops + finish)
    flops += paths * OPS["path_setup"]
    if not material_table:
        flops += (paths - e["first_miss"]) * OPS["dispersion"]
        flops += e["absorption_fill"] * OPS["band_fill"] * n_bands
    return float(flops)


def algorithmic_flops_uniform(events: dict, n_planes: int, paths: int, anisotropic: bool,
                              absorbing: bool, env_studio: bool, n_bands: int = 0) -> float:
    """The same events charged as if every segment ran the full 12-per-plane clip and every
    path evaluated the per-path setup, dispersion, finish and band sums (the shortcuts and
    tables ignored)."""
    e = dict(events)
    e["cert_miss"] = e["exterior_miss"] = e["first_miss"] = 0
    e["absorption_fill"] = paths if absorbing else 0
    flops = algorithmic_flops(e, n_planes, paths, anisotropic, absorbing, env_studio, n_bands)
    flops += paths * (OPS["clip_per_plane"] - OPS["clip_per_plane_camera"]) * n_planes
    flops += (paths - e["miss"]) * OPS["finish"]
    return float(flops)
