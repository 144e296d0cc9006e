"""The hero-offset grid the hero tables are indexed by (no GPU).

init_path forms hero_off = fma(hero, 400, 380) - 380 for a hero draw in [0, 1)
(transport.rs:651-722: the hero wavelength, then the comb's offset from 380 nm).  The tables
hold one entry per possible offset: multiples of 2^-15 in [0, 132) and of 2^-14 in
[132, 400], kHeroEntries = 132 * 2^15 + 268 * 2^14 + 1 in all.  These tests check that
grid against what the float arithmetic can produce: every float wavelength in [380, 780]
minus 380 is exact and on the grid, every such offset maps to its own entry and back, the
entries are exactly the offsets (a bijection), and the offsets of hero draws -- random ones,
the smallest positive floats and the largest below 1 -- are all keys."""
import numpy as np

K_LOW = 132 << 15
K_ENTRIES = K_LOW + (268 << 14) + 1


def hero_key(e):
    e = np.asarray(e, dtype=np.int64)
    return np.where(e < K_LOW, e.astype(np.float64) * 2.0 ** -15,
                    132.0 + (e - K_LOW).astype(np.float64) * 2.0 ** -14).astype(np.float32)


def hero_entry(h):
    h = np.asarray(h, dtype=np.float32)
    hd = np.nan_to_num(h.astype(np.float64), nan=-1.0)
    e = np.where(hd < 132.0, np.floor(hd * 32768.0), K_LOW + np.floor((hd - 132.0) * 16384.0)).astype(np.int64)
    ok = (hd >= 0.0) & (hd <= 400.0) & (e < K_ENTRIES) & (hero_key(np.clip(e, 0, K_ENTRIES - 1)) == h)
    return np.where(ok, e, K_ENTRIES)


def offsets_of_hero(hero):
    """fma(hero, 400, 380) rounded once to f32 (the product of a float and 400 and its sum
    with 380 are exact in float64), minus 380 in f32."""
    lam = (np.asarray(hero, dtype=np.float32).astype(np.float64) * 400.0 + 380.0).astype(np.float32)
    return (lam - np.float32(380.0)).astype(np.float32), lam


def test_every_float_wavelength_offset_is_a_key_and_the_map_is_a_bijection():
    bits = np.arange(0x43BE0000, 0x44430000 + 1, dtype=np.uint32)
    lam = bits.view(np.float32)
    assert lam[0] == 380.0 and lam[-1] == 780.0
    off = (lam - np.float32(380.0)).astype(np.float32)
    # The subtraction is exact (Sterbenz below 760, and the grid step above).
    assert np.array_equal(off.astype(np.float64), lam.astype(np.float64) - 380.0)
    e = hero_entry(off)
    assert len(e) == K_ENTRIES
    assert np.array_equal(e, np.arange(K_ENTRIES))
    assert np.array_equal(hero_key(e).view(np.uint32), off.view(np.uint32))


def test_hero_draw_offsets_are_keys():
    rng = np.random.default_rng(7)
    tiny = np.arange(1, 4097, dtype=np.uint32).view(np.float32)              # smallest positive floats
    near_one = (np.uint32(0x3F800000) - np.arange(1, 4097, dtype=np.uint32)).view(np.float32)
    draws = np.concatenate([rng.random(2_000_000, dtype=np.float32), tiny, near_one,
                            np.array([0.0, 0.5, 0.33, 0.33000001], np.float32)])
    off, lam = offsets_of_hero(draws)
    assert lam.min() >= 380.0 and lam.max() <= 780.0
    e = hero_entry(off)
    assert (e < K_ENTRIES).all()
    assert np.array_equal(hero_key(e).view(np.uint32), off.view(np.uint32))
    # The largest draws round up to 780: offset 400, the last entry.
    assert hero_entry(np.float32(400.0)) == K_ENTRIES - 1


def test_values_off_the_grid_are_not_keys():
    for h in (np.float32(1e-7), np.float32(131.99999), np.float32(132.00002), np.float32(-0.0) - np.float32(1e-3),
              np.float32(400.0001), np.float32(np.nan)):
        e = hero_entry(h)
        assert e == K_ENTRIES or hero_key(e) == h
    assert hero_entry(np.float32(np.nan)) == K_ENTRIES
    assert hero_entry(np.float32(401.0)) == K_ENTRIES
