"""Shared test helpers (f32 glam-order vector maths for building inputs)."""
import numpy as np


def f32_normalize(v):
    """glam Vec3::normalize in f32: v * (1 / sqrt(dot(v, v))), left-associative dot."""
    v = np.asarray(v, dtype=np.float32)
    d = np.float32(v[0] * v[0]) + np.float32(v[1] * v[1])
    d = np.float32(d + np.float32(v[2] * v[2]))
    r = np.float32(np.float32(1.0) / np.sqrt(d))
    return (v * r).astype(np.float32)


def unit_draw(h):
    return float(np.float32(np.uint32(h)) / np.float32(4294967295.0))


def bits(v):
    return int(np.float32(v).view(np.uint32))


def ulp_distance(a, b):
    """Signed-magnitude f32 ULP distance, elementwise."""
    a = np.asarray(a, dtype=np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, dtype=np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, np.int64(-2**31) - a, a)
    b = np.where(b < 0, np.int64(-2**31) - b, b)
    return np.abs(a - b)
