"""Seeded synthetic data, bit-identical to the device generator for 'uniform'.

h(seed, i) = mix(mix(seed + G) + (i + 1) * G) with splitmix64's finaliser ``mix``
and G = 0x9E3779B97F4A7C15 (all arithmetic mod 2^64).  uniform f32 = 2*(h>>40)*2^-24 - 1,
uniform f64 = 2*(h>>11)*2^-53 - 1, normal = Box-Muller in f64 on h(2i), h(2i+1).

Seeds follow ``0xC0A10000 + 16*config_id + operand_id`` (A=0, B=1, C=2, x=3, y=4).
"""
from __future__ import annotations

import numpy as np

G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

OPERAND = {"A": 0, "B": 1, "C": 2, "x": 3, "y": 4}


def seed_for(config_id: int, operand: str) -> int:
    return 0xC0A10000 + 16 * config_id + OPERAND[operand]


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def raw(seed: int, count: int, start: int = 0) -> np.ndarray:
    with np.errstate(over="ignore"):
        s = _mix(np.array([seed], dtype=np.uint64) + G)[0]
        i = np.arange(start, start + count, dtype=np.uint64)
        return _mix(s + (i + np.uint64(1)) * G)


def uniform(seed: int, count: int, dtype=np.float64) -> np.ndarray:
    h = raw(seed, count)
    if np.dtype(dtype) == np.float32:
        u = (h >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)
        return (np.float32(2.0) * u - np.float32(1.0)).astype(np.float32)
    u = (h >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return 2.0 * u - 1.0


def normal(seed: int, count: int, dtype=np.float64) -> np.ndarray:
    h = raw(seed, 2 * count)
    u1 = ((h[0::2] >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (h[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return z.astype(dtype)


def complex_uniform(seed: int, count: int, dtype=np.complex128) -> np.ndarray:
    real = np.float32 if np.dtype(dtype) == np.complex64 else np.float64
    v = uniform(seed, 2 * count, real)
    return v.view(dtype)
