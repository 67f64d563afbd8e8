"""BF16 helpers for the tests: round-to-nearest-even of float32 to the raw
16-bit pattern, and back to float64 (exact)."""

import numpy as np


def bf16_bits(a):
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_value(bits):
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_round(a):
    return bf16_value(bf16_bits(a))
