"""Rounding of FP64 values to the storage types of a demoted tile, as the
checker of the mixed-precision tests: FP32 by IEEE round to nearest even
(numpy's conversion), BF16 through FP32, then to nearest even on the upper
16 bits of the float (hicma_tlr_demote's documented rounding)."""
import numpy as np


def f32_round(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)


def bf16_round(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << np.uint32(16)
    return u.view(np.float32).astype(np.float64)
