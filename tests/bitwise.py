"""Bitwise equality of result arrays.

np.array_equal treats -0.0 and +0.0 as equal; the parity tests ask for the
same bits, so they also compare the sign of every floating-point value
(NaN payloads aside, equal values with equal signs are equal bit patterns)."""
import numpy as np


def same_bits(a, b) -> bool:
    a = np.asarray(a)
    b = np.asarray(b)
    if not np.array_equal(a, b):
        return False
    if np.issubdtype(a.dtype, np.floating) or np.issubdtype(b.dtype, np.floating):
        return bool(np.array_equal(np.signbit(a), np.signbit(b)))
    return True
