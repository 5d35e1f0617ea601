import numpy as np


def rel_err(a, b, floor=1e-300):
    """max_i |a_i - b_i| / max(|b_i|, floor)"""
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def scaled_err(a, b):
    """max_i |a_i - b_i| / max_i |b_i|  (error relative to the profile's scale)"""
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    s = float(np.max(np.abs(b))) if b.size else 1.0
    return float(np.max(np.abs(a - b))) / (s if s > 0 else 1.0)
