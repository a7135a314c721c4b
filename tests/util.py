"""Comparison helpers shared by the tests."""
import numpy as np

STATE = ("rho", "energy", "vx", "vy", "vz", "bx", "by", "bz")


def rel_l2(a, b):
    nb = float(np.linalg.norm(b))
    d = float(np.linalg.norm(a - b))
    if nb == 0.0:
        return d
    return d / nb


def field_errors(fa, fb, names=STATE, interior=True):
    out = {}
    for n in names:
        a = fa.interior(n) if interior else getattr(fa, n)
        b = fb.interior(n) if interior else getattr(fb, n)
        out[n] = rel_l2(a, b)
    return out
