"""Helpers for the GPU parity tests."""
import sys

import numpy as np

import oracle_py as O
from lbm_py import lbm

MACROS = ["phi", "phiPrev", "rho", "ux", "uy", "uz"]


def announce(*what):
    print("[gpu]", *what, flush=True)
    sys.stdout.flush()


def gpu_sphere(nx, ny, nz, cx, cy, cz, r, d):
    return lbm.make_sphere(nx, ny, nz, cx, cy, cz, r, d)


def oracle_sphere(nx, ny, nz, cx, cy, cz, r, d):
    return O.sphere(nx, ny, nz, cx, cy, cz, r, d)


def interior_fluid(tag, nx, ny, nz):
    t = tag.reshape(nz, ny, nx)
    m = np.zeros_like(t, dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return (m & (t == 0)).ravel()


def rel_err(a, b, mask=None):
    """max|a-b| / max|b| over the mask (the field-relative error)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if mask is not None:
        a, b = a[mask], b[mask]
    den = np.max(np.abs(b)) if b.size else 0.0
    num = np.max(np.abs(a - b)) if b.size else 0.0
    return num / den if den > 0 else num


def same_values(a, b):
    """Equal values (+0 == -0), NaNs equal."""
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def pop_fluid(arr, q, mask):
    return arr.reshape(-1, q)[mask]
