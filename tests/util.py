"""Shared helpers for the tests."""
import copy

import numpy as np


def cfield(arr, w):
    """Complex view [ny][nx][ne] of a float (re,im) wavefront array."""
    a = np.frombuffer(arr, dtype=np.float32).astype(np.float64)
    return (a[0::2] + 1j * a[1::2]).reshape(w.mesh.ny, w.mesh.nx, w.mesh.ne)


def rel_l2(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def clone_wfr(w):
    return copy.deepcopy(w)
