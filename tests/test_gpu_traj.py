"""GPU trajectory (srwlCalcPartTraj) against the FP64 oracle's serial RK4.

The device integrates each chain of RK4 steps with a parallel fixed-point
iteration; its converged result equals the serial recurrence up to the
association order of the sums.  Two orderings of the same recurrence differ
by the rounding of z (about sqrt(n) ulps, ~1e-14 m), which shifts the field
phase k_u*dz and so moves x' and x by ~1e-11 of their magnitude.  The
tolerance, 1e-9 of each component's magnitude, is far below the RK4
truncation error at these steps (~1e-7), so a different step, order or
branch would fail it."""
import numpy as np
import pytest

from srw_amd import workloads as W
from srw_amd.srwlib import SRWLPrtTrj
from tests import oracle_lib as oracle

pytestmark = pytest.mark.gpu

REL = 1e-9


def _traj(np_, ct0, ct1, x=0., xp=0., y=0., yp=0., z=0.):
    beam = W.electron_beam(3.0, 0.5, x, xp, y, yp, z)
    t = SRWLPrtTrj()
    t.allocate(np_)
    t.partInitCond = beam.partStatMom1
    t.ctStart, t.ctEnd = ct0, ct1
    return t


def _arrays(t):
    return np.array([t.arX, t.arXp, t.arY, t.arYp, t.arZ])


CASES = [
    ("both sides, on axis", 10001, -1.08, 1.08, {}),
    ("both sides, offsets", 4001, -1.08, 1.08, dict(x=2e-6, xp=3e-6, y=-1e-6, yp=1e-6)),
    ("uneven split", 3001, -0.4, 1.08, dict(xp=1e-6)),
    ("all below zero", 2001, -1.08, -0.3, dict(y=2e-6)),
    ("all above zero", 2001, 0.2, 1.08, {}),
    ("start at zero", 1001, 0.0, 1.08, dict(yp=2e-6)),
    ("short chain", 7, -0.01, 0.01, {}),
]


@pytest.mark.parametrize("name,np_,ct0,ct1,kw", CASES)
def test_parallel_rk4_equals_serial(name, np_, ct0, ct1, kw):
    from srw_amd import srwlpy
    print("case:", name, flush=True)
    mag = W.planar_undulator()
    tg, ts, to = _traj(np_, ct0, ct1, **kw), _traj(np_, ct0, ct1, **kw), _traj(np_, ct0, ct1, **kw)
    srwlpy.set_traj_serial(False)
    srwlpy.CalcPartTraj(tg, mag)
    srwlpy.set_traj_serial(True)
    try:
        srwlpy.CalcPartTraj(ts, mag)
    finally:
        srwlpy.set_traj_serial(False)
    assert oracle.calc_traj(to, mag) == 0
    g, s, o = _arrays(tg), _arrays(ts), _arrays(to)
    scale = np.maximum(np.max(np.abs(o), axis=1), 1e-30)[:, None]
    d_go = np.max(np.abs(g - o) / scale, axis=1)
    d_so = np.max(np.abs(s - o) / scale, axis=1)
    e_go, e_so = float(np.max(d_go)), float(np.max(d_so))
    print("rel dev per component (x, x', y, y', z): parallel", np.array2string(d_go, precision=2),
          "serial", np.array2string(d_so, precision=2), flush=True)
    assert e_so < REL
    assert e_go < REL
