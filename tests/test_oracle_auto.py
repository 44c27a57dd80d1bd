"""Pins of the oracle's adaptive SR integral (method 1, SRW's
RadIntegrationAuto1, cpp/src/core/srradint.cpp:1731-2046) on the CPU.

The reference's tests hold no numeric values for this path.  The restatement
is pinned by what fixes it independently:
* at its final level L the result is the composite rule with 4*2^L intervals,
  i.e. the method-0 result at that step (level bookkeeping, weights, end
  derivative terms and normalisation);
* the final level is the first whose phase steps stay within pi (phases
  computed here from the trajectory interpolation) and whose |E|^2 changed by
  at most relPrec*|E|^2 for the first point of a call, and by at most relPrec
  times the largest |E|^2 of the earlier points after it (|E|^2 per level
  from the method-0 composite rule);
* as relPrec -> 0 it converges to the trajectory-step quadrature.
"""
from array import array

import numpy as np
import pytest

from srw_amd import workloads as W
from srw_amd.srwlib import SRWLPrtTrj
from tests import oracle_lib as oracle
from tests.util import cfield, clone_wfr, rel_l2

RANGE = 2.16          # integration range [m]: the whole field extent of the undulator (+/-1.08 m)
K_PH = 3.141592653590 * 1.E+06 * 0.80654658   # pi/lambda per eV with SRW's SR-core constants


def _point(xs, y, rel=0.01, termin=1):
    beam = W.electron_beam()
    e1 = W.first_harmonic_eV(beam.partStatMom1.gamma, 1.5, 0.02)
    w = W.wavefront(beam, 1, e1, e1, len(xs), 1, 0, 0, 20.)
    w.mesh.xStart, w.mesh.xFin = xs[0], xs[-1]
    w.mesh.yStart = w.mesh.yFin = y
    return w, W.planar_undulator(), array('d', [1, rel, 0, 0, 10001, termin, 0])


def _manual_at_level(w, mag, L, termin=1):
    """Method-0 field with the step of level L (4*2^L intervals)."""
    wm = clone_wfr(w)
    step = RANGE / (4 * 2 ** L) * (1 - 1e-12)
    assert oracle.calc_field(wm, mag, array('d', [0, step, 0, 0, 10001, termin, 0])) == 0
    return wm


def _sq(w):
    ex, ey = cfield(w.arEx, w).ravel(), cfield(w.arEy, w).ravel()
    return np.abs(ex) ** 2 + np.abs(ey) ** 2


def _guard(w, mag, L, xo, zo):
    """True when the phase advances by at most pi between consecutive points of
    level L, starting from the phase at the range start (srradint.cpp:1991)."""
    trj = SRWLPrtTrj()
    trj.allocate(10001, True)
    trj.partInitCond = w.partBeam.partStatMom1
    trj.ctStart, trj.ctEnd = -RANGE / 2, RANGE / 2
    assert oracle.calc_traj(trj, mag) == 0
    step = RANGE / 2 ** (L + 1)
    s = np.concatenate([[-RANGE / 2], -RANGE / 2 + step * (0.5 + np.arange(2 ** (L + 1)))])
    v = oracle.trj_interp(trj, s)
    g = w.partBeam.partStatMom1.gamma * 0.51099890221e-3 * 1000. / 0.511003414
    X, IBx, Z, IBz = v[:, 1], v[:, 2], v[:, 4], v[:, 5]
    yo = w.mesh.zStart
    ph = K_PH * w.mesh.eStart * (s / g ** 2 + IBx + IBz + ((xo - X) ** 2 + (zo - Z) ** 2) / (yo - s))
    return bool(np.all(np.diff(ph) <= 3.141592653590))


@pytest.mark.parametrize("termin", [1, 0])
@pytest.mark.parametrize("x,y", [(0., 0.), (3e-4, 2e-4), (9e-4, -5e-4)])
def test_final_level_is_composite_rule(x, y, termin):
    w, mag, p = _point((x,), y, termin=termin)
    rc, lv = oracle.calc_field_levels(w, mag, p)
    assert rc == 0 and len(lv) == 1
    wm = _manual_at_level(w, mag, int(lv[0]), termin)
    for a, b in ((w.arEx, wm.arEx), (w.arEy, wm.arEy)):
        a, b = cfield(a, w), cfield(b, wm)
        assert np.abs(a - b).max() <= 1e-8 * max(np.abs(b).max(), 1e-30)


def _first_passing(S, guards, thr_of):
    for L in range(1, len(S)):
        if guards[L] and not abs(S[L] - S[L - 1]) > thr_of(S[L]):
            return L
    return None


@pytest.mark.parametrize("x,y,rel", [(0., 0., 0.01), (3e-4, 2e-4, 1e-3), (9e-4, -5e-4, 1e-4)])
def test_first_point_stops_at_first_relative_pass(x, y, rel):
    w, mag, p = _point((x,), y, rel=rel)
    rc, lv = oracle.calc_field_levels(w, mag, p)
    L = int(lv[0])
    S = [_sq(_manual_at_level(w, mag, k))[0] for k in range(L + 2)]
    guards = [True] + [_guard(w, mag, k, x, y) for k in range(1, L + 2)]
    assert L == _first_passing(S, guards, lambda s: rel * s)


@pytest.mark.parametrize("xs", [(0., 2e-3), (9e-4, 0.)])
def test_later_points_use_the_running_maximum(xs):
    """Second point of a call: the threshold is relPrec times the first point's
    final |E|^2 (looser for a dim point after a bright one, tighter for a
    bright point after a dim one)."""
    rel = 0.01
    w, mag, p = _point(xs, 0., rel=rel)
    rc, lv = oracle.calc_field_levels(w, mag, p)
    assert rc == 0 and len(lv) == 2
    S1 = _sq(_manual_at_level(_point((xs[0],), 0.)[0], mag, int(lv[0])))[0]
    w2 = _point((xs[1],), 0.)[0]
    top = int(lv[1]) + 1
    S = [_sq(_manual_at_level(w2, mag, k))[0] for k in range(top + 1)]
    guards = [True] + [_guard(w2, mag, k, xs[1], 0.) for k in range(1, top + 1)]
    assert int(lv[1]) == _first_passing(S, guards, lambda s: rel * S1)
    alone = oracle.calc_field_levels(_point((xs[1],), 0.)[0], mag, p)[1]
    # the carried maximum changes the outcome in both set-ups
    assert int(lv[1]) != int(alone[0])


def test_converges_to_trajectory_step_quadrature():
    w, mag, p = _point((3e-4,), 2e-4, rel=1e-7)
    assert oracle.calc_field_levels(w, mag, p)[0] == 0
    w0 = clone_wfr(w)
    assert oracle.calc_field(w0, mag, W.traj_prec(10001)) == 0
    assert rel_l2(cfield(w.arEx, w), cfield(w0.arEx, w0)) < 1e-6


def test_bad_relative_precision_is_rejected():
    w, mag, p = _point((0.,), 0., rel=0.)
    assert oracle.calc_field_levels(w, mag, p)[0] > 0


@pytest.mark.parametrize("f", [0.5, 1.0, 2.0])
def test_auto_mesh_samples_the_curvature_phase(f):
    """precPar[6] = f: the new mesh has at least as many intervals as the
    curvature phase pi (x - xc)^2 / (lambda R) needs to advance by at most
    pi/(1.2 f) per step at the mesh edges (the count rounded down, as the
    reference counts them), and the next smaller fast-FFT size has fewer."""
    w, mag, prec = W.single_electron(n=21)
    p = array('d', list(prec))
    p[6] = f
    rc, (nx, ny) = oracle.calc_field_auto_mesh(w, mag, p)
    assert rc == 0 and (w.mesh.nx, w.mesh.ny) == (nx, ny)
    lam = 1.239842e-6 / w.mesh.eStart
    for n, a0, a1, c in ((nx, w.mesh.xStart, w.mesh.xFin, w.xc), (ny, w.mesh.yStart, w.mesh.yFin, w.yc)):
        edge = max(abs(a0 - c), abs(a1 - c))
        d_max = lam * w.Rx / (2 * edge) / (1.2 * f)      # largest step meeting the condition
        need = int((a1 - a0) / d_max)                     # intervals
        assert n - 1 >= need
        smaller = max(m for m in range(8, n) if oracle.lib.orc_next_fft_size(m) == m)
        assert smaller - 1 < need
    # the field on the new mesh is the plain computation on that mesh
    w2 = clone_wfr(w)
    assert oracle.calc_field(w2, mag, prec) == 0
    assert rel_l2(cfield(w.arEx, w), cfield(w2.arEx, w2)) == 0.
