"""GPU parity of the adaptive SR integral (precPar[0] = 1, SRW's
RadIntegrationAuto1) against the oracle's sequential restatement.

Both the field and the final level of every point must agree: a point that
stopped one level apart would differ by up to relPrec in |E|^2, so equal
levels are what makes the 1e-4 field contract reachable with relPrec 0.01.
"""
from array import array

import numpy as np
import pytest

from srw_amd import workloads as W
from tests import oracle_lib as oracle
from tests.util import cfield, clone_wfr, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-4   # single-electron field contract (relative L2)


def _mesh(nx, ny, ne, x0, x1, y0, y1, de=0.0, z=20.):
    beam = W.electron_beam()
    e1 = W.first_harmonic_eV(beam.partStatMom1.gamma, 1.5, 0.02)
    e0, ef = (e1 * (1 - de), e1 * (1 + de)) if ne > 1 else (e1, e1)
    w = W.wavefront(beam, ne, e0, ef, nx, ny, 0, 0, z)
    w.mesh.xStart, w.mesh.xFin = x0, x1
    w.mesh.yStart, w.mesh.yFin = y0, y1
    return w


def _compare(w, mag, prec):
    from srw_amd import srwlpy
    wg, wo = clone_wfr(w), clone_wfr(w)
    print("case: mesh %dx%dx%d relPrec %g termin %d" % (w.mesh.nx, w.mesh.ny, w.mesh.ne, prec[1], prec[5]),
          flush=True)
    srwlpy.CalcElecFieldSR(wg, 0, mag, prec)
    lg = srwlpy.last_auto_levels()
    rc, lo = oracle.calc_field_levels(wo, mag, prec)
    assert rc == 0
    assert len(lg) == len(lo)
    mism = int(np.sum(lg != lo))
    ex = rel_l2(cfield(wg.arEx, wg), cfield(wo.arEx, wo))
    ig = np.abs(cfield(wg.arEx, wg)) ** 2 + np.abs(cfield(wg.arEy, wg)) ** 2
    io = np.abs(cfield(wo.arEx, wo)) ** 2 + np.abs(cfield(wo.arEy, wo)) ** 2
    ei = rel_l2(ig, io)
    print("levels %s..%s, mismatched %d; rel-L2 Ex %.2e I %.2e" % (lo.min(), lo.max(), mism, ex, ei), flush=True)
    assert mism == 0
    assert ex < TOL and ei < 2 * TOL
    assert wg.Rx == pytest.approx(wo.Rx, rel=1e-12)
    return lo


@pytest.mark.parametrize("case", [
    # symmetric mesh (half computed), mirrored
    dict(mesh=(21, 21, 1, -1e-3, 1e-3, -1e-3, 1e-3), rel=0.01),
    # off-centre mesh starting at a dim corner: thresholds tighten point by point
    dict(mesh=(17, 13, 3, 2e-4, 2.5e-3, -3e-4, 1.8e-3, 0.02), rel=0.01),
    # tighter precision, no terminating terms
    dict(mesh=(9, 7, 2, -8e-4, 6e-4, -4e-4, 5e-4, 0.05), rel=1e-3, termin=0),
    # very tight precision: levels beyond the device tables (trajectory
    # evaluated per point in the kernel)
    dict(mesh=(3, 2, 1, 1e-4, 3e-4, 1e-4, 2e-4), rel=1e-10),
])
def test_auto_matches_oracle(case):
    w = _mesh(*case["mesh"])
    prec = array('d', [1, case["rel"], 0, 0, 10001, case.get("termin", 1), 0])
    lo = _compare(w, W.planar_undulator(), prec)
    if case["rel"] <= 1e-10:
        assert lo.max() > 13


def test_auto_batch_matches_oracle():
    """Three requests in one batch call: each restarts the running maximum."""
    from srw_amd import srwlpy
    ws = [_mesh(11, 9, 1, -1e-3, 1e-3, -8e-4, 8e-4), _mesh(6, 5, 2, 5e-4, 2e-3, 0., 1e-3, 0.01),
          _mesh(7, 7, 1, -2e-3, 2e-3, -2e-3, 2e-3)]
    mags = [W.planar_undulator(K=1.5), W.planar_undulator(K=1.2), W.planar_undulator(K=1.5)]
    prec = array('d', [1, 0.01, 0, 0, 10001, 1, 0])
    wg = [clone_wfr(w) for w in ws]
    print("case: batch of 3 method-1 requests", flush=True)
    srwlpy.CalcElecFieldSRBatch(wg, mags, prec)
    lg = srwlpy.last_auto_levels()
    los = []
    for w, mag, g in zip(ws, mags, wg):
        wo = clone_wfr(w)
        rc, lo = oracle.calc_field_levels(wo, mag, prec)
        assert rc == 0
        los.append(lo)
        e = rel_l2(cfield(g.arEx, g), cfield(wo.arEx, wo))
        print("request rel-L2 Ex %.2e" % e, flush=True)
        assert e < TOL
    assert np.array_equal(lg, np.concatenate(los))


def test_auto_rejects_zero_precision():
    from srw_amd import srwlpy
    w = _mesh(3, 3, 1, -1e-3, 1e-3, -1e-3, 1e-3)
    with pytest.raises(RuntimeError):
        srwlpy.CalcElecFieldSR(w, 0, W.planar_undulator(), array('d', [1, 0., 0, 0, 10001, 1, 0]))


@pytest.mark.parametrize("meth,f", [(0, 1.0), (1, 0.5)])
def test_auto_mesh_matches_oracle(meth, f):
    """precPar[6] > 0: the library resizes the mesh from the wavefront radius
    and reallocates the Python wavefront through its callback."""
    from srw_amd import srwlpy
    w, mag, prec = W.single_electron(n=15)
    p = array('d', list(prec))
    p[0] = meth
    if meth == 1:
        p[1] = 0.01
    p[6] = f
    wg, wo = clone_wfr(w), clone_wfr(w)
    print("case: auto mesh, method %d, sampling factor %g" % (meth, f), flush=True)
    srwlpy.CalcElecFieldSR(wg, 0, mag, p)
    rc, sz = oracle.calc_field_auto_mesh(wo, mag, p)
    assert rc == 0
    assert (wg.mesh.nx, wg.mesh.ny) == sz == (wo.mesh.nx, wo.mesh.ny)
    assert len(wg.arEx) == 2 * sz[0] * sz[1] * w.mesh.ne
    e = rel_l2(cfield(wg.arEx, wg), cfield(wo.arEx, wo))
    print("mesh %dx%d rel-L2 Ex %.2e" % (sz[0], sz[1], e), flush=True)
    assert e < TOL
