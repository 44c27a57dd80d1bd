"""GPU parity of the single-electron SR field against the FP64 oracle."""
import numpy as np
import pytest

from srw_amd import workloads as W
from tests import oracle_lib as oracle
from tests.util import cfield, clone_wfr, rel_l2

pytestmark = pytest.mark.gpu

# Tolerance stated by the path's contract: single-electron fields within
# 1e-4 relative L2 of the FP64 reference computation.
TOL = 1e-4


def _run_pair(w, mag, prec):
    from srw_amd import srwlpy
    wg = clone_wfr(w)
    wo = clone_wfr(w)
    srwlpy.CalcElecFieldSR(wg, 0, mag, prec)
    rc = oracle.calc_field(wo, mag, prec)
    assert rc == 0
    return wg, wo


def _check(wg, wo, tol=TOL):
    exg, exo = cfield(wg.arEx, wg), cfield(wo.arEx, wo)
    eyg, eyo = cfield(wg.arEy, wg), cfield(wo.arEy, wo)
    ig = np.abs(exg) ** 2 + np.abs(eyg) ** 2
    io = np.abs(exo) ** 2 + np.abs(eyo) ** 2
    e_ex = rel_l2(exg, exo)
    e_i = rel_l2(ig, io)
    print("rel-L2 Ex %.3e  I %.3e" % (e_ex, e_i), flush=True)
    assert e_ex < tol
    assert e_i < 2 * tol
    if np.linalg.norm(eyo) > 1e-6 * np.linalg.norm(exo):
        assert rel_l2(eyg, eyo) < tol
    assert wg.Rx == pytest.approx(wo.Rx, rel=1e-12)
    assert wg.xc == pytest.approx(wo.xc, rel=1e-9, abs=1e-15)
    _check_moments(wg, wo)


def _check_moments(wg, wo):
    """Statistical moments (11 per energy and component) against the oracle.
    First moments are compared on the scale of the rms size / divergence."""
    for name in ("arMomX", "arMomY"):
        g = np.array(getattr(wg, name)).reshape(-1, 11)[:wg.mesh.ne]
        o = np.array(getattr(wo, name)).reshape(-1, 11)[:wo.mesh.ne]
        for ie in range(wg.mesh.ne):
            a, b = g[ie], o[ie]
            if b[0] == 0:
                assert a[0] == 0
                continue
            sx, sz = np.sqrt(abs(b[5])), np.sqrt(abs(b[8]))
            sxp, szp = np.sqrt(abs(b[7])), np.sqrt(abs(b[10]))
            d = [abs(a[0] - b[0]) / abs(b[0]), abs(a[1] - b[1]) / sx, abs(a[2] - b[2]) / sxp,
                 abs(a[3] - b[3]) / sz, abs(a[4] - b[4]) / szp,
                 abs(a[5] - b[5]) / b[5], abs(a[6] - b[6]) / (sx * sxp), abs(a[7] - b[7]) / b[7],
                 abs(a[8] - b[8]) / b[8], abs(a[9] - b[9]) / (sz * szp), abs(a[10] - b[10]) / b[10]]
            if ie == 0:
                print("%s moment deviations: max %.2e" % (name, max(d)), flush=True)
            assert max(d) < 1e-4, (name, ie, d)


def test_planar_undulator_small_mesh():
    print("case: 21x21 mesh, 1 energy, 10001 traj pts", flush=True)
    w, mag, prec = W.single_electron(n=21)
    _check(*_run_pair(w, mag, prec))


def test_planar_undulator_c1_full():
    print("case: 101x101 mesh, 1 energy (C1)", flush=True)
    w, mag, prec = W.single_electron(n=101)
    _check(*_run_pair(w, mag, prec))


def test_offset_electron_no_symmetry():
    print("case: offset/tilted electron, 32x24 mesh", flush=True)
    w, mag, prec = W.single_electron(n=32)
    w.mesh.ny = 24
    w.allocate(1, 32, 24)
    p = w.partBeam.partStatMom1
    p.x, p.xp, p.y, p.yp = 40e-6, 6e-6, -7e-6, 3e-6
    w.partBeam.partStatMom1.gamma *= 1.0008
    _check(*_run_pair(w, mag, prec))


def test_multi_energy_small():
    print("case: 11 energies +/-5%, 15x15 mesh", flush=True)
    w, mag, prec = W.single_electron(n=15, ne=11, de_rel=0.05)
    _check(*_run_pair(w, mag, prec))


def test_termin_modes_and_step():
    for termin in (0, 2, 3):
        print("case: termin", termin, flush=True)
        w, mag, prec = W.single_electron(n=11)
        prec[5] = termin
        prec[1] = prec[1] * 1.7   # integration step not equal to the trajectory step
        _check(*_run_pair(w, mag, prec))


def test_uneven_batch_matches_single_requests():
    # different meshes, harmonics and undulator strengths in one device pass
    from srw_amd import srwlpy
    print("case: uneven batch of 6 requests", flush=True)
    wfrs, mags, prec = W.uneven_requests(n=6, seed=3, meshes=(8, 12, 16))
    wg = [clone_wfr(w) for w in wfrs]
    srwlpy.CalcElecFieldSRBatch(wg, mags, prec)
    for w, g, m in zip(wfrs, wg, mags):
        wo = clone_wfr(w)
        assert oracle.calc_field(wo, m, prec) == 0
        _check(g, wo)
