"""GPU parity of the wavefront resampling (srwlResizeElecField and the resize
instructions of a propagation container) against the FP64/FP32 oracle."""
import numpy as np
import pytest

from srw_amd import workloads as W
from srw_amd.srwlib import SRWLOptC, SRWLOptD
from tests import oracle_lib as oracle
from tests.util import cfield, clone_wfr, rel_l2

pytestmark = pytest.mark.gpu

# Same float arithmetic on both sides; the device forms the bicubic sums with
# other fused operations in the double parts, so values agree to ~1e-7.
TOL = 1e-5


def _sr(n, ne=1, de=0.0):
    w, mag, prec = W.single_electron(n=n, ne=ne, de_rel=de)
    assert oracle.calc_field(w, mag, prec) == 0
    return w


def _compare(wg, wo, tol=TOL):
    assert (wg.mesh.nx, wg.mesh.ny) == (wo.mesh.nx, wo.mesh.ny)
    for a in ("xStart", "xFin", "yStart", "yFin"):
        assert getattr(wg.mesh, a) == pytest.approx(getattr(wo.mesh, a), rel=1e-12, abs=1e-18)
    ex = rel_l2(cfield(wg.arEx, wg), cfield(wo.arEx, wo))
    ey = rel_l2(cfield(wg.arEy, wg), cfield(wo.arEy, wo))
    print("mesh %dx%d  rel-L2 Ex %.3e Ey %.3e" % (wg.mesh.nx, wg.mesh.ny, ex, ey), flush=True)
    assert ex < tol and ey < tol


@pytest.mark.parametrize("n,ne,par", [
    (64, 1, [0, 2., 2., 2., 2.]),
    (64, 1, [0, 1., 3., 1., 3.]),
    (64, 1, [0, 0.5, 1., 0.5, 1.]),
    (64, 1, [0, 2., 1., 2., 1.]),
    (48, 3, [0, 1.5, 2., 1.2, 1.5]),
])
def test_resize_matches_oracle(n, ne, par):
    from srw_amd import srwlpy
    print("case: resize n=%d ne=%d par=%s" % (n, ne, par), flush=True)
    w = _sr(n, ne, 0.02 if ne > 1 else 0.)
    wg, wo = clone_wfr(w), clone_wfr(w)
    srwlpy.ResizeElecField(wg, 'c', par)
    rc, wo2 = oracle.resize(wo, par + [0.5, 0.5])
    assert rc == 0
    _compare(wg, wo2)


def test_propagation_with_resize_before_drift():
    # SR field 101x101 resampled to 512x512 (range x2, resolution x2.5347),
    # then a 20 m drift with the moments-based radius
    from srw_amd import srwlpy
    print("case: 101^2 -> 512^2 resize + 20 m drift", flush=True)
    w = _sr(101)
    pd = 512. / (2. * 101.)
    wg, wo = clone_wfr(w), clone_wfr(w)
    opt = SRWLOptC([SRWLOptD(20.)], [[0, 0, 1., 1, 0, 2., pd, 2., pd]])
    srwlpy.PropagElecField(wg, opt)
    rc, wo2 = oracle.resize(wo, [0, 2., pd, 2., pd, 0.5, 0.5])
    assert rc == 0 and wo2.mesh.nx == 512
    assert oracle.propagate(wo2, W.drift_container(20., 1)) == 0
    _compare(wg, wo2, tol=2e-5)


def test_propagation_multi_energy_resize_per_energy():
    # the 201-energy workload in small: 5 energies of a 41x41 SR field,
    # resampled to 128x128 (range x3.1, step kept) and drifted 20 m with the
    # moments-based radius; the error is reported per photon energy
    from srw_amd import srwlpy
    print("case: 5 energies, 41^2 -> 128^2 resize + 20 m drift", flush=True)
    w = _sr(41, ne=5, de=0.03)
    pm = 127. / 41.
    wg, wo = clone_wfr(w), clone_wfr(w)
    opt = SRWLOptC([SRWLOptD(20.)], [[0, 0, 1., 1, 0, pm, 1., pm, 1.]])
    srwlpy.PropagElecField(wg, opt)
    rc, wo2 = oracle.resize(wo, [0, pm, 1., pm, 1., 0.5, 0.5])
    assert rc == 0 and wo2.mesh.nx == 128
    assert oracle.propagate(wo2, W.drift_container(20., 1)) == 0
    _compare(wg, wo2, tol=2e-5)
    exg, exo = cfield(wg.arEx, wg), cfield(wo2.arEx, wo2)
    for ie in range(w.mesh.ne):
        e = rel_l2(exg[:, :, ie], exo[:, :, ie])
        print("energy %d: rel-L2 Ex %.3e" % (ie, e), flush=True)
        assert e < 2e-5


def _mesh(ne, e, nx, x0, x1, ny, y0, y1):
    from srw_amd.srwlib import SRWLRadMesh
    m = SRWLRadMesh()
    m.ne, m.nx, m.ny = ne, nx, ny
    m.eStart, m.eFin = e
    m.xStart, m.xFin, m.yStart, m.yFin = x0, x1, y0, y1
    return m


@pytest.mark.parametrize("n,ne,box,par", [
    (64, 1, (100, -8e-4, 8e-4, 90, -7e-4, 9e-4), [0, 1, 1]),     # finer, shifted
    (64, 1, (128, -2e-3, 1e-3, 128, -1.5e-3, 2e-3), [0, 1, 1]),  # larger: zeros outside
    (64, 1, (48, -5e-4, 5e-4, 48, -5e-4, 5e-4), [0, 0, 1]),      # coarser, no phase treatment
    (48, 3, (96, -9e-4, 9e-4, 80, -9e-4, 9e-4), [0, 1, 0]),      # no Re/Im correction, 3 energies
])
def test_resize_to_mesh_matches_oracle(n, ne, box, par):
    from srw_amd import srwlpy
    print("case: resize to mesh n=%d ne=%d box=%s par=%s" % (n, ne, box, par), flush=True)
    w = _sr(n, ne, 0.02 if ne > 1 else 0.)
    m = _mesh(ne, (w.mesh.eStart, w.mesh.eFin), *box)
    wg, wo = clone_wfr(w), clone_wfr(w)
    srwlpy.ResizeElecFieldMesh(wg, m, par)
    rc, wo2 = oracle.resize_to_mesh(wo, m, par)
    assert rc == 0
    _compare(wg, wo2)


@pytest.mark.parametrize("n,ne,par", [
    (64, 1, [0, 2., 2.5, 2., 2.5]),
    (48, 3, [0, 1.5, 2., 1.2, 1.5]),
    (64, 1, [0, 0.5, 1., 0.5, 1.]),
])
def test_resize_per_cell_equals_per_point(n, ne, par):
    """The per-cell kernel (forms built once per old cell) gives the same bits
    as the per-point kernel."""
    from srw_amd import srwlpy
    print("case: resize modes n=%d ne=%d par=%s" % (n, ne, par), flush=True)
    w = _sr(n, ne, 0.02 if ne > 1 else 0.)
    wa, wb = clone_wfr(w), clone_wfr(w)
    srwlpy.ResizeElecField(wa, 'c', par)
    srwlpy.set_resize_per_point(True)
    try:
        srwlpy.ResizeElecField(wb, 'c', par)
    finally:
        srwlpy.set_resize_per_point(False)
    assert (wa.mesh.nx, wa.mesh.ny) == (wb.mesh.nx, wb.mesh.ny)
    assert np.array_equal(np.frombuffer(wa.arEx, np.float32), np.frombuffer(wb.arEx, np.float32))
    assert np.array_equal(np.frombuffer(wa.arEy, np.float32), np.frombuffer(wb.arEy, np.float32))
