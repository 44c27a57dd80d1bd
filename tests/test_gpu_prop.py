"""GPU parity of the SRW 2D FFT, the drift propagation and intensity extraction."""
from array import array

import numpy as np
import pytest

from srw_amd import workloads as W
from tests import oracle_lib as oracle
from tests.util import cfield, clone_wfr, rel_l2

pytestmark = pytest.mark.gpu

# FP32 FFT of an FP32 field compared with an FP64 oracle: the error is set by
# FP32 rounding of O(log N) butterfly stages.
TOL_FFT = 2e-6
TOL_DRIFT = 5e-6


def _rand_data(n, ny, seed):
    rng = np.random.default_rng(seed)
    z = (rng.standard_normal((ny, n)) + 1j * rng.standard_normal((ny, n))).astype(np.complex64)
    return array('f', z.view(np.float32).ravel().tobytes())


@pytest.mark.parametrize("n,ny,xs,dirn", [(64, 64, None, 1), (128, 32, None, -1), (256, 256, 3e-4, 1),
                                         (1024, 16, -2e-4, -1), (2048, 64, None, 1)])
def test_uti_fft(n, ny, xs, dirn):
    from srw_amd import srwlpy
    print("case: UtiFFT", n, ny, xs, dirn, flush=True)
    d1 = _rand_data(n, ny, n + ny)
    d2 = array('f', d1)
    dx, dy = 1e-5, 2e-5
    x0 = -0.5 * n * dx if xs is None else xs
    mesh1 = [x0, dx, n, -0.5 * ny * dy, dy, ny]
    mesh2 = list(mesh1)
    srwlpy.UtiFFT(d1, mesh1, dirn)
    assert oracle.uti_fft(d2, mesh2, dirn) == 0
    a = np.frombuffer(d1, np.float32).view(np.complex64)
    b = np.frombuffer(d2, np.float32).view(np.complex64)
    err = rel_l2(a, b)
    print("rel-L2 %.3e" % err, flush=True)
    assert err < TOL_FFT
    assert mesh1 == pytest.approx(mesh2, rel=1e-12)


def _sr_wfr(n, ne=1, de=0.0):
    w, mag, prec = W.single_electron(n=n, ne=ne, de_rel=de)
    assert oracle.calc_field(w, mag, prec) == 0
    return w


@pytest.mark.parametrize("n,ne,mode,L", [(128, 1, 2, 20.), (256, 1, 0, 10.), (128, 3, 2, 30.), (64, 2, 0, 5.),
                                         (128, 1, 1, 20.), (128, 3, 1, 30.)])
def test_drift(n, ne, mode, L):
    from srw_amd import srwlpy
    print("case: drift n=%d ne=%d mode=%d L=%g" % (n, ne, mode, L), flush=True)
    w = _sr_wfr(n, ne, 0.03 if ne > 1 else 0.)
    for i in range(20):
        w.arElecPropMatr[i] = 1. if i in (0, 5, 10, 15) else 0.
    w.arElecPropMatr[16], w.arElecPropMatr[17] = 1e-6, 2e-6
    wg, wo = clone_wfr(w), clone_wfr(w)
    opt = W.drift_container(L, mode)
    srwlpy.PropagElecField(wg, opt)
    assert oracle.propagate(wo, opt) == 0
    for attr in ("xStart", "xFin", "yStart", "yFin"):
        assert getattr(wg.mesh, attr) == pytest.approx(getattr(wo.mesh, attr), rel=1e-12)
    assert wg.Rx == pytest.approx(wo.Rx)
    ex_e = rel_l2(cfield(wg.arEx, wg), cfield(wo.arEx, wo))
    ig = np.abs(cfield(wg.arEx, wg)) ** 2 + np.abs(cfield(wg.arEy, wg)) ** 2
    io = np.abs(cfield(wo.arEx, wo)) ** 2 + np.abs(cfield(wo.arEy, wo)) ** 2
    i_e = rel_l2(ig, io)
    print("rel-L2 Ex %.3e I %.3e" % (ex_e, i_e), flush=True)
    assert ex_e < TOL_DRIFT
    assert i_e < 2 * TOL_DRIFT
    # moments and the 4x4 matrix are carried through the drift the same way
    assert np.allclose(np.array(wg.arMomX), np.array(wo.arMomX), rtol=1e-12, atol=0)
    assert np.allclose(np.array(wg.arMomY), np.array(wo.arMomY), rtol=1e-12, atol=0)
    assert np.allclose(np.array(wg.arElecPropMatr), np.array(wo.arElecPropMatr), rtol=1e-14, atol=0)


@pytest.mark.parametrize("n,ne", [(128, 1), (64, 3)])
def test_drift_radius_from_recomputed_moments(n, ne):
    # arProp[3] = 1 with empty moment arrays: both sides compute the moments of
    # the input field first; the device sums (other order, FP32 input) move the
    # radius, hence the output mesh, by ~1e-7 at most
    from srw_amd import srwlpy
    print("case: drift subtype 1, moments recomputed, n=%d ne=%d" % (n, ne), flush=True)
    w = _sr_wfr(n, ne, 0.03 if ne > 1 else 0.)
    for i in range(len(w.arMomX)):
        w.arMomX[i] = 0.
        w.arMomY[i] = 0.
    wg, wo = clone_wfr(w), clone_wfr(w)
    opt = W.drift_container(25., 1)
    srwlpy.PropagElecField(wg, opt)
    assert oracle.propagate(wo, opt) == 0
    for attr in ("xStart", "xFin", "yStart", "yFin"):
        assert getattr(wg.mesh, attr) == pytest.approx(getattr(wo.mesh, attr), rel=1e-6)
    ig = np.abs(cfield(wg.arEx, wg)) ** 2 + np.abs(cfield(wg.arEy, wg)) ** 2
    io = np.abs(cfield(wo.arEx, wo)) ** 2 + np.abs(cfield(wo.arEy, wo)) ** 2
    i_e = rel_l2(ig, io)
    print("rel-L2 I %.3e" % i_e, flush=True)
    assert i_e < 1e-4


@pytest.mark.parametrize("pol", [0, 1, 2, 3, 4, 5, 6, -1, -2, -3, -4])
def test_intensity(pol):
    from srw_amd import srwlpy
    w = _sr_wfr(16, ne=3, de=0.02)
    e = w.mesh.eStart + 0.37 * (w.mesh.eFin - w.mesh.eStart)
    a1 = array('f', [0] * (16 * 16))
    a2 = array('f', [0] * (16 * 16))
    srwlpy.CalcIntFromElecField(a1, w, pol, 0, 3, e, 0, 0)
    assert oracle.intensity(a2, w, pol, 0, 3, e) == 0
    x, y = np.array(a1), np.array(a2)
    assert rel_l2(x, y) < 1e-6


@pytest.mark.parametrize("n,mode,tiles", [
    (256, 2, (1, 1024, 4, 256, 4, 512, 0)),
    (256, 0, (8, 256, 16, 1024, 2, 256, 0)),
    (128, 2, (16, 512, 32, 1024, 16, 1024, 0)),
    (1024, 2, (2, 256, 8, 1024, 2, 256, 1)),
    (256, 0, (4, 256, 4, 512, 4, 256, 1)),
])
def test_drift_tile_shapes_give_same_values(n, mode, tiles):
    """Any tile shape of the FFT passes (the autotuner's search space) gives
    the same bits as the default shape, and the autotuned shape as well."""
    from srw_amd import srwlpy
    print("case: drift tiles n=%d mode=%d tiles=%s" % (n, mode, tiles), flush=True)
    w, mag, prec = W.single_electron(n=n)
    srwlpy.CalcElecFieldSR(w, 0, mag, prec)   # the field only feeds the comparison of tile shapes
    opt = W.drift_container(20., mode)
    res = []
    for setting in ("default", "forced", "tuned"):
        wg = clone_wfr(w)
        if setting == "default":
            srwlpy.set_tuning(False)
        elif setting == "forced":
            srwlpy.set_fft_tiles(tiles)
        try:
            srwlpy.PropagElecField(wg, opt)
        finally:
            srwlpy.set_tuning(True)
            srwlpy.set_fft_tiles(None)
        res.append(np.frombuffer(wg.arEx, np.float32).copy())
    lg = int(np.log2(n))
    print("tuned tiles", srwlpy.fft_tiles(lg, lg, 3 if mode else 0, 1 if mode else 0), flush=True)
    assert np.array_equal(res[0], res[1])
    assert np.array_equal(res[0], res[2])


@pytest.mark.parametrize("n,ny,xs,dirn", [(120, 112, None, 1), (66, 52, 3e-4, -1), (390, 30, None, 1),
                                         (2000, 16, -1e-4, -1), (14, 26, None, 1), (1024, 96, None, 1)])
def test_uti_fft_mixed_radix(n, ny, xs, dirn):
    """Even sizes with factors 2, 3, 5, 7, 11, 13 (and one power-of-2 axis
    beside such a size): the mixed-radix passes."""
    from srw_amd import srwlpy
    print("case: UtiFFT mixed radix", n, ny, xs, dirn, flush=True)
    d1 = _rand_data(n, ny, n * 7 + ny)
    d2 = array('f', d1)
    dx, dy = 1e-5, 2e-5
    x0 = -0.5 * n * dx if xs is None else xs
    mesh1 = [x0, dx, n, -0.5 * ny * dy, dy, ny]
    mesh2 = list(mesh1)
    srwlpy.UtiFFT(d1, mesh1, dirn)
    assert oracle.uti_fft(d2, mesh2, dirn) == 0
    a = np.frombuffer(d1, np.float32).view(np.complex64)
    b = np.frombuffer(d2, np.float32).view(np.complex64)
    err = rel_l2(a, b)
    print("rel-L2 %.3e" % err, flush=True)
    assert err < TOL_FFT
    assert mesh1 == pytest.approx(mesh2, rel=1e-12)


@pytest.mark.parametrize("nx,ny,ne,mode,L", [(120, 112, 1, 2, 20.), (90, 66, 2, 0, 10.), (150, 104, 1, 1, 25.)])
def test_drift_mixed_radix(nx, ny, ne, mode, L):
    from srw_amd import srwlpy
    print("case: drift mixed radix %dx%d ne=%d mode=%d L=%g" % (nx, ny, ne, mode, L), flush=True)
    w, mag, prec = W.single_electron(n=nx, ne=ne, de_rel=0.03 if ne > 1 else 0.)
    w.allocate(ne, nx, ny)
    assert oracle.calc_field(w, mag, prec) == 0
    wg, wo = clone_wfr(w), clone_wfr(w)
    opt = W.drift_container(L, mode)
    srwlpy.PropagElecField(wg, opt)
    assert oracle.propagate(wo, opt) == 0
    for attr in ("xStart", "xFin", "yStart", "yFin"):
        assert getattr(wg.mesh, attr) == pytest.approx(getattr(wo.mesh, attr), rel=1e-6)
    ig = np.abs(cfield(wg.arEx, wg)) ** 2 + np.abs(cfield(wg.arEy, wg)) ** 2
    io = np.abs(cfield(wo.arEx, wo)) ** 2 + np.abs(cfield(wo.arEy, wo)) ** 2
    ex_e = rel_l2(cfield(wg.arEx, wg), cfield(wo.arEx, wo))
    i_e = rel_l2(ig, io)
    print("rel-L2 Ex %.3e I %.3e" % (ex_e, i_e), flush=True)
    assert ex_e < TOL_DRIFT and i_e < 2 * TOL_DRIFT


def test_resize_to_good_size_then_drift():
    """A resize to a non-power-of-2 fast-FFT size (SRW's own size rule),
    then the drift of that mesh, against the oracle."""
    from srw_amd import srwlpy
    from srw_amd.srwlib import SRWLOptC, SRWLOptD
    print("case: 101^2 resized x1.5 range, x1.3 resolution, 20 m drift", flush=True)
    w = _sr_wfr(101)
    wg, wo = clone_wfr(w), clone_wfr(w)
    opt = SRWLOptC([SRWLOptD(20.)], [[0, 0, 1., 1, 0, 1.5, 1.3, 1.5, 1.3]])
    srwlpy.PropagElecField(wg, opt)
    rc, wo = oracle.resize(wo, [0, 1.5, 1.3, 1.5, 1.3, 0.5, 0.5])
    assert rc == 0
    assert oracle.propagate(wo, W.drift_container(20., 1)) == 0
    print("mesh %dx%d" % (wg.mesh.nx, wg.mesh.ny), flush=True)
    assert (wg.mesh.nx, wg.mesh.ny) == (wo.mesh.nx, wo.mesh.ny)
    assert wg.mesh.nx & (wg.mesh.nx - 1) != 0
    ig = np.abs(cfield(wg.arEx, wg)) ** 2 + np.abs(cfield(wg.arEy, wg)) ** 2
    io = np.abs(cfield(wo.arEx, wo)) ** 2 + np.abs(cfield(wo.arEy, wo)) ** 2
    i_e = rel_l2(ig, io)
    print("rel-L2 I %.3e" % i_e, flush=True)
    assert i_e < 1e-4
