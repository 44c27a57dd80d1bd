"""GPU srwlCalcIntFromElecField against the oracle: every polarisation
component and Stokes parameter, intensity / Re E / Im E / phase / intensity
integrated over photon energy / single-electron flux, every dependence type,
at coordinates on, between and outside the mesh points, and the arMeth
running-average and sum modes.  The device arithmetic repeats the
reference's float/double steps with contraction off, so the outputs must be
bit-identical (the phase within one float ulp: its arctangent comes from the
device's libm)."""
from array import array

import numpy as np
import pytest

from tests import oracle_lib as oracle
from tests.test_oracle_extract import E0, E1, LENS, NE, NX, NY, X0, X1, Y0, Y1, _wfr

pytestmark = pytest.mark.gpu

POINTS = [(117.3, -4.1e-4, 7.7e-4), (E0, X1, Y0), (90., 5e-3, -3e-3)]


def _field(seed):
    rng = np.random.default_rng(seed)
    shp = (NY, NX, NE)
    return _wfr(rng.normal(size=shp) + 1j * rng.normal(size=shp), rng.normal(size=shp) + 1j * rng.normal(size=shp))


@pytest.mark.parametrize("it", [0, 2, 4, 5, 6, 7])
@pytest.mark.parametrize("dep", range(7))
def test_extract_bit_exact(it, dep):
    from srw_amd import srwlpy
    w = _field(11 + dep)
    pols = [0, 1, 2, 3, 4, 5, 6, -1, -2, -3, -4, -5]
    print("case: intType %d depType %d" % (it, dep), flush=True)
    typ = 'd' if it == 4 else 'f'
    dt = np.float64 if it == 4 else np.float32
    for pol in pols:
        for (e, x, y) in POINTS:
            n = LENS[dep] * (4 if (pol == -5 and it != 2) else 1)
            ag = array(typ, [0] * n)
            ao = array(typ, [0] * n)
            if it == 2 and dep != 0:
                # flux only vs photon energy: a warning and a zero first value
                srwlpy.CalcIntFromElecField(ag, w, pol, it, dep, e, x, y)
                assert srwlpy.last_code == oracle.intensity(ao, w, pol, it, dep, e, x, y) < 0
                assert ag[0] == ao[0] == 0
                continue
            srwlpy.CalcIntFromElecField(ag, w, pol, it, dep, e, x, y)
            assert oracle.intensity(ao, w, pol, it, dep, e, x, y) == 0
            g, o = np.frombuffer(ag, dt), np.frombuffer(ao, dt)
            if it == 4:
                # phase: float values stored as double; one float ulp for the arctangent
                assert np.all(np.abs(g - o) <= np.spacing(np.abs(o).astype(np.float32)).astype(np.float64) + 1e-300), \
                    (pol, e, x, y, float(np.max(np.abs(g - o))))
            else:
                assert np.array_equal(g, o), (pol, e, x, y, float(np.max(np.abs(g - o))))


@pytest.mark.parametrize("it", [0, 4, 5, 7])
def test_extract_averaging_and_sum_modes(it):
    """arMeth[0] = 1 (running average over arMeth[1] previous values) and 2
    (sum) update the caller's array for the dependence on x and y; other
    dependence types overwrite, as in the reference."""
    from srw_amd import srwlpy
    w = _field(5)
    rng = np.random.default_rng(9)
    typ, dt = ('d', np.float64) if it == 4 else ('f', np.float32)
    for meth in ([1, 3], [1, 0], [2], [0]):
        for dep in (3, 1):
            for pol in (6, 0, -5):
                n = LENS[dep] * (4 if pol == -5 else 1)
                pre = rng.normal(size=n).astype(dt)
                ag, ao = array(typ, pre.tobytes()), array(typ, pre.tobytes())
                srwlpy.CalcIntFromElecField(ag, w, pol, it, dep, 117.3, -4.1e-4, 7.7e-4, meth)
                assert oracle.intensity(ao, w, pol, it, dep, 117.3, -4.1e-4, 7.7e-4, meth=meth) == 0
                g, o = np.frombuffer(ag, dt), np.frombuffer(ao, dt)
                if it == 4:
                    assert np.allclose(g, o, rtol=1e-6, atol=1e-6)
                else:
                    assert np.array_equal(g, o), (meth, dep, pol)


def test_extract_single_energy_mesh():
    """ne = 1: the energy coordinate collapses to the one slice."""
    from srw_amd import srwlpy
    from srw_amd.srwlib import SRWLWfr
    rng = np.random.default_rng(2)
    w = SRWLWfr()
    w.allocate(1, 9, 4)
    w.mesh.eStart = w.mesh.eFin = 1000.
    w.mesh.xStart, w.mesh.xFin, w.mesh.yStart, w.mesh.yFin = X0, X1, Y0, Y1
    w.arEx[:] = array('f', rng.normal(size=72).astype(np.float32).tobytes())
    w.arEy[:] = array('f', rng.normal(size=72).astype(np.float32).tobytes())
    for dep, n in ((1, 9), (2, 4), (3, 36), (0, 1)):
        ag, ao = array('f', [0] * n), array('f', [0] * n)
        srwlpy.CalcIntFromElecField(ag, w, 6, 0, dep, 1234., 1e-4, 3e-4)
        assert oracle.intensity(ao, w, 6, 0, dep, 1234., 1e-4, 3e-4) == 0
        assert ag == ao


def test_extract_rejects_unsupported():
    from srw_amd import srwlpy
    w = _field(1)
    with pytest.raises(RuntimeError):
        srwlpy.CalcIntFromElecField(array('f', [0] * LENS[3]), w, 6, 1, 3, E0, 0, 0)   # multi-electron intensity
    with pytest.raises(RuntimeError):
        srwlpy.CalcIntFromElecField(array('f', [0] * LENS[0]), w, 6, 3, 0, E0, 0, 0)   # multi-electron flux
    with pytest.raises(RuntimeError):
        srwlpy.CalcIntFromElecField(array('f', [0] * LENS[3]), w, 6, 0, 7, E0, 0, 0)   # no such dependence
    with pytest.raises(RuntimeError):
        srwlpy.CalcIntFromElecField(array('f', [0] * LENS[3]), w, 6, 8, 3, E0, 0, 0)   # mutual intensity
