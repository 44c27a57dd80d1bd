"""GPU srwlCalcIntFromElecField against the oracle: every polarisation
component and Stokes parameter, intensity / Re E / Im E, every dependence
type, at coordinates on, between and outside the mesh points.  The device
arithmetic repeats the reference's float/double steps with contraction off,
so the outputs must be bit-identical."""
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


@pytest.mark.parametrize("it", [0, 5, 6])
@pytest.mark.parametrize("dep", range(7))
def test_extract_bit_exact(it, dep):
    from srw_amd import srwlpy
    w = _field(11 + dep)
    pols = [0, 1, 2, 3, 4, 5, 6, -1, -2, -3, -4] + ([-5] if it == 0 else [])
    print("case: intType %d depType %d" % (it, dep), flush=True)
    for pol in pols:
        for (e, x, y) in POINTS:
            n = LENS[dep] * (4 if pol == -5 else 1)
            ag = array('f', [0] * n)
            ao = array('f', [0] * n)
            srwlpy.CalcIntFromElecField(ag, w, pol, it, dep, e, x, y)
            assert oracle.intensity(ao, w, pol, it, dep, e, x, y) == 0
            g, o = np.frombuffer(ag, np.float32), np.frombuffer(ao, np.float32)
            assert np.array_equal(g, o), (pol, e, x, y, float(np.max(np.abs(g - o))))


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
        srwlpy.CalcIntFromElecField(array('f', [0] * LENS[3]), w, 6, 1, 3, E0, 0, 0)   # multi-electron
    with pytest.raises(RuntimeError):
        srwlpy.CalcIntFromElecField(array('f', [0] * LENS[3]), w, 6, 0, 7, E0, 0, 0)
