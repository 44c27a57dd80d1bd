"""GPU parity of containers with thin lenses, apertures and obstacles against
the oracle: fields, moments, wavefront radius, centre and 4x4 matrix."""
import numpy as np
import pytest

from srw_amd import workloads as W
from srw_amd.srwlib import SRWLOptA, SRWLOptC, SRWLOptD, SRWLOptL
from tests import oracle_lib as oracle
from tests.util import cfield, clone_wfr, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _prop(mode=0):
    return [0, 0, 1.0, mode, 0, 1.0, 1.0, 1.0, 1.0]


def _sr_wfr(n, ne=1, de=0.0):
    w, mag, prec = W.single_electron(n=n, ne=ne, de_rel=de)
    assert oracle.calc_field(w, mag, prec) == 0
    for i in range(20):
        w.arElecPropMatr[i] = 1. if i in (0, 5, 10, 15) else 0.
    return w


CASES = {
    "aperture-lens-drift": (1, lambda: SRWLOptC([SRWLOptA('r', 'a', 1.2e-3, 0.9e-3, 5e-5, 0.), SRWLOptL(10., 12.),
                                                 SRWLOptD(10.)], [_prop(), _prop(), _prop(2)])),
    "obstacle-drift": (3, lambda: SRWLOptC([SRWLOptA('c', 'o', 3e-4, 0., 0., 1e-4), SRWLOptD(5.)],
                                           [_prop(), _prop(0)])),
    "lens-drift-moments": (2, lambda: SRWLOptC([SRWLOptL(8., 8., 1e-5, -1e-5), SRWLOptD(12.)],
                                               [_prop(), _prop(1)])),
    "circ-aperture-lens": (1, lambda: SRWLOptC([SRWLOptA('c', 'a', 1.5e-3, 0., 0., 0.), SRWLOptL(-20., 30.)],
                                               [_prop(), _prop()])),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_container_matches_oracle(name):
    from srw_amd import srwlpy
    ne, mk = CASES[name]
    print("case: container %s, n=128 ne=%d" % (name, ne), flush=True)
    w = _sr_wfr(128, ne, 0.03 if ne > 1 else 0.)
    wg, wo = clone_wfr(w), clone_wfr(w)
    srwlpy.PropagElecField(wg, mk())
    assert oracle.propagate(wo, mk()) == 0
    for attr in ("xStart", "xFin", "yStart", "yFin"):
        assert getattr(wg.mesh, attr) == pytest.approx(getattr(wo.mesh, attr), rel=1e-9)
    for attr in ("Rx", "Ry", "dRx", "dRy", "xc", "yc"):
        assert getattr(wg, attr) == pytest.approx(getattr(wo, attr), rel=1e-9, abs=1e-15)
    ex = rel_l2(cfield(wg.arEx, wg), cfield(wo.arEx, wo))
    ig = np.abs(cfield(wg.arEx, wg)) ** 2 + np.abs(cfield(wg.arEy, wg)) ** 2
    io = np.abs(cfield(wo.arEx, wo)) ** 2 + np.abs(cfield(wo.arEy, wo)) ** 2
    ei = rel_l2(ig, io)
    print("rel-L2 Ex %.3e I %.3e" % (ex, ei), flush=True)
    assert ex < TOL and ei < 2 * TOL
    # moments: transformed exactly, or recomputed from the field (device sums)
    np.testing.assert_allclose(np.array(wg.arMomX), np.array(wo.arMomX), rtol=1e-4, atol=1e-30)
    np.testing.assert_allclose(np.array(wg.arMomY), np.array(wo.arMomY), rtol=1e-4, atol=1e-30)
    np.testing.assert_allclose(np.array(wg.arElecPropMatr), np.array(wo.arElecPropMatr), rtol=1e-14, atol=0)


def test_lens_with_linear_phase_term_is_rejected():
    """A lens centred far from the wavefront centre would make the reference
    subtract a linear phase term first; the library reports that case."""
    from srw_amd import srwlpy
    w = _sr_wfr(64)
    w.Rx = w.Ry = 20.
    w.dRx = w.dRy = 0.1
    opt = lambda: SRWLOptC([SRWLOptL(10., 10., 8e-4, 0.)], [_prop()])
    assert oracle.propagate(clone_wfr(w), opt()) != 0
    with pytest.raises(RuntimeError):
        srwlpy.PropagElecField(clone_wfr(w), opt())


def test_wavefront_without_moment_arrays_aperture_then_drift():
    """A wavefront with no moment arrays: the arProp[3] = 1 drift after an
    aperture takes its radius from moments of the clipped field, as the
    oracle (fresh scratch moments for every drift)."""
    from srw_amd import srwlpy
    w = _sr_wfr(128)
    w.arMomX = w.arMomY = None
    mk = lambda: SRWLOptC([SRWLOptA('r', 'a', 0.6e-3, 0.5e-3, 1e-4, 0.), SRWLOptD(10.)], [_prop(), _prop(1)])
    wg, wo = clone_wfr(w), clone_wfr(w)
    srwlpy.PropagElecField(wg, mk())
    assert oracle.propagate(wo, mk()) == 0
    for attr in ("xStart", "xFin", "yStart", "yFin"):
        assert getattr(wg.mesh, attr) == pytest.approx(getattr(wo.mesh, attr), rel=1e-9)
    assert rel_l2(cfield(wg.arEx, wg), cfield(wo.arEx, wo)) < TOL


def test_multi_electron_beamline_with_resize_is_refused():
    """srwlEmitPropMultiE does not resample per electron: a beamline whose
    arProp asks for resampling is refused instead of being run unresampled."""
    from srw_amd import srwlpy
    from srw_amd.srwlib import SRWLStokes
    beam = W.beam_for_multi_e()
    e1 = W.first_harmonic_eV(beam.partStatMom1.gamma, 1.5, 0.02)
    w0 = W.wavefront(beam, 1, e1, e1, 64, 64, 1e-3, 1e-3, 20.)
    st = SRWLStokes()
    st.allocate(1, 64, 64, _n_comp=1)
    st.mesh.eStart = st.mesh.eFin = e1
    st.mesh.xStart, st.mesh.xFin, st.mesh.yStart, st.mesh.yFin = -2e-3, 2e-3, -2e-3, 2e-3
    for opt in (SRWLOptC([SRWLOptL(10., 10.), SRWLOptD(5.)], [[0, 0, 1.0, 0, 0, 2.0, 1.0, 2.0, 1.0], _prop(2)]),
                SRWLOptC([SRWLOptD(5.)], [_prop(2), [0, 0, 1.0, 0, 0, 1.0, 2.0, 1.0, 2.0]])):
        with pytest.raises(RuntimeError):
            srwlpy.EmitPropMultiE(st, w0, W.planar_undulator(), opt, W.traj_prec(10001), 16, _seed=3, _batch=16)
