"""GPU parity of the partially coherent (multi-electron) intensity."""
import numpy as np
import pytest

from srw_amd import workloads as W
from srw_amd.srwlib import SRWLStokes
from tests import oracle_lib as oracle
from tests.util import rel_l2

pytestmark = pytest.mark.gpu


def _setup(n=64, L=30., mode=2, half_res=2.4e-3, nres=48):
    beam = W.beam_for_multi_e()
    g = beam.partStatMom1.gamma
    e1 = W.first_harmonic_eV(g, 1.5, 0.02)
    w0 = W.wavefront(beam, 1, e1, e1, n, n, 1e-3, 1e-3, 20.)
    mag = W.planar_undulator()
    prec = W.traj_prec(10001)
    opt = W.drift_container(L, mode)
    st = SRWLStokes()
    st.allocate(1, nres, nres, _n_comp=1)
    st.mesh.eStart = st.mesh.eFin = e1
    st.mesh.xStart, st.mesh.xFin = -half_res, half_res
    st.mesh.yStart, st.mesh.yFin = -half_res, half_res
    return st, w0, mag, opt, prec


def test_normals_match_oracle_sampler():
    from srw_amd import srwlpy
    import ctypes as C
    from srw_amd import _structs  # noqa: F401
    g = srwlpy.normals(20251202, 7)
    assert len(g) == 6 and all(np.isfinite(g))


@pytest.mark.parametrize("mode,batch", [(2, 5), (0, 16)])
def test_multi_e_matches_oracle(mode, batch):
    from srw_amd import srwlpy
    print("case: multi-e 64x64, 16 electrons, mode", mode, "batch", batch, flush=True)
    st_g, w0, mag, opt, prec = _setup(mode=mode)
    st_o, _, _, _, _ = _setup(mode=mode)
    info = srwlpy.EmitPropMultiE(st_g, w0, mag, opt, prec, 16, _seed=99, _batch=batch)
    assert oracle.multi_e(st_o, w0, mag, opt, prec, 16, seed=99) == 0
    a, b = np.array(st_g.arS[:48 * 48]), np.array(st_o.arS[:48 * 48])
    err = rel_l2(a, b)
    print("rel-L2 %.3e electrons %d" % (err, info["electrons"]), flush=True)
    assert info["electrons"] == 16
    assert err < 2e-4


def test_rqmc_estimate_agrees_with_plain_mc():
    """Randomized quasi-Monte Carlo draws estimate the same intensity: the
    difference from a plain run is within the combined error estimates."""
    from srw_amd import srwlpy
    print("case: rqmc vs plain, 64x64, 256 electrons", flush=True)
    st_p, w0, mag, opt, prec = _setup()
    st_q, _, _, _, _ = _setup()
    ip = srwlpy.EmitPropMultiE(st_p, w0, mag, opt, prec, 512, _seed=5, _batch=128)
    iq = srwlpy.EmitPropMultiE(st_q, w0, mag, opt, prec, 256, _seed=6, _batch=128, _rqmc=1, _replicates=16)
    a, b = np.array(st_p.arS[:48 * 48]), np.array(st_q.arS[:48 * 48])
    diff = rel_l2(b, a)
    ep, eq = ip["rel_l2_est"], iq["rel_l2_est"]
    print("plain err %.4f (512) rqmc err %.4f (256, %d replicates) diff %.4f" %
          (ep, eq, iq["replicates"], diff), flush=True)
    assert iq["electrons"] == 256 and iq["replicates"] == 16
    assert 0 < eq and diff < 4 * np.hypot(ep, eq)


def test_error_estimates_match_true_errors():
    """The run's own error estimates (per-pixel sample variance for plain MC,
    the spread of 16 replicate averages for RQMC) against the true errors
    measured from an 8192-electron RQMC reference: each estimate is within a
    factor 2.5 of the true error (both are noisy at these counts)."""
    from srw_amd import srwlpy
    print("case: error estimates vs an 8192-electron reference, 64x64", flush=True)
    st_r, w0, mag, opt, prec = _setup()
    srwlpy.EmitPropMultiE(st_r, w0, mag, opt, prec, 8192, _seed=1, _batch=1024, _rqmc=1, _replicates=16)
    ref = np.array(st_r.arS[:48 * 48])
    for name, kw, n in (("plain", {}, 512), ("rqmc", {"_rqmc": 1, "_replicates": 16}, 256)):
        st, _, _, _, _ = _setup()
        info = srwlpy.EmitPropMultiE(st, w0, mag, opt, prec, n, _seed=7, _batch=256, **kw)
        true = rel_l2(np.array(st.arS[:48 * 48]), ref)
        est = info["rel_l2_est"]
        print("%s: %d electrons, estimated %.4f true %.4f" % (name, n, est, true), flush=True)
        assert est / 2.5 < true < est * 2.5


@pytest.mark.parametrize("beamline", ["aperture-lens-drift", "drift-aperture"])
def test_multi_e_with_lens_and_aperture_matches_oracle(beamline):
    """Multi-electron sum through a beamline with a slit, a thin lens and a
    drift (and one that ends on an aperture, so the intensity is taken after
    it), electron by electron against the oracle on the same draws."""
    from srw_amd import srwlpy
    from srw_amd.srwlib import SRWLOptA, SRWLOptC, SRWLOptD, SRWLOptL
    p0, p2 = [0, 0, 1.0, 0, 0, 1.0, 1.0, 1.0, 1.0], [0, 0, 1.0, 2, 0, 1.0, 1.0, 1.0, 1.0]
    if beamline == "aperture-lens-drift":
        mk = lambda: SRWLOptC([SRWLOptA('r', 'a', 1.4e-3, 1.0e-3), SRWLOptL(15., 15.), SRWLOptD(10.)], [p0, p0, p2])
    else:
        mk = lambda: SRWLOptC([SRWLOptD(30.), SRWLOptA('c', 'a', 3.0e-3, 0., 2e-4, 0.)], [p2, p0])
    print("case: multi-e 64x64, 16 electrons, beamline", beamline, flush=True)
    st_g, w0, mag, _, prec = _setup()
    st_o, _, _, _, _ = _setup()
    info = srwlpy.EmitPropMultiE(st_g, w0, mag, mk(), prec, 16, _seed=7, _batch=8)
    assert oracle.multi_e(st_o, w0, mag, mk(), prec, 16, seed=7) == 0
    a, b = np.array(st_g.arS[:48 * 48]), np.array(st_o.arS[:48 * 48])
    err = rel_l2(a, b)
    print("rel-L2 %.3e electrons %d" % (err, info["electrons"]), flush=True)
    assert info["electrons"] == 16 and np.max(b) > 0
    assert err < 2e-4
