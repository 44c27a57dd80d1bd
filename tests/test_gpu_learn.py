"""GPU kernels of the learned proposal against the FP64 numpy oracle, and
the effect of the trained proposal on the multi-electron estimator."""
import numpy as np
import pytest

from oracle import proposal_oracle as PO
from srw_amd import workloads as W
from srw_amd.srwlib import SRWLStokes
from tests.util import rel_l2

pytestmark = pytest.mark.gpu


def test_proposal_gradient_matches_oracle():
    from srw_amd import srwlpy
    rng = np.random.default_rng(11)
    P = srwlpy.proposal_init()
    P = (P + rng.standard_normal(P.size).astype(np.float32) * 0.05).astype(np.float32)
    c = rng.standard_normal(6).astype(np.float32)
    g = (rng.standard_normal((200, 6)) * 1.2).astype(np.float32)
    wh = rng.random(200).astype(np.float32)
    wh /= wh.sum()
    G, out, loss = srwlpy.proposal_grad(P, c, g, wh, 0.25)
    lo, Go, oo = PO.loss_grad(P.astype(np.float64), c.astype(np.float64), g.astype(np.float64),
                              wh.astype(np.float64), 0.25)
    print("out rel %.2e grad rel %.2e loss %.6f vs %.6f" % (rel_l2(out, oo), rel_l2(G, Go), loss, lo), flush=True)
    assert rel_l2(out, oo) < 1e-5
    assert abs(loss - lo) < 1e-4 * abs(lo)
    assert rel_l2(G, Go) < 1e-3


def _setup(n=64, nres=48):
    beam = W.beam_for_multi_e()
    e1 = W.first_harmonic_eV(beam.partStatMom1.gamma, 1.5, 0.02)
    w0 = W.wavefront(beam, 1, e1, e1, n, n, 1e-3, 1e-3, 20.)
    st = SRWLStokes()
    st.allocate(1, nres, nres, _n_comp=1)
    st.mesh.eStart = st.mesh.eFin = e1
    st.mesh.xStart, st.mesh.xFin = -2.4e-3, 2.4e-3
    st.mesh.yStart, st.mesh.yFin = -2.4e-3, 2.4e-3
    return st, w0, W.planar_undulator(), W.drift_container(30., 2), W.traj_prec(10001)


def test_learned_proposal_is_unbiased_and_reports_gain():
    from srw_amd import srwlpy
    st_p, w0, mag, opt, prec = _setup()
    st_q, _, _, _, _ = _setup()
    ip = srwlpy.EmitPropMultiE(st_p, w0, mag, opt, prec, 512, _seed=5, _batch=64, _proposal=0)
    iq = srwlpy.EmitPropMultiE(st_q, w0, mag, opt, prec, 512, _seed=6, _batch=64, _proposal=1, _train=4)
    a, b = np.array(st_p.arS[:48 * 48]), np.array(st_q.arS[:48 * 48])
    diff = rel_l2(b, a)
    ep, eq = ip["rel_l2_est"], iq["rel_l2_est"]
    print("plain err %.4f learned err %.4f (ess %.1f) diff %.4f gain(var) %.2f" %
          (ep, eq, iq["ess"], diff, (ep / eq) ** 2), flush=True)
    # both estimate the same mean: the difference is within the combined statistical error
    assert diff < 4 * np.hypot(ep, eq)
    assert np.isfinite(iq["train_loss"])


def _ref_intensity(n_el=8192):
    from srw_amd import srwlpy
    st, w0, mag, opt, prec = _setup()
    srwlpy.EmitPropMultiE(st, w0, mag, opt, prec, n_el, _seed=99, _batch=512, _rqmc=1, _replicates=16)
    return np.array(st.arS[:48 * 48], dtype=np.float64)


def test_learned_proposal_through_rqmc_points_is_unbiased_and_never_needs_more_electrons_than_plain_mc():
    """The learned path (proposal sampled through randomized Sobol' points,
    used only where the electrons so far predict a gain) against plain
    Gaussian Monte Carlo, both run to a 3 % estimated rel-L2 on four seeds:
    every learned run lands within twice the target of an 8192-electron
    reference (its error estimate is honest) and never needs more electrons
    than the plain run of the same seed."""
    from srw_amd import srwlpy
    ref = _ref_intensity()
    target = 0.03
    for seed in (1, 2, 3, 4):
        st_p, w0, mag, opt, prec = _setup()
        st_q, _, _, _, _ = _setup()
        ip = srwlpy.EmitPropMultiE(st_p, w0, mag, opt, prec, 4096, _seed=seed, _batch=64, _proposal=0,
                                   _target=target)
        iq = srwlpy.EmitPropMultiE(st_q, w0, mag, opt, prec, 4096, _seed=seed, _batch=64, _proposal=1,
                                   _target=target, _train=8, _rqmc=1, _replicates=16)
        tp = rel_l2(np.array(st_p.arS[:48 * 48]), ref)
        tq = rel_l2(np.array(st_q.arS[:48 * 48]), ref)
        print("seed %d: plain %d electrons (est %.4f, true %.4f); learned %d electrons (est %.4f, true %.4f, "
              "mixture batches %d, predicted ratio %.3f)" % (seed, ip["electrons"], ip["rel_l2_est"], tp,
                                                             iq["electrons"], iq["rel_l2_est"], tq,
                                                             iq["mixture_batches"], iq["predicted_ratio"]),
              flush=True)
        assert tq < 2 * target
        assert iq["electrons"] <= ip["electrons"]
