"""The oracle against the reference SRW sources' own results on identical
inputs (tests/golden/ref_*.npz, written by tests/golden/make_ref_vectors.py
from the reference compiled by oracle/ref.mk).

These pin the parts of the oracle that mathematics alone does not fix:
* FindAverageDistanceToSource (srradstr.cpp:2310-2501): Rx, Ry, their
  errors and the source-point position xc, yc, on centred and offset
  electrons, near and far observation planes, limited integration ranges, a
  zero field, two undulators, a horizontal field;
* the terminating terms (ComputeNormalResidual, srradint.cpp:398-871): the
  fields with none, both, the left only and the right only, and the
  near-field left residual (:589-652) at photon energies where the
  far-field expansion is rejected;
* EstimateWfrRadToSub / EstimateWfrRadToSub2 (sroptdrf.cpp:1085-1381), the
  radii of the arProp[3] = 1 and 2 drifts.
Where the runner is built (this container only), further tests run the
reference live on randomised inputs.
"""
import json
import os
import random

import numpy as np
import pytest

from tests import oracle_lib as oracle
from tests import ref_lib
from tests.golden import cases
from tests.golden.compare import FIELD_TOL, RS, SR, _cfield, check_against_golden
from tests.util import cfield, clone_wfr, rel_l2


@pytest.mark.parametrize("case", cases.SR_CASES, ids=[c["name"] for c in cases.SR_CASES])
def test_oracle_sr_vs_reference(case):
    w, mag, prec = cases.sr_case(case)
    rc = oracle.calc_field(w, mag, prec, nthreads=1)
    check_against_golden(w, rc, case["name"])


def test_golden_case_list_matches():
    stored = json.loads(bytes(SR["_cases"]).decode())
    assert stored == json.loads(json.dumps(cases.SR_CASES))
    stored = json.loads(bytes(RS["_cases"]).decode())
    assert stored == json.loads(json.dumps(cases.RADSUB_CASES))


def test_near_field_cases_use_the_near_field_residual():
    """The near-field golden cases really exercise the left near-field branch:
    without it (left residual zero) the field differs from the reference's."""
    for name in ("near_field_edge_0.5eV", "near_field_edge_1eV_2m"):
        case = next(c for c in cases.SR_CASES if c["name"] == name)
        w, mag, prec = cases.sr_case(case)
        prec[5] = 3   # right edge only: no left residual at all
        oracle.calc_field(w, mag, prec, nthreads=1)
        rex = _cfield(SR[name + "/ex"], w)
        assert rel_l2(cfield(w.arEx, w), rex) > 1e-2


@pytest.mark.parametrize("case", cases.RADSUB_CASES, ids=[c["name"] for c in cases.RADSUB_CASES])
def test_oracle_drift_radius_vs_reference(case):
    w = cases.radsub_wavefront(case)
    effRx, effRz, mx, my = oracle.rad_to_sub(w, case["L"], case["sub"])
    ref = RS[case["name"] + "/eff"]
    assert effRx == pytest.approx(ref[0], rel=1e-6)
    assert effRz == pytest.approx(ref[1], rel=1e-6)


# ---- live reference (only where the runner is built) -----------------------

live = pytest.mark.skipif(not ref_lib.available(), reason="reference runner oracle/_ref/ref_run not built")


@live
def test_golden_vectors_match_the_live_reference():
    for case in cases.SR_CASES[:3] + [c for c in cases.SR_CASES if c["name"].startswith("near_field")][:2]:
        w, mag, prec = cases.sr_case(case)
        rc, _ = ref_lib.calc_field(w, mag, prec)
        assert np.array_equal(np.frombuffer(w.arEx, dtype=np.float32), SR[case["name"] + "/ex"])
        assert rc == int(SR[case["name"] + "/rc"][0])


@live
@pytest.mark.parametrize("seed", range(6))
def test_oracle_vs_live_reference_random(seed):
    """Randomised electron offsets, observation distance, mesh window, photon
    energy (from the near-field regime up to the third harmonic), s-limits
    and terminating terms, oracle against the reference run live."""
    rng = random.Random(1000 + seed)
    case = dict(name="random", x0=rng.uniform(-1e-4, 1e-4), xp0=rng.uniform(-3e-5, 3e-5),
                y0=rng.uniform(-5e-5, 5e-5), yp0=rng.uniform(-1e-5, 1e-5), z=rng.choice([3., 10., 20., 40.]),
                nx=rng.randint(3, 9), ny=rng.randint(3, 9), termin=rng.choice([0, 1, 2, 3]))
    half = 1e-3 * case["z"] / 20.
    x0 = rng.uniform(-half, 0.5 * half)
    case["x"] = (x0, x0 + rng.uniform(0.2, 1.5) * half)
    case["y"] = (-half, rng.uniform(-0.5, 1.) * half)
    case["e"] = rng.choice([0.8, 3., 50., 1000., 2010.99, 4000., 6030.])
    if rng.random() < 0.3:
        case["slim"] = (rng.uniform(-1.0, -0.3), rng.uniform(0.3, 1.0))
    w, mag, prec = cases.sr_case(case)
    wr = clone_wfr(w)
    rc_o = oracle.calc_field(w, mag, prec, nthreads=1)
    rc_r, _ = ref_lib.calc_field(wr, mag, prec)
    assert rc_o == rc_r
    assert rel_l2(cfield(w.arEx, w), cfield(wr.arEx, wr)) < FIELD_TOL
    assert rel_l2(cfield(w.arEy, w), cfield(wr.arEy, wr)) < FIELD_TOL or \
        np.linalg.norm(cfield(wr.arEy, wr)) < 1e-6 * np.linalg.norm(cfield(wr.arEx, wr))
    for a, b in zip((w.Rx, w.Ry, w.dRx, w.dRy), (wr.Rx, wr.Ry, wr.dRx, wr.dRy)):
        assert a == pytest.approx(b, rel=1e-12)
    assert w.xc == pytest.approx(wr.xc, rel=1e-12, abs=1e-15)
    assert w.yc == pytest.approx(wr.yc, rel=1e-12, abs=1e-15)
