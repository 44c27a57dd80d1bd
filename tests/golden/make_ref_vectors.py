"""Writes the reference golden vectors tests/golden/ref_sr.npz and
tests/golden/ref_radsub.npz by running the reference SRW sources (built by
oracle/ref.mk into oracle/_ref/ref_run) on the cases of tests/golden/cases.py.

    python tests/golden/make_ref_vectors.py

Only this container has the reference sources; the committed .npz files are
what the tests read elsewhere.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests import ref_lib  # noqa: E402
from tests.golden import cases  # noqa: E402


def main():
    ref_lib.build()
    assert ref_lib.available(), "reference runner not built"
    out = {}
    for i, d in enumerate(cases.SR_CASES):
        w, mag, prec = cases.sr_case(d)
        rc, _ = ref_lib.calc_field(w, mag, prec)
        key = d["name"]
        out[key + "/ex"] = np.frombuffer(w.arEx, dtype=np.float32).copy()
        out[key + "/ey"] = np.frombuffer(w.arEy, dtype=np.float32).copy()
        out[key + "/geo"] = np.array([w.Rx, w.Ry, w.dRx, w.dRy, w.xc, w.yc])
        out[key + "/mom"] = np.concatenate([np.array(w.arMomX), np.array(w.arMomY)])
        out[key + "/rc"] = np.array([rc], dtype=np.int32)
        print("%-28s rc %6d Rx %.6f xc %.3e |Ex| %.3e" % (key, rc, w.Rx, w.xc, np.abs(out[key + "/ex"]).max()))
    out["_cases"] = np.frombuffer(json.dumps(cases.SR_CASES).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "ref_sr.npz"), **out)

    rs = {}
    for d in cases.RADSUB_CASES:
        w = cases.radsub_wavefront(d)
        effRx, effRz, mx, my = ref_lib.rad_to_sub(w, d["L"], d["sub"])
        rs[d["name"] + "/eff"] = np.array([effRx, effRz])
        rs[d["name"] + "/mom"] = np.concatenate([mx, my])
        print("%-28s effRx %.9g effRz %.9g" % (d["name"], effRx, effRz))
    rs["_cases"] = np.frombuffer(json.dumps(cases.RADSUB_CASES).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "ref_radsub.npz"), **rs)

    # extraction: SHA-256 of each output's bytes (the checks are bit-exact),
    # from a prefill seeded by the case index (the averaging and sum modes read it)
    import hashlib
    w, (e, x, y) = cases.int_wavefront()
    digests, rcs = [], []
    for k, (it, dep, pol, meth) in enumerate(cases.int_cases()):
        n = cases.int_len(w, dep, pol, it)
        pre = cases.int_prefill(k, n, it)
        rc, out = ref_lib.calc_int(w, pol, it, dep, e, x, y, n, meth=meth, prefill=pre)
        digests.append(hashlib.sha256(out.tobytes()).hexdigest())
        rcs.append(rc)
    with open(os.path.join(HERE, "ref_int.json"), "w") as f:
        json.dump({"cases": [[it, dep, pol, meth] for (it, dep, pol, meth) in cases.int_cases()],
                   "rc": rcs, "sha256": digests}, f, indent=0)
    print("extraction cases:", len(cases.int_cases()))


if __name__ == "__main__":
    main()
