"""srwlCalcElecFieldSRNoAlloc (the allocation-free variant of
srwlCalcElecFieldSR over a fixed file-scope scratch array) against
srwlCalcElecFieldSR through the C ABI: two calls in a row, the second with
a smaller request, both equal to the allocating entry point's field; the
size limits and the missing-initialisation case are refused."""
import ctypes as C

import numpy as np
import pytest

from srw_amd import _marshal as M
from srw_amd import _structs as S
from srw_amd import workloads as W
from tests.util import cfield, clone_wfr, rel_l2

pytestmark = pytest.mark.gpu

SRWL_AMD_ERR_NOALLOC_INIT = 23303
SRWL_AMD_ERR_NOALLOC_SIZE = 23304


def _lib():
    from srw_amd import srwlpy
    lib = srwlpy.lib
    P = C.POINTER
    lib.srwlCalcElecFieldSRNoAlloc.argtypes = [P(S.SRWLWfr), P(S.SRWLPrtTrj), P(S.SRWLMagFldC), P(C.c_double),
                                               C.c_int]
    lib.srwlAmdNoAllocInit.argtypes = []
    return srwlpy, lib


def _noalloc(lib, w, mag, prec):
    keep = M.Keep()
    cw = M.wfr(w, keep)
    cm = M.mag_cnt(mag, keep)
    cp = M.dvec(prec, keep)
    rc = lib.srwlCalcElecFieldSRNoAlloc(C.byref(cw), None, C.byref(cm), cp, len(prec))
    M.wfr_back(cw, w)
    return rc


def test_no_alloc_matches_allocating_entry_point_twice():
    srwlpy, lib = _lib()
    assert lib.srwlAmdNoAllocInit() == 0
    # the largest request the scratch takes (256 x 256 = 65536 points), then
    # smaller ones: each call's result must not depend on the previous one's
    for n, ne, de in ((256, 1, 0.), (101, 1, 0.), (41, 3, 0.05)):
        print("case: no-alloc field %dx%d, %d energies" % (n, n, ne), flush=True)
        w, mag, prec = W.single_electron(n=n, ne=ne, de_rel=de)
        wa, wn = clone_wfr(w), clone_wfr(w)
        srwlpy.CalcElecFieldSR(wa, 0, mag, prec)
        assert _noalloc(lib, wn, mag, prec) == 0
        for a, b in ((wa.arEx, wn.arEx), (wa.arEy, wn.arEy)):
            ea, eb = cfield(a, wa), cfield(b, wn)
            assert rel_l2(eb, ea) < 1e-6, rel_l2(eb, ea)
        assert (wn.Rx, wn.xc) == (wa.Rx, wa.xc)
        np.testing.assert_allclose(np.array(wn.arMomX), np.array(wa.arMomX), rtol=1e-6, atol=1e-30)


def test_no_alloc_refuses_requests_beyond_its_scratch():
    srwlpy, lib = _lib()
    assert lib.srwlAmdNoAllocInit() == 0
    w, mag, prec = W.single_electron(n=257, ne=1)            # 66049 points > 65536
    assert _noalloc(lib, w, mag, prec) == SRWL_AMD_ERR_NOALLOC_SIZE
    w, mag, prec = W.single_electron(n=21, ne=1)
    prec[0] = 1                                               # method 1
    assert _noalloc(lib, w, mag, prec) == SRWL_AMD_ERR_NOALLOC_SIZE
    w, mag, prec = W.single_electron(n=21, ne=1, np_traj=30001)
    assert _noalloc(lib, w, mag, prec) == SRWL_AMD_ERR_NOALLOC_SIZE
    # the library stays usable after a refused call
    w, mag, prec = W.single_electron(n=21, ne=1)
    wa = clone_wfr(w)
    srwlpy.CalcElecFieldSR(wa, 0, mag, prec)
    assert _noalloc(lib, w, mag, prec) == 0
    assert rel_l2(cfield(w.arEx, w), cfield(wa.arEx, wa)) < 1e-6
