"""CPU tests of the host side: the C-ABI library exports every declared
entry point, host-only entry points behave like the reference's, and the
Python mirror marshals the reference's structures byte-for-byte."""
import ctypes as C
import glob
import os
import re

import pytest

from srw_amd import _structs as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "srw_amd", "libsrw_amd.so")


def _declared():
    names = []
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        txt = open(h).read()
        names += re.findall(r"SRWL_EXPORT\s+[A-Za-z_][\w\s\*]*?\b(srwl\w+)\s*\(", txt)
    return names


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libsrw_amd.so not built")
    return C.CDLL(LIB)


def test_every_declared_entry_point_is_exported(lib):
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_reference_entry_points_are_declared():
    # the reference's C API functions on the emit-and-propagate path (srwlib.h)
    need = {"srwlCalcPartTraj", "srwlCalcElecFieldSR", "srwlCalcIntFromElecField", "srwlPropagElecField",
            "srwlUtiFFT", "srwlUtiGPUProc", "srwlUtiVerNo", "srwlUtiGetErrText", "srwlUtiSetWfrModifFunc",
            "srwlUtiSetAllocArrayFunc", "srwlUtiSetProgrIndFunc"}
    assert need <= set(_declared())


def test_every_reference_entry_point_links_and_the_rest_report_unsupported(lib):
    # every entry point of the reference C API (srwlib.h:542-998) is exported,
    # so that a binding written against it links; those outside this
    # library's path return SRWL_AMD_ERR_UNSUPPORTED (23301) and touch nothing
    ref_api = {"srwlCalcElecFieldGaussian", "srwlCalcElecFieldPointSrc", "srwlCalcElecFieldSR",
               "srwlCalcIntFromElecField", "srwlCalcMagFld", "srwlCalcPartTraj", "srwlCalcPartTrajFromKickMatr",
               "srwlCalcPowDenSR", "srwlCalcStokesUR", "srwlCalcTransm", "srwlProcElecField", "srwlPropagElecField",
               "srwlPropagRadMultiE", "srwlResizeElecField", "srwlResizeElecFieldMesh", "srwlSetRepresElecField",
               "srwlUtiConvWithGaussian", "srwlUtiFFT", "srwlUtiGPUProc", "srwlUtiGetErrText", "srwlUtiIntInf",
               "srwlUtiIntProc", "srwlUtiUndFindMagFldInterpInds", "srwlUtiUndFromMagFldTab", "srwlUtiVerNo",
               "srwlUtiSetWfrModifFunc", "srwlUtiSetAllocArrayFunc", "srwlUtiSetProgrIndFunc"}
    assert ref_api <= set(_declared())
    for n in ref_api:
        assert hasattr(lib, n), n
    nulls = {"srwlCalcMagFld": 3, "srwlCalcPartTrajFromKickMatr": 4, "srwlCalcElecFieldGaussian": 3,
             "srwlCalcElecFieldPointSrc": 3, "srwlCalcStokesUR": 4, "srwlCalcPowDenSR": 5, "srwlProcElecField": 3,
             "srwlSetRepresElecField": 2, "srwlPropagRadMultiE": 5, "srwlCalcTransm": 6,
             "srwlUtiConvWithGaussian": 5, "srwlUtiIntInf": 6, "srwlUtiIntProc": 8, "srwlUtiUndFromMagFldTab": 3,
             "srwlUtiUndFindMagFldInterpInds": 6}
    for n, k in nulls.items():
        f = getattr(lib, n)
        f.restype = C.c_int
        f.argtypes = [C.c_void_p] * k
        assert f(*([None] * k)) == 23301, n


def test_error_texts_follow_reference_codes(lib):
    buf = C.create_string_buffer(2048)
    lib.srwlUtiGetErrText.argtypes = [C.c_char_p, C.c_int]
    # SRWL_ERR_IN_FFT = 23040, SRWL_INCORRECT_PARAM_FOR_SR_COMP = 23164, warning -12988
    for code in (23040, 23164, 23171, -13000 + 12):
        buf.value = b""
        lib.srwlUtiGetErrText(buf, code)
        assert len(buf.value) > 5, code
    buf.value = b""
    lib.srwlUtiGetErrText(buf, 23300)
    assert b"GPU" in buf.value or b"device" in buf.value.lower()


def test_version_string(lib):
    buf = C.create_string_buffer(64)
    lib.srwlUtiVerNo.argtypes = [C.c_char_p, C.c_int]
    lib.srwlUtiVerNo(buf, 0)
    assert re.match(rb"\d+\.\d+", buf.value)


def test_struct_sizes_match_reference_layout():
    # sizes of the reference's structures on LP64 (srwlib.h:75-532)
    want = {"SRWLParticle": 64, "SRWLPartBeam": 248, "SRWLMagFldH": 40, "SRWLMagFldU": 32,
            "SRWLMagFldC": 112, "SRWLPrtTrj": 160, "SRWLRadMesh": 160, "SRWLWfr": 536, "SRWLStokes": 208,
            "SRWLOptD": 16, "SRWLOptC": 48}
    for k, v in want.items():
        assert C.sizeof(getattr(S, k)) == v, k


def test_marshal_round_trip_of_wavefront():
    from srw_amd import _marshal as M
    from srw_amd import workloads as W
    w, mag, prec = W.single_electron(n=6, ne=3, de_rel=0.01)
    w.Rx, w.dRx, w.xc = 20., 0.1, 1e-6
    keep = M.Keep()
    c = M.wfr(w, keep)
    assert c.mesh.nx == 6 and c.mesh.ne == 3 and c.mesh.eFin == w.mesh.eFin
    assert c.Rx == 20. and c.dRx == 0.1 and c.xc == 1e-6
    assert c.partBeam.partStatMom1.gamma == w.partBeam.partStatMom1.gamma
    c.mesh.xStart = -5e-3
    c.Ry = 7.
    M.wfr_back(c, w)
    assert w.mesh.xStart == -5e-3 and w.Ry == 7.
    cm = M.mag_cnt(mag, keep)
    assert cm.nElem == 1 and cm.arMagFldTypes[0] == ord("u")


def test_product_path_refuses_to_run_without_the_library(monkeypatch, tmp_path):
    # the Python mirror fails loudly instead of falling back to a CPU path
    import srw_amd.srwlpy as P
    monkeypatch.setattr(P, "_LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(ImportError):
        P._load()
    for f in glob.glob(os.path.join(ROOT, "srw_amd", "*.py")):
        src = open(f).read()
        assert "import oracle" not in src and "oracle_lib" not in src and "liboracle" not in src, f
