"""GPU parity of the streamed download: a single request's field computed in
row bands (each band's SR kernels on alternating streams, its finalize and
FP64 passes, its device-to-host copy overlapping the next band) against the
same request in one pass.  Fields of 16 MB or more take the banded path by
default (the 201-energy workload); here the band count is forced on smaller
fields."""
import numpy as np
import pytest

from srw_amd import workloads as W
from tests.util import cfield, clone_wfr, rel_l2

pytestmark = pytest.mark.gpu


def _run(w, mag, prec, bands=0, cond=None):
    from srw_amd import srwlpy
    wg = clone_wfr(w)
    srwlpy.set_bands(bands)
    if cond is not None:
        srwlpy.set_cond_tol(cond)
    try:
        srwlpy.CalcElecFieldSR(wg, 0, mag, prec)
        fp64 = int(srwlpy.stats()["sr_fp64_points"])
    finally:
        srwlpy.set_bands(0)
        srwlpy.set_cond_tol(1e-4)
    return wg, fp64


@pytest.mark.parametrize("bands", [3, 8])
def test_banded_field_equals_one_pass(bands):
    """16 energies on 101 x 101 (energy tiles plus the single-energy kernel):
    fields, wavefront radius and moments of the banded pass against one pass
    (the partial sums of the s-range differ in number, so the FP64 sums agree
    to rounding)."""
    print("case: 101x101, 16 energies, %d bands vs one pass" % bands, flush=True)
    w, mag, prec = W.single_electron(n=101, ne=16, de_rel=0.05)
    w1, _ = _run(w, mag, prec)
    wb, _ = _run(w, mag, prec, bands=bands)
    for comp in ("arEx", "arEy"):
        e = rel_l2(cfield(getattr(wb, comp), wb), cfield(getattr(w1, comp), w1))
        print("%s rel-L2 %.3e" % (comp, e), flush=True)
        assert e < 1e-6
    assert wb.Rx == w1.Rx and wb.xc == w1.xc
    for name in ("arMomX", "arMomY"):
        a = np.array(getattr(wb, name)).reshape(-1, 11)[:w.mesh.ne]
        b = np.array(getattr(w1, name)).reshape(-1, 11)[:w.mesh.ne]
        assert np.allclose(a, b, rtol=1e-5, atol=1e-5 * np.abs(b).max())


def test_banded_fp64_long_lists_equal_one_pass():
    """Every point sent to the FP64 pass (condition tolerance 1e-12), so that
    each band's list is long (one lane per point, sums and terminating terms
    in one kernel) while the one-pass list runs with its sums in a kernel of
    their own: the same FP64 arithmetic per point, so the same bits."""
    print("case: 201x201, 16 energies, every point in FP64, 4 bands vs one pass", flush=True)
    w, mag, prec = W.single_electron(n=201, ne=16, de_rel=0.05)
    w1, n1 = _run(w, mag, prec, cond=1e-12)
    wb, nb = _run(w, mag, prec, bands=4, cond=1e-12)
    print("FP64 points: one pass %d, banded %d" % (n1, nb), flush=True)
    assert n1 == nb and n1 >= 4 * 32768
    assert np.array_equal(np.frombuffer(wb.arEx, np.float32), np.frombuffer(w1.arEx, np.float32))
    assert np.array_equal(np.frombuffer(wb.arEy, np.float32), np.frombuffer(w1.arEy, np.float32))


def _stream(w, mag, prec, on, bands=0, cond=None):
    from srw_amd import srwlpy
    srwlpy.set_stream_rows(on)
    try:
        wg, fp64 = _run(w, mag, prec, bands=bands, cond=cond)
        streamed = srwlpy.last_streamed()
    finally:
        srwlpy.set_stream_rows(True)
    return wg, fp64, streamed


def _same_bits(a, b):
    for comp in ("arEx", "arEy"):
        assert np.array_equal(np.frombuffer(getattr(a, comp), np.float32), np.frombuffer(getattr(b, comp), np.float32)), comp


@pytest.mark.parametrize("bands", [0, 3])
def test_streamed_rows_equal_one_pass(bands):
    """201 x 201 x 201 (mirror-symmetric: 101 x 101 computed points, 25 energy
    tiles and one single-energy item): the streamed download (tile blocks that
    finish their own points and flag themselves done, bands copied while the
    kernel runs, FP64 stores patched in afterwards) against one pass with one
    download.  Both sum the s-range in one piece, so the fields are the same
    bits."""
    print("case: 201x201x201 streamed rows, bands=%d, vs one pass" % bands, flush=True)
    w, mag, prec = W.single_electron(n=201, ne=201, de_rel=0.05)
    w1, n1, s1 = _stream(w, mag, prec, True, bands=1)
    ws, ns, ss = _stream(w, mag, prec, True, bands=bands)
    print("streamed %s (one pass %s), FP64 points %d / %d" % (ss, s1, ns, n1), flush=True)
    assert ss and not s1
    assert ns == n1 and ns > 0
    _same_bits(ws, w1)
    for name in ("arMomX", "arMomY"):
        assert np.array_equal(np.array(getattr(ws, name)), np.array(getattr(w1, name)))
    assert ws.Rx == w1.Rx and ws.xc == w1.xc


def test_streamed_rows_close_to_banded():
    """The streamed download against the banded one (stream off): the banded
    pass splits the s-range into more partial sums, so the two agree to
    rounding."""
    print("case: 201x201x201 streamed rows vs row bands", flush=True)
    w, mag, prec = W.single_electron(n=201, ne=201, de_rel=0.05)
    wb, _, sb = _stream(w, mag, prec, False)
    ws, _, ss = _stream(w, mag, prec, True)
    assert ss and not sb
    for comp in ("arEx", "arEy"):
        e = rel_l2(cfield(getattr(ws, comp), ws), cfield(getattr(wb, comp), wb))
        print("%s rel-L2 %.3e" % (comp, e), flush=True)
        assert e < 1e-6


def test_streamed_rows_patch_overflow_falls_back():
    """Every point sent to the FP64 pass (condition tolerance 1e-12): its stores
    outnumber the patch area, so the call downloads the finished field in one
    piece instead (the patch area holds 2^18 stores), with the same bits as
    one pass."""
    print("case: 201x201x201 every point in FP64, streamed rows fall back", flush=True)
    w, mag, prec = W.single_electron(n=201, ne=201, de_rel=0.05)
    w1, n1, _ = _stream(w, mag, prec, True, bands=1, cond=1e-12)
    ws, ns, ss = _stream(w, mag, prec, True, cond=1e-12)
    print("FP64 points %d / %d, streamed %s" % (ns, n1, ss), flush=True)
    assert not ss and ns == n1 and ns > (1 << 18)
    _same_bits(ws, w1)
