"""srwlpy-compatible Python entry points over libsrw_amd.so.

Method names and argument order follow the reference binding
(cpp/src/clients/python/srwlpy.cpp:5895-5917): CalcElecFieldSR (:4622),
CalcPartTraj, CalcIntFromElecField (:4873), PropagElecField (:5214), UtiFFT.
Errors raise RuntimeError with the library's error text, like the
reference's ProcRes (srwlpy.cpp:148); warnings are returned silently.

The compute path is the HIP library only: importing this module without the
built library raises ImportError.
"""
import ctypes as C
import os

from . import _marshal as M
from . import _structs as S

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsrw_amd.so")


def _load():
    if not os.path.exists(_LIB_PATH):
        raise ImportError("libsrw_amd.so is not built (run python -m srw_amd.build)")
    lib = C.CDLL(_LIB_PATH)
    P = C.POINTER
    lib.srwlCalcElecFieldSR.argtypes = [P(S.SRWLWfr), P(S.SRWLPrtTrj), P(S.SRWLMagFldC), P(C.c_double), C.c_int]
    lib.srwlCalcPartTraj.argtypes = [P(S.SRWLPrtTrj), P(S.SRWLMagFldC), P(C.c_double)]
    lib.srwlCalcElecFieldSRBatch.argtypes = [P(P(S.SRWLWfr)), P(P(S.SRWLMagFldC)), C.c_int, P(C.c_double), C.c_int]
    lib.srwlUtiGetErrText.argtypes = [C.c_char_p, C.c_int]
    lib.srwlUtiGPUProc.argtypes = [C.c_int, P(C.c_double)]
    for name in ("srwlPropagElecField",):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = [P(S.SRWLWfr), P(S.SRWLOptC), C.c_int, P(C.c_char_p), P(S.SRWLRadMesh),
                                           P(C.c_char_p), P(C.c_double)]
    if hasattr(lib, "srwlCalcIntFromElecField"):
        lib.srwlCalcIntFromElecField.argtypes = [C.c_void_p, P(S.SRWLWfr), C.c_byte, C.c_byte, C.c_byte,
                                                 C.c_double, C.c_double, C.c_double, P(C.c_double),
                                                 C.c_void_p, P(C.c_double)]
    if hasattr(lib, "srwlUtiFFT"):
        lib.srwlUtiFFT.argtypes = [C.c_void_p, C.c_char, P(C.c_double), C.c_int, C.c_int, P(C.c_double)]
    if hasattr(lib, "srwlEmitPropMultiE"):
        lib.srwlEmitPropMultiE.argtypes = [P(S.SRWLStokes), P(S.SRWLWfr), P(S.SRWLMagFldC), P(S.SRWLOptC),
                                           P(C.c_double), C.c_int, C.c_longlong, P(C.c_double), P(C.c_double)]
    lib.srwlResizeElecField.argtypes = [P(S.SRWLWfr), C.c_char, P(C.c_double)]
    lib.srwlResizeElecFieldMesh.argtypes = [P(S.SRWLWfr), P(S.SRWLRadMesh), P(C.c_double)]
    lib.srwlAmdGetStats.argtypes = [P(C.c_double), C.c_int]
    lib.srwlAmdSetTiming.argtypes = [C.c_int]
    lib.srwlAmdSetKeepOnDevice.argtypes = [C.c_int]
    lib.srwlAmdNormals.argtypes = [C.c_ulonglong, C.c_longlong, P(C.c_double)]
    lib.srwlAmdLastAutoLevels.argtypes = [P(C.c_int), C.c_longlong]
    lib.srwlAmdLastAutoLevels.restype = C.c_longlong
    return lib


lib = _load()
LIB_PATH = _LIB_PATH

# Wavefront arrays are reallocated by the library through this callback when
# a call changes the mesh size, as the reference binding's ModifySRWLWfr does
# (srwlpy.cpp:4141-4420, registered at module init :5934-5941).  The C
# structure passed to a call is mapped back to its Python wavefront here.
_WFR_MODIF = C.CFUNCTYPE(C.c_int, C.c_int, C.POINTER(S.SRWLWfr), C.c_char)
_active_wfr = {}


def _modify_wfr(action, pw, pol):
    try:
        w = _active_wfr.get(C.addressof(pw.contents))
        if w is None:
            return -1
        c = pw.contents
        if action in (2, 12):
            from array import array
            n = 2 * c.mesh.ne * c.mesh.nx * c.mesh.ny
            p = pol if isinstance(pol, int) else (ord(pol) if pol else 0)
            if p in (0, ord('x'), ord('X')):
                w.arEx = array('f', bytes(4 * n))
                c.arEx = M._addr(w.arEx)
            if p in (0, ord('y'), ord('Y'), ord('z'), ord('Z')):
                w.arEy = array('f', bytes(4 * n))
                c.arEy = M._addr(w.arEy)
            M.mesh_back(c.mesh, w.mesh)
            return 0
        if action == 0:
            return 0
        return -1
    except Exception:
        return -1


_wfr_modif_cb = _WFR_MODIF(_modify_wfr)
lib.srwlUtiSetWfrModifFunc.argtypes = [_WFR_MODIF]
lib.srwlUtiSetWfrModifFunc(_wfr_modif_cb)


def err_text(code):
    buf = C.create_string_buffer(2048)
    lib.srwlUtiGetErrText(buf, int(code))
    return buf.value.decode()


# Return code of the last library call: 0, or the first warning (< 0) of a
# call that succeeded with warnings (the reference's srwlpy drops warnings).
last_code = 0


def _check(code):
    global last_code
    last_code = int(code)
    if code > 0:
        raise RuntimeError(err_text(code))
    return code


def CalcPartTraj(_prt_trj, _mag, _prec_par=None):
    keep = M.Keep()
    ct = M.prt_trj(_prt_trj, keep)
    cm = M.mag_cnt(_mag, keep)
    _check(lib.srwlCalcPartTraj(C.byref(ct), C.byref(cm), None))
    M.prt_trj_back(ct, _prt_trj)
    return _prt_trj


def CalcElecFieldSR(_wfr, _trj, _mag, _prec_par):
    """Near-field SR electric field (reference srwlpy.cpp:4622-4680)."""
    keep = M.Keep()
    cw = M.wfr(_wfr, keep)
    ct = None
    if _trj is not None and _trj != 0 and getattr(_trj, "arX", None) is not None:
        ct = C.byref(M.prt_trj(_trj, keep))
    cm = C.byref(M.mag_cnt(_mag, keep)) if (_mag is not None and _mag != 0) else None
    prec = M.dvec(_prec_par, keep)
    # precPar[6] > 0 may resize the mesh through the modification callback
    _active_wfr[C.addressof(cw)] = _wfr
    try:
        rc = lib.srwlCalcElecFieldSR(C.byref(cw), ct, cm, prec, len(_prec_par))
    finally:
        _active_wfr.pop(C.addressof(cw), None)
    _check(rc)
    M.wfr_back(cw, _wfr)
    return _wfr


def CalcElecFieldSRBatch(_arWfr, _arMag, _prec_par):
    """Batch of independent single-electron fields in one device pass."""
    keep = M.Keep()
    n = len(_arWfr)
    cws = [M.wfr(w, keep) for w in _arWfr]
    cms = [M.mag_cnt(m, keep) for m in _arMag]
    pw = (C.POINTER(S.SRWLWfr) * n)(*[C.pointer(c) for c in cws])
    pm = (C.POINTER(S.SRWLMagFldC) * n)(*[C.pointer(c) for c in cms])
    prec = M.dvec(_prec_par, keep)
    _check(lib.srwlCalcElecFieldSRBatch(pw, pm, n, prec, len(_prec_par)))
    for c, w in zip(cws, _arWfr):
        M.wfr_back(c, w)
    return _arWfr


def stats():
    out = (C.c_double * 16)()
    lib.srwlAmdGetStats(out, 16)
    keys = ("sr_ms", "sr_evals", "sr_blocks", "sr_split", "fft_ms", "fft_bytes", "sum_sr_ms", "sum_sr_evals",
            "n_sr", "sum_fft_ms", "sum_fft_bytes", "n_fft", "resize_ms", "resize_bytes", "prop_device_ms",
            "sr_fp64_points")
    return dict(zip(keys, list(out)))


def set_timing(on=True):
    _check(lib.srwlAmdSetTiming(1 if on else 0))


def reset_stats():
    lib.srwlAmdResetStats()


def set_keep_on_device(on=True):
    lib.srwlAmdSetKeepOnDevice(1 if on else 0)


def set_traj_serial(on=True):
    lib.srwlAmdSetTrajSerial(1 if on else 0)


def set_stream_rows(on=True):
    """Streamed-rows download of a single request's field (default on)."""
    _check(lib.srwlAmdSetStreamRows(1 if on else 0))


def last_streamed():
    """Whether the last single-request call streamed its rows to the end."""
    return bool(lib.srwlAmdLastStreamed())


def set_bands(n=0):
    """Row bands of a single request's streamed download (0: automatic)."""
    _check(lib.srwlAmdSetBands(int(n)))


def set_energy_tiles(e=8):
    """Energies per workgroup of the energy-tile SR kernel: 8 (default), 4, or 0 (off)."""
    _check(lib.srwlAmdSetEnergyTiles(int(e)))


def set_cond_tol(tol=1e-4):
    """Tolerance of the FP32 condition check of the SR sums (0: off)."""
    _check(lib.srwlAmdSetCondTol(C.c_double(tol)))


def set_fft_tiles(t7=None):
    """Forces the FFT tile shapes (7 ints) of every pass; None returns to the autotuner."""
    if t7 is None:
        lib.srwlAmdSetFftTiles(None)
    else:
        lib.srwlAmdSetFftTiles((C.c_int * 7)(*[int(v) for v in t7]))


def set_tuning(on=True):
    lib.srwlAmdSetTuning(1 if on else 0)


def fft_tiles(lognx, lognz, mode, post_op):
    """Tile shapes the autotuner chose for a drift shape, or None."""
    out = (C.c_int * 7)()
    if lib.srwlAmdGetFftTiles(int(lognx), int(lognz), int(mode), int(post_op), out):
        return list(out)
    return None


def set_resize_per_point(on=True):
    lib.srwlAmdSetResizePerPoint(1 if on else 0)


def last_auto_levels():
    """Final levels of the points of the last method-1 field computation, in
    the reference's point order (rows, columns, energies of the computed region)."""
    import numpy as np
    n = lib.srwlAmdLastAutoLevels(None, 0)
    out = np.zeros(max(n, 1), dtype=np.int32)
    lib.srwlAmdLastAutoLevels(out.ctypes.data_as(C.POINTER(C.c_int)), n)
    return out[:n]


def ResizeElecField(_wfr, _type, _par):
    """Resampling of the wavefront mesh (reference srwlpy.cpp ResizeElecField,
    srwlib.cpp:860-900); _par = [0, pxm, pxd, pzm, pzd, 0.5, 0.5]."""
    keep = M.Keep()
    cw = M.wfr(_wfr, keep)
    par = list(_par) + [0.5] * max(0, 7 - len(_par))
    cp = M.dvec(par, keep)
    t = _type.encode() if isinstance(_type, str) else _type
    _active_wfr[C.addressof(cw)] = _wfr
    try:
        rc = lib.srwlResizeElecField(C.byref(cw), C.c_char(t), cp)
    finally:
        _active_wfr.pop(C.addressof(cw), None)
    _check(rc)
    M.wfr_back(cw, _wfr)
    return _wfr


def ResizeElecFieldMesh(_wfr, _mesh, _par):
    """Resampling of the wavefront onto the mesh _mesh (reference srwlpy.cpp
    ResizeElecFieldMesh, srwlib.cpp:914-931); _par = [FFT side (0), allow the
    quadratic-phase treatment, allow the Re/Im correction], defaults {0, 1, 1}
    for the elements not given, as in the reference binding."""
    keep = M.Keep()
    cw = M.wfr(_wfr, keep)
    cm = M.mesh(_mesh)
    par = [0., 1., 1.]
    for i, v in enumerate(list(_par)[:3]):
        par[i] = float(v)
    cp = M.dvec(par, keep)
    _active_wfr[C.addressof(cw)] = _wfr
    try:
        rc = lib.srwlResizeElecFieldMesh(C.byref(cw), C.byref(cm), cp)
    finally:
        _active_wfr.pop(C.addressof(cw), None)
    _check(rc)
    M.wfr_back(cw, _wfr)
    return _wfr


def PropagElecField(_wfr, _opt_cnt, _int_list=None, _gpu_idx=None):
    """Propagation through a container of drifts, thin lenses, apertures and obstacles (reference srwlpy.cpp:5214-5320)."""
    if _int_list is not None:
        raise RuntimeError(err_text(23301))
    keep = M.Keep()
    cw = M.wfr(_wfr, keep)
    co = M.opt_cnt(_opt_cnt, keep)
    _active_wfr[C.addressof(cw)] = _wfr     # resize instructions may reallocate the arrays
    try:
        rc = lib.srwlPropagElecField(C.byref(cw), C.byref(co), 0, None, None, None, None)
    finally:
        _active_wfr.pop(C.addressof(cw), None)
    _check(rc)
    M.wfr_back(cw, _wfr)
    return _wfr


def CalcIntFromElecField(_arI, _wfr, _pol, _int_type, _dep_type, _e, _x, _y, _meth=None, _fldtrj=None,
                         _gpu_idx=None):
    """Intensity from the electric field (reference srwlpy.cpp:4873).  _arI is
    float ('f'), or double ('d') for the phase (_int_type 4); _meth is arMeth
    (srwlib.h:727: [0] = 1 running average over [1] previous values, 2 sum)."""
    keep = M.Keep()
    cw = M.wfr(_wfr, keep)
    pm = M.dvec(_meth, keep) if _meth is not None else None
    _check(lib.srwlCalcIntFromElecField(M._addr(_arI), C.byref(cw), int(_pol), int(_int_type), int(_dep_type),
                                        float(_e), float(_x), float(_y), pm, None, None))
    return _arI


def UtiFFT(_data, _mesh, _dir):
    """SRW-convention 2D FFT of float complex data; _mesh = [xStart, xStep, nx, yStart, yStep, ny] is updated."""
    keep = M.Keep()
    cm = M.dvec(_mesh, keep)
    _check(lib.srwlUtiFFT(M._addr(_data), b"f", cm, len(_mesh), int(_dir), None))
    for i in range(len(_mesh)):
        _mesh[i] = cm[i]
    return _data


def EmitPropMultiE(_stokes, _wfr0, _mag, _opt_cnt, _prec_par, _n_macro, _seed=1, _batch=0, _proposal=0,
                   _target=0., _train=0, _ideal_gain=0, _rqmc=0, _replicates=16):
    """Partially coherent intensity averaged over _n_macro electrons drawn from _wfr0.partBeam
    (replaces the per-electron loop of srwl_wfr_emit_prop_multi_e, reference srwlib.py:8506).
    _stokes.arS (float, nx*ny of _stokes.mesh) receives the averaged total intensity.
    _proposal=1: learned importance-sampling proposal (a batch is drawn from
    it only where the electrons so far predict a variance gain, otherwise from
    the beam distribution itself); _rqmc=1: randomized quasi-Monte Carlo draws
    in _replicates independently shifted replicates (with _proposal=1 the
    proposal is sampled through the same points).  Returns a dict of run
    information."""
    keep = M.Keep()
    cs = M.stokes(_stokes, keep)
    cw = M.wfr(_wfr0, keep)
    cm = M.mag_cnt(_mag, keep)
    co = M.opt_cnt(_opt_cnt, keep) if _opt_cnt is not None else None
    prec = M.dvec(_prec_par, keep)
    par = M.dvec([_seed, _batch, _proposal, _target, _train, 1 if _ideal_gain else 0, 1 if _rqmc else 0,
                  _replicates], keep)
    out = (C.c_double * 16)()
    _check(lib.srwlEmitPropMultiE(C.byref(cs), C.byref(cw), C.byref(cm), C.byref(co) if co is not None else None,
                                  prec, len(_prec_par), int(_n_macro), par, out))
    keys = ("electrons", "ess", "sum_w", "seconds", "rel_l2_est", "batch", "train_loss", "train_steps",
            "ideal_gain", "replicates", "mixture_batches", "predicted_ratio")
    return dict(zip(keys, list(out)[:12]))


def normals(seed, i):
    g = (C.c_double * 6)()
    lib.srwlAmdNormals(C.c_ulonglong(seed), C.c_longlong(i), g)
    return list(g)


def proposal_grad(P, c6, g, wh, alpha=0.25):
    """Runs the proposal kernels (forward, mixture gradient, backward) once; returns (G, out, loss)."""
    import numpy as np
    P = np.ascontiguousarray(P, np.float32)
    c6 = np.ascontiguousarray(c6, np.float32)
    g = np.ascontiguousarray(g, np.float32)
    wh = np.ascontiguousarray(wh, np.float32)
    G = np.zeros_like(P)
    out = np.zeros(104, np.float32)
    loss = np.zeros(1, np.float32)
    fp = C.POINTER(C.c_float)
    _check(lib.srwlAmdProposalGrad(P.ctypes.data_as(fp), c6.ctypes.data_as(fp), g.ctypes.data_as(fp),
                                   wh.ctypes.data_as(fp), len(wh), C.c_float(alpha), G.ctypes.data_as(fp),
                                   out.ctypes.data_as(fp), loss.ctypes.data_as(fp)))
    return G, out, float(loss[0])


def proposal_init(seed=0x6c65617264):
    import numpy as np
    P = np.zeros(11368, np.float32)
    lib.srwlAmdProposalInit(P.ctypes.data_as(C.POINTER(C.c_float)), C.c_ulonglong(seed))
    return P
