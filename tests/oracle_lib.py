"""ctypes loader of the FP64 CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it is built by oracle/Makefile (or
__graft_entry__.build()) and only tests, smoke() and the CPU-baseline leg of
bench.py load it.
"""
import ctypes as C
import os
import subprocess

from srw_amd import _marshal as M
from srw_amd import _structs as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "liboracle.so")


def load():
    if not os.path.exists(PATH):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    lib = C.CDLL(PATH)
    P = C.POINTER
    lib.orc_srwlCalcElecFieldSR_mt.argtypes = [P(S.SRWLWfr), P(S.SRWLPrtTrj), P(S.SRWLMagFldC), P(C.c_double),
                                               C.c_int, C.c_int]
    lib.orc_srwlCalcElecFieldSR_ex.argtypes = [P(S.SRWLWfr), P(S.SRWLPrtTrj), P(S.SRWLMagFldC), P(C.c_double),
                                               C.c_int, C.c_int, P(C.c_int), P(C.c_long)]
    lib.orc_srwlCalcPartTraj.argtypes = [P(S.SRWLPrtTrj), P(S.SRWLMagFldC), P(C.c_double)]
    lib.orc_srwlPropagElecField.argtypes = [P(S.SRWLWfr), P(S.SRWLOptC)]
    lib.orc_srwlCalcIntFromElecField.argtypes = [C.c_void_p, P(S.SRWLWfr), C.c_byte, C.c_byte, C.c_byte, C.c_double,
                                                 C.c_double, C.c_double]
    lib.orc_srwlCalcIntFromElecField_m.argtypes = [C.c_void_p, P(S.SRWLWfr), C.c_byte, C.c_byte, C.c_byte,
                                                   C.c_double, C.c_double, C.c_double, P(C.c_double)]
    lib.orc_srwlUtiFFT.argtypes = [C.c_void_p, C.c_char, P(C.c_double), C.c_int, C.c_int]
    lib.orc_field_B.argtypes = [P(S.SRWLMagFldC), C.c_double, C.c_double, C.c_double, P(C.c_double)]
    lib.orc_trj_interp.argtypes = [P(S.SRWLPrtTrj), P(C.c_double), C.c_longlong, P(C.c_double)]
    lib.orc_struct_sizes.argtypes = [P(C.c_long)]
    lib.orc_beam_sample.argtypes = [P(S.SRWLPartBeam), C.c_ulonglong, C.c_longlong, P(S.SRWLParticle)]
    lib.orc_next_fft_size.argtypes = [C.c_long]
    lib.orc_next_fft_size.restype = C.c_long
    lib.orc_resize_mesh.argtypes = [P(S.SRWLWfr), P(C.c_double), C.c_int, P(C.c_double)]
    lib.orc_resize_fill.argtypes = [P(S.SRWLWfr), P(S.SRWLWfr), P(C.c_double), P(C.c_double)]
    lib.orc_resize_fill_mesh.argtypes = [P(S.SRWLWfr), P(S.SRWLWfr), P(C.c_double)]
    lib.orc_compute_rad_moments.argtypes = [P(S.SRWLWfr), C.c_void_p, C.c_void_p]
    lib.orc_rad_to_sub.argtypes = [P(S.SRWLWfr), C.c_double, C.c_int, C.c_void_p, C.c_void_p, P(C.c_double)]
    lib.orc_emit_prop_multi_e.argtypes = [P(S.SRWLStokes), P(S.SRWLWfr), P(S.SRWLMagFldC), P(S.SRWLOptC),
                                          P(C.c_double), C.c_int, C.c_longlong, C.c_ulonglong, C.c_longlong,
                                          C.c_int, P(C.c_double)]
    return lib


lib = load()


def calc_field(w, mag, prec, trj=None, nthreads=8):
    keep = M.Keep()
    cw = M.wfr(w, keep)
    cm = M.mag_cnt(mag, keep) if mag is not None else None
    ct = M.prt_trj(trj, keep) if trj is not None else None
    cp = M.dvec(prec, keep)
    rc = lib.orc_srwlCalcElecFieldSR_mt(C.byref(cw), C.byref(ct) if ct is not None else None,
                                        C.byref(cm) if cm is not None else None, cp, len(prec), nthreads)
    M.wfr_back(cw, w)
    return rc


def calc_field_levels(w, mag, prec, trj=None):
    """Method-1 field with the final level of every computed point, in the
    reference's order (rows, columns, energies of the computed region)."""
    import numpy as np
    keep = M.Keep()
    cw = M.wfr(w, keep)
    cm = M.mag_cnt(mag, keep) if mag is not None else None
    ct = M.prt_trj(trj, keep) if trj is not None else None
    cp = M.dvec(prec, keep)
    lv = np.full(w.mesh.nx * w.mesh.ny * w.mesh.ne, -1, dtype=np.int32)
    rc = lib.orc_srwlCalcElecFieldSR_ex(C.byref(cw), C.byref(ct) if ct is not None else None,
                                        C.byref(cm) if cm is not None else None, cp, len(prec), 1,
                                        lv.ctypes.data_as(C.POINTER(C.c_int)), None)
    M.wfr_back(cw, w)
    return rc, lv[lv >= 0]


WFR_EXT_FUNC_NOT_DEFINED = 23172


def calc_field_auto_mesh(w, mag, prec, nthreads=8):
    """Field with the automatic mesh of precPar[6] > 0: the oracle reports the
    sizes it needs, the wavefront is reallocated here (the caller's part in
    the reference) and the computation repeated on the new mesh."""
    sizes = (C.c_long * 2)(0, 0)

    def run():
        keep = M.Keep()
        cw = M.wfr(w, keep)
        cm = M.mag_cnt(mag, keep)
        cp = M.dvec(prec, keep)
        rc = lib.orc_srwlCalcElecFieldSR_ex(C.byref(cw), None, C.byref(cm), cp, len(prec), nthreads, None, sizes)
        M.wfr_back(cw, w)
        return rc
    rc = run()
    if rc == WFR_EXT_FUNC_NOT_DEFINED:
        w.allocate(w.mesh.ne, sizes[0], sizes[1])
        rc = run()
    return rc, (sizes[0], sizes[1])


def trj_interp(trj, s):
    """Interpolated trajectory data at the positions s: rows of (Btx, X,
    IBx, Btz, Z, IBz, Bx, Bz)."""
    import numpy as np
    keep = M.Keep()
    ct = M.prt_trj(trj, keep)
    s = np.ascontiguousarray(s, dtype=np.float64)
    out = np.zeros((len(s), 8))
    rc = lib.orc_trj_interp(C.byref(ct), s.ctypes.data_as(C.POINTER(C.c_double)), len(s),
                            out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    return out


def calc_traj(trj, mag):
    keep = M.Keep()
    ct = M.prt_trj(trj, keep)
    cm = M.mag_cnt(mag, keep)
    rc = lib.orc_srwlCalcPartTraj(C.byref(ct), C.byref(cm), None)
    M.prt_trj_back(ct, trj)
    return rc


def propagate(w, optc):
    keep = M.Keep()
    cw = M.wfr(w, keep)
    co = M.opt_cnt(optc, keep)
    rc = lib.orc_srwlPropagElecField(C.byref(cw), C.byref(co))
    M.wfr_back(cw, w)
    return rc


def intensity(arI, w, pol=6, int_type=0, dep_type=3, e=None, x=0., y=0., meth=None):
    keep = M.Keep()
    cw = M.wfr(w, keep)
    e = w.mesh.eStart if e is None else e
    pm = M.dvec(meth, keep) if meth is not None else None
    return lib.orc_srwlCalcIntFromElecField_m(M._addr(arI), C.byref(cw), pol, int_type, dep_type, e, x, y, pm)


def uti_fft(data, mesh, direction):
    keep = M.Keep()
    cm = M.dvec(mesh, keep)
    rc = lib.orc_srwlUtiFFT(M._addr(data), b"f", cm, len(mesh), direction)
    for i in range(len(mesh)):
        mesh[i] = cm[i]
    return rc


def field_B(mag, x, y, z):
    keep = M.Keep()
    cm = M.mag_cnt(mag, keep)
    out = (C.c_double * 3)()
    lib.orc_field_B(C.byref(cm), x, y, z, out)
    return list(out)


def multi_e(stokes, w0, mag, optc, prec, n_macro, seed=1, first=0, nthreads=8, sums=None):
    keep = M.Keep()
    cs = M.stokes(stokes, keep)
    cw = M.wfr(w0, keep)
    cm = M.mag_cnt(mag, keep)
    co = M.opt_cnt(optc, keep) if optc is not None else None
    cp = M.dvec(prec, keep)
    out2 = None
    if sums is not None:
        out2 = (C.c_double * (2 * stokes.mesh.nx * stokes.mesh.ny))()
    rc = lib.orc_emit_prop_multi_e(C.byref(cs), C.byref(cw), C.byref(cm), C.byref(co) if co is not None else None,
                                   cp, len(prec), n_macro, seed, first, nthreads, out2)
    if sums is not None:
        sums.extend(list(out2))
    return rc


def resize(w, par):
    """srwlResizeElecField(w, 'c', par) restated: returns the new wavefront
    (w's arrays are consumed by the in-place phase treatment)."""
    import copy
    from array import array
    keep = M.Keep()
    cw = M.wfr(w, keep)
    cp = M.dvec(par, keep)
    out = (C.c_double * 10)()
    rc = lib.orc_resize_mesh(C.byref(cw), cp, len(par), out)
    if rc:
        return rc, None
    nx, ny = int(out[0]), int(out[1])
    wn = copy.deepcopy(w)
    wn.allocate(w.mesh.ne, nx, ny)
    wn.mesh.nx, wn.mesh.ny = nx, ny
    wn.mesh.xStart, wn.mesh.xFin, wn.mesh.yStart, wn.mesh.yFin = out[2], out[3], out[4], out[5]
    cn = M.wfr(wn, keep)
    rc = lib.orc_resize_fill(C.byref(cw), C.byref(cn), cp, out)
    return rc, wn


def resize_to_mesh(w, mesh, par=(0, 1, 1)):
    """srwlResizeElecFieldMesh(w, mesh, par) restated for a transverse mesh:
    returns (rc, new wavefront); w's arrays are consumed by the phase
    treatment."""
    import copy
    keep = M.Keep()
    cw = M.wfr(w, keep)
    wn = copy.deepcopy(w)
    wn.allocate(w.mesh.ne, mesh.nx, mesh.ny)
    wn.mesh.nx, wn.mesh.ny = mesh.nx, mesh.ny
    wn.mesh.xStart, wn.mesh.xFin, wn.mesh.yStart, wn.mesh.yFin = mesh.xStart, mesh.xFin, mesh.yStart, mesh.yFin
    cn = M.wfr(wn, keep)
    cp = M.dvec(list(par), keep)
    rc = lib.orc_resize_fill_mesh(C.byref(cw), C.byref(cn), cp)
    return rc, wn


def rad_to_sub(w, L, subtype=1):
    """The drift's radius estimate of arProp[3] = 1 (subtype 1) or 2
    (subtype 2); returns (effRx, effRz, arMomX, arMomY as left by it)."""
    import numpy as np
    keep = M.Keep()
    cw = M.wfr(w, keep)
    mx = np.array(w.arMomX, dtype=np.float64)
    my = np.array(w.arMomY, dtype=np.float64)
    out = (C.c_double * 2)()
    lib.orc_rad_to_sub(C.byref(cw), float(L), int(subtype), mx.ctypes.data, my.ctypes.data, out)
    return out[0], out[1], mx, my


def edge_noise_mask(w, mag, prec, tol, nthreads=8):
    """Points of the field computation on wavefront w (not modified) whose
    result depends on the rounding noise of the field at the integration
    edges: the oracle is run with the edge field and its derivative as
    computed, zeroed and doubled (orc_edge_field_scale), and a point is
    marked where either variant moves Ex or Ey by more than tol times that
    component's largest magnitude.  Returns a (ny, nx, ne) bool array."""
    import numpy as np
    from tests.util import cfield, clone_wfr
    knob = C.c_double.in_dll(lib, "orc_edge_field_scale")
    out = []
    try:
        for sc in (1., 0., 2.):
            knob.value = sc
            wv = clone_wfr(w)
            calc_field(wv, mag, prec, nthreads=nthreads)
            out.append((cfield(wv.arEx, wv), cfield(wv.arEy, wv)))
    finally:
        knob.value = 1.
    mask = np.zeros(out[0][0].shape, dtype=bool)
    for k in (1, 2):
        for c in (0, 1):
            ref = out[0][c]
            scale = max(np.abs(ref).max(), 1e-300)
            mask |= np.abs(out[k][c] - ref) > tol * scale
    return mask
