"""Conversion between the Python SRWL* objects and the C structures.

The reference binding parses Python objects by attribute name into C structs
(srwlpy.cpp:3058 ParseSructSRWLWfr, :1470 ParseSructSRWLMagFldC) with
zero-copy views of the array('f'/'d') buffers, and copies scalar results back
(srwlpy.cpp:3599 UpdatePyWfr).  The same is done here with ctypes; the
Python arrays are shared with the library, never copied.
"""
import ctypes as C
from array import array

from . import _structs as S


def _addr(a):
    """Address of the first element of an array.array or a numpy array."""
    if a is None:
        return None
    if isinstance(a, array):
        return a.buffer_info()[0]
    if hasattr(a, "__array_interface__"):
        return a.__array_interface__["data"][0]
    if hasattr(a, "ctypes"):
        return a.ctypes.data
    raise TypeError("unsupported array type: %r" % type(a))


def _dptr(a):
    p = _addr(a)
    return C.cast(p, S.c_double_p) if p else S.c_double_p()


class Keep(list):
    """Holds ctypes objects alive for the duration of a library call."""


def particle(p):
    return S.SRWLParticle(p.x, p.y, p.z, p.xp, p.yp, p.gamma, p.relE0, int(p.nq))


def part_beam(b):
    cb = S.SRWLPartBeam()
    cb.Iavg = b.Iavg
    cb.nPart = b.nPart
    cb.partStatMom1 = particle(b.partStatMom1)
    for i in range(min(21, len(b.arStatMom2))):
        cb.arStatMom2[i] = b.arStatMom2[i]
    return cb


def mesh(m):
    cm = S.SRWLRadMesh()
    for k in ("eStart", "eFin", "xStart", "xFin", "yStart", "yFin", "zStart"):
        setattr(cm, k, float(getattr(m, k)))
    cm.ne, cm.nx, cm.ny = int(m.ne), int(m.nx), int(m.ny)
    cm.nvx, cm.nvy, cm.nvz = getattr(m, "nvx", 0), getattr(m, "nvy", 0), getattr(m, "nvz", 1)
    cm.hvx, cm.hvy, cm.hvz = getattr(m, "hvx", 1), getattr(m, "hvy", 0), getattr(m, "hvz", 0)
    return cm


def mesh_back(cm, m):
    for k in ("eStart", "eFin", "xStart", "xFin", "yStart", "yFin", "zStart"):
        setattr(m, k, getattr(cm, k))
    m.ne, m.nx, m.ny = cm.ne, cm.nx, cm.ny


def mag_cnt(mc, keep):
    """SRWLMagFldC of undulators ('u') -> C container."""
    n = len(mc.arMagFld)
    ptrs = (C.c_void_p * n)()
    types = C.create_string_buffer(n + 1)
    for i, el in enumerate(mc.arMagFld):
        if hasattr(el, "arHarm"):
            nh = len(el.arHarm)
            harms = (S.SRWLMagFldH * nh)()
            for k, h in enumerate(el.arHarm):
                hv = h.h_or_v.encode() if isinstance(h.h_or_v, str) else bytes([h.h_or_v])
                harms[k] = S.SRWLMagFldH(int(h.n), hv, float(h.B), float(h.ph), int(h.s), float(h.a))
            u = S.SRWLMagFldU(C.cast(harms, C.POINTER(S.SRWLMagFldH)), nh, float(el.per), int(el.nPer))
            keep += [harms, u]
            ptrs[i] = C.cast(C.pointer(u), C.c_void_p)
            types[i] = b"u"
        elif hasattr(el, "arMagFld"):
            sub = mag_cnt(el, keep)
            keep.append(sub)
            ptrs[i] = C.cast(C.pointer(sub), C.c_void_p)
            types[i] = b"c"
        else:
            raise TypeError("magnetic field element type not supported by this build")
    arrs = []
    for name in ("arXc", "arYc", "arZc", "arVx", "arVy", "arVz", "arAng"):
        a = getattr(mc, name, None)
        if a is None:
            arrs.append(None)
            continue
        ca = (C.c_double * n)(*[float(v) for v in a[:n]])
        keep.append(ca)
        arrs.append(ca)
    c = S.SRWLMagFldC()
    c.arMagFld = C.cast(ptrs, C.POINTER(C.c_void_p))
    c.arMagFldTypes = C.cast(types, C.c_char_p)
    for name, ca in zip(("arXc", "arYc", "arZc", "arVx", "arVy", "arVz", "arAng"), arrs):
        setattr(c, name, C.cast(ca, S.c_double_p) if ca is not None else S.c_double_p())
    c.nElem = n
    keep += [ptrs, types, c]
    return c


def prt_trj(t, keep):
    c = S.SRWLPrtTrj()
    for k in ("arX", "arXp", "arY", "arYp", "arZ", "arZp", "arBx", "arBy", "arBz"):
        setattr(c, k, _dptr(getattr(t, k, None)))
    c.np = int(t.np)
    c.ctStart, c.ctEnd = float(t.ctStart), float(t.ctEnd)
    c.partInitCond = particle(t.partInitCond)
    keep.append(c)
    return c


def prt_trj_back(c, t):
    t.ctStart, t.ctEnd = c.ctStart, c.ctEnd


def wfr(w, keep):
    c = S.SRWLWfr()
    c.arEx = _addr(w.arEx)
    c.arEy = _addr(w.arEy)
    c.mesh = mesh(w.mesh)
    for k in ("Rx", "Ry", "dRx", "dRy", "xc", "yc", "avgPhotEn"):
        setattr(c, k, float(getattr(w, k, 0.)))
    c.presCA = int(getattr(w, "presCA", 0))
    c.presFT = int(getattr(w, "presFT", 0))
    c.numTypeElFld = b"f"
    c.unitElFld = int(getattr(w, "unitElFld", 1))
    c.unitElFldAng = int(getattr(w, "unitElFldAng", 0))
    c.partBeam = part_beam(w.partBeam)
    c.arMomX = _dptr(getattr(w, "arMomX", None))
    c.arMomY = _dptr(getattr(w, "arMomY", None))
    c.arElecPropMatr = _dptr(getattr(w, "arElecPropMatr", None))
    c.arWfrAuxData = _dptr(getattr(w, "arWfrAuxData", None))
    keep.append(c)
    return c


def wfr_back(c, w):
    mesh_back(c.mesh, w.mesh)
    for k in ("Rx", "Ry", "dRx", "dRy", "xc", "yc", "avgPhotEn"):
        setattr(w, k, getattr(c, k))
    w.presCA = ord(c.presCA) if isinstance(c.presCA, bytes) else c.presCA
    w.presFT = ord(c.presFT) if isinstance(c.presFT, bytes) else c.presFT
    p = c.partBeam.partStatMom1
    q = w.partBeam.partStatMom1
    q.x, q.y, q.z, q.xp, q.yp, q.gamma = p.x, p.y, p.z, p.xp, p.yp, p.gamma


def opt_cnt(oc, keep):
    n = len(oc.arOpt)
    ptrs = (C.c_void_p * n)()
    types = (C.c_char_p * n)()
    for i, el in enumerate(oc.arOpt):
        if hasattr(el, "L"):
            d, t = S.SRWLOptD(float(el.L), int(getattr(el, "treat", 0))), b"drift"
        elif hasattr(el, "Fx"):
            d, t = S.SRWLOptL(float(el.Fx), float(el.Fy), float(el.x), float(el.y)), b"lens"
        elif hasattr(el, "ap_or_ob"):
            ob = str(el.ap_or_ob)
            d = S.SRWLOptA(str(el.shape).encode()[:1], ob.encode()[:1], float(el.Dx), float(el.Dy),
                           float(el.x), float(el.y))
            t = b"obstacle" if ob in ("o", "O") else b"aperture"
        else:
            raise TypeError("optical element %r is not supported (drift, lens, aperture, obstacle)" % (el,))
        keep.append(d)
        ptrs[i] = C.cast(C.pointer(d), C.c_void_p)
        types[i] = t
    nprop = len(oc.arProp)
    props = (S.c_double_p * max(nprop, 1))()
    propn = C.create_string_buffer(max(nprop, 1))
    for i, pp in enumerate(oc.arProp):
        ca = (C.c_double * max(len(pp), 17))(*[float(v) for v in pp])
        keep.append(ca)
        props[i] = C.cast(ca, S.c_double_p)
        propn[i] = bytes([min(len(pp), 127)])
    c = S.SRWLOptC()
    c.arOpt = C.cast(ptrs, C.POINTER(C.c_void_p))
    c.arOptTypes = C.cast(types, C.POINTER(C.c_char_p))
    c.nElem = n
    c.arProp = C.cast(props, C.POINTER(S.c_double_p)) if nprop else C.POINTER(S.c_double_p)()
    c.arPropN = C.cast(propn, C.c_char_p) if nprop else None
    c.nProp = nprop
    keep += [ptrs, types, props, propn, c]
    return c


def stokes(s, keep):
    c = S.SRWLStokes()
    c.arS0 = _addr(s.arS)
    c.mesh = mesh(s.mesh)
    c.numTypeStokes = b"f"
    keep.append(c)
    return c


def dvec(v, keep, n=None):
    if v is None:
        return None
    v = list(v)
    if n is not None and len(v) < n:
        v = v + [0.] * (n - len(v))
    ca = (C.c_double * len(v))(*[float(x) for x in v])
    keep.append(ca)
    return ca
