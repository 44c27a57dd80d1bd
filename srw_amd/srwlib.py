"""Python classes of the SRWLIB API used by the emit-and-propagate path.

The class names, constructor arguments and attribute names follow the
reference's env/python/srwpy/srwlib.py (SRWLParticle :67, SRWLPartBeam :109,
SRWLMagFldH :306, SRWLMagFldU :325, SRWLMagFldC :410, SRWLPrtTrj :623,
SRWLRadMesh :801, SRWLStokes :924, SRWLWfr :2200, SRWLOptD :2908,
SRWLOptC :4259), so scripts written against srwlib can build the same objects
and hand them to srw_amd.srwlpy.  Arrays are array.array('f'/'d') like the
reference's.
"""
from array import array


class SRWLParticle(object):
    def __init__(self, _x=0, _y=0, _z=0, _xp=0, _yp=0, _gamma=1, _relE0=1, _nq=-1):
        self.x, self.y, self.z = _x, _y, _z
        self.xp, self.yp = _xp, _yp
        self.gamma, self.relE0, self.nq = _gamma, _relE0, _nq

    def drift(self, _dist):
        self.x += self.xp * _dist
        self.y += self.yp * _dist
        self.z += _dist


class SRWLPartBeam(object):
    def __init__(self, _Iavg=0, _nPart=0, _partStatMom1=None, _arStatMom2=None):
        self.Iavg = _Iavg
        self.nPart = _nPart
        self.partStatMom1 = SRWLParticle() if _partStatMom1 is None else _partStatMom1
        self.arStatMom2 = array('d', [0] * 21) if _arStatMom2 is None else _arStatMom2


class SRWLMagFldH(object):
    def __init__(self, _n=1, _h_or_v='v', _B=0, _ph=0, _s=1, _a=1):
        self.n, self.h_or_v, self.B, self.ph, self.s, self.a = _n, _h_or_v, _B, _ph, _s, _a


class SRWLMagFldU(object):
    def __init__(self, _arHarm=None, _per=0, _nPer=0):
        self.arHarm = [] if _arHarm is None else _arHarm
        self.per = _per
        self.nPer = _nPer

    def allocate(self, _nHarm):
        self.arHarm = [SRWLMagFldH() for _ in range(_nHarm)]

    def get_K(self):
        """Deflection parameter of the first harmonic: K = 93.3729*per*B."""
        return 93.37290417576577 * self.per * self.arHarm[0].B


class SRWLMagFldC(object):
    def __init__(self, _arMagFld=None, _arXc=None, _arYc=None, _arZc=None, _arVx=None, _arVy=None,
                 _arVz=None, _arAng=None):
        self.arMagFld = [] if _arMagFld is None else list(_arMagFld)
        n = len(self.arMagFld)
        self.arXc = array('d', [0] * n) if _arXc is None else _arXc
        self.arYc = array('d', [0] * n) if _arYc is None else _arYc
        self.arZc = array('d', [0] * n) if _arZc is None else _arZc
        self.arVx = array('d', [0] * n) if _arVx is None else _arVx
        self.arVy = array('d', [0] * n) if _arVy is None else _arVy
        self.arVz = array('d', [1] * n) if _arVz is None else _arVz
        self.arAng = array('d', [0] * n) if _arAng is None else _arAng


class SRWLPrtTrj(object):
    def __init__(self, _arX=None, _arXp=None, _arY=None, _arYp=None, _arZ=None, _arZp=None,
                 _arBx=None, _arBy=None, _arBz=None, _np=0, _ctStart=0, _ctEnd=0, _partInitCond=None):
        self.arX, self.arXp, self.arY, self.arYp = _arX, _arXp, _arY, _arYp
        self.arZ, self.arZp = _arZ, _arZp
        self.arBx, self.arBy, self.arBz = _arBx, _arBy, _arBz
        self.np, self.ctStart, self.ctEnd = _np, _ctStart, _ctEnd
        self.partInitCond = SRWLParticle() if _partInitCond is None else _partInitCond

    def allocate(self, _np, _allB=False):
        self.np = _np
        z = [0] * _np
        self.arX, self.arXp, self.arY, self.arYp = array('d', z), array('d', z), array('d', z), array('d', z)
        self.arZ, self.arZp = array('d', z), array('d', z)
        if _allB:
            self.arBx, self.arBy, self.arBz = array('d', z), array('d', z), array('d', z)


class SRWLRadMesh(object):
    def __init__(self, _eStart=0, _eFin=0, _ne=1, _xStart=0, _xFin=0, _nx=1, _yStart=0, _yFin=0, _ny=1,
                 _zStart=0, _nvx=0, _nvy=0, _nvz=1, _hvx=1, _hvy=0, _hvz=0, _arSurf=None):
        self.eStart, self.eFin, self.ne = _eStart, _eFin, _ne
        self.xStart, self.xFin, self.nx = _xStart, _xFin, _nx
        self.yStart, self.yFin, self.ny = _yStart, _yFin, _ny
        self.zStart = _zStart
        self.nvx, self.nvy, self.nvz, self.hvx, self.hvy, self.hvz = _nvx, _nvy, _nvz, _hvx, _hvy, _hvz
        self.arSurf = _arSurf

    def set_from_other(self, _mesh):
        for k in ('eStart', 'eFin', 'ne', 'xStart', 'xFin', 'nx', 'yStart', 'yFin', 'ny', 'zStart'):
            setattr(self, k, getattr(_mesh, k))

    def is_equal(self, _mesh, _rel_tol=1e-12):
        for k in ('ne', 'nx', 'ny'):
            if getattr(self, k) != getattr(_mesh, k):
                return False
        for k in ('eStart', 'eFin', 'xStart', 'xFin', 'yStart', 'yFin'):
            a, b = getattr(self, k), getattr(_mesh, k)
            if abs(a - b) > _rel_tol * max(abs(a), abs(b), 1e-300):
                return False
        return True


class SRWLStokes(object):
    def __init__(self, _arS=None, _typeStokes='f', _eStart=0, _eFin=0, _ne=0, _xStart=0, _xFin=0, _nx=0,
                 _yStart=0, _yFin=0, _ny=0, _mutual=0, _n_comp=4):
        self.arS = _arS
        self.numTypeStokes = _typeStokes
        self.mesh = SRWLRadMesh(_eStart, _eFin, _ne, _xStart, _xFin, _nx, _yStart, _yFin, _ny)
        self.avgPhotEn = 0
        self.presCA = 0
        self.presFT = 0
        self.unitStokes = 1
        self.mutual = _mutual
        if (_arS is None) and (_ne > 0) and (_nx > 0) and (_ny > 0):
            self.allocate(_ne, _nx, _ny, _typeStokes, _mutual, _n_comp)

    def allocate(self, _ne, _nx, _ny, _typeStokes='f', _mutual=0, _n_comp=4):
        self.arS = array(_typeStokes, [0] * (_n_comp * _ne * _nx * _ny))
        self.numTypeStokes = _typeStokes
        self.mesh.ne, self.mesh.nx, self.mesh.ny = _ne, _nx, _ny


class SRWLWfr(object):
    def __init__(self, _arEx=None, _arEy=None, _typeE='f', _eStart=0, _eFin=0, _ne=0, _xStart=0, _xFin=0,
                 _nx=0, _yStart=0, _yFin=0, _ny=0, _zStart=0, _partBeam=None):
        self.arEx, self.arEy = _arEx, _arEy
        self.arExAux = self.arEyAux = None
        self.mesh = SRWLRadMesh(_eStart, _eFin, _ne, _xStart, _xFin, _nx, _yStart, _yFin, _ny, _zStart)
        self.Rx = self.Ry = self.dRx = self.dRy = 0.
        self.xc = self.yc = 0.
        self.avgPhotEn = 0.
        self.presCA = 0
        self.presFT = 0
        self.numTypeElFld = _typeE
        self.unitElFld = 1
        self.unitElFldAng = 0
        self.partBeam = SRWLPartBeam() if _partBeam is None else _partBeam
        self.arElecPropMatr = array('d', [0] * 20)
        self.arMomX = array('d', [0] * (11 * max(_ne, 1)))
        self.arMomY = array('d', [0] * (11 * max(_ne, 1)))
        self.arWfrAuxData = array('d', [0] * 30)
        if (_arEx is None) and (_ne > 0) and (_nx > 0) and (_ny > 0):
            self.allocate(_ne, _nx, _ny)

    def allocate(self, _ne, _nx, _ny, _EXNeeded=1, _EYNeeded=1, _typeE='f', _backupNeeded=0):
        n = 2 * _ne * _nx * _ny
        if _EXNeeded:
            self.arEx = array(_typeE, bytes(4 * n)) if _typeE == 'f' else array(_typeE, [0] * n)
        if _EYNeeded:
            self.arEy = array(_typeE, bytes(4 * n)) if _typeE == 'f' else array(_typeE, [0] * n)
        self.numTypeElFld = _typeE
        self.mesh.ne, self.mesh.nx, self.mesh.ny = _ne, _nx, _ny
        if len(self.arMomX) < 11 * _ne:
            self.arMomX = array('d', [0] * (11 * _ne))
            self.arMomY = array('d', [0] * (11 * _ne))


class SRWLOptD(object):
    def __init__(self, _L=0, _treat=0):
        self.L = _L
        self.treat = _treat


class SRWLOptA(object):
    """Aperture or obstacle (reference srwlib.py:2922-2936)."""
    def __init__(self, _shape='r', _ap_or_ob='a', _Dx=0, _Dy=0, _x=0, _y=0):
        self.shape = _shape
        self.ap_or_ob = _ap_or_ob
        self.Dx = _Dx
        self.Dy = _Dy
        self.x = _x
        self.y = _y


class SRWLOptL(object):
    """Thin lens (reference srwlib.py:2941-2951)."""
    def __init__(self, _Fx=1e+23, _Fy=1e+23, _x=0, _y=0):
        self.Fx = _Fx
        self.Fy = _Fy
        self.x = _x
        self.y = _y


class SRWLOptC(object):
    def __init__(self, _arOpt=None, _arProp=None):
        self.arOpt = [] if _arOpt is None else list(_arOpt)
        self.arProp = [] if _arProp is None else list(_arProp)
