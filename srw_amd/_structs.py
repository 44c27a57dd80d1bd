"""ctypes mirror of the SRWLIB C structures (include/srwlib.h).

Field order and types follow the reference's C API structs
(cpp/src/lib/srwlib.h:75-532) so the same memory can be handed to any
library exporting that ABI.
"""
import ctypes as C

c_double_p = C.POINTER(C.c_double)


class SRWLParticle(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
                ("xp", C.c_double), ("yp", C.c_double), ("gamma", C.c_double),
                ("relE0", C.c_double), ("nq", C.c_int)]


class SRWLPartBeam(C.Structure):
    _fields_ = [("Iavg", C.c_double), ("nPart", C.c_double),
                ("partStatMom1", SRWLParticle), ("arStatMom2", C.c_double * 21)]


class SRWLMagFldH(C.Structure):
    _fields_ = [("n", C.c_int), ("h_or_v", C.c_char), ("B", C.c_double),
                ("ph", C.c_double), ("s", C.c_int), ("a", C.c_double)]


class SRWLMagFldU(C.Structure):
    _fields_ = [("arHarm", C.POINTER(SRWLMagFldH)), ("nHarm", C.c_int),
                ("per", C.c_double), ("nPer", C.c_int)]


class SRWLMagFldC(C.Structure):
    _fields_ = [("arMagFld", C.POINTER(C.c_void_p)), ("arMagFldTypes", C.c_char_p),
                ("arXc", c_double_p), ("arYc", c_double_p), ("arZc", c_double_p),
                ("arVx", c_double_p), ("arVy", c_double_p), ("arVz", c_double_p),
                ("arAng", c_double_p), ("arPar1", c_double_p), ("arPar2", c_double_p),
                ("arPar3", c_double_p), ("arPar4", c_double_p), ("nElem", C.c_int)]


class SRWLPrtTrj(C.Structure):
    _fields_ = [("arX", c_double_p), ("arXp", c_double_p), ("arY", c_double_p),
                ("arYp", c_double_p), ("arZ", c_double_p), ("arZp", c_double_p),
                ("arBx", c_double_p), ("arBy", c_double_p), ("arBz", c_double_p),
                ("np", C.c_longlong), ("ctStart", C.c_double), ("ctEnd", C.c_double),
                ("partInitCond", SRWLParticle)]


class SRWLRadMesh(C.Structure):
    _fields_ = [("eStart", C.c_double), ("eFin", C.c_double), ("xStart", C.c_double),
                ("xFin", C.c_double), ("yStart", C.c_double), ("yFin", C.c_double),
                ("zStart", C.c_double), ("ne", C.c_long), ("nx", C.c_long), ("ny", C.c_long),
                ("nvx", C.c_double), ("nvy", C.c_double), ("nvz", C.c_double),
                ("hvx", C.c_double), ("hvy", C.c_double), ("hvz", C.c_double),
                ("arSurf", c_double_p), ("type", C.c_char),
                ("itStart", C.c_longlong), ("itFin", C.c_longlong)]


class SRWLWfr(C.Structure):
    _fields_ = [("arEx", C.c_void_p), ("arEy", C.c_void_p), ("arExAux", C.c_void_p),
                ("arEyAux", C.c_void_p), ("mesh", SRWLRadMesh),
                ("Rx", C.c_double), ("Ry", C.c_double), ("dRx", C.c_double), ("dRy", C.c_double),
                ("xc", C.c_double), ("yc", C.c_double), ("avgPhotEn", C.c_double),
                ("presCA", C.c_char), ("presFT", C.c_char), ("numTypeElFld", C.c_char),
                ("unitElFld", C.c_char), ("unitElFldAng", C.c_char),
                ("partBeam", SRWLPartBeam), ("arElecPropMatr", c_double_p),
                ("arMomX", c_double_p), ("arMomY", c_double_p), ("arWfrAuxData", c_double_p)]


class SRWLStokes(C.Structure):
    _fields_ = [("arS0", C.c_void_p), ("arS1", C.c_void_p), ("arS2", C.c_void_p),
                ("arS3", C.c_void_p), ("mesh", SRWLRadMesh), ("avgPhotEn", C.c_double),
                ("presCA", C.c_char), ("presFT", C.c_char), ("numTypeStokes", C.c_char),
                ("unitStokes", C.c_char)]


class SRWLOptD(C.Structure):
    _fields_ = [("L", C.c_double), ("treat", C.c_char)]


class SRWLOptA(C.Structure):
    _fields_ = [("shape", C.c_char), ("ap_or_ob", C.c_char), ("Dx", C.c_double), ("Dy", C.c_double),
                ("x", C.c_double), ("y", C.c_double)]


class SRWLOptL(C.Structure):
    _fields_ = [("Fx", C.c_double), ("Fy", C.c_double), ("x", C.c_double), ("y", C.c_double)]


class SRWLOptC(C.Structure):
    _fields_ = [("arOpt", C.POINTER(C.c_void_p)), ("arOptTypes", C.POINTER(C.c_char_p)),
                ("nElem", C.c_int), ("arProp", C.POINTER(c_double_p)),
                ("arPropN", C.c_char_p), ("nProp", C.c_int)]


STRUCTS = [SRWLParticle, SRWLPartBeam, SRWLMagFldH, SRWLMagFldU, SRWLMagFldC, SRWLPrtTrj,
           SRWLRadMesh, SRWLWfr, SRWLStokes, SRWLOptD, SRWLOptC]
