// other_api.cpp - the reference C API entry points outside this library's
// path (other sources, field and transmission utilities, Stokes parameters
// and power density, intensity processing).  They keep the reference's
// names and signatures (srwlib.h lines cited in include/srwlib.h) so that a
// binding built against the reference links unchanged, and report
// SRWL_AMD_ERR_UNSUPPORTED without touching their arguments.
#include "srw_api_internal.h"

extern "C" {

SRWL_EXPORT int srwlCalcMagFld(SRWLMagFldC*, SRWLMagFldC*, double*) { return SRWL_AMD_ERR_UNSUPPORTED; }
SRWL_EXPORT int srwlCalcPartTrajFromKickMatr(SRWLPrtTrj*, SRWLKickM*, int, double*) { return SRWL_AMD_ERR_UNSUPPORTED; }
SRWL_EXPORT int srwlCalcElecFieldGaussian(SRWLWfr*, SRWLGsnBm*, double*) { return SRWL_AMD_ERR_UNSUPPORTED; }
SRWL_EXPORT int srwlCalcElecFieldPointSrc(SRWLWfr*, SRWLPtSrc*, double*) { return SRWL_AMD_ERR_UNSUPPORTED; }
SRWL_EXPORT int srwlCalcStokesUR(SRWLStokes*, SRWLPartBeam*, SRWLMagFldU*, double*) { return SRWL_AMD_ERR_UNSUPPORTED; }
SRWL_EXPORT int srwlCalcPowDenSR(SRWLStokes*, SRWLPartBeam*, SRWLPrtTrj*, SRWLMagFldC*, double*)
{
    return SRWL_AMD_ERR_UNSUPPORTED;
}
SRWL_EXPORT int srwlProcElecField(SRWLWfr*, double*, SRWLWfr*) { return SRWL_AMD_ERR_UNSUPPORTED; }
SRWL_EXPORT int srwlSetRepresElecField(SRWLWfr*, char) { return SRWL_AMD_ERR_UNSUPPORTED; }
SRWL_EXPORT int srwlPropagRadMultiE(SRWLStokes*, SRWLWfr*, SRWLOptC*, double*, int (*)(int, SRWLStokes*))
{
    return SRWL_AMD_ERR_UNSUPPORTED;
}
SRWL_EXPORT int srwlCalcTransm(SRWLOptT*, const double*, const double*, double**, int, const double*)
{
    return SRWL_AMD_ERR_UNSUPPORTED;
}
SRWL_EXPORT int srwlUtiConvWithGaussian(char*, char, double*, int, double*) { return SRWL_AMD_ERR_UNSUPPORTED; }
SRWL_EXPORT int srwlUtiIntInf(double*, char*, char, SRWLRadMesh*, double*, int) { return SRWL_AMD_ERR_UNSUPPORTED; }
SRWL_EXPORT int srwlUtiIntProc(char*, char, SRWLRadMesh*, char*, char, SRWLRadMesh*, double*, int)
{
    return SRWL_AMD_ERR_UNSUPPORTED;
}
SRWL_EXPORT int srwlUtiUndFromMagFldTab(SRWLMagFldC*, SRWLMagFldC*, double*) { return SRWL_AMD_ERR_UNSUPPORTED; }
SRWL_EXPORT int srwlUtiUndFindMagFldInterpInds(int*, int*, double*, double*, int, double*)
{
    return SRWL_AMD_ERR_UNSUPPORTED;
}

}  // extern "C"
