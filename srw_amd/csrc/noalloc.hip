// noalloc.hip - srwlCalcElecFieldSR without device allocation, for callers
// that cannot allocate (code under HIP graph capture, tight serving loops).
//
// Every device buffer of the call comes from one fixed file-scope array
// (`scratch`, SRWL_AMD_NOALLOC_BYTES per GPU, held until the process ends)
// through the arena of srw_api_internal.h; the caller's host arrays receive
// the field by plain copies.  srwlAmdNoAllocInit, called once per device
// with that device current, sets up the library (streams, pinned flags) and
// looks up the array's device address.  Calls on one GPU share its scratch
// and keep the last call's contents; calls that overlap on one GPU race.
#include "srw_api_internal.h"

#define SRWL_AMD_NOALLOC_BYTES (64ull << 20)
#define SRWL_AMD_NOALLOC_MAX_POINTS 65536LL
#define SRWL_AMD_NOALLOC_MAX_TRAJ 20001LL
#define SRWL_AMD_NOALLOC_MAX_DEV 64

static __device__ __align__(256) char scratch[SRWL_AMD_NOALLOC_BYTES];

namespace {
void* g_scratchDev[SRWL_AMD_NOALLOC_MAX_DEV] = {nullptr};
}

using namespace srwamd;

extern "C" {

SRWL_EXPORT int srwlAmdNoAllocInit(void)
{
    int rc = ensure_device();
    if (rc) return rc;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SRWL_AMD_NOALLOC_MAX_DEV) return SRWL_AMD_ERR_GPU;
    if (g_scratchDev[dev]) return 0;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || ptr == nullptr) return SRWL_AMD_ERR_GPU;
    g_scratchDev[dev] = ptr;
    return 0;
}

SRWL_EXPORT int srwlCalcElecFieldSRNoAlloc(SRWLWfr* pWfr, SRWLPrtTrj* pTrj, SRWLMagFldC* pMagFld, double* precPar,
                                           int nPrecPar)
{
    if (pWfr == nullptr || precPar == nullptr || nPrecPar < 5) return SRWL_INCORRECT_PARAM_FOR_SR_COMP;
    const int dev = g_ctx.device;
    if (!g_ctx.ready || dev < 0 || dev >= SRWL_AMD_NOALLOC_MAX_DEV || g_scratchDev[dev] == nullptr)
        return SRWL_AMD_ERR_NOALLOC_INIT;
    // limits of the fixed scratch: method 0, field and trajectory sizes, no
    // automatic mesh (that would need the caller's reallocation)
    const SRWLRadMesh& m = pWfr->mesh;
    const long long pts = (long long)m.ne * m.nx * m.ny;
    const bool trjDef = pTrj != nullptr && pTrj->np > 0 && pTrj->arX != nullptr;
    const long long np = trjDef ? pTrj->np : (long long)precPar[4];
    if ((int)precPar[0] != 0 || pts <= 0 || pts > SRWL_AMD_NOALLOC_MAX_POINTS || np > SRWL_AMD_NOALLOC_MAX_TRAJ ||
        (nPrecPar > 6 && precPar[6] > 0.) || g_ctx.keepOnDevice)
        return SRWL_AMD_ERR_NOALLOC_SIZE;
    Arena arena;
    arena.base = (char*)g_scratchDev[dev];
    arena.cap = SRWL_AMD_NOALLOC_BYTES;
    g_arena = &arena;
    int rc = srwlCalcElecFieldSR(pWfr, pTrj, pMagFld, precPar, nPrecPar);
    g_arena = nullptr;
    if (arena.overflow) return SRWL_AMD_ERR_NOALLOC_SIZE;
    return rc;
}

}  // extern "C"
