/*
 * srwlib.h - C ABI of the MI355X-native SRW emit-and-propagate library
 * (libsrw_amd.so).
 *
 * The structures below are laid out exactly like SRW's own C API structures
 * (cpp/src/lib/srwlib.h:75-532 in the reference), so existing callers that
 * fill SRWLWfr / SRWLPartBeam / SRWLMagFldC / SRWLOptC and call the srwl*
 * functions can link against this library instead.  Only the entry points of
 * the single-electron emit-and-propagate path are implemented; the other
 * functions of the reference API are not exported.
 *
 * Error convention (reference srwlib.cpp:524-528, srercode.h):
 *   return 0 on success, > 0 on error (23000-based codes), < 0 on warning
 *   (-13000-based codes).  srwlUtiGetErrText() returns the text.
 *
 * Axis naming follows the C API: x horizontal, y vertical, z longitudinal.
 */
#ifndef SRW_AMD_SRWLIB_H
#define SRW_AMD_SRWLIB_H

#ifdef __cplusplus
extern "C" {
#endif

#define SRWL_EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------------- */
/* Structures (byte-compatible with the reference)                         */
/* ---------------------------------------------------------------------- */

/* reference srwlib.h:75-82 */
typedef struct SRWLStructParticle {
    double x, y, z;   /* instant coordinates [m] */
    double xp, yp;    /* instant angles [rad] */
    double gamma;     /* relative energy */
    double relE0;     /* rest mass in electron rest masses */
    int nq;           /* charge in electron charges (-1 for electron) */
} SRWLParticle;

/* reference srwlib.h:87-94 */
typedef struct SRWLStructParticleBeam {
    double Iavg;              /* average current [A] */
    double nPart;             /* number of particles */
    SRWLParticle partStatMom1;/* first-order moments */
    double arStatMom2[21];    /* second-order moments */
} SRWLPartBeam;

/* reference srwlib.h:140-147: undulator harmonic ('h') */
typedef struct SRWLStructMagneticFieldHarmonic {
    int n;        /* harmonic number */
    char h_or_v;  /* 'h' horizontal or 'v' vertical field */
    double B;     /* peak field [T] */
    double ph;    /* phase [rad] */
    int s;        /* 1: symmetric (cos), -1: anti-symmetric (sin) */
    double a;     /* transverse dependence coefficient */
} SRWLMagFldH;

/* reference srwlib.h:153-159: undulator ('u') */
typedef struct SRWLStructMagneticFieldUndulator {
    SRWLMagFldH *arHarm;
    int nHarm;
    double per;   /* period [m] */
    int nPer;     /* number of periods */
} SRWLMagFldU;

/* reference srwlib.h:165-182: field container ('c') */
typedef struct SRWLStructMagneticFieldContainer {
    void **arMagFld;
    char *arMagFldTypes;
    double *arXc, *arYc, *arZc;
    double *arVx, *arVy, *arVz;
    double *arAng;
    double *arPar1, *arPar2, *arPar3, *arPar4;
    int nElem;
} SRWLMagFldC;

/* reference srwlib.h:187-194 */
typedef struct SRWLStructParticleTrajectory {
    double *arX, *arXp, *arY, *arYp, *arZ, *arZp;
    double *arBx, *arBy, *arBz;
    long long np;
    double ctStart, ctEnd;
    SRWLParticle partInitCond;
} SRWLPrtTrj;

/* reference srwlib.h:239-247 */
typedef struct SRWLStructRadMesh {
    double eStart, eFin, xStart, xFin, yStart, yFin, zStart;
    long ne, nx, ny;
    double nvx, nvy, nvz, hvx, hvy, hvz;
    double *arSurf;
    char type;
    long long itStart, itFin;
} SRWLRadMesh;

/* reference srwlib.h:253-277 */
typedef struct SRWLStructWaveFront {
    char *arEx, *arEy;        /* float32 (re,im) interleaved, [iy][ix][ie] */
    char *arExAux, *arEyAux;
    SRWLRadMesh mesh;
    double Rx, Ry;
    double dRx, dRy;
    double xc, yc;
    double avgPhotEn;
    char presCA;              /* 0 coordinates, 1 angles */
    char presFT;              /* 0 frequency, 1 time */
    char numTypeElFld;        /* 'f' */
    char unitElFld;
    char unitElFldAng;
    SRWLPartBeam partBeam;
    double *arElecPropMatr;
    double *arMomX, *arMomY;
    double *arWfrAuxData;
} SRWLWfr;

/* reference srwlib.h:283-295 */
typedef struct SRWLStructStokes {
    char *arS0, *arS1, *arS2, *arS3;
    SRWLRadMesh mesh;
    double avgPhotEn;
    char presCA;
    char presFT;
    char numTypeStokes;
    char unitStokes;
} SRWLStokes;

/* reference srwlib.h:302-306: drift ("drift") */
typedef struct SRWLStructOpticsDrift {
    double L;     /* length [m] */
    char treat;   /* 1: keep absolute optical path in the phase */
} SRWLOptD;

/* reference srwlib.h:312-318: aperture or obstacle ("aperture" / "obstacle") */
typedef struct SRWLStructOpticsAperture {
    char shape;      /* 'r' rectangular, 'c' circular */
    char ap_or_ob;   /* 'a' aperture, 'o' obstacle */
    double Dx, Dy;   /* transverse sizes [m]; the diameter is Dx for 'c' */
    double x, y;     /* centre [m] */
} SRWLOptA;

/* reference srwlib.h:324-328: thin lens ("lens") */
typedef struct SRWLStructOpticsLens {
    double Fx, Fy;   /* focal lengths [m] */
    double x, y;     /* centre [m] */
} SRWLOptL;

/* reference srwlib.h:523-532 */
typedef struct SRWLStructOpticsContainer {
    void **arOpt;
    char **arOptTypes;
    int nElem;
    double **arProp;
    char *arPropN;
    int nProp;
} SRWLOptC;

/* ---------------------------------------------------------------------- */
/* Error and warning codes (reference srercode.h; same numbers)            */
/* ---------------------------------------------------------------------- */
#define SRWL_FIRST_ERR 23000
#define SRWL_WARN_OFFSET (-13000)
#define SRWL_ERR_MEMORY_ALLOCATION_FAILURE            (8 + SRWL_FIRST_ERR)
#define SRWL_ERR_OBS_WITHIN_INTEGR_LIMITS             (26 + SRWL_FIRST_ERR)
#define SRWL_ERR_ELEC_BEAM_NOT_ULTRARELATIVISTIC      (27 + SRWL_FIRST_ERR)
#define SRWL_ERR_S0_OUT_OF_FIELD_LIMITS               (28 + SRWL_FIRST_ERR)
#define SRWL_ERR_WAVELENGTH_COMPARABLE_WITH_DISTANCE  (29 + SRWL_FIRST_ERR)
#define SRWL_ERR_IN_FFT                               (40 + SRWL_FIRST_ERR)
#define SRWL_ERR_CAN_NOT_COMPUTE_RADIATION_INTEGRAL   (56 + SRWL_FIRST_ERR)
#define SRWL_ERR_INCORRECT_MAG_FIELD_PER              (124 + SRWL_FIRST_ERR)
#define SRWL_ERR_INCORRECT_PER_MAG_FIELD_LENGTH       (125 + SRWL_FIRST_ERR)
#define SRWL_NO_FUNC_ARG_DATA                         (160 + SRWL_FIRST_ERR)
#define SRWL_MAG_FLD_UNRECOGN                         (161 + SRWL_FIRST_ERR)
#define SRWL_INCORRECT_TRJ_STRUCT                     (162 + SRWL_FIRST_ERR)
#define SRWL_INCORRECT_PARAM_FOR_TRJ_COMP             (163 + SRWL_FIRST_ERR)
#define SRWL_INCORRECT_PARAM_FOR_SR_COMP              (164 + SRWL_FIRST_ERR)
#define SRWL_INCORRECT_PARAM_FOR_INT_EXTR             (166 + SRWL_FIRST_ERR)
#define SRWL_INCORRECT_WFR_STRUCT                     (167 + SRWL_FIRST_ERR)
#define SRWL_WFR_EXT_MODIF_FAILED                     (168 + SRWL_FIRST_ERR)
#define SRWL_INCORRECT_PARAM_FOR_RESIZE               (169 + SRWL_FIRST_ERR)
#define SRWL_INCORRECT_PARAM_FOR_WFR_PROP             (171 + SRWL_FIRST_ERR)
#define SRWL_WFR_EXT_FUNC_NOT_DEFINED                 (172 + SRWL_FIRST_ERR)
#define SRWL_INCORRECT_PARAM_FOR_FFT                  (180 + SRWL_FIRST_ERR)
#define SRWL_RESIZE_NOT_SUPPORTED_WITH_SUCH_PARAMS    (193 + SRWL_FIRST_ERR)
/* Codes of this build (no counterpart in the reference): */
#define SRWL_AMD_ERR_GPU                              (300 + SRWL_FIRST_ERR)
#define SRWL_AMD_ERR_UNSUPPORTED                      (301 + SRWL_FIRST_ERR)
#define SRWL_AMD_ERR_BATCH                            (302 + SRWL_FIRST_ERR)
#define SRWL_AMD_ERR_NOALLOC_INIT                     (303 + SRWL_FIRST_ERR)
#define SRWL_AMD_ERR_NOALLOC_SIZE                     (304 + SRWL_FIRST_ERR)
/* Warnings */
#define SRWL_WARN_BAD_RAD_INT_TERMINATIONS            (12 + SRWL_WARN_OFFSET)
#define SRWL_WARN_FLUX_ONLY_VS_PHOTON_ENERGY          (21 + SRWL_WARN_OFFSET)
#define SRWL_WARN_ELEC_BEAM_NOT_ULTRARELATIVISTIC     (23 + SRWL_WARN_OFFSET)

/* ---------------------------------------------------------------------- */
/* Entry points of the reference API implemented here                      */
/* ---------------------------------------------------------------------- */

/* reference srwlib.h:542 (srwlib.cpp:97-101) */
SRWL_EXPORT void srwlUtiSetWfrModifFunc(int (*pExtFunc)(int action, SRWLWfr* pWfrIn, char pol));
/* reference srwlib.h:549 (srwlib.cpp:105-108) */
SRWL_EXPORT void srwlUtiSetAllocArrayFunc(char* (*pExtFunc)(char type, long long len));
/* reference srwlib.h:556 */
SRWL_EXPORT void srwlUtiSetProgrIndFunc(int (*pExtFunc)(double curVal));
/* reference srwlib.h:565 (srwlib.cpp:78-93) */
SRWL_EXPORT int srwlUtiVerNo(char* verNoStr, int code);
/* reference srwlib.h:574 (srwlib.cpp:119-140) */
SRWL_EXPORT int srwlUtiGetErrText(char* t, int erNo);

/* reference srwlib.h:607 (srwlib.cpp:328-373): RK4 trajectory, FP64 on the GPU */
SRWL_EXPORT int srwlCalcPartTraj(SRWLPrtTrj* pTrj, SRWLMagFldC* pMagFld, double* precPar);

/* reference srwlib.h:639 (srwlib.cpp:456-540): near-field SR electric field.
 * precPar: [0] method (0 manual, 1 adaptive; 2 not in this build), [1] step
 * [m] (method 0) or relative precision (method 1), [2], [3] integration
 * limits, [4] trajectory points, [5] terminating terms, [6] sampling factor:
 * > 0 sets the mesh sizes from the wavefront radius (srradstr.cpp:1063-1119)
 * and reallocates the field arrays through the callback of
 * srwlUtiSetWfrModifFunc (action 2) when they change; without one the call
 * returns SRWL_WFR_EXT_FUNC_NOT_DEFINED. */
SRWL_EXPORT int srwlCalcElecFieldSR(SRWLWfr* pWfr, SRWLPrtTrj* pTrj, SRWLMagFldC* pMagFld,
                                    double* precPar, int nPrecPar);

/* reference srwlib.h:748 (srwlib.cpp:757-856): intensity from field.
 * Implemented: pol 0..6, Stokes -1..-4 and -5 (all four, one after another),
 * intType 0 (single-electron intensity), 5 (Re E), 6 (Im E); depType 0 (vs e),
 * 1 (vs x), 2 (vs y), 3 (vs x, y), 4 (vs e, x), 5 (vs e, y), 6 (vs e, x, y).
 * Coordinates that are not varied are taken at e, x, y, interpolated
 * linearly between mesh points (clamped at the mesh ends).  Output: float,
 * photon energy fastest, then x, then y.  Other intType values (multi-electron,
 * flux, phase) return SRWL_AMD_ERR_UNSUPPORTED. */
SRWL_EXPORT int srwlCalcIntFromElecField(char* pInt, SRWLWfr* pWfr, char pol, char intType,
                                         char depType, double e, double x, double y,
                                         double* arMeth, void* pFldTrj, double* arParGPU);

/* reference srwlib.h:821 (srwlib.cpp:1023-1074): propagation through a
 * container of drift spaces ("drift"), thin lenses ("lens", SRWLOptL) and
 * rectangular / circular apertures and obstacles ("aperture", "obstacle",
 * SRWLOptA); arProp[3] = 0 (angular representation)
 * or 1, 2 (quadratic phase treated analytically with the wavefront radius
 * from the moments or from Rx, Ry); arProp[5..8] resize the mesh first.
 * Mesh sizes: even, 8..2048, with factors 2, 3, 5, 7, 11, 13 (SRW's own
 * fast-FFT sizes); others return SRWL_ERR_IN_FFT. */
SRWL_EXPORT int srwlPropagElecField(SRWLWfr* pWfr, SRWLOptC* pOpt, int nInt, char** arID,
                                    SRWLRadMesh* arIM, char** arI, double* arParGPU);

/* reference srwlib.h:775 (srwlib.cpp:860-900): resampling of the transverse
 * mesh, type 'c'; par = [0 (no FFT-side resizing), pxm, pxd, pzm, pzd, 0.5,
 * 0.5] (range and resolution factors, centred).  A size change reallocates
 * the field arrays through the callback of srwlUtiSetWfrModifFunc (action 2);
 * without one the call returns SRWL_WFR_EXT_FUNC_NOT_DEFINED. */
SRWL_EXPORT int srwlResizeElecField(SRWLWfr* pWfr, char type, double* par);

/* reference srwlib.h:787 (srwlib.cpp:914-931): resampling onto the
 * transverse mesh pMesh (nx, ny, xStart, xFin, yStart, yFin; the photon-energy
 * mesh must be the wavefront's own).  par = [0 (no FFT-side resizing),
 * allow the quadratic-phase treatment (1/0), allow the Re/Im correction from
 * the interpolated intensity (1/0)]; the reference's Python binding passes
 * {0, 1, 1} by default.  New points outside the old mesh are zero.  Arrays
 * are reallocated as for srwlResizeElecField.  Resampling in photon energy
 * returns SRWL_AMD_ERR_UNSUPPORTED; both at once, as in the reference,
 * SRWL_RESIZE_NOT_SUPPORTED_WITH_SUCH_PARAMS. */
SRWL_EXPORT int srwlResizeElecFieldMesh(SRWLWfr* pWfr, SRWLRadMesh* pMesh, double* par);

/* reference srwlib.h:871 (srwlib.cpp:1090-1204): SRW-convention 2D FFT of
 * float complex data; arMesh = [xStart, xStep, nx, yStart, yStep, ny].
 * nx, ny even, 8..2048, with factors 2, 3, 5, 7, 11, 13; other even sizes
 * return SRWL_AMD_ERR_UNSUPPORTED. */
SRWL_EXPORT int srwlUtiFFT(char* pcData, char typeData, double* arMesh, int nMesh, int dir,
                           double* arParGPU);

/* reference srwlib.h:998 (srwlib.cpp:1607-1641): GPU control (op 0: release,
 * 1: init; arParGPU = [n, deviceIndex]). */
SRWL_EXPORT int srwlUtiGPUProc(int op, double* arParGPU);

/* ---------------------------------------------------------------------- */
/* The rest of the reference C API (other sources, magnetic field and     */
/* transmission utilities, Stokes and power density, intensity            */
/* processing).  These are outside this library's path: they keep the    */
/* reference's names and signatures, so that a binding written against    */
/* the reference (its srwlpy.cpp) links unchanged, and each returns       */
/* SRWL_AMD_ERR_UNSUPPORTED without touching its arguments.  The          */
/* structures they take are declared here by tag only.                    */
/* ---------------------------------------------------------------------- */
typedef struct SRWLStructKickMatrix SRWLKickM;
typedef struct SRWLStructGaussianBeam SRWLGsnBm;
typedef struct SRWLStructPointSource SRWLPtSrc;
typedef struct SRWLStructOpticsTransmission SRWLOptT;

/* reference srwlib.h:590 */
SRWL_EXPORT int srwlCalcMagFld(SRWLMagFldC* pDispMagFld, SRWLMagFldC* pMagFld, double* precPar);
/* reference srwlib.h:619 */
SRWL_EXPORT int srwlCalcPartTrajFromKickMatr(SRWLPrtTrj* pTrj, SRWLKickM* arKickM, int nKickM, double* precPar);
/* reference srwlib.h:649 */
SRWL_EXPORT int srwlCalcElecFieldGaussian(SRWLWfr* pWfr, SRWLGsnBm* pGsnBm, double* precPar);
/* reference srwlib.h:659 */
SRWL_EXPORT int srwlCalcElecFieldPointSrc(SRWLWfr* pWfr, SRWLPtSrc* pPtSrc, double* precPar);
/* reference srwlib.h:675 */
SRWL_EXPORT int srwlCalcStokesUR(SRWLStokes* pStokes, SRWLPartBeam* pElBeam, SRWLMagFldU* pUnd, double* precPar);
/* reference srwlib.h:692 */
SRWL_EXPORT int srwlCalcPowDenSR(SRWLStokes* pStokes, SRWLPartBeam* pElBeam, SRWLPrtTrj* pTrj, SRWLMagFldC* pMagFld,
                                 double* precPar);
/* reference srwlib.h:802 */
SRWL_EXPORT int srwlProcElecField(SRWLWfr* pWfr, double* par, SRWLWfr* pWfr2);
/* reference srwlib.h:811 */
SRWL_EXPORT int srwlSetRepresElecField(SRWLWfr* pWfr, char repr);
/* reference srwlib.h:838 */
SRWL_EXPORT int srwlPropagRadMultiE(SRWLStokes* pStokes, SRWLWfr* pWfr0, SRWLOptC* pOpt, double* precPar,
                                    int (*pExtFunc)(int action, SRWLStokes* pStokesInst));
/* reference srwlib.h:851 */
SRWL_EXPORT int srwlCalcTransm(SRWLOptT* pOpTr, const double* pDelta, const double* pAttenLen, double** arObjShapeDefs,
                               int nObj3D, const double* arPar);
/* reference srwlib.h:893 */
SRWL_EXPORT int srwlUtiConvWithGaussian(char* pcData, char typeData, double* arMesh, int nMesh, double* arSig);
/* reference srwlib.h:920 */
SRWL_EXPORT int srwlUtiIntInf(double* arInf, char* pcData, char typeData, SRWLRadMesh* pMesh, double* arPar, int nPar);
/* reference srwlib.h:954 */
SRWL_EXPORT int srwlUtiIntProc(char* pcI1, char typeI1, SRWLRadMesh* pMesh1, char* pcI2, char typeI2,
                               SRWLRadMesh* pMesh2, double* arPar, int nPar);
/* reference srwlib.h:968 */
SRWL_EXPORT int srwlUtiUndFromMagFldTab(SRWLMagFldC* pUndCnt, SRWLMagFldC* pMagCnt, double* arPrecPar);
/* reference srwlib.h:986 */
SRWL_EXPORT int srwlUtiUndFindMagFldInterpInds(int* arResInds, int* pnResInds, double* arGaps, double* arPhases,
                                               int nVals, double arPrecPar[5]);

/* ---------------------------------------------------------------------- */
/* Batch entry points (no counterpart in the reference C API; they replace */
/* the per-electron loop of srwl_wfr_emit_prop_multi_e, srwlib.py:8506)    */
/* ---------------------------------------------------------------------- */

/* Computes nWfr independent single-electron fields in one device pass.
 * Each request has its own mesh, electron and magnetic field; precPar is
 * shared (method 0 or 1; precPar[6] > 0 is not supported here).  Equivalent
 * to calling srwlCalcElecFieldSR per item. */
SRWL_EXPORT int srwlCalcElecFieldSRBatch(SRWLWfr** arWfr, SRWLMagFldC** arMagFld, int nWfr,
                                         double* precPar, int nPrecPar);

/* Partially coherent intensity: nMacroElec electrons drawn from the Gaussian
 * beam pWfr0->partBeam (moments as in srwlib.py:9000-9029, mapping
 * srwlib.py:9793-9803; pWfr0's mesh sizes powers of two, 16..2048),
 * each emitting on pWfr0's mesh, propagated through pOpt (drifts) and
 * bilinearly re-gridded onto pStokes->mesh (srwlib.py avg_update_interp),
 * averaged into pStokes->arS0 (float32, nx*ny).
 * arPar: [0] seed, [1] electrons per device batch (0: auto), [2] proposal
 *   mode (0 plain Gaussian MC, 1 learned mixture proposal), [3] target
 *   rel-L2 for early stop (0: none), [4] training steps per batch, [5] (plain
 *   mode) > 0: estimate the variance gain of the optimal proposal (arOut[8]),
 *   [6] > 0: randomized quasi-Monte Carlo draws (digitally shifted Sobol'
 *   points; not with the proposal), [7] its number of independent replicates
 *   (default 16; electrons are used in whole groups of one per replicate).
 *   arPar must hold 8 values.
 * arOut (optional, >= 8 doubles; >= 9 with arPar[5] > 0, >= 10 with arPar[6]
 *   > 0): [0] electrons used, [1] effective sample size, [2] sum of weights,
 *   [3] device seconds, [4] rel-L2 error estimate (per-pixel sample variance;
 *   with replicates, the spread of the replicate averages), [5] batch size
 *   used, [6] last training loss, [7] training steps taken, [8] variance
 *   ratio plain / optimal proposal, [9] replicates. */
SRWL_EXPORT int srwlEmitPropMultiE(SRWLStokes* pStokes, SRWLWfr* pWfr0, SRWLMagFldC* pMagFld,
                                   SRWLOptC* pOpt, double* precPar, int nPrecPar,
                                   long long nMacroElec, double* arPar, double* arOut);

/* No counterpart in the reference: srwlCalcElecFieldSR for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, semantics and
 * results as srwlCalcElecFieldSR; every device buffer comes from a fixed
 * 64 MiB array per GPU (held until the process ends), so the call makes no
 * hipMalloc / hipFree / pinned or event allocation.  Limits: method 0,
 * ne*nx*ny <= 65536 field points, <= 20001 trajectory points, no automatic
 * mesh (precPar[6] <= 0); larger requests return SRWL_AMD_ERR_NOALLOC_SIZE.
 * srwlAmdNoAllocInit must have been called once on the device (it sets the
 * library up and looks up the array's device address); otherwise
 * SRWL_AMD_ERR_NOALLOC_INIT.  Calls on one GPU share its array and keep the
 * last call's contents: calls that overlap on one GPU race. */
SRWL_EXPORT int srwlAmdNoAllocInit(void);
SRWL_EXPORT int srwlCalcElecFieldSRNoAlloc(SRWLWfr* pWfr, SRWLPrtTrj* pTrj, SRWLMagFldC* pMagFld, double* precPar,
                                           int nPrecPar);

#ifdef __cplusplus
}
#endif
#endif
