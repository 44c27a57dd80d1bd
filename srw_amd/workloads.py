"""Standard workloads of the emit-and-propagate path.

Each builder returns SRWL* objects (srw_amd.srwlib) for one workload, named
by what it is.  The undulator field uses SRW's K-to-B conversion
K = 93.3729*lambda_u[m]*B[T] (reference srmagfld.cpp:166): K = 1.5 with a
20 mm period gives B = 0.80323 T.
"""
import math
from array import array

from .srwlib import (SRWLMagFldC, SRWLMagFldH, SRWLMagFldU, SRWLOptC, SRWLOptD, SRWLPartBeam,
                     SRWLParticle, SRWLStokes, SRWLWfr)

ELEC_REST_GEV = 0.51099890221e-03
HC_EV_M = 1.239842e-06
K2B = 93.37290417576577


def undulator_B(K, per):
    return K / (K2B * per)


def first_harmonic_eV(gamma, K, per):
    """Resonant first-harmonic photon energy on axis."""
    lam1 = per / (2. * gamma * gamma) * (1. + 0.5 * K * K)
    return HC_EV_M / lam1


def planar_undulator(K=1.5, per=0.02, nper=100, zc=0.):
    und = SRWLMagFldU([SRWLMagFldH(1, 'v', undulator_B(K, per), 0, 1, 1)], per, nper)
    return SRWLMagFldC([und], array('d', [0]), array('d', [0]), array('d', [zc]))


def electron_beam(E_GeV=3.0, Iavg=0.5, x=0., xp=0., y=0., yp=0., z=0.):
    b = SRWLPartBeam()
    b.Iavg = Iavg
    b.partStatMom1 = SRWLParticle(x, y, z, xp, yp, E_GeV / ELEC_REST_GEV, 1, -1)
    return b


def gaussian_beam_moments(b, sx=30e-6, sxp=5e-6, sy=5e-6, syp=2e-6, se=1e-3):
    """Second moments as srwlib.py arStatMom2 ([0]=<x^2>, [2]=<x'^2>, [3]=<y^2>, [5]=<y'^2>, [10]=(dE/E)^2)."""
    m = b.arStatMom2
    m[0] = sx * sx; m[1] = 0; m[2] = sxp * sxp
    m[3] = sy * sy; m[4] = 0; m[5] = syp * syp
    m[10] = se * se
    return b


def wavefront(beam, ne, e0, e1, nx, ny, half_x=1e-3, half_y=1e-3, z=20.):
    w = SRWLWfr()
    w.allocate(ne, nx, ny)
    w.mesh.zStart = z
    w.mesh.eStart, w.mesh.eFin = e0, e1
    w.mesh.xStart, w.mesh.xFin = -half_x, half_x
    w.mesh.yStart, w.mesh.yFin = -half_y, half_y
    w.partBeam = beam
    return w


def traj_prec(np_traj=10001, half_range=None, per=0.02, nper=100, step=None, termin=1):
    """precPar of srwlCalcElecFieldSR for method 0 with the step equal to the trajectory step."""
    if half_range is None:
        half_range = 0.5 * nper * per + 4 * per
    if step is None:
        step = 2. * half_range / (np_traj - 1)
    return array('d', [0, step, 0, 0, np_traj, termin, 0])


def single_electron(n=101, ne=1, de_rel=0.0, np_traj=10001, K=1.5, per=0.02, nper=100, z=20.,
                    half=1e-3, E_GeV=3.0):
    """Planar undulator, 3 GeV zero-emittance electron, n x n mesh over +/-half at z,
    ne photon energies spanning +/-de_rel around the first harmonic."""
    beam = electron_beam(E_GeV)
    g = beam.partStatMom1.gamma
    e1 = first_harmonic_eV(g, K, per)
    e0, ef = (e1 * (1 - de_rel), e1 * (1 + de_rel)) if ne > 1 else (e1, e1)
    mag = planar_undulator(K, per, nper)
    w = wavefront(beam, ne, e0, ef, n, n, half, half, z)
    return w, mag, traj_prec(np_traj, per=per, nper=nper)


def drift_container(L, mode=2):
    """Drift of length L; mode = arProp[3] (0: angular representation, 2: analytic quadratic phase)."""
    return SRWLOptC([SRWLOptD(L)], [[0, 0, 1.0, mode, 0, 1.0, 1.0, 1.0, 1.0]])


def stokes_like(w):
    s = SRWLStokes()
    s.allocate(1, w.mesh.nx, w.mesh.ny, _n_comp=1)
    s.mesh.set_from_other(w.mesh)
    s.mesh.ne = 1
    return s


def uneven_requests(n=2048, seed=7, meshes=(128, 256, 512, 1024), half=1e-3, z=20., np_traj=10001,
                    per=0.02, nper=100, E_GeV=3.0, alloc=True):
    """Independent single-electron requests with mixed meshes, harmonics and
    undulator strengths: mesh N uniform over `meshes`, K ~ U[0.5, 2.5],
    photon energy = h * E1(K) with h uniform over 1..5; zero-emittance 3 GeV
    electron, +/-half at z.  Returns (wavefronts, fields, precPar); with
    alloc=False the field arrays are one-element placeholders (for runs that
    leave the fields in device memory)."""
    import random
    rng = random.Random(seed)
    wfrs, mags = [], []
    for _ in range(n):
        N = rng.choice(meshes)
        K = rng.uniform(0.5, 2.5)
        h = rng.randint(1, 5)
        beam = electron_beam(E_GeV)
        e = h * first_harmonic_eV(beam.partStatMom1.gamma, K, per)
        if alloc:
            w = wavefront(beam, 1, e, e, N, N, half, half, z)
        else:
            w = SRWLWfr()
            w.mesh.ne, w.mesh.nx, w.mesh.ny = 1, N, N
            w.mesh.zStart = z
            w.mesh.eStart = w.mesh.eFin = e
            w.mesh.xStart, w.mesh.xFin = -half, half
            w.mesh.yStart, w.mesh.yFin = -half, half
            w.partBeam = beam
            w.arEx = array('f', [0., 0.])
            w.arEy = array('f', [0., 0.])
        wfrs.append(w)
        mags.append(planar_undulator(K, per, nper))
    return wfrs, mags, traj_prec(np_traj, per=per, nper=nper)


def beam_for_multi_e(E_GeV=3.0, Iavg=0.5):
    return gaussian_beam_moments(electron_beam(E_GeV, Iavg))


def sigma_check(gamma):
    return math.isfinite(gamma)
