"""CPU tests that pin the FP64 oracle with closed forms, invariants, brute
force and independent solvers (no GPU needed).

The reference's own tests hold no numeric golden values for this path (they
only assert array lengths of example scripts), so every oracle function is
pinned here by mathematics that a plausible implementation mistake would
violate.
"""
import ctypes as C
import math
from array import array

import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.special import jv

from srw_amd import _marshal as M
from srw_amd import _structs as S
from srw_amd import workloads as W
from srw_amd.srwlib import SRWLPrtTrj, SRWLWfr
from tests import oracle_lib as oracle
from tests.util import cfield, clone_wfr, rel_l2

K, PER, NPER = 1.5, 0.02, 100


def test_struct_layouts_match_c_compiler():
    sizes = (C.c_long * 11)()
    oracle.lib.orc_struct_sizes(sizes)
    assert list(sizes) == [C.sizeof(t) for t in S.STRUCTS]


def test_philox_known_answers():
    # Random123 known-answer vectors for Philox4x32-10
    def ph(ctr, k):
        c = (C.c_uint * 4)(*ctr)
        o = (C.c_uint * 4)()
        oracle.lib.orc_philox(c, C.c_uint(k[0]), C.c_uint(k[1]), o)
        return list(o)
    assert ph([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert ph([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert ph([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_field_peak_period_and_extent():
    mag = W.planar_undulator(K, PER, NPER)
    B0 = W.undulator_B(K, PER)
    assert oracle.field_B(mag, 0, 0, 0)[1] == pytest.approx(B0, rel=1e-12)          # cos(0) at the centre
    for z in (0.1234, -0.5, 0.77):
        assert oracle.field_B(mag, 0, 0, z)[1] == pytest.approx(oracle.field_B(mag, 0, 0, z + PER)[1], abs=1e-9)
    assert oracle.field_B(mag, 0, 0, 1.0801) == [0., 0., 0.]                          # beyond N*per/2 + 4 per
    # vertical offset: cosh(k y) dependence
    y = 3e-4
    assert oracle.field_B(mag, 0, y, 0)[1] == pytest.approx(B0 * math.cosh(2 * math.pi / 6.2831853072 * 6.2831853072 /
                                                                               PER * y), rel=1e-9)


def _traj(np_=2001, x=0., xp=0., y=0., yp=0.):
    beam = W.electron_beam(3.0, 0.5, x, xp, y, yp)
    t = SRWLPrtTrj()
    t.allocate(np_)
    t.partInitCond = beam.partStatMom1
    return t


def _traj_vs_scipy(mag, np_):
    t = _traj(np_, 2e-6, 3e-6, -1e-6, 1e-6)
    assert oracle.calc_traj(t, mag) == 0
    s = np.linspace(t.ctStart, t.ctEnd, t.np)
    g = t.partInitCond.gamma * 0.51099890221e-03 * 1000. / 0.511003414
    mult = -1.602176462e-19 / (9.10938188e-31 * 2.99792458e+08 * g)

    def rhs(_s, F):        # Lorentz force in s, (x, x', y, y', z) with z' = sqrt(1 - 1/g^2 - x'^2 - y'^2)
        xd, yd = F[1], F[3]
        zd = math.sqrt(1 - (1 / g ** 2 + xd * xd + yd * yd))
        B = oracle.field_B(mag, F[0], F[2], F[4])
        return [xd, mult * (yd * B[2] - zd * B[1]), yd, mult * (zd * B[0] - xd * B[2]), zd]
    p = t.partInitCond
    y0 = [p.x, p.xp, p.y, p.yp, p.z]
    r = solve_ivp(rhs, (0, t.ctEnd), y0, rtol=1e-12, atol=1e-18, t_eval=s[s >= 0], method="DOP853")
    l = solve_ivp(rhs, (0, t.ctStart), y0, rtol=1e-12, atol=1e-18, t_eval=s[s < 0][::-1], method="DOP853")
    ref = np.concatenate([l.y[:, ::-1], r.y], axis=1)
    got = np.array([t.arX, t.arXp, t.arY, t.arYp, t.arZ])
    return t, s, g, np.max(np.abs(got - ref), axis=1)


def test_trajectory_is_fourth_order_runge_kutta_of_lorentz_force():
    mag = W.planar_undulator(K, PER, NPER)
    t, s, g, e1 = _traj_vs_scipy(mag, 4001)
    _, _, _, e2 = _traj_vs_scipy(mag, 8001)
    defl = K / t.partInitCond.gamma
    # global error of a 4th-order method drops ~16x when the step halves
    assert 10. < e1[1] / e2[1] < 40.
    assert e2[1] < 5e-8 * defl and e2[0] < 1e-12
    # deflection amplitude K/gamma (closed form of a planar undulator)
    mid = np.abs(s) < 0.5
    assert np.max(np.abs(np.array(t.arXp)[mid] - t.partInitCond.xp)) == pytest.approx(defl, rel=2e-3)
    # speed invariant |beta|^2 = 1 - 1/gamma^2
    b2 = np.array(t.arXp) ** 2 + np.array(t.arYp) ** 2 + np.array(t.arZp) ** 2
    assert np.max(np.abs(b2 - (1 - 1 / g ** 2))) < 1e-15


def test_interpolation_is_exact_for_quintic_and_integrates_squares():
    n = 301
    t = _traj(n)
    t.ctStart, t.ctEnd = -1.0, 1.0
    s = np.linspace(-1, 1, n)
    c = np.array([1e-6, 2e-6, -3e-6, 1e-6, 5e-7, -2e-7])
    X = np.polyval(c[::-1], s)
    dX = np.polyval(np.polyder(c[::-1]), s)
    for i in range(n):
        t.arX[i] = X[i]
        t.arXp[i] = dX[i]
    t.partInitCond.z = 0.
    q = np.linspace(-0.997, 0.993, 57)
    out = (C.c_double * (8 * len(q)))()
    kk = M.Keep()
    ct = M.prt_trj(t, kk)
    assert oracle.lib.orc_trj_interp(C.byref(ct), (C.c_double * len(q))(*q), len(q), out) == 0
    o = np.array(out).reshape(-1, 8)
    assert np.max(np.abs(o[:, 1] - np.polyval(c[::-1], q))) < 1e-18       # 6-point quintic fit is exact
    assert np.max(np.abs(o[:, 0] - np.polyval(np.polyder(c[::-1]), q))) < 1e-15
    # integral of x'^2 from s0 = 0: compare with the exact polynomial integral
    p2 = np.polyint(np.polymul(np.polyder(c[::-1]), np.polyder(c[::-1])))
    exact = np.polyval(p2, q) - np.polyval(p2, 0.)
    # exact up to Simpson's error beyond the first interval; the first interval
    # starts from f0 = 0 instead of the squared first angle (srtrjdat.cpp:1783-1793)
    inner = q > -1 + 2 * 2. / (n - 1)
    assert np.max(np.abs(o[inner, 2] - exact[inner])) < 1e-8 * np.max(np.abs(exact))
    assert np.max(np.abs(o[:, 2] - exact)) < 1e-3 * np.max(np.abs(exact))


def test_on_axis_far_field_matches_kim_formula():
    # Far-field on-axis flux density of a planar undulator at the first harmonic
    # (K.-J. Kim, AIP Conf. Proc. 184 (1989)): d2F/dOmega = alpha N^2 gamma^2 (dw/w) (I/e) F1(K),
    # F1 = K^2/(1+K^2/2)^2 [J0(xi)-J1(xi)]^2, xi = K^2/(4+2K^2), per rad^2; at distance L [m]
    # one mm^2 subtends 1e-6/L^2 rad^2.  The field ends add a fixed fraction of a
    # period, so the ratio to the formula is 1 + c/N: the N -> infinity limit,
    # 2 r(2N) - r(N), must be 1.
    L = 2000.
    xi = K * K / (4 + 2 * K * K)
    F1 = K * K / (1 + K * K / 2) ** 2 * (jv(0, xi) - jv(1, xi)) ** 2
    alpha, e = 1 / 137.035999, 1.602176634e-19
    r = []
    for nper in (NPER, 2 * NPER):
        w, mag, prec = W.single_electron(n=3, z=L, half=1e-7, nper=nper, np_traj=100 * nper + 1)
        assert oracle.calc_field(w, mag, prec, nthreads=4) == 0
        flux_mm2 = abs(cfield(w.arEx, w)[1, 1, 0]) ** 2
        g = w.partBeam.partStatMom1.gamma
        r.append(flux_mm2 / (alpha * nper ** 2 * g * g * 1e-3 * (w.partBeam.Iavg / e) * F1 * 1e-6 / (L * L)))
    assert 0.005 < r[1] - 1 < 0.05
    assert 2 * r[1] - r[0] == pytest.approx(1., abs=2e-3)


def test_quadrature_converges_and_matches_brute_force_without_end_terms():
    # method 0 at the trajectory step vs the same integral with a 4x finer step
    w, mag, prec = W.single_electron(n=9)
    prec[5] = 0
    w2 = clone_wfr(w)
    assert oracle.calc_field(w, mag, prec, nthreads=4) == 0
    prec2 = array('d', prec)
    prec2[1] = prec[1] / 4.
    assert oracle.calc_field(w2, mag, prec2, nthreads=4) == 0
    a, b = cfield(w.arEx, w), cfield(w2.arEx, w2)
    assert rel_l2(a, b) < 1e-4


def test_mesh_symmetry_shortcut_equals_full_computation():
    w, mag, prec = W.single_electron(n=11)
    w2 = clone_wfr(w)
    assert oracle.calc_field(w, mag, prec, nthreads=4) == 0           # mirrored half mesh
    w2.partBeam.partStatMom1.y = 1e-15                                 # breaks the shortcut, not the physics
    assert oracle.calc_field(w2, mag, prec, nthreads=4) == 0
    assert rel_l2(cfield(w.arEx, w), cfield(w2.arEx, w2)) < 1e-6
    ey = cfield(w2.arEy, w2)[:, :, 0]
    ex = cfield(w2.arEx, w2)[:, :, 0]
    assert rel_l2(ex[::-1, :], ex) < 1e-6                               # Ex even in y
    assert np.linalg.norm(ey + ey[::-1, :]) <= 1e-3 * np.linalg.norm(ex)  # Ey odd in y


@pytest.mark.parametrize("n,ny,xs,dirn", [(8, 6, None, 1), (16, 8, 3e-5, 1), (12, 10, -2e-5, -1), (6, 4, 1e-5, -1)])
def test_fft_equals_brute_force_fourier_integral(n, ny, xs, dirn):
    rng = np.random.default_rng(n * 7 + ny)
    f = (rng.standard_normal((ny, n)) + 1j * rng.standard_normal((ny, n))).astype(np.complex64)
    dx, dy = 1e-5, 1.5e-5
    x0 = -0.5 * n * dx if xs is None else xs
    y0 = -0.5 * ny * dy
    data = array('f', f.view(np.float32).ravel().tobytes())
    mesh = [x0, dx, n, y0, dy, ny]
    assert oracle.uti_fft(data, mesh, dirn) == 0
    g = np.frombuffer(data, np.float32).view(np.complex64).reshape(ny, n)
    # brute force F(qx, qy) = dx dy sum f(x, y) exp(-+ i 2pi (qx x + qy y)) on the output mesh
    x = x0 + dx * np.arange(n)
    y = y0 + dy * np.arange(ny)
    qx = mesh[0] + mesh[1] * np.arange(n)
    qy = mesh[3] + mesh[4] * np.arange(ny)
    assert mesh[0] == pytest.approx(-0.5 / dx) and mesh[1] == pytest.approx(1 / (n * dx))
    sgn = -1 if dirn > 0 else 1
    Ex = np.exp(sgn * 2j * np.pi * np.outer(qx, x))
    Ey = np.exp(sgn * 2j * np.pi * np.outer(qy, y))
    ref = dx * dy * (Ey @ f.astype(np.complex128) @ Ex.T)
    assert rel_l2(g, ref) < 2e-6


def _gauss(x, y, z, w0, lam):
    """TEM00 field at distance z from its waist, in the convention of the SR
    integral (a diverging wave carries the phase +pi r^2/(lambda R))."""
    zR = math.pi * w0 * w0 / lam
    wz = w0 * math.sqrt(1 + (z / zR) ** 2)
    invR = z / (z * z + zR * zR)
    r2 = x * x + y * y
    return (w0 / wz) * np.exp(-r2 / wz ** 2 + 1j * (math.pi * r2 * invR / lam - math.atan(z / zR)))


def _gauss_wfr(n, w0, e, half, z0=0.):
    lam = 1.239842e-06 / e
    w = SRWLWfr()
    w.allocate(1, n, n)
    w.mesh.eStart = w.mesh.eFin = e
    w.mesh.xStart, w.mesh.xFin = -half, half * (n - 2) / n        # FFT centre on x = 0
    w.mesh.yStart, w.mesh.yFin = -half, half * (n - 2) / n
    x = np.linspace(w.mesh.xStart, w.mesh.xFin, n)
    X, Y = np.meshgrid(x, x)
    E = _gauss(X, Y, z0, w0, lam)
    w.arEx = array('f', E.astype(np.complex64).view(np.float32).ravel().tobytes())
    w.arEy = array('f', (0.5j * E).astype(np.complex64).view(np.float32).ravel().tobytes())
    return w, lam


def _grid(w):
    x = np.linspace(w.mesh.xStart, w.mesh.xFin, w.mesh.nx)
    y = np.linspace(w.mesh.yStart, w.mesh.yFin, w.mesh.ny)
    return np.meshgrid(x, y)


def test_drift_angular_representation_matches_gaussian_beam():
    # drift of L from the waist: exact Fresnel solution of a TEM00 beam
    w0, L, n = 60e-6, 8., 256
    w, lam = _gauss_wfr(n, w0, 1000., 1.6e-3)
    P0 = np.sum(np.abs(cfield(w.arEx, w)) ** 2)
    assert oracle.propagate(w, W.drift_container(L, 0)) == 0
    X, Y = _grid(w)
    Ex, Ey = cfield(w.arEx, w)[:, :, 0], cfield(w.arEy, w)[:, :, 0]
    ref = _gauss(X, Y, L, w0, lam)
    assert rel_l2(Ex, ref) < 1e-4
    assert rel_l2(Ey, 0.5j * ref) < 1e-4
    assert np.sum(np.abs(Ex) ** 2) == pytest.approx(P0, rel=1e-5)      # unitary (Parseval)


def test_drift_analytic_quadratic_phase_matches_gaussian_beam():
    # beam 10 m after its waist (radius R known), drifted 20 m with the analytic
    # quadratic-phase treatment: output mesh scaled by (R+L)/R, exact Fresnel solution
    w0, z0, L, n = 20e-6, 10., 20., 128
    w, lam = _gauss_wfr(n, w0, 1000., 1.2e-3, z0)
    zR = math.pi * w0 * w0 / lam
    R0 = z0 * (1 + (zR / z0) ** 2)
    w.Rx = w.Ry = R0
    dx0 = (w.mesh.xFin - w.mesh.xStart) / (n - 1)
    P0 = np.sum(np.abs(cfield(w.arEx, w)) ** 2) * dx0 * dx0
    assert oracle.propagate(w, W.drift_container(L, 2)) == 0
    dx1 = (w.mesh.xFin - w.mesh.xStart) / (n - 1)
    assert dx1 == pytest.approx(dx0 * (R0 + L) / R0, rel=1e-12)
    assert w.Rx == pytest.approx(R0 + L)
    X, Y = _grid(w)
    Ex = cfield(w.arEx, w)[:, :, 0]
    assert rel_l2(Ex, _gauss(X, Y, z0 + L, w0, lam)) < 1e-4
    assert np.sum(np.abs(Ex) ** 2) * dx1 * dx1 == pytest.approx(P0, rel=1e-5)


def test_beam_sampler_reproduces_second_moments():
    # 6-D Gaussian beam with x-x' correlation: sample moments of the mapped
    # electrons match arStatMom2 within their statistical errors
    beam = W.gaussian_beam_moments(W.electron_beam(), sx=30e-6, sxp=5e-6, sy=5e-6, syp=2e-6, se=1e-3)
    beam.arStatMom2[1] = -0.4 * 30e-6 * 5e-6          # <x x'>
    cb = M.part_beam(beam)
    p = S.SRWLParticle()
    n = 40000
    a = np.empty((n, 5))
    for i in range(n):
        oracle.lib.orc_beam_sample(C.byref(cb), 20251202, i, C.byref(p))
        a[i] = (p.x, p.xp, p.y, p.yp, p.gamma)
    g0 = beam.partStatMom1.gamma
    m = beam.arStatMom2
    cov = np.cov(a[:, :4].T)
    rel = 4 * math.sqrt(2. / n)                       # 4 sigma of a sample variance
    assert cov[0, 0] == pytest.approx(m[0], rel=rel)
    assert cov[1, 1] == pytest.approx(m[2], rel=rel)
    assert cov[2, 2] == pytest.approx(m[3], rel=rel)
    assert cov[3, 3] == pytest.approx(m[5], rel=rel)
    assert cov[0, 1] == pytest.approx(m[1], abs=4 * math.sqrt(m[0] * m[2] / n))
    assert abs(cov[0, 2]) < 4 * math.sqrt(m[0] * m[3] / n)
    assert np.var(a[:, 4] / g0 - 1) == pytest.approx(m[10], rel=rel)
    assert abs(np.mean(a[:, 0])) < 4 * math.sqrt(m[0] / n)


def test_sobol_points_are_nets_and_inverse_normal_is_exact():
    """Sobol' points of the RQMC sampler: every dimension alone is a
    (0,1)-sequence (each of the 2^m intervals of width 2^-m holds one of the
    first 2^m points), dimensions 1-2 form a (0,2)-sequence (each elementary
    box of area 2^-8 holds one of the first 256 points); the inverse normal
    inverts Phi = erfc(-g/sqrt 2)/2."""
    import ctypes as C
    import math
    import os
    import subprocess
    import tempfile
    import numpy as np
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = r'''
#include <stdio.h>
#include "srw_rng.h"
int main(void) {
    uint32_t V[SRW_SOBOL_DIMS][32];
    srw_sobol_directions(V);
    for (unsigned k = 0; k < 1024; k++)
        for (int d = 0; d < SRW_SOBOL_DIMS; d++) printf("%u\n", srw_sobol_word(V, d, k));
    double us[] = {1e-9, 1e-4, 0.02, 0.31, 0.5, 0.77, 0.999, 1 - 1e-9};
    for (int i = 0; i < 8; i++) printf("%.17g %.17g\n", us[i], srw_norm_inv(us[i]));
    return 0;
}
'''
    with tempfile.TemporaryDirectory() as td:
        c, exe = os.path.join(td, "s.c"), os.path.join(td, "s")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-O2", "-I", os.path.join(root, "include"), c, "-o", exe, "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    w = np.array([int(x) for x in out[:1024 * 5]], dtype=np.uint64).reshape(1024, 5)
    for m in (4, 7, 10):
        for d in range(5):
            cells = w[:2 ** m, d] >> (32 - m)
            assert len(set(cells.tolist())) == 2 ** m
    for j in range(9):
        box = (w[:256, 0] >> (32 - j)) * (1 << (8 - j)) + (w[:256, 1] >> (32 - (8 - j)) if j < 8 else 0)
        assert len(set(box.tolist())) == 256
    for line in out[1024 * 5:1024 * 5 + 8]:
        u, g = (float(x) for x in line.split())
        assert 0.5 * math.erfc(-g / math.sqrt(2)) == pytest.approx(u, rel=1e-13)
