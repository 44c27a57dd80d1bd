"""CPU pins of the oracle's wavefront moments (SRW ComputeRadMoments) and of
the moment / 4x4-matrix updates of a drift, on Gaussian beams whose moments
are known in closed form."""
import ctypes as C
import math
from array import array

import numpy as np
import pytest
from scipy.stats import norm

from srw_amd import _marshal as M
from srw_amd import workloads as W
from srw_amd.srwlib import SRWLWfr
from tests import oracle_lib as oracle

E_EV = 1000.
LAM = 1.239842e-06 / E_EV


def _beam(n, half, w0, z0=0., tilt=0., R=1e23, dR=0.):
    """TEM00 field z0 past its waist (diverging-wave phase convention), tilted by `tilt` in x."""
    w = SRWLWfr()
    w.allocate(1, n, n)
    w.mesh.eStart = w.mesh.eFin = E_EV
    w.mesh.xStart, w.mesh.xFin = -half, half * (n - 2) / n
    w.mesh.yStart, w.mesh.yFin = -half, half * (n - 2) / n
    x = np.linspace(w.mesh.xStart, w.mesh.xFin, n)
    X, Y = np.meshgrid(x, x)
    zR = math.pi * w0 * w0 / LAM
    wz = w0 * math.sqrt(1 + (z0 / zR) ** 2)
    invR = z0 / (z0 * z0 + zR * zR)
    r2 = X * X + Y * Y
    E = (w0 / wz) * np.exp(-r2 / wz ** 2 + 1j * (math.pi * r2 * invR / LAM + 2 * math.pi * tilt * X / LAM))
    w.arEx = array('f', E.astype(np.complex64).view(np.float32).ravel().tobytes())
    w.arEy = array('f', (0.5j * E).astype(np.complex64).view(np.float32).ravel().tobytes())
    w.Rx = w.Ry = R
    w.dRx = w.dRy = dR
    return w, x, wz


def _moments(w):
    keep = M.Keep()
    cw = M.wfr(w, keep)
    oracle.lib.orc_compute_rad_moments.argtypes = [C.POINTER(M.S.SRWLWfr), C.c_void_p, C.c_void_p]
    oracle.lib.orc_compute_rad_moments(C.byref(cw), M._addr(w.arMomX), M._addr(w.arMomY))
    return np.array(w.arMomX[:11]), np.array(w.arMomY[:11])


def test_total_flux_and_tilt():
    w0, theta = 40e-6, 3e-6
    w, x, _ = _beam(256, 0.4e-3, w0, tilt=theta)
    mx, mz = _moments(w)
    # integral of |E|^2 over the plane = pi w0^2 / 2, reported per mm^2
    assert mx[0] == pytest.approx(math.pi * w0 * w0 / 2 * 1e6, rel=1e-4)
    assert mz[0] == pytest.approx(0.25 * mx[0], rel=1e-6)
    assert abs(mx[1]) < 1e-4 * w0 and abs(mx[3]) < 1e-4 * w0
    # backward-difference phase slope of a linear phase, (lambda / 2 pi h) sin(2 pi theta h / lambda),
    # times sum A(x) A(x-h) / sum A^2 = exp(-h^2 / 2 w0^2) for the Gaussian amplitude
    h = x[1] - x[0]
    want = LAM / (2 * math.pi * h) * math.sin(2 * math.pi * theta * h / LAM) * math.exp(-h * h / (2 * w0 * w0))
    assert mx[2] == pytest.approx(want, rel=1e-5)
    assert abs(mx[4]) < 1e-3 * theta


def _trunc_second_moment(x, sigma):
    """Second moments restricted to SRW's 90 % power rectangle for a separable
    Gaussian of rms sigma: rows keep all but 2.5 % of the total power on each
    side; columns, within those rows, all but 2.5 % of the TOTAL power per side.
    Returns (sum over rectangle of x^2 I / sum of I) on the grid x (both axes)."""
    p = norm.pdf(x / sigma)
    p = p / p.sum()
    c = np.cumsum(p)
    i0 = int(np.argmax(c > 0.025))
    i1 = len(x) - 1 - int(np.argmax(np.cumsum(p[::-1]) > 0.025))
    frac_rows = p[i0:i1 + 1].sum()
    q = p * frac_rows
    j0 = int(np.argmax(np.cumsum(q) > 0.025))
    j1 = len(x) - 1 - int(np.argmax(np.cumsum(q[::-1]) > 0.025))
    return frac_rows * np.sum(x[j0:j1 + 1] ** 2 * p[j0:j1 + 1]), frac_rows


def test_second_moments_of_diverging_gaussian_beam():
    # 10 m past a 20 um waist: local angle x' = x/R, so <xx'> = <xx>/R and
    # <x'x'> = <xx> ((lambda/(pi wz^2))^2 + 1/R^2) on any region (here the 90 % rectangle)
    w0, z0 = 20e-6, 10.
    zR = math.pi * w0 * w0 / LAM
    R = z0 * (1 + (zR / z0) ** 2)
    w, x, wz = _beam(512, 1.0e-3, w0, z0=z0, R=R, dR=0.01 * R)
    mx, _ = _moments(w)
    xx, xxp, xpxp = mx[5], mx[6], mx[7]
    assert xxp == pytest.approx(xx / R, rel=2e-3)
    assert xpxp == pytest.approx(xx * ((LAM / (math.pi * wz * wz)) ** 2 + 1 / R ** 2), rel=5e-3)
    # truncation to the 90 % power rectangle (sigma of the intensity = wz/2)
    want, _ = _trunc_second_moment(x, wz / 2)
    assert xx == pytest.approx(want, rel=2e-2)
    assert mx[8] == pytest.approx(want, rel=2e-2)
    # and not the untruncated moment
    assert xx < 0.8 * (wz / 2) ** 2


def test_moment_pass_restores_the_field():
    # the phase is removed and added back in place (float storage): field unchanged to float precision
    w0, z0 = 20e-6, 10.
    zR = math.pi * w0 * w0 / LAM
    R = z0 * (1 + (zR / z0) ** 2)
    w, _, _ = _beam(128, 1.0e-3, w0, z0=z0, R=R, dR=0.01 * R)
    before = np.frombuffer(w.arEx, np.float32).copy()
    _moments(w)
    after = np.frombuffer(w.arEx, np.float32)
    assert np.linalg.norm(after - before) < 2e-6 * np.linalg.norm(before)


def test_drift_updates_moments_radius_and_matrix():
    w0, theta, L = 40e-6, 3e-6, 8.
    w, _, _ = _beam(256, 1.6e-3, w0, tilt=theta)
    mx0, _ = _moments(w)
    for i in range(20):
        w.arElecPropMatr[i] = 0.
    for i in (0, 5, 10, 15):
        w.arElecPropMatr[i] = 1.
    w.arElecPropMatr[17] = 2e-6      # angle x' of the reference vector
    R0 = w.Rx
    assert oracle.propagate(w, W.drift_container(L, 0)) == 0
    m = np.array(w.arMomX[:11])
    assert m[1] == pytest.approx(mx0[1] + L * mx0[2], rel=1e-12, abs=1e-18)
    assert m[2] == mx0[2]
    assert m[5] == pytest.approx(mx0[5] + 2 * L * mx0[6] + L * L * mx0[7], rel=1e-12)
    assert w.Rx == R0 + L
    P = np.array(w.arElecPropMatr)
    assert P[1] == pytest.approx(L) and P[11] == pytest.approx(L) and P[0] == 1. and P[5] == 1.
    assert P[16] == pytest.approx(L * 2e-6) and P[17] == pytest.approx(2e-6)
    # the analytic first moments agree with the moments of the propagated field
    w.Rx = w.Ry = 1e23
    mx1, _ = _moments(w)
    assert mx1[1] == pytest.approx(L * theta, rel=2e-3)


def test_subtype1_radius_follows_moments():
    # arProp[3] = 1: radius s1 + smin^2/s1 with smin = lambda/(4 pi (0.2 sigma_x')^2); here
    # sigma_x' is small, so smin >> |s1| and the first branch gives R_eff = 2 smin
    w0, z0, L = 20e-6, 10., 20.
    zR = math.pi * w0 * w0 / LAM
    R = z0 * (1 + (zR / z0) ** 2)
    w, _, _ = _beam(128, 1.2e-3, w0, z0=z0, R=R, dR=0.01 * R)
    mx, _ = _moments(w)
    sxp = math.sqrt(max(mx[7] - mx[2] ** 2, 0.1 * mx[7]))
    smin = LAM / (4 * 3.14159265359 * (0.2 * sxp) ** 2)
    want = R + smin * smin / R if abs(R) >= smin else 2 * smin
    dx0 = (w.mesh.xFin - w.mesh.xStart) / 127
    assert oracle.propagate(w, W.drift_container(L, 1)) == 0
    dx1 = (w.mesh.xFin - w.mesh.xStart) / 127
    assert dx1 / dx0 == pytest.approx((want + L) / want, rel=1e-9)
