"""Pins of the oracle's srwlCalcIntFromElecField restatement (every
dependence type, Re/Im E, polarisation components, Stokes parameters).

What the mathematics fixes, independently of the restatement's formulas:
  * Re E / Im E of a field that is linear in (e, x, y) are reproduced by the
    (bi)linear interpolation at any off-mesh point inside the mesh, and take
    the end value outside it (clamping, srradmnp.cpp:35-79);
  * the components obey I(2) + I(3) = I(4) + I(5) = I(0) + I(1) = s0,
    I(0) - I(1) = s1, I(2) - I(3) = s2, I(4) - I(5) = s3, s0^2 = s1^2 + s2^2 + s3^2
    for one electron;
  * the dependence types are slices of the full (e, x, y) extraction;
  * Re E of polarisation 2 is (Re Ex + Re Ey) / sqrt(2).
"""
from array import array

import numpy as np
import pytest

from srw_amd.srwlib import SRWLWfr
from tests import oracle_lib as oracle

NE, NX, NY = 5, 7, 6
E0, E1, X0, X1, Y0, Y1 = 100., 140., -2e-3, 1e-3, -1e-3, 2e-3
LENS = {0: NE, 1: NX, 2: NY, 3: NX * NY, 4: NE * NX, 5: NE * NY, 6: NE * NX * NY}


def _wfr(fx, fy):
    """Wavefront from complex arrays [ny][nx][ne]."""
    w = SRWLWfr()
    w.allocate(NE, NX, NY)
    w.mesh.eStart, w.mesh.eFin = E0, E1
    w.mesh.xStart, w.mesh.xFin = X0, X1
    w.mesh.yStart, w.mesh.yFin = Y0, Y1
    for arr, f in ((w.arEx, fx), (w.arEy, fy)):
        v = np.empty(2 * f.size, np.float32)
        v[0::2], v[1::2] = f.real.ravel(), f.imag.ravel()
        arr[:] = array('f', v.tobytes())
    return w


def _grid():
    e = np.linspace(E0, E1, NE)
    x = np.linspace(X0, X1, NX)
    y = np.linspace(Y0, Y1, NY)
    Y, X, E = np.meshgrid(y, x, e, indexing="ij")
    return E, X, Y


def _extract(w, pol, it, dep, e, x, y):
    out = array('f', [0] * (4 * LENS[dep]))
    assert oracle.intensity(out, w, pol, it, dep, e, x, y) == 0
    n = LENS[dep] * (4 if pol == -5 else 1)
    return np.frombuffer(out, np.float32)[:n].astype(np.float64)


def _linear():
    E, X, Y = _grid()
    re = 1.0 + 0.01 * (E - E0) + 300. * X - 200. * Y
    im = -0.5 + 0.02 * (E - E0) - 100. * X + 400. * Y
    return re + 1j * im


def _lin_val(e, x, y):
    return (1.0 + 0.01 * (e - E0) + 300. * x - 200. * y, -0.5 + 0.02 * (e - E0) - 100. * x + 400. * y)


@pytest.mark.parametrize("dep", range(7))
@pytest.mark.parametrize("pt", [(117.3, -4.1e-4, 7.7e-4), (90., 5e-3, -3e-3), (150., -9e-3, 9e-3)])
def test_linear_field_interpolation(dep, pt):
    """Re/Im E of a linear field at (e, x, y): exact inside the mesh, the
    clamped end value outside, on every dependence type."""
    f = _linear()
    w = _wfr(f, np.zeros_like(f))
    e, x, y = pt
    ec, xc, yc = np.clip(e, E0, E1), np.clip(x, X0, X1), np.clip(y, Y0, Y1)
    ev, xv, yv = np.linspace(E0, E1, NE), np.linspace(X0, X1, NX), np.linspace(Y0, Y1, NY)
    grids = {0: (ev, xc, yc), 1: (ec, xv, yc), 2: (ec, xc, yv)}
    if dep in grids:
        ee, xx, yy = grids[dep]
    elif dep == 3:
        yy, xx = np.meshgrid(yv, xv, indexing="ij"); ee = ec
    elif dep == 4:
        xx, ee = np.meshgrid(xv, ev, indexing="ij"); yy = yc
    elif dep == 5:
        yy, ee = np.meshgrid(yv, ev, indexing="ij"); xx = xc
    else:
        ee, xx, yy = _grid()
    re_t, im_t = _lin_val(np.ravel(ee) * np.ones(LENS[dep]), np.ravel(xx) * np.ones(LENS[dep]),
                          np.ravel(yy) * np.ones(LENS[dep]))
    re = _extract(w, 0, 5, dep, e, x, y)
    im = _extract(w, 0, 6, dep, e, x, y)
    np.testing.assert_allclose(re, re_t, rtol=0, atol=2e-5)
    np.testing.assert_allclose(im, im_t, rtol=0, atol=2e-5)


def test_components_and_stokes_identities():
    rng = np.random.default_rng(3)
    shp = (NY, NX, NE)
    fx = rng.normal(size=shp) + 1j * rng.normal(size=shp)
    fy = rng.normal(size=shp) + 1j * rng.normal(size=shp)
    w = _wfr(fx, fy)
    g = lambda pol, it=0: _extract(w, pol, it, 6, E0, X0, Y0)
    s0, s1, s2, s3 = g(-1), g(-2), g(-3), g(-4)
    ex, ey = fx.ravel(), fy.ravel()
    np.testing.assert_allclose(s0, np.abs(ex) ** 2 + np.abs(ey) ** 2, rtol=1e-5)
    np.testing.assert_allclose(s3, 2 * np.imag(np.conj(ex) * ey), rtol=1e-4, atol=1e-5)
    tol = dict(rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(g(2) + g(3), s0, **tol)
    np.testing.assert_allclose(g(4) + g(5), s0, **tol)
    np.testing.assert_allclose(g(0) + g(1), s0, **tol)
    np.testing.assert_allclose(g(6), s0, **tol)
    np.testing.assert_allclose(g(0) - g(1), s1, **tol)
    np.testing.assert_allclose(g(2) - g(3), s2, **tol)
    np.testing.assert_allclose(g(4) - g(5), s3, **tol)
    np.testing.assert_allclose(s0 ** 2, s1 ** 2 + s2 ** 2 + s3 ** 2, rtol=1e-4)
    np.testing.assert_allclose(g(2, 5), (ex.real + ey.real) / np.sqrt(2), rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(g(1, 6), ey.imag.astype(np.float32))
    allst = g(-5)
    np.testing.assert_array_equal(allst, np.concatenate([s0, s1, s2, s3]))


@pytest.mark.parametrize("pol", [6, -3])
def test_dependences_are_slices_on_mesh(pol):
    """At mesh points every dependence type is a slice of the 3D extraction."""
    rng = np.random.default_rng(5)
    shp = (NY, NX, NE)
    w = _wfr(rng.normal(size=shp) + 1j * rng.normal(size=shp), rng.normal(size=shp) + 1j * rng.normal(size=shp))
    ev, xv, yv = np.linspace(E0, E1, NE), np.linspace(X0, X1, NX), np.linspace(Y0, Y1, NY)
    ie, ix, iy = 3, 2, 4
    full = _extract(w, pol, 0, 6, 0, 0, 0).reshape(NY, NX, NE)
    e, x, y = ev[ie], xv[ix], yv[iy]
    tol = dict(rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(_extract(w, pol, 0, 0, e, x, y), full[iy, ix, :], **tol)
    np.testing.assert_allclose(_extract(w, pol, 0, 1, e, x, y), full[iy, :, ie], **tol)
    np.testing.assert_allclose(_extract(w, pol, 0, 2, e, x, y), full[:, ix, ie], **tol)
    np.testing.assert_allclose(_extract(w, pol, 0, 3, e, x, y), full[:, :, ie].ravel(), **tol)
    np.testing.assert_allclose(_extract(w, pol, 0, 4, e, x, y), full[iy, :, :].ravel(), **tol)
    np.testing.assert_allclose(_extract(w, pol, 0, 5, e, x, y), full[:, ix, :].ravel(), **tol)
