"""CPU pins of the oracle's wavefront resampling (srwlResizeElecField, 'c')."""
import math
from array import array

import numpy as np
import pytest

from srw_amd.srwlib import SRWLWfr
from tests import oracle_lib as oracle
from tests.util import cfield, rel_l2

E_EV = 1000.
LAM = 1.239842e-06 / E_EV


def test_good_fft_sizes():
    f = oracle.lib.orc_next_fft_size
    # values of the reference's table of good sizes (gmfft.cpp:26-150)
    assert [f(n) for n in (1, 3, 4, 5, 201, 1023, 1024, 1025, 1041, 2049)] == \
        [4, 4, 4, 6, 208, 1024, 1024, 1040, 1050, 2058]
    for n in range(4, 3000):
        m = f(n)
        assert m >= n and m % 2 == 0
        r = m
        for p in (2, 3, 5, 7, 11, 13):
            while r % p == 0:
                r //= p
        assert r == 1


def _field(n, half, fn, R=1e23, dR=0.):
    w = SRWLWfr()
    w.allocate(1, n, n)
    w.mesh.eStart = w.mesh.eFin = E_EV
    w.mesh.xStart, w.mesh.xFin = -half, half
    w.mesh.yStart, w.mesh.yFin = -half, half
    x = np.linspace(-half, half, n)
    X, Y = np.meshgrid(x, x)
    E = fn(X, Y)
    w.arEx = array('f', E.astype(np.complex64).view(np.float32).ravel().tobytes())
    w.arEy = array('f', (0.3 * E).astype(np.complex64).view(np.float32).ravel().tobytes())
    w.Rx = w.Ry = R
    w.dRx = w.dRy = dR
    return w


def _grid(w):
    x = np.linspace(w.mesh.xStart, w.mesh.xFin, w.mesh.nx)
    y = np.linspace(w.mesh.yStart, w.mesh.yFin, w.mesh.ny)
    return np.meshgrid(x, y)


def test_range_only_enlargement_copies_points():
    g = lambda X, Y: np.exp(-(X ** 2 + Y ** 2) / (2e-4) ** 2) * np.exp(1j * X / 3e-5)
    w = _field(64, 4e-4, g)
    old = cfield(w.arEx, w)[:, :, 0].copy()
    rc, wn = oracle.resize(w, [0, 2., 1., 2., 1.])
    assert rc == 0
    assert wn.mesh.nx == 128 and wn.mesh.ny == 128
    dx0 = 8e-4 / 63
    assert (wn.mesh.xFin - wn.mesh.xStart) / 127 == pytest.approx(dx0, rel=1e-12)
    new = cfield(wn.arEx, wn)[:, :, 0]
    # old mesh sits at indices 32..95 (left margin rounded up), zeros elsewhere
    assert np.array_equal(new[32:96, 32:96], old)
    assert not new[:32].any() and not new[96:].any() and not new[:, :32].any()


def test_finer_mesh_reproduces_a_smooth_tilted_beam():
    k = 2 * math.pi / 2.0e-4
    g = lambda X, Y: np.exp(-(X ** 2 + Y ** 2) / (1.5e-4) ** 2 + 1j * k * (X + 0.5 * Y))
    w = _field(128, 4e-4, g)
    rc, wn = oracle.resize(w, [0, 1., 2., 1., 2.])
    assert rc == 0
    assert wn.mesh.nx == 256
    assert wn.mesh.xStart == pytest.approx(-4e-4) and wn.mesh.xFin == pytest.approx(4e-4)
    X, Y = _grid(wn)
    # bicubic interpolation error at k h = 0.2: below 1e-3 of the field
    assert rel_l2(cfield(wn.arEx, wn)[:, :, 0], g(X, Y)) < 1e-3
    assert rel_l2(cfield(wn.arEy, wn)[:, :, 0], 0.3 * g(X, Y)) < 1e-3


def test_quadratic_phase_is_removed_before_interpolation():
    # 10 m diverging wave: its curvature phase is undersampled for plain
    # interpolation, but with the radius known the field is resampled to 1e-3
    R = 10.
    g = lambda X, Y: np.exp(-(X ** 2 + Y ** 2) / (5e-4) ** 2 + 1j * math.pi * (X ** 2 + Y ** 2) / (LAM * R))
    w = _field(200, 8e-4, g, R=R, dR=0.01 * R)
    rc, wn = oracle.resize(w, [0, 2., 2., 2., 2.])
    assert rc == 0
    assert wn.mesh.nx == 800 and wn.mesh.ny == 800
    X, Y = _grid(wn)
    inside = (np.abs(X) <= 8e-4) & (np.abs(Y) <= 8e-4)
    got = cfield(wn.arEx, wn)[:, :, 0]
    assert rel_l2(got[inside], g(X, Y)[inside]) < 1e-3
    assert not got[~inside].any()
    # without the radius the same resampling is far off
    w2 = _field(200, 8e-4, g)
    _, wn2 = oracle.resize(w2, [0, 2., 2., 2., 2.])
    assert rel_l2(cfield(wn2.arEx, wn2)[:, :, 0][inside], g(X, Y)[inside]) > 1e-2


def test_cells_with_zeros_use_bilinear_interpolation():
    # field exactly zero outside a square: cells touching the zeros are
    # interpolated bilinearly, so points surrounded by zeros stay zero
    g = lambda X, Y: np.where((np.abs(X) < 2e-4) & (np.abs(Y) < 2e-4), 1. + 0.5j, 0.)
    w = _field(64, 4e-4, g)
    rc, wn = oracle.resize(w, [0, 1., 2., 1., 2.])
    assert rc == 0
    X, Y = _grid(wn)
    got = cfield(wn.arEx, wn)[:, :, 0]
    far = (np.abs(X) > 2.3e-4) | (np.abs(Y) > 2.3e-4)
    assert not got[far].any()
    core = (np.abs(X) < 1.6e-4) & (np.abs(Y) < 1.6e-4)
    assert np.allclose(got[core], 1. + 0.5j, rtol=1e-6)


def _mesh(n, x0, x1, z0=None, z1=None):
    from srw_amd.srwlib import SRWLRadMesh
    m = SRWLRadMesh()
    m.ne, m.nx, m.ny = 1, n, n
    m.eStart = m.eFin = E_EV
    m.xStart, m.xFin = x0, x1
    m.yStart, m.yFin = (x0, x1) if z0 is None else (z0, z1)
    return m


def test_resize_to_mesh_at_the_old_nodes_is_exact():
    # a new mesh on the old nodes (every other node, inside the old range):
    # the bicubic form at a node is the node value and the intensity rescale
    # is 1, so the values are copied to float rounding
    g = lambda X, Y: np.exp(-(X ** 2 + Y ** 2) / (2e-4) ** 2) * np.exp(1j * X / 5e-5)
    w = _field(65, 4e-4, g)
    old = cfield(w.arEx, w)[:, :, 0].copy()
    rc, wn = oracle.resize_to_mesh(w, _mesh(17, -2e-4, 2e-4), (0, 1, 1))
    assert rc == 0
    new = cfield(wn.arEx, wn)[:, :, 0]
    assert np.abs(new - old[16:49:2, 16:49:2]).max() <= 2e-6 * np.abs(old).max()


@pytest.mark.parametrize("phase,order_ratio", [(0., 3.5), (0.3, 12.)])
def test_resize_to_mesh_interpolates_smooth_field(phase, order_ratio):
    # off-node mesh, no curvature phase, with and without the intensity-based
    # Re/Im correction.  A real field has exact zeros (Im) in every cell and
    # takes the bilinear low-order path: the error falls ~4x per halved old
    # step.  A field with a constant phase takes the bicubic path: ~16x.
    g = lambda X, Y: np.exp(-(X ** 2 + Y ** 2) / (2e-4) ** 2) * np.exp(1j * phase)
    m = _mesh(50, -3.3e-4, 2.9e-4, -2.7e-4, 3.1e-4)
    for improve in (0, 1):
        errs = []
        for n in (41, 81, 161):
            w = _field(n, 8e-4, g)
            rc, wn = oracle.resize_to_mesh(w, m, (0, 1, improve))
            assert rc == 0
            X, Y = _grid(wn)
            errs.append(rel_l2(cfield(wn.arEx, wn)[:, :, 0], g(X, Y)))
            assert rel_l2(cfield(wn.arEy, wn)[:, :, 0], 0.3 * g(X, Y)) == pytest.approx(errs[-1], rel=1e-3)
        assert errs[2] < 1e-3
        assert errs[0] / errs[1] > order_ratio and errs[1] / errs[2] > order_ratio


def test_resize_to_mesh_zero_outside_old_range():
    # no exact zeros in the cells (a real field would take the reference's
    # low-order path, which extrapolates in the edge cells)
    g = lambda X, Y: np.exp(-(X ** 2 + Y ** 2) / (2e-4) ** 2) * np.exp(0.3j)
    w = _field(40, 4e-4, g)
    m = _mesh(60, -9e-4, 3e-4, -2e-4, 9e-4)
    rc, wn = oracle.resize_to_mesh(w, m, (0, 1, 1))
    assert rc == 0
    X, Y = _grid(wn)
    E = cfield(wn.arEx, wn)[:, :, 0]
    out = (X < -4e-4 - 1e-12) | (X > 4e-4 + 1e-12) | (Y < -4e-4 - 1e-12) | (Y > 4e-4 + 1e-12)
    assert np.all(E[out] == 0)
    assert np.all(np.abs(E[~out]) > 0)
    assert rel_l2(E[~out], g(X, Y)[~out]) < 1e-3


def test_resize_to_mesh_matches_factor_resize():
    # the mesh that srwlResizeElecField builds from (pxm, pxd), given as an
    # explicit mesh, is filled the same way (both treat the curvature phase)
    R = 20.
    g = lambda X, Y: np.exp(-(X ** 2 + Y ** 2) / (2e-4) ** 2) * np.exp(1j * math.pi / LAM * (X ** 2 + Y ** 2) / R)
    w1 = _field(64, 4e-4, g, R=R, dR=0.01)
    w2 = _field(64, 4e-4, g, R=R, dR=0.01)
    rc, wa = oracle.resize(w1, [0, 1.5, 2., 1.5, 2.])
    assert rc == 0
    m = _mesh(wa.mesh.nx, wa.mesh.xStart, wa.mesh.xFin, wa.mesh.yStart, wa.mesh.yFin)
    m.ny = wa.mesh.ny
    rc, wb = oracle.resize_to_mesh(w2, m, (0, 1, 1))
    assert rc == 0
    a, b = cfield(wa.arEx, wa), cfield(wb.arEx, wb)
    assert np.array_equal(a, b)
