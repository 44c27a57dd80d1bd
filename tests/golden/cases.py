"""Input cases of the reference golden vectors (tests/golden/ref_*.npz).

Each SR case is a plain dict; `sr_case(d)` builds the srw_amd.srwlib objects
(wavefront, field container, precPar) from it.  make_ref_vectors.py runs the
reference SRW sources on them and stores the results; the tests rebuild the
same inputs and compare the oracle and the device path with the stored
results.
"""
from array import array

from srw_amd import workloads as W
from srw_amd.srwlib import SRWLMagFldC, SRWLMagFldH, SRWLMagFldU

E1 = None   # first-harmonic energy of the 3 GeV, K = 1.5, 20 mm undulator (set below)


def _e1(K=1.5, per=0.02):
    b = W.electron_beam(3.0)
    return W.first_harmonic_eV(b.partStatMom1.gamma, K, per)


E1 = _e1()

# name, then overrides of the defaults in sr_case()
SR_CASES = [
    dict(name="c1_11x11"),
    dict(name="termin0_offaxis", x=(0.2e-3, 1.4e-3), nx=9, ny=7, termin=0),
    dict(name="termin1_offaxis", x=(0.2e-3, 1.4e-3), nx=9, ny=7, termin=1),
    dict(name="termin2_offaxis", x=(0.2e-3, 1.4e-3), nx=9, ny=7, termin=2),
    dict(name="termin3_offaxis", x=(0.2e-3, 1.4e-3), nx=9, ny=7, termin=3),
    dict(name="three_energies", ne=3, de=0.05, nx=9, ny=9),
    dict(name="electron_offset_tilt", x0=50e-6, xp0=10e-6, y0=-20e-6, yp0=-5e-6, nx=9, ny=9),
    dict(name="offset_far_mesh", x0=-0.3e-3, xp0=-30e-6, x=(-2e-3, 0.5e-3), y=(-0.4e-3, 1.2e-3), nx=8, ny=9),
    dict(name="near_5m", z=5.0, half=0.5e-3, nx=9, ny=9),
    dict(name="far_50m", z=50.0, half=2.5e-3, nx=9, ny=9),
    dict(name="s_limits", slim=(-0.5, 0.7), nx=9, ny=9),
    dict(name="s_limits_offset", slim=(-0.9, 0.2), x0=20e-6, xp0=-15e-6, nx=7, ny=7),
    dict(name="zero_field", K=0.0, e=2000.0, nx=9, ny=9),
    dict(name="K05_20per", K=0.5, nper=20, nx=9, ny=9),
    dict(name="K25", K=2.5, nx=9, ny=9, e_harm=3),
    dict(name="two_undulators", two=True, z=30.0, nx=9, ny=9),
    dict(name="horizontal_field", hv="h", nx=9, ny=9),
    dict(name="coarse_step", np_traj=2001, step_mult=2, nx=9, ny=9),
    dict(name="even_mesh_symmetric", nx=10, ny=10),
    dict(name="meth1_relprec_0.01", meth=1, rel=0.01, nx=9, ny=9),
    dict(name="meth1_offaxis", meth=1, rel=0.005, x=(0.1e-3, 1.1e-3), nx=7, ny=5),
    # edge_noise: at a few points of these two cases the reference's own
    # result depends on rounding noise: the field and its derivative at the
    # integration edges, outside the magnet, are the noise of the trajectory
    # interpolation's higher derivatives (1e-13 T, 5e-10 T/m here), and they
    # decide there whether the third term of the far-field expansion is
    # taken (srradint.cpp:497-503).  Setting them to zero or doubling them
    # moves those points by up to 65 % of the largest |Ey|.
    dict(name="near_field_edge_0.5eV", e=0.5, half=5e-3, nx=11, ny=11, edge_noise=True),
    dict(name="near_field_edge_1eV_2m", e=1.0, half=1e-3, z=2.0, nx=11, ny=11),
    dict(name="near_field_edge_2eV", e=2.0, half=5e-3, nx=11, ny=11, edge_noise=True),
    dict(name="near_field_edge_meth1", e=1.0, half=4e-3, meth=1, rel=0.01, nx=9, ny=9),
    dict(name="near_field_edge_termin2", e=0.7, half=5e-3, x=(0.5e-3, 5e-3), nx=8, ny=7, termin=2),
    # every point of the mesh lies where the left-edge expansion is rejected,
    # except the x = 0 column: the edge derivative carries from that column
    # into the next row, across more than one block of 256 points
    dict(name="near_field_carried_state", e=0.5, half=1e-3, nx=41, ny=41),
    dict(name="near_field_carried_energies", e_range=(0.45, 0.6), ne=4, half=3e-3, nx=21, ny=21),
    # method 1, where the carried derivative weighs on the coarse levels, on
    # a field-free straight line (no magnet: the edge fields are exactly
    # zero, not rounding noise): starting every point from a zero derivative
    # moves these fields by 5e-5 and 8e-5 relative L2 (tol: the tolerance
    # their tests use)
    dict(name="near_field_carried_meth1", K=0, e=1.0, half=4e-3, meth=1, rel=0.01, nx=41, ny=41, tol=1e-5),
    dict(name="near_field_carried_meth1_energies", K=0, e_range=(0.8, 1.2), ne=3, half=4e-3, meth=1, rel=0.01,
         nx=21, ny=21, tol=1e-5),
]


def sr_case(d):
    """(wavefront, field container, precPar) of one case dict."""
    K = d.get("K", 1.5)
    per = d.get("per", 0.02)
    nper = d.get("nper", 100)
    beam = W.electron_beam(3.0, 0.5, x=d.get("x0", 0.), xp=d.get("xp0", 0.), y=d.get("y0", 0.),
                           yp=d.get("yp0", 0.))
    if "e" in d:
        e0 = e1 = d["e"]
    elif "e_range" in d:
        e0, e1 = d["e_range"]
    else:
        eh = d.get("e_harm", 1) * _e1(K if K > 0 else 1.5, per)
        de = d.get("de", 0.)
        e0, e1 = eh * (1 - de), eh * (1 + de)
    ne = d.get("ne", 1)
    if ne == 1:
        e1 = e0
    half = d.get("half", 1e-3)
    w = W.wavefront(beam, ne, e0, e1, d.get("nx", 11), d.get("ny", 11), half, half, d.get("z", 20.))
    if "x" in d:
        w.mesh.xStart, w.mesh.xFin = d["x"]
    if "y" in d:
        w.mesh.yStart, w.mesh.yFin = d["y"]
    B = W.undulator_B(K, per) if K > 0 else 0.
    hv = d.get("hv", "v")
    if d.get("two"):
        u1 = SRWLMagFldU([SRWLMagFldH(1, hv, B, 0, 1, 1)], per, 40)
        u2 = SRWLMagFldU([SRWLMagFldH(1, hv, B, 0, 1, 1)], per, 40)
        mag = SRWLMagFldC([u1, u2], array('d', [0, 0]), array('d', [0, 0]), array('d', [-1.2, 1.2]))
        half_range = 1.2 + 0.5 * 40 * per + 4 * per
        np_traj = d.get("np_traj", 20001)
    else:
        mag = SRWLMagFldC([SRWLMagFldU([SRWLMagFldH(1, hv, B, 0, 1, 1)], per, nper)], array('d', [0]),
                          array('d', [0]), array('d', [0]))
        half_range = 0.5 * nper * per + 4 * per
        np_traj = d.get("np_traj", 10001)
    step = d.get("step_mult", 1) * 2. * half_range / (np_traj - 1)
    meth = d.get("meth", 0)
    prec = array('d', [meth, d["rel"] if meth == 1 else step, 0, 0, np_traj, d.get("termin", 1), 0])
    if "slim" in d:
        prec[2], prec[3] = d["slim"]
    if meth == 0 and "slim" in d:
        prec[1] = step
    return w, mag, prec


# Wavefronts for the drift's radius estimates (sroptdrf.cpp:1085-1381): a
# Gaussian beam of waist w0 observed z0 past its waist on an n x n mesh, with
# the radius (and its error) set as given; "mom" = 0 leaves the moment arrays
# empty so the call computes them.
RADSUB_CASES = [
    dict(name="waist_near", w0=20e-6, z0=10., n=64, half=1.2e-3, L=20., sub=1, R=None, dRf=0.01, mom=0),
    dict(name="waist_near_L_neg", w0=20e-6, z0=10., n=64, half=1.2e-3, L=-4., sub=1, R=None, dRf=0.01, mom=0),
    dict(name="far_large_spread", w0=5e-6, z0=20., n=64, half=2e-3, L=30., sub=1, R=None, dRf=0.01, mom=0),
    dict(name="converging", w0=10e-6, z0=-5., n=64, half=0.6e-3, L=3., sub=1, R=None, dRf=0.01, mom=0),
    dict(name="converging_past_focus", w0=10e-6, z0=-2., n=48, half=0.5e-3, L=6., sub=1, R=None, dRf=0.01, mom=0),
    dict(name="radius_zero", w0=30e-6, z0=8., n=48, half=1e-3, L=10., sub=1, R=0., dRf=0.0, mom=0),
    dict(name="ne3_avg_energy", w0=20e-6, z0=10., n=48, half=1.2e-3, L=20., sub=1, R=None, dRf=0.01, mom=0, ne=3),
    dict(name="sub2_small_radius", w0=20e-6, z0=10., n=48, half=1.2e-3, L=20., sub=2, R=0.02, dRf=1.0, mom=0),
    dict(name="sub2_near_minus_L", w0=20e-6, z0=10., n=48, half=1.2e-3, L=20., sub=2, R=-19.99, dRf=0.01, mom=0),
    dict(name="sub2_plain", w0=20e-6, z0=10., n=48, half=1.2e-3, L=20., sub=2, R=None, dRf=0.01, mom=0),
]


def radsub_wavefront(d, e_ev=1000.):
    """Gaussian-beam wavefront of a RADSUB case: TEM00 field z0 past its
    waist (diverging-wave phase convention), ne photon energies around e_ev
    (each with its own wavelength), Ey = 0.5i Ex; Rx = Ry = the beam's
    radius z0(1 + (zR/z0)^2) unless given, dRx = dRy = dRf*|R|."""
    import math
    import numpy as np
    from srw_amd.srwlib import SRWLWfr
    n, half, w0, z0, ne = d["n"], d["half"], d["w0"], d["z0"], d.get("ne", 1)
    w = SRWLWfr()
    w.allocate(ne, n, n)
    w.mesh.eStart, w.mesh.eFin = (e_ev, e_ev) if ne == 1 else (0.98 * e_ev, 1.02 * e_ev)
    w.mesh.xStart, w.mesh.xFin = -half, half * (n - 2) / n
    w.mesh.yStart, w.mesh.yFin = -half, half * (n - 2) / n
    w.mesh.zStart = 0.
    x = np.linspace(w.mesh.xStart, w.mesh.xFin, n)
    X, Y = np.meshgrid(x, x)
    r2 = X * X + Y * Y
    E = np.zeros((n, n, ne), dtype=np.complex128)
    energies = np.linspace(w.mesh.eStart, w.mesh.eFin, ne)
    for ie, e in enumerate(energies):
        lam = 1.239842e-06 / e
        zR = math.pi * w0 * w0 / lam
        wz = w0 * math.sqrt(1 + (z0 / zR) ** 2)
        invR = z0 / (z0 * z0 + zR * zR)
        E[:, :, ie] = (w0 / wz) * np.exp(-r2 / wz ** 2 + 1j * (math.pi * r2 * invR / lam))
    w.arEx = array('f', E.astype(np.complex64).view(np.float32).ravel().tobytes())
    w.arEy = array('f', (0.5j * E).astype(np.complex64).view(np.float32).ravel().tobytes())
    lam0 = 1.239842e-06 / e_ev
    zR0 = math.pi * w0 * w0 / lam0
    R = z0 * (1 + (zR0 / z0) ** 2) if d.get("R") is None else d["R"]
    w.Rx = w.Ry = R
    w.dRx = w.dRy = d["dRf"] * abs(R)
    w.avgPhotEn = e_ev
    return w


# Extraction (srwlCalcIntFromElecField) golden cases: a seeded random field on
# a 5-energy, 7 x 6 mesh, every polarisation, intType 0, 2, 4, 5, 6, 7, every
# dependence type, the averaging / sum modes of arMeth for the x-y dependence;
# at one off-mesh (e, x, y).  The reference's own bugs are left out: the
# phase vs (e, x), (e, y), (e, x, y) (written through a null float pointer)
# and the fourth Stokes block vs (e, x, y) (a copy of the third).
INT_POLS = [0, 1, 2, 3, 4, 5, 6, -1, -2, -3, -4, -5]
INT_TYPES = [0, 2, 4, 5, 6, 7]
INT_METHS = [None, [1, 3], [2]]
INT_POINT = (None, 0.13e-3, -0.21e-3)   # e: mid-band + 3.3 eV (set in int_wavefront)


def int_wavefront(seed=3):
    import numpy as np
    from srw_amd import workloads as W
    rng = np.random.default_rng(seed)
    w, _, _ = W.single_electron(n=7, ne=5, de_rel=0.05)
    w.allocate(5, 7, 6)
    w.mesh.ny = 6
    w.mesh.yFin = 0.8e-3
    n = 2 * 5 * 7 * 6
    w.arEx = array('f', rng.normal(size=n).astype(np.float32).tobytes())
    w.arEy = array('f', (0.3 * rng.normal(size=n)).astype(np.float32).tobytes())
    e = 0.5 * (w.mesh.eStart + w.mesh.eFin) + 3.3
    return w, (e, INT_POINT[1], INT_POINT[2])


def int_cases():
    """(intType, depType, pol, meth) of the extraction golden set."""
    out = []
    for it in INT_TYPES:
        for dep in range(7):
            if it == 4 and dep >= 4:
                continue
            for pol in INT_POLS:
                if dep == 6 and pol == -5:
                    continue
                for meth in INT_METHS:
                    if meth is not None and dep != 3:
                        continue
                    out.append((it, dep, pol, meth))
    return out


def int_len(w, dep, pol, it):
    m = w.mesh
    lens = [m.ne, m.nx, m.ny, m.nx * m.ny, m.ne * m.nx, m.ne * m.ny, m.ne * m.nx * m.ny]
    return lens[dep] * (4 if (pol == -5 and it != 2) else 1)


def int_prefill(k, n, it):
    """Prefilled output buffer of extraction case k (seeded by k)."""
    import numpy as np
    rng = np.random.default_rng(1000 + k)
    return rng.normal(size=n).astype(np.float64 if it == 4 else np.float32)
