"""Python classes of the SRWLIB API used by the emit-and-propagate path.

The class names, constructor arguments and attribute names follow the
reference's env/python/srwpy/srwlib.py (SRWLParticle :67, SRWLPartBeam :109,
SRWLMagFldH :306, SRWLMagFldU :325, SRWLMagFldC :410, SRWLPrtTrj :623,
SRWLRadMesh :801, SRWLStokes :924, SRWLWfr :2200, SRWLOptD :2908,
SRWLOptC :4259), so scripts written against srwlib can build the same objects
and hand them to srw_amd.srwlpy.  Arrays are array.array('f'/'d') like the
reference's.
"""
from array import array


class SRWLParticle(object):
    def __init__(self, _x=0, _y=0, _z=0, _xp=0, _yp=0, _gamma=1, _relE0=1, _nq=-1):
        self.x, self.y, self.z = _x, _y, _z
        self.xp, self.yp = _xp, _yp
        self.gamma, self.relE0, self.nq = _gamma, _relE0, _nq

    def drift(self, _dist):
        self.x += self.xp * _dist
        self.y += self.yp * _dist
        self.z += _dist


class SRWLPartBeam(object):
    def __init__(self, _Iavg=0, _nPart=0, _partStatMom1=None, _arStatMom2=None):
        self.Iavg = _Iavg
        self.nPart = _nPart
        self.partStatMom1 = SRWLParticle() if _partStatMom1 is None else _partStatMom1
        self.arStatMom2 = array('d', [0] * 21) if _arStatMom2 is None else _arStatMom2


class SRWLMagFldH(object):
    def __init__(self, _n=1, _h_or_v='v', _B=0, _ph=0, _s=1, _a=1):
        self.n, self.h_or_v, self.B, self.ph, self.s, self.a = _n, _h_or_v, _B, _ph, _s, _a


class SRWLMagFldU(object):
    def __init__(self, _arHarm=None, _per=0, _nPer=0):
        self.arHarm = [] if _arHarm is None else _arHarm
        self.per = _per
        self.nPer = _nPer

    def allocate(self, _nHarm):
        self.arHarm = [SRWLMagFldH() for _ in range(_nHarm)]

    def get_K(self):
        """Deflection parameter of the first harmonic: K = 93.3729*per*B."""
        return 93.37290417576577 * self.per * self.arHarm[0].B


class SRWLMagFldC(object):
    def __init__(self, _arMagFld=None, _arXc=None, _arYc=None, _arZc=None, _arVx=None, _arVy=None,
                 _arVz=None, _arAng=None):
        self.arMagFld = [] if _arMagFld is None else list(_arMagFld)
        n = len(self.arMagFld)
        self.arXc = array('d', [0] * n) if _arXc is None else _arXc
        self.arYc = array('d', [0] * n) if _arYc is None else _arYc
        self.arZc = array('d', [0] * n) if _arZc is None else _arZc
        self.arVx = array('d', [0] * n) if _arVx is None else _arVx
        self.arVy = array('d', [0] * n) if _arVy is None else _arVy
        self.arVz = array('d', [1] * n) if _arVz is None else _arVz
        self.arAng = array('d', [0] * n) if _arAng is None else _arAng


class SRWLPrtTrj(object):
    def __init__(self, _arX=None, _arXp=None, _arY=None, _arYp=None, _arZ=None, _arZp=None,
                 _arBx=None, _arBy=None, _arBz=None, _np=0, _ctStart=0, _ctEnd=0, _partInitCond=None):
        self.arX, self.arXp, self.arY, self.arYp = _arX, _arXp, _arY, _arYp
        self.arZ, self.arZp = _arZ, _arZp
        self.arBx, self.arBy, self.arBz = _arBx, _arBy, _arBz
        self.np, self.ctStart, self.ctEnd = _np, _ctStart, _ctEnd
        self.partInitCond = SRWLParticle() if _partInitCond is None else _partInitCond

    def allocate(self, _np, _allB=False):
        self.np = _np
        z = [0] * _np
        self.arX, self.arXp, self.arY, self.arYp = array('d', z), array('d', z), array('d', z), array('d', z)
        self.arZ, self.arZp = array('d', z), array('d', z)
        if _allB:
            self.arBx, self.arBy, self.arBz = array('d', z), array('d', z), array('d', z)


class SRWLRadMesh(object):
    def __init__(self, _eStart=0, _eFin=0, _ne=1, _xStart=0, _xFin=0, _nx=1, _yStart=0, _yFin=0, _ny=1,
                 _zStart=0, _nvx=0, _nvy=0, _nvz=1, _hvx=1, _hvy=0, _hvz=0, _arSurf=None):
        self.eStart, self.eFin, self.ne = _eStart, _eFin, _ne
        self.xStart, self.xFin, self.nx = _xStart, _xFin, _nx
        self.yStart, self.yFin, self.ny = _yStart, _yFin, _ny
        self.zStart = _zStart
        self.nvx, self.nvy, self.nvz, self.hvx, self.hvy, self.hvz = _nvx, _nvy, _nvz, _hvx, _hvy, _hvz
        self.arSurf = _arSurf

    def set_from_other(self, _mesh):
        for k in ('eStart', 'eFin', 'ne', 'xStart', 'xFin', 'nx', 'yStart', 'yFin', 'ny', 'zStart'):
            setattr(self, k, getattr(_mesh, k))

    def is_equal(self, _mesh, _rel_tol=1e-12):
        for k in ('ne', 'nx', 'ny'):
            if getattr(self, k) != getattr(_mesh, k):
                return False
        for k in ('eStart', 'eFin', 'xStart', 'xFin', 'yStart', 'yFin'):
            a, b = getattr(self, k), getattr(_mesh, k)
            if abs(a - b) > _rel_tol * max(abs(a), abs(b), 1e-300):
                return False
        return True


class SRWLStokes(object):
    def __init__(self, _arS=None, _typeStokes='f', _eStart=0, _eFin=0, _ne=0, _xStart=0, _xFin=0, _nx=0,
                 _yStart=0, _yFin=0, _ny=0, _mutual=0, _n_comp=4):
        self.arS = _arS
        self.numTypeStokes = _typeStokes
        self.mesh = SRWLRadMesh(_eStart, _eFin, _ne, _xStart, _xFin, _nx, _yStart, _yFin, _ny)
        self.avgPhotEn = 0
        self.presCA = 0
        self.presFT = 0
        self.unitStokes = 1
        self.mutual = _mutual
        if (_arS is None) and (_ne > 0) and (_nx > 0) and (_ny > 0):
            self.allocate(_ne, _nx, _ny, _typeStokes, _mutual, _n_comp)

    def allocate(self, _ne, _nx, _ny, _typeStokes='f', _mutual=0, _n_comp=4):
        self.arS = array(_typeStokes, [0] * (_n_comp * _ne * _nx * _ny))
        self.numTypeStokes = _typeStokes
        self.mesh.ne, self.mesh.nx, self.mesh.ny = _ne, _nx, _ny


class SRWLWfr(object):
    def __init__(self, _arEx=None, _arEy=None, _typeE='f', _eStart=0, _eFin=0, _ne=0, _xStart=0, _xFin=0,
                 _nx=0, _yStart=0, _yFin=0, _ny=0, _zStart=0, _partBeam=None):
        self.arEx, self.arEy = _arEx, _arEy
        self.arExAux = self.arEyAux = None
        self.mesh = SRWLRadMesh(_eStart, _eFin, _ne, _xStart, _xFin, _nx, _yStart, _yFin, _ny, _zStart)
        self.Rx = self.Ry = self.dRx = self.dRy = 0.
        self.xc = self.yc = 0.
        self.avgPhotEn = 0.
        self.presCA = 0
        self.presFT = 0
        self.numTypeElFld = _typeE
        self.unitElFld = 1
        self.unitElFldAng = 0
        self.partBeam = SRWLPartBeam() if _partBeam is None else _partBeam
        self.arElecPropMatr = array('d', [0] * 20)
        self.arMomX = array('d', [0] * (11 * max(_ne, 1)))
        self.arMomY = array('d', [0] * (11 * max(_ne, 1)))
        self.arWfrAuxData = array('d', [0] * 30)
        if (_arEx is None) and (_ne > 0) and (_nx > 0) and (_ny > 0):
            self.allocate(_ne, _nx, _ny)

    def allocate(self, _ne, _nx, _ny, _EXNeeded=1, _EYNeeded=1, _typeE='f', _backupNeeded=0):
        n = 2 * _ne * _nx * _ny
        if _EXNeeded:
            self.arEx = array(_typeE, bytes(4 * n)) if _typeE == 'f' else array(_typeE, [0] * n)
        if _EYNeeded:
            self.arEy = array(_typeE, bytes(4 * n)) if _typeE == 'f' else array(_typeE, [0] * n)
        self.numTypeElFld = _typeE
        self.mesh.ne, self.mesh.nx, self.mesh.ny = _ne, _nx, _ny
        if len(self.arMomX) < 11 * _ne:
            self.arMomX = array('d', [0] * (11 * _ne))
            self.arMomY = array('d', [0] * (11 * _ne))


class SRWLOptD(object):
    def __init__(self, _L=0, _treat=0):
        self.L = _L
        self.treat = _treat


class SRWLOptA(object):
    """Aperture or obstacle (reference srwlib.py:2922-2936)."""
    def __init__(self, _shape='r', _ap_or_ob='a', _Dx=0, _Dy=0, _x=0, _y=0):
        self.shape = _shape
        self.ap_or_ob = _ap_or_ob
        self.Dx = _Dx
        self.Dy = _Dy
        self.x = _x
        self.y = _y


class SRWLOptL(object):
    """Thin lens (reference srwlib.py:2941-2951)."""
    def __init__(self, _Fx=1e+23, _Fy=1e+23, _x=0, _y=0):
        self.Fx = _Fx
        self.Fy = _Fy
        self.x = _x
        self.y = _y


class SRWLOptC(object):
    def __init__(self, _arOpt=None, _arProp=None):
        self.arOpt = [] if _arOpt is None else list(_arOpt)
        self.arProp = [] if _arProp is None else list(_arProp)


# ---------------------------------------------------------------------------
# Result files and progress log of the multi-electron driver
# ---------------------------------------------------------------------------

def srwl_uti_save_intens_ascii(_ar_intens, _mesh, _file_path, _n_stokes=1,
                               _arLabels=('Photon Energy', 'Horizontal Position', 'Vertical Position', 'Intensity'),
                               _arUnits=('eV', 'm', 'm', 'ph/s/.1%bw/mm^2'), _mutual=0, _cmplx=0):
    """Intensity file in SRW's ASCII layout (reference srwlib.py:6815-6891):
    a header line naming the entity and the loop order, the mesh (start, end
    and point count vs photon energy, x and y, one '#'-prefixed value per
    line), the component count, then every value on its own line, photon
    energy fastest."""
    lab = [_arLabels[i] + ' [' + _arUnits[i] + ']' for i in range(4)]
    ent = lab[3]
    if _mutual == 1:
        ent = ('Complex Mutual' if _cmplx else 'Mutual') + ('' if ent.startswith(' ') else ' ') + ent
    lines = ['#' + ent + ' (C-aligned, inner loop is vs ' + _arLabels[0] + ', outer loop vs ' + _arLabels[2] + ')']
    for k, (a, b, n) in enumerate((('eStart', 'eFin', 'ne'), ('xStart', 'xFin', 'nx'), ('yStart', 'yFin', 'ny'))):
        lines.append('#' + repr(getattr(_mesh, a)) + ' #Initial ' + lab[k])
        lines.append('#' + repr(getattr(_mesh, b)) + ' #Final ' + lab[k])
        lines.append('#' + repr(getattr(_mesh, n)) + ' #Number of points vs ' + _arLabels[k])
    n_comp = 1
    if _n_stokes > 0:
        lines.append('#' + repr(_n_stokes) + ' #Number of components')
        n_comp = _n_stokes
    n_pt = _mesh.ne * _mesh.nx * _mesh.ny
    if _mutual > 0:
        n_pt *= n_pt
    n_val = n_pt * n_comp * (2 if _cmplx else 1)
    with open(_file_path, 'w') as f:
        f.write('\n'.join(lines) + '\n')
        f.write(''.join(' ' + repr(float(_ar_intens[i])) + '\n' for i in range(n_val)))


def srwl_uti_read_intens_ascii(_file_path, _num_type='f'):
    """Reads a file written by srwl_uti_save_intens_ascii (reference
    srwlib.py:6893-6923): returns (array of values, SRWLRadMesh)."""
    with open(_file_path) as f:
        lines = f.readlines()
    m = SRWLRadMesh()
    vals = [lines[i].split('#')[1] for i in range(1, 10)]
    m.eStart, m.eFin, m.ne = float(vals[0]), float(vals[1]), int(vals[2])
    m.xStart, m.xFin, m.nx = float(vals[3]), float(vals[4]), int(vals[5])
    m.yStart, m.yFin, m.ny = float(vals[6]), float(vals[7]), int(vals[8])
    i0 = 11 if lines[10].startswith('#') else 10
    return array(_num_type, [float(s) for s in lines[i0:] if len(s.strip()) > 0]), m


def srwl_uti_save_intens(_ar_intens, _mesh, _file_path, _n_stokes=1,
                         _arLabels=('Photon Energy', 'Horizontal Position', 'Vertical Position', 'Intensity'),
                         _arUnits=('eV', 'm', 'm', 'ph/s/.1%bw/mm^2'), _mutual=0, _cmplx=0, _form='ascii', _wfr=None):
    """Format dispatch of the intensity writer (reference srwlib.py:7139-7146).
    Only the ASCII layout is written here (h5py is not available)."""
    form = _form.upper()
    if form in ('ASCII', 'ASC', 'TXT'):
        srwl_uti_save_intens_ascii(_ar_intens, _mesh, _file_path, _n_stokes, _arLabels, _arUnits, _mutual, _cmplx)
    else:
        raise Exception('file format %r is not supported by this build (ASCII only)' % _form)


def srwl_uti_read_intens(_file_path, _num_type='f', _form='ascii'):
    form = _form.upper()
    if form in ('ASCII', 'ASC', 'TXT'):
        return srwl_uti_read_intens_ascii(_file_path, _num_type)
    return None


def srwl_uti_save_stat_wfr_emit_prop_multi_e(particle_number=0, total_num_of_particles=0,
                                             filename='srwl_stat_wfr_emit_prop_multi_e', cores=None,
                                             particles_per_iteration=None):
    """Progress files of a multi-electron run (reference srwlib.py:7746-7812):
    <filename>.log gets one line per call (a start line for particle 0, then
    'Running' / 'Finished' with the particle count and per cent done) and
    <filename>.json holds the latest status, replaced atomically."""
    import datetime
    import json
    import os
    import shutil
    import tempfile
    offset = len(str(total_num_of_particles))
    timestamp = '{:%Y-%m-%d %H:%M:%S}'.format(datetime.datetime.now())
    progress = float(particle_number) / float(total_num_of_particles) * 100.0
    status = 'Running' if progress < 100.0 else 'Finished'
    if particle_number == 0:
        mode = 'w'
        text = '[{}]: Calculation on {} core{} with averaging of {} particle{}/iteration.'.format(
            timestamp, cores, 's' if cores > 1 else '', particles_per_iteration,
            's' if particles_per_iteration > 1 else '')
    else:
        mode = 'a'
        text = '[{}]: {:8s} {:{offset}d} out of {:{offset}d} ({:6.2f}% complete)'.format(
            timestamp, status, particle_number, total_num_of_particles, progress, offset=offset)
    with open('{}.log'.format(filename), mode) as f:
        f.write(text + '\n')
    st = {'timestamp': timestamp, 'particle_number': particle_number,
          'total_num_of_particles': total_num_of_particles, 'progress': progress, 'status': status}
    json_file = '{}.json'.format(filename)
    tmp = tempfile.NamedTemporaryFile(delete=False, mode='w', dir=os.path.dirname(os.path.abspath(json_file)))
    json.dump(st, tmp, indent=4, separators=(',', ': '), sort_keys=True)
    tmp.close()
    shutil.move(tmp.name, json_file)


def srwl_uti_save_stat_wfr_emit_prop_multi_e_init(num_of_proc, num_part_avg_proc, _tot_num_part, _i_gr=None):
    """Creates ./__srwl_logs__/srwl_stat_wfr_emit_prop_multi_e_<timestamp>
    (reference srwlib.py:7814-7850), writes its start line; returns its path
    without extension."""
    import datetime
    import os
    log_dir = os.path.abspath('__srwl_logs__')
    os.makedirs(log_dir, exist_ok=True)
    timestamp = '{:%Y-%m-%d_%H-%M-%S}'.format(datetime.datetime.now())
    name = 'srwl_stat_wfr_emit_prop_multi_e_' + (repr(_i_gr) if _i_gr is not None else '') + timestamp
    path = os.path.join(log_dir, name)
    srwl_uti_save_stat_wfr_emit_prop_multi_e(0, _tot_num_part, filename=path, cores=num_of_proc,
                                             particles_per_iteration=num_part_avg_proc)
    return path


def srwl_wfr_fn_ext(_fp, _form, _char=0):
    """File name with the extension of the requested format (reference
    srwlib.py:8261-8298): '.dat' for ASCII, '.h5' for HDF5."""
    if _fp is None or _form is None:
        return _fp
    form = _form.upper()
    form = {'ASCII': 'ASC', 'HDF5': 'H5'}.get(form, form)
    if _char in (6, 61, 7):
        form = 'H5'
    ext = '.h5' if form == 'H5' else ('.dat' if form == 'ASC' else '')
    dot = _fp.rfind('.')
    core, orig = (_fp[:dot], _fp[dot:]) if dot >= 0 else (_fp, '')
    return _fp if orig == ext else core + ext


# ---------------------------------------------------------------------------
# Partially coherent emission and propagation (reference srwlib.py:8506)
# ---------------------------------------------------------------------------

def srwl_wfr_emit_prop_multi_e(_e_beam, _mag, _mesh, _sr_meth, _sr_rel_prec, _n_part_tot, _n_part_avg_proc=1,
                               _n_save_per=100, _file_path=None, _sr_samp_fact=-1, _opt_bl=None, _pres_ang=0,
                               _char=0, _x0=0, _y0=0, _e_ph_integ=0, _rand_meth=1, _tryToUseMPI=True, _wr=0.,
                               _wre=0., _det=None, _me_approx=0, _file_bkp=False, _rand_opt=False,
                               _file_form='ascii', _n_mpi=1, _n_cm=1000, _ms=0, _del_aux_files=False, _seed=None):
    """Total intensity of partially coherent SR emitted by the finite-emittance
    beam _e_beam in the field _mag, on _mesh, propagated through _opt_bl,
    averaged over _n_part_tot macro-electrons; same signature and result as
    the reference driver (srwlib.py:8506-11690) for the total intensity
    (_char = 0) in coordinates.

    The per-electron loop runs on the GPU (srwlEmitPropMultiE).  As in the
    reference: the result mesh is the mesh of the central electron's
    propagated wavefront (srwlib.py:9360-9395), the SR precision is
    [_sr_meth, _sr_rel_prec, 0, 0, 50000, 1, _sr_samp_fact] (:8878), the
    running average is written to _file_path (ASCII, extension '.dat') every
    _n_save_per electrons and at the end, with a progress log under
    ./__srwl_logs__ (:7746-7850, 10759-10881); returns the SRWLStokes result.
    _rand_meth 1 draws pseudo-random Gaussian electrons (a counter-based
    Philox stream; _seed fixes it, default from the clock), 2 and 3
    randomized quasi-Monte Carlo points (digitally shifted Sobol' points in
    16 replicates) in place of the reference's Halton / LPtau sequences.
    _n_part_avg_proc and _tryToUseMPI are accepted and have no effect (one
    GPU).  Not in this build (raises): _char other than 0, angular
    presentation, integration over photon energy, detectors, the emittance
    convolution approximation, randomized optics, HDF5 files, a mesh with
    more than one photon energy or whose sizes are not powers of two
    (16..2048)."""
    import time as _time
    from copy import deepcopy
    from . import srwlpy
    unsupported = []
    if _char != 0: unsupported.append('_char=%r' % _char)
    if _pres_ang != 0: unsupported.append('_pres_ang=%r' % _pres_ang)
    if _e_ph_integ != 0: unsupported.append('_e_ph_integ=%r' % _e_ph_integ)
    if _det is not None: unsupported.append('_det')
    if _me_approx != 0: unsupported.append('_me_approx=%r' % _me_approx)
    if _rand_opt: unsupported.append('_rand_opt')
    if _file_path is not None and _file_form.upper() not in ('ASCII', 'ASC', 'TXT'):
        unsupported.append('_file_form=%r' % _file_form)
    if _mesh.ne != 1: unsupported.append('_mesh.ne=%r' % _mesh.ne)
    pow2 = lambda n: n >= 16 and n <= 2048 and (n & (n - 1)) == 0
    if not (pow2(_mesh.nx) and pow2(_mesh.ny)): unsupported.append('_mesh %dx%d' % (_mesh.nx, _mesh.ny))
    if _sr_samp_fact is not None and _sr_samp_fact > 0: unsupported.append('_sr_samp_fact=%r' % _sr_samp_fact)
    if unsupported:
        raise Exception('srwl_wfr_emit_prop_multi_e: not supported by this build: ' + ', '.join(unsupported))
    n_tot = int(_n_part_tot)
    n_save = max(1, int(_n_save_per))
    prec = array('d', [_sr_meth, _sr_rel_prec, 0, 0, 50000, 1, _sr_samp_fact if _sr_samp_fact else -1])
    file_path = srwl_wfr_fn_ext(_file_path, _file_form, _char)
    wfr0 = SRWLWfr()
    wfr0.allocate(1, _mesh.nx, _mesh.ny)
    wfr0.mesh.set_from_other(_mesh)
    wfr0.partBeam = deepcopy(_e_beam)
    if _wr != 0.:
        wfr0.Rx = wfr0.Ry = _wr
    if _wre != 0.:
        wfr0.dRx = wfr0.dRy = _wre
    # result mesh: the central electron's wavefront after the beamline
    mesh_res = SRWLRadMesh()
    mesh_res.set_from_other(_mesh)
    if _opt_bl is not None:
        wc = deepcopy(wfr0)
        srwlpy.CalcElecFieldSR(wc, 0, _mag, prec)
        srwlpy.PropagElecField(wc, deepcopy(_opt_bl))
        mesh_res.set_from_other(wc.mesh)
    res = SRWLStokes(_arS=None, _typeStokes='f')
    res.allocate(1, mesh_res.nx, mesh_res.ny, _n_comp=1)
    res.mesh.set_from_other(mesh_res)
    res.mesh.ne = 1
    import numpy as np
    acc = np.zeros(mesh_res.nx * mesh_res.ny, dtype=np.float64)
    seed = int(_seed) if _seed is not None else int(_time.time() * 1e6) & 0xffffffff
    rqmc = 1 if _rand_meth in (2, 3) else 0
    log_path = srwl_uti_save_stat_wfr_emit_prop_multi_e_init(1, n_save, n_tot) if file_path is not None else None
    done, chunk = 0, 0
    while done < n_tot:
        n = min(n_save, n_tot - done)
        if rqmc:
            n = max(16, (n // 16) * 16)
        st = SRWLStokes(_arS=None, _typeStokes='f')
        st.allocate(1, mesh_res.nx, mesh_res.ny, _n_comp=1)
        st.mesh.set_from_other(mesh_res)
        st.mesh.ne = 1
        info = srwlpy.EmitPropMultiE(st, wfr0, _mag, deepcopy(_opt_bl) if _opt_bl is not None else None, prec, n,
                                     _seed=(seed + 0x9E3779B1 * chunk) & 0xffffffffffff, _rqmc=rqmc)
        used = int(info["electrons"])
        acc += np.frombuffer(st.arS, dtype=np.float32)[:acc.size] * float(used)
        done += used
        chunk += 1
        res.arS = array('f', (acc / done).astype(np.float32).tobytes())
        if file_path is not None:
            srwl_uti_save_intens(res.arS, res.mesh, file_path, 1,
                                 _arLabels=['Photon Energy', 'Horizontal Position', 'Vertical Position', 'Intensity'],
                                 _arUnits=['eV', 'm', 'm', 'ph/s/.1%bw/mm^2'], _mutual=0, _form=_file_form)
            srwl_uti_save_stat_wfr_emit_prop_multi_e(done, n_tot, filename=log_path)
    return res
