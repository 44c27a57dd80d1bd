"""The srwl_wfr_emit_prop_multi_e front end (srw_amd.srwlib): its result
files in SRW's ASCII layout (reference srwlib.py:6815-6923) and progress log
(:7746-7850) on the CPU, argument checks, and, on the GPU, a run with
periodic saves whose file equals the returned result and whose average
equals the oracle's multi-electron sum over the same draws."""
import json
import os

import numpy as np
import pytest

from srw_amd import srwlib as S
from srw_amd import workloads as W


def test_ascii_intensity_file_round_trip(tmp_path):
    m = S.SRWLRadMesh(1999.5, 1999.5, 1, -2.5e-3, 2.5e-3, 7, -1e-3, 1.5e-3, 5)
    vals = np.random.default_rng(0).random(35).astype(np.float32)
    fp = str(tmp_path / "res.dat")
    S.srwl_uti_save_intens_ascii(vals, m, fp)
    lines = open(fp).read().splitlines()
    assert lines[0] == ("#Intensity [ph/s/.1%bw/mm^2] (C-aligned, inner loop is vs Photon Energy, "
                        "outer loop vs Vertical Position)")
    assert lines[1] == "#1999.5 #Initial Photon Energy [eV]"
    assert lines[6] == "#7 #Number of points vs Horizontal Position"
    assert lines[10] == "#1 #Number of components"
    assert len(lines) == 11 + 35
    back, mb = S.srwl_uti_read_intens_ascii(fp)
    assert np.array_equal(np.array(back, dtype=np.float32), vals)
    assert mb.is_equal(m) and (mb.ne, mb.nx, mb.ny) == (1, 7, 5)


def test_progress_log(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    path = S.srwl_uti_save_stat_wfr_emit_prop_multi_e_init(1, 100, 250)
    assert os.path.dirname(path) == str(tmp_path / "__srwl_logs__")
    S.srwl_uti_save_stat_wfr_emit_prop_multi_e(100, 250, filename=path)
    S.srwl_uti_save_stat_wfr_emit_prop_multi_e(250, 250, filename=path)
    log = open(path + ".log").read().splitlines()
    assert log[0].endswith("Calculation on 1 core with averaging of 100 particles/iteration.")
    assert log[1].endswith("Running  100 out of 250 ( 40.00% complete)")
    assert log[2].endswith("Finished 250 out of 250 (100.00% complete)")
    st = json.load(open(path + ".json"))
    assert st["status"] == "Finished" and st["particle_number"] == 250 and st["progress"] == 100.0


def test_file_name_extension():
    assert S.srwl_wfr_fn_ext("out/res_int_pr_me.txt", "ascii") == "out/res_int_pr_me.dat"
    assert S.srwl_wfr_fn_ext("res.dat", "asc") == "res.dat"
    assert S.srwl_wfr_fn_ext("res", "hdf5") == "res.h5"
    assert S.srwl_wfr_fn_ext(None, "ascii") is None


def _beam_mesh(n=64):
    beam = W.beam_for_multi_e()
    e1 = W.first_harmonic_eV(beam.partStatMom1.gamma, 1.5, 0.02)
    mesh = S.SRWLRadMesh(e1, e1, 1, -1e-3, 1e-3, n, -1e-3, 1e-3, n, 20.)
    return beam, mesh


def test_unsupported_arguments_raise():
    beam, mesh = _beam_mesh()
    mag = W.planar_undulator()
    for kw in (dict(_char=1), dict(_pres_ang=1), dict(_e_ph_integ=1), dict(_me_approx=1), dict(_rand_opt=True),
               dict(_file_path="x.h5", _file_form="hdf5")):
        with pytest.raises(Exception, match="not supported"):
            S.srwl_wfr_emit_prop_multi_e(beam, mag, mesh, 0, 0.01, 10, **kw)
    m2 = S.SRWLRadMesh(mesh.eStart, mesh.eFin, 1, -1e-3, 1e-3, 100, -1e-3, 1e-3, 64, 20.)
    with pytest.raises(Exception, match="not supported"):
        S.srwl_wfr_emit_prop_multi_e(beam, mag, m2, 0, 0.01, 10)


@pytest.mark.gpu
def test_front_end_run_saves_and_matches_oracle(tmp_path, monkeypatch):
    from srw_amd import srwlpy  # noqa: F401  (the device library must load)
    from tests import oracle_lib as oracle
    from tests.util import rel_l2
    monkeypatch.chdir(tmp_path)
    beam, mesh = _beam_mesh(64)
    mag = W.planar_undulator()
    opt = W.drift_container(30., 2)
    fp = str(tmp_path / "res_int_pr_me.txt")
    res = S.srwl_wfr_emit_prop_multi_e(beam, mag, mesh, 0, 0.01, 48, _n_save_per=16, _file_path=fp, _opt_bl=opt,
                                       _seed=12345)
    vals, m = S.srwl_uti_read_intens_ascii(str(tmp_path / "res_int_pr_me.dat"))
    assert np.array_equal(np.array(vals, dtype=np.float32), np.frombuffer(res.arS, dtype=np.float32))
    assert m.is_equal(res.mesh)
    logs = os.listdir(tmp_path / "__srwl_logs__")
    log = [f for f in logs if f.endswith(".log")][0]
    lines = open(tmp_path / "__srwl_logs__" / log).read().splitlines()
    assert len(lines) == 4 and "Finished 48 out of 48" in lines[-1]
    # the same three 16-electron chunks through the oracle
    acc = np.zeros(res.mesh.nx * res.mesh.ny)
    for k in range(3):
        st = S.SRWLStokes()
        st.allocate(1, res.mesh.nx, res.mesh.ny, _n_comp=1)
        st.mesh.set_from_other(res.mesh)
        w0 = S.SRWLWfr()
        w0.allocate(1, mesh.nx, mesh.ny)
        w0.mesh.set_from_other(mesh)
        w0.partBeam = beam
        from array import array
        prec = array('d', [0, 0.01, 0, 0, 50000, 1, -1])
        seed = (12345 + 0x9E3779B1 * k) & 0xffffffffffff
        assert oracle.multi_e(st, w0, mag, W.drift_container(30., 2), prec, 16, seed=seed) == 0
        acc += np.frombuffer(st.arS, dtype=np.float32)[:acc.size] * 16
    e = rel_l2(np.frombuffer(res.arS, dtype=np.float32), acc / 48)
    print("front end vs oracle rel-L2 %.2e" % e)
    assert e < 2e-4
