"""Benchmark of the single-electron emit-and-propagate path on one MI355X.

Workload (one step): the SR electric field of a 3 GeV electron in a planar
undulator (lambda_u = 20 mm, 100 periods, K = 1.5), 10001-point FP64
trajectory, observed at 20 m on a 201 x 201 mesh over +/-1 mm at 201 photon
energies spanning +/-5 % around the first harmonic.  `value` is the
throughput of srwlCalcElecFieldSR as the API delivers it: the field in the
caller's host arrays on return (the rate with the field left in HBM is
reported beside it).

Prints one JSON line with the single-electron field throughput (points/s),
the roofline of the dominant kernel (the SR integral), the CPU baseline (the
reference SRW sources when their runner is built, else the FP64 oracle, on
the host cores, bounded sample) and the per-energy relative L2 error of the
device field against the FP64 oracle and the reference.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_VALU_PEAK_TFLOPS = 157.3      # MI355X published FP32 vector peak (packed FMA)
HBM_PEAK_GBS = 8000.0
# FP32 flops the k_sr_integral loop executes per integrand evaluation: two
# phase FMAs, two amplitude FMAs and four accumulating FMAs (8 FMA = 16 flops);
# the hardware sine and cosine are not counted.
FLOPS_PER_EVAL = 16
# the same for k_sr_integral_tile<8>, per energy of a group of 8: four
# accumulating FMAs (8), the two recurrence FMAs for 4 of the 8 energies
# (4, over 2), and the per-s work shared by the 8 (four phase FMAs, the
# second chain's start phase FMA and two adds, the doubled cosine, two
# amplitude FMAs: 17, over 8)
FLOPS_PER_EVAL_TILE = 8 + 4 * 4 / 8 + 17 / 8
VALU_LANE_ISSUE_PER_S = 256 * 4 * 16 * 2.4e9   # CUs x SIMDs x lanes per cycle x clock


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=201, help="mesh points per axis")
    ap.add_argument("--ne", type=int, default=201, help="photon energies")
    ap.add_argument("--cpu-energies", type=int, default=24, help="energies in the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--api", type=int, default=1,
                    help="time the API path into host arrays (0: only the device-resident leg, for profiling)")
    ap.add_argument("--multi-e", type=int, default=1, help="also measure the multi-electron time to 1%% rel-L2")
    ap.add_argument("--me-n", type=int, default=512, help="multi-electron emission mesh")
    ap.add_argument("--me-max", type=int, default=4096, help="electron cap per multi-electron run")
    ap.add_argument("--uneven", type=int, default=2048, help="requests of the uneven batch (0: skip)")
    ap.add_argument("--prop", type=int, default=1, help="also run the resample + drift leg on the SR wavefront")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain device-to-device copy of known size, 5 times, for scaling the
    HBM byte counters of a profiler run (read MIB MiB, write MIB MiB per copy)."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    b.copy_(a)
    e0.record()
    for _ in range(5):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 6, "ms_per_copy": ms,
                      "read_plus_write_GBps": 2 * n * 4 / (ms * 1e-3) / 1e9}))


def multi_e_runs(srwlpy, W, n, nmax, seeds=8, validate=True):
    """Time to 1 % rel-L2 of the partially coherent intensity (Gaussian 6-D beam,
    sigma_x 30 um, sigma_x' 5 urad, sigma_y 5 um, sigma_y' 2 urad, sigma_E/E 1e-3;
    n x n mesh at 20 m, 30 m drift) for plain Gaussian MC, the learned path
    (the proposal sampled through randomized Sobol' points and used only where
    the electrons so far predict a gain) and randomized quasi-Monte Carlo with
    p itself, each on `seeds` seeds.  A run stops at its own error estimate
    (per-pixel sample variance of w*I, or the spread of 16 independently
    shifted replicates); with validate, each run's true error against a
    16384-electron RQMC reference is reported, and the time to a true 1 % is
    estimated per run as seconds * (true / 0.01)^2 (the Monte Carlo rate; it
    overstates the time of the quasi-Monte Carlo runs, whose error falls
    faster)."""
    from srw_amd.srwlib import SRWLStokes
    import numpy as np
    beam = W.beam_for_multi_e()
    e1 = W.first_harmonic_eV(beam.partStatMom1.gamma, 1.5, 0.02)
    w0 = W.wavefront(beam, 1, e1, e1, n, n, 1e-3, 1e-3, 20.)
    mag, prec, opt = W.planar_undulator(), W.traj_prec(10001), W.drift_container(30., 2)

    def stokes():
        st = SRWLStokes()
        st.allocate(1, n, n, _n_comp=1)
        st.mesh.eStart = st.mesh.eFin = e1
        st.mesh.xStart, st.mesh.xFin = -2.5e-3, 2.5e-3
        st.mesh.yStart, st.mesh.yFin = -2.5e-3, 2.5e-3
        return st

    res = {}
    ref = None
    if validate:
        st = stokes()
        info = srwlpy.EmitPropMultiE(st, w0, mag, opt, prec, 16384, _seed=31, _batch=1024, _rqmc=1, _replicates=16)
        ref = np.array(st.arS[:n * n], dtype=np.float64)
        res["reference"] = {"electrons": 16384, "sampler": "rqmc, 16 replicates", "rel_l2_est": info["rel_l2_est"],
                            "seconds": info["seconds"]}
        log("multi-electron reference done")
    methods = (("plain_mc", dict(_proposal=0)),
               ("learned", dict(_proposal=1, _train=32, _rqmc=1, _replicates=16)),
               ("rqmc", dict(_rqmc=1, _replicates=16)))
    for name, kw in methods:
        runs = []
        for k in range(seeds):
            st = stokes()
            info = srwlpy.EmitPropMultiE(st, w0, mag, opt, prec, nmax, _seed=20251202 + k, _batch=256, _target=0.01,
                                         _ideal_gain=(name == "plain_mc" and k == 0), **kw)
            r = {"seed": 20251202 + k, "seconds": info["seconds"], "electrons": int(info["electrons"]),
                 "rel_l2_est": info["rel_l2_est"]}
            if name == "learned":
                r["mixture_batches"] = int(info["mixture_batches"])
                r["predicted_ratio"] = info["predicted_ratio"]
            if name == "plain_mc" and k == 0:
                # variance ratio of plain MC to the variance-optimal proposal q* ~ p ||I||:
                # the most any importance-sampling proposal can save on this beam
                res["ideal_proposal_gain"] = info["ideal_gain"]
            if ref is not None:
                out = np.array(st.arS[:n * n], dtype=np.float64)
                r["rel_l2_true"] = float(np.linalg.norm(out - ref) / np.linalg.norm(ref))
                r["seconds_to_true_1pct"] = r["seconds"] * (r["rel_l2_true"] / 0.01) ** 2
            runs.append(r)
        t = np.array([r["seconds"] for r in runs])
        agg = {"seconds_mean": float(t.mean()), "seconds_min": float(t.min()), "seconds_max": float(t.max()),
               "electrons_mean": float(np.mean([r["electrons"] for r in runs])), "runs": runs}
        if ref is not None:
            tt = np.array([r["seconds_to_true_1pct"] for r in runs])
            agg["rel_l2_true_mean"] = float(np.mean([r["rel_l2_true"] for r in runs]))
            agg["rel_l2_true_max"] = float(np.max([r["rel_l2_true"] for r in runs]))
            agg["seconds_to_true_1pct_mean"] = float(tt.mean())
            agg["seconds_to_true_1pct_spread"] = float(tt.std())
        res[name] = agg
        log("multi-electron %s: %.3f s mean over %d seeds" % (name, agg["seconds_mean"], seeds))
    p, q, r = res["plain_mc"], res["learned"], res["rqmc"]
    res["speedup_learned_vs_plain"] = p["seconds_mean"] / q["seconds_mean"]
    res["speedup_learned_vs_plain_min_over_seeds"] = min(a["seconds"] / b["seconds"] for a, b in zip(p["runs"], q["runs"]))
    res["speedup_rqmc_vs_plain"] = p["seconds_mean"] / r["seconds_mean"]
    if ref is not None:
        res["speedup_learned_vs_plain_true_1pct"] = p["seconds_to_true_1pct_mean"] / q["seconds_to_true_1pct_mean"]
    res["workload"] = ("Gaussian beam, %d-electron cap, %dx%d mesh at 20 m, 30 m drift (analytic quadratic phase), "
                       "intensity re-gridded to %dx%d over +/-2.5 mm; %d seeds per method" % (nmax, n, n, n, n, seeds))
    return res


def sr_traffic(n, ne):
    """HBM bytes per k_sr_integral launch from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy), when it was
    measured for this workload; None otherwise."""
    import re
    path = os.path.join(ROOT, "profiles", "k_sr_integral_%dx%dx%d_traffic.md" % (n, n, ne))
    if not os.path.exists(path):
        return None
    m = re.search(r"hbm_bytes_per_launch\s*=\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def fft_traffic(n, slices):
    """HBM bytes of one drift (its three FFT passes) from the committed PMC
    summary for this size, when measured; None otherwise."""
    import re
    path = os.path.join(ROOT, "profiles", "fft_drift_%dx%dx%d_traffic.md" % (n, n, slices))
    if not os.path.exists(path):
        return None
    m = re.search(r"hbm_bytes_per_drift\s*=\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def sr_valu_per_eval(n, ne):
    """(vector, transcendental) instructions per integrand evaluation of
    k_sr_integral from the committed counter summary (SQ_INSTS_VALU and
    SQ_INSTS_VALU_TRANS_F32 per 64 evaluations); (None, None) when not
    measured for this workload."""
    import re
    path = os.path.join(ROOT, "profiles", "k_sr_integral_%dx%dx%d_sq_counters.md" % (n, n, ne))
    if not os.path.exists(path):
        return None, None
    txt = open(path).read()
    m = re.search(r"valu_instr_per_eval\s*=\s*([0-9.eE+]+)", txt)
    t = re.search(r"valu_trans_per_eval\s*=\s*([0-9.eE+]+)", txt)
    return (float(m.group(1)) if m else None), (float(t.group(1)) if t else None)


def uneven_batch(srwlpy, W, n):
    """n independent requests (meshes 128^2..1024^2, harmonics 1-5, K in
    [0.5, 2.5]) computed in one device pass; fields left in HBM."""
    wfrs, mags, prec = W.uneven_requests(n=n, seed=7, alloc=False)
    for w in wfrs:
        w.arMomX = w.arMomY = None      # no moments: field computation only
    pts = sum(w.mesh.nx * w.mesh.ny for w in wfrs)
    srwlpy.set_keep_on_device(True)
    srwlpy.reset_stats()
    t0 = time.perf_counter()
    srwlpy.CalcElecFieldSRBatch(wfrs, mags, prec)
    dt = time.perf_counter() - t0
    st = srwlpy.stats()
    return {"requests": n, "points": pts, "seconds": dt, "points_per_s": pts / dt,
            "sr_kernel_ms": st["sr_ms"], "sr_evals": st["sr_evals"], "sr_fp64_points": int(st["sr_fp64_points"]),
            "sr_kernel_tflops": FLOPS_PER_EVAL * st["sr_evals"] / (st["sr_ms"] * 1e-3) / 1e12 if st["sr_ms"] else None,
            "workload": "%d requests, meshes 128^2-1024^2 over +/-1 mm at 20 m, harmonic 1-5 of K in [0.5, 2.5], "
                        "one device pass, fields left in HBM" % n}


def c2_propagation(srwlpy, w, n_out=1024, L=20.):
    """Propagation leg of the 201-energy workload: the SR wavefront (already in
    host memory) resampled from n x n to n_out x n_out (range x2, i.e. +/-2 mm,
    resolution x n_out/(2n)) and drifted L metres with the quadratic phase
    treated analytically (arProp[3] = 1, radius from the wavefront moments)."""
    from srw_amd.srwlib import SRWLOptC, SRWLOptD
    import copy
    n, ne = w.mesh.nx, w.mesh.ne
    pm, pd = 2.0, n_out / (2.0 * n)
    opt = SRWLOptC([SRWLOptD(L)], [[0, 0, 1.0, 1, 0, pm, pd, pm, pd]])
    # untimed first propagation: the FFT tile autotuner searches its shapes once per size
    srwlpy.PropagElecField(copy.deepcopy(w), opt)
    srwlpy.reset_stats()
    t0 = time.perf_counter()
    srwlpy.PropagElecField(w, opt)
    wall = time.perf_counter() - t0
    st = srwlpy.stats()
    if (w.mesh.nx, w.mesh.ny) != (n_out, n_out):
        raise RuntimeError("propagation produced a %dx%d mesh" % (w.mesh.nx, w.mesh.ny))
    slices = 2 * ne
    alg = 64. * n_out * n_out * slices        # 128 N^2 bytes per energy (2 components x 2 FFTs x 2 sweeps x 16 B)
    fft_ms = st["fft_ms"]
    return {
        "workload": "%d energies, %dx%d resampled to %dx%d (range x%g, resolution x%.3f), %g m drift, arProp[3] = 1"
                    % (ne, n, n, n_out, n_out, pm, pd, L),
        "seconds_wall_incl_host_copies": wall,
        "device_ms": st["prop_device_ms"],
        "resize_ms": st["resize_ms"],
        "resize_gbs": st["resize_bytes"] / (st["resize_ms"] * 1e-3) / 1e9 if st["resize_ms"] else None,
        "drift_ms": fft_ms,
        "roofline_fft": {"bound": "hbm", "achieved": alg / (fft_ms * 1e-3) / 1e9 if fft_ms else None,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": alg / (fft_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if fft_ms else None,
                         "traffic": fft_traffic(n_out, slices), "kernels": "k_rows_fwd + k_cols + k_rows_inv",
                         "algorithmic_bytes": alg,
                         "design_pass_bytes": st["fft_bytes"],
                         "tiles_autotuned": srwlpy.fft_tiles(n_out.bit_length() - 1, n_out.bit_length() - 1, 3, 1)},
    }


def c1_methods(srwlpy, W, reps=20):
    """The CPU-runnable single-electron case (101 x 101 mesh over +/-1 mm at
    20 m, one photon energy at the first harmonic) with method 0 (manual,
    trajectory step) and method 1 (adaptive, relative precision 0.01, SRW's
    usual setting); fields left in HBM, wall time per call."""
    from array import array
    out = {}
    for name, prec in (("method0", None), ("method1_relprec_0.01", array('d', [1, 0.01, 0, 0, 10001, 1, 0]))):
        w, mag, p0 = W.single_electron(n=101)
        p = p0 if prec is None else prec
        srwlpy.set_keep_on_device(True)
        srwlpy.CalcElecFieldSR(w, 0, mag, p)
        t0 = time.perf_counter()
        for _ in range(reps):
            srwlpy.CalcElecFieldSR(w, 0, mag, p)
        dt = (time.perf_counter() - t0) / reps
        out[name] = {"ms_per_call": 1e3 * dt, "points_per_s": 101 * 101 / dt}
        if prec is not None:
            import numpy as np
            lv = srwlpy.last_auto_levels()
            out[name]["levels"] = [int(lv.min()), int(lv.max())]
            out[name]["evaluations_per_point"] = float(np.mean(4.0 * 2.0 ** lv + 1))
    srwlpy.set_keep_on_device(False)
    out["workload"] = "101x101 mesh over +/-1 mm at 20 m, one photon energy (first harmonic), 10001-point trajectory"
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baselines(w_gpu, mag, prec, workers):
    """CPU legs on bounded samples of the same workload, beside the device.

    reference: the reference SRW sources' srwlCalcElecFieldSR (its runner
    built by oracle/ref.mk), one process per worker core, each computing one
    photon energy of the 201-energy grid on the 101 x 101 subgrid (every
    second point) of the 201 x 201 mesh; the reference's SR integral is
    serial, so cores scale by processes as its MPI ranks would.
    port: the FP64 oracle (oracle/, -O3 -march=x86-64-v3), one OpenMP thread per
    worker core, 4 energies on the 201 x 201 mesh; 1 warm-up, median of 5.
    Both also check the device field of the timed run: relative L2 of Ex, Ey
    and the intensity per sampled energy."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    from srw_amd import workloads as W
    from tests import oracle_lib as oracle
    from tests import ref_lib
    from tests.util import cfield, clone_wfr, rel_l2
    n, ne = w_gpu.mesh.nx, w_gpu.mesh.ne
    eg = np.linspace(w_gpu.mesh.eStart, w_gpu.mesh.eFin, ne)
    exg, eyg = cfield(w_gpu.arEx, w_gpu), cfield(w_gpu.arEy, w_gpu)

    def errs(ie, wr, step):
        a, b = exg[::step, ::step, ie], cfield(wr.arEx, wr)[:, :, 0]
        c, d = eyg[::step, ::step, ie], cfield(wr.arEy, wr)[:, :, 0]
        return {"energy_index": int(ie), "eV": float(eg[ie]), "Ex": rel_l2(a, b), "Ey": rel_l2(c, d),
                "I": rel_l2(np.abs(a) ** 2 + np.abs(c) ** 2, np.abs(b) ** 2 + np.abs(d) ** 2)}

    out = {}
    # the reference
    if ref_lib.available():
        picks = sorted(set(int(round(k * (ne - 1) / max(workers - 1, 1))) for k in range(workers)))
        sub = (n + 1) // 2

        def one(ie):
            w, _, _ = W.single_electron(n=sub, ne=1)
            w.mesh.eStart = w.mesh.eFin = float(eg[ie])
            rc, secs = ref_lib.calc_field(w, mag, prec)
            if rc > 0:
                raise RuntimeError("reference runner failed: %d" % rc)
            return ie, w, secs

        t0 = time.perf_counter()
        with ThreadPoolExecutor(len(picks)) as ex:
            runs = list(ex.map(one, picks))
        wall = time.perf_counter() - t0
        secs = sorted(r[2] for r in runs)
        med = secs[len(secs) // 2]
        out["reference"] = {"value": len(picks) * sub * sub / med, "unit": "points/s", "cores": len(picks),
                            "kind": "reference",
                            "sample": "%d processes, each one photon energy of the 201-energy grid on the %dx%d "
                                      "subgrid (every second point) of the %dx%d mesh; median per-process call "
                                      "%.2f s, wall %.2f s" % (len(picks), sub, sub, n, n, med, wall),
                            "cpu_model": cpu_model(),
                            "rel_l2_device_vs_reference": [errs(ie, w, 2) for ie, w, _ in runs]}
        log("cpu baseline (reference) done")
    # the FP64 oracle
    picks = [0, ne // 3, (2 * ne) // 3, ne - 1]
    w, _, _ = W.single_electron(n=n, ne=len(picks))
    times = []
    for rep in range(6):
        wo = clone_wfr(w)
        wo.mesh.eStart, wo.mesh.eFin = float(eg[picks[0]]), float(eg[picks[-1]])
        wo.mesh.ne = 2
        # four separate energies: one call per pair keeps the grid's own values
        t0 = time.perf_counter()
        res = []
        for a, b in ((picks[0], picks[1]), (picks[2], picks[3])):
            wp, _, _ = W.single_electron(n=n, ne=2)
            wp.mesh.eStart, wp.mesh.eFin = float(eg[a]), float(eg[b])
            rc = oracle.calc_field(wp, mag, prec, nthreads=workers)
            if rc > 0:
                raise RuntimeError("oracle failed: %d" % rc)
            res.append(wp)
        if rep > 0:
            times.append(time.perf_counter() - t0)
    med = sorted(times)[len(times) // 2]
    port_errs = []
    for wp, (a, b) in zip(res, ((picks[0], picks[1]), (picks[2], picks[3]))):
        for k, ie in enumerate((a, b)):
            w1 = clone_wfr(wp)
            w1.mesh.ne = 1
            w1.arEx = np.asarray(cfield(wp.arEx, wp)[:, :, k]).astype(np.complex64).view(np.float32).ravel()
            w1.arEy = np.asarray(cfield(wp.arEy, wp)[:, :, k]).astype(np.complex64).view(np.float32).ravel()
            port_errs.append(errs(ie, w1, 1))
    out["port"] = {"value": len(picks) * n * n / med, "unit": "points/s", "cores": workers, "kind": "port",
                   "sample": "4 photon energies of the 201-energy grid on the %dx%d mesh; 1 warm-up, median of 5 "
                             "(%.2f s)" % (n, n, med),
                   "cpu_model": cpu_model(),
                   "build": "oracle/Makefile: gcc -O3 -march=x86-64-v3 -ffp-contract=off -fopenmp (built off the box; AVX2 is the ISA level both hosts have; no FMA contraction, as the reference)",
                   "rel_l2_device_vs_oracle": port_errs}
    log("cpu baseline (port) done")
    return out


def main():
    a = parse()
    if a.copy_calib:
        copy_calibration(a.copy_calib)
        return
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (requested --gpus %d)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    from srw_amd import srwlpy
    from srw_amd import workloads as W

    w, mag, prec = W.single_electron(n=a.n, ne=a.ne, de_rel=0.05)
    pts = a.n * a.n * a.ne
    # the API's own path: the field in the caller's host arrays on return
    srwlpy.set_keep_on_device(False)
    srwlpy.set_timing(True)
    n_api = a.steps if a.api else 1
    for _ in range(a.warmup if a.api else 1):
        srwlpy.CalcElecFieldSR(w, 0, mag, prec)
    srwlpy.reset_stats()
    t0 = time.perf_counter()
    call_s, streamed = [], []
    for _ in range(n_api):
        tc = time.perf_counter()
        srwlpy.CalcElecFieldSR(w, 0, mag, prec)
        call_s.append(time.perf_counter() - tc)
        streamed.append(srwlpy.last_streamed())
    dt = (time.perf_counter() - t0) * a.steps / n_api     # time of a.steps calls
    st = srwlpy.stats()
    value = pts * a.steps / dt
    k_ms = st["sum_sr_ms"] / max(st["n_sr"], 1)
    evals = st["sum_sr_evals"] / max(st["n_sr"], 1)
    fp64_points = int(st["sr_fp64_points"])
    log("single-electron field into host arrays: %.4g points/s" % value)
    # the same with the field left in HBM (device throughput)
    srwlpy.set_keep_on_device(True)
    wd = W.single_electron(n=a.n, ne=a.ne, de_rel=0.05)[0]
    srwlpy.CalcElecFieldSR(wd, 0, mag, prec)
    srwlpy.reset_stats()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        srwlpy.CalcElecFieldSR(wd, 0, mag, prec)
    dt_dev = time.perf_counter() - t1
    st_dev = srwlpy.stats()
    srwlpy.set_keep_on_device(False)
    k_ms_dev = st_dev["sum_sr_ms"] / max(st_dev["n_sr"], 1)
    # executed FP32 flops per evaluation: the energy-tile kernel takes the
    # energies of whole groups of 8, the single-energy kernel the rest
    tiled = (a.ne // 8) * 8 / a.ne if a.ne >= 8 else 0.
    flops_per_eval = tiled * FLOPS_PER_EVAL_TILE + (1. - tiled) * FLOPS_PER_EVAL
    achieved = flops_per_eval * evals / (k_ms_dev * 1e-3) / 1e12
    vpe, tpe = sr_valu_per_eval(a.n, a.ne)
    # vector instructions issued per second over one wave64 instruction per 4
    # cycles per SIMD; and the same with each transcendental taking two slots
    # (they issue at a lower rate than FMAs)
    issue = vpe * evals / (k_ms_dev * 1e-3) / VALU_LANE_ISSUE_PER_S if vpe else None
    issue2 = (vpe + tpe) * evals / (k_ms_dev * 1e-3) / VALU_LANE_ISSUE_PER_S if vpe and tpe is not None else None
    c1 = c1_methods(srwlpy, W)
    log("101x101 methods 0 and 1 done")
    prop = None
    if a.prop:
        import copy
        prop = c2_propagation(srwlpy, copy.deepcopy(w))
        log("propagation leg done")
    me = multi_e_runs(srwlpy, W, a.me_n, a.me_max) if a.multi_e else None
    if me:
        log("multi-electron runs done")
    ub = uneven_batch(srwlpy, W, a.uneven) if a.uneven > 0 else None
    if ub:
        log("uneven batch done")
    cpu = None
    cpu_all = None
    if not a.no_cpu:
        workers = min(os.cpu_count() or 1, 16)
        cpu_all = cpu_baselines(w, mag, prec, workers)
        cpu = dict(cpu_all.get("reference") or cpu_all["port"])
    out = {
        "metric": "single-e field pts/s",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * dt / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "precision": ("FP32 integrand, 64-point FP32 block sums added with Kahan compensation per 256-point chunk, "
                      "FP64 chunk sums; FP64 trajectory, tables and per-chunk phase reference; points failing the "
                      "FP32 condition check recomputed in FP64 (%d in the timed calls' last pass)" % fp64_points),
        "data": "synthetic",
        "config": {"workload": "planar undulator 3 GeV, lambda_u 20 mm x 100, K 1.5; %dx%d mesh at 20 m, %d photon "
                               "energies +/-5%% around the 1st harmonic; 10001-point trajectory, method 0; field "
                               "delivered into the caller's host arrays" % (a.n, a.n, a.ne),
                   "mesh": [a.n, a.n], "energies": a.ne, "trajectory_points": 10001, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_VALU_PEAK_TFLOPS, "traffic": sr_traffic(a.n, a.ne),
                     "kernel": "k_sr_integral_tile<8, false> + k_sr_integral", "kernel_ms": k_ms_dev,
                     "evals_per_launch": evals, "flops_per_eval": flops_per_eval,
                     "flops_counted": ("executed FP32 flops, an FMA counted as 2 (sin/cos excluded): 16 per evaluation "
                                       "in the single-energy kernel; 12.125 in the energy-tile kernel (4 accumulating "
                                       "FMAs, 2 recurrence FMAs for half of the energies, 1/8 of the per-s phase, "
                                       "chain-start and amplitude work)"),
                     "valu_instr_per_eval": vpe, "valu_trans_per_eval": tpe, "valu_issue_frac": issue,
                     "valu_issue_frac_trans_two_slots": issue2,
                     "nonpacked_fma_ceiling_frac": 0.5},
        "device_resident_points_per_s": pts * a.steps / dt_dev,
        "sr_kernel_ms_in_api_run": k_ms,
        "api_call_ms": {"min": 1e3 * min(call_s), "median": 1e3 * sorted(call_s)[len(call_s) // 2],
                        "max": 1e3 * max(call_s), "streamed_rows": sum(streamed), "calls": len(call_s)},
        "single_electron_101x101": c1,
        "propagation": prop,
        "multi_electron": me,
        "uneven_batch": ub,
        "cpu_baseline": cpu,
        "cpu_legs": cpu_all,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
