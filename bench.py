#!/usr/bin/env python3
"""Benchmark of the MI355X fdfault time-stepping core.

Metric: Mpt-updates/s per RK stage = N_points * nstages * steps / t (whole
job, inputs resident in HBM, outputs off), plus the HBM-roofline fraction of
the dominant kernel (the fused bulk stage kernel) and an in-run CPU baseline
(the project's FP64 CPU oracle on a bounded sample of the same workload).
The default line reports the FP64 path as `value` and, in the same run, the
FP32-bulk / FP64-fault path with its speed ratio and its error against FP64
after --check-steps steps.  Other workloads: --workload planar2d, rough2d,
ensemble, stz3d, tpv4 (the reference's own
example problems/tpv4.in run whole through the drop-in executable, against
the run time the reference quotes for it).

Default workload: 3-D mixed-mode planar-fault rupture, two blocks of
401x201x201 points (grid 401x402x201), dx = 100 m, slip-weakening friction,
70 MPa shear + seeded Gaussian random field (std 5 MPa, 2 km correlation)
over -120 MPa normal stress, 3 km square nucleation patch, RK4.
A "step" is one time step (all RK stages) over the whole grid.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fdfault_amd.generators import gaussian_random_field, write_load_file  # noqa: E402
from fdfault_amd.problems import Pert, two_block_fault  # noqa: E402

HBM_PEAK_GBS = 8000.0


def workload_3d(d, nx=401, ny=201, nz=201, dx=0.1, sbporder=3, nt=1500, seed=1234):
    """The 3-D planar-fault workload, written as a .in file plus load file."""
    Lx, Lz = (nx - 1) * dx, (nz - 1) * dx
    tau = 70.0 + gaussian_random_field((nx, nz), dx, 2.0, 5.0, seed)
    lf = write_load_file(os.path.join(d, "load3d.bin"), np.full((nx, nz), -120.0), tau, np.zeros((nx, nz)))
    loads = [Pert("boxcar", 0., 0.5 * Lx, 1.5, 0.5 * Lz, 1.5, (0., 11.6, 0.))]
    p = two_block_fault(ndim=3, nx=nx, ny=ny, nz=nz, dx=dx, nt=nt, sbporder=sbporder, loads=loads,
                        datadir=os.path.join(d, "data") + "/", name="bench3d", loadfile=lf, front=("V", 0.001))
    os.makedirs(os.path.join(d, "data"), exist_ok=True)
    return p.write(os.path.join(d, "bench3d.in"))


def workload_2d(d, nx=801, ny=401, dx=0.04, sbporder=3, nt=2000, outputs=False):
    """2-D mode II planar fault: two blocks nx x ny, -120 MPa normal / 70 MPa
    shear, 3 km nucleation patch (+11.6 MPa) at x = 16 km, slip weakening.
    outputs: slip and slip rate along the whole fault at every step (two
    output units) besides the rupture front."""
    from fdfault_amd.problems import Output
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
             Pert("boxcar", 0., 16.0, 1.5, 0., 0., (0., 11.6, 0.))]
    outs = [Output("slip", "U", (0, nt, 1), (0, nx - 1, 1), (ny, ny, 1), (0, 0, 1)),
            Output("sliprate", "V", (0, nt, 1), (0, nx - 1, 1), (ny, ny, 1), (0, 0, 1))] if outputs else []
    p = two_block_fault(ndim=2, nx=nx, ny=ny, dx=dx, nt=nt, sbporder=sbporder, loads=loads, outputs=outs,
                        datadir=os.path.join(d, "data") + "/", name="planar2d", front=("V", 0.001))
    os.makedirs(os.path.join(d, "data"), exist_ok=True)
    return p.write(os.path.join(d, "planar2d.in"))


def workload_rough2d(d, nx=1601, ny=801, dx=0.025, sbporder=4, nt=4000, seed=7):
    """2-D mode II rough fault with Drucker-Prager plasticity: fault profile
    self-similar (Hurst 1, amplitude/wavelength 1e-2, wavelengths 0.2-20 km),
    curvilinear blocks nx x ny, s0 = (sxx -100, sxy 52, syy -120, szz -55) MPa
    (szz = nu (sxx + syy) for plane strain), plastic mu 0.7, c 0, beta 0.35,
    eta = 0.05 s x G, slip weakening, 3 km nucleation patch (+30 MPa shear) at
    the middle of the fault."""
    from fdfault_amd.generators import fault_surfaces, self_similar_profile
    os.makedirs(os.path.join(d, "data"), exist_ok=True)
    h = self_similar_profile(nx, dx, 1.0, 1e-2, 0.2, 20.0, seed)
    surf = fault_surfaces(d, nx, ny, dx, h)
    Lx = (nx - 1) * dx
    loads = [Pert("boxcar", 0., 0.5 * Lx, 1.5, 0., 0., (0., 30., 0.))]
    p = two_block_fault(ndim=2, nx=nx, ny=ny, dx=dx, nt=nt, sbporder=sbporder, loads=loads, surf=surf,
                        material="plastic", plastic=dict(mu=0.7, c=0., beta=0.35, eta=0.05 * 32.04),
                        s0=(-100., 52., 0., -120., 0., -55.), datadir=os.path.join(d, "data") + "/",
                        name="rough2d", front=("V", 0.001))
    return p.write(os.path.join(d, "rough2d.in"))


def single_workload(args):
    """Plain single-problem workloads on one GPU: 2-D planar (FP64 by default)
    and 2-D rough-fault plastic (FP32 bulk / FP64 fault by default)."""
    from fdfault_amd import lib
    rough = args.workload == "rough2d"
    if "--precision" not in sys.argv:
        args.precision = "fp32" if rough else "fp64"
    prec = lib.FP32 if args.precision == "fp32" else lib.FP64
    w = 4 if prec == lib.FP32 else 8
    with tempfile.TemporaryDirectory() as d:
        path = workload_rough2d(d) if rough else workload_2d(d)
        # kernel times from eager launches (HIP events around each launch)
        g = lib.Solver(path, precision=prec)
        print("opened: N=%d dt=%.6g tiled blocks %d" % (g.N, g.dt, g.info.tiled_blocks), flush=True)
        g.step(args.warmup)
        g.bulk_kernel_ms()
        g.fault_kernel_ms()
        g.step(args.steps)
        kms, nl = g.bulk_kernel_ms()
        fms, _ = g.fault_kernel_ms()
        npts, nst, nb, nt_ = g.N, g.nstages, g.info.nblocks, g.nt
        tiled = g.info.tiled_blocks == nb
        curv = g.info.curv_blocks == nb
        fused = g.info.fused_faults > 0  # the stage's fault solve inside the bulk launch
        g.close()
        # throughput: whole time steps replayed as hipGraphs
        # (a step of these grids takes 0.1-1 ms: at least 200 steps are timed,
        # so that the figure does not hang on a few ms of host timing)
        nts = max(args.steps, 200)
        g = lib.Solver(path, precision=prec, graph=True)
        g.step(args.warmup)  # (step returns after the stream has drained)
        t0 = time.perf_counter()
        g.step(nts)
        t = time.perf_counter() - t0
        g.close()
        with_out = None
        if not rough:
            # the same run writing its output units (slip and slip rate along
            # the fault at every step) and front: steps, then the final drain
            # of the staged output copies and the front file (finish)
            od = os.path.join(d, "out")
            os.makedirs(od)
            po = workload_2d(od, outputs=True)
            g = lib.Solver(po, precision=prec, graph=True, write_outputs=True)
            g.step(args.warmup)
            no = max(args.steps, 200)  # (finish's front file amortized over 200 steps or more)
            t0 = time.perf_counter()
            g.step(no)
            g.finish()
            to = time.perf_counter() - t0
            g.close()
            nbytes = sum(os.path.getsize(os.path.join(od, "data", fn)) for fn in os.listdir(os.path.join(od, "data")))
            with_out = {"value": round(npts * nst * no / to / 1e6, 2), "ms_per_step": round(to / no * 1e3, 3),
                        "steps": no, "outputs": "slip and slip rate along the whole fault at every step, rupture front",
                        "bytes_written": nbytes}
        cpu = None
        if not args.no_cpu:
            threads, host = cpu_threads()
            cv, cn = cpu_baseline(path, args.cpu_steps, threads)
            cpu = {"value": round(cv, 3), "unit": "Mpt-updates/s", "cores": threads, "kind": "port",
                   "sample": "the same problem (%d pts), %d steps, FP64 CPU oracle, OpenMP on %s"
                             % (cn, args.cpu_steps, host)}
    nf = 6 if rough else 5
    # the bulk kernel's own algorithmic bytes (the plastic fields are read and
    # written once per step by the plasticity kernel, not by this kernel)
    B = bytes_per_point(nf, w, curvilinear=rough, ndim=2)
    B_need = bytes_per_point(nf, w, curvilinear=rough, ndim=2, needed=True)
    B_stage = bytes_per_point(nf, w, curvilinear=rough, ndim=2, plastic_words=2 * (4 + 2) if rough else 0)
    # point updates per bulk launch (blocks may share one launch)
    ppl = npts * nst * args.steps / nl
    # one launch per stage covers every block when the blocks are merged
    merged = nb > 1 and nl == nst * args.steps
    if tiled:
        kernel = "bulk_fault2d_batch" if fused else ("bulk_tiled2d_batch" if merged else "bulk_tiled2d")
    elif curv:
        kernel = "bulk_curv2d_multi" if merged else "bulk_curv2d"
    else:
        kernel = "bulk_kernel"
    wname = "rough2d_plastic_2x1601x801" if rough else "planar2d_2x801x401"
    from fdfault_amd.profiling import lookup_traffic
    tr = lookup_traffic("%s/%s/sbp%d/%s" % (wname, args.precision, 4 if rough else 3, kernel))
    line = {
        "metric": "Mpt-updates/s per RK stage (%s)" % ("2-D rough-fault plastic" if rough else "2-D planar fault"),
        "value": round(npts * nst * nts / t / 1e6, 2), "unit": "Mpt-updates/s", "n_gpus": 1,
        "steps": nts, "warmup": args.warmup, "ms_per_step": round(t / nts * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64" if w == 8 else "f32",
        "data": "synthetic",
        "config": {"workload": wname, "points": npts,
                   "rk_stages": nst, "sbporder": 4 if rough else 3, "precision": args.precision, "nt": nt_,
                   "minutes_per_full_run": round(t / nts * nt_ / 60., 3)},
        "roofline": {"bound": "hbm", "achieved": round(B * ppl / (kms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(B * ppl / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": tr["traffic"] if tr else None, "traffic_source": tr["summary"] if tr else None,
                     "kernel": kernel, "avg_launch_ms": round(kms, 4),
                     "launches": nl, "bytes_per_point": B, "bytes_per_point_stage_with_plasticity": B_stage,
                     "bytes_per_point_needed": B_need,
                     "frac_at_bytes_needed": round(B_need * ppl / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "points_per_launch": int(ppl),
                     "fault_ms_per_stage": round(fms, 4),
                     "fault_solve_in_bulk_launch": fused,
                     "note": "working set %.0f MB: may be resident in the 256 MB Infinity Cache" % (
                         npts * (4 * nf + (5 if rough else 0)) * w / 1e6)},
        "cpu_baseline": cpu,
    }
    if with_out:
        line["with_outputs"] = with_out
    print(json.dumps(line), flush=True)


def compare_precision(args):
    """FP32 bulk / FP64 fault against FP64: per-field relative error of f,
    final slip and slip-rate error, front error in steps.  The 3-D planar
    workload, or (--workload planar2d / rough2d) the 2-D planar and the 2-D
    rough-fault plastic ones over their full run lengths by default."""
    from fdfault_amd import lib
    res = {}
    w2 = args.workload in ("planar2d", "rough2d")
    with tempfile.TemporaryDirectory() as d:
        if args.workload == "planar2d":
            path, wname = workload_2d(d), "planar2d_2x801x401"
            names = ["vx", "vy", "sxx", "sxy", "syy"]
        elif args.workload == "rough2d":
            path, wname = workload_rough2d(d), "rough2d_plastic_2x1601x801"
            names = ["vx", "vy", "sxx", "sxy", "syy", "szz", "lambda", "gammap"]
        else:
            path = workload_3d(d, args.nx, args.ny, args.nz, sbporder=args.sbporder)
            wname = "3d_planar_fault_slipweak_2x%dx%dx%d" % (args.nx, args.ny, args.nz)
            names = ["vx", "vy", "vz", "sxx", "sxy", "sxz", "syy", "syz", "szz"]
        if args.compare_steps <= 0:
            args.compare_steps = 1500
            if w2:
                s0 = lib.Solver(path)
                args.compare_steps = s0.nt
                s0.close()
        for prec in (lib.FP64, lib.FP32):
            g = lib.Solver(path, precision=prec, graph=True)
            done = 0
            while done < args.compare_steps:
                k = min(100, args.compare_steps - done)
                g.step(k)
                done += k
                print("precision %d: %d steps" % (prec, done), flush=True)
            res[prec] = {"f": g.array("f"), "U": g.iface(0, "U"), "V": g.iface(0, "V"), "front": g.front(0),
                         "dt": g.dt}
            g.close()
    a, b = res[lib.FP32], res[lib.FP64]
    per = {n: float(np.abs(a["f"][q] - b["f"][q]).max() / max(np.abs(b["f"][q]).max(), 1e-300))
           for q, n in enumerate(names[:a["f"].shape[0]])}
    both = (a["front"] >= 0) & (b["front"] >= 0)
    line = {"metric": "FP32 bulk / FP64 fault vs FP64, " + wname,
            "config": {"workload": wname, "steps": args.compare_steps},
            "per_field_rel_err": per,
            "final_slip_rel_err": float(np.abs(a["U"] - b["U"]).max() / max(np.abs(b["U"]).max(), 1e-300)),
            "slip_rate_rel_err": float(np.abs(a["V"] - b["V"]).max() / max(np.abs(b["V"]).max(), 1e-300)),
            "front_err_steps": float(np.abs(a["front"][both] - b["front"][both]).max() / b["dt"]) if both.any() else 0.,
            "front_points_ruptured_in_one_only": int(((a["front"] >= 0) != (b["front"] >= 0)).sum()),
            "ruptured_points": int(both.sum())}
    print(json.dumps(line), flush=True)


STZ_PARAMS_3D = (1.e-6, 0.3, 0.01, 0.1, 1.e4, 1.e-3, 0.1, 0.4, 1.e4)  # v0 f0 a muy c0 R beta chiw v1


def workload_stz3d(d, seed, nx=301, ny=151, nz=151, dx=0.1, sbporder=3, nt=1000):
    """3-D STZ rupture: two blocks nx x ny x nz, -120 MPa normal / 70 MPa shear
    load, effective temperature seeded uniform on [0.06, 0.08] (state file)."""
    os.makedirs(os.path.join(d, "data"), exist_ok=True)
    chi = np.random.default_rng(seed).uniform(0.06, 0.08, size=(nx, nz))
    sp = os.path.join(d, "state.bin")
    chi.tofile(sp)
    stz = dict(chi0=0., statefile=sp, perts=[Pert("constant", 0., 0., 0., 0., 0., STZ_PARAMS_3D)])
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.))]
    p = two_block_fault(ndim=3, nx=nx, ny=ny, nz=nz, dx=dx, nt=nt, sbporder=sbporder, law="stz", stz=stz,
                        loads=loads, datadir=os.path.join(d, "data") + "/", name="stz3d", front=("V", 0.001))
    return p.write(os.path.join(d, "stz3d.in"))


def stz_proposal(args):
    """Learned STZ root proposal on the 3-D STZ workload: log roots from plain
    runs (4 seeds, one step each at different times), train on the GPU, then
    compare plain and proposal solves on a fresh seed."""
    from fdfault_amd import lib
    prec = lib.FP64 if args.precision == "fp64" else lib.FP32
    rows = []
    with tempfile.TemporaryDirectory() as d:
        for k, seed in enumerate((11, 12, 13, 14)):
            path = workload_stz3d(os.path.join(d, "s%d" % seed), seed, args.stz_nx, args.stz_ny, args.stz_nz)
            s = lib.Solver(path, precision=prec)
            s.step(8 * k)
            s.stz_log(400000)
            s.step(1)
            rows.append(s.stz_log_read())
            s.close()
            print("logged seed %d: %d solves" % (seed, rows[-1].shape[0]), flush=True)
        rows = np.concatenate(rows)
        t0 = time.perf_counter()
        P, losses = lib.mlp_train(rows, epochs=args.mlp_epochs, batch=1024, lr=3e-3, seed=0)
        ttrain = time.perf_counter() - t0
        print("trained on %d pairs in %.1f s, loss %.3g -> %.3g" % (rows.shape[0], ttrain, losses[0], losses[-1]),
              flush=True)
        path = workload_stz3d(os.path.join(d, "s99"), 99, args.stz_nx, args.stz_ny, args.stz_nz)
        res = {}
        for mode in ("plain", "exact_derivative_no_network", "proposal"):
            if mode == "exact_derivative_no_network":
                os.environ["FDF_STZ_NEWTON"] = "exact"
            s = lib.Solver(path, precision=prec)
            os.environ.pop("FDF_STZ_NEWTON", None)
            if mode == "proposal":
                s.stz_set_model(P)
            s.step(args.warmup)
            s.stz_stats()
            s.fault_kernel_ms()
            s.bulk_kernel_ms()
            t0 = time.perf_counter()
            s.step(args.steps)
            t = time.perf_counter() - t0
            st = s.stz_stats()
            fms, _ = s.fault_kernel_ms()
            bms, _ = s.bulk_kernel_ms()
            res[mode] = {"solves": st[0], "mean_newton_iterations": round(st[1] / max(st[0], 1), 3),
                         "proposals_used": st[2], "fallbacks": st[3], "fault_ms_per_stage": round(fms, 4),
                         "bulk_ms_per_launch": round(bms, 4), "ms_per_step": round(t / args.steps * 1e3, 3),
                         "U": s.iface(0, "U")}
            s.close()
            print(mode, {k: v for k, v in res[mode].items() if k != "U"}, flush=True)
        du = float(np.abs(res["proposal"]["U"] - res["plain"]["U"]).max() / max(np.abs(res["plain"]["U"]).max(), 1e-300))
        for m in res.values():
            del m["U"]
    npts = 2 * args.stz_nx * args.stz_ny * args.stz_nz
    line = {
        "metric": "STZ root solve: Newton iterations per solve, plain / learned proposal",
        "value": round(res["plain"]["mean_newton_iterations"] / max(res["proposal"]["mean_newton_iterations"], 1e-9), 3),
        "unit": "x fewer iterations", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": True, "vs_baseline": None, "dtype": "f64" if prec == lib.FP64 else "f32",
        "data": "synthetic",
        "config": {"workload": "stz3d_2x%dx%dx%d" % (args.stz_nx, args.stz_ny, args.stz_nz), "points": npts,
                   "stz": dict(zip(("v0", "f0", "a", "muy", "c0", "R", "beta", "chiw", "v1"), STZ_PARAMS_3D)),
                   "chi0": "uniform [0.06, 0.08], seed 99", "training_pairs": int(rows.shape[0]),
                   "training_seeds": [11, 12, 13, 14], "epochs": args.mlp_epochs, "train_s": round(ttrain, 2),
                   "final_train_loss": float(losses[-1])},
        "plain": res["plain"], "exact_derivative_no_network": res["exact_derivative_no_network"],
        "proposal": res["proposal"],
        "note": "plain = the reference's Newton (bracket midpoint start, the reference's df/dmu whose muy term has "
                "the wrong sign, stz.cpp:479); exact_derivative_no_network = same start, exact df/dmu; proposal = "
                "network start, exact df/dmu, fallback to plain",
        "final_slip_rel_diff_proposal_vs_plain": du,
    }
    print(json.dumps(line), flush=True)


def stage_field_words(nstages=5):
    """Words per evolved field per point, averaged over the stages of a time
    step, that a low-storage RK stage must move: read f and write f always,
    read df except where A = 0 (the first stage of every scheme of
    rk.cpp:42-91), write df except on the last stage (its only reader is the
    next step's first stage, which scales it by A = 0).  RK4 (5 stages):
    (3 + 4 + 4 + 4 + 3) / 5 = 3.6."""
    if nstages == 1:
        return 2.0
    return (3 + 4 * (nstages - 2) + 3) / nstages


def bytes_per_point(nf, w, curvilinear=False, ndim=3, plastic_words=0, nstages=5, needed=False):
    """Algorithmic bytes per point per RK stage: 4 words per field (read f
    and df, write df and f: a fused low-storage stage with on-chip stencil
    reuse) + geometry words (curvilinear) + plastic words per step spread
    over the stages.  needed=True: the f/df words a stage must move averaged
    over a step instead (stage_field_words: no df read where A = 0, no df
    write on the last stage)."""
    G = (ndim * ndim + 1) if curvilinear else 0
    fw = stage_field_words(nstages) if needed else 4.0
    return w * (fw * nf + G) + w * plastic_words / nstages


def cpu_threads():
    """Threads for the CPU baseline: every core this process may run on (its
    affinity mask, what `nproc` counts), limited by OMP_NUM_THREADS when the
    host sets it (this job's CPU share); returns the count and a description
    of the host."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    n = min(aff, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else aff
    return n, "%d threads (affinity %d cores, %d logical CPUs on the host, OMP_NUM_THREADS=%s)" % (
        n, aff, os.cpu_count() or 0, omp)


def cpu_baseline(sample_in, steps, threads):
    """Oracle (FP64 CPU restatement) timed on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import oracle_lib as O
    o = O.Oracle(sample_in)
    t0 = time.perf_counter()
    o.step(steps)
    t = time.perf_counter() - t0
    mpts = o.N * o.nstages * steps / t / 1e6
    n = o.N
    o.close()
    return mpts, n


def ensemble(args):
    """Batched ensemble of uneven 2-D ruptures: ruptures/hour, whole-ensemble
    Mpt-updates/s and, with --check-error, each instance's final-slip and
    rupture-front error against a bit-faithful FP64 single run."""
    from fdfault_amd import lib
    from fdfault_amd.ensemble import ensemble_workload
    prec = lib.FP64 if args.precision == "fp64" else lib.FP32
    w = 8 if prec == lib.FP64 else 4
    with tempfile.TemporaryDirectory() as d:
        paths, shapes = ensemble_workload(d, n=args.instances, nt=args.ens_nt)
        print("ensemble written: %d instances, %d points" % (len(paths), sum(a * b for a, b in shapes)), flush=True)
        t0 = time.perf_counter()
        b = lib.Batch(paths, precision=prec, graph=True)
        topen = time.perf_counter() - t0
        tun = b.tuning()
        print("opened in %.1f s, tuning %s" % (topen, tun), flush=True)
        b.step(args.warmup)
        t0 = time.perf_counter()
        b.step(args.steps)
        t = time.perf_counter() - t0
        npts = sum(b.points(i) for i in range(b.n))
        nst = b.nstages
        stage_ms = t / args.steps / nst * 1e3
        B = bytes_per_point(5, w)
        B_need = bytes_per_point(5, w, needed=True)
        achieved = B * npts / (stage_ms * 1e-3) / 1e9
        rph = b.n / (t / args.steps * args.ens_nt) * 3600.
        err = None
        if args.check_error:
            b.step(args.err_steps - args.warmup - args.steps)
            worst_slip, worst_front, unmatched = 0., 0., 0
            per = []
            for i, p in enumerate(paths):
                s = lib.Solver(p, precision=lib.FP64, exact=True)
                s.step(args.err_steps)
                u64, u32 = s.iface(0, "U"), b.iface(i, 0, "U")
                es = float(np.abs(u32 - u64).max() / max(np.abs(u64).max(), 1e-300))
                f64, f32 = s.front(0), b.front(i, 0)
                both = (f64 >= 0) & (f32 >= 0)
                ef = float(np.abs(f64[both] - f32[both]).max() / s.dt) if both.any() else 0.
                unmatched += int(((f64 >= 0) != (f32 >= 0)).sum())
                per.append({"slip_rel": es, "front_steps": ef})
                print("instance %d: slip err %.3g, front err %.3g steps" % (i, es, ef), flush=True)
                worst_slip, worst_front = max(worst_slip, es), max(worst_front, ef)
                s.close()
            err = {"steps": args.err_steps, "max_final_slip_rel_err": worst_slip,
                   "max_front_err_steps": worst_front, "front_points_ruptured_in_one_only": unmatched,
                   "reference": "FP64 bit-faithful single runs", "per_instance": per}
        b.close()
    from fdfault_amd.profiling import lookup_traffic
    tr = lookup_traffic("ensemble_2d_slipweak_%d_instances/%s/sbp3/bulk_tiled2d_batch_tw%d_cx%d"
                        % (len(paths), args.precision, tun["tile_width"], tun["chunk"])) \
        if not tun["persistent"] else None
    line = {
        "metric": "ruptures/hour (batched ensemble of uneven 2-D ruptures)",
        "value": round(rph, 1), "unit": "ruptures/hour", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if w == 8 else "f32", "data": "synthetic",
        "config": {"workload": "ensemble_2d_slipweak_%d_instances" % len(paths), "instances": len(paths),
                   "points": npts, "nt_per_rupture": args.ens_nt, "sbporder": 3, "precision": args.precision,
                   "tile_width": tun["tile_width"], "x_chunk": tun["chunk"],
                   "launch": "persistent grid" if tun["persistent"] else "one workgroup per tile",
                   "tuned_from_cache": tun["from_cache"]},
        "mpt_updates_per_s": round(npts * nst * args.steps / t / 1e6, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": tr["traffic"] if tr else None, "traffic_source": tr["summary"] if tr else None,
                     "traffic_note": "HBM bytes of one bulk_tiled2d_batch launch over every instance "
                                     "(the stage's fault kernel adds about 1.6 B per point)" if tr else None,
                     "kernel": "bulk_tiled2d_batch + fault_batch_kernel (whole stage)",
                     "stage_ms": round(stage_ms, 4), "bytes_per_point": B, "bytes_per_point_needed": B_need,
                     "frac_at_bytes_needed": round(B_need * npts / (stage_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "error_vs_fp64": err,
    }
    print(json.dumps(line), flush=True)


# The reference quotes one run time: problems/tpv4.in, "about 1 hour" on an
# 8-core Mac Pro using all processors (docs/examples/tpv4.rst:20), that is
# 4 100 802 points x 701 steps x 5 RK stages / 3600 s = 4.0 Mpt-updates/s.
TPV4_REF_WALL_S = 3600.0
TPV4_REF_MPTS = 4100802 * 701 * 5 / TPV4_REF_WALL_S / 1e6


def reference_tpv4(args):
    """The reference's own example problems/tpv4.in (kept verbatim as
    tests/golden/tpv4.in: 3-D, two 201x101x101 blocks, sbporder 4, 701
    steps, a fault output unit every 10 steps and the rupture front), run
    whole through the drop-in executable: process start, open, every step,
    the output and front files, files closed.  value: Mpt-updates/s per RK
    stage of that whole run; vs_baseline: against the rate the reference's
    quoted run time implies.  Also the stepping alone (library, graph
    replay, no outputs) and the FP64 CPU oracle on a few steps."""
    from fdfault_amd import lib
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from cases import reference_example
    nt = args.tpv4_steps
    exe = os.path.join(ROOT, "fdfault_amd", "_lib", "fdfault")
    with tempfile.TemporaryDirectory() as d:
        dd = os.path.join(d, "data") + "/"
        os.makedirs(dd)
        path = reference_example(d, "tpv4", nt, dd)
        env = dict(os.environ, FDFAULT_PRECISION=args.precision)
        t0 = time.perf_counter()
        r = subprocess.run([exe, path], capture_output=True, text=True, env=env, timeout=1200)
        wall = time.perf_counter() - t0
        if r.returncode != 0 or "Problem finished" not in r.stdout:
            raise RuntimeError("drop-in run failed: %s %s" % (r.stdout[-500:], r.stderr[-500:]))
        files = sorted(os.listdir(dd))
        nbytes = sum(os.path.getsize(os.path.join(dd, f)) for f in files)
        print("drop-in run: %.2f s, %d files, %d bytes" % (wall, len(files), nbytes), flush=True)
        g = lib.Solver(path, precision=lib.FP32 if args.precision == "fp32" else lib.FP64, graph=True)
        npts, nst = g.N, g.nstages
        g.step(args.warmup)
        t1 = time.perf_counter()
        g.step(args.steps)
        ts = time.perf_counter() - t1
        g.close()
        cpu = None
        if not args.no_cpu:
            threads, host = cpu_threads()
            cv, cn = cpu_baseline(path, args.cpu_steps, threads)
            cpu = {"value": round(cv, 3), "unit": "Mpt-updates/s", "cores": threads, "kind": "port",
                   "sample": "the same problem (%d pts), %d steps, FP64 CPU oracle, OpenMP on %s"
                             % (cn, args.cpu_steps, host)}
    value = npts * nst * nt / wall / 1e6
    line = {
        "metric": "Mpt-updates/s per RK stage, whole drop-in run of the reference's problems/tpv4.in",
        "value": round(value, 2), "unit": "Mpt-updates/s", "n_gpus": 1, "steps": nt, "warmup": 0,
        "ms_per_step": round(wall / nt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": round(value / TPV4_REF_MPTS, 1) if nt == 701 else None,
        "dtype": "f64" if args.precision == "fp64" else "f32", "data": "the reference's example input",
        "config": {"workload": "tpv4_2x201x101x101_sbp4", "points": npts, "rk_stages": nst, "nt": nt,
                   "precision": args.precision, "outputs": "fault slip rate every 10 steps, rupture front"},
        "wall_s": round(wall, 3), "files_written": len(files), "bytes_written": nbytes,
        "reference_quoted": {"wall_s": TPV4_REF_WALL_S, "mpt_updates_per_s": round(TPV4_REF_MPTS, 2),
                             "hardware": "8 core Mac Pro desktop, all processors (CPU model and year unstated)",
                             "source": "docs/examples/tpv4.rst:20"},
        "stepping_only": {"value": round(npts * nst * args.steps / ts / 1e6, 2),
                          "ms_per_step": round(ts / args.steps * 1e3, 3), "steps": args.steps,
                          "note": "library, graph replay, outputs off"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--sbporder", type=int, default=3)
    ap.add_argument("--nx", type=int, default=401)
    ap.add_argument("--ny", type=int, default=201)
    ap.add_argument("--nz", type=int, default=201)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--tpv4-steps", type=int, default=701, help="tpv4: time steps of the drop-in run")
    ap.add_argument("--workload", choices=["planar3d", "ensemble", "stz3d", "planar2d", "rough2d", "drift", "setup",
                                           "tpv4"],
                    default="planar3d")
    ap.add_argument("--compare-precision", action="store_true",
                    help="planar3d, planar2d, rough2d: FP32 vs FP64 errors after --compare-steps steps")
    ap.add_argument("--compare-steps", type=int, default=0,
                    help="0: the workload's run length (1500 steps for planar3d, nt of the 2-D workloads)")
    ap.add_argument("--no-stream", action="store_true",
                    help="planar3d: skip the copy / same-shape stream bandwidth references")
    ap.add_argument("--only", action="store_true",
                    help="planar3d: time only --precision (no second precision, no error check)")
    ap.add_argument("--check-steps", type=int, default=200,
                    help="planar3d: FP32 vs FP64 error check after this many steps (0: skip)")
    ap.add_argument("--stz-nx", type=int, default=301)
    ap.add_argument("--stz-ny", type=int, default=151)
    ap.add_argument("--stz-nz", type=int, default=151)
    ap.add_argument("--mlp-epochs", type=int, default=20)
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--ens-nt", type=int, default=2000, help="time steps of one rupture (ensemble)")
    ap.add_argument("--check-error", action="store_true", help="ensemble: compare with FP64 single runs")
    ap.add_argument("--err-steps", type=int, default=2000)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: fdfault_amd runs on one GPU only (the whole grid fits one MI355X); "
              "a multi-GPU slab decomposition is designed but not built", file=sys.stderr)
        sys.exit(2)
    if args.compare_precision and args.workload in ("planar3d", "planar2d", "rough2d"):
        compare_precision(args)
        return
    if args.workload in ("planar2d", "rough2d"):
        single_workload(args)
        return
    if args.workload == "drift":
        fp64_drift(args)
        return
    if args.workload == "tpv4":
        reference_tpv4(args)
        return
    if args.workload == "setup":
        setup_times(args)
        return
    if args.workload == "stz3d":
        if "--precision" not in sys.argv:
            args.precision = "fp32"
        stz_proposal(args)
        return
    if args.workload == "ensemble":
        if "--precision" not in sys.argv:
            args.precision = "fp32"
        ensemble(args)
        return
    from fdfault_amd import lib
    headline = args.precision
    other = "fp32" if headline == "fp64" else "fp64"
    workload = "3d_planar_fault_slipweak_2x%dx%dx%d" % (args.nx, args.ny, args.nz)
    res, err, cpu, exact = {}, None, None, None
    with tempfile.TemporaryDirectory() as d:
        path = workload_3d(d, args.nx, args.ny, args.nz, sbporder=args.sbporder)
        print("workload written", flush=True)
        for pname in ((headline,) if args.only else (headline, other)):
            res[pname] = run_planar(lib, path, pname, args, workload)
            print(pname, {k: v for k, v in res[pname].items() if k != "roofline"}, flush=True)
        exact = None
        if not args.only:
            g = lib.Solver(path, exact=True)
            assert g.info.tiled_blocks == 0
            g.step(2)
            t0 = time.perf_counter()
            g.step(10)
            te = time.perf_counter() - t0
            exact = {"value": round(g.N * g.nstages * 10 / te / 1e6, 2), "ms_per_step": round(te / 10 * 1e3, 3),
                     "steps": 10, "note": "FP64 with the reference's per-point association in every kernel "
                                          "(bit-faithful mode), eager launches"}
            g.close()
            print("exact fp64", exact, flush=True)
        if args.check_steps > 0 and not args.only:
            err = fp32_error(lib, path, args.check_steps)
            print("fp32 vs fp64 after %d steps: %s" % (args.check_steps, err), flush=True)
        if not args.no_cpu:
            sd = os.path.join(d, "cpu")
            os.makedirs(sd)
            snx, sny, snz = 201, 101, 101
            sp = workload_3d(sd, snx, sny, snz, sbporder=args.sbporder)
            threads, host = cpu_threads()
            cv, cn = cpu_baseline(sp, args.cpu_steps, threads)
            cpu = {"value": round(cv, 3), "unit": "Mpt-updates/s", "cores": threads, "kind": "port",
                   "sample": "same 3-D workload at 2x%dx%dx%d points (%d pts), %d steps, FP64 CPU oracle, "
                             "OpenMP on %s" % (snx, sny, snz, cn, args.cpu_steps, host)}
    h, o = res[headline], res.get(other)
    line = {
        "metric": "Mpt-updates/s per RK stage & % HBM peak; FP32 vs FP64 at <=1e-4 slip err",
        "value": h["value"], "unit": "Mpt-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": h["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if headline == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": workload, "points": h["points"], "rk_stages": h["rk_stages"],
                   "sbporder": args.sbporder, "precision": headline},
        "roofline": h["roofline"],
        "cpu_baseline": cpu,
        "%s_path" % other: {"value": o["value"], "ms_per_step": o["ms_per_step"], "roofline": o["roofline"]} if o else None,
        "fp32_over_fp64": round(res["fp32"]["value"] / res["fp64"]["value"], 3) if o else None,
        "fp64_bit_faithful_path": exact,
        "fp64_fast_vs_bit_faithful_drift": "profiles/fp64_fast_vs_exact_drift_full_runs.md",
        "error_fp32_vs_fp64": err,
    }
    print(json.dumps(line), flush=True)


def run_planar(lib, path, pname, args, workload):
    """Timed steps of one precision on the 3-D planar workload; the bulk
    kernel's average launch time from HIP events on its stream."""
    prec = lib.FP32 if pname == "fp32" else lib.FP64
    w = 4 if prec == lib.FP32 else 8
    g = lib.Solver(path, precision=prec)
    g.step(args.warmup)
    g.bulk_kernel_ms()
    t0 = time.perf_counter()
    g.step(args.steps)
    t = time.perf_counter() - t0
    kms, nlaunch = g.bulk_kernel_ms()
    npts, nst, nblocks = g.N, g.nstages, g.info.nblocks
    ntiled, nrows = g.info.tiled_blocks, g.info.row_blocks
    g.close()
    B = bytes_per_point(9, w)
    B_need = bytes_per_point(9, w, needed=True)
    # point updates per bulk launch (both blocks may share one launch)
    ppl = npts * nst * args.steps / nlaunch
    achieved = B * ppl / (kms * 1e-3) / 1e9
    kernel = ("bulk_rows3d" if nrows == nblocks else "bulk_tiled3d") if ntiled == nblocks else "bulk_kernel"
    from fdfault_amd.profiling import lookup_traffic
    tr = lookup_traffic("%s/%s/sbp%d/%s" % (workload, pname, args.sbporder, kernel))
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": tr["traffic"] if tr else None,
            "traffic_source": tr["summary"] if tr else None, "kernel": kernel,
            "avg_launch_ms": round(kms, 4), "launches": nlaunch, "bytes_per_point": B,
            "points_per_launch": int(ppl),
            # the same launches counted at the bytes a stage must move,
            # averaged over a step (3.6 nf words: no df read where A = 0, no
            # df write on the last stage)
            "bytes_per_point_needed": B_need,
            "frac_at_bytes_needed": round(B_need * ppl / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    ref = stream_reference(int(ppl), w) if not args.no_stream else None
    if ref:
        roof.update(ref)
        roof["frac_of_stream36"] = round(achieved / ref["stream36_GBs"], 4)
    return {"value": round(npts * nst * args.steps / t / 1e6, 2), "ms_per_step": round(t / args.steps * 1e3, 3),
            "points": npts, "rk_stages": nst, "roofline": roof}


def setup_times(args):
    """Open time (fdf_open / fdf_open_batch: parse, grid, metric, Jacobian,
    normals, dt, upload, t = 0 interface solve) with the grid built on the
    device and on the host, for the 3-D planar workload in FP64 and FP32 and
    for the 64-instance ensemble (bulk tile fixed, so no autotuning is
    timed)."""
    from fdfault_amd import lib
    from fdfault_amd.ensemble import ensemble_workload
    out = {}

    def timed(fn, host):
        if host:
            os.environ["FDF_HOST_SETUP"] = "1"
        else:
            os.environ.pop("FDF_HOST_SETUP", None)
        best = 1e30
        for _ in range(2):
            t0 = time.perf_counter()
            h = fn()
            best = min(best, time.perf_counter() - t0)
            h.close()
        os.environ.pop("FDF_HOST_SETUP", None)
        return round(best, 3)

    threads, host_desc = cpu_threads()
    with tempfile.TemporaryDirectory() as d:
        path = workload_3d(d, args.nx, args.ny, args.nz, sbporder=args.sbporder)
        lib.Solver(path).close()  # first HIP use (runtime and code-object load) outside the timing
        for pname, prec in (("fp64", lib.FP64), ("fp32", lib.FP32)):
            for host in (False, True):
                key = "3d_planar_2x%dx%dx%d_%s_%s" % (args.nx, args.ny, args.nz, pname, "host" if host else "device")
                out[key] = timed(lambda: lib.Solver(path, precision=prec), host)
                print(key, out[key], flush=True)
        ed = os.path.join(d, "ens")
        paths, shapes = ensemble_workload(ed, n=args.instances, nt=args.ens_nt)
        os.environ["FDF_BATCH_TW"], os.environ["FDF_BATCH_CX"] = "128", "4"
        for host in (False, True):
            key = "ensemble_%d_instances_fp32_%s" % (args.instances, "host" if host else "device")
            out[key] = timed(lambda: lib.Batch(paths, precision=lib.FP32, graph=True), host)
            print(key, out[key], flush=True)
    print(json.dumps({"workload": "open_time_seconds", "host": host_desc, "results": out}), flush=True)


def fp64_drift(args):
    """FP64 fast kernels (block-constant metric, fused multiply-adds) against
    the bit-faithful FP64 path over the full run of the 2-D planar workload
    (2x801x401, 2000 steps) and the 3-D planar workload (2x401x201x201, 1500
    steps): max relative difference of the fields, slip, slip rate and
    traction, and of the rupture times in time steps; also the bit-faithful
    path's throughput (whole steps, eager launches)."""
    from fdfault_amd import lib
    out = {}
    rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    with tempfile.TemporaryDirectory() as d:
        cases = (("2d_planar_2x801x401", workload_2d(d, nt=2000), 2000),
                 ("3d_planar_2x401x201x201", workload_3d(d, sbporder=3), 1500))
        for name, path, nt in cases:
            res = {}
            for exact in (False, True):
                g = lib.Solver(path, exact=exact, graph=not exact)
                t0 = time.perf_counter()
                done = 0
                while done < nt:
                    k = min(100, nt - done)
                    g.step(k)
                    done += k
                    print(name, "exact" if exact else "fast", "step", done, flush=True)
                t = time.perf_counter() - t0
                res[exact] = dict(f=g.array("f"), U=g.iface(0, "U"), V=g.iface(0, "V"), S=g.iface(0, "S"),
                                  front=g.front(0), dt=g.dt, mpts=g.N * g.nstages * nt / t / 1e6,
                                  tiled=g.info.tiled_blocks, blocks=g.info.nblocks)
                g.close()
            fa, ex = res[False], res[True]
            assert fa["tiled"] == fa["blocks"] and ex["tiled"] == 0
            both = (fa["front"] >= 0) & (ex["front"] >= 0)
            out[name] = {
                "steps": nt,
                "max_field_rel_diff": max(rel(fa["f"][q], ex["f"][q]) for q in range(ex["f"].shape[0])),
                "slip_rel_diff": rel(fa["U"], ex["U"]), "slip_rate_rel_diff": rel(fa["V"], ex["V"]),
                "traction_rel_diff": rel(fa["S"], ex["S"]),
                "front_diff_steps": float(np.abs(fa["front"][both] - ex["front"][both]).max() / ex["dt"])
                if both.any() else 0.,
                "front_points_ruptured_in_one_only": int(((fa["front"] >= 0) != (ex["front"] >= 0)).sum()),
                "ruptured_points": int(both.sum()),
                "fast_mpts": round(fa["mpts"], 1), "exact_mpts": round(ex["mpts"], 1),
            }
            print(name, json.dumps(out[name]), flush=True)
    print(json.dumps({"workload": "fp64_fast_vs_exact_drift", "results": out}), flush=True)


def stream_reference(points, w):
    """Achievable-bandwidth references measured on this GPU by the copy probe
    (fdfault_amd/_lib/copy_bw, a separate process): a plain 1 GiB copy, and a
    stream over the 3-D stage's own array shape (18 arrays read and 18
    written, `points` words of w bytes each: the algorithmic bytes of one
    block launch, with no stencil or re-reads); each the best of a small
    sweep of launch shapes (and, for the copy, non-temporal accesses)."""
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fdfault_amd", "_lib", "copy_bw")
    if not os.path.exists(exe):
        return None
    out = {}
    for key, argv in (("copy_GBs", [exe, "copybest", "1024", "20"]), ("stream36_GBs", [exe, "soa36", str(points * w // 4)])):
        r = subprocess.run(argv, capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            return None
        out[key] = float(r.stdout.strip().split("\n")[-1].split()[-2])
    return out


def fp32_error(lib, path, steps):
    """FP32 bulk / FP64 fault against FP64 after `steps` steps: max relative
    error of each field, of final slip and slip rate; front error in steps."""
    out = {}
    for prec in (lib.FP64, lib.FP32):
        g = lib.Solver(path, precision=prec, graph=True)
        g.step(steps)
        out[prec] = (g.array("f"), g.iface(0, "U"), g.iface(0, "V"), g.front(0), g.dt)
        g.close()
    (f64, u64, v64, t64, dt), (f32, u32, v32, t32, _) = out[lib.FP64], out[lib.FP32]
    rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    both = (t64 >= 0) & (t32 >= 0)
    return {"steps": steps, "max_field_rel_err": max(rel(f32[q], f64[q]) for q in range(f64.shape[0])),
            "final_slip_rel_err": rel(u32, u64), "slip_rate_rel_err": rel(v32, v64),
            "front_err_steps": float(np.abs(t32[both] - t64[both]).max() / dt) if both.any() else 0.,
            "front_points_ruptured_in_one_only": int(((t64 >= 0) != (t32 >= 0)).sum())}


if __name__ == "__main__":
    main()
