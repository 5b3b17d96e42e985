#!/usr/bin/env python3
"""Benchmark of the exact-diagonalisation hot path on one MI355X.

Workload: 3-orbital Kanamori impurity (U=2.5, J=0.4) + 9 bath sites (Ns=12 per spin),
seeded bath, sector (N_up, N_dn) = (6, 6), dimension 853,776 ("Ns=12 Sz=0").

A step is one Lanczos run of 128 iterations (the reference's Nitermax) in that
sector with the vectors resident in HBM; `value` is FP64 Lanczos iterations per
second.  Reported beside it: the mixed-precision (FP32 H.v, FP64 recurrence) rate,
the ground-state solve to convergence with |dE0| between the two, the G(i w_n)
excitation runs (6 c / c+ runs, 36 Lanczos steps, 2048 Matsubara points at
beta=100) batched vs one by one, the roofline of the fused H.v kernel, and the CPU
restatement of the reference algorithm timed on this host.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import json
from math import comb
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Lanczos iters/s (Ns=12 Sz=0) + batched G(iω) solves/s vs plain, |ΔE0|<1e-8"
HBM_PEAK_GBS = 8000.0
# Measured HBM bytes in the (6,6) sector (dim 853,776): rocprofv3 --pmc FETCH_SIZE and
# --pmc WRITE_SIZE passes over `bench.py --steps 1 --warmup 0 --no-cpu --no-green --no-scan
# --no-grad`, FETCH_SIZE scaled by 65536 / 32780.3 (a 64 MiB device copy reads as 32,780.3
# KiB), WRITE_SIZE by 1.0 (it reads 65,536 KiB); summary in profiles/lanczos_ns12_66_pmc.md.
# FP64 fused H.v per launch, by H.v schedule (8 and 16 share one kernel instance and were
# not separated: null):
HV_TRAFFIC_F64 = {18: 22816400, 19: 22834620, 20: 22855556}
# multi-column persistent Lanczos kernel k_lz_persist_cols<double, 8, 4> (4 columns per
# workgroup), per step; same passes with --schedule 20
PERSIST_TRAFFIC_F64_STEP = 21923560
# multi-run persistent kernel k_lz_persist_runs<double, 12, 4> of the batched G(i w_n) runs
# (6 runs of dim 731,808), per job-step (one run advanced one step)
RUNS_TRAFFIC_F64_JOB_STEP = 45275093


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_info():
    """Host CPU model, logical CPUs, the process's affinity mask and the thread count the
    CPU baseline uses: every core of the affinity mask, capped by OMP_NUM_THREADS when the
    environment sets it (a share of a larger host)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    threads = aff
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity_cpus": aff,
            "omp_num_threads": omp or None, "threads": max(1, threads)}


def cpu_baseline(model, nu, nd, e0_gpu, reps=5, niter=200):
    """The oracle (C + OpenMP restatement of HAIMupdo_multr + lanc_opt, same row split as
    openmp_split_array) on the host cores: 5 warm-up H.v, then the median of `reps` runs of
    `niter` Lanczos iterations; and the ground state + G(i w_n) of the metric workload timed
    end to end on the CPU (Lanczos_fast_diagonalize to 1e-12, the second pass for psi, and
    the 6 c / c+ excitation runs of 36 steps with the continued fraction on 2048 w_n)."""
    import oracle
    info = cpu_info()
    cores = info["threads"]
    o = oracle.OracleSector(model, nu, nd, nthreads=cores)
    start = oracle.start_vector(o.dim)
    for _ in range(5):
        o.hmult(start)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        o.lanczos_run(niter, start)
        times.append(time.perf_counter() - t0)
    dt = sorted(times)[reps // 2]
    # ground state + G on the CPU path
    wn = oracle.matsubara(100.0, 2048)
    t0 = time.perf_counter()
    ro = o.lanczos(nitermax=128, neigen=1, tol=1e-12)
    psi = o.ground_state(ro["iters"], ro["vecp"][0])
    t_gs = time.perf_counter() - t0
    op = oracle.OracleSector(model, nu + 1, nd, nthreads=cores)
    oh = oracle.OracleSector(model, nu - 1, nd, nthreads=cores)
    t0 = time.perf_counter()
    g = []
    for c in range(model.nc):
        for dst, iph, pm in ((op, 1, +1), (oh, 2, -1)):
            phi = oracle.apply_c(o, dst, c, 0, pm, psi)
            g.append(dst.dynamic(iph, ro["evals"][0], phi, 36, wn)[0])
    t_green = time.perf_counter() - t0
    for x in (o, op, oh):
        x.close()
    return {"value": niter / dt, "unit": "Lanczos iterations/s", "cores": cores, "kind": "port",
            "sample": f"median of {reps} runs of {niter} FP64 Lanczos iterations (lanc_opt + HAIMupdo_multr "
                      f"restated in C/OpenMP, oracle/ed_oracle.c) in the ({nu},{nd}) sector, dim {o.dim}, after 5 "
                      f"warm-up H.v: {dt:.3f} s per run",
            "cpu": info, "repeats_s": times,
            "gs_plus_green": {"t_gs_ms": t_gs * 1e3, "t_green_ms": t_green * 1e3,
                              "t_total_ms": (t_gs + t_green) * 1e3, "gs_iters": ro["iters"],
                              "abs_dE0_vs_gpu": abs(float(ro["evals"][0]) - e0_gpu),
                              "G_iw0": [float(sum(x[0] for x in g).real / model.nc),
                                        float(sum(x[0] for x in g).imag / model.nc)]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--niter", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-green", action="store_true")
    ap.add_argument("--no-scan", action="store_true", help="skip the whole-Fock-space sector scan (block Lanczos)")
    ap.add_argument("--no-grad", action="store_true", help="skip the Ns=14 gradient workload")
    ap.add_argument("--schedule", type=int, default=0,
                    help="force this H.v schedule instead of the autotuner's choice (counter runs, where the "
                         "profiler's serialisation would make the tuner's timings meaningless)")
    ap.add_argument("--only", choices=["ns14"], default=None,
                    help="ns14: only the Ns=14 (7,7) H.v and Lanczos step in FP64 and FP32 (for profiler runs), "
                         "one JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs one sector (or one batch of sectors) per GPU; "
              "multi-GPU sharding is designed but not built, run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import ed_solver_amd as ed
    from ed_solver_amd._lib import check

    check(ed.lib().ed_init(0))
    if args.only == "ns14":
        # Ns=14 (7,7), dim 11,778,624: the fused H.v (256 launches between an event pair) and
        # fixed-length two-kernel Lanczos runs of 32 steps, FP64 and FP32, under the tuned
        # H.v schedule (or --schedule)
        mE, secE = ed.synthetic("E")
        sE = ed.Sector(mE, *secE)
        schedE, msE = sE.tune()
        if args.schedule:
            sE.set_schedule(args.schedule)
            schedE = args.schedule
        out = {"workload": f"Ns=14: 2-orbital Kanamori + 12 bath sites, sector (7,7), dim {sE.dim}",
               "schedule": schedE, "schedule_ms": msE}
        for prec in ("f64", "f32"):
            ms, by = sE.hmult_benchmark(prec, reps=max(1, args.steps) * 8)
            rms, rby, _ = sE.lanczos_benchmark(prec, niter=32, reps=max(1, args.steps))
            out[prec] = {"hv_us": ms * 1e3, "hv_GBps": by / (ms * 1e-3) / 1e9, "step_us": rms / 32 * 1e3,
                         "step_GBps": rby / (rms * 1e-3) / 1e9}
            log(f"Ns=14 {prec}: H.v {ms * 1e3:.1f} us, Lanczos step {rms / 32 * 1e3:.1f} us")
        sE.close()
        print(json.dumps(out), flush=True)
        return
    model, (nu, nd) = ed.synthetic("C")
    t0 = time.perf_counter()
    sec = ed.Sector(model, nu, nd)
    t_build = time.perf_counter() - t0
    log(f"sector (6,6) dim {sec.dim} built in {t_build*1e3:.1f} ms")
    D = sec.dim
    sched, sched_ms = sec.tune()
    if args.schedule:
        sec.set_schedule(args.schedule)
        sched = args.schedule
    log(f"H.v schedule {sched} (ms for schedules 1/2/4 row kernel, 8 fast, 16 fast XCD-grouped, 18/19/20 fast XCD-grouped with 2/3/4 rows per thread: {sched_ms})")

    def rate(prec):
        for _ in range(args.warmup):
            sec.lanczos_run(args.niter, precision=prec)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sec.lanczos_run(args.niter, precision=prec)
        dt = time.perf_counter() - t0
        return args.steps * args.niter / dt, dt

    it64, dt64 = rate("f64")
    log(f"FP64 Lanczos: {it64:.0f} iterations/s")
    it32, dt32 = rate("f32")
    log(f"FP32-H.v Lanczos: {it32:.0f} iterations/s")

    # the persistent kernel's two forms on a sector whose threads own one row (workload B,
    # (5,5), d_up = 252): one grid barrier per step (the default there) against two
    mB, secB = ed.synthetic("B")
    sB = ed.Sector(mB, *secB)
    sB.tune()
    forms_b = {"workload": f"Ns=10 (5,5), dim {sB.dim}, {args.niter}-step FP64 runs"}
    for form in ("one-barrier", "two-barrier"):
        sB.set_persistent(True, form)
        ms_b, _, pk_b = sB.lanczos_benchmark("f64", args.niter, 5)
        forms_b[form] = {"us_per_step": ms_b / args.niter * 1e3, "persistent": pk_b}
    forms_b["speedup"] = forms_b["two-barrier"]["us_per_step"] / forms_b["one-barrier"]["us_per_step"]
    log(f"workload B persistent Lanczos: one barrier {forms_b['one-barrier']['us_per_step']:.2f} us/step, "
        f"two barriers {forms_b['two-barrier']['us_per_step']:.2f} us/step")
    sB.close()

    def gs_timed(prec, tol, reps=5):
        """Ground state with the reference convergence test: one warm-up, then the median
        over `reps` runs of the eigenvalue solve and of solve + psi (copied to the host)."""
        sec.lanczos(nitermax=128, neigen=1, tolerance=tol, precision=prec)
        te, tg = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            r = sec.lanczos(nitermax=128, neigen=1, tolerance=tol, precision=prec)
            te.append(time.perf_counter() - t0)
            p = sec.ground_state(r)
            tg.append(time.perf_counter() - t0)
        return r, p, sorted(te)[reps // 2], sorted(tg)[reps // 2]

    # ground state to convergence (reference convergence test, tolerance 1e-12); the
    # eigenvalue solve and the ground-state vector (copied to the host) timed apart
    r64, psi, t_e64, t_gs64 = gs_timed("f64", 1e-12)
    r32, psi32, t_e32, t_gs32 = gs_timed("f32", 1e-12)
    e_mixed = float(r32.evals[0])
    de0 = abs(e_mixed - r64.evals[0])
    de0_rq = abs(sec.rayleigh(psi32) - r64.evals[0])
    log(f"GS FP64 E0={r64.evals[0]:.12f} ({r64.iters} it, E0 {t_e64*1e3:.1f} ms, +psi {t_gs64*1e3:.1f} ms); "
        f"mixed E0={e_mixed:.12f} ({r32.iters_f32} FP32 + {r32.iters} FP64 it, E0 {t_e32*1e3:.1f} ms, "
        f"+psi {t_gs32*1e3:.1f} ms) |dE0|={de0:.2e}")

    green = None
    if not args.no_green:
        wn = ed.matsubara(100.0, 2048)
        E0 = r64.evals[0]
        secs, phis, iphs = [], [], []
        s_p = ed.Sector(model, nu + 1, nd)
        s_h = ed.Sector(model, nu - 1, nd)
        for site in range(model.nc):
            secs.append(s_p); phis.append(ed.apply_C(sec, s_p, site, 0, +1, psi)); iphs.append(1)
            secs.append(s_h); phis.append(ed.apply_C(sec, s_h, site, 0, -1, psi)); iphs.append(2)
        ens = [E0] * len(secs)
        jobs = [(secs[k], iphs[k], [(k // 2, 1.0)]) for k in range(len(secs))]
        ed.compute_dynamic_batched(secs, phis, iphs, ens, 36, wn)   # warm-up
        ed.dynamic_from_state(sec, psi, E0, jobs, 0, 36, wn)
        reps = 3
        # one by one from host excitation vectors (the reference's call pattern)
        t0 = time.perf_counter()
        for _ in range(reps):
            for k in range(len(secs)):
                ed.compute_dynamic(secs[k], iphs[k], E0, phis[k], 36, wn)
        t_plain = (time.perf_counter() - t0) / reps
        # batched from host excitation vectors
        t0 = time.perf_counter()
        for _ in range(reps):
            dyn_h, _, _ = ed.compute_dynamic_batched(secs, phis, iphs, ens, 36, wn)
        t_host = (time.perf_counter() - t0) / reps
        # batched from the ground state (psi copied once, c / c+ on the device): the solver's path
        t0 = time.perf_counter()
        for _ in range(reps):
            dyn_b = ed.dynamic_from_state(sec, psi, E0, jobs, 0, 36, wn)
        t_batch = (time.perf_counter() - t0) / reps
        # the batch kernel of that last call: one launch for all runs, HIP event pair
        n_launch, _, lastk = ed.batched_stats(last=True)
        batch_kernel = None
        if n_launch and lastk["job_steps"]:
            ach_b = lastk["bytes"] / (lastk["ms"] * 1e-3) / 1e9
            batch_kernel = {"kernel": "k_lz_persist_runs<double, 12, 4> (all excitation runs in one cooperative "
                                      "launch)", "avg_us": lastk["ms"] * 1e3, "job_steps": lastk["job_steps"],
                            "us_per_job_step": lastk["ms"] * 1e3 / lastk["job_steps"],
                            "bytes_per_launch": lastk["bytes"], "achieved": ach_b, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": ach_b / HBM_PEAK_GBS,
                            "traffic": RUNS_TRAFFIC_F64_JOB_STEP * lastk["job_steps"]}
            log(f"batch kernel: {lastk['ms']*1e3:.0f} us for {lastk['job_steps']:.0f} job-steps "
                f"({lastk['ms']*1e3/lastk['job_steps']:.2f} us per job-step) -> {ach_b:.0f} GB/s")
        # the same runs with FP32 vectors and H.v (FP64 scalars and continued fraction): rate,
        # and the relative error at each w_n against the FP64 device result (which the GPU
        # tests hold to 1e-10 of the oracle)
        ed.dynamic_from_state(sec, psi, E0, jobs, 0, 36, wn, precision="f32")
        t0 = time.perf_counter()
        for _ in range(reps):
            dyn_32 = ed.dynamic_from_state(sec, psi, E0, jobs, 0, 36, wn, precision="f32")
        t_32 = (time.perf_counter() - t0) / reps
        per_w = np.max(np.abs(dyn_32 - dyn_b) / np.abs(dyn_b), axis=0)
        green_f32 = {"batched_solves_per_s": len(secs) / t_32, "vs_fp64": t_batch / t_32,
                     "per_w_rel_err_max": float(per_w.max()), "per_w_rel_err_w1": float(per_w[0]),
                     "per_w_rel_err_median": float(np.median(per_w)), "per_w_rel_err_last": float(per_w[-1])}
        log(f"G(iw) FP32: {len(secs)/t_32:.1f} solves/s; per-w relative error max {per_w.max():.2e} "
            f"(w_1 {per_w[0]:.2e}, median {np.median(per_w):.2e})")
        g = dyn_b.sum(axis=0) / model.nc
        green = {"solves": len(secs), "nitergreen": 36, "n_matsubara": 2048, "beta": 100.0,
                 "excited_dim": s_p.dim, "plain_solves_per_s": len(secs) / t_plain,
                 "batched_solves_per_s": len(secs) / t_batch, "batched_vs_plain": t_plain / t_batch,
                 "batched_host_vectors_solves_per_s": len(secs) / t_host,
                 "max_abs_diff_device_vs_host_vectors": float(np.max(np.abs(dyn_b - dyn_h))),
                 "G_iw0": [float(g[0].real), float(g[0].imag)], "batch_kernel": batch_kernel,
                 "fp32": green_f32}
        log(f"G(iw): plain {len(secs)/t_plain:.1f} solves/s, batched from host vectors {len(secs)/t_host:.1f}, "
            f"batched from the state on the device {len(secs)/t_batch:.1f} solves/s")
        s_p.close(); s_h.close()

    # mixed precision at the accuracy the target is stated for (|dE0| < 1e-8): both
    # ground-state solves with the convergence tolerance 1e-9 (GS vector included)
    r64b, _, t_e64b, t_gs64b = gs_timed("f64", 1e-9)
    r32b, _, t_e32b, t_gs32b = gs_timed("f32", 1e-9)
    rrqb, _, t_erqb, t_gsrqb = gs_timed("mixed_rq", 1e-9)
    rq_acc, rq_r2, rq_gap = sec.rq_info()
    mixed_1e8 = {"tolerance": 1e-9, "gs_ms_fp64": t_gs64b * 1e3, "gs_ms_mixed": t_gs32b * 1e3,
                 "gs_speedup": t_gs64b / t_gs32b, "e0_ms_fp64": t_e64b * 1e3, "e0_ms_mixed": t_e32b * 1e3,
                 "e0_speedup": t_e64b / t_e32b, "abs_dE0_fp64": abs(float(r64b.evals[0]) - float(r64.evals[0])),
                 "abs_dE0_mixed": abs(float(r32b.evals[0]) - float(r64.evals[0])), "iters_fp64": r64b.iters,
                 "iters_mixed_fp32_phase": r32b.iters_f32, "iters_mixed_fp64_phase": r32b.iters,
                 # the Rayleigh form: FP32 Lanczos, then E0 from one FP64 H.v of its Ritz vector
                 "rayleigh_form": {"e0_ms": t_erqb * 1e3, "e0_speedup": t_e64b / t_erqb, "gs_ms": t_gsrqb * 1e3,
                                   "gs_speedup": t_gs64b / t_gsrqb,
                                   "abs_dE0": abs(float(rrqb.evals[0]) - float(r64.evals[0])),
                                   "accepted": rq_acc, "res2": rq_r2, "ritz_gap": rq_gap,
                                   "iters_fp32": rrqb.iters_f32}}
    log(f"GS at tolerance 1e-9: FP64 E0 {t_e64b*1e3:.1f} ms (+psi {t_gs64b*1e3:.1f}), mixed E0 {t_e32b*1e3:.1f} ms "
        f"(+psi {t_gs32b*1e3:.1f}) (|dE0| {mixed_1e8['abs_dE0_mixed']:.1e}); Rayleigh form E0 {t_erqb*1e3:.1f} ms "
        f"({t_e64b / t_erqb:.2f}x, accepted {rq_acc}, |dE0| {mixed_1e8['rayleigh_form']['abs_dE0']:.1e})")

    # ARPACK_diago at (6,6): the implicitly restarted Lanczos on the device (basis, H.v,
    # reorthogonalisation, restart rotation in HBM) against SciPy's ARPACK driving the
    # device H.v through host vectors (two PCIe copies per H.v, as Hmultr); 4 lowest, ncv 24
    from ed_solver_amd.arpack import arpack_diago
    arp = {}
    for backend in ("device", "scipy"):
        arpack_diago(sec, 5, tolerance=1e-12, ncv=24, which="SA", backend=backend)   # warm-up
        t0 = time.perf_counter()
        ra = arpack_diago(sec, 5, tolerance=1e-12, ncv=24, which="SA", backend=backend)
        arp[backend] = {"ms": (time.perf_counter() - t0) * 1e3, "matvecs": ra.matvecs, "nconv": ra.nconv,
                        "E": [float(x) for x in ra.evals], "max_rel_resid": float(ra.rel_resid.max())}
    arp["device_vs_scipy"] = arp["scipy"]["ms"] / arp["device"]["ms"]
    arp["abs_dE0_vs_lanczos"] = abs(arp["device"]["E"][0] - float(r64.evals[0]))
    log(f"ARPACK_diago (4 lowest, ncv 24): device {arp['device']['ms']:.1f} ms ({arp['device']['matvecs']} H.v), "
        f"SciPy over host vectors {arp['scipy']['ms']:.1f} ms ({arp['scipy']['matvecs']} H.v)")

    scan = None
    if not args.no_scan:
        # every sector of the Ns=12 model with Boltzmann weight > 1e-6 at beta = 100, 8 states
        # each by block Lanczos of width 8, then G(i w_n) of all of them in one batched run
        from ed_solver_amd.solver import compute_green, sector_scan
        t0 = time.perf_counter()
        states, info = sector_scan(model, beta=100.0, weight_cutoff=1e-6, block_size=8, neigen=8)
        t_scan = time.perf_counter() - t0
        t0 = time.perf_counter()
        G = compute_green(model, states, 100.0, ed.matsubara(100.0, 2048), nitergreen=36)
        t_g = time.perf_counter() - t0
        scan = {"sectors": info["sectors"], "kept_sectors": [list(k) for k in info["kept_sectors"]],
                "states": info["states"], "E0": info["E0"], "t_lowest_all_sectors_s": info["t_pass1_s"],
                "t_lowest_parts_s": {k: round(v, 4) for k, v in info["t_pass1_parts"].items()},
                "t_block_lanczos_s": info["t_pass2_s"], "t_total_s": t_scan, "t_green_s": t_g,
                "block_steps": info.get("block_steps"), "block_full_passes": info.get("block_full_passes"),
                "block_ms": info.get("block_ms"), "block_ritz_ms": info.get("block_ritz_ms"),
                "green_runs_one_call": True, "G_iw0": [float(G[0, 0].real), float(G[0, 0].imag)]}
        log(f"sector scan: {info['sectors']} sectors in {info['t_pass1_s']:.2f} s, block Lanczos on "
            f"{len(info['kept_sectors'])} sectors ({info['states']} states) in {info['t_pass2_s']:.2f} s, "
            f"G in {t_g:.2f} s")

    grad = None
    if not args.no_grad:
        # Ns=14: 2 orbitals + 12 bath sites, sector (7,7), dim 11,778,624: ground state and
        # dE0/d(eps_k, V_k, U) by Hellmann-Feynman; bath-distance gradient on 4096 w_n
        from ed_solver_amd.grad import bath_distance, hf_gradient
        mE, secE = ed.synthetic("E")
        t0 = time.perf_counter()
        sE = ed.Sector(mE, *secE)
        t_b = time.perf_counter() - t0
        t0 = time.perf_counter()
        rE = sE.lanczos(nitermax=300, neigen=1, tolerance=1e-12)
        psiE = sE.ground_state(rE)
        t_gsE = time.perf_counter() - t0
        t0 = time.perf_counter()
        gE = hf_gradient(sE, psiE)
        t_hf = time.perf_counter() - t0
        dimE = sE.dim
        # mixed precision where the step is bandwidth-bound (d_up = 3432: the two-kernel
        # step, 11.8M-element vectors): E0 and psi at tolerance 1e-9, FP64 against mixed,
        # |dE0| against the tolerance-1e-12 E0 above
        mixE = {}
        for prec in ("f64", "f32", "mixed_rq"):
            sE.lanczos(nitermax=300, neigen=1, tolerance=1e-9, precision=prec)   # warm-up
            t0 = time.perf_counter()
            rr = sE.lanczos(nitermax=300, neigen=1, tolerance=1e-9, precision=prec)
            t_e = time.perf_counter() - t0
            sE.ground_state(rr)
            mixE[prec] = (t_e, time.perf_counter() - t0, rr)
        mixed_e = {"tolerance": 1e-9, "e0_ms_fp64": mixE["f64"][0] * 1e3, "e0_ms_mixed": mixE["f32"][0] * 1e3,
                   "e0_speedup": mixE["f64"][0] / mixE["f32"][0], "gs_ms_fp64": mixE["f64"][1] * 1e3,
                   "gs_ms_mixed": mixE["f32"][1] * 1e3, "gs_speedup": mixE["f64"][1] / mixE["f32"][1],
                   "abs_dE0_fp64": abs(float(mixE["f64"][2].evals[0]) - float(rE.evals[0])),
                   "abs_dE0_mixed": abs(float(mixE["f32"][2].evals[0]) - float(rE.evals[0])),
                   "iters_fp64": mixE["f64"][2].iters, "iters_mixed_fp32_phase": mixE["f32"][2].iters_f32,
                   "iters_mixed_fp64_phase": mixE["f32"][2].iters}
        acc, r2, gap = sE.rq_info()
        mixed_e["rayleigh_form"] = {"e0_ms": mixE["mixed_rq"][0] * 1e3,
                                    "e0_speedup": mixE["f64"][0] / mixE["mixed_rq"][0],
                                    "gs_ms": mixE["mixed_rq"][1] * 1e3,
                                    "abs_dE0": abs(float(mixE["mixed_rq"][2].evals[0]) - float(rE.evals[0])),
                                    "accepted": acc, "res2": r2, "ritz_gap": gap,
                                    "iters_fp32": mixE["mixed_rq"][2].iters_f32}
        log(f"Ns=14 mixed at tolerance 1e-9: E0 FP64 {mixed_e['e0_ms_fp64']:.1f} ms ({mixed_e['iters_fp64']} it), "
            f"mixed {mixed_e['e0_ms_mixed']:.1f} ms ({mixed_e['iters_mixed_fp32_phase']} FP32 + "
            f"{mixed_e['iters_mixed_fp64_phase']} FP64 it) -> {mixed_e['e0_speedup']:.2f}x, "
            f"|dE0| {mixed_e['abs_dE0_mixed']:.1e}")
        sE.close()
        import numpy as np
        nbE, ncE = mE.nb, mE.nc
        wn4 = ed.matsubara(100.0, 4096)
        epsE = np.diag(mE.eb[0]).copy()
        vE = mE.vbc[0].copy()
        tgt = np.zeros((4096, 2, ncE, ncE), complex)
        for k in range(nbE):   # target: the bath with every level shifted by 0.1
            c = k % ncE
            tgt[:, :, c, c] += (vE[k, c] ** 2 / (1j * wn4 - epsE[k] - 0.1))[:, None]
        bath_distance(epsE, vE, wn4, tgt)   # warm-up
        t0 = time.perf_counter()
        dist, ge, gv = bath_distance(epsE, vE, wn4, tgt)
        t_bd = time.perf_counter() - t0
        # the bath fit from that start (the quasi-Newton loop of the reference with the
        # analytic gradient): 24 parameters, 4096 w_n
        from ed_solver_amd.bathfit import fit_bath
        t0 = time.perf_counter()
        _, _, fit = fit_bath(epsE, vE, wn4, tgt, dist_max=1e-14, maxfn=2000)
        t_fit = time.perf_counter() - t0
        grad = {"workload": "Ns=14: 2-orbital Kanamori (U=3, J=0.5) + 12 bath sites, sector (7,7), "
                            f"dim {dimE}", "t_build_s": t_b, "t_gs_s": t_gsE, "gs_iters": rE.iters,
                "E0": float(rE.evals[0]), "t_hf_gradient_s": t_hf, "n_params": 2 * nbE + 1,
                "dE0_dU": gE.u, "bath_distance": dist, "t_bath_distance_and_gradient_ms": t_bd * 1e3,
                "n_matsubara": 4096, "bath_fit": {"t_s": t_fit, "device_calls": fit.nfev, "iterations": fit.niter,
                                                  "distance_start": dist, "distance_end": fit.f},
                "mixed_precision_at_1e-8": mixed_e}
        log(f"Ns=14 (7,7): GS {t_gsE:.2f} s ({rE.iters} it), HF gradient {t_hf*1e3:.1f} ms, "
            f"bath distance + gradient (4096 w_n) {t_bd*1e3:.2f} ms; bath fit {dist:.2e} -> {fit.f:.2e} in "
            f"{t_fit*1e3:.0f} ms ({fit.nfev} device calls)")

    # roofline of the dominant kernel of the timed region: the fixed-length FP64 run of
    # args.niter steps (the persistent Lanczos kernel where it applies: one launch per run),
    # a HIP event pair around the run on the sector's stream; algorithmic bytes 40 D per
    # step (H.v: read u_j, write H u_j; recurrence: read u_{j-1}, u_j, write u_{j+1})
    run_ms, run_bytes, persistent = sec.lanczos_benchmark("f64", niter=args.niter, reps=5)
    ach = run_bytes / (run_ms * 1e-3) / 1e9
    # the fused H.v of the two-kernel step (ground-state runs with a convergence test),
    # 256 launches back to back between one event pair
    avg_ms, bytes_l = sec.hmult_benchmark("f64", reps=256)
    ach_hv = bytes_l / (avg_ms * 1e-3) / 1e9
    avg32, bytes32 = sec.hmult_benchmark("f32", reps=256)
    log(f"Lanczos run of {args.niter} steps: {run_ms:.3f} ms ({'persistent kernel' if persistent else 'two-kernel steps'}) "
        f"-> {ach:.0f} GB/s; fused H.v FP64 {avg_ms*1e3:.2f} us -> {ach_hv:.0f} GB/s; FP32 {avg32*1e3:.2f} us")
    rows = {8: 1, 16: 1, 18: 2, 19: 3, 20: 4}.get(sched)
    kname_hv = (f"k_hv_fast<double, 1, 8, {rows}> (fused Lanczos H.v, schedule {sched})" if rows
                else f"k_hv_rows<double, 1, ...> (fused Lanczos H.v, schedule {sched})")
    ps = sec.persistent_shape("f64")
    if persistent and ps["cols_per_wg"] > 1:
        kname = (f"k_lz_persist_cols<double, {ps['width']}, {ps['rows']}, false> (persistent Lanczos, "
                 f"{ps['cols_per_wg']} columns per workgroup, {ps['workgroups']} workgroups, {args.niter} steps "
                 "per launch)")
    elif persistent:
        kname = f"k_lz_persist<double, {ps['width']}, {ps['rows']}, false> (persistent Lanczos, {args.niter} steps per launch)"
    else:
        kname = f"{kname_hv} + k_lz_update<double>"
    traffic = (PERSIST_TRAFFIC_F64_STEP * args.niter
               if (persistent and ps["cols_per_wg"] == 4 and ps["width"] == 8 and ps["rows"] == 4) else None)
    roofline = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": traffic, "kernel": kname, "avg_us": run_ms * 1e3, "bytes_per_launch": run_bytes,
                "steps_per_launch": args.niter,
                "hv_fused": {"kernel": kname_hv, "achieved": ach_hv, "frac": ach_hv / HBM_PEAK_GBS,
                             "avg_us": avg_ms * 1e3, "bytes_per_launch": bytes_l,
                             "traffic": HV_TRAFFIC_F64.get(sched), "fp32_avg_us": avg32 * 1e3,
                             "fp32_achieved": bytes32 / (avg32 * 1e-3) / 1e9},
                "schedule": sched, "schedule_ms": sched_ms}

    cpu = None
    if not args.no_cpu:
        log("timing the CPU restatement on the host cores ...")
        cpu = cpu_baseline(model, nu, nd, float(r64.evals[0]))
        log(f"CPU ({cpu['cpu']['model']}, {cpu['cores']} threads): {cpu['value']:.2f} iterations/s; "
            f"GS + G {cpu['gs_plus_green']['t_total_ms']:.0f} ms")

    out = {
        "metric": METRIC, "value": it64, "unit": "Lanczos iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt64 / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "Ns=12: 3-orbital Kanamori (U=2.5, J=0.4) + 9 bath sites, sector (6,6), "
                               f"dim {D}", "iterations_per_step": args.niter, "start_vector": "deterministic"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed_precision": {"iters_per_s": it32, "vs_fp64": it32 / it64, "gs_ms_fp64": t_gs64 * 1e3,
                            "gs_ms_mixed": t_gs32 * 1e3, "gs_speedup": t_gs64 / t_gs32,
                            "e0_ms_fp64": t_e64 * 1e3, "e0_ms_mixed": t_e32 * 1e3, "e0_speedup": t_e64 / t_e32,
                            "E0_fp64": float(r64.evals[0]), "E0_mixed": e_mixed,
                            "abs_dE0": de0, "abs_dE0_rayleigh": de0_rq, "iters_fp64": r64.iters,
                            "iters_mixed_fp32_phase": r32.iters_f32, "iters_mixed_fp64_phase": r32.iters,
                            "steps_run_fp64": r64.steps_run, "steps_run_mixed": r32.steps_run},
        "mixed_precision_at_1e-8": mixed_1e8,
        "green": green,
        "speedup_vs_cpu": (it64 / cpu["value"]) if cpu else None,
        "sector_scan": scan,
        "arpack": arp,
        "persistent_forms_b": forms_b,
        "gradients_ns14": grad,
    }
    if cpu and green:
        # ground state (E0 and psi) + the 6 excitation runs with their continued fractions,
        # both timed: the device path (monitored persistent solve, psi, batched runs from the
        # state) against the CPU restatement of the same steps
        t_gpu = t_gs64 + green["solves"] / green["batched_solves_per_s"]
        t_cpu = cpu["gs_plus_green"]["t_total_ms"] * 1e-3
        out["gs_plus_green"] = {"t_gpu_ms": t_gpu * 1e3, "t_cpu_ms": t_cpu * 1e3, "speedup": t_cpu / t_gpu,
                                "cpu_threads": cpu["cores"]}
    sec.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
