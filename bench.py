"""ELBO+gradient evaluations per second of the sparse GP-LVM fit on one MI355X.

Workload: N=100,000 cells, G=2,000 genes, Q=10 latent dims, M=256 inducing
points, no covariates (P = M+1 = 257), synthetic counts generated as the
workload recipe states (about 92% zeros).  One step = one ELBO evaluation
(Kuf, Kmm factor, weighted Grams, batched Cholesky, likelihood) plus one
gradient evaluation (beta-tilde, inverses, delta/Omega/X/T/rho gradients) at
the same parameters, after two IRLS steps have set W, nu and phi; inputs are
resident in HBM when the timed region starts.

Prints one JSON line (the bench contract) with a roofline entry for the
dominant kernel (HIP-event timed on the library's stream) and a CPU baseline:
the FP64 oracle timed on the host on a bounded gene sample at the same N, M,
extrapolated linearly in the gene count.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6e12, "mixed": 2.5e15}   # MI355X dense FP64 / FP16 MFMA


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def gram_flops(N, J, P):
    return float(J) * N * P * (P + 1)


def gradcontract_flops(N, J, P):
    return 2.0 * J * N * P * P


def host_cpu():
    """Host CPU model, logical CPUs of the machine, CPUs this process may run
    on, and the OpenMP thread setting of the environment."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"model": model, "nproc": os.cpu_count(), "affinity": aff,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_threads():
    """Every CPU this process is allowed to use, capped by OMP_NUM_THREADS when
    the environment sets it (the host's CPU share for this GPU)."""
    h = host_cpu()
    n = h["affinity"]
    if h["omp_num_threads"]:
        try:
            n = min(n, int(h["omp_num_threads"]))
        except ValueError:
            pass
    return max(1, n)


def cpu_baseline(ds, genes=(2, 4, 8), threads=None, ard_only=True):
    """FP64 oracle: one ELBO+grad unit on all cells and the first g genes for
    each g in genes; the unit time is fitted as a + b*g by least squares and
    extrapolated to all genes."""
    from oracle.oracle import OracleState
    from gaspacho2_amd import lib
    threads = threads or cpu_threads()
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    times = []
    for g in genes:
        L = int(ds.Yp[g])
        o = OracleState(ds.N, g, ds.M, ds.Q, ds.nh, ds.Yp[: g + 1], ds.Yi[:L], ds.Yx[:L],
                        ds.Zcov, ds.ls, ds.X0, T0, ard_only=ard_only, nthreads=threads)
        o.init_fit()
        o.irls_step(0)
        o.ql_start()
        t0 = time.perf_counter()
        o.elbo(1)
        o.grad()
        times.append(time.perf_counter() - t0)
        o.close()
        log(f"cpu oracle: {g} genes, one ELBO+grad unit {times[-1]:.2f} s ({threads} threads)")
    gs, ts = np.asarray(genes, float), np.asarray(times)
    b, a = np.polyfit(gs, ts, 1)
    resid = ts - (a + b * gs)
    t_full = a + b * ds.J
    host = host_cpu()
    return {"value": 1.0 / t_full, "unit": "evals/s", "cores": threads, "kind": "port",
            "host": host,
            "fit": {"genes": list(genes), "seconds": [round(t, 3) for t in times],
                    "intercept_s": a, "per_gene_s": b,
                    "max_abs_residual_s": float(np.max(np.abs(resid))),
                    "seconds_per_unit_at_G": t_full},
            "sample": f"FP64 oracle (C restatement of the reference kernels, OpenMP, {threads} threads "
                      f"on {host['model']}), one ELBO+grad unit at N={ds.N}, M={ds.M}, Q={ds.Q} on "
                      f"{', '.join(str(g) for g in genes)} genes ({', '.join(f'{t:.1f}' for t in times)} s), "
                      f"least-squares line extrapolated to G={ds.J}: {t_full:.0f} s per unit"}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_mixed_N100K_G2K_M256.txt")


def measured_traffic(group: str, precision: str, shape):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (lines 'traffic <group> <bytes>');
    only for the workload that summary was measured on."""
    if precision != "mixed" or tuple(shape) != (100000, 2000, 10, 256) or not os.path.exists(TRAFFIC_FILE):
        return None, None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic" and parts[1] == group:
                return float(parts[2]), os.path.relpath(TRAFFIC_FILE, ROOT)
    return None, None


def copy_calibration(mib: int, reps: int = 5) -> None:
    """A plain device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE
    counter readings give the scale factor from counter units to HBM bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_mib": mib, "bytes_read": n * 4, "bytes_written": n * 4,
                      "ms": dt * 1e3, "GBps": 2 * n * 4 / dt / 1e9}), flush=True)


def main_batched(args):
    """The uneven per-gene GLM/Laplace workload: G genes, each on its own cell
    subset (sizes log-uniform in [nmin, nmax]) of N cells, p_g ~ U{2..16}
    covariate columns (intercept included) plus the M Knm columns.  A step is
    one IRLS step of every gene (persistent pass, batched Cholesky, update),
    replayed from a recorded HIP graph."""
    from gaspacho2_amd import lib, synth
    t0 = time.perf_counter()
    ds = synth.make_dataset(args.N, 1, args.Q, args.M, seed=0, pca=False)
    bd = synth.make_batched(ds, args.G, args.nmin, args.nmax, seed=0)
    npos = int(bd.off[-1])
    log(f"batched data: N={ds.N} G={args.G} positions={npos} ({time.perf_counter() - t0:.1f} s)")
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0, T0,
                      ard_only=args.kernel == "ard")
    g.init_fit()
    bf = lib.BatchedFit(g, bd.off, bd.cell, bd.ncov, bd.cov, bd.yoff, bd.ypos, bd.yval)
    del bd.cov
    t1 = time.perf_counter()
    bf.init()
    log(f"init (unweighted fits of every gene): {time.perf_counter() - t1:.3f} s")
    grid = None
    if not args.no_tune:
        grid = bf.autotune(verbose=args.verbose)
        log(f"autotuned: {grid} persistent workgroups per CU")
    for w in range(args.warmup):
        bf.run(1, 0.0, graph=True)
    t1 = time.perf_counter()
    its, obj = bf.run(args.steps, 0.0, graph=True)
    dt = time.perf_counter() - t1
    sizes = (bd.off[1:] - bd.off[:-1]).astype(np.float64)
    PA = bd.ncov.astype(np.float64) + ds.M + 1
    flops = float(np.sum(sizes * PA * (PA + 1)))          # symmetric Gram of [z | nu]
    ach = flops / (dt / its)
    out = {"metric": "batched uneven-subset IRLS steps/s (G genes, each on its own cells)",
           "value": its / dt, "unit": "steps/s", "n_gpus": 1, "steps": its, "warmup": args.warmup,
           "ms_per_step": dt / its * 1e3, "higher_is_better": True, "scaling": "strong",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seed 0)",
           "config": {"workload": f"uneven per-gene GLM/Laplace fits: N={ds.N} cells, G={args.G} genes, "
                                  f"subsets log-uniform [{args.nmin}, {args.nmax}], M={ds.M}, Q={ds.Q}, "
                                  f"p_g in 2..16", "positions": npos},
           "objective": obj, "tuning": {"workgroups_per_cu": grid},
           "roofline": {"kernel": "k_bf_pass (Gram part)", "bound": "mfma", "achieved": ach / 1e12,
                        "peak": PEAK["f64"] / 1e12, "unit": "TFLOP/s", "frac": ach / PEAK["f64"],
                        "traffic": None, "flops_per_step": flops}}
    bf.close()
    g.close()
    print(json.dumps(out), flush=True)


def _trace_summary(o, sec, kind):
    """Seconds spent in IRLS steps, in the quasi-likelihood starts and in
    L-BFGS steps, from a fit trace (each entry closes the interval since the
    previous one)."""
    prev = 0.0
    spent = {0: 0.0, 1: 0.0, 2: 0.0}
    counts = {0: 0, 1: 0, 2: 0}
    for t, k in zip(sec, kind):
        spent[int(k)] += t - prev
        counts[int(k)] += 1
        prev = t
    return {"irls_s": spent[0], "irls_steps": counts[0], "ql_start_s": spent[1],
            "lbfgs_s": spent[2], "lbfgs_steps": counts[2]}


def _ctx_options(g, args):
    """Apply the --set NAME=VALUE context options to a new GPContext."""
    for item in args.set:
        name, _, val = item.partition("=")
        g.set_option(name, int(val))


def _run_fit(args, ds, prec):
    """One gp_fit of ds in precision prec ("f64", "mixed", "f64w") from its
    PCA latent: wall time, final objective, trace and fitted latent means."""
    from gaspacho2_amd import lib
    pmap = {"f64": lib.PREC_FP64, "mixed": lib.PREC_MIXED, "f64w": lib.PREC_FP64_WHITENED}
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0, T0,
                      ard_only=args.kernel == "ard", precision=pmap[prec])
    _ctx_options(g, args)
    if prec == "mixed" and not args.no_tune:
        g.set_option("autotune", 1)
    log(f"{prec} fit: {args.fit_itr} outer iteration(s), L-BFGS cap {args.fit_cap or 'none'}")
    t1 = time.perf_counter()
    obj = g.fit(args.fit_itr, args.fit_cap, verbose=args.verbose)
    dt = time.perf_counter() - t1
    o, sec, kind = g.fit_trace()
    _, _, info = g.fit_trace_steps()
    r = {"seconds": dt, "final": obj, "trace": (o, sec, kind), "info": info, **_trace_summary(o, sec, kind)}
    r["X"] = g.get("X")[:ds.N * ds.Q].reshape(ds.Q, ds.N).T.copy()
    g.close()
    return r


def _lbfgs_phases(o, kind):
    """L-BFGS steps and last relative change of each outer iteration's phase
    from a fit trace's objectives and kinds (a phase starts at its
    quasi-likelihood start, kind 1)."""
    o, kind = np.asarray(o), np.asarray(kind)
    starts = list(np.flatnonzero(kind == 1)) + [len(kind)]
    out = []
    for a, b in zip(starts[:-1], starts[1:]):
        v = o[a:b][(kind[a:b] == 1) | (kind[a:b] == 2)]
        d = abs(v[-1] - v[-2]) / abs(v[-1]) if len(v) > 1 else float("nan")
        out.append({"steps": int(len(v) - 1), "last_rel_change": float(d), "converged": bool(d < 1e-8)})
    return out


def _phases(o, info):
    """Per outer iteration: IRLS and L-BFGS steps, and whether each phase
    ended on its convergence rule (1e-7 IRLS, 1e-8 L-BFGS relative change of
    the objective, ggplvm.cu:1183, 1292) rather than on its step limit."""
    out = []
    for itr in np.unique(info[:, 1]):
        row = {"itr": int(itr)}
        for kind, name, tol in ((0, "irls", 1e-7), (2, "lbfgs", 1e-8)):
            v = o[(info[:, 1] == itr) & (info[:, 0] == kind)]
            if kind == 2:     # the L-BFGS phase starts from the quasi-likelihood objective
                v = np.concatenate([o[(info[:, 1] == itr) & (info[:, 0] == 1)], v])
            d = abs(v[-1] - v[-2]) / abs(v[-1]) if len(v) > 1 else float("nan")
            row[name] = {"steps": int(len(v) - (kind == 2)), "last_rel_change": float(d),
                         "converged": bool(d < tol), "end_objective": float(v[-1])}
        out.append(row)
    return out


def main_fit(args):
    """End-to-end fit (gp_fit: init, IRLS phase, quasi-likelihood start,
    L-BFGS phase with its line search, per outer iteration) from the PCA
    initial latent, in FP64 and in mixed precision on the same data; reports
    wall times, the final objectives and when the mixed fit reaches the FP64
    fit's final objective (1e-6 relative)."""
    from gaspacho2_amd import lib, synth
    t0 = time.perf_counter()
    ds = synth.make_dataset(args.N, args.G, args.Q, args.M, seed=0)
    log(f"data: N={ds.N} G={ds.J} ({time.perf_counter() - t0:.1f} s)")
    res = {}
    precs = args.fit_precs.split(",")
    if args.fit_f64w and "f64w" not in precs:
        precs.append("f64w")
    tag = f"N{ds.N}_G{ds.J}_M{ds.M}_itr{args.fit_itr}_cap{args.fit_cap}"
    for prec in precs:
        path = os.path.join(args.fit_load, f"fit_{prec}_{tag}.npz") if args.fit_load else None
        if path and os.path.exists(path):
            # a fit of the same data and schedule saved by an earlier run (--fit-save)
            z = np.load(path)
            res[prec] = {"seconds": float(z["seconds"]), "final": float(z["final"]),
                         "trace": (z["obj"], z["sec"], z["kind"]), "info": z["info"], "X": z["X"],
                         **_trace_summary(z["obj"], z["sec"], z["kind"])}
            log(f"{prec} fit loaded from {path}")
        else:
            res[prec] = _run_fit(args, ds, prec)
            if args.fit_save:
                os.makedirs(args.fit_save, exist_ok=True)
                r = res[prec]
                np.savez(os.path.join(args.fit_save, f"fit_{prec}_{tag}.npz"), seconds=r["seconds"],
                         final=r["final"], obj=r["trace"][0], sec=r["trace"][1], kind=r["trace"][2],
                         info=r["info"], X=r["X"])
        r = res[prec]
        r["phases"] = _phases(r["trace"][0], r["info"])
        log(f"{prec} fit: {r['seconds']:.2f} s, final objective {r['final']:.10e}, "
            f"{r['lbfgs_steps']} L-BFGS steps ({r['lbfgs_s']:.2f} s), "
            f"{r['irls_steps']} IRLS steps ({r['irls_s']:.2f} s); phases {r['phases']}")
    if "f64" not in res or "mixed" not in res:
        log("the comparison needs the f64 and mixed fits (run them, or load them with --fit-load)")
        return
    f64, mx = res["f64"], res["mixed"]

    # latent posterior means after orthogonal Procrustes alignment onto the
    # FP64 fit's
    def xdiff(a):
        U, _, Vt = np.linalg.svd(a.T @ f64["X"])
        xa = a @ (U @ Vt)
        n = np.linalg.norm(f64["X"])
        return float(np.linalg.norm(a - f64["X"]) / n), float(np.linalg.norm(xa - f64["X"]) / n)
    x_rel_raw, x_rel = xdiff(mx["X"])
    log(f"latent means, mixed vs FP64 fit: {x_rel_raw:.2e} relative, {x_rel:.2e} after Procrustes")
    xw = None
    if "f64w" in res:
        xw = xdiff(res["f64w"]["X"])
        log(f"latent means, FP64 whitened-basis vs FP64 fit: {xw[0]:.2e} relative, {xw[1]:.2e} after "
            f"Procrustes; final objectives differ by "
            f"{(res['f64w']['final'] - f64['final']) / abs(f64['final']):.2e}")
    from gaspacho2_amd.warmstart import steps_to_reach
    n_hit, t_hit = steps_to_reach(mx["trace"], f64["final"])
    gap = (mx["final"] - f64["final"]) / abs(f64["final"])

    def phase_gaps(r):
        # relative gap of each outer iteration's last L-BFGS objective to the FP64 fit's
        return [(a["lbfgs"]["end_objective"] - b["lbfgs"]["end_objective"]) / abs(b["lbfgs"]["end_objective"])
                for a, b in zip(r["phases"], f64["phases"])]
    log(f"end-of-phase objective gaps to the FP64 fit: mixed {phase_gaps(mx)}"
        + (f", FP64 whitened {phase_gaps(res['f64w'])}" if "f64w" in res else ""))
    out = {"metric": "end-to-end fit seconds at N=100K,G=2K,M=256: mixed vs FP64 plain",
           "value": mx["seconds"], "unit": "s", "n_gpus": 1, "steps": mx["lbfgs_steps"], "warmup": 0,
           "ms_per_step": mx["seconds"] / max(mx["lbfgs_steps"], 1) * 1e3, "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None, "dtype": "fp16x3/f32 MFMA + f64",
           "data": "synthetic (seed 0), PCA initial latent",
           "config": {"workload": f"sparse GP-LVM fit, N={ds.N} cells, G={ds.J} genes, Q={ds.Q}, M={ds.M}, "
                                  f"{args.fit_itr} outer iteration(s), L-BFGS cap {args.fit_cap}"},
           "mixed": {k: v for k, v in mx.items() if k not in ("trace", "X", "info")},
           "fp64_plain": {k: v for k, v in f64.items() if k not in ("trace", "X", "info")},
           "final_rel_gap_mixed_vs_fp64": gap, "phase_end_gaps_mixed_vs_fp64": phase_gaps(mx),
           "latent_rel_diff_mixed_vs_fp64": {"raw": x_rel_raw, "procrustes": x_rel},
           "fp64_whitened_vs_fp64": (None if xw is None else
                                     {"latent_raw": xw[0], "latent_procrustes": xw[1],
                                      "final_rel_gap": (res["f64w"]["final"] - f64["final"]) / abs(f64["final"]),
                                      "seconds": res["f64w"]["seconds"], "phases": res["f64w"]["phases"],
                                      "phase_end_gaps": phase_gaps(res["f64w"])}),
           "mixed_reaches_fp64_final": {"seconds": t_hit, "lbfgs_steps": n_hit,
                                        "speedup_vs_fp64_fit": (f64["seconds"] / t_hit) if t_hit else None},
           "speedup_same_schedule": f64["seconds"] / mx["seconds"]}
    print(json.dumps(out), flush=True)


def main_fitcheck(args):
    """The mixed fit's schedule (gp_fit's, ggplvm.cu:1146-1330: IRLS phase to
    1e-7, quasi-likelihood start, L-BFGS with the reference line search) driven
    step by step through the C ABI; at every accepted L-BFGS iterate the ELBO
    and the gradient are evaluated again in FP64 at the same parameters and
    the relative errors of the mixed values are recorded.  The mixed state is
    restored after each check, so the trajectory is the mixed fit's own."""
    from gaspacho2_amd import lib, synth
    t0 = time.perf_counter()
    ds = synth.make_dataset(args.N, args.G, args.Q, args.M, seed=0)
    log(f"data: N={ds.N} G={ds.J} ({time.perf_counter() - t0:.1f} s)")
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0, T0,
                      ard_only=args.kernel == "ard", precision=lib.PREC_MIXED)
    _ctx_options(g, args)
    scratch = 49999                              # an objective slot the schedule never reaches
    lk = {}
    jtrmax = min(200, args.fit_cap) if args.fit_cap > 0 else 200   # first outer iteration
    g.init_fit()
    for jtr in range(jtrmax):
        lk[jtr] = g.irls_step(jtr)
        if jtr > 0 and (jtr == jtrmax - 1 or abs(lk[jtr - 1] - lk[jtr]) / abs(lk[jtr]) < 1e-7):
            break
    lk[0] = g.ql_start()
    rows = []

    def check(it, slot):
        g.grad()
        gm = g.get("grad").copy()
        em = lk[slot]
        g.set_precision(lib.PREC_FP64)
        e64 = g.elbo(scratch)
        g.grad()
        g64 = g.get("grad").copy()
        g.set_precision(lib.PREC_MIXED)
        g.elbo(slot)                             # the mixed state back (same bits as before)
        gerr = float(np.max(np.abs(gm - g64)) / np.max(np.abs(g64)))
        rows.append({"iter": it, "elbo_mixed": em, "elbo_fp64": e64,
                     "elbo_rel_err": abs(em - e64) / abs(e64), "grad_rel_err": gerr})
        log(f"iterate {it}: ELBO mixed {em:.10e} fp64 {e64:.10e} rel {abs(em - e64) / abs(e64):.2e}, "
            f"gradient rel (max entry) {gerr:.2e}")

    check(0, 0)
    t1 = time.perf_counter()
    for jtr in range(jtrmax):
        g.grad()
        g.lbfgs_step(jtr, 0)
        sr = 1.0
        for ktr in range(1, 8):
            lk[jtr + 1] = g.elbo(jtr + 1)
            v = lk[jtr + 1]
            if v > lk[jtr] or ktr == 7:
                break
            if not np.isfinite(v):
                sr *= 0.01
            elif v < lk[jtr]:
                sr *= 0.1
            g.downdate(jtr, sr, 0)
        check(jtr + 1, jtr + 1)
        if jtr > 0 and abs(lk[jtr + 1] - lk[jtr]) < abs(lk[jtr + 1]) * 1e-8:
            break
    dt = time.perf_counter() - t1
    g.close()
    ee = [r["elbo_rel_err"] for r in rows]
    ge = [r["grad_rel_err"] for r in rows]
    out = {"metric": "max relative ELBO error of the mixed fit vs FP64 at every L-BFGS iterate",
           "value": max(ee), "unit": "relative", "n_gpus": 1, "steps": len(rows) - 1, "warmup": 0,
           "ms_per_step": dt / max(len(rows) - 1, 1) * 1e3, "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None, "dtype": "fp16x3/f32 MFMA + f64 vs f64",
           "data": "synthetic (seed 0), PCA initial latent",
           "config": {"workload": f"mixed fit with FP64 checks, N={ds.N} cells, G={ds.J} genes, "
                                  f"Q={ds.Q}, M={ds.M}, L-BFGS cap {args.fit_cap}"},
           "elbo_rel_err_max": max(ee), "elbo_rel_err_median": float(np.median(ee)),
           "grad_rel_err_max": max(ge), "grad_rel_err_median": float(np.median(ge)),
           "iterates": rows}
    print(json.dumps(out), flush=True)


def _make_ds(args_tuple):
    from gaspacho2_amd import synth
    N, G, Q, M, seed, model_seed = args_tuple
    return seed, synth.make_dataset(N, G, Q, M, seed=seed, model_seed=model_seed)


def hip_mem_info():
    """(free, total) device bytes from the HIP runtime the library uses."""
    import ctypes as C
    rt = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    f, t = C.c_size_t(), C.c_size_t()
    rt.hipMemGetInfo(C.byref(f), C.byref(t))
    return f.value, t.value


def main_score(args):
    """Genetic-mapping score tests along the latent space: N cells from D
    donors (uniform assignment), G genes, Q latent dims, M inducing points;
    each gene has its own S cis SNPs with donor dosages 0/1/2 drawn under
    Hardy-Weinberg equilibrium (MAF ~ U(0.05, 0.5)); every gene x cis-SNP
    pair is tested with the interaction column dosage x X[:, 0] (the first
    latent coordinate), centred, as ggplvm_de.cu:989-994 builds it.  The
    fit state is the mixed-precision IRLS phase (K steps) at the latent the
    counts were drawn from; the statistics and p-values of a gene sample are
    checked against an FP64 context on those genes alone (genes are
    independent in the IRLS phase)."""
    from gaspacho2_amd import lib, synth
    t0 = time.perf_counter()
    N, G, Q, M = args.N, args.G, args.Q, args.M
    last = [time.perf_counter()]

    def progress(done, total):
        if time.perf_counter() - last[0] > 20 or done == total:
            log(f"data: {done} of {total} genes ({time.perf_counter() - t0:.0f} s)")
            last[0] = time.perf_counter()

    ds = synth.make_dataset(N, G, Q, M, seed=0, gen_M=64, pca=False, block=128, progress=progress)
    if ds.L >= 2 ** 31:
        raise SystemExit(f"{ds.L} non-zero counts do not fit the 32-bit SparseMat offsets")
    log(f"data: N={N} G={G} L={ds.L} zeros={ds.zero_fraction():.3f} ({time.perf_counter() - t0:.1f} s)")
    rng = np.random.default_rng(11)
    D, S = args.donors, args.snps_per_gene
    donor = rng.integers(0, D, N).astype(np.int32)
    maf = rng.uniform(0.05, 0.5, G * S)
    dos = rng.binomial(2, np.broadcast_to(maf, (D, G * S))).astype(np.uint8)   # D x (G S)
    gene = np.repeat(np.arange(G), S).astype(np.int32)
    snp = np.arange(G * S).astype(np.int32)
    factor = np.ascontiguousarray(np.asarray(ds.X0)[:, 0])
    T0 = lib.init_ta_with_per(M, Q)
    f0, tot = hip_mem_info()
    t1 = time.perf_counter()
    g = lib.GPContext(N, G, M, Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0, T0,
                      ard_only=True, precision=lib.PREC_MIXED)
    g.set_option("gram_splits", args.gram_splits)     # FP32 Gram partials: segments x genes x blocks
    _ctx_options(g, args)
    g.init_fit()
    irls = []
    for jtr in range(args.irls):
        t2 = time.perf_counter()
        irls.append(g.irls_step(jtr))
        log(f"mixed IRLS step {jtr}: {time.perf_counter() - t2:.2f} s, objective {irls[-1]:.10e}")
    f1, _ = hip_mem_info()
    log(f"context + {args.irls} IRLS steps: {time.perf_counter() - t1:.1f} s; device memory in use "
        f"{(f0 - f1) / 2**30:.1f} GiB of {tot / 2**30:.0f}")
    g.score_tests_donor(donor, dos, factor, gene[:64], snp[:64])      # warm-up
    t2 = time.perf_counter()
    stat, pval = g.score_tests_donor(donor, dos, factor, gene, snp)
    dt = time.perf_counter() - t2
    log(f"{len(gene)} interaction score tests: {dt:.2f} s ({len(gene) / dt:.0f} tests/s); "
        f"p < 0.05 in {np.mean(pval < 0.05):.3f} of them; finite {np.all(np.isfinite(stat))}")
    g.close()
    # FP64 on a gene sample (same cells, latent, inducing points; its own IRLS)
    pick = np.sort(rng.choice(G, args.check_genes, replace=False))
    sel = np.concatenate([np.arange(ds.Yp[j], ds.Yp[j + 1]) for j in pick])
    Yp = np.zeros(len(pick) + 1, np.int64)
    Yp[1:] = np.cumsum(np.diff(ds.Yp)[pick])
    h = lib.GPContext(N, len(pick), M, Q, ds.nh, Yp.astype(np.int32), ds.Yi[sel], ds.Yx[sel], ds.Zcov,
                      ds.ls, ds.X0, T0, ard_only=True, precision=lib.PREC_FP64)
    h.init_fit()
    for jtr in range(args.irls):
        h.irls_step(jtr)
    gl = np.repeat(np.arange(len(pick)), S).astype(np.int32)
    sl = (pick[:, None] * S + np.arange(S)[None, :]).ravel().astype(np.int32)
    s64, p64 = h.score_tests_donor(donor, dos, factor, gl, sl)
    h.close()
    sm = stat[sl]
    pm = pval[sl]
    ea = float(np.max(np.abs(sm - s64)))
    es = float(np.max(np.abs(sm - s64) / np.maximum(np.abs(s64), 1e-300)))
    ep = float(np.max(np.abs(np.log(pm) - np.log(p64))))
    log(f"FP64 check on {len(pick)} genes ({len(sl)} tests): statistic max abs err {ea:.2e} "
        f"(max rel {es:.2e}), p-value max |log ratio| {ep:.2e}")
    P = ds.P
    flops = 2.0 * len(gene) * N * P                      # b = Z^T D h per test
    out = {"metric": "gene x cis-SNP interaction score tests/s (genotype x latent coordinate)",
           "value": len(gene) / dt, "unit": "tests/s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
           "dtype": "f64 (statistics) on mixed-precision IRLS factors",
           "data": "synthetic (SE-ARD GP counts, HWE dosages per donor, seed 0/11)",
           "config": {"workload": f"score tests N={N} cells from {D} donors, G={G} genes x {S} cis SNPs, "
                                  f"Q={Q}, M={M}, P={P}, SE-ARD, {args.irls} mixed IRLS steps",
                      "N": N, "G": G, "Q": Q, "M": M, "donors": D, "tests": len(gene)},
           "irls_objectives": irls, "p_below_0.05": float(np.mean(pval < 0.05)),
           "fp64_check": {"genes": int(len(pick)), "tests": int(len(sl)), "stat_max_abs_err": ea,
                          "stat_max_rel_err": es, "pval_max_abs_log_ratio": ep},
           "device_GiB_in_use": (f0 - f1) / 2 ** 30, "gflops_b": flops / dt / 1e9}
    print(json.dumps(out), flush=True)


def main_warmstart(args):
    """Amortization-encoder warm start (gaspacho2_amd.warmstart).  Fits
    (mixed precision, PCA initial latent) on seeds 0..S-1 that share one set
    of gene functions give (counts, fitted X) pairs; an encoder is trained on
    them for each target in --enc-target (the first is the headline), and a
    fresh seed is fitted with the same schedule from the PCA latent and from
    each encoder's warm start.  Reports the objectives and the L-BFGS steps
    the warm start needs to reach the PCA-started fit's final objective."""
    from multiprocessing import get_context
    from gaspacho2_amd import lib, warmstart
    S = args.train_seeds
    targets = args.enc_target.split(",")
    t0 = time.perf_counter()
    with get_context("fork").Pool(min(S + 1, 8)) as pool:
        data = dict(pool.imap_unordered(_make_ds, [(args.N, args.G, args.Q, args.M, s, 1000)
                                                   for s in range(S + 1)]))
    log(f"data: {S + 1} seeds of N={args.N} G={args.G} ({time.perf_counter() - t0:.1f} s)")
    T0 = lib.init_ta_with_per(args.M, args.Q)
    train = []
    t_itr = args.fit_itr if args.train_itr is None else args.train_itr
    t_cap = args.fit_cap if args.train_cap is None else args.train_cap
    for s in range(S):
        X, obj, dt, _ = warmstart.fit_from(data[s], data[s].X0, T0, t_itr, t_cap,
                                              ard_only=args.kernel == "ard")
        train.append((data[s], X))
        log(f"seed {s}: PCA-started fit {dt:.1f} s, objective {obj:.10e}, "
            f"rms(X - X_pca) {np.sqrt(np.mean((X - data[s].X0) ** 2)):.3f}")
    ds = data[S]
    Xp, obj_p, dt_p, tr_p = warmstart.fit_from(ds, ds.X0, T0, args.fit_itr, args.fit_cap,
                                                ard_only=args.kernel == "ard")
    n_pca = int(np.sum(tr_p[2] == 2))
    first = lambda tr: float(tr[0][np.argmax(tr[2] == 1)])
    log(f"fresh seed, PCA start: {obj_p:.10e} after {n_pca} L-BFGS steps ({dt_p:.1f} s)")

    def procrustes(a, b):
        U, _, Vt = np.linalg.svd(a.T @ b)
        return float(np.linalg.norm(a @ (U @ Vt) - b) / np.linalg.norm(b))
    ref = None
    if args.fp64_ref:
        # the FP64 fit of the fresh seed from its PCA latent, same schedule: the
        # reference the mixed fits are measured against
        log("fresh seed, FP64 plain fit from PCA (every objective printed)")
        X64, obj64, dt64, tr64 = warmstart.fit_from(ds, ds.X0, T0, args.fit_itr, args.fit_cap,
                                                    precision=lib.PREC_FP64, ard_only=args.kernel == "ard",
                                                    verbose=True)
        n64, t64 = warmstart.steps_to_reach(tr_p, obj64)
        ref = {"final_objective": obj64, "fit_seconds": dt64, "lbfgs_steps": int(np.sum(tr64[2] == 2)),
               "lbfgs_phases": _lbfgs_phases(tr64[0], tr64[2]),
               "pca_mixed": {"final_rel_gap": (obj_p - obj64) / abs(obj64),
                             "latent_procrustes": procrustes(Xp, X64),
                             "reaches_fp64_final": {"lbfgs_steps": n64, "seconds": t64,
                                                    "speedup": (dt64 / t64) if t64 else None}}}
        log(f"fresh seed, FP64 plain from PCA: {obj64:.10e} after {ref['lbfgs_steps']} L-BFGS steps "
            f"({dt64:.1f} s); the PCA-started mixed fit reaches it after {n64} steps ({t64} s)")
    variants = {}
    for target in targets:
        t1 = time.perf_counter()
        enc, losses = warmstart.train_encoder(train, target, epochs=args.enc_epochs, log=log)
        t_train = time.perf_counter() - t1
        X0e = warmstart.warm_start(enc, ds, target)
        enc.close()
        log(f"{target}: encoder {S * args.N} cells, {args.enc_epochs} epochs, {t_train:.1f} s, "
            f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
        Xe, obj_e, dt_e, tr_e = warmstart.fit_from(ds, X0e, T0, args.fit_itr, args.fit_cap,
                                                   ard_only=args.kernel == "ard")
        U, _, Vt = np.linalg.svd(Xe.T @ Xp)
        x_proc = float(np.linalg.norm(Xe @ (U @ Vt) - Xp) / np.linalg.norm(Xp))
        n_hit, t_hit = warmstart.steps_to_reach(tr_e, obj_p)
        log(f"{target}: warm-started fit {obj_e:.10e} ({dt_e:.1f} s), reaches the PCA-started final after "
            f"{n_hit} of {n_pca} L-BFGS steps")
        vs64 = None
        if ref is not None:
            n6, t6 = warmstart.steps_to_reach(tr_e, ref["final_objective"])
            vs64 = {"final_rel_gap": (obj_e - ref["final_objective"]) / abs(ref["final_objective"]),
                    "latent_procrustes": procrustes(Xe, X64),
                    "reaches_fp64_final": {"lbfgs_steps": n6, "seconds": t6,
                                           "speedup": (ref["fit_seconds"] / t6) if t6 else None}}
            log(f"{target}: reaches the FP64 fit's final after {n6} L-BFGS steps ({t6} s): "
                f"{vs64['reaches_fp64_final']['speedup']}x less wall time than the FP64 fit")
        variants[target] = {
            "steps_saved": (n_pca - n_hit) if n_hit else None,
            "final_objective": obj_e, "rel_gap_vs_pca_start": (obj_e - obj_p) / abs(obj_p),
            "objective_after_ql_start": first(tr_e), "fit_seconds": dt_e,
            "lbfgs_steps": int(np.sum(tr_e[2] == 2)), "latent_procrustes_vs_pca_start": x_proc,
            "lbfgs_phases": _lbfgs_phases(tr_e[0], tr_e[2]),
            "reaches_pca_final": {"lbfgs_steps": n_hit, "seconds": t_hit},
            "encoder_train_seconds": t_train, "encoder_epoch_loss": [float(losses[0]), float(losses[-1])],
            "warm_start_rms_to_pca_fit": float(np.sqrt(np.mean((X0e - Xp) ** 2))),
            "vs_fp64_plain": vs64}
    head = variants[targets[0]]
    out = {"metric": "L-BFGS steps saved by the encoder warm start (fresh seed, mixed fit)",
           "value": head["steps_saved"], "unit": "steps", "n_gpus": 1,
           "steps": n_pca, "warmup": 0, "ms_per_step": head["fit_seconds"] / max(n_pca, 1) * 1e3,
           "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
           "dtype": "bf16/f32 MFMA encoder, fp16x3/f32 + f64 fit", "data": f"synthetic, seeds 0..{S}",
           "config": {"workload": f"encoder warm start: {S} training seeds x N={args.N}, G={args.G}, "
                                  f"Q={args.Q}, M={args.M}; MLP {args.G}-512-256-{2 * args.Q}; "
                                  f"{args.fit_itr} outer iteration(s), L-BFGS cap {args.fit_cap} "
                                  f"(training fits: {t_itr}, {t_cap})",
                      "target": targets[0]},
           "pca_start": {"final_objective": obj_p, "objective_after_ql_start": first(tr_p), "fit_seconds": dt_p,
                         "lbfgs_steps": n_pca, "lbfgs_phases": _lbfgs_phases(tr_p[0], tr_p[2]),
                         "pca_rms_to_pca_fit": float(np.sqrt(np.mean((ds.X0 - Xp) ** 2)))},
           "fp64_plain": ref, "variants": variants}
    print(json.dumps(out), flush=True)


def main_irls(args):
    """GLM-phase IRLS steps (getWForBeta, getPsi, cholPsi, getBetaPhi,
    getLkhdPois: ggplvm.cu:1156-1181) at the ELBO workload's size: the mixed
    step with the IRLS weight pass fused into the Gram kernel (the default),
    the same step on separate kernels, and the FP64 step, each from init_fit
    in its own context.  Reports per-step times, the Gram kernel's HIP-event
    time and each mixed form's objective error against FP64 at every step."""
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(args.N, args.G, args.Q, args.M, seed=0)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    P = ds.P
    nsteps = args.warmup + args.steps
    res = {}
    for name, prec, fused in (("f64", lib.PREC_FP64, 0), ("fused", lib.PREC_MIXED, 1),
                              ("separate", lib.PREC_MIXED, 0)):
        g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls,
                          ds.X0, T0, ard_only=args.kernel == "ard", precision=prec)
        g.set_option("irls_fused", fused)
        g.init_fit()
        obj, ms = [], []
        for jtr in range(nsteps):
            if jtr == args.warmup:
                g.timing(True)
                g.timing_reset()
            t1 = time.perf_counter()
            obj.append(g.irls_step(jtr))      # returns after the objective's read-back
            ms.append((time.perf_counter() - t1) * 1e3)
        g.timing(False)
        kern = {k: g.timing_get(k)[0] for k in ("gram", "gram_kernel", "chol")}
        g.close()
        res[name] = {"obj": obj, "ms": ms[args.warmup:], "kern": kern}
        log(f"[irls {name}] {np.median(ms[args.warmup:]):.2f} ms per step, Gram kernel "
            f"{kern['gram_kernel']:.2f} ms, objective {obj[-1]:.12e}")
    ref = np.array(res["f64"]["obj"])
    for k in ("fused", "separate"):
        res[k]["obj_rel_err"] = float(np.max(np.abs(np.array(res[k]["obj"]) - ref) / np.abs(ref)))
    f = res["fused"]
    ms_step = float(np.median(f["ms"]))
    kt = f["kern"]["gram_kernel"]
    fl = gram_flops(ds.N, ds.J, P)
    T = ds.M // 16 if int(ds.nh[-2]) == 1 and ds.M % 16 == 0 else (P + 15) // 16
    out = {"metric": f"mixed IRLS steps/s at N={ds.N},G={ds.J},M={ds.M}", "value": 1e3 / ms_step,
           "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms_step, "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
           "dtype": "f16x2-mfma/f32-acc/f64",
           "data": "synthetic (Poisson counts from the SE-ARD GP recipe, seed 0; PCA-initialised X)",
           "config": {"workload": f"GLM-phase IRLS steps, N={ds.N} cells, G={ds.J} genes, Q={ds.Q}, "
                                  f"M={ds.M}, P={P}, no covariates",
                      "N": ds.N, "G": ds.J, "Q": ds.Q, "M": ds.M, "P": P},
           "roofline": {"kernel": f"k_gram_irls<{T}> (fused IRLS weights + Gram)", "bound": "mfma",
                        "achieved": fl / (kt * 1e-3) / 1e12, "peak": PEAK["mixed"] / 1e12,
                        "unit": "TFLOP/s", "frac": fl / (kt * 1e-3) / PEAK["mixed"],
                        "traffic": None, "kernel_ms": kt, "flops_per_launch": fl},
           "forms": {k: {"ms_per_step": float(np.median(v["ms"])), "kernels_ms": v["kern"],
                         "objective_rel_err_vs_fp64": v.get("obj_rel_err")} for k, v in res.items()},
           "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["f64", "mixed"], default="mixed")
    ap.add_argument("--N", type=int, default=None, help="cells (100000; batched: 200000)")
    ap.add_argument("--G", type=int, default=None, help="genes (2000; batched: 20000)")
    ap.add_argument("--Q", type=int, default=10)
    ap.add_argument("--M", type=int, default=None, help="inducing points (256; batched: 128)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", choices=["ard", "per-ard"], default="ard",
                    help="ard: SE-ARD on every latent dimension (kernel_funcs.cu:221-255, the "
                         "kernel the workloads name); per-ard: periodic dimension 0 + SE-ARD "
                         "(kernel_funcs.cu:61-97, the reference driver's default)")
    ap.add_argument("--no-tune", action="store_true", help="skip the mixed-path autotuner")
    ap.add_argument("--only", action="store_true",
                    help="run only --precision (no FP64 comparison run)")
    ap.add_argument("--fit-itr", type=int, default=1, help="fit workload: outer iterations")
    ap.add_argument("--verbose", action="store_true", help="fit workload: print every objective")
    ap.add_argument("--fit-cap", type=int, default=100, help="fit workload: L-BFGS steps per outer iteration")
    ap.add_argument("--fit-precs", default="f64,mixed",
                    help="fit workload: precisions to fit (f64, mixed, f64w), comma separated")
    ap.add_argument("--fit-save", default="", metavar="DIR", help="fit workload: save each fit to DIR")
    ap.add_argument("--fit-load", default="", metavar="DIR",
                    help="fit workload: take fits of the same data and schedule saved in DIR")
    ap.add_argument("--fit-f64w", action="store_true",
                    help="fit workload: also fit in FP64 on the whitened basis (rounding-only divergence)")
    ap.add_argument("--train-seeds", type=int, default=8, help="warmstart workload: training fits")
    ap.add_argument("--train-itr", type=int, default=None,
                    help="warmstart workload: outer iterations of the training fits (default --fit-itr)")
    ap.add_argument("--train-cap", type=int, default=None,
                    help="warmstart workload: L-BFGS cap of the training fits (default --fit-cap)")
    ap.add_argument("--fp64-ref", action="store_true",
                    help="warmstart workload: also fit the fresh seed in FP64 and measure the mixed fits against it")
    ap.add_argument("--enc-epochs", type=int, default=30, help="warmstart workload: encoder epochs")
    ap.add_argument("--enc-target", default="residual,aligned",
                    help="warmstart workload: encoder targets, comma-separated (first = headline)")
    ap.add_argument("--workload", choices=["elbo", "batched", "fit", "fitcheck", "warmstart", "score",
                                           "irls"],
                    default="elbo",
                    help="elbo: the ELBO+gradient unit (default); batched: uneven per-gene GLM fits; "
                         "fitcheck: mixed-vs-FP64 ELBO and gradient errors at every L-BFGS iterate; "
                         "irls: GLM-phase IRLS steps, fused and separate weight pass")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="context option (gp_set_option) for the contexts of the elbo, fit, fitcheck and "
                         "score workloads (the score workload's mixed context); repeatable")
    ap.add_argument("--nmin", type=int, default=500)
    ap.add_argument("--donors", type=int, default=200, help="score workload: donors")
    ap.add_argument("--snps-per-gene", type=int, default=10, help="score workload: cis SNPs per gene")
    ap.add_argument("--irls", type=int, default=6, help="score workload: IRLS steps before the tests")
    ap.add_argument("--gram-splits", type=int, default=8, help="score workload: mixed Gram cell segments")
    ap.add_argument("--check-genes", type=int, default=16, help="score workload: genes checked in FP64")
    ap.add_argument("--nmax", type=int, default=50000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="time a plain device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "the gene-sharded multi-GPU design is described but not built.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    batched = args.workload == "batched"
    if args.workload == "score":
        args.N = args.N or 1000000
        args.G = args.G or 5000
        args.M = args.M or 512
        if args.Q == 10:
            args.Q = 20
        main_score(args)
        return
    args.N = args.N or (200000 if batched else 100000)
    args.G = args.G or (20000 if batched else 2000)
    args.M = args.M or (128 if batched else 256)
    if batched:
        main_batched(args)
        return
    if args.workload == "fit":
        main_fit(args)
        return
    if args.workload == "fitcheck":
        main_fitcheck(args)
        return
    if args.workload == "warmstart":
        main_warmstart(args)
        return
    if args.workload == "irls":
        main_irls(args)
        return

    from gaspacho2_amd import lib, synth
    t0 = time.perf_counter()
    ds = synth.make_dataset(args.N, args.G, args.Q, args.M, seed=0)
    log(f"data: N={ds.N} G={ds.J} L={ds.L} zeros={ds.zero_fraction():.3f} "
        f"({time.perf_counter() - t0:.1f} s)")
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls,
                      ds.X0, T0, ard_only=args.kernel == "ard", precision=lib.PREC_FP64)
    _ctx_options(g, args)
    g.init_fit()
    for jtr in range(2):
        g.irls_step(jtr)
    e0 = g.ql_start()
    log(f"setup done, ELBO at start {e0:.10e}")
    P = ds.P

    tuning = None

    def run(prec_name, steps, warmup):
        nonlocal tuning
        g.set_precision(lib.PREC_MIXED if prec_name == "mixed" else lib.PREC_FP64)
        if prec_name == "mixed" and not args.no_tune:
            tuning = g.autotune()
            log(f"[mixed] autotuned {tuning}")
        for w in range(warmup):
            t1 = time.perf_counter()
            g.elbo(1)
            g.grad()
            log(f"[{prec_name}] warmup {w}: {time.perf_counter() - t1:.3f} s")
        # timed: the unit launched eagerly (one host call per kernel group)
        t1 = time.perf_counter()
        e = None
        for s in range(steps):
            e = g.elbo(1)
            g.grad()
            log(f"[{prec_name}] step {s}: ELBO {e:.12e}")
        t_all = time.perf_counter() - t1
        grad = g.get("grad")
        # timed: the same unit replayed as a recorded HIP graph
        g.eval_graph(1)
        t1 = time.perf_counter()
        eg = None
        for s in range(steps):
            eg = g.eval_graph(1)
        t_graph = time.perf_counter() - t1
        log(f"[{prec_name}] graph replay: {t_graph / steps * 1e3:.1f} ms per unit, ELBO {eg:.12e}")
        # per-kernel-group device times (HIP events on the library stream), one extra unit
        g.timing(True)
        g.timing_reset()
        g.elbo(1)
        g.grad()
        g.timing(False)
        kern = {}
        for name in ("gram", "gradcontract", "chol", "kuf", "elbo", "grad", "gram_kernel",
                     "gradcontract_kernel"):
            ms, cnt = g.timing_get(name)
            kern[name] = ms
        return {"t": min(t_all, t_graph), "t_eager": t_all, "t_graph": t_graph, "elbo": e,
                "elbo_graph": eg, "kern": kern, "grad": grad}

    res = {}
    order = ["f64", "mixed"] if args.precision == "mixed" else ["f64"]
    if args.only:
        order = [args.precision]
    for pn in order:
        res[pn] = run(pn, args.steps, args.warmup)
    main_p = args.precision
    r = res[main_p]
    kern = r["kern"]
    peak = PEAK["mixed" if main_p == "mixed" else "f64"]
    # the main kernel of each group alone (its preparation and reduction
    # kernels are in the group times and in the unit time, not here)
    cands = {
        "gram": gram_flops(ds.N, ds.J, P) / (kern["gram_kernel"] * 1e-3),
        "gradcontract": gradcontract_flops(ds.N, ds.J, P) / (kern["gradcontract_kernel"] * 1e-3),
    }
    peaks = {"gram": peak, "gradcontract": peak}
    dom = max(("gram", "gradcontract"), key=lambda k: kern[k + "_kernel"])
    achieved = cands[dom]
    traffic, tsrc = measured_traffic(dom + "_kernel", main_p, (ds.N, ds.J, ds.Q, ds.M))
    # the intercept-peeled forms apply when the intercept is the only column
    # before the GP block and M fills whole 16- (Gram) or 32-column (contraction) tiles
    g0 = int(ds.nh[-2])
    peel_gram = main_p == "mixed" and g0 == 1 and ds.M % 16 == 0 and g.get_option("gram_peel") == 1
    peel_gc = (main_p == "mixed" and g0 == 1 and ds.M % 32 == 0 and ds.M <= 256 and
               g.get_option("gc_form") == 2 and g.get_option("gc_peel") == 1)
    if main_p == "mixed":
        gname = "k_gram_sym4" if g.get_option("gram_waves") == 4 else "k_gram_sym"
        gname += f"<{ds.M // 16 if peel_gram else (P + 15) // 16}>"
        if peel_gc:
            rg = 8 if g.get_option("gc_waves") == 4 else g.get_option("gc_wave_rows")
            cg = 4 if g.get_option("gc_waves") == 4 else 16 // rg
            pd = min(g.get_option("gc_prefetch"), 1 if rg * cg > 16 else 7)
            cname = f"k_gc_gp<{ds.M // 16},{rg},{cg},{pd}>"
        else:
            cname = {0: "k_gc_mixed", 1: "k_gc_rows", 2: "k_gc_rows2",
                     3: "k_gc_rows4"}[g.get_option("gc_form")] + f"<{(P + 15) // 16}>"
        kname = {"gram": gname, "gradcontract": cname}[dom]
    else:
        kname = {"gram": "k_gram_mfma_f64", "gradcontract": "k_gc_mfma_f64"}[dom] + f"<{(P + 15) // 16}>"
    roofline = {"kernel": f"{kname} ({dom})", "bound": "mfma", "achieved": achieved / 1e12,
                "peak": peaks[dom] / 1e12, "unit": "TFLOP/s", "frac": achieved / peaks[dom],
                "traffic": traffic, "traffic_source": tsrc,
                "flops_per_launch": gram_flops(ds.N, ds.J, P) if dom == "gram" else gradcontract_flops(ds.N, ds.J, P),
                "kernel_ms": kern[dom + "_kernel"],
                # HBM rate of the same kernel: measured bytes per launch over its time
                "hbm_gbs": (traffic / (kern[dom + "_kernel"] * 1e-3) / 1e9) if traffic else None,
                "hbm_frac": (traffic / (kern[dom + "_kernel"] * 1e-3) / 8e12) if traffic else None,
                "others": {k: {"achieved_tflops": v / 1e12, "frac": v / peaks[k], "avg_ms": kern[k + "_kernel"]}
                           for k, v in cands.items()}}
    ms_step = r["t"] / args.steps * 1e3
    out = {
        "metric": "ELBO+grad evals/s at N=100K,G=2K,M=256",
        "value": args.steps / r["t"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if main_p == "f64" else "f16x2-mfma/f32-acc/f64",
        "data": "synthetic (Poisson counts from the SE-ARD GP recipe, seed 0; PCA-initialised X)",
        "config": {"workload": f"sparse GP-LVM ELBO+gradient, N={ds.N} cells, G={ds.J} genes, "
                               f"Q={ds.Q}, M={ds.M}, P={P}, no covariates, "
                               + ("SE-ARD kernel on every latent dimension" if args.kernel == "ard"
                                  else "periodic dim 0 + SE-ARD kernel"),
                   "kernel": args.kernel,
                   "N": ds.N, "G": ds.J, "Q": ds.Q, "M": ds.M, "P": P,
                   "precision": main_p, "nnz": ds.L},
        "elbo": r["elbo"],
        "kernels_ms": kern,
        "roofline": roofline,
        "cpu_baseline": None,
        "tuning": tuning,
        "launch": {"eager_ms_per_step": r["t_eager"] / args.steps * 1e3,
                   "graph_ms_per_step": r["t_graph"] / args.steps * 1e3,
                   "value_from": "graph" if r["t_graph"] <= r["t_eager"] else "eager"},
    }
    if "mixed" in res and "f64" in res:
        f, m = res["f64"], res["mixed"]
        gd = m["grad"] - f["grad"]
        out["fp64_plain"] = {"value": args.steps / f["t"], "ms_per_step": f["t"] / args.steps * 1e3,
                             "elbo": f["elbo"], "kernels_ms": f["kern"]}
        out["mixed_vs_fp64"] = {"speedup": f["t"] / m["t"],
                                "elbo_rel_err": abs(m["elbo"] - f["elbo"]) / abs(f["elbo"]),
                                "grad_rel_err": float(np.max(np.abs(gd)) / np.max(np.abs(f["grad"])))}
        # the same unit with the contraction's S_j in FP16 alone (option gc_slo 0:
        # two MFMAs per tile instead of three), for reference: its time and its
        # gradient's distance from FP64, whole and in the inducing-point block
        g.set_precision(lib.PREC_MIXED)
        g.set_option("gc_slo", 0)
        g.elbo(1)
        g.grad()
        t1 = time.perf_counter()
        for s in range(args.steps):
            e0 = g.elbo(1)
            g.grad()
        t0s = time.perf_counter() - t1
        g0d = g.get("grad") - f["grad"]
        npar = len(ds.nh) - 1
        a = ds.N + npar + ds.N * ds.Q
        b = a + ds.M * ds.Q
        g.timing(True)
        g.timing_reset()
        g.elbo(1)
        g.grad()
        g.timing(False)
        out["option_gc_slo0"] = {
            "value": args.steps / t0s, "ms_per_step": t0s / args.steps * 1e3,
            "gradcontract_kernel_ms": g.timing_get("gradcontract_kernel")[0],
            "elbo_rel_err": abs(e0 - f["elbo"]) / abs(f["elbo"]),
            "grad_rel_err": float(np.max(np.abs(g0d)) / np.max(np.abs(f["grad"]))),
            "grad_T_block_rel_err": float(np.max(np.abs(g0d[a:b])) / np.max(np.abs(f["grad"][a:b]))),
            "grad_T_block_rel_err_default": float(np.max(np.abs(gd[a:b])) / np.max(np.abs(f["grad"][a:b])))}
        g.set_option("gc_slo", 1)
    g.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ds, ard_only=args.kernel == "ard")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
