"""Benchmark: 8-D double Gaussian peak on one MI355X, FP32-compensated vs plain FP64.

Workload (the headline configuration): f(x) = 1/2 [g(x; 0.33) + g(x; 0.67)],
g a normalised isotropic Gaussian with sigma = 0.05 on [0,1]^8, nbins = 1024,
n_strat = 8 (16,777,216 hypercubes), 1e8 evaluations per iteration, alpha = 0.5,
beta = 0.75.  A step is one VEGAS+ iteration (sampling pass over all hypercubes,
map refinement, reallocation); --warmup iterations adapt the map and are left
out of the average (skip), --steps iterations are timed.  Defaults 5 + 15 = the
20 iterations of the workload.

Prints one JSON line: value = evaluations/s of the FP32-compensated form over
the timed iterations, with the plain FP64 form, time to 1e-4 relative error,
errors against the closed form, the roofline of the sampling kernel and the
CPU oracle timed on a bounded sample.  Under "configs" the same line carries
the other configurations of the path, each at its own size: the 16-D
oscillatory integrand (osc16), the 6-D mixture forward and gradient (mix6)
and the 512-integral batch against sequential launches (batch512).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def _log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def _time_to_target(I_k, S2, iter_ms, warmup, exact, target=1e-4, S2chi=None):
    """Device time (ms, iterations from the first) and index of the first
    iteration at which the running inverse-variance mean of the iterations >=
    warmup is within `target` relative of the closed form with chi2/dof < 2
    (chi2 from the variances S2chi, S2 by default); (None, None) if none is."""
    S2chi = S2 if S2chi is None else S2chi
    cum = 0.0
    for k in range(len(I_k)):
        cum += iter_ms[k]
        if k < warmup:
            continue
        sl = slice(warmup, k + 1)
        w = 1.0 / S2[sl]
        mean = float((I_k[sl] * w).sum() / w.sum())
        K = k + 1 - warmup
        chi2 = float((((I_k[sl] - mean) ** 2) / S2chi[sl]).sum() / (K - 1)) if K > 1 else 0.0
        if abs(mean - exact) / abs(exact) <= target and chi2 < 2.0:
            return cum, k
    return None, None


def time_to_target_seeds(itg, n_eval_it, n_strat, nbins, warmup, iters, seeds):
    """Time to 1e-4 of both forms over several seeds (one path per seed and
    form; the single path of the headline run says little, as the two forms'
    estimates scatter independently and an iteration's sigma is about the
    target itself): for the seeds that reach it within `iters` iterations,
    the mean and median device time and the mean iteration index, and how
    many do -- with chi2/dof < 2 from the reference's per-iteration variances
    (the metric's definition) and from the unbiased ones."""
    import numpy as np
    from cuvegas_amd import Integrator
    out = {}
    for prec in ("fp64", "fp32"):
        got = {"reference": ([], []), "unbiased": ([], [])}
        prog = None
        for seed in seeds:
            st = Integrator(itg.source, itg.n_dim, n_eval_it, n_intervals=nbins, n_strat=n_strat,
                            max_it=iters, skip=warmup, seed=seed, precision=prec, program=prog)
            prog = st.program
            st.set_timing(True)
            st.run(iters, use_graph=False)
            I_k, S2, _ = st.history()
            S2u = st.history_unbiased()
            _, iter_ms = st.timing()
            st.close()
            for key, chi in (("reference", S2), ("unbiased", S2u)):
                t, k = _time_to_target(I_k, S2, iter_ms, warmup, itg.exact, S2chi=chi)
                if t is not None:
                    got[key][0].append(t)
                    got[key][1].append(k)
        out[prec] = {key: {"mean_ms": float(np.mean(t)) if t else None,
                           "median_ms": float(np.median(t)) if t else None,
                           "mean_iteration": float(np.mean(k)) if k else None,
                           "reached": len(t)}
                     for key, (t, k) in got.items()}
    a, b = out["fp32"]["reference"]["mean_ms"], out["fp64"]["reference"]["mean_ms"]
    out["fp64_over_fp32_mean"] = b / a if a and b else None
    out["iterations"] = iters
    out["seeds"] = [min(seeds), max(seeds)]
    out["n_seeds"] = len(seeds)
    return out


def run_form(precision, itg, n_eval_it, warmup, steps, n_strat, nbins, seed, use_timing=True, options=None,
             run_len=0):
    import numpy as np
    from cuvegas_amd import Integrator
    iters = warmup + steps
    t0 = time.perf_counter()
    st = Integrator(itg.source, itg.n_dim, n_eval_it, n_intervals=nbins, n_strat=n_strat,
                    max_it=iters, skip=warmup, seed=seed, precision=precision, options=options,
                    run_len=run_len)
    setup_s = time.perf_counter() - t0
    st.set_timing(use_timing)
    st.run(warmup, use_graph=False)
    t1 = time.perf_counter()
    st.run(steps, use_graph=False)
    wall_timed = time.perf_counter() - t1
    r = st.result()
    I_k, S2, ne = st.history()
    fill_ms, iter_ms = st.timing()
    st.close()
    timed_evals = int(ne[warmup:].sum())
    timed_ms = float(iter_ms[warmup:].sum())
    t_to, it_to = _time_to_target(I_k, S2, iter_ms, warmup, itg.exact)
    return {
        "precision": precision,
        "evals_per_s": timed_evals / (timed_ms * 1e-3),
        "ms_per_step": timed_ms / steps,
        "wall_ms_per_step": wall_timed * 1e3 / steps,
        "fill_ms_avg": float(np.mean(fill_ms[warmup:])),
        "evals_per_fill_avg": float(np.mean(ne[warmup:])),
        "result": r["result"], "error": r["error"], "chi2_dof": r["chi2_dof"],
        "error_unbiased": r["error_unbiased"], "chi2_dof_unbiased": r["chi2_dof_unbiased"],
        "rel_err_vs_exact": abs(r["result"] - itg.exact) / abs(itg.exact),
        "pull": (r["result"] - itg.exact) / r["error"],
        "pull_unbiased": (r["result"] - itg.exact) / r["error_unbiased"],
        "time_to_1e-4_ms": t_to, "iteration_reached_1e-4": it_to,
        "tot_nevals": int(r["tot_nevals"]), "compile_ms": r["compile_ms"],
        "setup_s": setup_s,
    }


def _form_difference(forms):
    """The FP32-compensated result against the FP64 form of the same run: the
    difference, relative to the FP64 value, and in units of the two forms'
    combined error (the forms draw the same Philox stream, but their maps and
    allocations diverge after the first refinement, so they are compared
    statistically)."""
    if "fp32" not in forms or "fp64" not in forms:
        return None
    a, b = forms["fp32"], forms["fp64"]
    diff = a["result"] - b["result"]
    return {"difference": diff, "rel_difference": diff / abs(b["result"]),
            "pull": diff / math.hypot(a["error"], b["error"])}


def cpu_baseline(itg, family, nbins, n_strat, n_eval_it, seed):
    """The CPU oracle (FP64 restatement of the reference algorithm) on a bounded
    sample of the same workload (n_eval_it samples per iteration; the caller
    may scale it down to keep the CPU time bounded), on every host core this
    process may use (OpenMP over sample windows): a run of 4 iterations and
    one of the first iteration alone, whose difference times the 3 adapted
    iterations (the value: after the first refinement the allocation and the
    sample count are those of the rest of the run); on one core, the first
    iteration only."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    fid = {"double_gaussian": orc.F_DGAUSS, "oscillatory": orc.F_OSC, "gaussian_peak": orc.F_GAUSS}[family]
    # every core of the affinity mask, up to the process's OpenMP thread budget
    # (OMP_NUM_THREADS: the host share a one-GPU job is given; the machine's
    # own count is recorded beside it)
    affinity = len(os.sched_getaffinity(0))
    budget = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    cores = min(affinity, budget) if budget > 0 else affinity

    def timed(threads, iters):
        p, keep = orc.make_params(itg.n_dim, nbins, n_strat, n_eval_it, iters, fid,
                                  list(itg.params), skip=0, seed=seed, threads=threads)
        t0 = time.perf_counter()
        r = orc.run(p, keep, n_strat ** itg.n_dim, snapshots=False)
        return time.perf_counter() - t0, [int(v) for v in r["hist_n"]]

    t1, n1 = timed(cores, 1)
    t4, n4 = timed(cores, 4)
    adapted = sum(n4[1:])
    v = adapted / max(t4 - t1, 1e-9)
    s1, m1 = timed(1, 1)
    return {"value": v, "unit": "evals/s", "cores": cores, "kind": "port",
            "sample": f"iterations 2-4 of the same workload ({adapted} evals, FP64, {t4 - t1:.1f} s on "
                      f"{cores} threads: a 4-iteration run less the first iteration's run); first "
                      f"iteration {n1[0] / t1:.3e} evals/s on {cores} threads, {m1[0] / s1:.3e} evals/s "
                      f"on 1 thread ({s1:.1f} s)",
            "value_first_iteration": n1[0] / t1, "value_1core": m1[0] / s1, "nproc": os.cpu_count(),
            "affinity_cores": affinity, "omp_num_threads": budget or None}


def _device_cu_clock():
    """(compute units, clock in Hz) of the current device."""
    import ctypes as C
    from cuvegas_amd import lib
    n_cu, clk = C.c_int(0), C.c_int(0)
    arch = C.create_string_buffer(64)
    if lib.cvg_device_info(C.byref(n_cu), C.byref(clk), arch, 64) != 0:
        sys.exit("no HIP device: " + lib.cvg_last_error().decode())
    return n_cu.value, clk.value * 1e3


def _workloads():
    with open(os.path.join(ROOT, "tests", "golden", "workloads.json")) as f:
        return json.load(f)


def _batch_fill_roofline(prog, items, theta, fams, iters, skip):
    """Roofline of the batched sampling kernel: an eager run of the same batch
    with HIP events around each iteration's one sampling launch; achieved =
    the algorithmic VALU slots of the iteration's evaluations (each item's
    slots per evaluation for its family and dimension, cuvegas_amd.roofline)
    over the launch's duration, against the issue peak.  The kernel's own
    overheads (runtime strata division, per-task dimension switch) are not
    counted as work."""
    import numpy as np
    from cuvegas_amd.batch import Batch
    from cuvegas_amd.roofline import slots_per_eval, valu_peak
    fam_name = {"gaussian": "gaussian_peak", "oscillatory": "oscillatory", "corner": "corner_peak"}
    S = np.array([slots_per_eval(d, fam_name[f]) for (d, _, _), f in zip(items, fams)], dtype=np.float64)
    b = Batch(prog, items, theta, max_it=iters, skip=skip, seed=0)
    b.set_timing(True)
    b.run(iters, use_graph=False)
    fill_ms, iter_ms = b.timing()
    ne = np.array([b.history(i)[2] for i in range(len(items))], dtype=np.float64)  # [item, iteration]
    b.close()
    slots = (S[:, None] * ne).sum(axis=0)  # per iteration
    peak = valu_peak(*_device_cu_clock())
    achieved = float(slots.sum() / (fill_ms.sum() * 1e-3))
    return {"bound": "valu", "kernel": "vg_bfill", "achieved": achieved / 1e12, "peak": peak / 1e12,
            "unit": "Tslot/s", "frac": achieved / peak, "fill_ms_avg": float(fill_ms.mean()),
            "fill_share_of_iteration": float(fill_ms.sum() / iter_ms.sum()),
            "slots_per_eval_mean": float(slots.sum() / ne.sum())}


def batch_bench(args, sequential=True):
    """512 independent integrals from the committed fixture (d ~ U{2..12},
    Gaussian / oscillatory / corner peak, budgets log-uniform in [1e4, 1e7] per
    iteration), 10 iterations, FP32-compensated: one batched sampling launch
    per iteration in a hipGraph replay, against the same kernels launched one
    integral at a time.  Returns the result dict."""
    import numpy as np
    from cuvegas_amd.autotune import autotune_batch, program
    from cuvegas_amd.batch import FAMILY_SOURCE, Batch, workload_from_items
    iters, skip = 10, 5
    items, theta, exact, fams = workload_from_items(_workloads()["batch_512"])
    if args.batch_config:
        blk, rl, tw = (int(v) for v in args.batch_config.split(","))
        tune = {"block": blk, "run_len": rl, "task_windows": tw, "table": []}
    else:
        _log("autotune")
        tune = autotune_batch(FAMILY_SOURCE, items, theta, max_dim=12, n_intervals=256, precision="fp32",
                              use_cache=False, verbose=True)
    extra = (getattr(args, "fill_options", "") or "").strip()
    if extra:  # same-box A/B of program options: a program of its own
        from cuvegas_amd.batch import BatchProgram
        prog = BatchProgram(FAMILY_SOURCE, 12, 256, "fp32", run_len=tune["run_len"],
                            options=f"block={tune['block']} task_windows={tune['task_windows']} {extra}")
    else:
        prog = program(FAMILY_SOURCE, 12, 256, "fp32", tune["block"], tune["run_len"], tune["task_windows"])
    _log("batched run", tune["block"], tune["run_len"], tune["task_windows"])
    b = Batch(prog, items, theta, max_it=iters, skip=skip, seed=0)
    t0 = time.perf_counter()
    b.run(iters, use_graph=True)
    r = b.results()
    batch_wall = time.perf_counter() - t0
    b.close()
    evals = int(r["nevals"].sum())
    pull = (r["result"] - exact) / r["error"]
    # A Gaussian peak is resolvable when uniform sampling expects at least 20
    # points inside it: N * (sigma sqrt(2 pi))^d >= 20.
    ok = np.array([fam != "gaussian" or N * (theta[off + 1] * math.sqrt(2 * math.pi)) ** d >= 20
                   for (d, N, off), fam in zip(items, fams)])
    out = {
        "metric": "batch of 512 integrals: total evals/s and p50/p99 per-integral completion (ms)",
        "value": evals / batch_wall, "unit": "evals/s", "steps": iters, "dtype": "f32",
        "data": "synthetic (seed 3 parameter draws, committed fixture tests/golden/workloads.json)",
        "config": {"workload": "512 integrals, d~U{2..12}, Gaussian/oscillatory/corner, neval "
                               "log-U[1e4,1e7]/iteration, nbins=256, 10 iterations, FP32-compensated",
                   "block": tune["block"], "run_len": tune["run_len"],
                   "task_windows": tune["task_windows"]},
        # the items advance in lockstep, one launch per iteration for all of
        # them: every item's last iteration ends in the same launch, so every
        # per-integral completion time is the batch's wall time
        "batched": {"wall_ms": batch_wall * 1e3, "device_loop_ms": r["loop_ms"], "evals": evals,
                    "completion_p50_ms": batch_wall * 1e3, "completion_p99_ms": batch_wall * 1e3,
                    "completion": "lockstep: all 512 complete with the last iteration's launch"},
        "accuracy": {"pull_rms_resolvable": float(np.sqrt(np.mean(pull[ok] ** 2))),
                     "pull_rms_resolvable_unbiased": float(np.sqrt(np.mean(
                         ((r["result"] - exact) / r["error_unbiased"])[ok] ** 2))),
                     "chi2_dof_median": float(np.median(r["chi2_dof"])),
                     "chi2_dof_unbiased_median": float(np.median(r["chi2_dof_unbiased"])),
                     "n_resolvable": int(ok.sum()),
                     "frac_within_3sigma_resolvable": float(np.mean(np.abs(pull[ok]) < 3)),
                     "frac_within_3sigma": float(np.mean(np.abs(pull) < 3)),
                     "median_rel_err": float(np.median(np.abs(r["result"] - exact) /
                                                       np.maximum(np.abs(exact), 1e-300)))},
        "autotune": tune["table"],
    }
    out["roofline"] = _batch_fill_roofline(prog, items, theta, fams, iters, skip)
    if not sequential:
        return out
    # Sequential baseline: the same integrals one after another through the same
    # kernels.  Each integral runs once, so its iterations are launched eagerly
    # (a graph per integral costs its capture; measured slower), and the
    # baseline takes the launch shape that suits it best among the autotuner's
    # candidates (small integrals want small tasks), not the batch's choice.
    # Completion of integral k = wall time from the start of the sequence to
    # the end of its last iteration (integrals in fixture order).
    from cuvegas_amd.autotune import DEFAULT_CANDIDATES
    _log("sequential runs")
    seq = None
    for blk, rl, tw in DEFAULT_CANDIDATES:
        sprog = program(FAMILY_SOURCE, 12, 256, "fp32", blk, rl, tw)
        seqs = [Batch(sprog, [items[i]], theta, max_it=iters, skip=skip, seed=0) for i in range(len(items))]
        done = []
        t1 = time.perf_counter()
        for sb in seqs:
            sb.run(iters, use_graph=False)
            done.append(time.perf_counter() - t1)
        seq_evals = sum(int(sb.results()["nevals"].sum()) for sb in seqs)
        for sb in seqs:
            sb.close()
        _log("sequential", blk, rl, tw, f"{done[-1] * 1e3:.1f} ms")
        if seq is None or done[-1] < seq["wall"]:
            seq = {"wall": done[-1], "done": np.array(done) * 1e3, "evals": seq_evals, "config": (blk, rl, tw)}
    out["sequential"] = {"wall_ms": seq["wall"] * 1e3, "evals": seq["evals"],
                         "evals_per_s": seq["evals"] / seq["wall"], "launch": "eager",
                         "block": seq["config"][0], "run_len": seq["config"][1],
                         "task_windows": seq["config"][2],
                         "completion_p50_ms": float(np.percentile(seq["done"], 50)),
                         "completion_p99_ms": float(np.percentile(seq["done"], 99))}
    out["speedup_vs_sequential"] = seq["wall"] / batch_wall
    return out


def cpu_config_bench(args):
    """The CPU-runnable configuration at its own size: 4-D normalised Gaussian
    (sigma = 0.1), 1024 bins, n_strat 22 (234,256 cubes), 1e6 evaluations per
    iteration, 10 iterations (5 averaged), FP64, seed 0.  The whole 10-iteration
    job on the GPU (FP64 form, one graph replay, after a first run that loads
    the program) and in the CPU oracle (the FP64 restatement of the reference
    algorithm, on every core of the process's OpenMP budget and on one core),
    each with its error against the closed form.  At this size a GPU iteration
    is a few launches' latency, not throughput."""
    from cuvegas_amd import Integrator
    from cuvegas_amd import integrands as I
    d, nbins, n_strat, N, iters, skip = 4, 1024, 22, 1_000_000, 10, 5
    itg = I.gaussian_peak(d, 0.5, 0.1)
    runs = []
    prog = None
    for _ in range(2):
        st = Integrator(itg.source, d, N, n_intervals=nbins, n_strat=n_strat, max_it=iters, skip=skip,
                        seed=0, precision="fp64", program=prog)
        prog = st.program
        st.run(iters, use_graph=True)
        runs.append(st.result())
        st.close()
    r = runs[-1]
    gpu = {"loop_ms": r["loop_ms"], "evals": int(r["tot_nevals"]),
           "evals_per_s": r["tot_nevals"] / (r["loop_ms"] * 1e-3), "result": r["result"], "error": r["error"],
           "pull": (r["result"] - itg.exact) / r["error"]}
    out = {"workload": f"4-D Gaussian sigma=0.1, nbins={nbins}, n_strat={n_strat} ({n_strat ** d:,} cubes), "
                       f"{N:.0e} evals/iteration, {iters} iterations ({skip} skipped), FP64, seed 0",
           "exact": itg.exact, "gpu_fp64": gpu}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        affinity = len(os.sched_getaffinity(0))
        budget = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        cores = min(affinity, budget) if budget > 0 else affinity
        for threads in (cores, 1):
            p, keep = orc.make_params(d, nbins, n_strat, N, iters, orc.F_GAUSS, list(itg.params), skip=skip,
                                      seed=0, threads=threads)
            t0 = time.perf_counter()
            o = orc.run(p, keep, n_strat ** d, snapshots=False)
            dt = time.perf_counter() - t0
            out[f"cpu_oracle_{threads}_threads" if threads > 1 else "cpu_oracle_1_thread"] = {
                "wall_ms": dt * 1e3, "evals": int(o["tot_nevals"]), "evals_per_s": o["tot_nevals"] / dt,
                "result": o["result"], "error": o["error"], "pull": (o["result"] - itg.exact) / o["error"]}
        out["cores"] = cores
        out["gpu_over_cpu_all_cores"] = out[f"cpu_oracle_{cores}_threads" if cores > 1 else
                                            "cpu_oracle_1_thread"]["wall_ms"] / gpu["loop_ms"]
    return out


def mix6_bench(args, neval=None):
    """6-D four-component Gaussian mixture (32 parameters from the committed
    fixture, numpy default_rng(2)): 5e7 evaluations over 10 iterations (5e6 per
    iteration, 5 averaged), forward with recording, then the backward replay of
    the 5 recorded iterations, d value / d theta.  Reports forward evals/s per
    form, backward time (HIP events around the replay), and the gradient's pulls
    against the closed-form erf gradient.  Returns the result dict."""
    import numpy as np
    from cuvegas_amd import Integrator
    from cuvegas_amd import integrands as I
    theta = np.array(_workloads()["mixture_6d"]["theta"])
    d, nbins, iters, skip = 6, 1024, 10, 5
    n_eval_it = int(neval or 5e6)
    itg = I.gaussian_mixture_param(d, theta)
    exact_grad = np.array(itg.exact_grad(theta))
    forms = {}
    for prec in (["fp64", "fp32"] if args.only is None else [args.only]):
        _log("mixture", prec)
        st = Integrator(itg.source, d, n_eval_it, n_intervals=nbins, max_it=iters, skip=skip, seed=0,
                        precision=prec, theta=theta,
                        options=("grad " + (getattr(args, "fill_options", "") or "")).strip())
        st.set_record(True)
        st.set_timing(True)
        st.run(iters, use_graph=False)
        r = st.result()
        I_k, S2, ne = st.history()
        fill_ms, iter_ms = st.timing()
        st.backward()  # first call allocates the replay buffers
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            v, vv, g, gv = st.backward()
        bw_ms = (time.perf_counter() - t0) * 1e3 / reps
        bw_dev = st.backward_timing()
        replayed = int(ne[skip:].sum())
        pull = (g - exact_grad) / np.sqrt(gv)
        # roofline of the replay kernel (vg_grad and its unit reduction): the
        # replayed samples' algorithmic slots over the kernels' device time
        from cuvegas_amd.roofline import slots_per_replayed_sample, valu_peak
        s_grad = slots_per_replayed_sample(d, "gaussian_mixture", int(theta.size))
        peak = valu_peak(*_device_cu_clock())
        grad_roof = {"bound": "valu", "kernel": "vg_grad", "slots_per_sample": s_grad,
                     "achieved": s_grad * replayed / (bw_dev * 1e-3) / 1e12, "peak": peak / 1e12,
                     "unit": "Tslot/s", "frac": s_grad * replayed / (bw_dev * 1e-3) / peak}
        forms[prec] = {
            "precision": prec, "n_strat": st.n_strat, "n_cubes": st.n_cubes,
            "forward_evals_per_s": float(ne.sum() / (iter_ms.sum() * 1e-3)),
            "forward_ms_per_iteration": float(iter_ms.mean()), "fill_ms_avg": float(fill_ms.mean()),
            "result": r["result"], "error": r["error"], "chi2_dof": r["chi2_dof"],
            "error_unbiased": r["error_unbiased"], "chi2_dof_unbiased": r["chi2_dof_unbiased"],
            "pull_value": (r["result"] - itg.exact) / r["error"],
            "backward_ms": bw_ms, "backward_replay_kernel_ms": bw_dev,
            "backward_replayed_evals": replayed,
            "backward_evals_per_s": replayed / (bw_ms * 1e-3),
            "grad_pull_rms": float(np.sqrt(np.mean(pull ** 2))),
            "grad_pull_max": float(np.max(np.abs(pull))), "n_theta": int(theta.size),
            "grad_within_3sigma": bool(np.all(np.abs(pull) < 3)),
            "backward_roofline": grad_roof,
        }
        _log(prec, json.dumps(forms[prec]))
        st.close()
    head = forms.get("fp32") or forms["fp64"]
    return {
        "metric": "6-D Gaussian mixture: forward evals/s, backward (32-parameter gradient) time",
        "value": head["forward_evals_per_s"], "unit": "evals/s", "steps": iters,
        "dtype": "f32" if head is forms.get("fp32") else "f64",
        "data": "synthetic (mixture parameters from numpy default_rng(2), committed fixture)",
        "config": {"workload": f"6-D 4-component Gaussian mixture, nbins={nbins}, {n_eval_it:.0e} "
                               f"evals/iteration, {iters} iterations ({iters - skip} averaged and "
                               "replayed by the backward)", "exact": itg.exact},
        "forms": forms}


def copy_probe():
    """Three device-to-device copies of 1 GiB: the calibration kernel for the
    FETCH_SIZE / WRITE_SIZE counters (bytes known exactly)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = 1 << 30
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(n)) != 0:
            sys.exit("hipMalloc failed")
    if hip.hipMemset(a, 1, C.c_size_t(n)) != 0:
        sys.exit("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, C.c_size_t(n), 3) != 0:  # device to device
            sys.exit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)


def _traffic_from_profile(workload):
    """HBM bytes per vg_fill launch (FP32 form) from the committed PMC summary of
    the workload (profiles/vg_fill_<workload>_*.md, line traffic_bytes_per_launch)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"vg_fill_{workload}_*.md"))):
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    return None


def single_bench(args, workload, warmup, steps, with_cpu=True, neval=None):
    """One single-integral workload in both forms (dg8, osc16 or gauss4):
    evals/s over the timed iterations, time to 1e-4, errors against the closed
    form, the sampling kernel's roofline and the CPU oracle on a bounded
    sample.  Returns the result dict (the bench line's fields)."""
    from cuvegas_amd import integrands as I
    from cuvegas_amd import lib
    from cuvegas_amd.roofline import device_rates, lds_roofline, slots_per_eval, valu_peak
    import ctypes as C

    seed = 0
    if workload == "dg8":
        d, nbins, n_strat, family = 8, 1024, 8, "double_gaussian"
        n_eval_it = int(neval or 1e8)
        itg = I.double_gaussian(d, 0.33, 0.67, 0.05)
        desc = (f"8-D double Gaussian peak sigma=0.05, nbins=1024, n_strat=8 (16,777,216 cubes), "
                f"{n_eval_it:.0e} evals/iteration")
        cpu_neval = n_eval_it
    elif workload == "gauss4":
        # the CPU configuration's integrand (4-D normalised Gaussian, sigma = 0.1, nbins 1024)
        # at 1e8 evaluations per iteration; n_strat by the reference's formula
        d, nbins, family = 4, 1024, "gaussian_peak"
        n_eval_it = int(neval or 1e8)
        n_strat = max(1, int(math.pow(n_eval_it / 4.0, 1.0 / d)))
        itg = I.gaussian_peak(d, 0.5, 0.1)
        desc = (f"4-D Gaussian peak sigma=0.1, nbins=1024, n_strat={n_strat} ({n_strat ** d:,} cubes), "
                f"{n_eval_it:.0e} evals/iteration")
        cpu_neval = n_eval_it
    else:  # osc16: parameters drawn once (numpy default_rng(1)), committed as a fixture
        osc = _workloads()["oscillatory_16d"]
        d, nbins, n_strat, family = 16, 512, 2, "oscillatory"
        n_eval_it = int(neval or 2e8)
        itg = I.oscillatory(d, osc["u"], osc["a"])
        desc = (f"16-D oscillatory cos(2 pi u + sum a_i x_i), a_i ~ U(0,5), nbins=512, n_strat=2 "
                f"(65,536 cubes), {n_eval_it:.0e} evals/iteration")
        cpu_neval = n_eval_it // 10  # bounded CPU sample: ~1/10 of an iteration

    n_cu, clk = C.c_int(0), C.c_int(0)
    arch = C.create_string_buffer(64)
    rc = lib.cvg_device_info(C.byref(n_cu), C.byref(clk), arch, 64)
    if rc != 0:
        sys.exit("no HIP device: " + lib.cvg_last_error().decode())
    clock_hz = clk.value * 1e3
    _log(workload, "device", arch.value.decode(), n_cu.value, "CUs", clock_hz / 1e9, "GHz")

    forms = {}
    for prec in (["fp64", "fp32"] if args.only is None else [args.only]):
        _log(workload, "running", prec)
        forms[prec] = run_form(prec, itg, n_eval_it, warmup, steps, n_strat, nbins, seed,
                               options=getattr(args, "fill_options", None) or None,
                               run_len=getattr(args, "run_len", 0) or 0)
        _log(workload, prec, json.dumps(forms[prec]))
    head = forms.get("fp32") or forms["fp64"]

    S = slots_per_eval(d, family)
    peak = valu_peak(n_cu.value, clock_hz)
    achieved = S * head["evals_per_fill_avg"] / (head["fill_ms_avg"] * 1e-3)
    roof = {"bound": "valu", "kernel": "vg_fill",
            "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "Tslot/s",
            "frac": achieved / peak, "slots_per_eval": S,
            "traffic": _traffic_from_profile(workload) if head is forms.get("fp32") else None,
            "fill_ms_avg": head["fill_ms_avg"], "evals_per_fill": head["evals_per_fill_avg"]}
    # the same launch against the LDS: d bin updates per evaluation at the rate the
    # device's LDS sustains them at random addresses (measured in this run)
    rates = device_rates()
    roof["lds"] = lds_roofline(head["evals_per_fill_avg"], d, head["fill_ms_avg"] * 1e-3, n_cu.value,
                               clock_hz, rates["lds_pair_random"])
    if roof["traffic"]:
        # the same launch against the HBM roofline: measured bytes per launch / its duration
        from cuvegas_amd.roofline import MI355X_HBM_BPS
        hbm = roof["traffic"] / (head["fill_ms_avg"] * 1e-3)
        roof["hbm"] = {"achieved": hbm / 1e9, "peak": MI355X_HBM_BPS / 1e9, "unit": "GB/s",
                       "frac": hbm / MI355X_HBM_BPS}

    cpu = None
    if with_cpu:
        _log(workload, "cpu baseline")
        cpu = cpu_baseline(itg, family, nbins, n_strat, cpu_neval, seed)

    return {
        "metric": "integrand evals/s + time to 1e-4 rel. error, FP32-comp vs plain FP64, 1 GPU",
        "value": head["evals_per_s"], "unit": "evals/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if head is forms.get("fp32") else "f64",
        "data": "synthetic (closed-form integrand, Philox4x32-10 seed 0)",
        "config": {"workload": f"{desc}, alpha=0.5, beta=0.75, {warmup}+{steps} iterations",
                   "n_eval_it": n_eval_it, "exact": itg.exact},
        "forms": forms,
        "fp32_over_fp64": (forms["fp32"]["evals_per_s"] / forms["fp64"]["evals_per_s"]
                           if "fp32" in forms and "fp64" in forms else None),
        "fp32_vs_fp64": _form_difference(forms),
        "roofline": roof,
        "cpu_baseline": cpu,
        "device": {"arch": arch.value.decode(), "n_cu": n_cu.value, "clock_hz": clock_hz,
                   "cycles_per_wave64_instruction": rates},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=15)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--neval", type=float, default=None,
                    help="evaluations per iteration (default: the workload's)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extras", action="store_true",
                    help="dg8: leave out the osc16, mix6 and batch512 workloads that the default line carries")
    ap.add_argument("--only", choices=["fp32", "fp64"], default=None)
    ap.add_argument("--fill-options", default="",
                    help="sampling-program options (e.g. 'bin_copies=1') for A/B runs")
    ap.add_argument("--run-len", type=int, default=0,
                    help="dg8/osc16/gauss4: samples per lane run (power of two; default: the library's 16)")
    ap.add_argument("--batch-config", default="",
                    help="batch512: 'block,run_len,task_windows' instead of autotuning")
    ap.add_argument("--skip-sequential", action="store_true",
                    help="batch512: time the batched run only (for profiling)")
    ap.add_argument("--workload", choices=["dg8", "osc16", "gauss4", "mix6", "batch512", "copy"], default="dg8",
                    help="dg8: the 8-D double Gaussian headline (with the other configurations under "
                         "extra keys unless --no-extras); osc16: 16-D oscillatory; "
                         "gauss4: the 4-D Gaussian of the CPU configuration at 1e8 evals/iteration "
                         "(a cheap integrand: bin atomics dominate); "
                         "mix6: 6-D mixture forward + gradient; batch512: 512-integral batch; "
                         "copy: 1 GiB device copies for counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding of the sampling pass is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        sys.exit("steps must be >= 1 and warmup >= 0")
    if args.workload == "batch512":
        out = batch_bench(args, sequential=not args.skip_sequential)
        out.update(n_gpus=1, warmup=0, ms_per_step=out["batched"]["wall_ms"] / out["steps"],
                   higher_is_better=True, scaling="weak", vs_baseline=None)
        print(json.dumps(out), flush=True)
        return
    if args.workload == "copy":
        copy_probe()
        return
    if args.workload == "mix6":
        out = mix6_bench(args, args.neval)
        out.update(n_gpus=1, warmup=0, ms_per_step=out["forms"][next(iter(out["forms"]))]["forward_ms_per_iteration"],
                   higher_is_better=True, scaling="weak", vs_baseline=None)
        print(json.dumps(out), flush=True)
        return
    out = single_bench(args, args.workload, args.warmup, args.steps, with_cpu=not args.no_cpu, neval=args.neval)
    if args.workload == "dg8" and not args.no_extras:
        # time to 1e-4 of both forms over 64 more seeds (5 + 15 iterations each)
        from cuvegas_amd import integrands as I
        _log("dg8 time to 1e-4 over seeds")
        out["time_to_1e-4_seeds"] = time_to_target_seeds(
            I.double_gaussian(8, 0.33, 0.67, 0.05), int(args.neval or 1e8), 8, 1024, 5, 20, range(1, 65))
        # the other configurations of the hot path, each at its own size and
        # iteration count, in the same line (their own timed regions)
        extras = {}
        osc = single_bench(args, "osc16", 5, 10, with_cpu=False)
        extras["osc16"] = {k: osc[k] for k in ("value", "unit", "ms_per_step", "dtype", "config", "forms",
                                              "fp32_over_fp64", "fp32_vs_fp64", "roofline")}
        extras["mix6"] = mix6_bench(args)
        extras["batch512"] = batch_bench(args)
        _log("cpu configuration")
        extras["cpu_config"] = cpu_config_bench(args)
        out["configs"] = extras
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
