#!/usr/bin/env python3
"""Benchmark of the Neurite cable-equation solvers on one MI355X.

Headline (`value`): config 4, the batched parameter sweep: 4,096 seeded trees with
compartment counts log-uniform in [500, 50,000] (~42.7 M compartments), per-tree
strain and stimulus, backward Euler at dt = 25 us, all trees advanced together by
the batched tree-solve kernel. A "step" is one time step of every tree (HH gating,
matrix refresh, right-hand side and the exact solve). Inputs are on the device
before the timed region; `value` = compartments x timed steps / wall time of the
timed steps. The form reported is the fastest of FP64 and FP32 + FP64 refinement
whose max |dV| against the oracle's exact FP64 solve (16 trees across the size
range, same steps) is <= 1e-4 mV.

The same JSON line carries the metric's latency part, config 5 (a served stream:
Poisson arrivals at 200 requests/s, compartments log-normal around 20,000,
5..50 ms each at dt = 25 us) with p50/p99 request latency and its own error
against the oracle on a sample of requests, and legs for configs 1-3. Every GPU
form's error is taken against the oracle (test infrastructure, imported here only
for those checks and for the CPU baselines, which time the oracle's restatement
of the reference CPU method on this host).
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "168861_amd")]

METRIC = "compartment-steps/s at max|ΔV|≤1e-4 mV vs FP64; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
TOL_MV = 1e-4


# ------------------------------------------------------------ algorithmic bytes
def tree_bytes(N, H, nnz, steps, solves, vbytes):
    """Algorithmic bytes of tree-solved steps (k_implicit_tree*, k_batch_tree), per
    element, counting every chunk as one that reuses its factors:
    update phase per step: 144 B per HH element (V, gates r/w, parameters, the
    refreshed diagonal) and per element V 8 + code 4 + c_m dx 8 + K dx 8 + input 8
    + the right-hand side written 8 (44 B);
    per tree solve: the forward sweep reads r, the multiplier and the weight and
    writes the partial right-hand side r'; the back-substitution reads r', the pivot
    reciprocal, the coupling to the next row and the fill-in g and writes V (FP64: 8
    values of 8 B + V 8 = 72 B; FP32: 8 values of 4 B + V read and written 16 = 48 B);
    refinement only, per residual evaluation (solves + steps of them): the matrix
    (12 nnz + 4 (N+1)), the gathered V (8 nnz), b (8 N), the residual written into
    the tree system (vbytes N)."""
    per_step = 144 * H + 44 * N
    if vbytes == 8:
        return steps * per_step + solves * 72 * N
    per_solve = (8 * vbytes + 16) * N
    per_resid = 12 * nnz + 4 * (N + 1) + 8 * nnz + 8 * N + vbytes * N
    return steps * per_step + solves * per_solve + (solves + steps) * per_resid


def bicgstab_bytes(N, H, nnz, steps, iters):
    """Per step 144 B per HH element + 40 B per element (update phase); per
    BiCGSTAB iteration 2 SpMV (12 nnz + 4 (N+1) + 16 N) + 21 vector passes of 8 N."""
    spmv = 12 * nnz + 4 * (N + 1) + 16 * N
    return steps * (144 * H + 40 * N) + iters * (2 * spmv + 21 * 8 * N)


def gmres_bytes(N, H, nnz, steps, iters, cycles, vbytes):
    """GMRES(m) steps (k_implicit_gmres<double|float>), cycles of a run taken as
    equally long (iters / cycles Arnoldi steps): update phase per step 144 B per HH
    element + 40 B per element; per cycle the FP64 residual (12 nnz + 4 (N+1) +
    24 N) and the update x += V y (kk vbytes N + 16 N); per Arnoldi step the SpMV
    (vbytes nnz + 4 nnz + 4 (N+1) + 3 vbytes N) and CGS2 ((4 (k+1) + 6) vbytes N)."""
    if cycles <= 0:
        return steps * (144 * H + 40 * N)
    kk = iters / cycles
    per_cycle = (12 * nnz + 4 * (N + 1) + 24 * N) + kk * vbytes * N + 16 * N
    spmv = vbytes * nnz + 4 * nnz + 4 * (N + 1) + 3 * vbytes * N
    cgs = (2 * kk * (kk + 1) + 6 * kk) * vbytes * N
    return steps * (144 * H + 40 * N) + cycles * (per_cycle + kk * spmv + cgs)


def explicit_bytes(N, H):
    """52 B per passive element (V r/w, a, r_a dx, W dx, K dx, code) and 156 B per
    HH element (V r/w, a, r_a dx, m/h/n r/w, 9 parameters, code)."""
    return 52 * (N - H) + 156 * H


def roofline(nbytes, kernel_ms, kernel, launches, traffic=None):
    ach = nbytes / max(launches, 1) / (kernel_ms / max(launches, 1) / 1e3) / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
            "launches": launches, "kernel_ms_per_launch": kernel_ms / max(launches, 1),
            "algorithmic_bytes_per_launch": nbytes / max(launches, 1)}


def measured_traffic(kernel, workload, steps_per_launch):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE + WRITE_SIZE
    runs of this kernel on this workload (per step, copy-calibrated), scaled to the
    launch's steps; None when no run covers it."""
    tfile = os.path.join(REPO, "profiles", "traffic.json")
    if not os.path.exists(tfile):
        return None
    t = json.load(open(tfile)).get(f"{kernel} | {workload}")
    return None if not t else t["hbm_bytes_per_step"] * steps_per_launch


# -------------------------------------------------------------------- helpers
def say(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_info():
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable,
            "threads_used": min(16, usable)}


def _oracle():
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle  # the checker: errors of the GPU forms and the CPU baselines only
    oracle.lib()
    return oracle


def oracle_exact(orc, p, steps, a0, a1):
    ref = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT)
    ref.run(1, steps, a0, a1, 0, -1)
    return ref.snapshot()


def mask_of(p):
    m = np.ones(p.N, bool)
    m[[0, p.N - 1]] = False
    m[np.asarray(p.terminals)] = False
    return m


def err_mv(v, ref, p):
    return float(np.max(np.abs(np.asarray(v) - ref)[mask_of(p)])) * 1e3


def split_launches(total, per):
    out = [per] * (total // per)
    if total % per:
        out.append(total % per)
    return out


def cpu_pool(orc, jobs, threads, method):
    """Runs jobs (problem, steps, a0, a1) on `threads` host threads, one problem per
    job, with the oracle (its calls release the interpreter lock, so the threads
    run in parallel); returns (compartment-steps, wall seconds, failures)."""
    def one(job):
        p, n, a0, a1 = job
        if method == "gmres":
            s = orc.Implicit(p, variant=orc.VAR_CPU, method=orc.M_GMRES, restart=150,
                             tol=1e-14, tol_mode=0)
        else:
            s = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT)
        bad = 0
        for i in range(1, n + 1):
            s.update(a0 <= i <= a1)
            bad += s.calculate() != 0
        s.close()
        return p.N * n, bad

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(one, jobs))
    wall = time.perf_counter() - t0
    return sum(r[0] for r in res), wall, sum(r[1] for r in res)


# ------------------------------------------------------- config 4: the sweep
def sweep_problems(nh, wl, count):
    t0 = time.perf_counter()
    geoms = [wl.sweep_tree(i) for i in range(count)]
    imp = [nh.setup(g, dT=r.dT)[0] for g, r in geoms]
    gen_s = time.perf_counter() - t0
    return geoms, imp, gen_s


def run_sweep_form(nh, probs, kw, warmup, steps, per_launch, tune_rows=0, shapes=None):
    """Adds every problem as a slot for warmup + steps steps; untimed warm-up
    launches, then the timed steps. Implicit: the launch shapes (chunk lengths and
    workgroup size per size class) are `shapes` when given (a tune() result kept
    from an earlier run), else come from nrt_batch_tune on a sample of the same
    problems (tune_rows compartments per class) before the slots are added; with
    neither, the first four launches pick each class's workgroup size."""
    warmup = max(warmup, 4)
    b = nh.Batch(**kw)
    tuned = None
    if shapes is not None:
        b.set_shapes(shapes["classes"])
        tuned = dict(shapes, reused=True)
    elif tune_rows > 0:
        tuned = b.tune(probs, steps=per_launch, rows_per_class=tune_rows)
    t0 = time.perf_counter()
    ids = [b.add(p, warmup + steps, 1, 20) for p in probs]
    add_s = time.perf_counter() - t0
    for k in split_launches(warmup, max(1, warmup // 5)):
        b.step(k)
    implicit = kw["kind"] == nh.IMPLICIT
    inner0 = [b.info(i)["inner_iters"] for i in ids] if implicit else None
    st0 = b.stats()
    t0 = time.perf_counter()
    for k in split_launches(steps, per_launch):
        b.step(k)
    wall = time.perf_counter() - t0
    st1 = b.stats()
    out = {"wall_s": wall, "kernel_ms": st1["kernel_ms"] - st0["kernel_ms"],
           "launches": len(split_launches(steps, per_launch)),
           "compartment_steps": st1["compartment_steps"] - st0["compartment_steps"],
           "add_s": add_s, "device_bytes": st1["device_bytes"],
           "threads": st1["threads"] if implicit else 256,
           "class_threads": list(st1["class_threads"]) if implicit else None,
           "tune": tuned}
    if implicit:
        infos = [b.info(i) for i in ids]
        out["solves"] = [inf["inner_iters"] - s0 for inf, s0 in zip(infos, inner0)]
        out["nnz"] = [inf["nnz"] for inf in infos]
        out["status"] = max(inf["status"] for inf in infos)
    return b, ids, out


def load_shapes(path):
    """Launch shapes kept by an earlier run ({form: tune() result}), if any."""
    if not path or not os.path.exists(path):
        return {}
    return json.load(open(path))


def save_shapes(path, form, tuned):
    kept = load_shapes(path)
    kept[form] = tuned
    json.dump(kept, open(path, "w"), indent=1)


def sweep_leg(nh, wl, orc, a, info):
    say(f"config 4: building {a.sweep_trees} trees")
    geoms, imp, gen_s = sweep_problems(nh, wl, a.sweep_trees)
    Ntot = sum(p.N for p in imp)
    H = sum(p.H for p in imp)
    order = np.argsort([p.N for p in imp])
    sample = [int(order[k]) for k in np.linspace(0, len(imp) - 1, 16).astype(int)]
    per_launch = min(a.steps, 10)
    wname = f"sweep {a.sweep_trees} trees"
    leg = {"workload": f"{a.sweep_trees} seeded trees, compartments log-uniform 500..50,000, depth "
                       "U{4..12} (capped at >= 2 compartments per segment), per-tree strain "
                       "U[0,0.3], stimulus U[1,20] nA at the soma for 0.5 ms (steps 1-20); "
                       "implicit dt = 25 us; explicit dt = 0.8 x each tree's critical step",
           "trees": a.sweep_trees, "compartments": Ntot, "hh_elements": H, "setup_s": gen_s,
           "timed_steps": a.steps, "warmup_steps": a.warmup, "steps_per_launch": per_launch,
           "oracle_sample": f"{len(sample)} trees across the size range, "
                            f"N {imp[sample[0]].N}..{imp[sample[-1]].N}"}
    forms = {"implicit_fp64": (dict(kind=nh.IMPLICIT, krylov=nh.DIRECT), 8),
             "implicit_mixed": (dict(kind=nh.IMPLICIT, krylov=nh.DIRECT,
                                     precision=nh.FP32_REFINE, refine_tol=1e-10), 4)}
    for name, (kw, vb) in forms.items():
        if name not in a.sweep_forms:
            continue
        say(f"config 4: {name}")
        kept = load_shapes(a.shapes).get(name)
        b, ids, r = run_sweep_form(nh, imp, kw, a.warmup, a.steps, per_launch, a.tune_rows,
                                   kept)
        if a.shapes and kept is None and r["tune"] is not None:
            save_shapes(a.shapes, name, r["tune"])
        snaps = {i: b.snapshot(ids[i]) for i in sample}
        done = b.info(ids[sample[0]])["steps_done"]
        b.close()
        nb = sum(tree_bytes(p.N, p.H, nz, a.steps, sv, vb)
                 for p, nz, sv in zip(imp, r["nnz"], r["solves"]))
        # one k_batch_tree launch per size class, concurrent: the roofline is the
        # group's (all classes' bytes over the time from the first start to the last end)
        kern = f"k_batch_tree<{'double' if vb == 8 else 'float'},*>"
        rec = {"compartment_steps_per_s": r["compartment_steps"] / r["wall_s"],
               "ms_per_step": r["wall_s"] / a.steps * 1e3, "status": r["status"],
               "solves_per_step": sum(r["solves"]) / len(imp) / a.steps,
               "workgroup_threads": r["threads"],
               "workgroup_threads_by_size_class": dict(zip(
                   ["N<4096", "4096<=N<16384", "N>=16384"], r["class_threads"])),
               "launch_shape_tuning": r["tune"],
               "add_s": r["add_s"],
               "device_bytes": r["device_bytes"],
               "roofline": roofline(nb, r["kernel_ms"], kern, r["launches"],
                                    measured_traffic(kern, wname, per_launch))}
        if orc is not None:
            rec["max_abs_dV_mV_vs_oracle"] = max(
                err_mv(snaps[i], oracle_exact(orc, imp[i], done, 1, 20), imp[i]) for i in sample)
        leg[name] = rec
    if "explicit" not in a.sweep_forms:
        return leg
    # explicit: every tree at 0.8 x its own critical step
    say("config 4: explicit")
    exp_ = [nh.setup(g, dT=None, dt_factor=0.8)[0] for g, _ in geoms]
    b, ids, r = run_sweep_form(nh, exp_, dict(kind=nh.EXPLICIT), a.warmup, a.steps, per_launch)
    ex_snap = {i: b.snapshot(ids[i]) for i in sample[::4]}
    ex_done = b.info(ids[sample[0]])["steps_done"]
    b.close()
    nb = explicit_bytes(sum(p.N for p in exp_), sum(p.H for p in exp_)) * a.steps
    rec = {"compartment_steps_per_s": r["compartment_steps"] / r["wall_s"],
           "ms_per_step": r["wall_s"] / a.steps * 1e3,
           "roofline": roofline(nb, r["kernel_ms"], "k_explicit_batch", a.steps,
                                measured_traffic("k_explicit_batch", wname, 1))}
    if orc is not None:
        errs = []
        for i, v in ex_snap.items():
            ex = orc.Explicit(exp_[i])
            ex.run(1, ex_done, 1, 20)
            errs.append(err_mv(v, ex.snapshot(), exp_[i]))
        rec["max_abs_dV_mV_vs_oracle"] = max(errs)
        rec["oracle_sample"] = "4 trees of the sample, smallest to largest"
    leg["explicit"] = rec
    if orc is not None and not a.no_cpu_baseline:
        say("config 4: CPU baseline")
        th = info["threads_used"]
        jobs = [(imp[i], 1, 1, 20) for i in range(min(64, len(imp)))]
        cs, wall, bad = cpu_pool(orc, jobs, th, "gmres")
        leg["cpu_baseline"] = {
            "value": cs / wall, "unit": "compartment-steps/s", "cores": th, "kind": "port",
            "sample": f"the first step of each of the first {len(jobs)} trees of the sweep "
                      f"({sum(j[0].N for j in jobs)} compartments), one tree per thread on "
                      f"{th} threads of {info['cpu_model']}, with the reference CPU method: "
                      "restarted GMRES(150), x0 = 0, ||r|| <= 1e-14 A (solverIm-cpu.cpp "
                      "semantics), oracle C++ build" + (f"; {bad} steps not converged" if bad else ""),
            "wall_s": wall}
        cs, wall, _ = cpu_pool(orc, [(imp[i], 20, 1, 20) for i in range(min(64, len(imp)))], th,
                               "direct")
        leg["cpu_exact_elimination"] = {
            "value": cs / wall, "cores": th,
            "sample": "20 steps of each of the same trees with the exact (Hines) elimination, "
                      "one tree per thread"}
    return leg


# ------------------------------------------------------ config 5: the stream
def served_requests(nh, wl, count):
    reqs = []
    for i in range(count):
        g, r, gap = wl.served_request(i)
        p = nh.setup(g, dT=r.dT)[0]
        reqs.append((p, r.steps(r.dT), gap))
    return reqs


def served_stream(nh, reqs, precision, fetch_ids):
    kw = dict(kind=nh.IMPLICIT, krylov=nh.DIRECT, precision=precision, refine_tol=1e-10)
    srv = nh.Server(steps_per_launch=16, **kw)
    t0 = time.perf_counter()
    due = 0.0
    rids = []
    for p, n, gap in reqs:
        due += gap
        wait = t0 + due - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        rids.append(srv.submit(p, n, 1, 20))
    srv.wait_idle(600.0)
    wall = time.perf_counter() - t0
    st = srv.stats()
    fetched = {i: srv.fetch(rids[i])[1] for i in fetch_ids}
    srv.close()
    return fetched, {
        "requests": len(reqs), "completed": st["completed"], "failed": st["failed"],
        "arrival_window_s": due, "wall_s": wall,
        "compartment_steps_per_s": st["compartment_steps"] / wall,
        "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
        "launches": st["launches"],
        "loop_ms": {k: st[k] for k in ("launch_ms", "admit_ms", "wait_ms", "retire_ms",
                                       "batch_kernel_ms")}}


def served_leg(nh, wl, orc, a, info):
    reqs = served_requests(nh, wl, a.serve_requests)
    # oracle sample: every 33rd request (its full run)
    sample = list(range(0, len(reqs), max(1, len(reqs) // 12)))[:12]
    leg = {"workload": "Poisson arrivals 200/s, compartments log-normal (median 20000, sigma 1, "
                       "cap 2M), 5..50 ms at dt 25 us, stimulus 10 nA for 0.5 ms",
           "largest_request": max(p.N for p, _, _ in reqs),
           "offered_compartment_steps": sum(p.N * n for p, n, _ in reqs),
           "oracle_sample": f"{len(sample)} requests (every {max(1, len(reqs) // 12)}th)"}
    for name, prec in (("fp64", nh.FP64), ("mixed", nh.FP32_REFINE)):
        say(f"config 5: served {name}")
        fetched, rec = served_stream(nh, reqs, prec, sample)
        if orc is not None:
            rec["max_abs_dV_mV_vs_oracle"] = max(
                err_mv(fetched[i], oracle_exact(orc, reqs[i][0], reqs[i][1], 1, 20), reqs[i][0])
                for i in sample)
        leg[name] = rec
    # the plain FP64 GPU form (the reference's method: unpreconditioned BiCGSTAB,
    # relative 1e-10) on a sample of the same requests, one handle each
    plain, fails = [], 0
    for p, n, _ in reqs[:4]:
        with nh.Solver(p, kind=nh.IMPLICIT, krylov=nh.BICGSTAB, tol=1e-10,
                       tol_mode=nh.TOL_RELATIVE) as s:
            t0 = time.perf_counter()
            try:
                s.run(1, 2, 1, 20)
            except nh.NrtError:
                fails += 1
            s.sync()
            plain.append((p.N * 2, time.perf_counter() - t0, s.stats()["krylov_iters"] / 2))
    leg["fp64_plain_bicgstab"] = {
        "compartment_steps_per_s": sum(x[0] for x in plain) / sum(x[1] for x in plain),
        "iterations_per_step": [x[2] for x in plain], "not_converged_runs": fails,
        "sample": "the first 4 requests of the stream, 2 steps each, alone"}
    if orc is not None and not a.no_cpu_baseline:
        th = info["threads_used"]
        jobs = [(p, 1, 1, 20) for p, _, _ in reqs[:48] if p.N <= 60000][:32]
        cs, wall, bad = cpu_pool(orc, jobs, th, "gmres")
        leg["cpu_baseline"] = {
            "value": cs / wall, "unit": "compartment-steps/s", "cores": th, "kind": "port",
            "sample": f"the first step of each of the first {len(jobs)} requests of at most "
                      f"60,000 compartments, one request per thread on {th} threads, reference "
                      "CPU method (GMRES(150), x0 = 0, 1e-14 A)" +
                      (f"; {bad} steps not converged" if bad else ""),
            "wall_s": wall}
    return leg


# ------------------------------------------------------------ configs 1..3
def run_solver(nh, prob, kind, opts, steps, a0, a1):
    with nh.Solver(prob, kind=kind, device=0, **opts) as s:
        s.sync()
        st0 = s.stats()
        t0 = time.perf_counter()
        s.run(1, steps, a0, a1)
        s.sync()
        wall = time.perf_counter() - t0
        st1 = s.stats()
        v = s.snapshot()
    return v, wall, st1, st0


def config1_leg(nh, wl, orc, a):
    """Config 1: the 1,000-compartment squid axon, explicit FP64, 10 ms."""
    geom, run = wl.squid_axon()
    _, crit = nh.setup(geom, dT=run.dT)
    leg = {"workload": "unbranched HH cable 1 cm, d 1 um, dx 10 um (1,000 compartments + 2 "
                       "ghosts), 10 nA for 0.5 ms at compartment 1, 10 ms",
           "critical_dT_s": crit}
    say("config 1")
    for label, dT in (("dt_1us_literal", run.dT), ("dt_0.8_critical", 0.8 * crit)):
        prob, _ = nh.setup(geom, dT=dT)
        steps = run.steps(dT)
        a0, a1 = nh.active_range(dT, steps, run.init_time, run.end_input)
        v, wall, st, _ = run_solver(nh, prob, nh.EXPLICIT, {}, steps, a0, a1)
        rec = {"dT_s": dT, "steps": steps, "compartment_steps_per_s": prob.N * steps / wall,
               "ms_per_step": wall / steps * 1e3, "finite": bool(np.all(np.isfinite(v))),
               "max_abs_V_mV": float(np.max(np.abs(v))) * 1e3}
        if st["kernel_ms"] > 0:
            nb = explicit_bytes(prob.N, prob.H) * steps
            launches = st["kernel_launches"]
            kern = "k_explicit_resident" if launches == 1 else "k_explicit_fp64"
            rec["roofline"] = roofline(nb, st["kernel_ms"], kern, launches)
        if orc is not None:
            ex = orc.Explicit(prob)
            t0 = time.perf_counter()
            ex.run(1, steps, a0, a1)
            cpu = time.perf_counter() - t0
            rv = ex.snapshot()
            rec["max_abs_dV_mV_vs_oracle"] = (err_mv(v, rv, prob) if np.all(np.isfinite(rv))
                                              else None)
            rec["identical_to_oracle"] = bool(np.array_equal(v, rv, equal_nan=True))
            if not a.no_cpu_baseline:
                rec["cpu_baseline"] = {"value": prob.N * steps / cpu, "cores": 1, "kind": "port",
                                       "sample": "the full run, oracle explicit step (solverEx-"
                                                 "cpu.cpp semantics), 1 thread"}
        leg[label] = rec
    return leg


def implicit_forms(nh, prob, forms, a0, a1, orc):
    out = {}
    refs = {}
    for name, (steps, opts) in forms.items():
        say(f"N={prob.N}: {name}")
        try:
            v, wall, st, st0 = run_solver(nh, prob, nh.IMPLICIT, opts, steps, a0, a1)
        except nh.NrtError as e:
            out[name] = {"steps": steps, "error": str(e)}
            continue
        inner = st["krylov_iters"] - st0["krylov_iters"]
        rec = {"compartment_steps_per_s": prob.N * steps / wall, "ms_per_step": wall / steps * 1e3,
               "steps": steps, "inner_per_step": inner / steps, "workgroups": st["num_blocks"]}
        kr = opts.get("krylov")
        vb = 8 if opts.get("precision", nh.FP64) == nh.FP64 else 4
        if kr == nh.DIRECT:
            nb = tree_bytes(prob.N, prob.H, st["matrix_nnz"], steps, inner, vb)
            kern = nh.TREE_KERNELS[st["tree_form"]].format(T="double" if vb == 8 else "float")
        elif kr == nh.GMRES:
            cycles = st["refine_sweeps"] if vb == 4 else st["restarts"] + steps
            nb = gmres_bytes(prob.N, prob.H, st["matrix_nnz"], steps, inner, cycles, vb)
            kern = f"k_implicit_gmres<{'double' if vb == 8 else 'float'}>"
            rec["cycles_per_step"] = cycles / steps
        else:
            nb = bicgstab_bytes(prob.N, prob.H, st["matrix_nnz"], steps, inner)
            kern = "k_implicit_bicgstab_fp64"
        rec["roofline"] = roofline(nb, st["kernel_ms"], kern, 1)
        if orc is not None:
            if steps not in refs:
                refs[steps] = oracle_exact(orc, prob, steps, a0, a1)
            rec["max_abs_dV_mV_vs_oracle"] = err_mv(v, refs[steps], prob)
        out[name] = rec
    return out


def config2_leg(nh, wl, orc, a):
    geom, run = wl.myelinated_axon()
    prob, _ = nh.setup(geom, dT=run.dT)
    a0, a1 = nh.active_range(run.dT, run.steps(run.dT), run.init_time, run.end_input)
    n = a.config2_steps
    forms = {"fp64_tree_direct": (n, dict(krylov=nh.DIRECT)),
             "mixed_tree_ir": (n, dict(krylov=nh.DIRECT, precision=nh.FP32_REFINE,
                                       refine_tol=1e-10)),
             "fp64_plain_bicgstab": (a.plain_steps, dict(krylov=nh.BICGSTAB, tol=1e-10,
                                                         tol_mode=nh.TOL_RELATIVE))}
    leg = {"workload": "myelinated axon, 200 HH nodes (1 um) + 200 x 1 mm passive internodes "
                       "(100 x 10 um), d 10 um, N 20,202, dt 10 us, stimulus 10 nA x 0.5 ms",
           "N": prob.N, "hh_elements": prob.H}
    leg.update(implicit_forms(nh, prob, forms, a0, a1, orc))
    if orc is not None and not a.no_cpu_baseline:
        im = orc.Implicit(prob, variant=orc.VAR_CPU, method=orc.M_GMRES, restart=150, tol=1e-14,
                          tol_mode=0)
        t0 = time.perf_counter()
        k = 0
        while k < 3 and time.perf_counter() - t0 < 15.0:
            k += 1
            im.update(a0 <= k <= a1)
            im.calculate()
        leg["cpu_baseline"] = {"value": prob.N * k / (time.perf_counter() - t0), "cores": 1,
                               "kind": "port", "sample": f"the first {k} steps, reference CPU "
                               "method (GMRES(150), x0 = 0, 1e-14 A), 1 thread"}
        cs, wall, _ = cpu_pool(orc, [(prob, 200, a0, a1)], 1, "direct")
        leg["cpu_exact_elimination"] = {"value": cs / wall, "cores": 1,
                                        "sample": "200 steps, exact elimination, 1 thread"}
    return leg


def config3_leg(nh, wl, orc, a):
    geom, run = wl.random_tree(wl.seed_for(3), depth=14)
    prob, _ = nh.setup(geom, dT=run.dT)
    a0, a1 = nh.active_range(run.dT, run.steps(run.dT), run.init_time, run.end_input)
    n, kn = a.config3_steps, a.krylov3_steps
    gm = dict(krylov=nh.GMRES, restart=30, max_iter=a.krylov3_max_iter)
    forms = {"fp64_tree_direct": (n, dict(krylov=nh.DIRECT)),
             "mixed_tree_ir": (n, dict(krylov=nh.DIRECT, precision=nh.FP32_REFINE,
                                       refine_tol=1e-10)),
             "fp64_gmres30": (kn, dict(gm, tol=1e-10, tol_mode=nh.TOL_RELATIVE)),
             "mixed_gmres30_ir": (kn, dict(gm, precision=nh.FP32_REFINE, refine_tol=1e-12))}
    leg = {"workload": "seeded binary dendritic tree, depth 14 (16,383 segments of U[20,200] um), "
                       "Rall 3/2 diameters from 4 um, dx 2 um, strain U[0,0.3] per segment, HH "
                       "soma, dt 25 us, stimulus 1 nA x 0.5 ms at the soma",
           "N": prob.N, "hh_elements": prob.H}
    leg.update(implicit_forms(nh, prob, forms, a0, a1, orc))
    if orc is not None and not a.no_cpu_baseline:
        # the reference CPU method cannot finish one step of this tree in the
        # budget: time a bounded number of its Arnoldi iterations, which bounds its
        # rate from above (a step needs more than that many)
        it = a.cpu3_iters
        im = orc.Implicit(prob, variant=orc.VAR_CPU, method=orc.M_GMRES, restart=150, tol=1e-14,
                          tol_mode=0, max_iter=it)
        t0 = time.perf_counter()
        im.update(True)
        st = im.calculate()
        t = time.perf_counter() - t0
        leg["cpu_baseline"] = {"value_upper_bound": prob.N / t, "cores": 1, "kind": "port",
                               "sample": f"{it} Arnoldi iterations of step 1 with the reference "
                                         f"CPU method (GMRES(150), x0 = 0, 1e-14 A), {t:.1f} s, "
                                         f"step {'converged' if st == 0 else 'not converged'}: "
                                         "the rate is below N / t", "converged": st == 0}
        cs, wall, _ = cpu_pool(orc, [(prob, 5, a0, a1)], 1, "direct")
        leg["cpu_exact_elimination"] = {"value": cs / wall, "cores": 1,
                                        "sample": "5 steps, exact elimination, 1 thread"}
    return leg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps of the headline sweep")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--legs", default="sweep,served,config1,config2,config3")
    ap.add_argument("--sweep-trees", type=int, default=4096)
    ap.add_argument("--sweep-forms", default="implicit_fp64,implicit_mixed,explicit")
    ap.add_argument("--shapes", default=None,
                    help="JSON file of launch shapes per sweep form: used when it holds "
                         "the form, else the tuned shapes are written to it")
    ap.add_argument("--tune-rows", type=int, default=4_000_000,
                    help="compartments per size class the sweep's launch-shape tuner times "
                         "(0: no tuning)")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--config2-steps", type=int, default=1000)
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--config3-steps", type=int, default=200)
    ap.add_argument("--krylov3-steps", type=int, default=3)
    ap.add_argument("--krylov3-max-iter", type=int, default=4000)
    ap.add_argument("--cpu3-iters", type=int, default=60)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle", action="store_true", help="skip every oracle check")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): the solvers have "
              "no multi-GPU path (each neurite or batch runs on one device)", file=sys.stderr)
        sys.exit(2)
    import neurite_hip as nh
    import workloads as wl

    legs = set(a.legs.split(","))
    info = host_info()
    orc = None if a.no_oracle else _oracle()
    out = {"metric": METRIC, "value": None, "unit": "compartment-steps/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": None, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": None,
           "data": "synthetic (seeded trees; no datasets)"}
    res = {"host": info}
    if "sweep" in legs:
        res["sweep"] = sweep_leg(nh, wl, orc, a, info)
    if "served" in legs:
        res["served"] = served_leg(nh, wl, orc, a, info)
    if "config1" in legs:
        res["config1"] = config1_leg(nh, wl, orc, a)
    if "config2" in legs:
        res["config2"] = config2_leg(nh, wl, orc, a)
    if "config3" in legs:
        res["config3"] = config3_leg(nh, wl, orc, a)
    if "sweep" in legs:
        sw = res["sweep"]
        cand = {k: sw[k] for k in ("implicit_fp64", "implicit_mixed")
                if k in sw and sw[k].get("status", 0) == 0 and
                (orc is None or sw[k].get("max_abs_dV_mV_vs_oracle", 1e9) <= TOL_MV)}
        best = max(cand, key=lambda k: cand[k]["compartment_steps_per_s"]) if cand else None
    if "sweep" in legs and best is not None:
        b = sw[best]
        out.update({
            "value": b["compartment_steps_per_s"], "ms_per_step": b["ms_per_step"],
            "dtype": "f64" if best == "implicit_fp64" else "f32+f64-refine",
            "config": {"workload": f"config 4 batched sweep: {sw['workload']}",
                       "trees": sw["trees"], "compartments": sw["compartments"], "form": best,
                       "max_abs_dV_mV_vs_oracle": b.get("max_abs_dV_mV_vs_oracle"),
                       "oracle_sample": sw["oracle_sample"]},
            "roofline": b["roofline"],
            "cpu_baseline": sw.get("cpu_baseline")})
    if "served" in legs:
        sv = res["served"]
        okf = [k for k in ("fp64", "mixed") if sv[k]["failed"] == 0 and
               (orc is None or sv[k].get("max_abs_dV_mV_vs_oracle", 1e9) <= TOL_MV)]
        if okf:
            bf = max(okf, key=lambda k: sv[k]["compartment_steps_per_s"])
            out["p50_ms"] = sv[bf]["p50_ms"]
            out["p99_ms"] = sv[bf]["p99_ms"]
            out["served_form"] = bf
            out["served_compartment_steps_per_s"] = sv[bf]["compartment_steps_per_s"]
            out["speedup_served_mixed_vs_plain_fp64_bicgstab"] = (
                sv["mixed"]["compartment_steps_per_s"] /
                sv["fp64_plain_bicgstab"]["compartment_steps_per_s"])
    out.update(res)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
