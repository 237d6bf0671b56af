#!/usr/bin/env python3
"""Benchmark of the implicit cable-equation step on one MI355X.

Workload (the configuration the headline metric is quoted on): a myelinated
axon of 200 HH nodes of Ranvier (1 um) alternating with 1 mm passive internodes
(100 x 10 um, Cm = 0.005 uF/cm^2, gL = 0.001 mS/cm^2), d = 10 um, N = 20,202
elements, backward Euler at dt = 10 us, stimulus 10 nA for 0.5 ms at the first
node. A "step" is one time step: HH gating, matrix refresh, right-hand side and
the solve, all inside one persistent kernel launch per run; inputs are resident
on the device.

Forms measured on the same workload:
  fp64_tree_direct     exact tree partition solve in FP64
  mixed_tree_ir        FP32 tree solves inside FP64 iterative refinement
                       (residual b - A x and x += d in FP64, Jacobi-scaled
                       relative residual 1e-10)
  fp64_plain_bicgstab  the reference method: unpreconditioned BiCGSTAB,
                       relative residual 1e-10 (few steps: it is slow)
`value` is the fastest form whose max |dV| against the FP64 exact solve after
the same steps is <= 1e-4 mV. Prints ONE JSON line.
"""
import argparse
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


def bicgstab_bytes(N, H, nnz, steps, iters):
    """Per step 144 B per HH element + 40 B per element (update phase); per
    BiCGSTAB iteration 2 SpMV (12 nnz + 4 (N+1) + 16 N) + 21 vector passes of 8 N."""
    spmv = 12 * nnz + 4 * (N + 1) + 16 * N
    return steps * (144 * H + 40 * N) + iters * (2 * spmv + 21 * 8 * N)


def tree_bytes(N, H, nnz, steps, solves, vbytes):
    """Algorithmic bytes of the tree-solved steps (k_implicit_tree, k_batch_tree),
    counting every chunk as one that reuses its factors (chunks holding an HH row
    are refactored each step and move more; they are a small share):
    update phase per step: 144 B per HH element (V, gates r/w, parameters, the
      refreshed diagonal) and per element V 8 + code 4 + c_m dx 8 + K dx 8 +
      input 8 + the right-hand side written (FP64 b for the refinement: 8; into the
      tree system: vbytes);
    per tree solve: the forward sweep reads r and the multiplier and writes the
      partial right-hand side, the backward sweep reads it with the pivot
      reciprocal and the coupling and writes alpha, the back-substitution reads
      alpha, beta, gamma (10 values) and writes V (8; the refinement also reads V);
    refinement only, per residual evaluation (solves + steps of them): the matrix
      (12 nnz + 4 (N+1)), the gathered V (8 nnz), b (8 N), the residual written
      into the tree system (vbytes N)."""
    if vbytes == 8:
        per_step = 144 * H + (8 + 4 + 8 + 8 + 8 + 8) * N
        per_solve = (10 * 8 + 8) * N
        return steps * per_step + solves * per_solve
    per_step = 144 * H + (8 + 4 + 8 + 8 + 8 + 8) * N
    per_solve = (10 * vbytes + 16) * N
    per_resid = 12 * nnz + 4 * (N + 1) + 8 * nnz + 8 * N + vbytes * N
    return steps * per_step + solves * per_solve + (solves + steps) * per_resid


def tree_kernel(nh, name, f):
    """Name of the kernel a form's launches ran (as rocprofv3 lists it)."""
    fixed = {"fp64_plain_bicgstab": "k_implicit_bicgstab_fp64",
             "fp64_gmres30": "k_implicit_gmres<double>", "mixed_gmres30_ir": "k_implicit_gmres<float>"}
    if name in fixed:
        return fixed[name]
    return nh.TREE_KERNELS[f["tree_form"]].format(T="double" if name.startswith("fp64") else "float")


def run_form(nh, prob, opts, first, steps, a0, a1):
    s = nh.Solver(prob, kind=nh.IMPLICIT, device=0, **opts)
    if first > 1:
        s.run(1, first - 1, a0, a1)
    s.sync()
    st0 = s.stats()
    t0 = time.perf_counter()
    s.run(first, steps, a0, a1)
    s.sync()
    t1 = time.perf_counter()
    st1 = s.stats()
    v = s.snapshot()
    s.close()
    return {"wall": t1 - t0, "kernel_ms": st1["kernel_ms"],
            "inner": st1["krylov_iters"] - st0["krylov_iters"], "nnz": st1["matrix_nnz"],
            "blocks": st1["num_blocks"], "v": v, "tree_form": st1["tree_form"],
            # GMRES restart cycles (FP32 forms: the refinement sweeps; FP64: the
            # restarts plus one per step)
            "cycles": (st1["refine_sweeps"] - st0["refine_sweeps"]) if st1["refine_sweeps"]
            else (st1["restarts"] - st0["restarts"]) + steps}


def gmres_bytes(N, H, nnz, steps, iters, cycles, vbytes):
    """Algorithmic bytes of GMRES(m) steps (k_implicit_gmres<double|float>), with
    the cycles of a run taken as equally long (iters / cycles Arnoldi steps):
    update phase per step: 144 B per HH element + 40 B per element;
    per cycle start: the FP64 residual (matrix 12 nnz + 4 (N+1), V, b, r0 written:
      24 N) and at its end x += V y (kk basis vectors of vbytes N, x read and
      written 16 N);
    per Arnoldi step k: the SpMV (values vbytes nnz + columns 4 nnz + rowptr
      4 (N+1), the previous vector read, v_k and w written: 3 vbytes N) and CGS2
      (two projections and two corrections each read v_0..v_k and w, the
      corrections write w: (4 (k+1) + 6) vbytes N)."""
    if cycles <= 0:
        return steps * (144 * H + 40 * N)
    kk = iters / cycles
    per_cycle = (12 * nnz + 4 * (N + 1) + 24 * N) + kk * vbytes * N + 16 * N
    spmv = vbytes * nnz + 4 * nnz + 4 * (N + 1) + 3 * vbytes * N
    cgs = (2 * kk * (kk + 1) + 6 * kk) * vbytes * N
    return steps * (144 * H + 40 * N) + cycles * (per_cycle + kk * spmv + cgs)


def explicit_bytes(N, H):
    """48 B per passive element (V r/w, a, r_a dx, W dx, K dx) + 4 B code; HH
    elements 152 B (V r/w, a, r_a dx, m/h/n r/w, 9 parameters)."""
    return 52 * (N - H) + 156 * H


def batched_sweep(nh, wl, count, steps):
    """Config 4: `count` seeded trees (sizes log-uniform 500..50,000), one batch per
    form. Returns the per-form throughput and the accuracy of the mixed form."""
    t0 = time.perf_counter()
    geoms = [wl.sweep_tree(i) for i in range(count)]
    imp = [nh.setup(g, dT=r.dT)[0] for g, r in geoms]
    exp_ = [nh.setup(g, dT=None, dt_factor=0.8)[0] for g, _ in geoms]
    gen_s = time.perf_counter() - t0
    Ntot = sum(p.N for p in imp)
    H = sum(p.H for p in imp)
    out = {"workload": f"{count} seeded trees, compartments log-uniform 500..50000, depth "
                       "U{4..12} (capped at >= 2 compartments per segment), per-tree strain "
                       "U[0,0.3] and stimulus U[1,20] nA at the soma for 0.5 ms; implicit "
                       "dt = 25 us, explicit dt = 0.8 x each tree's critical step",
           "trees": count, "compartments": Ntot, "hh_elements": H, "setup_s": gen_s}
    sample = list(range(0, count, max(1, count // 16)))
    snaps = {}
    for name, kw, probs in [
            ("implicit_fp64", dict(kind=nh.IMPLICIT, krylov=nh.DIRECT), imp),
            ("implicit_mixed", dict(kind=nh.IMPLICIT, krylov=nh.DIRECT,
                                    precision=nh.FP32_REFINE, refine_tol=1e-10), imp),
            ("explicit", dict(kind=nh.EXPLICIT), exp_)]:
        b = nh.Batch(**kw)
        t0 = time.perf_counter()
        ids = [b.add(p, steps + 8, 1, 20) for p in probs]
        add_s = time.perf_counter() - t0
        b.step(4)  # warm-up (and the workgroup-size autotuner's first launches)
        b.step(4)
        st0 = b.stats()
        t0 = time.perf_counter()
        for _ in range(steps // 10):
            b.step(10)
        wall = time.perf_counter() - t0
        st1 = b.stats()
        cs = st1["compartment_steps"] - st0["compartment_steps"]
        kms = st1["kernel_ms"] - st0["kernel_ms"]
        rec = {"compartment_steps_per_s": cs / wall, "ms_per_step": wall / (steps // 10 * 10) * 1e3,
               "kernel_ms": kms, "add_s": add_s, "device_bytes": st1["device_bytes"],
               "workgroup_threads": st1["threads"] if name != "explicit" else 256}
        timed = steps // 10 * 10
        if name == "explicit":
            nb = explicit_bytes(Ntot, H) * timed
            kern = "k_explicit_batch"
        else:
            snaps[name] = [b.snapshot(ids[i]) for i in sample]
            # solves in the timed steps: scaled from each slot's total (the warm-up
            # steps take as many solves per step as the timed ones, within a sweep)
            nb = 0
            vb = 8 if name == "implicit_fp64" else 4
            for sid, p in zip(ids, probs):
                inf = b.info(sid)
                solves = inf["inner_iters"] * timed / max(1, inf["steps_done"])
                nb += tree_bytes(p.N, p.H, inf["nnz"], timed, solves, vb)
            kern = f"k_batch_tree<{'double' if vb == 8 else 'float'},{st1['threads']}>"
        ach = nb / (kms / 1e3) / 1e9
        rec["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "kernel": kern,
                           "algorithmic_bytes": nb}
        out[name] = rec
        b.close()
    err = 0.0
    for i, (vm, vd) in enumerate(zip(snaps["implicit_mixed"], snaps["implicit_fp64"])):
        m = np.ones(len(vd), bool)
        m[[0, len(vd) - 1]] = False
        err = max(err, float(np.max(np.abs(vm - vd)[m])) * 1e3)
    out["implicit_mixed"]["max_abs_dV_mV_vs_fp64"] = err
    out["implicit_mixed"]["compared_trees"] = len(sample)
    return out


def served_requests(nh, wl, count):
    reqs = []
    for i in range(count):
        g, r, gap = wl.served_request(i)
        p = nh.setup(g, dT=r.dT)[0]
        reqs.append((p, r.steps(r.dT), gap))
    return reqs


def served_stream(nh, reqs, precision):
    """Config 5: requests with Poisson arrivals (200/s), sizes log-normal (median
    20,000, sigma 1, cap 2M), 5..50 ms simulated at dt = 25 us, submitted in real
    time to the request server."""
    count = len(reqs)
    kw = dict(kind=nh.IMPLICIT, krylov=nh.DIRECT, precision=precision, refine_tol=1e-10)
    srv = nh.Server(steps_per_launch=16, **kw)
    t0 = time.perf_counter()
    due = 0.0
    for p, n, gap in reqs:
        due += gap
        wait = t0 + due - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        srv.submit(p, n, 1, 20)
    srv.wait_idle(600.0)
    wall = time.perf_counter() - t0
    st = srv.stats()
    res = srv.poll()
    srv.close()
    return {"requests": count, "completed": st["completed"], "failed": st["failed"],
            "arrival_window_s": due, "wall_s": wall,
            "compartment_steps_per_s": st["compartment_steps"] / wall,
            "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
            "largest_request": max(p.N for p, _, _ in reqs),
            "launches": st["launches"], "results": len(res),
            "loop_ms": {k: st[k] for k in ("launch_ms", "admit_ms", "wait_ms", "retire_ms",
                                           "batch_kernel_ms")}}


def dendritic_tree(nh, wl, tree_steps, krylov_steps, max_iter):
    """Config 3: the seeded depth-14 dendritic tree (16,383 segments, ~1.8M
    compartments, dx = 2 um, per-segment strain U[0, 0.3], HH soma, dt = 25 us).
    GMRES(30) in FP64 (relative 1e-10) and in FP32 inside FP64 refinement (scaled
    relative 1e-12) for `krylov_steps` steps (each step takes hundreds of Arnoldi
    steps), the tree solves for `tree_steps`; every form's error is taken against
    the FP64 tree solve after the same steps."""
    t0 = time.perf_counter()
    geom, run = wl.random_tree(wl.seed_for(3), depth=14)
    prob, _ = nh.setup(geom, dT=run.dT)
    setup_s = time.perf_counter() - t0
    N, H = prob.N, prob.H
    a0, a1 = nh.active_range(run.dT, run.steps(run.dT), run.init_time, run.end_input)
    gm = dict(krylov=nh.GMRES, restart=30, max_iter=max_iter)
    specs = {
        "fp64_tree_direct": (tree_steps, dict(krylov=nh.DIRECT, precision=nh.FP64)),
        "mixed_tree_ir": (tree_steps, dict(krylov=nh.DIRECT, precision=nh.FP32_REFINE,
                                           refine_tol=1e-10)),
        "fp64_gmres30": (krylov_steps, dict(gm, precision=nh.FP64, tol=1e-10,
                                            tol_mode=nh.TOL_RELATIVE)),
        # the scaled 2-norm over ~900k rows: 1e-12 keeps every element within 1e-4 mV
        "mixed_gmres30_ir": (krylov_steps, dict(gm, precision=nh.FP32_REFINE, refine_tol=1e-12)),
    }
    refs = {}
    for st in {tree_steps, krylov_steps} - {0}:
        refs[st] = run_form(nh, prob, dict(krylov=nh.DIRECT, precision=nh.FP64), 1, st, a0,
                            a1)["v"]
    mask = np.ones(N, bool)
    mask[[0, N - 1]] = False
    kernels = {"fp64_gmres30": "k_implicit_gmres<double>",
               "mixed_gmres30_ir": "k_implicit_gmres<float>"}
    out = {"workload": "seeded binary dendritic tree, depth 14 (16,383 segments of "
                       "U[20,200] um), Rall 3/2 diameters from 4 um, dx 2 um, strain U[0,0.3] "
                       "per segment, HH soma, dt 25 us, stimulus 1 nA x 0.5 ms at the soma",
           "N": N, "hh_elements": H, "setup_s": setup_s}
    for name, (steps, opts) in specs.items():
        if steps <= 0:
            continue
        try:
            f = run_form(nh, prob, opts, 1, steps, a0, a1)
        except nh.NrtError as e:
            out[name] = {"steps": steps, "error": str(e)}
            continue
        vb = 8 if name.startswith("fp64") else 4
        if "gmres" in name:
            nbytes = gmres_bytes(N, H, f["nnz"], steps, f["inner"], f["cycles"], vb)
        else:
            nbytes = tree_bytes(N, H, f["nnz"], steps, f["inner"], vb)
        achieved = nbytes / (f["kernel_ms"] / 1e3) / 1e9
        out[name] = {"compartment_steps_per_s": N * steps / f["wall"],
                     "ms_per_step": f["wall"] / steps * 1e3, "steps": steps,
                     "max_abs_dV_mV_vs_fp64": float(np.max(np.abs(f["v"] - refs[steps])[mask]))
                     * 1e3,
                     "inner_per_step": f["inner"] / steps,
                     "cycles_per_step": f["cycles"] / steps, "workgroups": f["blocks"],
                     "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                                  "kernel": tree_kernel(nh, name, f), "kernel_ms": f["kernel_ms"],
                                  "algorithmic_bytes": nbytes}}
        if "gmres" not in name:
            del out[name]["cycles_per_step"]
    return out


def cpu_baseline(prob, budget_s=15.0, max_steps=3):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle  # checker library: timed here as the CPU baseline only
    im = oracle.Implicit(prob, variant=oracle.VAR_CPU, method=oracle.M_GMRES, restart=150,
                         tol=1e-14, tol_mode=0)
    t0 = time.perf_counter()
    n = 0
    while n < max_steps and time.perf_counter() - t0 < budget_s:
        n += 1
        im.update(True)
        im.calculate()
    dt = time.perf_counter() - t0
    out = {"value": prob.N * n / dt, "unit": "compartment-steps/s", "cores": 1, "kind": "port",
           "sample": f"first {n} steps of the same workload (N={prob.N}) with the reference "
                     "CPU method: restarted GMRES(150), x0=0, ||r||<=1e-14 A "
                     "(solverIm-cpu.cpp semantics), oracle C++ build, 1 thread"}
    ex = oracle.Implicit(prob, variant=oracle.VAR_GPU, method=oracle.M_DIRECT)
    t0 = time.perf_counter()
    n = 0
    while n < 2000 and time.perf_counter() - t0 < 5.0:
        n += 1
        ex.update(n <= 50)
        ex.calculate()
    out["exact_tree_elimination_1core"] = prob.N * n / (time.perf_counter() - t0)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep-trees", type=int, default=4096)
    ap.add_argument("--sweep-steps", type=int, default=100)
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--tree3-steps", type=int, default=200,
                    help="depth-14 dendritic tree: tree-solve steps (0 skips the leg)")
    ap.add_argument("--krylov3-steps", type=int, default=4,
                    help="depth-14 dendritic tree: GMRES(30) steps")
    ap.add_argument("--krylov3-max-iter", type=int, default=4000,
                    help="depth-14 dendritic tree: Arnoldi steps allowed per time step")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the solvers have "
              "no multi-GPU path (each neurite or batch runs on one device)", file=sys.stderr)
        sys.exit(2)

    import neurite_hip as nh
    import workloads as wl

    geom, run = wl.myelinated_axon()
    prob, _ = nh.setup(geom, dT=run.dT)
    N, H = prob.N, prob.H
    a0, a1 = nh.active_range(run.dT, run.steps(run.dT), run.init_time, run.end_input)
    first = a.warmup + 1
    forms = {}
    ref = run_form(nh, prob, dict(krylov=nh.DIRECT, precision=nh.FP64), first, a.steps, a0, a1)
    forms["fp64_tree_direct"] = ref
    forms["mixed_tree_ir"] = run_form(nh, prob, dict(krylov=nh.DIRECT, precision=nh.FP32_REFINE,
                                                     refine_tol=1e-10), first, a.steps, a0, a1)
    if not a.no_plain:
        plain = run_form(nh, prob, dict(krylov=nh.BICGSTAB, precision=nh.FP64, tol=1e-10,
                                        tol_mode=nh.TOL_RELATIVE), first, a.plain_steps, a0, a1)
        pref = run_form(nh, prob, dict(krylov=nh.DIRECT, precision=nh.FP64), first,
                        a.plain_steps, a0, a1)
        plain["ref_v"] = pref["v"]
        forms["fp64_plain_bicgstab"] = plain
    mask = np.ones(N, bool)
    mask[[0, N - 1]] = False
    summary = {}
    for name, f in forms.items():
        steps = a.plain_steps if name == "fp64_plain_bicgstab" else a.steps
        rv = f.get("ref_v", ref["v"])
        err_mv = float(np.max(np.abs(f["v"] - rv)[mask])) * 1e3
        if name == "fp64_plain_bicgstab":
            nbytes = bicgstab_bytes(N, H, f["nnz"], steps, f["inner"])
        else:
            nbytes = tree_bytes(N, H, f["nnz"], steps, f["inner"], 8 if name.startswith("fp64")
                                else 4)
        summary[name] = {"compartment_steps_per_s": N * steps / f["wall"],
                         "ms_per_step": f["wall"] / steps * 1e3, "steps": steps,
                         "max_abs_dV_mV_vs_fp64": err_mv, "inner_per_step": f["inner"] / steps,
                         "kernel_ms": f["kernel_ms"], "algorithmic_bytes": nbytes,
                         "workgroups": f["blocks"]}
    ok = {k: v for k, v in summary.items() if v["max_abs_dV_mV_vs_fp64"] <= TOL_MV}
    best = max(ok, key=lambda k: ok[k]["compartment_steps_per_s"])
    b = summary[best]
    kernel = tree_kernel(nh, best, forms[best])
    achieved = b["algorithmic_bytes"] / (b["kernel_ms"] / 1e3) / 1e9
    # HBM bytes measured with rocprofv3 counters for this kernel and workload (per step,
    # calibrated with a copy), scaled to this launch's steps
    traffic = None
    tfile = os.path.join(REPO, "profiles", "traffic.json")
    tkey = f"{kernel} myelinated axon N={N}"
    if os.path.exists(tfile):
        t = json.load(open(tfile)).get(tkey)
        if t:
            traffic = t["hbm_bytes_per_step"] * b["steps"]
    out = {
        "metric": METRIC,
        "value": b["compartment_steps_per_s"],
        "unit": "compartment-steps/s",
        "n_gpus": 1,
        "steps": b["steps"],
        "warmup": a.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best.startswith("fp64") else "f32+f64-refine",
        "data": "synthetic",
        "config": {"workload": "myelinated axon, 200 HH nodes (1 um) + 200 x 1 mm passive "
                               "internodes (100 x 10 um), d=10 um, N=20202, backward Euler "
                               "dt=10 us, stimulus 10 nA x 0.5 ms",
                   "N": N, "hh_elements": H, "nnz": forms[best]["nnz"], "form": best,
                   "tolerance": "relative 1e-10 (refinement: Jacobi-scaled; BiCGSTAB: plain)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
                     "launches": 1, "kernel_ms": b["kernel_ms"],
                     "algorithmic_bytes": b["algorithmic_bytes"]},
        "forms": summary,
    }
    if "fp64_plain_bicgstab" in summary:
        out["speedup_vs_fp64_plain"] = (b["compartment_steps_per_s"] /
                                        summary["fp64_plain_bicgstab"]["compartment_steps_per_s"])
    if a.tree3_steps > 0:
        out["dendritic_tree"] = dendritic_tree(nh, wl, a.tree3_steps, a.krylov3_steps,
                                               a.krylov3_max_iter)
    if a.sweep_trees > 0:
        out["batched_sweep"] = batched_sweep(nh, wl, a.sweep_trees, a.sweep_steps)
    if a.serve_requests > 0:
        reqs = served_requests(nh, wl, a.serve_requests)
        out["served"] = {"workload": "Poisson arrivals 200/s, compartments log-normal "
                                     "(median 20000, sigma 1, cap 2M), 5..50 ms at dt 25 us",
                         "fp64": served_stream(nh, reqs, nh.FP64),
                         "mixed": served_stream(nh, reqs, nh.FP32_REFINE)}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(prob)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
