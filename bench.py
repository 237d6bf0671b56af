#!/usr/bin/env python3
"""Benchmark of the implicit cable-equation step on one MI355X.

Workload (the configuration the headline metric is quoted on): a myelinated
axon of 200 HH nodes of Ranvier (1 um) alternating with 1 mm passive internodes
(100 x 10 um, Cm = 0.005 uF/cm^2, gL = 0.001 mS/cm^2), d = 10 um, N = 20,202
elements, backward Euler at dt = 10 us, Krylov relative residual 1e-10,
stimulus 10 nA for 0.5 ms at the first node. A "step" is one time step: HH
gating, matrix refresh, right-hand side and the full Krylov solve, all inside
the library's persistent kernel; inputs are resident on the device.

Prints ONE JSON line. `roofline` uses algorithmic bytes (per element and per
Krylov iteration, formulas below) over the kernel's device time measured with
HIP events on the library's stream. `cpu_baseline` times the reference CPU
method (restarted GMRES(150), x0 = 0, ||r|| <= 1e-14 A) on this host's cores
for a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "168861_amd")]

METRIC = "compartment-steps/s at max|ΔV|≤1e-4 mV vs FP64; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0


def implicit_bytes_per_run(N, H, nnz, steps, iters):
    """Algorithmic HBM bytes of `steps` implicit FP64 steps with `iters` BiCGSTAB
    iterations in total: per step 144 B per HH element (V, gates r/w,
    parameters, diagonal) + 40 B per element (right-hand side); per iteration
    2 SpMV (12 nnz + 4 (N+1) + 16 N bytes each) + 21 vector passes of 8 N."""
    spmv = 12 * nnz + 4 * (N + 1) + 16 * N
    return steps * (144 * H + 40 * N) + iters * (2 * spmv + 21 * 8 * N)


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
    return {"value": prob.N * n / dt, "unit": "compartment-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"first {n} steps of the same workload (N={prob.N}) with the reference "
                      "CPU method: restarted GMRES(150), x0=0, ||r||<=1e-14 A "
                      "(solverIm-cpu.cpp semantics), oracle C++ build, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU "
              "decomposition is described in DESIGN.md and not built", file=sys.stderr)
        sys.exit(2)

    import neurite_hip as nh
    import workloads as wl

    geom, run = wl.myelinated_axon()
    prob, _ = nh.setup(geom, dT=run.dT)
    a0, a1 = nh.active_range(run.dT, run.steps(run.dT), run.init_time, run.end_input)
    s = nh.Solver(prob, kind=nh.IMPLICIT, precision=nh.FP64, krylov=nh.BICGSTAB, tol=a.tol,
                  tol_mode=nh.TOL_RELATIVE, device=0)
    if a.warmup > 0:
        s.run(1, a.warmup, a0, a1)
    s.sync()
    st0 = s.stats()
    t0 = time.perf_counter()
    s.run(a.warmup + 1, a.steps, a0, a1)
    s.sync()
    t1 = time.perf_counter()
    st1 = s.stats()
    wall = t1 - t0
    iters = st1["krylov_iters"] - st0["krylov_iters"]
    N, H, nnz = prob.N, prob.H, st1["matrix_nnz"]
    kernel_s = st1["kernel_ms"] / 1e3
    nbytes = implicit_bytes_per_run(N, H, nnz, a.steps, iters)
    achieved = nbytes / kernel_s / 1e9
    out = {
        "metric": METRIC,
        "value": N * a.steps / wall,
        "unit": "compartment-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "myelinated axon, 200 HH nodes (1 um) + 200 x 1 mm passive "
                               "internodes (100 x 10 um), d=10 um, N=20202, backward Euler "
                               "dt=10 us, stimulus 10 nA x 0.5 ms",
                   "N": N, "hh_elements": H, "nnz": nnz, "solver": "BiCGSTAB",
                   "tolerance": f"relative {a.tol:g}", "form": "fp64-plain",
                   "num_blocks": st1["num_blocks"]},
        "krylov_iters_per_step": iters / a.steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_implicit_bicgstab_fp64", "launches": st1["kernel_launches"],
                     "kernel_ms": st1["kernel_ms"], "algorithmic_bytes": nbytes},
        "accuracy": {"form": "fp64-plain", "max_abs_dV_mV_vs_fp64": 0.0,
                     "note": "plain FP64 form; parity vs the FP64 oracle is in tests/"},
    }
    s.close()
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(prob)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
