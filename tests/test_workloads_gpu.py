"""Parity on the benchmark workloads themselves (configs 3-5), every result against
the oracle's exact solve of the same problem for the same steps, on non-ghost
elements: FP64 within 1e-9 V, FP32 with FP64 refinement within 1e-7 V (1e-4 mV,
the metric's tolerance).

  config 3: the seeded depth-14 dendritic tree (N = 902,037), dt 25 us: the FP64
            and the FP32 + refinement tree solves (local mode, 64 workgroups) for
            20 steps, and FP32 GMRES(30) cycles inside FP64 refinement for 3 steps;
  config 4: 64 trees of the seeded sweep (its sizes span 500..50,000) in one batch,
            both precisions, and the full 4,096-tree batch for 10 FP64 steps with
            16 slots across the size range checked;
  config 5: the first 40 requests of the served stream plus its largest of 400
            (N = 344,689), submitted with their Poisson gaps, each capped at 100
            steps so the oracle stays within seconds.
"""
import time

import numpy as np
import pytest

import neurite_hip as nh
import workloads as wl
from common import TOL_V, ghost_mask, max_abs_diff

pytestmark = pytest.mark.gpu
TOL_MIXED = 1e-7


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _exact(orc, p, steps, a0, a1):
    ref = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT)
    _, st = ref.run(1, steps, a0, a1, 0, -1)
    assert st == 0
    return ref.snapshot()


@pytest.fixture(scope="module")
def tree3(nrt):
    geom, run = wl.random_tree(wl.seed_for(3), depth=14)
    prob, _ = nrt.setup(geom, dT=run.dT)
    a0, a1 = nrt.active_range(run.dT, run.steps(run.dT), run.init_time, run.end_input)
    return prob, a0, a1


@pytest.mark.parametrize("form,steps,opts,tol", [
    ("fp64_tree", 20, dict(krylov=nh.DIRECT), TOL_V),
    ("mixed_tree", 20, dict(krylov=nh.DIRECT, precision=nh.FP32_REFINE, refine_tol=1e-10),
     TOL_MIXED),
    ("mixed_gmres30", 3, dict(krylov=nh.GMRES, restart=30, precision=nh.FP32_REFINE,
                              refine_tol=1e-12, max_iter=4000), TOL_MIXED),
])
def test_config3_depth14_tree(form, steps, opts, tol, tree3, orc):
    prob, a0, a1 = tree3
    assert prob.N == 902037
    _say(f"config 3 {form} N={prob.N} steps={steps}")
    with nh.Solver(prob, kind=nh.IMPLICIT, **opts) as s:
        s.run(1, steps, a0, a1)
        v = s.snapshot()
        st = s.stats()
    err = max_abs_diff(v, _exact(orc, prob, steps, a0, a1), ghost_mask(prob))
    _say(f"  max|dV| = {err:.3e} V, workgroups {st['num_blocks']}, inner {st['krylov_iters']}")
    assert err <= tol


def _sweep(nrt, indices):
    out = []
    for i in indices:
        geom, run = wl.sweep_tree(i)
        p, _ = nrt.setup(geom, dT=run.dT)
        out.append(p)
    return out


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_config4_sweep_64_trees(form, nrt, orc):
    probs = _sweep(nrt, range(64))
    sizes = sorted(p.N for p in probs)
    assert sizes[0] < 1000 and sizes[-1] > 30000  # the sweep's size range
    opts = dict(krylov=nh.DIRECT) if form == "fp64" else dict(
        krylov=nh.DIRECT, precision=nh.FP32_REFINE, refine_tol=1e-10)
    steps, a0, a1 = 20, 1, 20  # the sweep's stimulus: 0.5 ms at dt 25 us
    _say(f"config 4 {form} 64 trees, {sum(sizes)} compartments")
    with nh.Batch(kind=nh.IMPLICIT, **opts) as b:
        ids = [b.add(p, steps, a0, a1) for p in probs]
        b.step(8)
        b.step(12)
        worst = 0.0
        for sid, p in zip(ids, probs):
            info = b.info(sid)
            assert info["status"] == 0 and info["steps_done"] == steps
            err = max_abs_diff(b.snapshot(sid), _exact(orc, p, steps, a0, a1), ghost_mask(p))
            worst = max(worst, err)
    _say(f"  worst max|dV| = {worst:.3e} V")
    assert worst <= (TOL_V if form == "fp64" else TOL_MIXED)


def test_config4_full_sweep_4096_trees(nrt, orc):
    t0 = time.perf_counter()
    probs = _sweep(nrt, range(4096))
    total = sum(p.N for p in probs)
    _say(f"config 4 fp64 4096 trees, {total} compartments (built in "
         f"{time.perf_counter() - t0:.1f} s)")
    assert 40e6 < total < 46e6
    order = np.argsort([p.N for p in probs])
    check = [int(order[k]) for k in np.linspace(0, 4095, 16).astype(int)]
    steps, a0, a1 = 10, 1, 20
    with nh.Batch(kind=nh.IMPLICIT, krylov=nh.DIRECT) as b:
        ids = [b.add(p, steps, a0, a1) for p in probs]
        b.step(steps)
        assert b.stats()["active_slots"] == 0
        for i in check:
            p = probs[i]
            info = b.info(ids[i])
            assert info["status"] == 0 and info["steps_done"] == steps
            err = max_abs_diff(b.snapshot(ids[i]), _exact(orc, p, steps, a0, a1), ghost_mask(p))
            assert err <= TOL_V, (i, p.N, err)


def _stream(nrt):
    sizes = [wl.served_request_target(i) for i in range(400)]
    indices = list(range(40)) + [int(np.argmax(sizes))]
    reqs = []
    for i in indices:
        geom, run, gap = wl.served_request(i)
        p, _ = nrt.setup(geom, dT=run.dT)
        reqs.append((p, min(100, run.steps(run.dT)), gap))
    return reqs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config5_served_stream(precision, nrt, orc):
    reqs = _stream(nrt)
    assert max(p.N for p, _, _ in reqs) == 344689
    kw = dict(krylov=nh.DIRECT)
    if precision == "mixed":
        kw.update(precision=nh.FP32_REFINE, refine_tol=1e-10)
    _say(f"config 5 {precision}: {len(reqs)} requests, sizes {min(p.N for p, _, _ in reqs)}.."
         f"{max(p.N for p, _, _ in reqs)}")
    with nh.Server(kind=nh.IMPLICIT, steps_per_launch=16, **kw) as srv:
        ids = []
        t0 = time.perf_counter()
        due = 0.0
        for p, n, gap in reqs:
            due += gap
            wait = t0 + due - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            ids.append(srv.submit(p, n, 1, 20))
        srv.wait_idle(120.0)
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        worst = 0.0
        for rid, (p, n, _) in zip(ids, reqs):
            status, v, _ = srv.fetch(rid)
            assert status == 0
            err = max_abs_diff(v, _exact(orc, p, n, 1, 20), ghost_mask(p))
            worst = max(worst, err)
    _say(f"  worst max|dV| = {worst:.3e} V, p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms")
    assert worst <= (TOL_V if precision == "fp64" else TOL_MIXED)
