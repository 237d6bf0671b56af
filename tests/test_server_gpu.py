"""Request server (nrt_server_*) on the GPU: requests submitted while others run
are admitted, batched and retired; each result is the same alone and in the mix
(FP64: within 1e-12 V of the single-problem tree solve, traces too; FP32 with
FP64 refinement: within 1e-7 V of the oracle's exact solve)."""
import time

import numpy as np
import pytest

import neurite_hip as nh
import workloads as wl
from common import ghost_mask, load_golden, max_abs_diff, problem_from_golden

pytestmark = pytest.mark.gpu
TOL_SAME = 1e-12
TOL_MIXED = 1e-7


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _requests(nrt):
    out = []
    for i, name in enumerate(("ex3_hh_y", "ex6_ct_tree", "ex8_neuron")):
        g = load_golden(name)
        out.append((problem_from_golden(g, dT=g["dT_crit"] * 50), 40 + 7 * i, 1, 15, 1, 3))
    for i, depth in enumerate((6, 8, 9, 10)):
        geom, run = wl.random_tree(wl.seed_for(5, i), depth=depth, amp=4e-9)
        out.append((nrt.setup(geom, dT=run.dT)[0], 30 + 11 * i, 2, 20, 1, 0))
    return out


def _alone(p, n, a0, a1, te, ev, **opt):
    with nh.Solver(p, kind=nh.IMPLICIT, krylov=nh.DIRECT, **opt) as s:
        tr = s.run(1, n, a0, a1, te, ev)
        return s.snapshot(), tr


@pytest.mark.parametrize("k", [1, 8])
def test_server_results_same_alone_and_in_mix(k, nrt):
    reqs = _requests(nrt)
    sizes = [r[0].N for r in reqs]
    # two requests above the threshold take the grid-wide path (one in 4-workgroup
    # cooperative mode: N >= 16384)
    thr = sorted(sizes)[-3]
    _say(f"server k={k} sizes={sizes} big_threshold={thr}")
    with nh.Server(kind=nh.IMPLICIT, krylov=nh.DIRECT, steps_per_launch=k,
                   big_threshold=thr, batch_threads=256) as srv:
        ids = []
        for p, n, a0, a1, te, ev in reqs:
            ids.append(srv.submit(p, n, a0, a1, te, ev))
            time.sleep(0.003)  # arrivals while earlier requests run
        srv.wait_idle(60.0)
        done = {r["id"]: r for r in srv.poll()}
        assert sorted(done) == sorted(ids)
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert 0 < st["p50_ms"] <= st["p99_ms"]
        for rid, (p, n, a0, a1, te, ev) in zip(ids, reqs):
            info = done[rid]
            assert info["status"] == 0 and info["steps_done"] == n
            assert info["submit_s"] <= info["start_s"] <= info["finish_s"]
            status, v, tr = srv.fetch(rid)
            ref, rtr = _alone(p, n, a0, a1, te, ev)
            assert max_abs_diff(v, ref, ghost_mask(p)) <= TOL_SAME
            assert len(tr) == len(rtr) and np.max(np.abs(tr - rtr)) <= TOL_SAME


def test_server_mixed_precision(nrt, orc):
    reqs = _requests(nrt)[:5]
    _say("server fp32+refinement")
    with nh.Server(kind=nh.IMPLICIT, krylov=nh.DIRECT, precision=nh.FP32_REFINE,
                   refine_tol=1e-10, steps_per_launch=10) as srv:
        ids = [srv.submit(p, n, a0, a1) for p, n, a0, a1, _, _ in reqs]
        srv.wait_idle(60.0)
        for rid, (p, n, a0, a1, _, _) in zip(ids, reqs):
            status, v, _ = srv.fetch(rid)
            assert status == 0
            ex = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT)
            ex.run(1, n, a0, a1, 0, -1)
            assert max_abs_diff(v, ex.snapshot(), ghost_mask(p)) <= TOL_MIXED
