"""BiCGSTAB (the reference GPU solver's method, solverIm-gpu.cpp:725-843)
right-preconditioned by the tree solve (NRT_PRECOND_TREE) on the GPU, against the
oracle's exact solve on non-ghost elements. FP64 preconditioner (exact): one
iteration a step, within 1e-9 V. FP32 preconditioner (the FP32 tree solve inside
FP64 BiCGSTAB): a few iterations, within 1e-7 V (1e-4 mV). Single handles in block
and grid launches, batch slots and served requests."""
import numpy as np
import pytest

import neurite_hip as nh
import workloads as wl
from common import TOL_V, ghost_mask, load_golden, max_abs_diff, problem_from_golden

pytestmark = pytest.mark.gpu
TOL_MIXED = 1e-7
PC64 = dict(krylov=nh.BICGSTAB, precond=nh.PRECOND_TREE, tol=1e-12, tol_mode=nh.TOL_RELATIVE)
PC32 = dict(PC64, precision=nh.FP32_REFINE)


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _exact(orc, p, steps, a0, a1):
    ref = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT)
    _, st = ref.run(1, steps, a0, a1, 0, -1)
    assert st == 0
    return ref.snapshot()


@pytest.mark.parametrize("form", ["fp64", "fp32"])
@pytest.mark.parametrize("name,dtf", [("ex3_hh_y", 100), ("ex4_ct_cable", 100),
                                      ("ex6_ct_tree", 10), ("ex8_neuron", 100)])
def test_precond_bicgstab_golden(form, name, dtf, nrt, orc):
    g = load_golden(name)
    p = problem_from_golden(g, dT=g["dT_crit"] * dtf)
    a0, a1 = nrt.active_range(p.dT, 10 ** 9, g["init_time"], g["end_input"])
    steps = 100
    _say(f"precond bicgstab {form} {name} N={p.N}")
    with nrt.Solver(p, kind=nh.IMPLICIT, **(PC64 if form == "fp64" else PC32)) as s:
        s.run(1, steps, a0, a1)
        v = s.snapshot()
        st = s.stats()
    err = max_abs_diff(v, _exact(orc, p, steps, a0, a1), ghost_mask(p))
    its = st["krylov_iters"] / steps
    _say(f"  max|dV| = {err:.3e} V, {its:.2f} iterations a step, {st['refine_sweeps']} solves")
    if form == "fp64":
        assert its <= 1.0 and err <= TOL_V
    else:
        assert its <= 5.0 and err <= TOL_MIXED


@pytest.mark.parametrize("form", ["fp64", "fp32"])
@pytest.mark.parametrize("nblocks", [0, 4])
def test_precond_bicgstab_config2(form, nblocks, nrt, orc):
    """Config 2 (N = 20,202): one workgroup, and a 4-workgroup grid launch."""
    geom, run = wl.myelinated_axon()
    prob, _ = nrt.setup(geom, dT=run.dT)
    a0, a1 = nrt.active_range(run.dT, run.steps(run.dT), run.init_time, run.end_input)
    steps, probe = 300, 1011
    _say(f"precond bicgstab {form} config2 nblocks={nblocks}")
    with nrt.Solver(prob, kind=nh.IMPLICIT, num_blocks=nblocks,
                    **(PC64 if form == "fp64" else PC32)) as s:
        tr = s.run(1, steps, a0, a1, trace_elem=probe, interval=0)
        v = s.snapshot()
        st = s.stats()
    ref = orc.Implicit(prob, variant=orc.VAR_GPU, method=orc.M_DIRECT)
    rtr, _ = ref.run(1, steps, a0, a1, probe, 0)
    tol = TOL_V if form == "fp64" else TOL_MIXED
    assert max_abs_diff(v, ref.snapshot(), ghost_mask(prob)) <= tol
    assert max_abs_diff(tr, rtr) <= tol
    assert st["krylov_iters"] <= (1 if form == "fp64" else 5) * steps
    assert st["tree_form"] == (nh.TREE_BLOCK if nblocks == 0 else nh.TREE_GRID)


def _problems(nrt):
    out = []
    for name in ("ex3_hh_y", "ex6_ct_tree", "ex8_neuron"):
        g = load_golden(name)
        out.append(problem_from_golden(g, dT=g["dT_crit"] * 50))
    for i, depth in enumerate((5, 8, 10)):
        geom, run = wl.random_tree(wl.seed_for(4, i), depth=depth, amp=(i + 1) * 3e-9)
        out.append(nrt.setup(geom, dT=run.dT)[0])
    return out


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_precond_bicgstab_batch(form, nrt, orc):
    probs = _problems(nrt)
    steps, a0, a1 = 40, 1, 20
    opts = PC64 if form == "fp64" else PC32
    _say(f"precond bicgstab batch {form} sizes={[p.N for p in probs]}")
    with nh.Batch(kind=nh.IMPLICIT, batch_threads=256, **opts) as b:
        ids = [b.add(p, steps, a0, a1) for p in probs]
        b.step(15)
        b.step(25)
        for sid, p in zip(ids, probs):
            info = b.info(sid)
            assert info["status"] == 0 and info["steps_done"] == steps
            v = b.snapshot(sid)
            exact = _exact(orc, p, steps, a0, a1)
            if form == "fp64":
                assert info["inner_iters"] <= steps  # at most one iteration a step
                with nh.Solver(p, kind=nh.IMPLICIT, **opts) as s:
                    s.run(1, steps, a0, a1)
                    assert max_abs_diff(v, s.snapshot(), ghost_mask(p)) <= 1e-12
                assert max_abs_diff(v, exact, ghost_mask(p)) <= TOL_V
            else:
                assert max_abs_diff(v, exact, ghost_mask(p)) <= TOL_MIXED


def test_precond_bicgstab_server(nrt, orc):
    probs = _problems(nrt)
    _say("precond bicgstab served, fp32 preconditioner")
    with nh.Server(kind=nh.IMPLICIT, steps_per_launch=8, **PC32) as srv:
        ids = [srv.submit(p, 30 + 5 * i, 1, 15) for i, p in enumerate(probs)]
        srv.wait_idle(60.0)
        for i, (rid, p) in enumerate(zip(ids, probs)):
            status, v, _ = srv.fetch(rid)
            assert status == 0
            assert max_abs_diff(v, _exact(orc, p, 30 + 5 * i, 1, 15), ghost_mask(p)) <= TOL_MIXED
