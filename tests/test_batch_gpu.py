"""Batches of independent neurites (nrt_batch_*) on the GPU.

Every slot must give what the same problem gives alone: FP64 implicit slots within
1e-12 V of the single-problem tree solve and within 1e-9 V of the oracle's exact
solve; FP32-with-refinement slots within 1e-7 V (1e-4 mV) of the oracle; explicit
slots bit-identical to the single-problem explicit step (the same element code on
the same inputs). Slots join at different batch steps (continuous batching) and
leave when removed; traces continue across calls.
"""
import numpy as np
import pytest

import neurite_hip as nh
import workloads as wl
from common import TOL_V, ghost_mask, load_golden, max_abs_diff, problem_from_golden

pytestmark = pytest.mark.gpu
TOL_MIXED = 1e-7
TOL_SAME = 1e-12


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _problems(nrt):
    """Uneven sizes: golden problems, random trees of depth 5..9, a squid axon."""
    out = []
    for name in ("ex3_hh_y", "ex6_ct_tree", "ex8_neuron"):
        g = load_golden(name)
        out.append(problem_from_golden(g, dT=g["dT_crit"] * 50))
    for i, depth in enumerate((5, 7, 9)):
        geom, run = wl.random_tree(wl.seed_for(4, i), depth=depth, amp=(i + 1) * 3e-9)
        out.append(nrt.setup(geom, dT=run.dT)[0])
    geom, run = wl.squid_axon(n_comp=600)
    out.append(nrt.setup(geom, dT=25e-6)[0])
    return out


def _alone(p, steps, a0, a1, **opt):
    with nh.Solver(p, kind=nh.IMPLICIT, krylov=nh.DIRECT, **opt) as s:
        s.run(1, steps, a0, a1)
        return s.snapshot()


def _oracle(orc, p, steps, a0, a1):
    ref = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT)
    ref.run(1, steps, a0, a1, 0, -1)
    return ref.snapshot()


TUNED = (128, 256, 512, 1024)


@pytest.mark.parametrize("threads", [128, 256, 1024, 0])
def test_implicit_batch_matches_alone(threads, nrt, orc):
    probs = _problems(nrt)
    steps, a0, a1 = 60, 1, 25
    _say(f"implicit batch fp64 threads={threads} sizes={[p.N for p in probs]}")
    with nh.Batch(kind=nh.IMPLICIT, krylov=nh.DIRECT, batch_threads=threads) as b:
        ids = [b.add(p, steps, a0, a1) for p in probs]
        for _ in range(5):  # 5 calls of 12 steps (the autotuner times calls 2-5)
            b.step(12)
        st = b.stats()
        assert st["active_slots"] == 0
        assert st["compartment_steps"] == steps * sum(p.N for p in probs)
        if threads == 0:
            assert st["threads"] in TUNED
            assert all(t in TUNED for t in st["class_threads"])
        else:
            assert st["class_threads"] == [threads] * 3
        for sid, p in zip(ids, probs):
            v = b.snapshot(sid)
            info = b.info(sid)
            assert info["status"] == 0 and info["steps_done"] == steps
            assert info["inner_iters"] == steps  # one exact solve per step, every slot
            ref = _alone(p, steps, a0, a1)
            assert max_abs_diff(v, ref, ghost_mask(p)) <= TOL_SAME
            exact = _oracle(orc, p, steps, a0, a1)
            assert max_abs_diff(v, exact, ghost_mask(p)) <= TOL_V


def _class_sample(nrt, per_class):
    """Sweep trees, `per_class` of each size class (< 4,096 rows, < 16,384, the rest)."""
    picked = {}
    for i in range(400):
        geom, run = wl.sweep_tree(i)
        p = nrt.setup(geom, dT=run.dT)[0]
        c = 0 if p.N < 4096 else (1 if p.N < 16384 else 2)
        if len(picked.get(c, [])) < per_class:
            picked.setdefault(c, []).append(p)
        if all(len(picked.get(c, [])) == per_class for c in range(3)):
            break
    return [p for c in range(3) for p in picked[c]]


def test_implicit_batch_size_classes(nrt, orc):
    """Slots of every size class; with this few slots the classes share one launch
    (each class alone would fill less than half the device); every slot gives what it
    gives alone. (The 4,096-tree sweep, which runs a launch per class, is checked
    against the oracle in test_workloads_gpu.)"""
    probs = _class_sample(nrt, 2)
    steps, a0, a1 = 30, 1, 20
    _say(f"implicit batch size classes sizes={[p.N for p in probs]}")
    with nh.Batch(kind=nh.IMPLICIT, krylov=nh.DIRECT) as b:
        ids = [b.add(p, steps, a0, a1) for p in probs]
        for _ in range(6):
            b.step(5)
        st = b.stats()
        assert st["active_slots"] == 0
        assert st["launches"] == 6  # one launch per call: the classes are merged
        assert all(t in TUNED for t in st["class_threads"])
        for sid, p in zip(ids, probs):
            info = b.info(sid)
            assert info["status"] == 0 and info["steps_done"] == steps
            v = b.snapshot(sid)
            assert max_abs_diff(v, _alone(p, steps, a0, a1), ghost_mask(p)) <= TOL_SAME
            assert max_abs_diff(v, _oracle(orc, p, steps, a0, a1), ghost_mask(p)) <= TOL_V


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_batch_tune_per_class(form, nrt, orc):
    """nrt_batch_tune times every (chunk plan, workgroup size) pair per size class on
    a sample; slots added afterwards use the class's choice and still match the
    oracle (FP64: and the single-problem solve)."""
    probs = _class_sample(nrt, 6)
    kw = dict(kind=nh.IMPLICIT, krylov=nh.DIRECT)
    if form == "mixed":
        kw.update(precision=nh.FP32_REFINE, refine_tol=1e-10)
    steps, a0, a1 = 16, 1, 20
    _say(f"batch tune {form} sizes={[p.N for p in probs]}")
    with nh.Batch(**kw) as b:
        r = b.tune(probs, steps=2, rows_per_class=60_000)
        _say(f"  tune: {r}")
        assert r["trials"] == 3 * 7 * 4
        assert set(r["classes"]) == {"N<4096", "4096<=N<16384", "N>=16384"}
        for c in r["classes"].values():
            assert c["threads"] in TUNED
            assert (c["chunk"], c["chunk_hi"]) in {("auto", "auto"), (8, 8), (12, 8), (16, 8),
                                                   (24, 8), (16, 5), (16, 12)}
            assert 0 < c["ns_per_compartment_step"] <= c["auto_256_ns_per_compartment_step"]
            assert 1 <= c["sample_trees"] <= 6
        st = b.stats()
        assert st["class_threads"] == [r["classes"][k]["threads"] for k in nh.SIZE_CLASSES]
        with nh.Batch(**kw) as b2:  # the same shapes kept and set on another batch
            b2.set_shapes(r["classes"])
            assert b2.stats()["class_threads"] == st["class_threads"]
        ids = [b.add(p, steps, a0, a1) for p in probs]
        b.step(steps)
        for sid, p in zip(ids, probs):
            info = b.info(sid)
            assert info["status"] == 0 and info["steps_done"] == steps
            v = b.snapshot(sid)
            if form == "fp64":
                assert max_abs_diff(v, _alone(p, steps, a0, a1), ghost_mask(p)) <= TOL_SAME
            assert max_abs_diff(v, _oracle(orc, p, steps, a0, a1), ghost_mask(p)) <= (
                TOL_V if form == "fp64" else TOL_MIXED)


def test_implicit_batch_mixed_precision(nrt, orc):
    probs = _problems(nrt)
    steps, a0, a1 = 50, 1, 20
    _say("implicit batch fp32+refinement")
    with nh.Batch(kind=nh.IMPLICIT, krylov=nh.DIRECT, precision=nh.FP32_REFINE,
                  refine_tol=1e-10, batch_threads=512) as b:
        ids = [b.add(p, steps, a0, a1) for p in probs]
        b.step(steps)
        # a step whose previous V already meets the residual test takes no sweep
        assert sum(b.info(sid)["inner_iters"] for sid in ids) > 0
        for sid, p in zip(ids, probs):
            info = b.info(sid)
            assert info["status"] == 0 and info["steps_done"] == steps
            assert info["inner_iters"] <= 8 * steps  # refine_max sweeps per step at most
            exact = _oracle(orc, p, steps, a0, a1)
            assert max_abs_diff(b.snapshot(sid), exact, ghost_mask(p)) <= TOL_MIXED


def test_continuous_batching_join_leave(nrt, orc):
    """Slots join at batch steps 0, 7 and 20, one leaves early; each slot's state
    after its own steps equals the problem run alone for that many steps, and its
    trace equals the single-problem trace."""
    probs = _problems(nrt)[:5]
    lengths = [40, 33, 25, 50, 12]
    with nh.Batch(kind=nh.IMPLICIT, krylov=nh.DIRECT, batch_threads=256) as b:
        ids = {}
        ids[0] = b.add(probs[0], lengths[0], 1, 10, trace_elem=1, trace_every=2)
        ids[1] = b.add(probs[1], lengths[1], 3, 12, trace_elem=2, trace_every=0)
        b.step(7)
        ids[2] = b.add(probs[2], lengths[2], 1, 5, trace_elem=1, trace_every=4)
        ids[3] = b.add(probs[3], lengths[3], 2, 30)
        b.step(13)
        ids[4] = b.add(probs[4], lengths[4], 1, 12)
        b.remove(ids[3])  # leaves after 13 of its steps
        while b.stats()["active_slots"] > 0:
            b.step(6)
        for k in (0, 1, 2, 4):
            p, n = probs[k], lengths[k]
            a0, a1 = {0: (1, 10), 1: (3, 12), 2: (1, 5), 4: (1, 12)}[k]
            te, ev = {0: (1, 2), 1: (2, 0), 2: (1, 4), 4: (0, -1)}[k]
            with nh.Solver(p, kind=nh.IMPLICIT, krylov=nh.DIRECT) as s:
                tr = s.run(1, n, a0, a1, te, ev)
                v = s.snapshot()
            assert b.info(ids[k])["steps_done"] == n
            assert max_abs_diff(b.snapshot(ids[k]), v, ghost_mask(p)) <= TOL_SAME
            if ev >= 0:
                bt = b.trace(ids[k])
                assert len(bt) == len(tr) == n // (ev + 1)
                assert np.max(np.abs(bt - tr)) <= TOL_SAME


def test_explicit_batch_bit_identical(nrt):
    probs = []
    for name in ("ex2_hh_axon", "ex3_hh_y", "ex6_ct_tree", "ex8_neuron"):
        g = load_golden(name)
        probs.append(problem_from_golden(g, dT=g["dT_crit"] * 0.8))
    geom, _ = wl.squid_axon(n_comp=700)
    p, crit = nrt.setup(geom, dT=None, dt_factor=0.8)
    probs.append(p)
    steps = 300
    _say(f"explicit batch sizes={[q.N for q in probs]}")
    with nh.Batch(kind=nh.EXPLICIT) as b:
        ids = [b.add(q, steps, 1, 100, trace_elem=1, trace_every=9) for q in probs[:3]]
        b.step(40)
        ids += [b.add(q, steps, 1, 100, trace_elem=1, trace_every=9) for q in probs[3:]]
        while b.stats()["active_slots"] > 0:
            b.step(40)  # the same graph is replayed
        for sid, q in zip(ids, probs):
            with nh.Solver(q, kind=nh.EXPLICIT) as s:
                tr = s.run(1, steps, 1, 100, 1, 9)
                v = s.snapshot()
            assert np.array_equal(b.snapshot(sid), v)
            assert np.array_equal(b.trace(sid), tr)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_implicit_batch_virtual_grids(form, nrt, orc):
    """Slots above 24,576 rows run on several co-resident workgroups with their own
    grid barrier inside the batched launch; results must not change."""
    probs = []
    for depth in (12, 11):
        geom, run = wl.random_tree(wl.seed_for(4, 40 + depth), depth=depth, amp=5e-9)
        probs.append(nrt.setup(geom, dT=run.dT)[0])
    geom, run = wl.myelinated_axon()
    probs.append(nrt.setup(geom, dT=run.dT)[0])
    geom, run = wl.squid_axon(n_comp=500)
    probs.append(nrt.setup(geom, dT=25e-6)[0])
    assert max(p.N for p in probs) > 8 * 24576
    opts = dict(krylov=nh.DIRECT) if form == "fp64" else dict(
        krylov=nh.DIRECT, precision=nh.FP32_REFINE, refine_tol=1e-10)
    steps, a0, a1 = 30, 1, 20
    _say(f"implicit batch virtual grids {form} sizes={[p.N for p in probs]}")
    with nh.Batch(kind=nh.IMPLICIT, batch_threads=256, **opts) as b:
        ids = [b.add(p, steps, a0, a1) for p in probs]
        b.step(10)
        b.step(20)
        for sid, p in zip(ids, probs):
            info = b.info(sid)
            assert info["status"] == 0 and info["steps_done"] == steps
            v = b.snapshot(sid)
            exact = _oracle(orc, p, steps, a0, a1)
            if form == "fp64":
                assert info["inner_iters"] == steps
                assert max_abs_diff(v, _alone(p, steps, a0, a1), ghost_mask(p)) <= TOL_SAME
                assert max_abs_diff(v, exact, ghost_mask(p)) <= TOL_V
            else:
                assert max_abs_diff(v, exact, ghost_mask(p)) <= TOL_MIXED
