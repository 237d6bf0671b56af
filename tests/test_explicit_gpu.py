"""Explicit FP64 plain form on the GPU (through the C ABI) against the reference.

Bar: max |dV| <= 1e-9 V (1e-6 mV) over every element and identical spike steps.
References: the golden trajectories written by the reference's own CPU build,
and the oracle (CPU restatement, itself bitwise-equal to those fixtures).
"""
import numpy as np
import pytest

import neurite_hip as nh
import workloads as wl
from common import GOLDEN, TOL_V, load_golden, max_abs_diff, problem_from_golden

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.mark.parametrize("name", GOLDEN)
def test_golden_run(name, nrt):
    g = load_golden(name)
    _say(f"explicit golden run {name} N={g['N']} steps={g['steps_run']}")
    p = problem_from_golden(g)
    with nrt.Solver(p, kind=nh.EXPLICIT) as s:
        tr = s.run(1, g["steps_run"], g["active_first"], g["active_last"], g["dist_mes_points"],
                   g["interval"])
        v = s.snapshot()
    assert len(tr) == len(g["trace_v"])
    assert max_abs_diff(tr, g["trace_v"]) <= TOL_V
    assert max_abs_diff(v, g["final_v"]) <= TOL_V


@pytest.mark.parametrize("name", ["ex1_only_hh", "ex3_hh_y", "ex6_ct_tree", "ex8_neuron"])
def test_golden_solver_interface(name, nrt):
    """update / calculate / snapshot, one step at a time (inc/solver.h)."""
    g = load_golden(name)
    _say(f"explicit golden step-by-step {name}")
    p = problem_from_golden(g)
    cps = dict(zip(g["checkpoint_steps"], g["checkpoints"]))
    last = max(k for k in cps if k <= g["steps_run"])
    with nrt.Solver(p, kind=nh.EXPLICIT) as s:
        for i in range(1, last + 1):
            s.update(g["active_first"] <= i <= g["active_last"])
            s.calculate()
            if i in cps:
                assert max_abs_diff(s.snapshot(), cps[i]) <= TOL_V, f"step {i}"


def _spike_steps(v, thr=0.0):
    v = np.asarray(v)
    return np.nonzero((v[1:] >= thr) & (v[:-1] < thr))[0] + 1


def test_config1_stable_dt(nrt, orc):
    """Config 1 (1,000-compartment HH axon) at dt = 0.8 x critical, 10 ms."""
    geom, run = wl.squid_axon()
    prob, crit = nrt.setup(geom)
    dT = 0.8 * crit
    _say(f"explicit config1 dT={dT:.4e} N={prob.N}")
    prob = prob.with_dt(dT)
    n = run.steps(dT) - 2
    a0, a1 = nrt.active_range(dT, run.steps(dT), run.init_time, run.end_input)
    with nrt.Solver(prob, kind=nh.EXPLICIT) as s:
        tr = s.run(1, n, a0, a1, run.mes, 0)
        v = s.snapshot()
    ref = orc.Explicit(prob)
    rtr, st = ref.run(1, n, a0, a1, run.mes, 0)
    assert st == 0
    assert max_abs_diff(tr, rtr) <= TOL_V
    assert max_abs_diff(v, ref.snapshot()) <= TOL_V
    sp, rsp = _spike_steps(tr), _spike_steps(rtr)
    assert len(rsp) >= 1, "the stimulus must evoke a spike"
    assert np.array_equal(sp, rsp)


def test_config1_literal_dt_diverges_like_reference(nrt, orc):
    """dt = 1 us is 1.41x the reference's own stability bound: FTCS diverges.
    GPU and oracle agree while finite and both blow up."""
    geom, run = wl.squid_axon()
    prob, crit = nrt.setup(geom)
    assert 1e-6 > crit
    prob = prob.with_dt(1e-6)
    a0, a1 = nrt.active_range(1e-6, run.steps(1e-6), run.init_time, run.end_input)
    _say("explicit config1 literal dt=1us")
    steps, early = 400, 5
    with nrt.Solver(prob, kind=nh.EXPLICIT) as s:
        s.run(1, early, a0, a1)
        v5 = s.snapshot()
        s.run(early + 1, steps - early, a0, a1)
        v = s.snapshot()
    ref = orc.Explicit(prob)
    ref.run(1, early, a0, a1)
    r5 = ref.snapshot()
    ref.run(early + 1, steps - early, a0, a1)
    rv = ref.snapshot()
    # the unstable mode grows ~1.83x per step, so an ulp of exp() difference is
    # amplified; agreement is checked before the blow-up
    assert np.all(np.isfinite(v5)) and max_abs_diff(v5, r5) <= TOL_V
    assert (not np.all(np.isfinite(rv))) or np.max(np.abs(rv)) > 1.0
    assert (not np.all(np.isfinite(v))) or np.max(np.abs(v)) > 1.0


def test_device_setup_matches_oracle(nrt, orc):
    """nrt_setup (device setup chain) against the oracle's restatement."""
    for geom, _ in [wl.squid_axon(), wl.myelinated_axon(), wl.random_tree(wl.seed_for(3), depth=7)]:
        _say(f"device setup N={geom.N}")
        H = nh.count_hh(geom)
        s, bufs = nh.setup_out_buffers(geom.N, H)
        gs = geom.struct()
        orc.setup(gs, s)
        prob, crit = nrt.setup(geom)
        assert crit == pytest.approx(s.dT_crit, rel=1e-15)
        for f in ["dx", "c_m", "r_a", "W", "K"]:
            np.testing.assert_allclose(getattr(prob, f), bufs[f], rtol=4e-16, atol=0)
        for f, o in [("hh_m", "hh_m"), ("E_L", "E_L"), ("g_Na", "g_Na"), ("ls_Na", "ls_Na")]:
            np.testing.assert_allclose(getattr(prob, f), bufs[o][:H], rtol=1e-15, atol=0)


def test_tree_matches_oracle(nrt, orc):
    """Branched seeded tree with strain (records, branch points, sealed ends)."""
    geom, run = wl.random_tree(wl.seed_for(3), depth=8)
    prob, crit = nrt.setup(geom)
    dT = 0.8 * crit
    prob = prob.with_dt(dT)
    _say(f"explicit tree N={prob.N} dT={dT:.3e}")
    n = 3000
    with nrt.Solver(prob, kind=nh.EXPLICIT) as s:
        s.run(1, n, 1, 1000)
        v = s.snapshot()
    ref = orc.Explicit(prob)
    ref.run(1, n, 1, 1000)
    assert max_abs_diff(v, ref.snapshot()) <= TOL_V


@pytest.mark.parametrize("name", ["ex3_hh_y", "ex6_ct_tree", "ex8_neuron"])
def test_resident_run_bit_identical_to_per_step(name, nrt):
    """nrt_run on up to 8,192 elements is one launch of the LDS-resident kernel;
    update/calculate launch the per-element kernel once per step: the two give
    the same bits, the trace included, and a run continues where the steps left
    off (and the reverse)."""
    g = load_golden(name)
    p = problem_from_golden(g)
    a0, a1 = g["active_first"], g["active_last"]
    n, ev, te = 700, 6, g["dist_mes_points"]
    _say(f"explicit resident vs per-step {name} N={p.N}")
    with nrt.Solver(p, kind=nh.EXPLICIT) as s:
        tr = s.run(1, n, a0, a1, te, ev)
        assert s.stats()["kernel_launches"] == 1
        v_run = s.snapshot()
        for i in range(n + 1, n + 51):
            s.update(a0 <= i <= a1)
            s.calculate()
        v_mix = s.snapshot()
    with nrt.Solver(p, kind=nh.EXPLICIT) as s:
        trace = []
        for i in range(1, n + 1):
            s.update(a0 <= i <= a1)
            s.calculate()
            if i % (ev + 1) == 0:
                trace.append(s.snapshot()[te])
        assert np.array_equal(s.snapshot(), v_run)
        s.run(n + 1, 50, a0, a1)
        assert np.array_equal(s.snapshot(), v_mix)
    assert np.array_equal(tr, np.array(trace))


def test_large_cable_per_step_path(nrt, orc):
    """Above 8,192 elements nrt_run launches a kernel a step: a 12,000-compartment
    HH cable against the oracle."""
    geom, run = wl.squid_axon(n_comp=12000, length=1.2e-1)
    prob, crit = nrt.setup(geom)
    dT = 0.8 * crit
    prob = prob.with_dt(dT)
    steps = 400
    with nrt.Solver(prob, kind=nh.EXPLICIT) as s:
        s.run(1, steps, 1, 300)
        assert s.stats()["kernel_launches"] == steps
        v = s.snapshot()
    ref = orc.Explicit(prob)
    ref.run(1, steps, 1, 300)
    assert max_abs_diff(v, ref.snapshot()) <= TOL_V
