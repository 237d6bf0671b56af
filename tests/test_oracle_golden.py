"""Pin the oracle against the reference's own CPU build (tests/golden/*).

The fixtures were written by oracle/_ref/neurite_ref: the reference setup chain,
space.cpp builders and SolverExCPU compiled from the reference sources
(tests/golden/make_golden.py). Both sides are g++ -O3 double arithmetic without
contraction, so the oracle must agree bit for bit.
"""
import numpy as np
import pytest

import neurite_hip as nh
from common import GOLDEN, geometry_from_golden, load_golden, problem_from_golden

SETUP_N = ["dx", "c_m", "r_a", "r_m", "W", "K", "crit_dT", "rest"]
SETUP_H = [("hh_m", "hh_m"), ("hh_h", "hh_h"), ("hh_n", "hh_n"), ("E_Na", "E_Na"),
           ("E_K", "E_K"), ("E_L", "E_L"), ("g_Na", "g_Na"), ("g_K", "g_K"), ("G_L", "G_L"),
           ("ls_Na", "ls_Na"), ("ls_K", "ls_K"), ("hh_rest", "hh_rest")]


@pytest.mark.parametrize("name", GOLDEN)
def test_setup_chain_bitwise(name, orc):
    g = load_golden(name)
    geom = geometry_from_golden(g)
    H = len(g["hh_elem"])
    s, bufs = nh.setup_out_buffers(geom.N, H)
    gs = geom.struct()
    orc.setup(gs, s)
    for f in SETUP_N:
        ref = np.asarray(g[f])
        got = bufs[f]
        if f in ("W", "K"):  # HH elements keep W = K = 0 before the first update
            pass
        assert np.array_equal(got, ref), f"{name}:{f} max diff {np.max(np.abs(got - ref))}"
    for ours, theirs in SETUP_H:
        assert np.array_equal(bufs[ours][:H], np.asarray(g[theirs])), f"{name}:{ours}"
    assert np.array_equal(bufs["hh_elem"][:H], np.asarray(g["hh_elem"]))
    assert s.dT_crit == g["dT_crit"]


@pytest.mark.parametrize("name", GOLDEN)
def test_explicit_trajectory_bitwise(name, orc):
    g = load_golden(name)
    p = problem_from_golden(g)
    ex = orc.Explicit(p)
    steps = g["steps_run"]
    cps = dict(zip(g["checkpoint_steps"], g["checkpoints"]))
    trace = []
    gl = 0
    for i in range(1, steps + 1):
        active = g["active_first"] <= i <= g["active_last"]
        ex.update(active)
        ex.calculate()
        if gl + 1 > g["interval"]:
            gl = 0
            trace.append(ex.snapshot()[g["dist_mes_points"]])
        else:
            gl += 1
        if i in cps:
            assert np.array_equal(ex.snapshot(), np.asarray(cps[i])), f"{name} step {i}"
    assert np.array_equal(np.asarray(trace), np.asarray(g["trace_v"]))
    assert np.array_equal(ex.snapshot(), np.asarray(g["final_v"]))


@pytest.mark.parametrize("name", GOLDEN)
def test_run_loop_matches(name, orc):
    """orc_run (the neurite.cpp loop restated) gives the same trace and final V."""
    g = load_golden(name)
    p = problem_from_golden(g)
    ex = orc.Explicit(p)
    tr, st = ex.run(1, g["steps_run"], g["active_first"], g["active_last"],
                    g["dist_mes_points"], g["interval"])
    assert st == 0
    assert np.array_equal(tr, np.asarray(g["trace_v"]))
    assert np.array_equal(ex.snapshot(), np.asarray(g["final_v"]))


@pytest.mark.parametrize("name", GOLDEN)
def test_active_range(name, orc):
    g = load_golden(name)
    a0, a1 = nh.active_range(g["dT"], g["numStepsT"], g["init_time"], g["end_input"])
    b0, b1 = orc.active_range(g["dT"], g["numStepsT"], g["init_time"], g["end_input"])
    assert (a0, a1) == (b0, b1)
    if g["active_first"] >= 0:
        assert a0 == g["active_first"]
        assert a1 >= g["active_last"]
