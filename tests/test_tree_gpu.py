"""Implicit step with the exact tree partition solve on the GPU.

FP64 direct (NRT_DIRECT, FP64): bar max |dV| <= 1e-9 V against the oracle's exact
solve on non-ghost elements. FP32 tree solves inside FP64 iterative refinement
(NRT_DIRECT, FP32_REFINE): bar max |dV| <= 1e-7 V (1e-4 mV, the metric's
tolerance) against the same FP64 truth.
"""
import numpy as np
import pytest

import neurite_hip as nh
import workloads as wl
from common import TOL_V, ghost_mask, load_golden, max_abs_diff, problem_from_golden

pytestmark = pytest.mark.gpu
TOL_MIXED = 1e-7


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _exact(orc, p, steps, a0, a1, trace_elem=0, interval=-1):
    ref = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT)
    tr, st = ref.run(1, steps, a0, a1, trace_elem, interval)
    assert st == 0
    return ref.snapshot(), tr


FORMS = {"fp64": dict(krylov=nh.DIRECT, precision=nh.FP64),
         "mixed": dict(krylov=nh.DIRECT, precision=nh.FP32_REFINE, refine_tol=1e-10)}


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("name,dtf", [("ex3_hh_y", 100), ("ex4_ct_cable", 100),
                                      ("ex6_ct_tree", 10), ("ex8_neuron", 100)])
def test_tree_golden(form, name, dtf, nrt, orc):
    g = load_golden(name)
    dT = g["dT_crit"] * dtf
    p = problem_from_golden(g, dT=dT)
    a0, a1 = nrt.active_range(dT, 10 ** 9, g["init_time"], g["end_input"])
    _say(f"tree {form} {name} N={p.N}")
    with nrt.Solver(p, kind=nh.IMPLICIT, **FORMS[form]) as s:
        s.run(1, 200, a0, a1)
        v = s.snapshot()
        st = s.stats()
    ref, _ = _exact(orc, p, 200, a0, a1)
    tol = TOL_V if form == "fp64" else TOL_MIXED
    err = max_abs_diff(v, ref, ghost_mask(p))
    _say(f"  max|dV| = {err:.3e} V, inner solves {st['krylov_iters']}")
    assert err <= tol, st


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("nblocks", [0, 4, 3, 16])
def test_tree_config2(form, nblocks, nrt, orc):
    """Config 2 (N = 20,202): default (one workgroup) and forced grids."""
    geom, run = wl.myelinated_axon()
    prob, _ = nrt.setup(geom, dT=run.dT)
    a0, a1 = nrt.active_range(run.dT, run.steps(run.dT), run.init_time, run.end_input)
    steps = 300
    _say(f"tree {form} config2 nblocks={nblocks}")
    probe = 1011  # node 11: the spike passes it at ~1.1 ms
    with nrt.Solver(prob, kind=nh.IMPLICIT, num_blocks=nblocks, **FORMS[form]) as s:
        tr = s.run(1, steps, a0, a1, trace_elem=probe, interval=0)
        v = s.snapshot()
    ref, rtr = _exact(orc, prob, steps, a0, a1, probe, 0)
    tol = TOL_V if form == "fp64" else TOL_MIXED
    assert max_abs_diff(v, ref, ghost_mask(prob)) <= tol
    assert max_abs_diff(tr, rtr) <= tol
    assert np.max(rtr) > 0.0, "the spike must pass node 11"


@pytest.mark.parametrize("chunk", [3, 8, 64])
def test_tree_chunk_sizes(chunk, nrt, orc):
    geom, run = wl.random_tree(wl.seed_for(3), depth=9)
    prob, _ = nrt.setup(geom, dT=run.dT)
    _say(f"tree fp64 random depth 9 N={prob.N} chunk={chunk}")
    with nrt.Solver(prob, kind=nh.IMPLICIT, krylov=nh.DIRECT, chunk=chunk) as s:
        s.run(1, 100, 1, 20)
        v = s.snapshot()
    ref, _ = _exact(orc, prob, 100, 1, 20)
    assert max_abs_diff(v, ref, ghost_mask(prob)) <= TOL_V


@pytest.mark.parametrize("form", sorted(FORMS))
def test_tree_large_grid_mode(form, nrt, orc):
    """A depth-12 tree (N ~ 2.2e5) takes the grid mode by default."""
    geom, run = wl.random_tree(wl.seed_for(3), depth=12)
    prob, _ = nrt.setup(geom, dT=run.dT)
    assert prob.N > (1 << 17)
    _say(f"tree {form} random depth 12 N={prob.N}")
    with nrt.Solver(prob, kind=nh.IMPLICIT, **FORMS[form]) as s:
        s.run(1, 40, 1, 20)
        v = s.snapshot()
        st = s.stats()
    assert st["num_blocks"] > 1
    ref, _ = _exact(orc, prob, 40, 1, 20)
    tol = TOL_V if form == "fp64" else TOL_MIXED
    assert max_abs_diff(v, ref, ghost_mask(prob)) <= tol


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("groups,chunk", [(2, 0), (8, 0), (8, 32), (16, 0), (37, 0)])
def test_tree_local_mode_config2(form, groups, chunk, nrt, orc):
    """Config 2 in local mode: one workgroup per row group, the top of the final
    forest (a chain across the group boundaries) solved between two grid barriers."""
    geom, run = wl.myelinated_axon()
    prob, _ = nrt.setup(geom, dT=run.dT)
    a0, a1 = nrt.active_range(run.dT, run.steps(run.dT), run.init_time, run.end_input)
    steps = 300
    _say(f"tree {form} config2 local groups={groups} chunk={chunk}")
    probe = 1011
    with nrt.Solver(prob, kind=nh.IMPLICIT, tree_groups=groups, chunk=chunk,
                    **FORMS[form]) as s:
        tr = s.run(1, steps, a0, a1, trace_elem=probe, interval=0)
        v = s.snapshot()
        st = s.stats()
    assert st["num_blocks"] == groups
    ref, rtr = _exact(orc, prob, steps, a0, a1, probe, 0)
    tol = TOL_V if form == "fp64" else TOL_MIXED
    assert max_abs_diff(v, ref, ghost_mask(prob)) <= tol
    assert max_abs_diff(tr, rtr) <= tol
    assert np.max(rtr) > 0.0


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("name,dtf", [("ex3_hh_y", 100), ("ex6_ct_tree", 10),
                                      ("ex8_neuron", 100)])
def test_tree_local_mode_golden(form, name, dtf, nrt, orc):
    g = load_golden(name)
    dT = g["dT_crit"] * dtf
    p = problem_from_golden(g, dT=dT)
    a0, a1 = nrt.active_range(dT, 10 ** 9, g["init_time"], g["end_input"])
    _say(f"tree {form} local groups=3 {name} N={p.N}")
    with nrt.Solver(p, kind=nh.IMPLICIT, tree_groups=3, chunk=4, **FORMS[form]) as s:
        s.run(1, 200, a0, a1)
        v = s.snapshot()
    ref, _ = _exact(orc, p, 200, a0, a1)
    tol = TOL_V if form == "fp64" else TOL_MIXED
    assert max_abs_diff(v, ref, ghost_mask(p)) <= tol


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("groups", [32, 128, 256])
def test_tree_local_mode_large(form, groups, nrt, orc):
    """A depth-12 tree in local mode: 32 and 128 groups stage the top in LDS, 256
    groups leave a top too large for it (solved from the plan's global lists)."""
    geom, run = wl.random_tree(wl.seed_for(3), depth=12)
    prob, _ = nrt.setup(geom, dT=run.dT)
    _say(f"tree {form} local groups={groups} random depth 12 N={prob.N}")
    with nrt.Solver(prob, kind=nh.IMPLICIT, tree_groups=groups, **FORMS[form]) as s:
        s.run(1, 40, 1, 20)
        v = s.snapshot()
        st = s.stats()
    assert st["num_blocks"] == groups
    ref, _ = _exact(orc, prob, 40, 1, 20)
    tol = TOL_V if form == "fp64" else TOL_MIXED
    assert max_abs_diff(v, ref, ghost_mask(prob)) <= tol
