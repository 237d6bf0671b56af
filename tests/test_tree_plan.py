"""The tree partition solve (TPS) plan and routines, run on the host with the same
per-chunk / per-node code the device kernels execute, against the oracle's
sequential exact elimination and dense numpy solves (CPU only)."""
import ctypes as C

import numpy as np
import pytest

import neurite_hip as nh
import workloads as wl
from common import load_golden, problem_from_golden
from test_oracle_math import _problem, _setup

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def tps_host(rp, col, val, rhs, chunk, chunk_hi=0, dyn=None, rhs_factor=None, groups=0,
             cut_out=None, check=True):
    L = nh.lib()
    f = L.nrt_test_tree_solve_host
    f.argtypes = [C.c_int, _ip, _ip, _dp, _dp, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                  C.c_int, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    rp = np.ascontiguousarray(rp, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    rhs = np.ascontiguousarray(rhs, np.float64)
    x = np.empty(len(rhs))
    lv, h, cut = C.c_int(0), C.c_int(0), C.c_int(0)
    dyn = None if dyn is None else np.ascontiguousarray(dyn, np.uint8)
    rf = None if rhs_factor is None else np.ascontiguousarray(rhs_factor, np.float64)
    st = f(len(rhs), rp.ctypes.data_as(_ip), col.ctypes.data_as(_ip), val.ctypes.data_as(_dp),
           rhs.ctypes.data_as(_dp), chunk, chunk_hi,
           None if dyn is None else dyn.ctypes.data, None if rf is None else rf.ctypes.data,
           groups, x.ctypes.data_as(_dp), C.byref(lv), C.byref(h), C.byref(cut))
    if not check and st != 0:
        return None
    assert st == 0
    if cut_out is not None:
        cut_out.append(cut.value)
    return x, lv.value, h.value


def _system(orc, p):
    im = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT)
    im.update(True)
    return im.matrix()


@pytest.mark.parametrize("name", ["ex3_hh_y", "ex4_ct_cable", "ex5_ct_y", "ex6_ct_tree",
                                  "ex8_neuron"])
@pytest.mark.parametrize("chunk,chunk_hi", [(3, 0), (4, 0), (7, 0), (32, 0), (64, 0), (20, 6),
                                            (5, 40)])
def test_tps_golden_systems(name, chunk, chunk_hi, orc):
    g = load_golden(name)
    p = problem_from_golden(g, dT=g["dT_crit"] * 10)
    rp, col, val, rhs = _system(orc, p)
    x, lv, h = tps_host(rp, col, val, rhs, chunk, chunk_hi)
    ref = orc.tree_solve(rp, col, val, rhs)
    assert np.max(np.abs(x - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("chunk,chunk_hi", [(4, 0), (16, 0), (64, 0), (20, 8)])
def test_tps_long_chain_and_tree(chunk, chunk_hi, orc):
    for geom, dT in [(wl.myelinated_axon()[0], 1e-5),
                     (wl.random_tree(wl.seed_for(3), depth=8)[0], 2.5e-5)]:
        s, bufs, H = _setup(orc, geom)
        p = _problem(geom, s, bufs, H, dT)
        rp, col, val, rhs = _system(orc, p)
        x, lv, h = tps_host(rp, col, val, rhs, chunk, chunk_hi)
        ref = orc.tree_solve(rp, col, val, rhs)
        assert np.max(np.abs(x - ref)) <= 1e-11 * np.max(np.abs(ref))
        # the compressed forest must be shallow: at most 2 nodes per chain remain
        assert lv >= 1


def test_tps_random_forest_dense(orc):
    """Random diagonally dominant forests (parent < child), against numpy."""
    rng = np.random.default_rng(7)
    for trial in range(20):
        N = int(rng.integers(2, 400))
        par = [-1] + [int(rng.integers(-1, i)) if rng.random() < 0.3 else i - 1
                      for i in range(1, N)]
        A = np.zeros((N, N))
        for i in range(N):
            A[i, i] = 4.0 + rng.random()
            if par[i] >= 0:
                A[i, par[i]] = -rng.random()
                A[par[i], i] = -rng.random()
        # make it diagonally dominant
        for i in range(N):
            A[i, i] = 1.0 + np.sum(np.abs(A[i])) - abs(A[i, i]) + A[i, i]
        rp, col, val = [0], [], []
        for i in range(N):
            for j in np.nonzero(A[i])[0]:
                col.append(j)
                val.append(A[i, j])
            rp.append(len(col))
        b = rng.standard_normal(N)
        for chunk in (3, 5, 32):
            x, _, _ = tps_host(rp, col, val, b, chunk)
            np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("name", ["ex3_hh_y", "ex6_ct_tree", "ex8_neuron"])
@pytest.mark.parametrize("chunk,chunk_hi", [(4, 0), (20, 8), (7, 3)])
def test_tps_reused_factors_identical(name, chunk, chunk_hi, orc):
    """Chunks without a changing row reuse the factors of an elimination done with
    another right-hand side: the solution must be bit-identical to a full solve."""
    g = load_golden(name)
    p = problem_from_golden(g, dT=g["dT_crit"] * 10)
    rp, col, val, rhs = _system(orc, p)
    rng = np.random.default_rng(3)
    for frac in (0.0, 0.05, 0.5):
        dyn = (rng.random(len(rhs)) < frac).astype(np.uint8)
        full, _, _ = tps_host(rp, col, val, rhs, chunk, chunk_hi)
        reuse, _, _ = tps_host(rp, col, val, rhs, chunk, chunk_hi, dyn=dyn,
                               rhs_factor=rng.standard_normal(len(rhs)))
        assert np.array_equal(full, reuse), frac
        ref = orc.tree_solve(rp, col, val, rhs)
        assert np.max(np.abs(reuse - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("index", [0, 5, 176, 504, 400])
def test_tps_sweep_trees(index, orc):
    """Trees of the 4,096-tree batched sweep (603 to 46,979 rows; 176 and 504 have
    final forests of 5,118 and 10,238 nodes, past the LDS of a workgroup) with the
    chunk lengths its size classes use (auto 8..16, 16/8, 24/8, 16/12), HH rows
    dynamic and factors reused: the hook also solves each final forest both staged
    (LDS layout) and by the task-indexed lists and requires the same result to the
    bit, and the reused-factor solve must equal the full one to the bit and the exact
    solve to 1e-12."""
    geom, run = wl.sweep_tree(index)
    s, bufs, H = _setup(orc, geom)
    p = _problem(geom, s, bufs, H, run.dT)
    rp, col, val, rhs = _system(orc, p)
    n = len(rhs)
    ref = orc.tree_solve(rp, col, val, rhs)
    dyn = np.zeros(n, np.uint8)
    dyn[np.asarray(bufs["hh_elem"][:H])] = 1
    rng = np.random.default_rng(index)
    for chunk, chunk_hi in [(min(16, max(8, -(-n // 256))), 8), (16, 8), (24, 8), (16, 12)]:
        full, lv, _ = tps_host(rp, col, val, rhs, chunk, chunk_hi)
        reuse, _, _ = tps_host(rp, col, val, rhs, chunk, chunk_hi, dyn=dyn,
                               rhs_factor=rng.standard_normal(n))
        assert lv >= 1
        assert np.array_equal(full, reuse), (chunk, chunk_hi)
        assert np.max(np.abs(full - ref)) <= 1e-12 * np.max(np.abs(ref)), (chunk, chunk_hi)


@pytest.mark.parametrize("groups", [2, 7, 64, 256])
def test_final_forest_split(groups, orc):
    """Grid launches split the final forest into subtrees per workgroup group; the
    hook checks the split (children stay in their parent's group, every node placed
    once) and solves in the launch's order: the solution is bit-identical to the
    unsplit one."""
    for bushy, geom, dT in [(True, wl.random_tree(wl.seed_for(3), depth=10)[0], 2.5e-5),
                            (False, wl.myelinated_axon()[0], 1e-5)]:
        s, bufs, H = _setup(orc, geom)
        p = _problem(geom, s, bufs, H, dT)
        rp, col, val, rhs = _system(orc, p)
        x0, _, h = tps_host(rp, col, val, rhs, 8, 8)
        cuts = []
        x1, _, _ = tps_host(rp, col, val, rhs, 8, 8, groups=groups, cut_out=cuts)
        assert np.array_equal(x0, x1)
        cut, staged = cuts[0] & 0xFFFF, cuts[0] >> 30
        if bushy:  # splits, and its top (LDS-staged) is small
            assert 0 < cut <= h and staged, (cuts, h)
        ref = orc.tree_solve(rp, col, val, rhs)
        assert np.max(np.abs(x1 - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_final_forest_split_random_forests():
    """Random forests of every shape, split for several group counts."""
    rng = np.random.default_rng(11)
    for trial in range(30):
        N = int(rng.integers(2, 3000))
        par = [-1] + [int(rng.integers(-1, i)) if rng.random() < 0.4 else i - 1
                      for i in range(1, N)]
        rp, col, val = [0], [], []
        for i in range(N):
            ent = {i: 10.0 + rng.random()}
            if par[i] >= 0:
                ent[par[i]] = -rng.random()
            for j in range(i + 1, N):
                if par[j] == i:
                    ent[j] = -rng.random()
            for j in sorted(ent):
                col.append(j)
                val.append(ent[j])
            rp.append(len(col))
        b = rng.standard_normal(N)
        x0, _, _ = tps_host(rp, col, val, b, 4)
        for groups in (2, 3, 16):
            x1, _, _ = tps_host(rp, col, val, b, 4, groups=groups)
            assert np.array_equal(x0, x1), (trial, groups)


@pytest.mark.parametrize("G", [2, 3, 8, 33])
def test_local_mode(G, orc):
    """Local mode: rows split into G contiguous groups, chains broken at group
    boundaries, every chunk and final node owned by one group except the top. The
    hook checks the plan's ownership rules and solves in the launch's order (groups
    alone, last group first; the top; groups back down), with reused factors over
    two solves; the solution matches the oracle's sequential elimination."""
    cases = [(wl.random_tree(wl.seed_for(3), depth=10)[0], 2.5e-5),
             (wl.myelinated_axon()[0], 1e-5)]
    for geom, dT in cases:
        s, bufs, H = _setup(orc, geom)
        p = _problem(geom, s, bufs, H, dT)
        rp, col, val, rhs = _system(orc, p)
        dyn = np.zeros(len(rhs), np.uint8)
        dyn[np.asarray(p.hh_elem)] = 1
        rng = np.random.default_rng(G)
        ref = orc.tree_solve(rp, col, val, rhs)
        for chunk, chunk_hi in ((8, 8), (32, 8), (128, 8)):
            tops = []
            x, lv, h = tps_host(rp, col, val, rhs, chunk, chunk_hi, dyn=dyn,
                                rhs_factor=rng.standard_normal(len(rhs)), groups=-G,
                                cut_out=tops)
            assert lv >= 1
            assert np.max(np.abs(x - ref)) <= 1e-11 * np.max(np.abs(ref)), (chunk, G)
            top_n, staged = tops[0] & 0xFFFF, tops[0] >> 30
            assert top_n > 0 and staged, tops
            # group-resident form: the same operations on per-group buffers, except
            # that chains longer than 128 nodes above level 0 (the cable's) are
            # eliminated by parallel cyclic reduction instead of a serial sweep
            xr, _, _ = tps_host(rp, col, val, rhs, chunk, chunk_hi, dyn=dyn,
                                rhs_factor=rng.standard_normal(len(rhs)),
                                groups=-(4096 + G), cut_out=tops)
            assert np.max(np.abs(xr - ref)) <= 1e-11 * np.max(np.abs(ref)), (chunk, G)
            assert np.max(np.abs(xr - x)) <= 1e-12 * np.max(np.abs(ref)), (chunk, G)
            assert 0 < tops[1] < len(rhs)


@pytest.mark.parametrize("G,chunk,chunk_hi", [(7, 6, 5), (7, 6, 8), (6, 8, 8), (8, 5, 5)])
def test_local_mode_cable_default_plans(G, chunk, chunk_hi, orc):
    """The plans the automatic choice gives the 20,202-row cable (FP64: 7 groups, 6-row
    level-0 chunks, 5-row reduced chunks; FP32: 6 groups, 8 and 8) and their neighbours:
    one level-0 chunk per thread (at most 512 per group), and the local and
    group-resident forms both match the oracle's sequential elimination."""
    geom = wl.myelinated_axon()[0]
    s, bufs, H = _setup(orc, geom)
    p = _problem(geom, s, bufs, H, 1e-5)
    rp, col, val, rhs = _system(orc, p)
    assert (len(rhs) + G - 1) // G <= 512 * chunk
    dyn = np.zeros(len(rhs), np.uint8)
    dyn[np.asarray(p.hh_elem)] = 1
    rng = np.random.default_rng(G * 100 + chunk)
    ref = orc.tree_solve(rp, col, val, rhs)
    tops = []
    x, lv, _ = tps_host(rp, col, val, rhs, chunk, chunk_hi, dyn=dyn,
                        rhs_factor=rng.standard_normal(len(rhs)), groups=-G, cut_out=tops)
    assert lv >= 1
    assert np.max(np.abs(x - ref)) <= 1e-11 * np.max(np.abs(ref))
    xr, _, _ = tps_host(rp, col, val, rhs, chunk, chunk_hi, dyn=dyn,
                        rhs_factor=rng.standard_normal(len(rhs)), groups=-(4096 + G),
                        cut_out=tops)
    assert np.max(np.abs(xr - ref)) <= 1e-11 * np.max(np.abs(ref))
    assert 0 < tops[1] < len(rhs)


@pytest.mark.parametrize("name", ["ex3_hh_y", "ex5_ct_y", "ex6_ct_tree", "ex8_neuron"])
def test_local_mode_golden_systems(name, orc):
    g = load_golden(name)
    p = problem_from_golden(g, dT=g["dT_crit"] * 10)
    rp, col, val, rhs = _system(orc, p)
    ref = orc.tree_solve(rp, col, val, rhs)
    for G in (2, 4, 7):
        x = tps_host(rp, col, val, rhs, 4, 4, groups=-G, check=False)
        if x is None:  # no reduction level: local mode does not apply
            continue
        assert np.max(np.abs(x[0] - ref)) <= 1e-12 * np.max(np.abs(ref)), (name, G)
        xr = tps_host(rp, col, val, rhs, 4, 4, groups=-(4096 + G))
        assert np.max(np.abs(xr[0] - x[0])) <= 1e-12 * np.max(np.abs(ref)), (name, G)


def test_local_mode_random_forests():
    """Random forests of every shape, local mode for several group counts (forests
    left with no reduction level are refused, as the solver then keeps the other
    launch forms)."""
    rng = np.random.default_rng(12)
    solved = 0
    for trial in range(30):
        N = int(rng.integers(2, 3000))
        par = [-1] + [int(rng.integers(-1, i)) if rng.random() < 0.4 else i - 1
                      for i in range(1, N)]
        children = [[] for _ in range(N)]
        for i in range(1, N):
            if par[i] >= 0:
                children[par[i]].append(i)
        rp, col, val = [0], [], []
        for i in range(N):
            ent = {i: 10.0 + rng.random()}
            if par[i] >= 0:
                ent[par[i]] = -rng.random()
            for j in children[i]:
                ent[j] = -rng.random()
            for j in sorted(ent):
                col.append(j)
                val.append(ent[j])
            rp.append(len(col))
        b = rng.standard_normal(N)
        x0, _, _ = tps_host(rp, col, val, b, 4)
        for G in (2, 5, 16):
            r = tps_host(rp, col, val, b, 4, groups=-G, check=False)
            if r is None:
                continue
            solved += 1
            assert np.max(np.abs(r[0] - x0)) <= 1e-12 * np.max(np.abs(x0)), (trial, G)
            xr = tps_host(rp, col, val, b, 4, groups=-(4096 + G))
            assert np.max(np.abs(xr[0] - r[0])) <= 1e-12 * np.max(np.abs(x0)), (trial, G)
    assert solved >= 30


def test_wide_chunks_pcr():
    """A long unbranched cable: above level 0 each group's chain stays one wide chunk
    (plan levels drop), and its parallel cyclic reduction, emulated with every step
    reading all rows before writing any, matches the sequential elimination."""
    rng = np.random.default_rng(5)
    for N, G in ((3000, 2), (9000, 3), (5000, 1)):
        rp, col, val = [0], [], []
        for i in range(N):
            ent = {i: 2.5 + rng.random()}
            if i > 0:
                ent[i - 1] = -rng.random()
            if i + 1 < N:
                ent[i + 1] = -rng.random()
            for j in sorted(ent):
                col.append(j)
                val.append(ent[j])
            rp.append(len(col))
        b = rng.standard_normal(N)
        x0, lv0, _ = tps_host(rp, col, val, b, 8, 8)
        xr, lv, _ = tps_host(rp, col, val, b, 8, 8, groups=-(4096 + G))
        assert lv < lv0, (lv, lv0)
        assert np.max(np.abs(xr - x0)) <= 1e-12 * np.max(np.abs(x0)), (N, G)
