"""Pin the oracle where the reference build cannot: known answers, closed forms,
independent dense constructions and exact solves (CPU only)."""
import math

import numpy as np
import pytest

import neurite_hip as nh
import workloads as wl
from common import geometry_from_golden, ghost_mask, load_golden, problem_from_golden


def _setup(orc, geom):
    H = nh.count_hh(geom)
    s, bufs = nh.setup_out_buffers(geom.N, H)
    gs = geom.struct()
    orc.setup(gs, s)
    return s, bufs, H


def _problem(geom, s, bufs, H, dT):
    rest = bufs["rest"]
    return nh.Problem(dT, geom.mother, geom.dr, geom.dl, geom.branching, geom.terminals,
                      bufs["hh_elem"][:H], potential0=rest, potential1=rest, dx=bufs["dx"],
                      c_m=bufs["c_m"], r_a=bufs["r_a"], W=bufs["W"], K=bufs["K"],
                      input=geom.input, rest_pot=bufs["hh_rest"][:H],
                      **{k: bufs[k][:H] for k in ["hh_m", "hh_h", "hh_n", "E_Na", "E_K", "E_L",
                                                   "g_Na", "g_K", "G_L", "ls_Na", "ls_K"]})


def test_hh_steady_state_known_answers(orc):
    """At rest (u = 0 mV) the HH steady-state gates are the textbook values
    m = 0.0529, h = 0.5961, n = 0.3177 (Hodgkin & Huxley 1952, rates at V = 0),
    and E_L balances the resting currents (-54.4 mV for Vr = -65 mV)."""
    geom, _ = wl.squid_axon(n_comp=4)
    s, bufs, H = _setup(orc, geom)
    assert H == 4
    assert bufs["hh_m"][0] == pytest.approx(0.0529325, abs=5e-7)
    assert bufs["hh_h"][0] == pytest.approx(0.5961208, abs=5e-7)
    assert bufs["hh_n"][0] == pytest.approx(0.3176769, abs=5e-7)
    # exact balance of the resting currents; HH quote Vr + 10.613 mV = -54.387 mV
    # from rounded gates, the configuration states -54.4 mV
    assert bufs["E_L"][0] * 1e3 == pytest.approx(-54.4, abs=0.02)
    assert bufs["E_Na"][0] * 1e3 == pytest.approx(50.0, abs=1e-9)
    assert bufs["E_K"][0] * 1e3 == pytest.approx(-77.0, abs=1e-9)
    # densities: g / (pi d) = sigma / t = 120, 36, 0.3 mS/cm^2 = 1200, 360, 3 S/m^2
    d = geom.diameter[1]
    assert bufs["g_Na"][0] / (3.14159265 * d) == pytest.approx(1200.0, rel=1e-12)
    assert bufs["g_K"][0] / (3.14159265 * d) == pytest.approx(360.0, rel=1e-12)
    assert bufs["G_L"][0] / (3.14159265 * d) == pytest.approx(3.0, rel=1e-12)


def test_critical_step_closed_form(orc):
    """init.cpp:141 for HH: r_a c_m dx^2 / 2; config 1 gives 0.708 us."""
    geom, _ = wl.squid_axon()
    s, bufs, H = _setup(orc, geom)
    ra = 4 * 0.354 / (3.14159265 * 1e-12)
    cm = 4e-11 * 3.14159265 * 1e-6 / 4e-9
    assert s.dT_crit == pytest.approx(ra * cm * (1e-5) ** 2 / 2, rel=1e-14)
    assert s.dT_crit == pytest.approx(0.7079e-6, rel=1e-3)


def test_passive_rest_is_stationary(orc):
    """W*Vr + K = 0 for cable elements: with no stimulus V stays at rest."""
    g = load_golden("ex6_ct_tree")
    p = problem_from_golden(g)
    ex = orc.Explicit(p)
    ex.run(1, 2000)
    v = ex.snapshot()
    assert np.max(np.abs(v - p.potential0)) < 1e-15


def _passive_cable(n, I=1e-10):
    geom, run = wl.squid_axon(n_comp=n, amp=I)
    mats = [wl.material(0), wl.material(0)]
    return nh.Geometry(geom.mother, geom.dr, geom.dl, geom.branching, geom.diameter, geom.dx0,
                       geom.epsilon, geom.input, geom.material, mats, geom.terminals, 0)


def _discrete_steady_state(bufs, geom):
    """Independent dense construction of the sealed-end passive steady state:
    g_a (V_{j-1} - V_j) + g_a' (V_{j+1} - V_j) - (dx/r_m)(V_j - Vr) + I_j = 0."""
    N = geom.N
    ga = 1.0 / (bufs["r_a"] * bufs["dx"])
    gm = bufs["dx"] / bufs["r_m"]
    A = np.zeros((N - 2, N - 2))
    rhs = np.zeros(N - 2)
    for k, j in enumerate(range(1, N - 1)):
        A[k, k] -= gm[j]
        rhs[k] -= gm[j] * bufs["rest"][j] + geom.input[j]
        if j > 1:                 # mother coupling (element 0 mirrors element 1)
            A[k, k] -= ga[j]
            A[k, k - 1] += ga[j]
        if j < N - 2:             # right daughter (element N-1 mirrors N-2)
            A[k, k] -= ga[j + 1]
            A[k, k + 1] += ga[j + 1]
    return np.linalg.solve(A, rhs)


def test_explicit_reaches_discrete_steady_state(orc):
    """Explicit FTCS on a passive sealed cable with constant injection converges
    to the steady state of an independently built linear system, and that state
    matches the continuous cable's closed form to discretisation accuracy."""
    geom = _passive_cable(60)
    s, bufs, H = _setup(orc, geom)
    p = _problem(geom, s, bufs, H, 0.9 * s.dT_crit)
    ex = orc.Explicit(p)
    ex.run(1, 60000, 1, 10 ** 9)
    v = ex.snapshot()[1:-1]
    ref = _discrete_steady_state(bufs, geom)
    assert np.max(np.abs(v - ref)) < 1e-9 * np.max(np.abs(ref - bufs["rest"][1]))
    # continuous sealed cable, current into x = 0: V(x)-Vr = I r_a lam cosh((L-x)/lam)/sinh(L/lam)
    ra, rm, dx = bufs["r_a"][1], bufs["r_m"][1], bufs["dx"][1]
    lam = math.sqrt(rm / ra)
    L = dx * 60
    x = (np.arange(60) + 0.5) * dx
    cont = 1e-10 * ra * lam * np.cosh((L - x) / lam) / np.sinh(L / lam)
    np.testing.assert_allclose(v - bufs["rest"][1], cont, rtol=2e-3)


def _dense(rowptr, col, val, N):
    A = np.zeros((N, N))
    for i in range(N):
        for k in range(rowptr[i], rowptr[i + 1]):
            A[i, col[k]] += val[k]
    return A


def _first_principles(p, W):
    """Backward-Euler matrix from the cable equation: for a non-ghost row j,
    A[j,j] = c_m dx/dT - W dx + sum of the axial conductances to its neighbours,
    A[j,nb] = -conductance (1/(r_a dx) of the daughter side, of j for the mother),
    sealed ends drop the coupling to a terminal (its V mirrors j), ghosts are 1."""
    N = p.N
    A = np.zeros((N, N))
    term = set(int(t) for t in p.terminals)
    g = 1.0 / (p.r_a * p.dx)
    for j in range(N):
        if j in (0, N - 1) or j in term:
            A[j, j] = 1.0
            continue
        A[j, j] = p.c_m[j] * p.dx[j] / p.dT - W[j] * p.dx[j]
        nbrs = [(int(p.mother[j]), g[j])]
        nbrs.append((int(p.dr[j]), g[p.dr[j]]))
        if p.branching[j]:
            nbrs.append((int(p.dl[j]), g[p.dl[j]]))
        for nb, gg in nbrs:
            if nb == j:
                continue
            if nb in term or nb == 0:     # sealed end / fictitious element: mirrored
                continue
            A[j, j] += gg
            A[j, nb] -= gg
    return A


@pytest.mark.parametrize("name", ["ex4_ct_cable", "ex5_ct_y", "ex6_ct_tree", "ex8_neuron"])
def test_assembly_matches_first_principles(name, orc):
    g = load_golden(name)
    p = problem_from_golden(g, dT=g["dT_crit"] * 10)
    ref = _first_principles(p, p.W)
    scale = np.max(np.abs(ref[1:-1, 1:-1]))
    for var in (orc.VAR_GPU, orc.VAR_CPU):
        im = orc.Implicit(p, variant=var, method=orc.M_DIRECT)
        im.update(False)
        rp, col, val, rhs = im.matrix()
        A = _dense(rp, col, val, p.N)
        assert np.max(np.abs(A - ref)) <= 1e-13 * scale, var


@pytest.mark.parametrize("name", ["ex3_hh_y", "ex6_ct_tree", "ex8_neuron"])
def test_exact_solve_matches_numpy(name, orc):
    g = load_golden(name)
    p = problem_from_golden(g, dT=g["dT_crit"] * 10)
    im = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT)
    im.update(True)
    rp, col, val, rhs = im.matrix()
    x = orc.tree_solve(rp, col, val, rhs)
    xd = np.linalg.solve(_dense(rp, col, val, p.N), rhs)
    assert np.max(np.abs(x - xd)) <= 1e-12 * np.max(np.abs(xd))


@pytest.mark.parametrize("method", ["bicgstab", "gmres"])
def test_krylov_restatements_converge_to_exact(method, orc):
    g = load_golden("ex6_ct_tree")
    p = problem_from_golden(g, dT=g["dT_crit"] * 10)
    if method == "bicgstab":
        kr = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_BICGSTAB, tol=1e-13, tol_mode=1,
                          x0_mode=0, check_every=1)
        ex = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT)
    else:
        kr = orc.Implicit(p, variant=orc.VAR_CPU, method=orc.M_GMRES, restart=150, tol=1e-18)
        ex = orc.Implicit(p, variant=orc.VAR_CPU, method=orc.M_DIRECT)
    for i in range(1, 21):
        for s in (kr, ex):
            s.update(i < 10)
            assert s.calculate() == 0
    m = np.ones(p.N, bool)
    m[[0, -1]] = False
    m[np.asarray(p.terminals)] = False
    assert np.max(np.abs(kr.snapshot() - ex.snapshot())[m]) < 1e-10


def test_variants_agree_on_trajectory(orc):
    """CPU createMatrix vs GPU assembly: same physics, FP op order only."""
    g = load_golden("ex3_hh_y")
    p = problem_from_golden(g, dT=g["dT_crit"] * 50)
    a = orc.Implicit(p, variant=orc.VAR_CPU, method=orc.M_DIRECT)
    b = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT)
    for i in range(1, 101):
        for s in (a, b):
            s.update(i < 60)
            assert s.calculate() == 0
    m = np.ones(p.N, bool)
    m[[0, -1]] = False
    m[np.asarray(p.terminals)] = False
    assert np.max(np.abs(a.snapshot() - b.snapshot())[m]) < 1e-12


def test_tree_generator_numbering():
    geom, run = wl.random_tree(wl.seed_for(3), depth=6)
    N = geom.N
    mo, dr, dl, br = geom.mother, geom.dr, geom.dl, geom.branching
    assert mo[0] == -1 and all(mo[j] < j for j in range(1, N))
    term = set(geom.terminals.tolist())
    assert len(term) == 1 + 2 ** 5 and N - 1 in term
    for j in range(1, N):
        if j in term:
            assert dr[j] == j
        else:
            assert dr[j] == j + 1 and mo[j + 1] == j
        if br[j]:
            assert mo[dl[j]] == j and dl[j] > j


def test_bicgstab_from_rhs_converges_on_hh_y(orc):
    """The configuration smoke() runs on the GPU: BiCGSTAB (the reference GPU solver's
    method) from x0 = b at relative 1e-12, tested every iteration, on the HH Y-branch at
    100x the critical step. Every step converges within max_iter = N, and V stays within
    1e-9 V of the exact tree solve after 50 steps."""
    g = load_golden("ex3_hh_y")
    p = problem_from_golden(g, dT=g["dT_crit"] * 100)
    ex = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT)
    ex.run(1, 50, 1, 50)
    it = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_BICGSTAB, tol=1e-12, tol_mode=1,
                      x0_mode=1, check_every=1)
    for step in range(1, 51):
        it.update(True)
        assert it.calculate() == 0, f"step {step}"
        assert it.stats()[0] < p.N
    m = ghost_mask(p)
    assert np.max(np.abs(it.snapshot() - ex.snapshot())[m]) <= 1e-9
