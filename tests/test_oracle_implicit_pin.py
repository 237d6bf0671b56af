"""Pin the oracle's implicit step (CPU only) where no reference build can: the
reference implicit CPU solver needs MKL headers and the GPU one CUDA, so neither
runs here. Three checks fixed by the mathematics, each sensitive to the plausible
mistakes of a restatement (a dropped K dx or stimulus term, a sign, the wrong time
unit in the gating, a wrong c_m dx / dT):

1. rest invariant: with no stimulus V stays at rest, since W Vr + K = 0 for cable
   elements and E_L balances the resting HH currents (hhIconds);
2. first principles: the backward-Euler system of the cable equation satisfies
   A V - b = -F(V) at the state it is built from, where F is the membrane + axial
   current of the explicit step (solverEx-cpu.cpp:92-108), taken from the explicit
   oracle, which is bit-identical to the reference build; and A equals the dense
   first-principles matrix with W from that explicit gating;
3. time refinement: backward Euler converges at first order in dT to the explicit
   trajectory of the same problem (error ratio ~2 per halving of dT).
Both matrix variants are checked: GPU assembly + Update_Mat/Update_B
(solverIm-gpu.cpp:103-171,199-369) and CPU createMatrix (solverIm-cpu.cpp:29-183).
"""
import copy

import numpy as np
import pytest

from common import GOLDEN, ghost_mask, load_golden, problem_from_golden

# ex1_only_hh (N = 3) has no implicit form: row 1 is both the first and the last
# interior row (the solvers require N >= 4)
HH_SCEN = ["ex2_hh_axon", "ex3_hh_y", "ex7_hh_axon_strain"]
IMPLICIT_SCEN = [n for n in GOLDEN if n != "ex1_only_hh"]
ULP_V = 2.0 ** -56  # ulp of |V| < 0.125 V


def _dense(rowptr, col, val, N):
    A = np.zeros((N, N))
    for i in range(N):
        for k in range(rowptr[i], rowptr[i + 1]):
            A[i, col[k]] += val[k]
    return A


def _variants(orc):
    return [(orc.VAR_GPU, "gpu-assembly"), (orc.VAR_CPU, "cpu-createMatrix")]


@pytest.mark.parametrize("name", IMPLICIT_SCEN)
def test_implicit_rest_invariant(name, orc):
    """No stimulus, 200 backward-Euler steps at 10x the critical step: every
    non-ghost element stays at rest to 1e-13 V (measured <= 1.3e-14 V)."""
    g = load_golden(name)
    p = problem_from_golden(g, dT=g["dT_crit"] * 10)
    m = ghost_mask(p)
    for var, label in _variants(orc):
        im = orc.Implicit(p, variant=var, method=orc.M_DIRECT)
        for _ in range(200):
            im.update(False)
            assert im.calculate() == 0
        drift = float(np.max(np.abs(im.snapshot() - p.potential0)[m]))
        assert drift <= 1e-13, (label, drift)


def test_rest_invariant_has_teeth(orc):
    """The invariant is sensitive: E_L (hence K dx) off by 1e-6 relative on one HH
    element (a mistake planted in the oracle's input) moves V far beyond 1e-13 V."""
    g = load_golden("ex3_hh_y")
    p = problem_from_golden(g, dT=g["dT_crit"] * 10)
    q = copy.copy(p)
    q.E_L = p.E_L.copy()
    q.E_L[3] *= 1 + 1e-6
    im = orc.Implicit(q, variant=orc.VAR_GPU, method=orc.M_DIRECT)
    for _ in range(200):
        im.update(False)
        im.calculate()
    assert np.max(np.abs(im.snapshot() - p.potential0)[ghost_mask(p)]) > 1e-11


def _excited_state(g, steps):
    """A state off rest: the explicit oracle run through the stimulus onset. V and
    the gates are taken from it (V satisfies the sealed-end copies)."""
    import oracle
    p = problem_from_golden(g)
    ex = oracle.Explicit(p)
    ex.run(1, steps, 1, steps)
    V = ex.snapshot()
    m, h, n = ex.gates()
    q = copy.copy(p)
    q.potential0, q.potential1 = V.copy(), V.copy()
    q.hh_m, q.hh_h, q.hh_n = m, h, n
    # really off rest: somewhere more than 0.1 mV away
    assert np.max(np.abs(V - p.potential0)[ghost_mask(p)]) > 1e-4
    return q


@pytest.mark.parametrize("name", HH_SCEN + ["ex6_ct_tree"])
@pytest.mark.parametrize("dtf", [1.0, 10.0])
@pytest.mark.parametrize("active", [True, False])
def test_backward_euler_system_first_principles(name, dtf, active, orc):
    g = load_golden(name)
    q = _excited_state(g, 3000)
    q.dT = g["dT_crit"] * dtf
    N = q.N
    m = ghost_mask(q)
    # F(V): the explicit step's current per element, with W, K from its own gating
    ex = orc.Explicit(q)
    ex.update(active)
    gm, gh, gn = ex.gates()
    ex.calculate()
    F = (ex.snapshot() - q.potential0) * (q.c_m * q.dx) / q.dT
    # W of every element after that gating (HH: textbook conductances from the gates)
    W = q.W.copy()
    hh = np.asarray(q.hh_elem)
    W[hh] = -(q.g_Na * gm ** 3 * gh + q.g_K * gn ** 4 + q.G_L)
    Aref = _first_principles(q, W)
    cdt = q.c_m * q.dx / q.dT
    tol_F = 64 * ULP_V * float(np.max(cdt)) + 1e-12 * float(np.max(np.abs(F[m])))
    for var, label in _variants(orc):
        im = orc.Implicit(q, variant=var, method=orc.M_DIRECT)
        im.update(active)
        rp, col, val, rhs = im.matrix()
        A = _dense(rp, col, val, N)
        scale = float(np.max(np.abs(Aref[1:-1, 1:-1])))
        dA = np.abs(A - Aref)[m]
        assert float(np.max(dA)) <= 1e-13 * scale, (label, float(np.max(dA)) / scale)
        res = A @ q.potential0 - rhs
        err = float(np.max(np.abs(res + F)[m]))
        assert err <= tol_F, (label, err, tol_F)


def _first_principles(p, W):
    """Dense backward-Euler matrix from the cable equation: row j of a non-ghost
    element holds c_m dx/dT - W dx + the axial conductances to its neighbours on the
    diagonal and -conductance at each neighbour (1/(r_a dx) of the daughter for a
    daughter, of j itself for the mother); a neighbour that is a sealed end
    (terminal) or the fictitious element 0 mirrors j, so its coupling drops out;
    ghost rows are identity rows."""
    N = p.N
    A = np.zeros((N, N))
    term = set(int(t) for t in p.terminals)
    g = 1.0 / (p.r_a * p.dx)
    for j in range(N):
        if j in (0, N - 1) or j in term:
            A[j, j] = 1.0
            continue
        A[j, j] = p.c_m[j] * p.dx[j] / p.dT - W[j] * p.dx[j]
        nbrs = [(int(p.mother[j]), g[j]), (int(p.dr[j]), g[p.dr[j]])]
        if p.branching[j]:
            nbrs.append((int(p.dl[j]), g[p.dl[j]]))
        for nb, gg in nbrs:
            if nb == j or nb in term or nb == 0:
                continue
            A[j, j] += gg
            A[j, nb] -= gg
    return A


def _run(orc, g, dT, T, t_on, t_off, implicit):
    p = problem_from_golden(g, dT=dT)
    s = orc.Implicit(p, variant=orc.VAR_GPU, method=orc.M_DIRECT) if implicit else orc.Explicit(p)
    n = int(round(T / dT))
    for i in range(1, n + 1):
        s.update(t_on <= i * dT <= t_off)
        s.calculate()
    return s.snapshot()


@pytest.mark.parametrize("name", ["ex2_hh_axon", "ex3_hh_y", "ex7_hh_axon_strain"])
def test_backward_euler_converges_first_order(name, orc):
    """Implicit oracle at dT = 0.8, 0.4, 0.2 x critical against the explicit oracle
    at 0.025 x critical, after 640 critical steps with a stimulus over the first
    160: errors halve with dT (measured ratios 2.06-2.07, then 2.14-2.16 as the
    explicit reference's own O(dT) error begins to show)."""
    g = load_golden(name)
    base = 0.8 * g["dT_crit"]
    T, t_on, t_off = base * 800, base, base * 200
    ref = _run(orc, g, base / 32, T, t_on, t_off, implicit=False)
    m = ghost_mask(problem_from_golden(g))
    errs = [float(np.max(np.abs(_run(orc, g, base / f, T, t_on, t_off, True) - ref)[m]))
            for f in (1, 2, 4)]
    assert errs[0] > 1e-7  # the trajectory moved (a response, not a rest state)
    for a, b in zip(errs, errs[1:]):
        assert 1.8 <= a / b <= 2.4, errs
