"""Seeded synthetic neurites for the benchmark configurations.

Each builder returns (Geometry, RunSpec): raw geometry in the reference's
element model (element 0 is the fictitious first element, every branch end is a
sealed terminal, numbering as the reference builders use it: a segment's
compartments are consecutive, the right daughter continues the numbering and
the left daughter subtree follows the whole right subtree) and the run
parameters (dT, duration, stimulus window, measurement element).

RNG: SplitMix64 with 53-bit uniforms; seed = 0x4E52540000000000 + config index +
0x100 * tree index.
"""
from dataclasses import dataclass

import numpy as np

import neurite_hip as nh

SEED_BASE = 0x4E52540000000000
MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def integers(self, lo, hi):  # inclusive
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo

    def normal(self):  # Box-Muller
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def seed_for(config, tree=0):
    return SEED_BASE + config + 0x100 * tree


def material(kind, rho_a=0.354, rest_pot=-0.065, per_m=4e-11, rho_m=2.5e9, thickness=4e-9,
             layers_my=0):
    """A material in the reference's parameterisation. Defaults: squid-axon cable
    (Ri = 35.4 ohm cm, Cm = 1 uF/cm^2, Vr = -65 mV); HH densities 120/36/0.3
    mS/cm^2 (sigma / thickness), E_Na = Vr+115 mV, E_K = Vr-12 mV."""
    return nh.Material(kind=kind, layers_my=layers_my, rho_a=rho_a, rho_m=rho_m, per_m=per_m,
                       thickness=thickness, rest_pot=rest_pot, thickness_my=18e-9,
                       per_my=1.08e-10, ro_my=4.44e6, sigma_Na=4.8e-6, sigma_K=1.44e-6,
                       sigma_L=1.2e-8, dE_Na=0.115, dE_K=-0.012, n_damage=2.0)


@dataclass
class RunSpec:
    dT: float            # time step [s] (None -> dt_factor * critical step)
    tspan: float         # simulated time [s]
    init_time: float     # stimulus window [s]
    end_input: float
    mes: int             # measurement element
    dt_factor: float = 1.0

    def steps(self, dT):
        return int(self.tspan / dT)


class _Builder:
    """Accumulates elements in reference numbering."""

    def __init__(self):
        self.mother, self.dr, self.dl, self.br = [], [], [], []
        self.diam, self.dx0, self.eps, self.inp, self.mat = [], [], [], [], []
        self.terminals = [0]

    def add(self, mother, diam, dx, mat, eps=0.0):
        j = len(self.mother)
        self.mother.append(mother)
        self.dr.append(j + 1)
        self.dl.append(0)
        self.br.append(0)
        self.diam.append(diam)
        self.dx0.append(dx)
        self.eps.append(eps)
        self.inp.append(0.0)
        self.mat.append(mat)
        return j

    def seal(self, j):
        self.dr[j] = j
        self.terminals.append(j)

    def geometry(self, materials, myelinate, E_MAX=1.0):
        return nh.Geometry(self.mother, self.dr, self.dl, self.br, self.diam, self.dx0,
                           self.eps, self.inp, self.mat, materials, self.terminals, myelinate,
                           E_MAX)


def squid_axon(n_comp=1000, length=1e-2, diameter=1e-6, dT=1e-6, tspan=10e-3, amp=10e-9,
               pulse=0.5e-3):
    """Config 1: unbranched HH cable, 1 cm, d = 1 um, dx = 10 um (1,000 compartments
    + 2 ghosts), standard HH, Vr = -65 mV, 10 nA for 0.5 ms at compartment 1."""
    b = _Builder()
    dx = length / n_comp
    b.add(-1, diameter, dx, 0)
    for _ in range(n_comp):
        b.add(len(b.mother) - 1, diameter, dx, 1)
    last = b.add(len(b.mother) - 1, diameter, dx, 0)
    b.seal(last)
    b.inp[1] = amp
    geom = b.geometry([material(0), material(1)], myelinate=0)
    return geom, RunSpec(dT, tspan, 0.0, pulse, mes=n_comp // 2)


def myelinated_axon(nodes=200, inter=100, node_dx=1e-6, inter_dx=10e-6, diameter=10e-6,
                    dT=10e-6, tspan=20e-3, amp=10e-9, pulse=0.5e-3):
    """Config 2: 200 nodes of Ranvier (1 um, HH) alternating with 1 mm internodes
    (passive, Cm = 0.005 uF/cm^2, gL = 0.001 mS/cm^2; 100 x 10 um), d = 10 um,
    N = 20,202; stimulus 10 nA for 0.5 ms at the first node."""
    node = material(1)
    # Cm = per_m/thickness = 5e-5 F/m^2; gL = 1/(rho_m*thickness) = 0.01 S/m^2
    internode = material(0, per_m=5e-5 * 4e-9, rho_m=1.0 / (0.01 * 4e-9))
    b = _Builder()
    b.add(-1, diameter, node_dx, 0)
    for _ in range(nodes):
        b.add(len(b.mother) - 1, diameter, node_dx, 1)
        for _ in range(inter):
            b.add(len(b.mother) - 1, diameter, inter_dx, 0)
    last = b.add(len(b.mother) - 1, diameter, inter_dx, 0)
    b.seal(last)
    b.inp[1] = amp
    geom = b.geometry([internode, node], myelinate=0)
    N = len(b.mother)
    return geom, RunSpec(dT, tspan, 0.0, pulse, mes=N // 2)


def random_tree(seed, depth=14, len_range=(20e-6, 200e-6), root_diam=4e-6, dx=2e-6,
                strain=(0.0, 0.3), soma_len=20e-6, amp=1e-9, pulse=0.5e-3, dT=25e-6,
                tspan=50e-3, hh_soma=True):
    """Config 3 style tree: an HH soma compartment (the root segment) with binary
    dendritic subtrees down to `depth` levels (2**depth - 1 segments in all),
    segment lengths U[len_range], child diameter = parent * 2**(-2/3) (Rall 3/2),
    compartments of about dx (at least 2 per segment), per-segment axial strain
    U[strain]; stimulus at the soma."""
    rng = SplitMix64(seed)
    b = _Builder()
    b.add(-1, root_diam, soma_len, 0)
    soma = b.add(0, root_diam, soma_len, 1 if hh_soma else 0)
    b.inp[soma] = amp
    rall = 2.0 ** (-2.0 / 3.0)

    def segment(parent, diam):
        L = rng.uniform(*len_range)
        eps = rng.uniform(*strain)
        n = max(2, int(round(L / dx)))
        prev, first = parent, -1
        for _ in range(n):
            prev = b.add(prev, diam, L / n, 0, eps)
            if first < 0:
                first = prev
        return first, prev

    # Depth-first, right daughter (numbering continues) before the left daughter
    # (numbered after the whole right subtree), as the reference builders number.
    work = [("children", soma, 0, root_diam)]
    while work:
        kind, node, level, diam = work.pop()
        if kind == "children":
            if level >= depth - 1:
                b.seal(node)
                continue
            work.append(("left", node, level + 1, diam * rall))
            work.append(("right", node, level + 1, diam * rall))
            continue
        first, last = segment(node, diam)
        if kind == "right":
            b.dr[node] = first
        else:
            b.dl[node] = first
            b.br[node] = 1
        work.append(("children", last, level, diam))
    geom = b.geometry([material(0), material(1)], myelinate=0)
    return geom, RunSpec(dT, tspan, 0.0, pulse, mes=soma)


def _tree_of_size(seed, n_target, depth, dx, strain, amp, dT, tspan):
    """A random_tree whose compartment count is about n_target: depth is capped so
    that every segment keeps at least two compartments, and segment lengths are
    uniform in [0.5, 1.5] x the mean that n_target implies."""
    depth = int(max(2, min(depth, np.floor(np.log2(n_target / 2.0 + 1.0)))))
    segs = 2 ** depth - 1
    mean_len = max(2.0, n_target / segs) * dx
    return random_tree(seed, depth=depth, len_range=(0.5 * mean_len, 1.5 * mean_len), dx=dx,
                       strain=strain, amp=amp, dT=dT, tspan=tspan)


def sweep_tree(index, n_range=(500, 50000), depth_range=(4, 12), strain=(0.0, 0.3),
               amp=(1e-9, 20e-9), dx=2e-6, dT=25e-6, steps=200):
    """Config 4 member `index` (of a batched parameter sweep): compartment count
    log-uniform in n_range, depth uniform in depth_range (capped as in
    _tree_of_size), one axial strain per tree uniform in `strain`, stimulus
    amplitude uniform in `amp`; implicit dt = 25 us, 200 steps."""
    rng = SplitMix64(seed_for(4, index))
    n_target = float(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
    depth = rng.integers(*depth_range)
    eps = rng.uniform(*strain)
    a = rng.uniform(*amp)
    return _tree_of_size(seed_for(4, index) ^ 0x5A5A5A5A, n_target, depth, dx, (eps, eps), a,
                         dT, steps * dT)


def served_request_target(index, median=20000.0, sigma=1.0, cap=2_000_000):
    """The compartment count request `index` of config 5 aims at (its first draw),
    without building the tree."""
    rng = SplitMix64(seed_for(5, index))
    return min(float(cap), median * float(np.exp(sigma * rng.normal())))


def served_request(index, median=20000.0, sigma=1.0, cap=2_000_000, duration=(5e-3, 50e-3),
                   rate=200.0, dx=2e-6, dT=25e-6):
    """Config 5 request `index`: compartment count log-normal (median, sigma),
    capped; simulated duration uniform in `duration`; arrival gap exponential with
    `rate` per second. Returns (geometry, run spec, inter-arrival gap [s])."""
    rng = SplitMix64(seed_for(5, index))
    n_target = min(float(cap), median * float(np.exp(sigma * rng.normal())))
    tspan = rng.uniform(*duration)
    gap = -np.log(max(rng.uniform(), 1e-300)) / rate
    depth = int(np.clip(np.floor(np.log2(n_target / 50.0 + 1.0)), 2, 14))
    eps = rng.uniform(0.0, 0.3)
    geom, run = _tree_of_size(seed_for(5, index) ^ 0x5A5A5A5A, n_target, depth, dx, (eps, eps),
                              10e-9, dT, tspan)
    return geom, run, gap
