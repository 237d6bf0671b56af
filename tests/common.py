"""Shared helpers for the tests: golden fixtures -> problems / geometries."""
import gzip
import json
import os

import numpy as np

import neurite_hip as nh

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN_DIR = os.path.join(HERE, "golden")
GOLDEN = sorted(f[:-8] for f in os.listdir(GOLDEN_DIR) if f.endswith(".json.gz"))

# Tolerance of the FP64 plain form against the reference trajectory: 1e-6 mV.
TOL_V = 1e-9


def load_golden(name):
    with gzip.open(os.path.join(GOLDEN_DIR, name + ".json.gz"), "rt") as f:
        return json.load(f)


def problem_from_golden(g, dT=None):
    """The derived arrays the reference GPU constructors would copy."""
    return nh.Problem(
        g["dT"] if dT is None else dT, g["mother"], g["dr"], g["dl"], g["branching"],
        g["terminals"], g["hh_elem"], potential0=g["potential0"], potential1=g["potential1"],
        dx=g["dx"], c_m=g["c_m"], r_a=g["r_a"], W=g["W"], K=g["K"], input=g["input"],
        hh_m=g["hh_m"], hh_h=g["hh_h"], hh_n=g["hh_n"], E_Na=g["E_Na"], E_K=g["E_K"],
        E_L=g["E_L"], g_Na=g["g_Na"], g_K=g["g_K"], G_L=g["G_L"], ls_Na=g["ls_Na"],
        ls_K=g["ls_K"], rest_pot=g["hh_rest"])


def reference_materials():
    """Defaults of the reference element classes (discretization.cpp:19-97)."""
    base = dict(rho_a=1.87, rho_m=2.5e9, per_m=4e-11, thickness=4 * pow(10, -9),
                rest_pot=-0.0655, layers_my=45, thickness_my=18e-9, per_my=1.08e-10,
                ro_my=4.44e6, sigma_Na=4.8e-6, sigma_K=1.44e-6, sigma_L=1.2e-8, dE_Na=0.115,
                dE_K=-0.012, n_damage=2.0)
    ct = nh.Material(kind=0, **base)
    hh = nh.Material(kind=1, **base)
    return [ct, hh]


def geometry_from_golden(g):
    is_hh = np.asarray(g["is_hh"], dtype=np.int32)
    return nh.Geometry(g["mother"], g["dr"], g["dl"], g["branching"], g["diameter0"], g["dx0"],
                       g["epsilon"], g["input"], is_hh, reference_materials(), g["terminals"],
                       g["myelinate"], g["E_MAX"])


def max_abs_diff(a, b, mask=None):
    a, b = np.asarray(a), np.asarray(b)
    d = np.abs(a - b)
    if mask is not None:
        d = d[mask]
    return float(np.max(d)) if d.size else 0.0


def ghost_mask(problem):
    """Non-ghost elements: interior 1..N-2 that are not terminals."""
    m = np.ones(problem.N, dtype=bool)
    m[0] = False
    m[-1] = False
    m[np.asarray(problem.terminals)] = False
    return m
