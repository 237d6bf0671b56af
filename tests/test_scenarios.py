"""Scenario builders, loaders and the command line of the drop-in driver
(nrt_scenario_*, CPU only).

Pinned by the reference itself: every golden fixture was dumped by the reference's
own builders (space.cpp) and main (neurite.cpp) on the example command lines, so
the raw geometry (topology, lengths, diameters, strains, stimuli, terminals), the
measurement point, the time span and, through the oracle's setup chain, the time
step must all come out bit-identical. The input files of examples 7 and 8 are the
reference's own data files (tests/golden/inputs).
"""
import os

import numpy as np
import pytest

import neurite_hip as nh
from common import load_golden

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "inputs")
NAMES = ["ex1_only_hh", "ex2_hh_axon", "ex3_hh_y", "ex4_ct_cable", "ex5_ct_y", "ex6_ct_tree",
         "ex7_hh_axon_strain", "ex8_neuron"]


def _build(g, extra=()):
    sc = g["scenario"]
    wd = os.path.join(INPUTS, sc["example"])
    return nh.scenario(sc["cli"].split() + list(extra), simulation=sc["simulation"],
                       working_directory=wd)


@pytest.mark.parametrize("name", NAMES)
def test_builder_geometry_matches_reference(name):
    g = load_golden(name)
    s = _build(g)
    geo = s["geometry"]
    assert geo.N == g["N"]
    for mine, ref in [(geo.mother, "mother"), (geo.dr, "dr"), (geo.dl, "dl"),
                      (geo.branching, "branching"), (geo.material, "is_hh"),
                      (geo.terminals, "terminals")]:
        assert np.array_equal(mine, np.asarray(g[ref])), ref
    for mine, ref in [(geo.diameter, "diameter0"), (geo.dx0, "dx0"), (geo.epsilon, "epsilon"),
                      (geo.input, "input")]:
        assert np.array_equal(mine, np.asarray(g[ref], dtype=np.float64)), ref
    assert geo.myelinate == g["myelinate"]
    assert geo.E_MAX == g["E_MAX"]
    # run values the main derives from the command line and the builder
    assert s["mes_point"] == g["dist_mes_points"]
    assert s["dist_mes"] == g["dist_mes"]
    assert s["tspan"] == g["tspan"]
    assert s["dt_factor"] == g["dt_factor"]
    assert s["init_time"] == g["init_time"]
    assert s["end_input"] == g["end_input"]
    assert s["epsilon"] == g["in_epsilon"]
    assert s["explicit_mode"] == 1 and s["implicit_mode"] == 0
    assert s["out_path"] == "o.dat" and s["log_path"] == "o.log"


@pytest.mark.parametrize("name", NAMES)
def test_time_step_and_step_count_match_reference(name, orc):
    """The oracle's setup chain on the built geometry gives the reference's dT;
    the step count and the trace interval follow (neurite.cpp:171-182)."""
    from test_oracle_math import _setup
    g = load_golden(name)
    s = _build(g)
    out, bufs, H = _setup(orc, s["geometry"])
    dT = out.dT_crit * s["dt_factor"]
    assert dT == g["dT"]
    steps = int(s["tspan"] / dT)
    assert steps == g["numStepsT"]
    assert int(steps / 1000) == g["interval"]
    assert np.array_equal(bufs["dx"], np.asarray(g["dx"]))


def test_mechanical_model_strain():
    """Example 7: the macroscopic strain from input_data.txt (case 3, K = 0.5,
    strains 1.0 / 0.25488959 at t0, lower bound 0.11195694) scales the
    measurement distance: K eps + (eps_macro_t0 - K eps_membrane_t0)."""
    g = load_golden("ex7_hh_axon_strain")
    s = _build(g)
    expect = 0.5 * 0.11195694 + (1.0 - 0.5 * 0.25488959)
    assert s["epsilon_macro"] == pytest.approx(expect, rel=1e-15)
    assert s["dist_mes"] == 0.010 * (1 + s["epsilon_macro"])
    # strain outside the model's range is rejected, as the reference exits
    with pytest.raises(nh.NrtError):
        nh.scenario("e Tt 0.1 str 0.5 out o.dat log o.log".split(), simulation="HH_axon",
                    working_directory=os.path.join(INPUTS, "example_7"))
    # and a strain without the data file is an error too
    with pytest.raises(nh.NrtError, match="input_data"):
        nh.scenario("e str 0.2 out o.dat log o.log".split(), simulation="HH_axon",
                    working_directory=HERE)


def test_command_line_semantics():
    """input.cpp: a leading '-' is skipped, a value is taken only when it starts
    with a digit, the solver flags are found among the tokens' first letters."""
    s = nh.scenario("-i -Tt 0.05 -dtf 50 out a.dat log b.log IRE 7 NRE 2".split(),
                    simulation="HH_axon")
    assert s["implicit_mode"] == 1 and s["explicit_mode"] == 0
    assert s["tspan"] == 0.05 and s["dt_factor"] == 50
    assert s["out_path"] == "a.dat" and s["log_path"] == "b.log"
    geo = s["geometry"]
    # 7 internode + 2 node elements per section
    parts = int(2 / (0.0008 + 2.1e-6) + 1)
    assert geo.N == parts * 9
    assert int(geo.material.sum()) == parts * 2
    # a non-numeric value is not consumed: Tt keeps its builder default
    s = nh.scenario("e Tt x out a.dat log b.log".split(), simulation="CT_cable")
    assert s["tspan"] == 0.025
    # the viewer flag is recorded, the viewer itself is not part of this build
    s = nh.scenario("e g out a.dat log b.log".split(), simulation="CT_cable")
    assert s["viewer"] == 1
    # extension tokens
    s = nh.scenario("i method direct prec mixed steps 100 tol 1e-9 restart 40 out a log b"
                    .split(), simulation="CT_cable")
    assert (s["method"], s["precision"], s["max_steps"]) == (nh.DIRECT, nh.FP32_REFINE, 100)
    assert s["restart"] == 40 and s["tol"] == 1e-9
    # a value is only taken when it starts with a digit: ".5" is not
    s = nh.scenario("i tol .5 out a log b".split(), simulation="CT_cable")
    assert s["tol"] == 0.0


@pytest.mark.parametrize("argv,sim,match", [
    ("out a log b", "CT_cable", "no solver"),
    ("e i out a log b", "CT_cable", "both solvers"),
    ("i out a log b", "Only_HH_Model", "isolated ion channel"),
    ("e out a log b", "Purkinje", "unknown simulation"),
    ("e method cg out a log b", "CT_cable", "method"),
])
def test_command_line_errors(argv, sim, match):
    with pytest.raises(nh.NrtError, match=match):
        nh.scenario(argv.split(), simulation=sim)


def test_neuron_loader_end_of_file(tmp_path):
    """The reference reads rows until end-of-file and drops the last read: with a
    trailing newline every row is an element; without one the last row is lost
    (space.cpp:62-75)."""
    src = open(os.path.join(INPUTS, "example_8", "neuron_for_neurite.txt")).read()
    rows = src.strip().split("\n")
    for trailing, n_elem in ((True, len(rows) + 1), (False, len(rows))):
        d = tmp_path / ("t" if trailing else "n")
        d.mkdir()
        (d / "neuron_for_neurite.txt").write_text("\n".join(rows) + ("\n" if trailing else ""))
        s = nh.scenario("e out a log b".split(), simulation="CT_segmented_neuron",
                        working_directory=str(d))
        assert s["geometry"].N == n_elem
