"""Regenerate the golden fixtures from the reference's own CPU solver.

Runs ``oracle/_ref/neurite_ref`` -- the reference sources (setup chain, the
space.cpp builders, SolverExCPU) compiled in place by ``oracle/Makefile`` and
driven by ``oracle/ref_harness.cpp`` -- over the reference's example scenarios
(applications/example_*/example_*.sh options, truncated to a few thousand
steps) and stores each dump as ``<name>.json.gz``: raw builder geometry, the
derived per-element arrays, dT, the .dat trace and full-V checkpoints.

Needs the reference tree (only in the build container); the fixtures it writes
are data and travel with the repository.

    python tests/golden/make_golden.py [--ref /root/reference]
"""
import argparse
import gzip
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))

# name -> (simulation env, example dir for working_directory, CLI, max steps)
# CLIs follow applications/example_N/example_N.sh (explicit solver, dtf as there).
SCENARIOS = {
    "ex1_only_hh": ("Only_HH_Model", "example_1", "e out o.dat log o.log", -1),
    "ex2_hh_axon": ("HH_axon", "example_2",
                    "e Tt 0.025 IRE 20 NRE 1 dtf 0.8 xspan 0.02 out o.dat log o.log", 3000),
    "ex3_hh_y": ("HH_Y_branching", "example_3", "e Tt 0.1 dtf 0.8 out o.dat log o.log", 30000),
    "ex4_ct_cable": ("CT_cable", "example_4", "e Tt 0.1 dtf 0.8 out o.dat log o.log", 6000),
    "ex5_ct_y": ("CT_Y_branching", "example_5", "e Tt 0.1 dtf 0.8 out o.dat log o.log", -1),
    "ex6_ct_tree": ("CT_symmetric_tree", "example_6", "e Tt 0.1 dtf 0.8 out o.dat log o.log",
                    6000),
    "ex7_hh_axon_strain": ("HH_axon", "example_7",
                           "e Tt 0.1 dtf 0.6 xspan 0.02 IRE 20 NRE 1 MAXE 0.1 str 0.11195694 "
                           "out o.dat log o.log", 3000),
    "ex8_neuron": ("CT_segmented_neuron", "example_8", "e Tt 0.8 dtf 0.6 out o.dat log o.log",
                   6000),
}
CHECKPOINTS = [1, 2, 10, 100, 1000, 2999]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", default=None)
    a = ap.parse_args()
    exe = os.path.join(REPO, "oracle", "_ref", "neurite_ref")
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref", f"REF={a.ref}"],
                   check=True)
    if not os.path.exists(exe):
        sys.exit("oracle/_ref/neurite_ref not built (reference tree missing?)")
    for name, (sim, exdir, cli, max_steps) in SCENARIOS.items():
        if a.only and name != a.only:
            continue
        with tempfile.TemporaryDirectory() as td:
            dump = os.path.join(td, "dump.json")
            env = dict(os.environ, simulation=sim, REF_DUMP=dump,
                       working_directory=os.path.join(a.ref, "applications", exdir),
                       REF_MAX_STEPS=str(max_steps),
                       REF_CHECKPOINTS=",".join(str(c) for c in CHECKPOINTS))
            subprocess.run([exe] + cli.split(), env=env, cwd=td, check=True,
                           stdout=subprocess.DEVNULL)
            with open(dump) as f:
                d = json.load(f)
        d["scenario"] = {"simulation": sim, "example": exdir, "cli": cli,
                         "max_steps": max_steps}
        out = os.path.join(HERE, name + ".json.gz")
        with gzip.open(out, "wt") as f:
            json.dump(d, f, separators=(",", ":"))
        print(f"{name}: N={d['N']} hh={len(d['hh_elem'])} steps={d['steps_run']} "
              f"dT={d['dT']:.3e} trace={len(d['trace_v'])} -> {os.path.getsize(out)} B")


if __name__ == "__main__":
    main()
