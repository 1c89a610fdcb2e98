"""Writes the deck fixtures the deck-driver tests read.

Reads five of the reference's benchmark decks (configs/benchmarks/*.json) and
its published results table (docs/results/benchmarks.csv) as data, and writes

* decks/<case>.json: each deck as plain JSON (comments removed, nothing else
  changed), so that the tests can run where the reference is absent;
* benchmarks_published.json: the published figures of the same cases.

Run from the repository root: python tests/golden/make_deck_fixtures.py
"""
import csv
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from sparlab_amd.deck import read_deck  # noqa: E402

REF = "/root/reference"
CASES = ("cantilever_beam", "mbb_beam", "aerospace_bracket", "bracket_3d", "wing_rib", "mbb_beam_projected",
         "bracket_3d_projected", "mbb_beam_robust", "bracket_3d_large")
FIELDS = {"dim": int, "elements": int, "dofs": int, "volume_fraction_target": float,
          "volume_fraction_achieved": float, "iterations": int, "stop_reason": str, "compliance_J": float,
          "initial_compliance_J": float, "full_solid_compliance_J": float,
          "equal_mass_plate_compliance_J": float, "grey_level": float, "linear_solver": str,
          "f1_solid_Hz": float, "mass_solid_kg": float, "f1_topology_Hz": float, "mass_topology_kg": float, "final_beta": float}


def main():
    os.makedirs(os.path.join(HERE, "decks"), exist_ok=True)
    for case in CASES:
        doc = read_deck(os.path.join(REF, "configs", "benchmarks", case + ".json"))
        with open(os.path.join(HERE, "decks", case + ".json"), "w") as f:
            json.dump(doc, f, indent=1)
            f.write("\n")
    rows = {r["case"]: r for r in csv.DictReader(open(os.path.join(REF, "docs", "results", "benchmarks.csv")))}
    out = {}
    for case in CASES:
        r = rows[case]
        out[case] = {k: (t(r[k]) if r[k] != "" else None) for k, t in FIELDS.items()}
    with open(os.path.join(HERE, "benchmarks_published.json"), "w") as f:
        json.dump({"source": "docs/results/benchmarks.csv of the reference", "cases": out}, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
