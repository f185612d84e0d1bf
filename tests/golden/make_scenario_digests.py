"""Writes scenario_digests.json: SHA-256 of the arrays each seeded scenario
generator produces at small sizes (the bytes a seed fixes).

    python tests/golden/make_scenario_digests.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from mfcm_amd import scenarios as S  # noqa: E402

CASES = {
    "hillslope_rain_n64_seed0": lambda: S.hillslope_rain(seed=0, n=64),
    "storm_catchment_n96_seed1": lambda: S.storm_catchment(seed=1, n=96),
    "breach_front_n128_seed2": lambda: S.breach_front(seed=2, n=128, wall_x=40, breach=10),
    "member_seed5": lambda: S.ensemble_member(5, side_lo=32, side_hi=96),
    "dambreak_64x48": lambda: S.dam_break_bump(64, 48),
}


def digests():
    return {k: f().digest() for k, f in CASES.items()}


if __name__ == "__main__":
    with open(os.path.join(HERE, "scenario_digests.json"), "w") as fh:
        json.dump(digests(), fh, indent=1, sort_keys=True)
        fh.write("\n")
    print("wrote scenario_digests.json")
