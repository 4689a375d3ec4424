"""The elastic-beam fixture (tests/golden/elastic_beam_20k.*) against the
oracle that wrote it: the first recorded tip deflection (500 steps) is
reproduced bit for bit, so the fixture and the oracle cannot drift apart."""
from __future__ import annotations

import json
import os

import numpy as np

import oracle
from mysph_amd import cases

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_fixture_matches_oracle_first_500_steps():
    with open(os.path.join(GOLDEN, "elastic_beam_20k.json")) as f:
        meta = json.load(f)
    gold = np.load(os.path.join(GOLDEN, "elastic_beam_20k.npz"))
    case = cases.elastic_beam()
    assert case["n_real"] == meta["n_particles"]
    assert case["config"].dt == meta["dt"]
    x = case["x"][:, 0]
    tip = x > x.max() - 0.5 * (x.max() - np.sort(np.unique(x))[-2])
    assert tip.sum() == 40
    oracle.set_threads(os.cpu_count() or 1)
    o = oracle.OracleSim(case)
    o.run(meta["tip_every"])
    first = float(np.mean(o["x"][tip, 1] - case["x"][tip, 1]))
    assert first == gold["tip"][0]
    # the deflection grows along the first quarter period of the mode
    assert np.all(np.diff(gold["tip"]) > 0)
