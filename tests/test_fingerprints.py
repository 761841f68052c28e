"""Per-layer fingerprints of whole circuits (tests/golden/fingerprints.json,
written by tests/golden/make_fingerprints.py from the oracle).

CPU: the oracle still reproduces them (a change to the oracle that moves any
term set or coefficient bit fails here).  GPU: the HIP engine reproduces them
layer by layer through the C ABI: identical key sets, identical coefficient
bits, overlap within 1e-10.
"""

import json
import os

import pytest

from tests.helpers import fingerprint_workloads, gpu_for, layer_fingerprints, oracle_for

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fingerprints.json")
NAMES = ["chain_tfim_n20", "lattice_tfim_10x10_prefix6", "random_circuit_127q", "batched_chain_0",
         "batched_chain_1", "batched_chain_2", "batched_chain_3"]


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)


def _compare(got, want, name):
    assert len(got) == len(want), name
    for li, (g, w) in enumerate(zip(got, want)):
        assert g["terms"] == w["terms"], f"{name} layer {li}: {g['terms']} vs {w['terms']} terms"
        assert g["keys"] == w["keys"], f"{name} layer {li}: key sets differ"
        assert g["coefs"] == w["coefs"], f"{name} layer {li}: coefficient bits differ"
        assert abs(g["overlap"] - w["overlap"]) <= 1e-10, f"{name} layer {li}: overlap"


def test_fixture_covers_workloads():
    assert sorted(_golden()) == sorted(NAMES)
    assert sorted(fingerprint_workloads()) == sorted(NAMES)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fingerprints(name):
    w = fingerprint_workloads()[name]
    o = oracle_for(w)
    _compare(layer_fingerprints(o, w, o.export), _golden()[name], name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_fingerprints(name):
    w = fingerprint_workloads()[name]
    g = gpu_for(w)
    _compare(layer_fingerprints(g, w, g._export), _golden()[name], name)
