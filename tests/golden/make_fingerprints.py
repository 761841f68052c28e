"""Writes tests/golden/fingerprints.json: per-layer fingerprints (term count,
hash of the sorted keys, hash of the coefficient bits in key order, overlap)
of the oracle on the measured workloads: the 20-qubit chain TFIM (all layers),
the first 6 Trotter steps of the 10x10 lattice TFIM, the 127-qubit random
Clifford + rz circuit (all 30 layers) and four of the batched chains.

No reference fixture covers whole circuits at these sizes, so these fingerprints
are the oracle's own; the oracle itself is pinned by tests/test_oracle.py.

    python tests/golden/make_fingerprints.py
"""

import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests.helpers import fingerprint_workloads, layer_fingerprints, oracle_for  # noqa: E402


def main():
    out = {}
    for name, w in fingerprint_workloads().items():
        o = oracle_for(w)
        out[name] = layer_fingerprints(o, w, o.export)
        print(name, len(out[name]), "layers, final terms", out[name][-1]["terms"], flush=True)
    path = os.path.join(ROOT, "tests", "golden", "fingerprints.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
