"""Writes tests/golden/lossy_clifford_words.json: the (gate, input word, output
word) vectors of the reference's lossy-word Clifford tests
(crates/ppvm-pauli-word/src/loss/clifford.rs, the `#[cfg(test)]` tables),
read from the reference checkout as text.  Only the data is kept.

    python tests/golden/make_lossy_vectors.py [/path/to/reference]
"""

import json
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = open(os.path.join(ref, "crates/ppvm-pauli-word/src/loss/clifford.rs")).read()
    out = {}
    # each test: fn test_<gate>() { for (input, target) in [ ("..", ".."), ... ] { ... output.<gate>(...)
    for m in re.finditer(r"fn test_(\w+)\(\)\s*\{(.*?)\n    \}", src, re.S):
        body = m.group(2)
        pairs = re.findall(r'\("([IXYZL]+)",\s*"([IXYZL]+)"\)', body)
        call = re.search(r"output\.(\w+)\(([^)]*)\)", body)
        if not pairs or not call:
            continue
        args = [int(a) for a in call.group(2).split(",") if a.strip()]
        out[call.group(1)] = {"qubits": args, "pairs": pairs}
    path = os.path.join(ROOT, "tests", "golden", "lossy_clifford_words.json")
    with open(path, "w") as f:  # one gate per line
        f.write("{\n" + ",\n".join(json.dumps(k) + ": " + json.dumps(out[k]) for k in sorted(out)) + "\n}\n")
    print({k: len(v["pairs"]) for k, v in out.items()})


if __name__ == "__main__":
    main()
