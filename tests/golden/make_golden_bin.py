"""Writes tests/golden/ref_bin_digests.json: SHA-256 digests and sizes of the
BIN pointers/blocks files that the reference's own output (src/output.f90,
OUTBIN branch, oracle/_ref) writes for the states of
tests/golden/bin_inputs.py.  Run from the repo root:
python tests/golden/make_golden_bin.py"""
import json
import os
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.dirname(__file__)]

import bin_inputs as B  # noqa: E402
from wxoracle import Oracle, RefLib  # noqa: E402


def main():
    O = Oracle()
    out = {}
    for variant, seed, steps in B.BIN_CASES:
        cfg, pp, m = B.state(O, variant, seed, steps)
        d = tempfile.mkdtemp()
        rp, rb = B.reference_files(RefLib(variant), pp, m, d)
        out[f"{variant}_{seed}_{steps}"] = {"pointers": [B.digest(rp), os.path.getsize(rp)],
                                             "blocks": [B.digest(rb), os.path.getsize(rb)]}
    path = os.path.join(os.path.dirname(__file__), "ref_bin_digests.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
