"""Writes thresholds.json: for every temperature of padd.json, the bits of p_add (from the compiled
reference, padd.json) and the integer bond threshold t with "u < t" <=> "2^-32f * (float)u < p_add" for all
2^32 draws u, checked exhaustively with the oracle. Run in the container that built oracle/liboracle.so:

    python tests/golden/make_thresholds.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from oracle import oracle  # noqa: E402


def main() -> int:
    padd = json.loads((HERE / "padd.json").read_text())
    out = {}
    for t, bits in padd.items():
        p = float(np.array([bits], dtype=np.uint32).view(np.float32)[0])
        thr = int(oracle.lib().or_threshold(p))
        assert oracle.lib().or_threshold_mismatches(p, thr) == 0, t
        out[t] = {"p_add_bits": bits, "threshold": thr}
    (HERE / "thresholds.json").write_text(json.dumps(out, indent=1) + "\n")
    print("wrote thresholds.json for", len(out), "temperatures")
    return 0


if __name__ == "__main__":
    sys.exit(main())
