"""Writes ref_cpu_stats.json: statistics of the reference CPU path's own runs, for the engine's statistical
parity tests. For 128 x 128 at T in {2.0, 2.2691853, 2.6}, the reference's protocol (ising/src/ising_2d.cpp:
122-162; README.md:66-82): hot start, 1000 warmup sweeps, then 100000 sweeps with a sample (E, M) every 10,
NUM_THREADS=32, lcg32 width 16 (AVX-512 build). oracle/_ref/ref_sw (the reference sources compiled where they
lie, with our harness main) writes the (E, M) series; this script reduces it to the means the driver prints
(<e>, C/N, <|m|>) and to error bars: binning over 50 bins for <e> and <|m|>, jackknife over the bins for C/N.
At T_c the run is the README's, so its means reproduce the published -1.41921, 2.54411, 0.550833.

    make -C oracle ref && python tests/golden/make_refstats.py
"""
import json
import os
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = HERE.parent.parent / "oracle" / "_ref" / "ref_sw"
TEMPERATURES = ["2.0", "2.2691853", "2.6"]
L, WARMUP, SWEEPS, SEP, BINS, THREADS = 128, 1000, 100000, 10, 50, 32


def stats(E: np.ndarray, M: np.ndarray, n: int, t: float) -> dict:
    e = -E.astype(np.float64) / n
    m = np.abs(M.astype(np.float64) / n)
    k = e.size // BINS
    be, bm = e[: k * BINS].reshape(BINS, k), m[: k * BINS].reshape(BINS, k)
    s1, s2 = be.sum(axis=1), (be * be).sum(axis=1)
    tot1, tot2, cnt = s1.sum(), s2.sum(), k * BINS
    cj = np.array([((tot2 - s2[i]) / (cnt - k) - ((tot1 - s1[i]) / (cnt - k)) ** 2) * n / (t * t) for i in range(BINS)])
    t2 = float(np.float32(t) * np.float32(t))   # the driver squares the float temperature in float
    samples = SWEEPS // SEP
    return {"samples": int(e.size), "bins": BINS,
            "e": float(e.sum() / samples), "e_err": float(be.mean(axis=1).std(ddof=1) / np.sqrt(BINS)),
            "abs_m": float(m.sum() / samples), "abs_m_err": float(bm.mean(axis=1).std(ddof=1) / np.sqrt(BINS)),
            "c": float(((e * e).sum() / samples - (e.sum() / samples) ** 2) * n / t2),
            "c_err": float(np.sqrt((BINS - 1) / BINS * ((cj - cj.mean()) ** 2).sum()))}


def main() -> int:
    out = {"source": "oracle/_ref/ref_sw series (reference sources, reference flags, NUM_THREADS=32, W=16)",
           "protocol": f"{L}x{L}, hot start srand48(1), {WARMUP} warmup + {SWEEPS} sweeps, a sample every {SEP}",
           "values": {}}
    env = dict(os.environ, NUM_THREADS=str(THREADS))
    with tempfile.TemporaryDirectory() as tmp:
        for t in TEMPERATURES:
            f = Path(tmp) / f"series_{t}.bin"
            subprocess.run([str(REF), "series", str(L), str(L), t, str(WARMUP), str(SWEEPS), str(f)], env=env, check=True)
            x = np.fromfile(f, dtype=np.int64).reshape(-1, 2)
            out["values"][t] = stats(x[:, 0], x[:, 1], L * L, float(np.float32(float(t))))
            print(t, out["values"][t], flush=True)
    (HERE / "ref_cpu_stats.json").write_text(json.dumps(out, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
