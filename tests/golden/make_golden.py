#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden from the compiled reference (oracle/_ref).

Runs only where /root/reference exists: `make -C oracle ref` compiles the reference's own sources in
place (oracle/Makefile) into oracle/_ref/, then this script runs those binaries and stores their
outputs as data:

  lcg32_streams.npz     LCG32_State<16>/<64> states after Init and after every Update, several seeds
  lcg32_buffered.json   LCG32<CPU>::NextInteger(vector) calls of mixed sizes (W = 16)
  padd.json             bits of p_add = 1.0f - (float)exp(-2.0f / T) under the reference's flags
  refcpu.npz            spin snapshots + (E, M) sums of the reference CPU path (Lattice<2>::Update with
                        SwendsenWang_2D<LCG32, CPU>) after given sweep counts, several extents,
                        temperatures and NUM_THREADS
  protocol.json         the report lines (E/N, C/N, |m|) of the compiled reference driver Ising2D
                        (ising/src/ising_2d.cpp: 1000 warmup sweeps, a sample every 10) for short runs

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
REF = ROOT / "oracle" / "_ref"

SEEDS = [1, 2, 3, 7, 1000, 123456789, 4294967295]
UPDATES = 61
BUFFERED_SIZES = [16, 24, 16, 5, 40, 24, 1, 64, 17]
TEMPERATURES = ["2.0", "2.2691853", "2.6", "1.5", "1.8", "2.8", "3.5", "0.5", "10.0", "2.2691853142130221"]
# (n0, n1, T, NUM_THREADS, sweep counts)
# (NUM_THREADS, n0, n1, T, measured sweeps) for the driver's report lines
PROTOCOL_CASES = [(32, 128, 128, "2.2691853", 3000), (3, 64, 32, "2.0", 2000), (5, 96, 64, "2.6", 2000)]
REFCPU_CASES = [
    (32, 32, "2.2691853", 1, [0, 1, 2, 5, 20]),
    (64, 48, "2.0", 4, [0, 1, 3, 10]),
    (128, 128, "2.2691853", 32, [0, 1, 2, 10]),
    (48, 34, "2.6", 3, [0, 1, 4, 12]),
    (16, 16, "1.5", 2, [0, 1, 7]),
    (96, 64, "3.5", 5, [0, 2, 6]),
]


def run(args, env=None) -> str:
    return subprocess.run(args, check=True, capture_output=True, text=True, env=env).stdout


def main() -> int:
    if not Path("/root/reference").exists():
        print("the reference sources are not available here; nothing to do", file=sys.stderr)
        return 1
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)

    streams = {}
    for w in (16, 64):
        for seed in SEEDS:
            out = json.loads(run([str(REF / "ref_lcg32"), "streams", str(w), str(seed), str(UPDATES)]))
            streams[f"w{w}_seed{seed}"] = np.array(out["states"], dtype=np.uint32)
    np.savez_compressed(HERE / "lcg32_streams.npz", **streams)

    buffered = []
    for seed in (1, 2, 9):
        out = json.loads(run([str(REF / "ref_lcg32"), "buffered", str(seed)] + [str(s) for s in BUFFERED_SIZES]))
        buffered.append({"seed": seed, "w": out["w"], "sizes": BUFFERED_SIZES, "calls": out["calls"]})
    (HERE / "lcg32_buffered.json").write_text(json.dumps(buffered) + "\n")

    padd = json.loads(run([str(REF / "ref_lcg32"), "padd"] + TEMPERATURES))
    (HERE / "padd.json").write_text(json.dumps(padd, indent=1) + "\n")

    arrays = {}
    meta = []
    with tempfile.TemporaryDirectory() as tmp:
        for n0, n1, t, nthreads, ks in REFCPU_CASES:
            name = f"n{n0}x{n1}_T{t}_t{nthreads}"
            path = Path(tmp) / f"{name}.bin"
            env = dict(os.environ, NUM_THREADS=str(nthreads))
            run([str(REF / "ref_sw"), "dump", str(n0), str(n1), t, str(path)] + [str(k) for k in ks], env=env)
            raw = np.fromfile(path, dtype=np.uint8)
            rec = n0 * n1 + 16
            assert raw.size == rec * len(ks)
            spins = np.stack([raw[i * rec:i * rec + n0 * n1] for i in range(len(ks))])
            sums = np.stack([raw[i * rec + n0 * n1:(i + 1) * rec].view(np.int64) for i in range(len(ks))])
            arrays[f"{name}_spins"] = spins
            arrays[f"{name}_sums"] = sums
            meta.append({"name": name, "n0": n0, "n1": n1, "temperature": t, "num_threads": nthreads, "sweeps": ks})
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(HERE / "refcpu.npz", **arrays)

    # report lines of the reference driver for short runs of its measurement protocol
    cases = []
    for nthreads, n0, n1, t, sweeps in PROTOCOL_CASES:
        env = dict(os.environ, NUM_THREADS=str(nthreads))
        out = run([str(REF / "Ising2D"), f"--extent={n0}x{n1}", "--target=cpu", f"--temperature={t}",
                   f"--num_sweeps={sweeps}"], env=env)
        vals = {}
        for key, name in (("Internal energy per spin", "e"), ("Heat capacity per spin", "c"),
                          ("Absolute Magnetization per spin", "m")):
            line = next(l for l in out.splitlines() if l.startswith(key + ":"))
            vals[name] = line.split(":", 1)[1].strip()
        cases.append({"num_threads": nthreads, "n0": n0, "n1": n1, "temperature": t, "sweeps": sweeps, "report": vals})
    (HERE / "protocol.json").write_text(json.dumps({"source": "oracle/_ref/Ising2D (reference sources, reference flags)",
                                                    "cases": cases}, indent=1) + "\n")
    print("wrote", sorted(p.name for p in HERE.glob("*") if p.suffix in (".npz", ".json")))
    return 0


if __name__ == "__main__":
    sys.exit(main())
