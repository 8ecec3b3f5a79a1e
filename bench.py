#!/usr/bin/env python3
"""Benchmark of the MI355X Swendsen-Wang engine (one JSON line on stdout).

Headline workload: plain Swendsen-Wang on a 4096 x 4096 periodic 2D Ising lattice at
T = 2.2691853 (float), hot start (srand48(1)), per-site lcg32 seeded from seed 0, everything resident in
HBM. A "step" is one sweep (bond activation + labelling + merge + flip of all 16.8M spins); `value` is
spin-updates/s = steps * N / seconds between two device synchronisations.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--n 4096] [--temperature T] [--no-cpu-baseline]
                    [--no-batched] [--no-served]

Two more workloads are reported in the same JSON line:
  batched  256 independent lattices, L = 2^k with k uniform in {5..10}, T uniform in [1.8, 2.8], one seed
           each (workload drawn from numpy default_rng(0)), 1000 warmup + 5000 sweeps in one engine (one
           work list, graph-captured sweep loops, autotuned tile height per L), against the same lattices
           run one by one through the plain path (sampled: 100 timed sweeps per lattice).
  served   a Poisson stream of jobs (L = 2^k, k in {6..12}; T uniform in [1.5, 3.5]; sweeps log-uniform in
           [1e3, 1e5]; a tenth of them warmup) at an offered load of 0.8 of the plain 4096^2 rate, through the
           continuous-batching server; p50/p99 completion latency and aggregate spin-updates/s.

Roofline: the dominant kernel (bond activation + tile labelling, which also applies the previous sweep's
cluster flips as it loads the spins) is HBM-bound; its algorithmic bytes are 18 per site per launch (spin
read 1 and write 1, lcg32 state read+write 8, label read 4 (the previous sweep's) and write 4). The whole
sweep's figure is 19 bytes per site (bond phase 13, flip phase: label read 4, spin read+write 2).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0          # MI355X HBM3E
BOND_LABEL_BYTES_PER_SITE = 18  # spin r/w 2 + lcg32 r/w 8 + label r/w 8 (the flip is applied on load)
SWEEP_BYTES_PER_SITE = 19       # bond phase 13 + flip phase: label r 4 + spin r/w 2
T_C = 2.2691853


def host_cpu() -> dict:
    """The host's CPU model, the cores this process may run on, and the physical cores among them."""
    model, cores_of = "", set()
    phys = core = None
    allowed = os.sched_getaffinity(0)
    try:
        cpu = None
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            key, _, val = line.partition(":")
            key, val = key.strip(), val.strip()
            if key == "processor":
                cpu = int(val)
            elif key == "model name" and not model:
                model = val
            elif key == "physical id":
                phys = val
            elif key == "core id":
                core = val
                if cpu in allowed:
                    cores_of.add((phys, core))
    except OSError:
        pass
    return {"model": model, "logical_cpus": len(allowed), "physical_cores": len(cores_of) or len(allowed)}


def cpu_threads() -> int:
    """Threads for the CPU baseline: the physical cores this process may use, within the host's thread
    budget for one GPU (OMP_NUM_THREADS where it is set)."""
    cores = host_cpu()["physical_cores"]
    budget = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(cores, budget) if budget > 0 else cores)


def cpu_baseline(n: int, temperature: float, warmup: int = 10, sweeps: int = 300) -> dict | None:
    """The reference CPU path (oracle/_ref/ref_sw, built from the reference sources) on the host cores,
    on a bounded sample of the same workload; falls back to the single-thread C restatement."""
    cores = cpu_threads()
    ref = ROOT / "oracle" / "_ref" / "ref_sw"
    if ref.exists():
        env = dict(os.environ, NUM_THREADS=str(cores))
        try:
            out = subprocess.run([str(ref), "run", str(n), str(n), str(temperature), str(warmup), str(sweeps)], env=env,
                                 capture_output=True, text=True, timeout=300, check=True).stdout
            ns = float(re.search(r"Update time per spin: ([0-9.eE+-]+) ns", out).group(1))
            return {"value": 1e9 / ns, "unit": "spin-updates/s", "cores": cores, "kind": "reference", "cpu": host_cpu(),
                    "sample": f"{n}x{n} T={temperature}, {warmup} warmup + {sweeps} sweeps with observables every 10 "
                              f"sweeps (ising_2d.cpp protocol), NUM_THREADS={cores}, lcg32 W=16 (AVX-512 build)"}
        except (subprocess.SubprocessError, AttributeError, ValueError) as exc:
            sys.stderr.write(f"reference CPU baseline failed: {exc}\n")
    try:
        import numpy as np
        from oracle import oracle
        r = oracle.RefCpu(n, n, 16, cores)
        s = oracle.hot_start(n * n, 1)
        k = 3
        t0 = time.perf_counter()
        for _ in range(k):
            r.sweep(s, temperature)
        dt = time.perf_counter() - t0
        return {"value": k * n * n / dt, "unit": "spin-updates/s", "cores": 1, "kind": "port", "cpu": host_cpu(),
                "sample": f"{n}x{n} T={temperature}, {k} sweeps of the single-threaded C restatement"}
    except Exception as exc:  # the baseline is informative only
        sys.stderr.write(f"CPU baseline unavailable: {exc}\n")
        return None


README_CONFIG1 = {"ns_per_spin": 3.32034, "e": -1.41921, "c": 2.54411, "abs_m": 0.550833}   # README.md:66-82


def config1_leg(cpa) -> dict:
    """The reference's own published run, exactly (README.md:66-82): 128x128, T = 2.2691853, 1000 warmup +
    100000 sweeps, a sample every 10, NUM_THREADS=32, through the compiled reference driver (oracle/_ref/Ising2D)
    on the host; and the same protocol on the GPU engine (swmi_run, samples reduced on the device)."""
    out = {"workload": "128x128 T=2.2691853, 1000 warmup + 100000 sweeps, sample every 10 (ising_2d.cpp defaults)",
           "readme": README_CONFIG1}
    ref = ROOT / "oracle" / "_ref" / "Ising2D"
    if ref.exists():
        env = dict(os.environ, NUM_THREADS="32")
        t0 = time.perf_counter()
        text = subprocess.run([str(ref), "--extent=128x128", "--num_sweeps=100000"], env=env, capture_output=True, text=True,
                              timeout=600, check=True).stdout
        wall = time.perf_counter() - t0
        got = {k: float(re.search(pat + r": ([-0-9.eE+]+)", text).group(1)) for k, pat in (
            ("ns_per_spin", "Update time per spin"), ("e", "Internal energy per spin"), ("c", "Heat capacity per spin"),
            ("abs_m", "Absolute Magnetization per spin"))}
        out["cpu_reference"] = {**got, "spin_updates_per_s": 1e9 / got["ns_per_spin"], "threads": 32, "cpu": host_cpu(),
                                "wall_s": wall, "matches_readme_digits": all(f"{got[k]:g}" == f"{README_CONFIG1[k]:g}"
                                                                              for k in ("e", "c", "abs_m"))}
    cfg = cpa.LatticeConfig(128, 128, T_C, seed=0, spin_seed=1, max_samples=10000)
    with cpa.Engine(cfg) as e:
        e.sweep(1000)
        e.reset_measurements()
        e.synchronize()
        t0 = time.perf_counter()
        e.run(100000, 10)
        e.synchronize()
        dt = time.perf_counter() - t0
        eb = e.error_bars(0, 50)
        mom = e.moments()
    t2 = float(T_C) ** 2
    out["gpu"] = {"spin_updates_per_s": 100000 * 128 * 128 / dt, "ns_per_spin": dt * 1e9 / (100000 * 128 * 128),
                  "e": mom[1] / mom[0], "c": (mom[2] / mom[0] - (mom[1] / mom[0]) ** 2) * 128 * 128 / t2,
                  "abs_m": mom[3] / mom[0], "e_err": eb["e_err"], "c_err": eb["c_err"], "abs_m_err": eb["abs_m_err"],
                  "note": "the engine's lcg32 layout (64-site groups) differs from the reference CPU's (per-thread "
                          "W=16 streams): the chains agree statistically, not sweep for sweep"}
    return out


def log(msg: str) -> None:
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def load_traffic(n: int) -> float | None:
    """HBM bytes per k_bond_label launch from the committed rocprofv3 PMC summary (scaled FETCH+WRITE)."""
    f = ROOT / "profiles" / f"traffic_bond_label_{n}.json"
    if f.exists():
        try:
            return float(json.loads(f.read_text())["bytes_per_launch"])
        except (ValueError, KeyError):
            return None
    return None


def batch_workload(cpa) -> list:
    """The batched leg's 256 lattices: L = 2^k with k uniform in {5..10}, T uniform in [1.8, 2.8] (float),
    drawn from numpy default_rng(0); lattice i has lcg32 seed 1000 + i and hot-start seed 1 + i."""
    import numpy as np
    rng = np.random.default_rng(0)
    ls = (2 ** rng.integers(5, 11, size=256)).tolist()
    ts = rng.uniform(1.8, 2.8, size=256).astype(np.float32).tolist()
    return [cpa.LatticeConfig(L, L, t, seed=1000 + i, spin_seed=1 + i) for i, (L, t) in enumerate(zip(ls, ts))]


def batched_leg(cpa, warmup: int = 1000, sweeps: int = 5000, alone_sweeps: int = 500, alone_warmup=(256, 32)) -> dict:
    cfgs = batch_workload(cpa)
    ls = [c.n0 for c in cfgs]
    tuned = {}
    for L in sorted(set(ls)):
        tuned[L] = cpa.autotune(cpa.LatticeConfig(L, L, 2.2691853, seed=1), sweeps=10)[0]
    sites = sum(L * L for L in ls)
    # the batch is tuned as a batch (one tile height for all lattices); the one-by-one baseline below uses
    # each lattice's own best height (tuned alone above)
    batch_rows, batch_ms = cpa.autotune_batch(cfgs, sweeps=20)
    for c in cfgs:
        c.tile_rows = batch_rows
    with cpa.Engine(cfgs) as eng:
        eng.sweep(warmup)
        eng.synchronize()
        t0 = time.perf_counter()
        eng.sweep(sweeps)
        eng.synchronize()
        dt = time.perf_counter() - t0
        flags = max(eng.error_flags(i) for i in range(len(cfgs)))
    batched = sweeps * sites / dt
    # the same lattices one by one through the plain path (one engine each; timed sample of the sweeps). The
    # untimed warmup loops (256 sweeps, then 32) capture the engine's 128- and 16-sweep graphs, so the timed
    # loop only replays them.
    t_alone = 0.0
    for c in cfgs:
        c.tile_rows = 0   # the height tuned for this extent alone
        with cpa.Engine(c) as e1:
            for w in alone_warmup:
                e1.sweep(w)
            e1.synchronize()
            t0 = time.perf_counter()
            e1.sweep(alone_sweeps)
            e1.synchronize()
            t_alone += time.perf_counter() - t0
    alone = alone_sweeps * sites / t_alone
    return {"value": batched, "unit": "spin-updates/s", "lattices": len(cfgs), "sites": sites, "warmup": warmup,
            "sweeps": sweeps, "seconds": dt, "one_by_one": alone, "one_by_one_sample_sweeps": alone_sweeps,
            "one_by_one_warmup_sweeps": sum(alone_warmup),
            "speedup_vs_one_by_one": batched / alone, "tile_rows_alone_by_L": tuned, "batch_tile_rows": batch_rows,
            "batch_autotune_ms_per_sweep": dict(zip(["1", "2", "4", "8", "16", "32", "64"], batch_ms)), "error_flags": flags,
            "workload": "256 lattices, L=2^k k~U{5..10}, T~U[1.8,2.8], default_rng(0); one engine, graph-captured"}


def accuracy_leg(cpa, n: int = 4096, warmup: int = 1000, sweeps: int = 4000, sep: int = 10) -> dict:
    """E/N and C/N of the plain path against the exact finite-lattice values (Ferdinand-Fisher,
    tests/golden/exact_values.json), with binning/jackknife error bars from the on-device series; both
    flip rules. Outside the timed region."""
    exact = {(v["L"], v["T"]): v for v in json.loads((ROOT / "tests" / "golden" / "exact_values.json").read_text())["values"]}
    ex = exact[(n, "2.2691853")]
    out = {"workload": f"{n}x{n} T=2.2691853, {warmup} warmup + {sweeps} sweeps, a sample every {sep}, 20 bins",
           "exact_e_per_spin": ex["U_per_spin"], "exact_c_per_spin": ex["C_per_spin"]}
    for rule, name in ((0, "parity_rule"), (1, "random_rule")):
        cfg = cpa.LatticeConfig(n, n, T_C, seed=7, max_samples=sweeps // sep, flip_rule=rule)
        with cpa.Engine(cfg) as e:
            e.sweep(warmup)
            e.reset_measurements()
            e.run(sweeps, sep)
            eb = e.error_bars(0, 20)
            flags = e.error_flags()
        out[name] = {"e_per_spin": eb["e"], "e_err": eb["e_err"], "z_e": (eb["e"] - ex["U_per_spin"]) / eb["e_err"],
                     "c_per_spin": eb["c"], "c_err": eb["c_err"], "abs_m": eb["abs_m"], "tau_int_e": eb["tau_int_e"],
                     "samples": int(eb["samples"]), "error_flags": flags}
    return out


def plain_configs_leg(cpa, warmup=(256, 32), sweeps: int = 300) -> dict:
    """The other plain-path configurations: 1024^2 at T_c with the default and the autotuned tile height,
    4096^2 and 8192^2 at T in {2.0, 2.2691853, 2.6}; sweep-only rates and the 4096^2 rate of the reference
    protocol (observables every 10 sweeps inside the timed region). The untimed warmup loops capture the
    graphs the timed loop replays (128- and 16-sweep units; with observables, the sample-group graph)."""
    def rate(cfg, n_sweeps, sep=0):
        with cpa.Engine(cfg) as e:
            for w in warmup:
                e.sweep(w)
            if sep:
                e.run(128, sep)
            e.synchronize()
            t0 = time.perf_counter()
            if sep:
                e.run(n_sweeps, sep)
            else:
                e.sweep(n_sweeps)
            e.synchronize()
            dt = time.perf_counter() - t0
            assert e.error_flags() == 0
        return n_sweeps * cfg.n0 * cfg.n1 / dt
    out = {}
    default = rate(cpa.LatticeConfig(1024, 1024, T_C, seed=0, tile_rows=16), 1000)
    best, ms = cpa.autotune(cpa.LatticeConfig(1024, 1024, T_C, seed=0), sweeps=20)
    tuned = rate(cpa.LatticeConfig(1024, 1024, T_C, seed=0, tile_rows=best), 1000)
    out["1024_Tc"] = {"tile_rows_16": default, "autotuned": tuned, "autotuned_tile_rows": best,
                      "autotune_ms_per_sweep": dict(zip(["1", "2", "4", "8", "16", "32", "64"], ms))}
    for n in (4096, 8192):
        for t in (2.0, T_C, 2.6):
            out[f"{n}_T{t}"] = rate(cpa.LatticeConfig(n, n, t, seed=0), sweeps if n == 4096 else sweeps // 2)
    out["4096_Tc_with_observables_every_10"] = rate(cpa.LatticeConfig(4096, 4096, T_C, seed=0, max_samples=0), 1000, sep=10)
    return out


def served_leg(cpa, capacity: float, n_jobs: int = 32, load: float = 0.8, seed: int = 1) -> dict:
    import numpy as np
    rng = np.random.default_rng(seed)
    ks = rng.integers(6, 13, size=n_jobs)
    temps = rng.uniform(1.5, 3.5, size=n_jobs).astype(np.float32)
    total = np.floor(10 ** rng.uniform(3, 5, size=n_jobs)).astype(np.int64)
    warm = (total // 10 // 10) * 10
    work = np.array([(4 ** int(k)) * int(t) for k, t in zip(ks, total)], dtype=np.float64)
    mean_work = float(np.mean([(4 ** k) for k in range(6, 13)])) * (1e5 - 1e3) / np.log(100.0)
    rate = load * capacity / mean_work          # jobs per second at the stated offered load
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_jobs))
    jobs = [cpa.Job(2 ** int(k), 2 ** int(k), float(t), seed=5000 + i, spin_seed=1 + i, warmup=int(w), sweeps=int(n - w),
                    tag=i) for i, (k, t, n, w) in enumerate(zip(ks, temps, total, warm))]
    with cpa.Server(memory_budget=200 * 10 ** 9) as srv:
        t0 = time.perf_counter()
        for j, a in zip(jobs, arrivals):
            now = time.perf_counter() - t0
            if a > now:
                time.sleep(a - now)
            srv.submit(j)
        left = srv.wait(600000)
        res = srv.poll(n_jobs + 8)
        st = srv.stats()
    lat = np.array([r["t_finish"] - r["t_submit"] for r in res])
    first = min(r["t_submit"] for r in res)
    last = max(r["t_finish"] for r in res)
    return {"p50_latency_s": float(np.percentile(lat, 50)), "p90_latency_s": float(np.percentile(lat, 90)),
            "p99_latency_s": float(np.percentile(lat, 99)), "max_latency_s": float(lat.max()),
            "spin_updates_per_s": float(st["spin_updates"] / (last - first)), "jobs": len(res), "unfinished": left,
            "failed": sum(1 for r in res if r["status"] != 0), "offered_load": load, "arrival_rate_per_s": rate,
            "makespan_s": last - first, "work_spin_updates": float(work.sum()),
            "workload": f"{n_jobs} Poisson jobs, L=2^k k~U{{6..12}}, T~U[1.5,3.5], sweeps~logU[1e3,1e5] (10% warmup), "
                        f"default_rng({seed}); 200 GB budget, 4 size-class lanes"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--temperature", type=float, default=T_C)
    ap.add_argument("--kernel-sweeps", type=int, default=100, help="sweeps timed kernel by kernel with HIP events")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--served-jobs", type=int, default=200)
    ap.add_argument("--no-config1", action="store_true", help="skip the reference's published 128^2 run (CPU and GPU)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="also run a device copy of BYTES bytes (scale for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this engine shards independent lattices per GPU (no collectives); "
              "multi-GPU runs are not part of this build.", file=sys.stderr)
        return 2

    import computational_physics_amd as cpa

    n = args.n
    sites = n * n
    info = cpa.device_info()
    copy_ms = cpa.copy_calibration(args.copy_calibration, 3) if args.copy_calibration else None
    eng = cpa.Engine(cpa.LatticeConfig(n, n, args.temperature, seed=0, spin_seed=1))
    eng.sweep(args.warmup)
    eng.synchronize()
    t0 = time.perf_counter()
    eng.sweep(args.steps)
    eng.synchronize()
    dt = time.perf_counter() - t0
    value = args.steps * sites / dt
    ms_per_step = dt * 1e3 / args.steps

    # per-kernel device time with HIP events on the engine stream
    k_ms = eng.time_kernels(args.kernel_sweeps)
    flags = eng.error_flags()
    eng.close()
    bond_label_s = k_ms[0] * 1e-3
    achieved = BOND_LABEL_BYTES_PER_SITE * sites / bond_label_s / 1e9
    sweep_gbps = SWEEP_BYTES_PER_SITE * sites * args.steps / dt / 1e9

    result = {
        "metric": "spin-updates/s at 4096^2 T_c (plain vs batched mixed-L); p99 served latency",
        "value": value,
        "unit": "spin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (hot start from srand48(1), lcg32 seed 0)",
        "config": {"workload": f"plain Swendsen-Wang {n}x{n} periodic, T={args.temperature}, int8 spins, uint32 labels, "
                               f"per-site lcg32, tile 64x16", "n0": n, "n1": n, "temperature": args.temperature},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": load_traffic(n),
                     "kernel": "k_bond_label", "bytes_per_site": BOND_LABEL_BYTES_PER_SITE,
                     "avg_launch_ms": k_ms[0]},
        "sweep_roofline": {"bytes_per_site": SWEEP_BYTES_PER_SITE, "achieved_GBps": sweep_gbps,
                           "frac": sweep_gbps / HBM_PEAK_GBPS},
        "kernels_ms": {"k_bond_label": k_ms[0], "k_merge": k_ms[1], "k_resolve": k_ms[2]},
        "device": info["name"] or info["arch"],
        "copy_calibration": None if copy_ms is None else {"bytes": args.copy_calibration, "ms": copy_ms,
                                                          "GBps": 2 * args.copy_calibration / copy_ms / 1e6},
        "error_flags": flags,
    }
    log(f"plain {n}x{n}: {value:.3e} spin-updates/s, {ms_per_step * 1e3:.1f} us/sweep")
    if not args.no_configs:
        result["accuracy"] = accuracy_leg(cpa)
        log(f"accuracy: z(E/N) parity {result['accuracy']['parity_rule']['z_e']:.2f}, "
            f"random {result['accuracy']['random_rule']['z_e']:.2f}")
        result["plain_configs"] = plain_configs_leg(cpa)
        log("plain configs done")
    if not args.no_batched:
        result["batched"] = batched_leg(cpa)
        log(f"batched: {result['batched']['value']:.3e} spin-updates/s, "
            f"{result['batched']['speedup_vs_one_by_one']:.2f}x one by one")
    if not args.no_served:
        result["served"] = served_leg(cpa, capacity=value, n_jobs=args.served_jobs)
        log(f"served: p50 {result['served']['p50_latency_s']:.3f} s, p99 {result['served']['p99_latency_s']:.3f} s")
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(n, args.temperature)
        log("cpu baseline done")
        if not args.no_config1:
            result["config1"] = config1_leg(cpa)
            log("config 1 (128^2, 1000 + 100000 sweeps) done")
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
