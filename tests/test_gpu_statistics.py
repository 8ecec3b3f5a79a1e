"""Statistical correctness of the engine's Markov chain, within 2 standard errors:

  * E/N and C/N of long runs against the exact finite-lattice values (Ferdinand-Fisher,
    tests/golden/exact_values.json);
  * E/N, C/N and <|m|> against the reference CPU path's own runs (tests/golden/ref_cpu_stats.json: the
    compiled reference, 128^2, 1000 + 100000 sweeps, NUM_THREADS=32; at T_c it is the README's published
    run, README.md:66-82), within the combined error of both runs.

Error bars: binning over 50 bins (E/N, |m|), jackknife over the bins (C/N). The engine's lcg32 layout
differs from the reference CPU's, so the two chains agree in distribution, not sweep for sweep. The runs are
deterministic (fixed seeds): a failure is reproducible, not a fluke."""
import json

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

EXACT = {(v["L"], v["T"]): v for v in json.loads((GOLDEN / "exact_values.json").read_text())["values"]}
REF_CPU = json.loads((GOLDEN / "ref_cpu_stats.json").read_text())["values"]
BINS = 50


def binned(x: np.ndarray, nbins: int = BINS):
    k = x.size // nbins
    b = x[: k * nbins].reshape(nbins, k).mean(axis=1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nbins)


def jackknife_c(e: np.ndarray, n_sites: int, t: float, nbins: int = BINS):
    k = e.size // nbins
    x = e[: k * nbins].reshape(nbins, k)
    s1, s2 = x.sum(axis=1), (x * x).sum(axis=1)
    tot1, tot2, m = s1.sum(), s2.sum(), k * nbins
    c_all = (tot2 / m - (tot1 / m) ** 2) * n_sites / (t * t)
    cj = np.array([((tot2 - s2[i]) / (m - k) - ((tot1 - s1[i]) / (m - k)) ** 2) * n_sites / (t * t) for i in range(nbins)])
    err = np.sqrt((nbins - 1) / nbins * ((cj - cj.mean()) ** 2).sum())
    return c_all, err


def _run(swmi, L, t, sweeps, seed, rule):
    n_samples = sweeps // 10
    cfg = swmi.LatticeConfig(L, L, temperature=t, seed=seed, max_samples=n_samples, flip_rule=rule)
    with swmi.Engine(cfg) as eng:
        eng.sweep(1000)
        eng.reset_measurements()
        eng.run(sweeps, 10)
        E, M = eng.series()
        mom = eng.moments()
        assert eng.error_flags() == 0
    assert E.size == n_samples and mom[0] == n_samples
    n = L * L
    e = -E.astype(np.float64) / n
    m = np.abs(M.astype(np.float64) / n)
    # the on-device FP64 moments are the sums of the same series
    assert mom[1] / mom[0] == pytest.approx(e.mean(), rel=1e-12)
    assert mom[3] / mom[0] == pytest.approx(m.mean(), rel=1e-12)
    return e, m


@pytest.mark.parametrize("L,T,sweeps,seed,rule", [
    (128, "2.2691853", 200000, 3, 0),
    (128, "2.0", 100000, 4, 0),
    (128, "2.6", 100000, 5, 0),
    (1024, "2.2691853", 20000, 6, 0),
    (128, "2.2691853", 200000, 7, 1),
    (64, "2.6", 100000, 8, 1),
    (4096, "2.2691853", 5000, 9, 0),
])
def test_energy_and_heat_capacity_within_two_sigma_of_exact(swmi, L, T, sweeps, seed, rule):
    """rule 0: the reference parity flip; rule 1: one random bit per cluster."""
    print(f"statistics {L}^2 T {T} rule {rule}: {sweeps} sweeps", flush=True)
    t = float(np.float32(float(T)))
    e, _ = _run(swmi, L, t, sweeps, seed, rule)
    ex = EXACT[(L, T)]
    mean, err = binned(e)
    print(f"  E/N {mean:.6f} +- {err:.6f} (exact {ex['U_per_spin']:.6f}, z {(mean - ex['U_per_spin']) / err:+.2f})", flush=True)
    assert err < 2e-3
    assert abs(mean - ex["U_per_spin"]) < 2 * err, (mean, err, ex["U_per_spin"])
    c, cerr = jackknife_c(e, L * L, t)
    print(f"  C/N {c:.5f} +- {cerr:.5f} (exact {ex['C_per_spin']:.5f}, z {(c - ex['C_per_spin']) / cerr:+.2f})", flush=True)
    assert abs(c - ex["C_per_spin"]) < 2 * cerr, (c, cerr, ex["C_per_spin"])


@pytest.mark.parametrize("T,seed,rule", [("2.2691853", 11, 0), ("2.0", 12, 0), ("2.6", 13, 0), ("2.2691853", 14, 1)])
def test_engine_agrees_with_the_reference_cpu_run(swmi, T, seed, rule):
    """128^2, the reference's protocol (1000 + 100000 sweeps, a sample every 10): E/N, C/N and <|m|> within
    2 combined standard errors of the compiled reference CPU run of the same protocol."""
    print(f"vs reference CPU run 128^2 T {T} rule {rule}", flush=True)
    t = float(np.float32(float(T)))
    e, m = _run(swmi, 128, t, 100000, seed, rule)
    ref = REF_CPU[T]
    got = {"e": binned(e), "abs_m": binned(m), "c": jackknife_c(e, 128 * 128, t)}
    for k, (v, err) in got.items():
        comb = np.hypot(err, ref[k + "_err"])
        print(f"  {k}: engine {v:.6f} +- {err:.6f}, reference {ref[k]:.6f} +- {ref[k + '_err']:.6f}, "
              f"z {(v - ref[k]) / comb:+.2f}", flush=True)
        assert abs(v - ref[k]) < 2 * comb, (k, v, err, ref[k], ref[k + "_err"])
