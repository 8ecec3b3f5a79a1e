"""Statistical correctness of the engine's Markov chain: E/N and C/N of long runs against the exact
finite-lattice values (Ferdinand-Fisher, tests/golden/exact_values.json), with error bars from binning
(E/N) and jackknife over bins (C/N). The runs are deterministic (fixed seeds), so a failure is
reproducible, not a fluke; the bound is 4 standard errors.

The reference's own sample (README.md:71-81, 128^2 at T_c, 100000 sweeps on its CPU path) reports
E/N = -1.41921 and C/N = 2.54411 against the exact -1.4190763 and 2.5363312."""
import json

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

EXACT = {(v["L"], v["T"]): v for v in json.loads((GOLDEN / "exact_values.json").read_text())["values"]}


def binned(e: np.ndarray, nbins: int):
    k = e.size // nbins
    b = e[: k * nbins].reshape(nbins, k).mean(axis=1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nbins)


def jackknife_c(e: np.ndarray, n_sites: int, t: float, nbins: int):
    k = e.size // nbins
    x = e[: k * nbins].reshape(nbins, k)
    s1, s2 = x.sum(axis=1), (x * x).sum(axis=1)
    tot1, tot2, m = s1.sum(), s2.sum(), k * nbins
    c_all = (tot2 / m - (tot1 / m) ** 2) * n_sites / (t * t)
    cj = np.array([((tot2 - s2[i]) / (m - k) - ((tot1 - s1[i]) / (m - k)) ** 2) * n_sites / (t * t) for i in range(nbins)])
    err = np.sqrt((nbins - 1) / nbins * ((cj - cj.mean()) ** 2).sum())
    return c_all, err


@pytest.mark.parametrize("L,T,sweeps,seed,rule", [
    (128, "2.2691853", 200000, 3, 0),
    (128, "2.0", 50000, 4, 0),
    (128, "2.6", 50000, 5, 0),
    (1024, "2.2691853", 20000, 6, 0),
    (128, "2.2691853", 100000, 7, 1),
    (64, "2.6", 50000, 8, 1),
    (4096, "2.2691853", 2000, 9, 0),
])
def test_energy_and_heat_capacity_match_exact(swmi, L, T, sweeps, seed, rule):
    """rule 0: the reference parity flip; rule 1: one random bit per cluster."""
    t = float(np.float32(float(T)))
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
    # the on-device FP64 moments are the sums of the same series
    assert mom[1] / mom[0] == pytest.approx(e.mean(), rel=1e-12)
    assert mom[3] / mom[0] == pytest.approx(np.abs(M / n).mean(), rel=1e-12)
    ex = EXACT[(L, T)]
    mean, err = binned(e, 50 if n_samples >= 1000 else 20)
    assert err < 2e-3
    assert abs(mean - ex["U_per_spin"]) < 4 * err, (mean, err, ex["U_per_spin"])
    if n_samples >= 1000:
        c, cerr = jackknife_c(e, n, t, 50)
        assert abs(c - ex["C_per_spin"]) < 4 * cerr + 1e-3, (c, cerr, ex["C_per_spin"])
