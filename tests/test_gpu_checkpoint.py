"""Checkpoint / resume and the on-device series' error bars."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TC = 2.2691853


@pytest.mark.parametrize("n0,n1,rule", [(128, 128, 0), (256, 64, 1), (32, 32, 0)])
def test_restore_continues_bit_for_bit(swmi, n0, n1, rule):
    """30 sweeps, checkpoint, 40 more (crossing lcg32 rotations) == restore into a fresh lattice + 40."""
    cfg = swmi.LatticeConfig(n0, n1, 2.1, seed=5, spin_seed=7, flip_rule=rule, max_samples=16)
    with swmi.Engine(cfg) as a:
        a.run(30, 10)
        ck = a.checkpoint()
        a.run(40, 10)
        want_spins, want_rng, want_mom, want_ser = a.spins(), a.rng_state(), a.moments(), a.series()
    other = swmi.LatticeConfig(n0, n1, 2.9, seed=99, spin_seed=3, flip_rule=1 - rule, max_samples=16)
    with swmi.Engine([swmi.LatticeConfig(64, 64, 2.0), other]) as b:
        b.restore(ck, 1)
        b.run(40, 10)
        np.testing.assert_array_equal(b.spins(1), want_spins)
        got = b.rng_state(1)
        for k in ("states", "masks", "rot"):
            np.testing.assert_array_equal(got[k], want_rng[k])
        assert got["iteration"] == want_rng["iteration"] == 140
        np.testing.assert_array_equal(b.moments(1), want_mom)
        np.testing.assert_array_equal(b.series(1)[0], want_ser[0])


def test_restore_rejects_a_mismatched_lattice(swmi):
    with swmi.Engine(swmi.LatticeConfig(64, 64)) as a:
        ck = a.checkpoint()
    with swmi.Engine(swmi.LatticeConfig(128, 64)) as b:
        with pytest.raises(swmi.SwmiError):
            b.restore(ck)
        with pytest.raises(swmi.SwmiError):
            b.restore(b"not a checkpoint" * 8)


def test_error_bars_match_the_series(swmi):
    L, n_bins = 64, 20
    cfg = swmi.LatticeConfig(L, L, TC, seed=2, max_samples=2000)
    with swmi.Engine(cfg) as e:
        e.sweep(200)
        e.reset_measurements()
        e.run(20000, 10)
        eb = e.error_bars(0, n_bins)
        E, M = e.series()
    n = L * L
    ev = -E.astype(np.float64) / n
    mv = np.abs(M.astype(np.float64)) / n
    k = ev.size // n_bins
    b = ev[: k * n_bins].reshape(n_bins, k).mean(axis=1)
    bm = mv[: k * n_bins].reshape(n_bins, k).mean(axis=1)
    assert eb["samples"] == k * n_bins
    assert eb["e"] == pytest.approx(ev[: k * n_bins].mean(), rel=1e-12)
    assert eb["e_err"] == pytest.approx(b.std(ddof=1) / np.sqrt(n_bins), rel=1e-9)
    assert eb["abs_m"] == pytest.approx(mv[: k * n_bins].mean(), rel=1e-12)
    assert eb["abs_m_err"] == pytest.approx(bm.std(ddof=1) / np.sqrt(n_bins), rel=1e-9)
    t = float(np.float32(TC))
    assert eb["c"] == pytest.approx(ev[: k * n_bins].var() * n / t ** 2, rel=1e-9)
    assert 0 < eb["c_err"] < eb["c"]
    assert 0.1 < eb["tau_int_e"] < 50
