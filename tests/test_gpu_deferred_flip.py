"""The last cluster flip of a sweep is applied lazily (by the next sweep's bond pass, or before spins are
read). These tests interleave every call that reads or replaces spins with sweeps and check each state
against the oracle's per-site sweep, for 64-wide and narrow tiles and both flip rules."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

TC = 2.2691853


def _ref(n0, n1, T, seed, spin_seed, rule, sweeps):
    r = oracle.PerSite(n0, n1, T, seed=seed, spin_seed=spin_seed, flip_rule=rule)
    for _ in range(sweeps):
        r.sweep()
    return r


@pytest.mark.parametrize("n0,n1,rule", [(128, 64, 0), (128, 64, 1), (32, 32, 0), (32, 32, 1)])
def test_interleaved_reads_and_writes_match_oracle(swmi, n0, n1, rule):
    print(f"deferred flip {n0}x{n1} rule {rule}", flush=True)
    T, seed, spin_seed = TC, 41, 9
    cfg = swmi.LatticeConfig(n0, n1, T, seed=seed, spin_seed=spin_seed, flip_rule=rule)
    ref = _ref(n0, n1, T, seed, spin_seed, rule, 0)
    with swmi.Engine(cfg) as e:
        def advance(k):
            e.sweep(k)
            for _ in range(k):
                ref.sweep()

        advance(3)
        np.testing.assert_array_equal(e.spins(), ref.spins, err_msg="spins after 3 sweeps")
        advance(1)
        assert e.energy_magnetization() == oracle.observables(n0, n1, ref.spins), "observables after 4 sweeps"
        advance(2)
        ckpt = e.checkpoint()                       # taken at 6 sweeps, with the 6th flip pending
        advance(2)
        np.testing.assert_array_equal(e.spins(), ref.spins, err_msg="spins after 8 sweeps")
        e.restore(ckpt)
        at6 = _ref(n0, n1, T, seed, spin_seed, rule, 6)
        np.testing.assert_array_equal(e.spins(), at6.spins, err_msg="spins restored at 6 sweeps")
        e.sweep(2)
        np.testing.assert_array_equal(e.spins(), ref.spins, err_msg="8 sweeps again after the restore")
        e.labels()                                  # inspection: canonical labels of the last sweep
        advance(1)
        np.testing.assert_array_equal(e.spins(), ref.spins, err_msg="spins after a label download")
        new = np.random.default_rng(n0 + rule).integers(0, 2, n0 * n1).astype(np.uint8)
        advance(1)                                  # a flip is pending when the spins are replaced
        e.set_spins(new)
        ref.spins[:] = new
        advance(3)
        np.testing.assert_array_equal(e.spins(), ref.spins, err_msg="spins after an upload and 3 sweeps")
        assert e.error_flags() == 0


@pytest.mark.parametrize("rule", [0, 1])
def test_stopped_lattice_keeps_its_pending_flip(swmi, rule):
    """A lattice whose schedule ends keeps its last flip pending through the batch's further sweeps; the
    flip applied when its spins are read is the one of its own last sweep (same flip key)."""
    cfgs = [swmi.LatticeConfig(128, 128, TC, seed=1, spin_seed=2, flip_rule=rule),
            swmi.LatticeConfig(64, 64, 2.0, seed=2, spin_seed=3, flip_rule=rule)]
    with swmi.Engine(cfgs) as e:
        e.set_schedule(0, sweeps=7)
        e.sweep(20)
        assert e.progress(0)[0] == 7 and e.progress(1)[0] == 20
        np.testing.assert_array_equal(e.spins(0), _ref(128, 128, TC, 1, 2, rule, 7).spins)
        np.testing.assert_array_equal(e.spins(1), _ref(64, 64, 2.0, 2, 3, rule, 20).spins)
        assert e.error_flags(0) == 0 and e.error_flags(1) == 0
