"""Every even extent the reference accepts (ising/src/swendsen_wang.cpp:45-48 rejects only odd ones;
swendsen_wang_cpu.cpp:23-36,70-71 clips partial tiles), bit-exact against the oracle.

Extents that the wave64 tiling does not take as they are (a width that is neither a multiple nor a divisor
of 64, or a divisor of 64 with a height the row groups do not divide) are held with their rows padded to a
pitch P (swmi_get_layout). The padding is internal: spins, bonds and labels cross the ABI unpadded, and the
lcg32 groups are 64 consecutive padded sites (the oracle's or_ps_pitch layout)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TC = 2.2691853

# (n0, n1): the reference-trajectory extents of the oracle fixtures (48x34, 130x18), squares that are not
# multiples of 64, narrow widths that need padding, and widths a divisor of 64 with too few rows
PADDED = [(96, 96), (100, 100), (48, 34), (130, 18), (80, 66), (6, 8), (2, 2), (34, 6), (18, 10), (62, 2),
          (126, 4), (200, 2), (16, 6), (1000, 6)]


def _cfg(swmi, n0, n1, T=TC, **kw):
    return swmi.LatticeConfig(n0, n1, T, **kw)


@pytest.mark.parametrize("n0,n1", PADDED)
def test_layout_rng_init_and_hot_start(swmi, oracle, n0, n1):
    print(f"padded layout {n0}x{n1}", flush=True)
    with swmi.Engine(_cfg(swmi, n0, n1, seed=5, spin_seed=9)) as e:
        pitch, groups = e.layout()
        st = e.rng_state()
        s = e.spins()
        assert e.error_flags() == 0
    assert pitch == oracle.pitch(n0, n1) and groups == pitch * n1 // 64
    ref = oracle.PerSite(n0, n1, TC, seed=5, spin_seed=9)
    np.testing.assert_array_equal(st["states"], ref.states)
    np.testing.assert_array_equal(st["masks"], ref.masks)
    np.testing.assert_array_equal(s, oracle.hot_start(n0 * n1, 9))


@pytest.mark.parametrize("n0,n1", PADDED)
@pytest.mark.parametrize("tile_rows", [0, 1, 64])
def test_labels_of_injected_bonds_padded(swmi, oracle, n0, n1, tile_rows):
    print(f"padded labels {n0}x{n1} tile_rows {tile_rows}", flush=True)
    rng = np.random.default_rng(n0 * 7919 + n1 + tile_rows)
    with swmi.Engine(_cfg(swmi, n0, n1, tile_rows=tile_rows)) as e:
        for density in (0.0, 0.4, 0.5, 0.8, 1.0):
            bonds = ((rng.random(n0 * n1) < density).astype(np.uint8) |
                     ((rng.random(n0 * n1) < density).astype(np.uint8) << 1))
            np.testing.assert_array_equal(e.label_bonds(bonds), oracle.label(n0, n1, bonds), err_msg=f"density {density}")
        assert e.error_flags() == 0


@pytest.mark.parametrize("n0,n1", PADDED)
@pytest.mark.parametrize("flip_rule", [0, 1])
def test_thirty_sweeps_bit_exact_padded(swmi, oracle, n0, n1, flip_rule):
    """30 sweeps under both flip rules: bonds, labels and spins of the first sweeps, spins every 5th sweep,
    the lcg32 state at the end."""
    print(f"padded sweeps {n0}x{n1} rule {flip_rule}", flush=True)
    ref = oracle.PerSite(n0, n1, TC, seed=11, spin_seed=5, flip_rule=flip_rule)
    with swmi.Engine(_cfg(swmi, n0, n1, seed=11, spin_seed=5, flip_rule=flip_rule)) as e:
        e.record_bonds(True)
        for k in range(1, 31):
            e.sweep(1)
            if k <= 3:
                rb, rl = ref.sweep(record=True)
                np.testing.assert_array_equal(e.bonds(), rb, err_msg=f"bonds, sweep {k}")
                np.testing.assert_array_equal(e.labels(), rl, err_msg=f"labels, sweep {k}")
            else:
                ref.sweep()
            if k <= 3 or k % 5 == 0:
                np.testing.assert_array_equal(e.spins(), ref.spins, err_msg=f"spins, sweep {k}")
        st = e.rng_state()
        np.testing.assert_array_equal(st["states"], ref.states)
        np.testing.assert_array_equal(st["rot"], ref.rot)
        assert st["iteration"] == ref.iteration.value == 60
        assert e.energy_magnetization() == oracle.observables(n0, n1, ref.spins)
        assert e.error_flags() == 0


@pytest.mark.parametrize("n0,n1,tile_rows", [(100, 100, 1), (130, 18, 4), (80, 66, 2), (1000, 6, 0), (48, 34, 8)])
def test_padded_tile_heights_give_the_same_chain(swmi, oracle, n0, n1, tile_rows):
    ref = oracle.PerSite(n0, n1, 2.0, seed=3)
    with swmi.Engine(_cfg(swmi, n0, n1, 2.0, seed=3, tile_rows=tile_rows)) as e:
        for k in range(10):
            e.sweep(1)
            ref.sweep()
        np.testing.assert_array_equal(e.spins(), ref.spins)
        assert e.error_flags() == 0


def test_padded_and_unpadded_lattices_in_one_batch_equal_alone(swmi):
    cfgs = [_cfg(swmi, 100, 100, 2.0, seed=1, spin_seed=2), _cfg(swmi, 64, 64, 2.4, seed=2),
            _cfg(swmi, 48, 34, TC, seed=3, flip_rule=1), _cfg(swmi, 32, 32, 2.9, seed=4),
            _cfg(swmi, 130, 18, 1.8, seed=5), _cfg(swmi, 6, 8, 3.0, seed=6)]
    print("padded batch", flush=True)
    with swmi.Engine(cfgs) as batch:
        batch.run(40, 10)
        got = [(batch.spins(i), batch.moments(i)) for i in range(len(cfgs))]
    for i, c in enumerate(cfgs):
        with swmi.Engine(c) as alone:
            alone.run(40, 10)
            np.testing.assert_array_equal(alone.spins(), got[i][0], err_msg=f"lattice {i}")
            np.testing.assert_array_equal(alone.moments(), got[i][1], err_msg=f"lattice {i}")


@pytest.mark.parametrize("n0,n1", [(100, 100), (34, 6)])
def test_padded_upload_checkpoint_restore(swmi, oracle, n0, n1):
    rng = np.random.default_rng(4)
    spins = (rng.random(n0 * n1) < 0.5).astype(np.uint8)
    ref = oracle.PerSite(n0, n1, TC, seed=8)
    ref.spins[:] = spins
    with swmi.Engine(_cfg(swmi, n0, n1, seed=8, max_samples=16)) as e:
        e.set_spins(spins)
        np.testing.assert_array_equal(e.spins(), spins)
        e.sweep(7)
        for _ in range(7):
            ref.sweep()
        np.testing.assert_array_equal(e.spins(), ref.spins)
        ck = e.checkpoint()
        e.sweep(5)
        after = e.spins()
        e.restore(ck)
        e.sweep(5)
        np.testing.assert_array_equal(e.spins(), after)
        assert e.error_flags() == 0
