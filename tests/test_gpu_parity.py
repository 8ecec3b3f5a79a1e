"""Bit-exact parity of the HIP engine against the oracle (CPU restatement of the reference path).

Levels:
  1. lcg32 streams: per-site states, selector planes, rotations and iteration equal the oracle's
     LCG32_State<64> restatement for the same seeds (the oracle is pinned to the compiled reference).
  2. Given identical spins and bond fields, labels (minimum site index per cluster) and flipped spins
     equal the oracle labeller (restated reference labeller, pinned to BFS and reference trajectories).
  3. Whole sweeps: spins after k sweeps equal the oracle's per-site sweep, bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TC = 2.2691853


def _engine(swmi, n0, n1, T=TC, seed=0, spin_seed=1, tile_rows=0, max_samples=0):
    return swmi.Engine(swmi.LatticeConfig(n0, n1, T, seed=seed, spin_seed=spin_seed, tile_rows=tile_rows,
                                          max_samples=max_samples))


def _perm_labels_check(labels, n0, n1, bonds):
    """Cheap structural checks: root labels point to themselves and are minimal along bonds."""
    idx = np.arange(n0 * n1)
    assert np.all(labels <= idx)
    assert np.all(labels[labels] == labels)


EXTENTS = [(64, 64), (128, 96), (32, 32), (16, 64), (64, 2), (192, 40), (256, 256), (2, 32), (8, 16)]


@pytest.mark.parametrize("n0,n1", EXTENTS)
@pytest.mark.parametrize("seed", [0, 7])
def test_rng_init_matches_oracle(swmi, oracle, n0, n1, seed):
    print(f"rng init {n0}x{n1} seed {seed}", flush=True)
    with _engine(swmi, n0, n1, seed=seed) as e:
        st = e.rng_state()
        assert e.error_flags() == 0
    ref = oracle.PerSite(n0, n1, TC, seed=seed)
    np.testing.assert_array_equal(st["states"], ref.states)
    np.testing.assert_array_equal(st["masks"], ref.masks)
    np.testing.assert_array_equal(st["rot"], ref.rot)
    assert st["iteration"] == 0


@pytest.mark.parametrize("n0,n1,spin_seed", [(64, 64, 1), (1024, 1024, 1), (32, 32, 77), (320, 64, 123456)])
def test_hot_start_matches_oracle(swmi, oracle, n0, n1, spin_seed):
    print(f"hot start {n0}x{n1} spin_seed {spin_seed}", flush=True)
    with _engine(swmi, n0, n1, spin_seed=spin_seed) as e:
        s = e.spins()
    np.testing.assert_array_equal(s, oracle.hot_start(n0 * n1, spin_seed))


@pytest.mark.parametrize("n0,n1", EXTENTS)
@pytest.mark.parametrize("tile_rows", [0, 1, 5, 64])
def test_labels_of_injected_bonds(swmi, oracle, n0, n1, tile_rows):
    print(f"labels {n0}x{n1} tile_rows {tile_rows}", flush=True)
    rng = np.random.default_rng(n0 * 1000 + n1 + tile_rows)
    with _engine(swmi, n0, n1, tile_rows=tile_rows) as e:
        for density in (0.0, 0.3, 0.5, 0.75, 1.0):
            bonds = ((rng.random(n0 * n1) < density).astype(np.uint8) |
                     ((rng.random(n0 * n1) < density).astype(np.uint8) << 1))
            got = e.label_bonds(bonds)
            want = oracle.label(n0, n1, bonds)
            np.testing.assert_array_equal(got, want, err_msg=f"density {density}")
            _perm_labels_check(got, n0, n1, bonds)
        assert e.error_flags() == 0


@pytest.mark.parametrize("n0,n1,T", [(64, 64, TC), (128, 96, 2.0), (32, 32, 2.6), (16, 64, 1.5), (64, 2, 3.5),
                                     (192, 40, TC), (2, 32, TC), (8, 16, 2.0)])
def test_bonds_match_oracle(swmi, oracle, n0, n1, T):
    print(f"bonds {n0}x{n1} T {T}", flush=True)
    ref = oracle.PerSite(n0, n1, T, seed=3)
    with _engine(swmi, n0, n1, T=T, seed=3) as e:
        e.record_bonds(True)
        for k in range(4):
            e.sweep(1)
            b = e.bonds()
            rb, rl = ref.sweep(record=True)
            np.testing.assert_array_equal(b, rb, err_msg=f"bonds, sweep {k}")
            np.testing.assert_array_equal(e.labels(), rl, err_msg=f"labels, sweep {k}")
            np.testing.assert_array_equal(e.spins(), ref.spins, err_msg=f"spins, sweep {k}")
        assert e.error_flags() == 0


@pytest.mark.parametrize("n0,n1,T,tile_rows", [(64, 64, TC, 0), (128, 128, 2.0, 4), (256, 128, 2.6, 16),
                                               (32, 32, TC, 0), (512, 64, TC, 64), (128, 256, 1.0, 8)])
def test_sweeps_bit_exact(swmi, oracle, n0, n1, T, tile_rows):
    print(f"sweeps {n0}x{n1} T {T} tile_rows {tile_rows}", flush=True)
    ref = oracle.PerSite(n0, n1, T, seed=11, spin_seed=5)
    with _engine(swmi, n0, n1, T=T, seed=11, spin_seed=5, tile_rows=tile_rows) as e:
        for k in range(1, 31):
            e.sweep(1)
            ref.sweep()
            if k % 5 == 0:
                np.testing.assert_array_equal(e.spins(), ref.spins, err_msg=f"sweep {k}")
        st = e.rng_state()
        np.testing.assert_array_equal(st["states"], ref.states)
        np.testing.assert_array_equal(st["rot"], ref.rot)
        assert st["iteration"] == ref.iteration.value == 60
        assert e.error_flags() == 0


@pytest.mark.parametrize("n0,n1,tile_rows", [(64, 32, 16), (64, 16, 8), (32, 32, 8), (32, 32, 16), (64, 64, 32),
                                             (32, 64, 32)])
def test_blocks_with_empty_tiles_bit_exact(swmi, oracle, n0, n1, tile_rows):
    """A lattice shorter than one block (4 tiles of tile_rows wave rows) leaves the block's last tiles without
    rows, in every kernel class (64-wide and narrow tiles, heights 8, 16, 32); their wavefronts take no part in
    the labelling."""
    print(f"empty tiles {n0}x{n1} tile_rows {tile_rows}", flush=True)
    ref = oracle.PerSite(n0, n1, TC, seed=13, spin_seed=7)
    with _engine(swmi, n0, n1, seed=13, spin_seed=7, tile_rows=tile_rows) as e:
        assert e.tile_rows() == tile_rows
        for k in range(1, 21):
            e.sweep(1)
            ref.sweep()
            if k % 5 == 0:
                np.testing.assert_array_equal(e.spins(), ref.spins, err_msg=f"sweep {k}")
        assert e.error_flags() == 0


def test_batch_equals_alone(swmi):
    cfgs = [swmi.LatticeConfig(64, 64, 2.0, seed=1, spin_seed=2),
            swmi.LatticeConfig(32, 32, 2.9, seed=2, spin_seed=3, tile_rows=4),
            swmi.LatticeConfig(256, 128, TC, seed=3, spin_seed=4),
            swmi.LatticeConfig(128, 128, 1.8, seed=4, spin_seed=5, tile_rows=32)]
    print("batch vs alone", flush=True)
    with swmi.Engine(cfgs) as batch:
        batch.sweep(25)
        got = [batch.spins(i) for i in range(len(cfgs))]
    for i, c in enumerate(cfgs):
        with swmi.Engine(swmi.LatticeConfig(c.n0, c.n1, c.temperature, seed=c.seed, spin_seed=c.spin_seed)) as alone:
            alone.sweep(25)
            np.testing.assert_array_equal(alone.spins(), got[i], err_msg=f"lattice {i}")


@pytest.mark.parametrize("n0,n1", [(64, 64), (32, 32), (256, 64), (2, 32)])
def test_observables_match_oracle(swmi, oracle, n0, n1):
    with _engine(swmi, n0, n1) as e:
        for _ in range(3):
            e.sweep(3)
            assert e.energy_magnetization() == oracle.observables(n0, n1, e.spins())


def test_run_series_and_moments(swmi, oracle):
    n0 = n1 = 64
    ref = oracle.PerSite(n0, n1, TC, seed=9)
    with _engine(swmi, n0, n1, seed=9, max_samples=64) as e:
        e.run(200, 10)
        es, ms = e.series()
        mom = e.moments()
    want_e, want_m = [], []
    for k in range(200):
        ref.sweep()
        if (k + 1) % 10 == 0:
            ee, mm = oracle.observables(n0, n1, ref.spins)
            want_e.append(ee)
            want_m.append(mm)
    np.testing.assert_array_equal(es, want_e)
    np.testing.assert_array_equal(ms, want_m)
    ev = -np.array(want_e, dtype=np.float64) / (n0 * n1)
    mv = np.array(want_m, dtype=np.float64) / (n0 * n1)
    assert mom[0] == 20
    np.testing.assert_allclose(mom[1:], [ev.sum(), (ev * ev).sum(), np.abs(mv).sum(), (mv * mv).sum(), (mv ** 4).sum()],
                               rtol=1e-13)


@pytest.mark.parametrize("n,T", [(4096, TC), (4096, 2.0), (4096, 2.6), (8192, 2.0), (8192, TC), (8192, 2.6)])
def test_large_lattice_one_sweep_bit_exact(swmi, oracle, n, T):
    """Full-size lattices (the 4096^2 and 8192^2 configurations at all three temperatures): one sweep, bonds +
    labels + spins bit-exact against the oracle. T = 2.0 has one giant cluster (atomic contention in the
    merge), T = 2.6 many small ones."""
    n0 = n1 = n
    print(f"{n}^2 T {T} one sweep", flush=True)
    ref = oracle.PerSite(n0, n1, T, seed=1)
    with _engine(swmi, n0, n1, T=T, seed=1) as e:
        e.record_bonds(True)
        e.sweep(1)
        rb, rl = ref.sweep(record=True)
        np.testing.assert_array_equal(e.bonds(), rb)
        np.testing.assert_array_equal(e.labels(), rl)
        np.testing.assert_array_equal(e.spins(), ref.spins)
        assert e.error_flags() == 0


@pytest.mark.parametrize("T", [TC, 2.0])
def test_large_padded_lattice_one_sweep_bit_exact(swmi, oracle, T):
    """4100 x 4098: rows padded to a pitch of 4160 and a last block row of 2 lattice rows; 65 x 65 = 4225 blocks,
    so the merge unites each edge's root pairs in LDS before its lattice-wide unions (launches of at least 4096
    blocks). One sweep: bonds, labels and spins bit-exact against the oracle; then two more sweeps' spins."""
    n0, n1 = 4100, 4098
    print(f"{n0}x{n1} T {T} padded, local unions", flush=True)
    ref = oracle.PerSite(n0, n1, T, seed=5)
    with _engine(swmi, n0, n1, T=T, seed=5) as e:
        assert e.layout()[0] == 4160
        e.record_bonds(True)
        e.sweep(1)
        rb, rl = ref.sweep(record=True)
        np.testing.assert_array_equal(e.bonds(), rb)
        np.testing.assert_array_equal(e.labels(), rl)
        np.testing.assert_array_equal(e.spins(), ref.spins)
        e.sweep(2)
        ref.sweep()
        ref.sweep()
        np.testing.assert_array_equal(e.spins(), ref.spins)
        assert e.error_flags() == 0


@pytest.mark.parametrize("n0,n1,T,tile_rows", [(64, 64, TC, 0), (128, 96, 2.0, 4), (32, 32, 2.6, 0), (256, 128, TC, 16),
                                               (16, 64, 1.5, 2), (512, 64, 3.0, 64)])
def test_random_flip_rule_bit_exact(swmi, oracle, n0, n1, T, tile_rows):
    """SWMI_FLIP_RANDOM: spins of every sweep equal the oracle's restatement; site 0 is not frozen."""
    print(f"random flips {n0}x{n1} T {T}", flush=True)
    ref = oracle.PerSite(n0, n1, T, seed=21, spin_seed=3, flip_rule=1)
    s0 = []
    with swmi.Engine(swmi.LatticeConfig(n0, n1, T, seed=21, spin_seed=3, tile_rows=tile_rows, flip_rule=1)) as e:
        for k in range(1, 21):
            e.sweep(1)
            ref.sweep()
            s = e.spins()
            np.testing.assert_array_equal(s, ref.spins, err_msg=f"sweep {k}")
            s0.append(int(s[0]))
        assert e.error_flags() == 0
    assert len(set(s0)) == 2, "site 0 never changed under the random rule"


def test_parity_rule_freezes_site_zero_on_the_gpu(swmi):
    with _engine(swmi, 128, 128, seed=4) as e:
        first = int(e.spins()[0])
        for _ in range(10):
            e.sweep(3)
            assert int(e.spins()[0]) == first


@pytest.mark.parametrize("n,T,sweeps", [(4096, 2.0, 40), (8192, TC, 10)])
def test_large_lattice_invariants_over_sweeps(swmi, oracle, n, T, sweeps):
    """Size-independent properties at full size after many sweeps: canonical labels (root = minimum
    index, idempotent), the parity rule's frozen site 0, observables equal to a brute-force sum."""
    with _engine(swmi, n, n, T=T, seed=2) as e:
        s0 = int(e.spins()[0])
        e.sweep(sweeps)
        labels = e.labels()
        spins = e.spins()
        assert e.error_flags() == 0
        idx = np.arange(n * n, dtype=np.uint32)
        assert np.all(labels <= idx)
        assert np.all(labels[labels] == labels)
        assert int(spins[0]) == s0
        assert e.energy_magnetization() == oracle.observables(n, n, spins)


def test_engine_threshold_equals_the_reference_fixture(swmi):
    """p_add and the integer threshold the engine derives on this host (glibc expf under the engine's
    flags) equal the bits of the compiled reference (padd.json) and the exhaustively checked thresholds
    (thresholds.json), for every fixture temperature."""
    import json
    from conftest import GOLDEN
    fixture = json.loads((GOLDEN / "thresholds.json").read_text())
    cfgs = [swmi.LatticeConfig(64, 64, float(np.float32(float(t))), seed=1) for t in fixture]
    with swmi.Engine(cfgs) as e:
        for i, (t, want) in enumerate(fixture.items()):
            p, thr = e.threshold(i)
            assert int(np.array([p], dtype=np.float32).view(np.uint32)[0]) == want["p_add_bits"], t
            assert thr == want["threshold"], t
            e.set_temperature(float(np.float32(float(t))) * 1.0, i)
            assert e.threshold(i)[1] == want["threshold"], t
