"""The oracle (CPU restatement of the reference path) against the reference's golden vectors and against
what the mathematics fixes. CPU only.

Each check is chosen so that a plausible mistake in the restatement fails it: a wrong drand48 scaling,
a lane rotation off by one, a skipped state[0], a bond decided on the wrong draw, a dropped merge, a
flip of the wrong parity.
"""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from conftest import GOLDEN, ROOT

TEMPS_REF = ["2.0", "2.2691853", "2.6"]


# ------------------------------------------------------------------------------------------------
# drand48: restated from POSIX, checked against glibc's own srand48/drand48 (library routine)
# ------------------------------------------------------------------------------------------------
def test_drand48_matches_glibc(oracle):
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    for seed in (0, 1, 2, 77, 123456789, 2**32 - 1):
        libc.srand48(seed)
        x = ctypes.c_uint64()
        oracle.lib().or_srand48(ctypes.byref(x), seed & 0xFFFFFFFF)
        for _ in range(1000):
            assert oracle.lib().or_drand48(ctypes.byref(x)) == libc.drand48()


def test_hot_start_matches_reference_lattice(oracle):
    """Lattice<2> constructor (lattice.cpp:15-27): the sweep-0 snapshots of the reference runs."""
    g = np.load(GOLDEN / "refcpu.npz")
    meta = json.loads(bytes(g["meta"]).decode())
    for m in meta:
        spins0 = g[m["name"] + "_spins"][0]
        np.testing.assert_array_equal(oracle.hot_start(m["n0"] * m["n1"], 1), spins0)


# ------------------------------------------------------------------------------------------------
# lcg32
# ------------------------------------------------------------------------------------------------
def test_lcg32_streams_match_reference(oracle):
    g = np.load(GOLDEN / "lcg32_streams.npz")
    for key in g.files:
        w = int(key.split("_")[0][1:])
        seed = int(key.split("seed")[1])
        ref = g[key]
        got = np.array(oracle.lcg32_states(w, seed, ref.shape[0] - 1), dtype=np.uint32)
        np.testing.assert_array_equal(got, ref, err_msg=key)


def test_lcg32_buffered_matches_reference(oracle):
    for case in json.loads((GOLDEN / "lcg32_buffered.json").read_text()):
        got = oracle.lcg32_buffered(case["w"], case["seed"], case["sizes"])
        assert got == case["calls"]


def test_lcg32_shuffle_invariant(oracle):
    """Between shuffles every lane is a plain LCG with its own (a, c); every 15th update rotates (a, c)."""
    s = oracle.Lcg32()
    oracle.lib().or_lcg32_init(ctypes.byref(s), 64, 5)
    for it in range(1, 60):
        before = list(s.state)
        a_before = list(s.a)
        oracle.lib().or_lcg32_update(ctypes.byref(s))
        if it % 15:
            assert list(s.a) == a_before
            for i in range(64):
                assert s.state[i] == (a_before[i] * before[i] + s.c[i]) & 0xFFFFFFFF
        else:
            shift = (before[it & 63] + (0 if it & 1 else 1)) & 63
            assert list(s.a) == [a_before[(i + shift) & 63] for i in range(64)]


def test_per_site_layout_equals_full_lcg32_state(oracle):
    """The engine's compact group state (selector planes + rotation) reproduces LCG32_State<64>."""
    n_groups = 5
    seed = 31
    states = np.empty(64 * n_groups, dtype=np.uint32)
    masks = np.empty(3 * n_groups, dtype=np.uint64)
    rot = np.empty(n_groups, dtype=np.uint32)
    oracle.lib().or_ps_init(n_groups, seed, states.ctypes.data, masks.ctypes.data, rot.ctypes.data)
    full = []
    for g in range(n_groups):
        s = oracle.Lcg32()
        oracle.lib().or_lcg32_init(ctypes.byref(s), 64, oracle.lib().or_group_seed(seed, g))
        full.append(s)
    for it in range(200):
        oracle.lib().or_ps_rng_update(n_groups, states.ctypes.data, masks.ctypes.data, rot.ctypes.data, it)
        for g in range(n_groups):
            oracle.lib().or_lcg32_update(ctypes.byref(full[g]))
            np.testing.assert_array_equal(states[64 * g:64 * g + 64], np.array(full[g].state[:], dtype=np.uint32))
    assert oracle.lib().or_group_seed(0, 0) == 1 and oracle.lib().or_group_seed(0, 9) == 10


@pytest.mark.parametrize("n0,n1", [(64, 64), (96, 96), (100, 100), (48, 34), (130, 18), (80, 66), (2, 2), (6, 8),
                                   (16, 6), (32, 32), (62, 2), (4, 16), (8, 24), (1000, 10)])
def test_pitch_is_the_smallest_width_the_wave64_tiling_takes(oracle, n0, n1):
    """The padded row pitch P: P >= n0, and either P % 64 == 0 (64-site groups within a row) or P divides
    64 with n1 a multiple of 64 / P (groups of whole padded rows); no smaller width of that form exists."""
    def ok(p):
        return p >= n0 and (p % 64 == 0 or (64 % p == 0 and n1 % (64 // p) == 0))
    p = oracle.pitch(n0, n1)
    assert ok(p) and p % 2 == 0
    assert not any(ok(q) for q in range(n0, p))
    if n0 % 64 == 0 or (64 % n0 == 0 and n1 % (64 // n0) == 0):
        assert p == n0   # every extent without padding keeps its layout


@pytest.mark.parametrize("n0,n1,t", [(100, 6, 2.2691853), (48, 34, 2.0), (130, 4, 2.6), (6, 8, 2.2691853), (2, 2, 1.5)])
def test_padded_engine_layout_draws_from_the_padded_groups(oracle, n0, n1, t):
    """The engine layout's bonds at an extent that needs padding, re-derived independently: every padded
    group is a full LCG32_State<64> (pinned against the compiled reference), the draws of site (x, y) are
    lane (y*P + x) % 64 of group (y*P + x) // 64 after the first and second update, and a bond needs equal
    spins and a draw below the threshold."""
    seed = 7
    ps = oracle.PerSite(n0, n1, t, seed=seed)
    P = ps.pitch
    spins0 = ps.spins.copy()
    full = []
    for g in range(P * n1 // 64):
        st = oracle.Lcg32()
        oracle.lib().or_lcg32_init(ctypes.byref(st), 64, oracle.lib().or_group_seed(seed, g))
        full.append(st)
    draws = []
    for _ in range(2):
        for st in full:
            oracle.lib().or_lcg32_update(ctypes.byref(st))
        draws.append(np.concatenate([np.array(st.state[:], dtype=np.uint32) for st in full]).reshape(n1, P)[:, :n0])
    bonds, labels = ps.sweep(record=True)
    s = spins0.reshape(n1, n0)
    thr = oracle.threshold(t)
    want = ((s == np.roll(s, -1, axis=1)) & (draws[0] < thr)).astype(np.uint8)
    want |= ((s == np.roll(s, -1, axis=0)) & (draws[1] < thr)).astype(np.uint8) << 1
    np.testing.assert_array_equal(bonds.reshape(n1, n0), want)
    np.testing.assert_array_equal(labels, components_min_label(n0, n1, bonds))
    np.testing.assert_array_equal(ps.states.reshape(-1, 64), np.array([st.state[:] for st in full], dtype=np.uint32))


# ------------------------------------------------------------------------------------------------
# bond probability
# ------------------------------------------------------------------------------------------------
def test_p_add_bits_match_reference(oracle):
    ref = json.loads((GOLDEN / "padd.json").read_text())
    for t, bits in ref.items():
        got = np.array([oracle.p_add(float(np.float32(float(t))))], dtype=np.float32).view(np.uint32)[0]
        assert int(got) == bits, t
    # bit patterns quoted in the survey of this path (glibc 2.35 expf)
    assert ref["2.2691853"] == 0x3F15F61A and ref["2.0"] == 0x3F21D2A7 and ref["2.6"] == 0x3F0960A0


@pytest.mark.parametrize("t", TEMPS_REF)
def test_threshold_exhaustive(oracle, t):
    """u < threshold equals 2^-32f * (float)u < p_add for all 2^32 draws."""
    p = oracle.lib().or_p_add(float(t))
    thr = oracle.lib().or_threshold(p)
    assert oracle.lib().or_threshold_mismatches(p, thr) == 0
    known = {"2.2691853": 2515933568, "2.0": 2714936961, "2.6": 2304810880}
    assert thr == known[t]


# ------------------------------------------------------------------------------------------------
# labeller: the restated reference labeller against connected components found independently
# ------------------------------------------------------------------------------------------------
def components_min_label(n0, n1, bonds):
    idx = np.arange(n0 * n1).reshape(n1, n0)
    right = np.roll(idx, -1, axis=1).ravel()
    down = np.roll(idx, -1, axis=0).ravel()
    bx = (bonds & 1).astype(bool)
    by = (bonds & 2).astype(bool)
    rows = np.concatenate([np.arange(n0 * n1)[bx], np.arange(n0 * n1)[by]])
    cols = np.concatenate([right[bx], down[by]])
    g = coo_matrix((np.ones(rows.size), (rows, cols)), shape=(n0 * n1, n0 * n1))
    _, comp = connected_components(g, directed=False)
    mins = np.full(comp.max() + 1, n0 * n1, dtype=np.int64)
    np.minimum.at(mins, comp, np.arange(n0 * n1))
    return mins[comp].astype(np.uint32)


@pytest.mark.parametrize("n0,n1", [(2, 2), (4, 6), (16, 8), (32, 32), (48, 34), (64, 64), (130, 18), (2, 64)])
@pytest.mark.parametrize("tile", [(16, 8), (4, 4), (64, 2)])
def test_label_equals_connected_components(oracle, n0, n1, tile):
    rng = np.random.default_rng(n0 * 7 + n1 + tile[0])
    for density in (0.0, 0.2, 0.5, 0.8, 1.0):
        bonds = ((rng.random(n0 * n1) < density).astype(np.uint8) | ((rng.random(n0 * n1) < density).astype(np.uint8) << 1))
        np.testing.assert_array_equal(oracle.label(n0, n1, bonds, tile), components_min_label(n0, n1, bonds))


def test_observables_bruteforce(oracle):
    rng = np.random.default_rng(3)
    for n0, n1 in [(2, 2), (8, 6), (64, 32)]:
        s = rng.integers(0, 2, n0 * n1).astype(np.uint8)
        v = 2 * s.reshape(n1, n0).astype(np.int64) - 1
        e = int((v * (np.roll(v, -1, 1) + np.roll(v, -1, 0))).sum())
        assert oracle.observables(n0, n1, s) == (e, int(v.sum()))


# ------------------------------------------------------------------------------------------------
# the whole reference CPU path, bit-exact against the compiled reference's trajectories
# ------------------------------------------------------------------------------------------------
def test_reference_cpu_trajectories(oracle):
    g = np.load(GOLDEN / "refcpu.npz")
    meta = json.loads(bytes(g["meta"]).decode())
    for m in meta:
        n0, n1, T = m["n0"], m["n1"], float(np.float32(float(m["temperature"])))
        spins = oracle.hot_start(n0 * n1, 1)
        r = oracle.RefCpu(n0, n1, 16, m["num_threads"])
        done = 0
        for i, k in enumerate(m["sweeps"]):
            while done < k:
                r.sweep(spins, T)
                done += 1
            np.testing.assert_array_equal(spins, g[m["name"] + "_spins"][i], err_msg=f"{m['name']} after {k} sweeps")
            assert oracle.observables(n0, n1, spins) == tuple(int(v) for v in g[m["name"] + "_sums"][i])


def test_parity_flip_freezes_site_zero(oracle):
    """Site 0 is always the minimum of its cluster (even root), so the reference rule never flips it."""
    p = oracle.PerSite(64, 64, 2.2691853, seed=4)
    s0 = p.spins[0]
    for _ in range(50):
        p.sweep()
        assert p.spins[0] == s0


@pytest.mark.skipif(not (ROOT / "oracle" / "_ref" / "ref_sw").exists(), reason="reference binaries are built only where the reference sources exist")
def test_reference_cpu_fresh_run(oracle, tmp_path):
    """Directly against a fresh run of the compiled reference (beyond the committed fixtures)."""
    n0, n1, T, nt, ks = 80, 66, "2.35", 6, [1, 9]
    out = tmp_path / "d.bin"
    subprocess.run([str(ROOT / "oracle" / "_ref" / "ref_sw"), "dump", str(n0), str(n1), T, str(out)] + [str(k) for k in ks],
                   check=True, capture_output=True, env=dict(os.environ, NUM_THREADS=str(nt)))
    raw = np.fromfile(out, dtype=np.uint8)
    spins = oracle.hot_start(n0 * n1, 1)
    r = oracle.RefCpu(n0, n1, 16, nt)
    done = 0
    for i, k in enumerate(ks):
        while done < k:
            r.sweep(spins, float(np.float32(float(T))))
            done += 1
        np.testing.assert_array_equal(spins, raw[i * (n0 * n1 + 16):i * (n0 * n1 + 16) + n0 * n1])


# ------------------------------------------------------------------------------------------------
# exact finite-lattice values (the statistical oracle)
# ------------------------------------------------------------------------------------------------
def test_exact_values_fixture():
    ex = json.loads((GOLDEN / "exact_values.json").read_text())
    # README.md:98 quotes U/N at the critical coupling on 128^2
    assert abs(ex["critical_128"]["U_per_spin"] - (-1.419076272084983)) < 1e-12
    vals = {(v["L"], v["T"]): v for v in ex["values"]}
    # the values quoted for this path at K = 1/float(T)
    assert abs(vals[(128, "2.2691853")]["U_per_spin"] - (-1.419076296360968)) < 1e-12
    assert abs(vals[(1024, "2.2691853")]["U_per_spin"] - (-1.414821447346456)) < 1e-12
    assert abs(vals[(4096, "2.2691853")]["C_per_spin"] - 4.25155843809) < 1e-8
    # away from T_c the finite-size effects vanish: every L gives the same U/N at T = 2.0
    assert abs(vals[(128, "2.0")]["U_per_spin"] - vals[(8192, "2.0")]["U_per_spin"]) < 1e-12


@pytest.mark.skipif(not (ROOT / "oracle" / "_ref" / "ff_exact").exists(), reason="reference binaries are built only where the reference sources exist")
def test_exact_values_match_reference_tool():
    """The reference's long-double closed-form tool agrees with the numerical-derivative route."""
    ex = json.loads((GOLDEN / "exact_values.json").read_text())
    for v in ex["values"]:
        if v["L"] > 128:
            continue
        out = subprocess.run([str(ROOT / "oracle" / "_ref" / "ff_exact"), repr(v["K"]), str(v["L"]), str(v["L"])],
                             check=True, capture_output=True, text=True).stdout.split()
        assert abs(float(out[1]) - v["U_per_spin"]) < 1e-9
        assert abs(float(out[2]) - v["C_per_spin"]) < 1e-6


# MurmurHash3_x86_32 test vectors published with SMHasher (Austin Appleby) and widely reproduced.
MURMUR3_KATS = [(b"", 0, 0x00000000), (b"", 1, 0x514E28B7), (b"", 0xFFFFFFFF, 0x81F16F39),
                (b"\x00\x00\x00\x00", 0, 0x2362F9DE), (b"\xff\xff\xff\xff", 0, 0x76293B50),
                (b"\x21\x43\x65\x87", 0, 0xF55B516B), (b"a", 0x9747B28C, 0x7FA09EA6), (b"aa", 0x9747B28C, 0x5D211726),
                (b"aaa", 0x9747B28C, 0x283E0130), (b"aaaa", 0x9747B28C, 0x5A97808A), (b"abcd", 0x9747B28C, 0xF0478627),
                (b"Hello, world!", 0x9747B28C, 0x24884CBA),
                (b"The quick brown fox jumps over the lazy dog", 0x9747B28C, 0x2FA826CD)]


@pytest.mark.parametrize("data,seed,want", MURMUR3_KATS)
def test_murmur3_matches_published_vectors(oracle, data, seed, want):
    assert oracle.murmur3_32(data, seed) == want


def test_flip_key_is_murmur3_of_the_iteration(oracle):
    for seed, it in [(0, 0), (7, 2), (0xDEADBEEF, 123456), (1, 0xFFFFFFFE)]:
        assert oracle.flip_key(seed, it) == oracle.murmur3_32(int(it).to_bytes(4, "little"), seed)


def test_random_flip_flips_whole_clusters_half_of_the_time(oracle):
    """Every site of a cluster flips together, the decision depends on the root only, and over many roots
    about half of the clusters flip."""
    import ctypes
    rng = np.random.default_rng(5)
    n0, n1 = 64, 64
    bonds = ((rng.random(n0 * n1) < 0.4).astype(np.uint8) | ((rng.random(n0 * n1) < 0.4).astype(np.uint8) << 1))
    labels = oracle.label(n0, n1, bonds)
    spins = rng.integers(0, 2, n0 * n1).astype(np.uint8)
    key = oracle.flip_key(3, 10)
    out = spins.copy()
    oracle.lib().or_flip_random(out.size, out.ctypes.data, labels.ctypes.data, key)
    flipped = out != spins
    roots = np.unique(labels)
    for r in roots:
        members = labels == r
        assert flipped[members].all() or not flipped[members].any()
        assert bool(flipped[members][0]) == bool(oracle.fmix32(int(r) ^ key) >> 31)
    frac = np.mean([flipped[labels == r][0] for r in roots])
    assert 0.4 < frac < 0.6, frac
    # site 0 is not frozen under this rule: over keys its cluster flips about half of the time
    hits = [oracle.fmix32(int(labels[0]) ^ oracle.flip_key(3, 2 * k)) >> 31 for k in range(400)]
    assert 0.4 < np.mean(hits) < 0.6


PROTOCOL = json.loads((GOLDEN / "protocol.json").read_text())


@pytest.mark.parametrize("case", PROTOCOL["cases"], ids=lambda c: f"{c['n0']}x{c['n1']}-T{c['temperature']}-t{c['num_threads']}")
def test_reference_driver_protocol(oracle, case):
    """The restated reference CPU path under the reference driver's measurement loop (1000 warmup sweeps,
    a sample every 10, FP64 sums) prints the same report lines as the compiled reference Ising2D."""
    n0, n1 = case["n0"], case["n1"]
    r = oracle.RefCpu(n0, n1, 16, case["num_threads"])
    s = oracle.hot_start(n0 * n1, 1)
    t = float(np.float32(float(case["temperature"])))
    e, c, m = r.measure(s, t, 1000, case["sweeps"], 10)
    assert ("%g" % e, "%g" % c, "%g" % m) == (case["report"]["e"], case["report"]["c"], case["report"]["m"])


@pytest.mark.skipif(os.environ.get("SWMI_SLOW_TESTS") != "1", reason="3 minutes on one core; set SWMI_SLOW_TESTS=1")
def test_reference_readme_sample(oracle):
    """The reference's published sample (README.md:66-82): NUM_THREADS=32, 128x128 at T_c, 1000 warmup +
    100000 sweeps on its CPU path prints E/N -1.41921, C/N 2.54411, |m| 0.550833."""
    r = oracle.RefCpu(128, 128, 16, 32)
    s = oracle.hot_start(128 * 128, 1)
    e, c, m = r.measure(s, float(np.float32(2.2691853142130221)), 1000, 100000, 10)
    assert ("%g" % e, "%g" % c, "%g" % m) == ("-1.41921", "2.54411", "0.550833")


def test_threshold_fixture_is_the_boundary_of_the_float_compare(oracle):
    """thresholds.json (written with an exhaustive check over all 2^32 draws) holds, for every p_add of the
    compiled reference, the first draw u at which 2^-32f * (float)u < p_add fails: (float)u is monotone in u,
    so the compare holds below t and fails from t on exactly when it holds at t - 1 and fails at t."""
    fixture = json.loads((GOLDEN / "thresholds.json").read_text())
    padd = json.loads((GOLDEN / "padd.json").read_text())
    two32 = np.float32(2.3283064370807974e-10)
    for t, v in fixture.items():
        assert v["p_add_bits"] == padd[t]
        p = np.array([v["p_add_bits"]], dtype=np.uint32).view(np.float32)[0]
        thr = v["threshold"]
        assert two32 * np.float32(thr - 1) < p and not (thr < 2 ** 32 and two32 * np.float32(thr) < p), t
        assert oracle.threshold(float(np.float32(float(t)))) == thr


def test_reference_cpu_statistics_fixture_reproduces_the_readme_sample():
    """ref_cpu_stats.json holds the compiled reference's own 128^2 runs; at T_c it is the README's published
    run (README.md:66-82: E/N -1.41921, C/N 2.54411, |m| 0.550833 at 6 significant digits)."""
    ref = json.loads((GOLDEN / "ref_cpu_stats.json").read_text())["values"]
    tc = ref["2.2691853"]
    assert (f"{tc['e']:g}", f"{tc['c']:g}", f"{tc['abs_m']:g}") == ("-1.41921", "2.54411", "0.550833")
    exact = {(v["L"], v["T"]): v for v in json.loads((GOLDEN / "exact_values.json").read_text())["values"]}
    for t, v in ref.items():
        assert v["samples"] == 10000 and v["e_err"] > 0 and v["c_err"] > 0 and v["abs_m_err"] > 0
        # the reference's chain is itself within 3 standard errors of the exact values
        ex = exact[(128, t)]
        assert abs(v["e"] - ex["U_per_spin"]) < 3 * v["e_err"], t
        assert abs(v["c"] - ex["C_per_spin"]) < 3 * v["c_err"], t
