"""Batched, graph-captured, scheduled and served runs give the same per-lattice results as the plain path.

Spins and lcg32 streams are per lattice and labels are canonical, so a lattice's trajectory does not depend
on which other lattices share its launches, on the tile height, on graph capture or on the chunking of
the serving loop. These tests check that bit for bit, against lone runs on the GPU and against the
oracle's restatement of the engine's per-site sweep.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TC = 2.2691853


def _cfg(swmi, n0, n1, T=TC, seed=0, spin_seed=1, tile_rows=0, max_samples=0):
    return swmi.LatticeConfig(n0, n1, T, seed=seed, spin_seed=spin_seed, tile_rows=tile_rows, max_samples=max_samples)


MIXED = [(32, 32, 2.9, 4), (64, 64, 2.0, 8), (128, 128, TC, 16), (256, 256, 1.8, 32), (16, 64, 2.5, 2),
         (512, 512, 2.6, 64), (64, 128, TC, 1), (1024, 1024, 2.4, 0)]


def test_graph_replays_equal_eager(swmi):
    """Graph replays (16-sweep units, and sample groups) with eager work and syncs in between, each checked."""
    cfgs = [_cfg(swmi, n0, n1, T, seed=i + 1, spin_seed=i + 2, tile_rows=tr, max_samples=64)
            for i, (n0, n1, T, tr) in enumerate(MIXED[:5])]
    print("graph vs eager", flush=True)
    with swmi.Engine(cfgs) as g, swmi.Engine(cfgs) as e:
        e.set_graphs(False)
        for rep in range(3):
            g.sweep(32)
            e.sweep(32)
            for i in range(len(cfgs)):
                np.testing.assert_array_equal(g.spins(i), e.spins(i), err_msg=f"replay {rep} lattice {i}")
                assert g.energy_magnetization(i) == e.energy_magnetization(i)
            g.synchronize()
        for rep in range(2):
            g.run(130, 10)   # 13 groups: two 6-group graph replays and one eager group
            e.run(130, 10)
            for i in range(len(cfgs)):
                np.testing.assert_array_equal(g.moments(i), e.moments(i), err_msg=f"run {rep} lattice {i}")
                ge, gm = g.series(i)
                ee, em = e.series(i)
                np.testing.assert_array_equal(ge, ee)
                np.testing.assert_array_equal(gm, em)
        for i in range(len(cfgs)):
            assert g.rng_state(i)["iteration"] == e.rng_state(i)["iteration"] == 2 * (96 + 260)
            assert g.error_flags(i) == 0


def test_long_graph_replays_equal_eager(swmi):
    """A long sweep loop replays 128-sweep graphs and finishes with 16-sweep graphs and eager sweeps: 300 sweeps
    are two long replays, two short ones and 12 eager sweeps, twice in a row on the same graphs."""
    cfgs = [_cfg(swmi, n0, n1, T, seed=i + 31, spin_seed=i + 32, tile_rows=tr) for i, (n0, n1, T, tr) in enumerate(MIXED[:4])]
    print("long graph vs eager", flush=True)
    with swmi.Engine(cfgs) as g, swmi.Engine(cfgs) as e:
        e.set_graphs(False)
        for rep in range(2):
            g.sweep(300)
            e.sweep(300)
            for i in range(len(cfgs)):
                np.testing.assert_array_equal(g.spins(i), e.spins(i), err_msg=f"sweep loop {rep} lattice {i}")
        for i in range(len(cfgs)):
            assert g.rng_state(i)["iteration"] == e.rng_state(i)["iteration"] == 2 * 600
            assert g.error_flags(i) == 0


def test_graph_cache_eviction_equals_eager(swmi):
    """More graph shapes than the engine keeps (plain 16- and 128-sweep units, and sampled runs with six
    different n_sep): the cache is emptied while earlier replays may still be queued, and every result still
    equals the eager engine's, spins, moments and series alike."""
    cfgs = [_cfg(swmi, n0, n1, T, seed=i + 41, spin_seed=i + 42, tile_rows=tr, max_samples=512)
            for i, (n0, n1, T, tr) in enumerate(MIXED[:3])]
    print("graph cache eviction vs eager", flush=True)
    with swmi.Engine(cfgs) as g, swmi.Engine(cfgs) as e:
        e.set_graphs(False)
        steps = [("sweep", 32), ("sweep", 256)] + [("run", sep) for sep in (2, 3, 4, 5, 6, 7)]
        for kind, k in steps:
            for eng in (g, e):
                if kind == "sweep":
                    eng.sweep(k)
                else:
                    eng.run(128, k)
            for i in range(len(cfgs)):
                np.testing.assert_array_equal(g.spins(i), e.spins(i), err_msg=f"{kind} {k} lattice {i}")
                np.testing.assert_array_equal(g.moments(i), e.moments(i), err_msg=f"{kind} {k} lattice {i}")
                ge, gm = g.series(i)
                ee, em = e.series(i)
                np.testing.assert_array_equal(ge, ee)
                np.testing.assert_array_equal(gm, em)
        for i in range(len(cfgs)):
            assert g.error_flags(i) == 0


def test_mixed_batch_equals_alone_and_oracle(swmi, oracle):
    """Lattices of every kernel class (narrow and 64-wide tiles, tile heights 1..64) in one engine."""
    cfgs = [_cfg(swmi, n0, n1, T, seed=10 + i, spin_seed=20 + i, tile_rows=tr) for i, (n0, n1, T, tr) in enumerate(MIXED)]
    print("mixed batch", flush=True)
    with swmi.Engine(cfgs) as batch:
        batch.sweep(40)
        got = [batch.spins(i) for i in range(len(cfgs))]
        assert all(batch.error_flags(i) == 0 for i in range(len(cfgs)))
    for i, c in enumerate(cfgs):
        with swmi.Engine(_cfg(swmi, c.n0, c.n1, c.temperature, seed=c.seed, spin_seed=c.spin_seed)) as alone:
            alone.sweep(40)
            np.testing.assert_array_equal(alone.spins(), got[i], err_msg=f"lattice {i} ({c.n0}x{c.n1})")
        if c.n0 * c.n1 <= 128 * 128:
            ref = oracle.PerSite(c.n0, c.n1, c.temperature, seed=c.seed, spin_seed=c.spin_seed)
            for _ in range(40):
                ref.sweep()
            np.testing.assert_array_equal(ref.spins, got[i], err_msg=f"oracle, lattice {i}")


def test_schedule_stops_one_lattice(swmi):
    cfgs = [_cfg(swmi, 128, 128, seed=1), _cfg(swmi, 64, 64, 2.0, seed=2)]
    with swmi.Engine(cfgs) as e:
        e.set_schedule(0, sweeps=7)
        e.sweep(20)
        assert e.progress(0)[0] == 7 and e.progress(1)[0] == 20
        s0 = e.spins(0)
    with swmi.Engine(cfgs[0]) as alone:
        alone.sweep(7)
        np.testing.assert_array_equal(alone.spins(), s0)


@pytest.mark.parametrize("sweeps,warmup,extra", [(100, 20, 0), (100, 30, 47), (60, 0, 13)])
def test_scheduled_samples_equal_plain_protocol(swmi, sweeps, warmup, extra):
    """warmup + sweeps with a sample every 10th, run inside a longer loop, equals the plain protocol."""
    c = _cfg(swmi, 128, 64, 2.1, seed=5, spin_seed=6, max_samples=32)
    with swmi.Engine([c, _cfg(swmi, 64, 64, seed=7)]) as e:
        e.set_schedule(0, sweeps=warmup + sweeps, warmup=warmup, n_sep=10)
        e.run_scheduled(warmup + sweeps + extra, 10)
        mom_s = e.moments(0)
        ser_s = e.series(0)
        assert e.progress(0) == (warmup + sweeps, sweeps // 10)
    with swmi.Engine(c) as p:
        p.sweep(warmup)
        p.reset_measurements()
        p.run(sweeps, 10)
        np.testing.assert_array_equal(p.moments(0), mom_s)
        np.testing.assert_array_equal(p.series(0)[0], ser_s[0])


def test_autotune_picks_a_height_and_results_do_not_depend_on_it(swmi):
    c = _cfg(swmi, 256, 256, seed=3)
    best, ms = swmi.autotune(c, sweeps=10)
    print(f"autotune 256^2: best {best}, ms {ms}", flush=True)
    assert best in (1, 2, 4, 8, 16, 32, 64)
    assert min(m for m in ms if m > 0) == pytest.approx(ms[[1, 2, 4, 8, 16, 32, 64].index(best)])
    with swmi.Engine(c) as tuned, swmi.Engine(_cfg(swmi, 256, 256, seed=3, tile_rows=16 if best != 16 else 4)) as other:
        assert tuned.tile_rows() == best
        tuned.sweep(20)
        other.sweep(20)
        np.testing.assert_array_equal(tuned.spins(), other.spins())


def _measured(sweeps, n_sep=10):
    """Measured sweeps of a served job: whole groups of n_sep (ising_2d.cpp:132 loops while i < n_sweeps)."""
    return -(-sweeps // n_sep) * n_sep


def _plain_job_result(swmi, job):
    """The same job through the plain path: warmup, reset, then samples every 10 sweeps."""
    c = _cfg(swmi, job.n0, job.n1, job.temperature, seed=job.seed, spin_seed=job.spin_seed)
    with swmi.Engine(c) as p:
        p.sweep(job.warmup)
        p.reset_measurements()
        p.run(job.sweeps, 10)
        return p.moments(0), p.energy_magnetization(0)


def _serve(swmi, jobs, delays=None, **kw):
    with swmi.Server(**kw) as srv:
        ids = {}
        for k, j in enumerate(jobs):
            if delays is not None:
                time.sleep(delays[k])
            ids[srv.submit(j)] = k
        assert srv.wait(120000) == 0
        res = srv.poll(len(jobs) + 8)
        st = srv.stats()
    assert len(res) == len(jobs)
    out = {ids[r["job_id"]]: r for r in res}
    for k, r in out.items():
        assert r["status"] == 0 and r["error_flags"] == 0, r
        assert r["sweeps_done"] == jobs[k].warmup + _measured(jobs[k].sweeps)
        assert r["t_submit"] <= r["t_admit"] <= r["t_finish"]
    return out, st


def test_server_job_alone_equals_in_mix_equals_plain(swmi):
    target = swmi.Job(128, 128, TC, seed=9, spin_seed=4, warmup=50, sweeps=300, tag=1)
    alone, _ = _serve(swmi, [target])
    mix_jobs = [swmi.Job(64, 64, 2.0, seed=1, warmup=0, sweeps=700),
                swmi.Job(1024, 1024, 2.5, seed=2, warmup=20, sweeps=90),
                target,
                swmi.Job(32, 32, 3.1, seed=3, warmup=10, sweeps=1234),
                swmi.Job(256, 512, 1.7, seed=4, warmup=0, sweeps=333),
                swmi.Job(128, 128, TC, seed=9, spin_seed=4, warmup=50, sweeps=300, tag=2),
                swmi.Job(2048, 2048, TC, seed=5, warmup=0, sweeps=40)]
    print("server mix", flush=True)
    mixed, st = _serve(swmi, mix_jobs, delays=[0, 0.002, 0.01, 0, 0.003, 0.02, 0])
    assert st["finished"] == len(mix_jobs) and st["active"] == 0 and st["queued"] == 0
    a = alone[0]
    for k in (2, 5):
        m = mixed[k]
        np.testing.assert_array_equal(m["moments"], a["moments"], err_msg=f"job {k}")
        assert (m["energy"], m["magnetization"]) == (a["energy"], a["magnetization"])
    mom, em = _plain_job_result(swmi, target)
    np.testing.assert_array_equal(a["moments"], mom)
    assert (a["energy"], a["magnetization"]) == em
    # the other jobs of the mix against the plain path too, including sweeps that are not a multiple of
    # n_sep (both run whole groups of 10 sweeps, as the reference's loop does)
    for k in (0, 1, 3, 4):
        mom, em = _plain_job_result(swmi, mix_jobs[k])
        np.testing.assert_array_equal(mixed[k]["moments"], mom, err_msg=f"job {k}")
        assert (mixed[k]["energy"], mixed[k]["magnetization"]) == em


def test_server_admits_under_a_small_memory_budget(swmi):
    """A budget for about two 512^2 jobs: the rest wait in the queue and are admitted as memory frees up."""
    jobs = [swmi.Job(512, 512, 2.0 + 0.1 * k, seed=k, warmup=0, sweeps=60 + 10 * k) for k in range(6)]
    budget = 2 * 512 * 512 * 12 + 200000
    out, st = _serve(swmi, jobs, memory_budget=budget, n_lanes=1)
    admits = sorted(r["t_admit"] for r in out.values())
    finishes = sorted(r["t_finish"] for r in out.values())
    # at most two resident at once: the third admission comes after the first completion
    assert admits[2] >= finishes[0]
    assert st["bytes_in_use"] <= budget
    for k in (0, 3):
        mom, em = _plain_job_result(swmi, jobs[k])
        np.testing.assert_array_equal(out[k]["moments"], mom)
        assert (out[k]["energy"], out[k]["magnetization"]) == em


def test_batch_autotune_returns_a_height_and_results_do_not_depend_on_it(swmi):
    cfgs = [_cfg(swmi, n0, n1, T, seed=30 + i, spin_seed=40 + i) for i, (n0, n1, T, _) in enumerate(MIXED[:6])]
    best, ms = swmi.autotune_batch(cfgs, sweeps=3)
    print(f"batch autotune: best {best}, ms {ms}", flush=True)
    assert best in (1, 2, 4, 8, 16, 32, 64) and all(m > 0 for m in ms)
    tuned = [_cfg(swmi, c.n0, c.n1, c.temperature, seed=c.seed, spin_seed=c.spin_seed, tile_rows=best) for c in cfgs]
    with swmi.Engine(tuned) as a, swmi.Engine(cfgs) as b:
        a.sweep(12)
        b.sweep(12)
        for i in range(len(cfgs)):
            np.testing.assert_array_equal(a.spins(i), b.spins(i), err_msg=f"lattice {i}")


def test_server_destroy_does_not_wait_for_long_jobs(swmi):
    """Destroying a server with a long resident job returns after the chunk in flight, not the job."""
    srv = swmi.Server(chunk_sweeps=10)
    srv.submit(swmi.Job(1024, 1024, TC, seed=1, sweeps=10 ** 6))   # about 30 s of device time
    time.sleep(0.2)
    t0 = time.perf_counter()
    srv.close()
    assert time.perf_counter() - t0 < 5.0


def test_two_engines_on_two_threads_equal_sequential(swmi):
    """Engines are independent: two host threads driving two engines at once give the results of running
    them one after the other."""
    import threading
    cfgs = [_cfg(swmi, 256, 256, 2.1, seed=50), _cfg(swmi, 512, 128, 2.4, seed=51)]
    out = [None, None]

    def work(i):
        with swmi.Engine(cfgs[i]) as e:
            e.run(300, 10)
            out[i] = (e.spins(), e.moments())

    threads = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for i, c in enumerate(cfgs):
        with swmi.Engine(c) as e:
            e.run(300, 10)
            np.testing.assert_array_equal(e.spins(), out[i][0])
            np.testing.assert_array_equal(e.moments(), out[i][1])


def test_server_many_small_jobs_from_two_threads(swmi):
    """200 small jobs submitted from two threads while earlier ones run: all finish with the plain-path
    results (checked for a sample) and no job is lost or duplicated."""
    import threading
    jobs = [swmi.Job(64 * (1 + k % 3), 64, 1.8 + 0.01 * k, seed=100 + k, warmup=10 * (k % 4), sweeps=50 + 10 * (k % 7), tag=k)
            for k in range(200)]
    with swmi.Server() as srv:
        def submit(part):
            for j in part:
                srv.submit(j)
        threads = [threading.Thread(target=submit, args=(jobs[i::2],)) for i in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert srv.wait(120000) == 0
        res = []
        while True:
            got = srv.poll(64)
            if not got:
                break
            res.extend(got)
        st = srv.stats()
    assert sorted(r["tag"] for r in res) == list(range(200))
    assert st["submitted"] == st["finished"] == 200 and st["active"] == 0
    by_tag = {r["tag"]: r for r in res}
    for k in (0, 37, 199):
        r = by_tag[k]
        assert r["status"] == 0 and r["sweeps_done"] == jobs[k].warmup + _measured(jobs[k].sweeps)
        mom, em = _plain_job_result(swmi, jobs[k])
        np.testing.assert_array_equal(r["moments"], mom)
        assert (r["energy"], r["magnetization"]) == em


@pytest.mark.parametrize("bad", [dict(n0=63), dict(sweeps=0), dict(warmup=15), dict(temperature=-1.0),
                                 dict(n0=65536, n1=65536)])
def test_server_rejects_invalid_jobs(swmi, bad):
    base = dict(n0=64, n1=64, temperature=2.0, seed=1, warmup=10, sweeps=100)
    base.update(bad)
    with swmi.Server() as srv:
        with pytest.raises(swmi.SwmiError):
            srv.submit(swmi.Job(**base))
        assert srv.stats()["submitted"] == 0


def test_bench_batch_equals_lone_runs_and_oracle(swmi, oracle):
    """The bench's batched workload itself (bench.batch_workload: 256 lattices of 32^2 .. 1024^2, default_rng(0)),
    in one engine at one tile height: after 6 sweeps every lattice equals its lone run (each at the height tuned
    for its extent), and every lattice up to 128^2 equals the oracle's restatement of the per-site sweep."""
    import bench
    cfgs = bench.batch_workload(swmi)
    assert len(cfgs) == 256
    n = 6
    for c in cfgs:
        c.tile_rows = 16
    print(f"bench batch: {len(cfgs)} lattices, {sum(c.n0 * c.n1 for c in cfgs)} sites", flush=True)
    with swmi.Engine(cfgs) as batch:
        batch.sweep(n)
        got = [batch.spins(i) for i in range(len(cfgs))]
        assert all(batch.error_flags(i) == 0 for i in range(len(cfgs)))
    small = 0
    for i, c in enumerate(cfgs):
        c.tile_rows = 0
        with swmi.Engine(c) as e:
            e.sweep(n)
            np.testing.assert_array_equal(got[i], e.spins(), err_msg=f"lattice {i} ({c.n0}^2)")
        if c.n0 <= 128:
            ref = oracle.PerSite(c.n0, c.n1, c.temperature, seed=c.seed, spin_seed=c.spin_seed)
            for _ in range(n):
                ref.sweep()
            np.testing.assert_array_equal(got[i], ref.spins, err_msg=f"lattice {i} ({c.n0}^2) vs oracle")
            small += 1
    assert small > 64
