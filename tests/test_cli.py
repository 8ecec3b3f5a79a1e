"""The ising2d driver: flags and report lines of the reference driver (ising/src/ising_2d.cpp), error
behaviour on the CPU, and on the GPU the host-side measurement loop (the reference's) against the
on-device one."""
import re
import subprocess

import pytest

from conftest import ROOT

CLI = ROOT / "computational_physics_amd" / "bin" / "ising2d"


def run(*args, timeout=120):
    return subprocess.run([str(CLI), *args], capture_output=True, text=True, timeout=timeout)


def report(out: str) -> dict:
    keys = {"Update time per spin": "ns", "Internal energy per spin": "e", "Heat capacity per spin": "c",
            "Absolute Magnetization per spin": "m"}
    res = {}
    for k, v in keys.items():
        m = re.search(re.escape(k) + r": ([-0-9.eE+]+)", out)
        assert m, (k, out)
        res[v] = float(m.group(1))
    return res


def test_help_lists_reference_flags():
    assert CLI.exists(), "build first: python -m computational_physics_amd.build"
    out = run("--help").stdout
    for flag in ("--algorithm", "--extent=N0xN1", "--num_sweeps", "--rng=<name>", "--target=<name>", "--temperature"):
        assert flag in out


@pytest.mark.parametrize("args,msg", [
    (["--target=cpu"], "Unknown target device."),
    (["--rng=mt19937"], "Unknown RNG name."),
    (["--algorithm=metropolis"], "Unknown algorithm."),
    (["--flip_rule=coin"], "Unknown flip rule."),
])
def test_errors_like_reference(args, msg):
    r = run(*args)
    assert r.returncode != 0 and msg in r.stderr


def test_no_device_fails_loudly():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a device is present")
    r = run("--extent=64x64", "--num_sweeps=10")
    assert r.returncode != 0 and "gfx950" in r.stderr


@pytest.mark.gpu
def test_cli_host_and_device_measurement_agree(swmi):
    args = ["--extent=128x128", "--num_warmup=200", "--num_sweeps=2000", "--seed=5"]
    dev = run(*args, "--measure=device")
    host = run(*args, "--measure=host")
    assert dev.returncode == 0, dev.stderr
    assert host.returncode == 0, host.stderr
    # the reference's report lines (ising_2d.cpp:101,120-121,152-162): a GPU target prints "AMD GPU"
    lines = dev.stdout.splitlines()
    assert lines[:3] == ["Target name: AMD GPU", "WARMUP: 200 sweeps", "MEASUREMENT: 2000 sweeps"], lines
    assert [l.split(":")[0] for l in lines[3:]] == ["Update time per spin", "Internal energy per spin",
                                                    "Heat capacity per spin", "Absolute Magnetization per spin"]
    rd, rh = report(dev.stdout), report(host.stdout)
    # the same configurations are sampled in the same order; both sums are sequential FP64
    assert rd["e"] == rh["e"] and rd["m"] == rh["m"] and rd["c"] == rh["c"]
    assert -1.5 < rd["e"] < -1.3


@pytest.mark.gpu
def test_cli_matches_python_engine(swmi):
    """The C++ mirror and the Python binding drive the same engine: same seed, same moments."""
    import numpy as np
    r = run("--extent=64x64", "--num_warmup=50", "--num_sweeps=500", "--seed=9")
    assert r.returncode == 0, r.stderr
    rep = report(r.stdout)
    with swmi.Engine(swmi.LatticeConfig(64, 64, seed=9)) as e:
        e.sweep(50)
        e.reset_measurements()
        e.run(500, 10)
        mom = e.moments()
    assert mom[0] == 50
    # printed with the stream's default precision (6 significant digits, as %g)
    assert rep["e"] == float(f"{mom[1] / 50:.6g}")
    assert rep["m"] == float(f"{mom[3] / 50:.6g}")
    t = np.float32(2.2691853142130221)
    c = (mom[2] / 50 - (mom[1] / 50) ** 2) * 4096 / float(t * t)
    assert rep["c"] == pytest.approx(float(f"{c:.6g}"), rel=1e-5)


@pytest.mark.gpu
def test_cli_random_flip_rule_matches_python_engine(swmi):
    r = run("--extent=64x64", "--num_warmup=50", "--num_sweeps=300", "--seed=4", "--flip_rule=random")
    assert r.returncode == 0, r.stderr
    rep = report(r.stdout)
    with swmi.Engine(swmi.LatticeConfig(64, 64, seed=4, flip_rule=swmi.FLIP_RANDOM)) as e:
        e.sweep(50)
        e.reset_measurements()
        e.run(300, 10)
        mom = e.moments()
    assert rep["e"] == float(f"{mom[1] / 30:.6g}")
    assert rep["m"] == float(f"{mom[3] / 30:.6g}")
