"""The C ABI library: it loads, exports every function include/swmi.h declares, and validates its
arguments before it touches a device. CPU only (no compute call is made here)."""
import ctypes
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "swmi.h"
LIB = ROOT / "computational_physics_amd" / "lib" / "libswmi.so"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(swmi_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_header_declares_the_interface():
    names = declared_functions()
    assert len(names) >= 26
    for must in ("swmi_create", "swmi_sweep", "swmi_run", "swmi_observables", "swmi_download_labels", "swmi_destroy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert LIB.exists(), "build the library first: python -m computational_physics_amd.build"
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], check=True, capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(declared_functions()) <= exported


def test_library_is_built_for_gfx950_only():
    targets = set(re.findall(rb"amdgcn-amd-amdhsa-[a-z0-9:+\-]*gfx[0-9a-f]+", LIB.read_bytes()))
    assert targets == {b"amdgcn-amd-amdhsa--gfx950"}, targets


def _config(**kw):
    import computational_physics_amd as cpa
    c = cpa._Config()
    base = dict(n0=64, n1=64, temperature=2.2691853, seed=0, spin_seed=1, tile_rows=0, flip_rule=0, max_samples=0)
    base.update(kw)
    for k, v in base.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("bad", [
    dict(n0=63), dict(n0=0), dict(n1=3), dict(n0=96, n1=95), dict(n0=2, n1=0), dict(temperature=0.0),
    dict(temperature=-1.0), dict(temperature=float("inf")), dict(temperature=float("nan")), dict(flip_rule=7),
    dict(n0=65536, n1=32768), dict(n0=65534, n1=32768), dict(n0=1 << 29, n1=2), dict(n0=(1 << 29) - 2, n1=2),
])
def test_create_rejects_invalid_configs(bad):
    import computational_physics_amd as cpa
    lib = cpa.load_library()
    h = ctypes.c_void_p()
    c = _config(**bad)
    assert lib.swmi_create(ctypes.byref(c), 1, ctypes.byref(h)) == 1  # SWMI_ERR_INVALID
    assert lib.swmi_last_error()
    assert not h.value


@pytest.mark.parametrize("n0,n1", [(96, 96), (100, 100), (48, 34), (130, 18), (80, 66), (2, 2), (6, 8), (16, 6),
                                   (16, 2), (8, 4), (65534, 2)])
def test_every_even_extent_passes_validation(n0, n1):
    """The reference rejects odd extents only (ising/src/swendsen_wang.cpp:45-48): every even extent gets
    past argument validation and fails only for want of a device here."""
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a device is present")
    import computational_physics_amd as cpa
    lib = cpa.load_library()
    h = ctypes.c_void_p()
    c = _config(n0=n0, n1=n1)
    assert lib.swmi_create(ctypes.byref(c), 1, ctypes.byref(h)) == 3  # SWMI_ERR_NO_DEVICE, not SWMI_ERR_INVALID


def test_null_arguments_are_errors_not_crashes():
    import computational_physics_amd as cpa
    lib = cpa.load_library()
    assert lib.swmi_create(None, 1, None) == 1
    assert lib.swmi_sweep(None, 1) == 1
    assert lib.swmi_destroy(None) in (0, 1)
    assert lib.swmi_version().decode()


def test_no_device_is_reported(monkeypatch):
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a device is present")
    import computational_physics_amd as cpa
    lib = cpa.load_library()
    h = ctypes.c_void_p()
    c = _config()
    assert lib.swmi_create(ctypes.byref(c), 1, ctypes.byref(h)) == 3  # SWMI_ERR_NO_DEVICE
    with pytest.raises(cpa.SwmiError):
        cpa.device_info()
    with pytest.raises(cpa.SwmiError):
        cpa.Engine([cpa.LatticeConfig(64, 64)])


def test_python_mirror_names():
    """The Python mirror exposes the reference's entry points under the same meaning."""
    import computational_physics_amd as cpa
    for name in ("sweep", "run", "energy_magnetization", "moments", "series", "spins", "set_spins", "labels"):
        assert hasattr(cpa.Engine, name)
    assert cpa.REFERENCE_TEMPERATURE == pytest.approx(2.2691853, abs=1e-6)


@pytest.mark.parametrize("n0,n1,why", [(63, 64, "even"), (2048, 4096, "at most"), (0, 64, "even")])
def test_noalloc_labelling_rejects_bad_arguments_before_the_device(n0, n1, why):
    """swmi_label_bonds_noalloc checks its arguments (extents, the 2^22-site limit of its static scratch
    array) before it looks at a device; the arrays are never dereferenced on an error."""
    import computational_physics_amd as cpa
    lib = cpa.load_library()
    fake = ctypes.c_void_p(0x1000)
    assert lib.swmi_label_bonds_noalloc(fake, n0, n1, fake, None) == 1  # SWMI_ERR_INVALID
    assert why in lib.swmi_last_error().decode()
    assert lib.swmi_label_bonds_noalloc(None, 64, 64, fake, None) == 1
    assert lib.swmi_label_bonds_noalloc(fake, 64, 64, None, None) == 1
