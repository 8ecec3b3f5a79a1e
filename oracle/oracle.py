"""ctypes wrapper of liboracle.so -- the CPU restatement of the reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, as the checker. The product (computational_physics_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_DIR = HERE / "_ref"

_lib = None


class Lcg32(ctypes.Structure):
    _fields_ = [("w", ctypes.c_uint32), ("iteration", ctypes.c_uint32), ("state", ctypes.c_uint32 * 64),
                ("a", ctypes.c_uint32 * 64), ("c", ctypes.c_uint32 * 64)]


class Lcg32Cpu(ctypes.Structure):
    _fields_ = [("s", Lcg32), ("current", ctypes.c_uint32)]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = ctypes.CDLL(str(LIB_PATH))
    P, u8p, u32, u64, vp = ctypes.POINTER, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p
    L.or_srand48.argtypes = [P(u64), u32]
    L.or_drand48.argtypes = [P(u64)]
    L.or_drand48.restype = ctypes.c_double
    L.or_lcg32_init.argtypes = [P(Lcg32), u32, u32]
    L.or_lcg32_update.argtypes = [P(Lcg32)]
    L.or_lcg32_cpu_init.argtypes = [P(Lcg32Cpu), u32, u32]
    L.or_lcg32_cpu_next_integers.argtypes = [P(Lcg32Cpu), vp, ctypes.c_size_t]
    L.or_hot_start.argtypes = [vp, ctypes.c_size_t, u32]
    L.or_p_add.argtypes = [ctypes.c_float]
    L.or_p_add.restype = ctypes.c_float
    L.or_threshold.argtypes = [ctypes.c_float]
    L.or_threshold.restype = u64
    L.or_threshold_mismatches.argtypes = [ctypes.c_float, u64]
    L.or_threshold_mismatches.restype = u64
    L.or_label.argtypes = [u32, u32, u32, u32, vp, vp]
    L.or_flip.argtypes = [ctypes.c_size_t, vp, vp]
    L.or_fmix32.argtypes = [u32]
    L.or_fmix32.restype = u32
    L.or_murmur3_32.argtypes = [ctypes.c_char_p, ctypes.c_size_t, u32]
    L.or_murmur3_32.restype = u32
    L.or_flip_key.argtypes = [u32, u32]
    L.or_flip_key.restype = u32
    L.or_flip_random.argtypes = [ctypes.c_size_t, vp, vp, u32]
    L.or_observables.argtypes = [u32, u32, vp, P(ctypes.c_int64), P(ctypes.c_int64)]
    L.or_refcpu_measure.argtypes = [vp, vp, ctypes.c_float, u32, u32, u32, P(ctypes.c_double)]
    L.or_refcpu_create.argtypes = [u32] * 6
    L.or_refcpu_create.restype = vp
    L.or_refcpu_destroy.argtypes = [vp]
    L.or_refcpu_sweep.argtypes = [vp, vp, ctypes.c_float, vp, vp]
    L.or_ps_pitch.argtypes = [u32, u32]
    L.or_ps_pitch.restype = u32
    L.or_group_seed.argtypes = [u32, u32]
    L.or_group_seed.restype = u32
    L.or_ps_init.argtypes = [u32, u32, vp, vp, vp]
    L.or_ps_rng_update.argtypes = [u32, vp, vp, vp, u32]
    L.or_ps_bonds.argtypes = [u32, u32, vp, vp, vp, u64, vp]
    L.or_ps_sweep.argtypes = [u32, u32, u64, vp, vp, vp, vp, P(u32), vp, vp]
    L.or_ps_sweep_rule.argtypes = [u32, u32, u64, vp, vp, vp, vp, P(u32), vp, vp, u32, u32]
    _lib = L
    return L


# ---- convenience wrappers ----
def p_add(temperature: float) -> float:
    return float(lib().or_p_add(temperature))


def threshold(temperature: float) -> int:
    return int(lib().or_threshold(lib().or_p_add(temperature)))


def hot_start(n: int, seed: int = 1) -> np.ndarray:
    s = np.empty(n, dtype=np.uint8)
    lib().or_hot_start(s.ctypes.data, n, seed)
    return s


def label(n0: int, n1: int, bonds: np.ndarray, tile=(16, 8)) -> np.ndarray:
    b = np.ascontiguousarray(bonds, dtype=np.uint8)
    out = np.empty(n0 * n1, dtype=np.uint32)
    lib().or_label(n0, n1, tile[0], tile[1], b.ctypes.data, out.ctypes.data)
    return out


def pitch(n0: int, n1: int) -> int:
    """Row pitch of the engine's padded layout."""
    return int(lib().or_ps_pitch(n0, n1))


def observables(n0: int, n1: int, spins: np.ndarray) -> tuple[int, int]:
    s = np.ascontiguousarray(spins, dtype=np.uint8)
    e, m = ctypes.c_int64(), ctypes.c_int64()
    lib().or_observables(n0, n1, s.ctypes.data, ctypes.byref(e), ctypes.byref(m))
    return e.value, m.value


def lcg32_states(w: int, seed: int, updates: int) -> list[list[int]]:
    s = Lcg32()
    lib().or_lcg32_init(ctypes.byref(s), w, seed)
    out = [list(s.state[:w])]
    for _ in range(updates):
        lib().or_lcg32_update(ctypes.byref(s))
        out.append(list(s.state[:w]))
    return out


def lcg32_buffered(w: int, seed: int, sizes) -> list[list[int]]:
    g = Lcg32Cpu()
    lib().or_lcg32_cpu_init(ctypes.byref(g), w, seed)
    out = []
    for n in sizes:
        buf = np.empty(n, dtype=np.uint32)
        lib().or_lcg32_cpu_next_integers(ctypes.byref(g), buf.ctypes.data, n)
        out.append([int(v) for v in buf])
    return out


class RefCpu:
    """The reference CPU path (thread-chunked lcg32 streams), restated."""

    def __init__(self, n0, n1, w=16, num_threads=1, tile=None):
        tile = tile or (w, 8)
        self.n0, self.n1 = n0, n1
        self._h = lib().or_refcpu_create(n0, n1, w, num_threads, tile[0], tile[1])
        if not self._h:
            raise ValueError("invalid reference CPU configuration")

    def sweep(self, spins: np.ndarray, temperature: float) -> None:
        lib().or_refcpu_sweep(self._h, spins.ctypes.data, temperature, None, None)

    def measure(self, spins: np.ndarray, temperature: float, warmup: int, sweeps: int, n_sep: int = 10):
        """The reference driver's measurement loop; returns (<e>, C/N, <|m|>)."""
        out = (ctypes.c_double * 3)()
        lib().or_refcpu_measure(self._h, spins.ctypes.data, temperature, warmup, sweeps, n_sep, out)
        return tuple(out[:])

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_refcpu_destroy(self._h)
            self._h = None


class PerSite:
    """The MI355X engine's per-site lcg32 layout and sweep, restated on the CPU."""

    def __init__(self, n0, n1, temperature, seed=0, spin_seed=1, flip_rule=0):
        self.n0, self.n1 = n0, n1
        self.seed, self.flip_rule = seed, flip_rule
        self.n = n0 * n1
        self.pitch = pitch(n0, n1)
        self.groups = self.pitch * n1 // 64
        self.states = np.empty(self.groups * 64, dtype=np.uint32)
        self.masks = np.empty(3 * self.groups, dtype=np.uint64)
        self.rot = np.empty(self.groups, dtype=np.uint32)
        self.iteration = ctypes.c_uint32(0)
        lib().or_ps_init(self.groups, seed, self.states.ctypes.data, self.masks.ctypes.data, self.rot.ctypes.data)
        self.spins = hot_start(self.n, spin_seed)
        self.set_temperature(temperature)

    def set_temperature(self, temperature):
        self.temperature = temperature
        self.threshold = threshold(temperature)

    def sweep(self, record=False):
        bonds = np.empty(self.n, dtype=np.uint8) if record else None
        labels = np.empty(self.n, dtype=np.uint32) if record else None
        lib().or_ps_sweep_rule(self.n0, self.n1, self.threshold, self.spins.ctypes.data, self.states.ctypes.data,
                               self.masks.ctypes.data, self.rot.ctypes.data, ctypes.byref(self.iteration),
                               None if bonds is None else bonds.ctypes.data, None if labels is None else labels.ctypes.data,
                               self.flip_rule, self.seed & 0xFFFFFFFF)
        return bonds, labels


def murmur3_32(data: bytes, seed: int) -> int:
    return int(lib().or_murmur3_32(data, len(data), seed & 0xFFFFFFFF))


def fmix32(h: int) -> int:
    return int(lib().or_fmix32(h & 0xFFFFFFFF))


def flip_key(seed: int, iteration: int) -> int:
    return int(lib().or_flip_key(seed & 0xFFFFFFFF, iteration & 0xFFFFFFFF))
