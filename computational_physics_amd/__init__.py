"""MI355X-native Swendsen-Wang engine for the 2D Ising model (Python binding of include/swmi.h).

The binding is a thin ctypes layer over ``lib/libswmi.so`` (hand-written gfx950 HIP kernels plus the
C ABI). There is no fallback: if the library is missing or no gfx950 device is present, every call
raises :class:`SwmiError`.

Reference interface mirrored here (flwende/computational_physics):
  * ``Engine`` + ``Engine.sweep``     <- ``Lattice<2>::Update<SwendsenWang_2D, LCG32>(T, target)``
                                         (ising/include/lattice.hpp:46-58, ising/src/swendsen_wang.cpp:94-112)
  * ``Engine.energy_magnetization``   <- ``Lattice<2>::GetEnergyAndMagnetization`` (ising/src/lattice.cpp:29-79)
  * ``Engine.run``                    <- the measurement loop of ising/src/ising_2d.cpp:126-162
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

__all__ = ["SwmiError", "LatticeConfig", "Engine", "Job", "Server", "library_path", "load_library", "device_info",
           "copy_calibration", "autotune", "autotune_batch", "label_bonds_noalloc", "REFERENCE_TEMPERATURE"]

FLIP_PARITY = 0   # flip a cluster iff its canonical root index is odd (the reference rule)
FLIP_RANDOM = 1   # flip a cluster on one random bit of its root (MurmurHash3 of root, sweep, seed)

# ising/src/ising_2d.cpp:33, as the float the reference stores it in.
REFERENCE_TEMPERATURE = float(np.float32(2.2691853142130221))

_PKG = Path(__file__).resolve().parent
_LIB_PATH = _PKG / "lib" / "libswmi.so"
_lib = None


class SwmiError(RuntimeError):
    """Error reported by the engine (status code and message of swmi_last_error)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[swmi {code}] {message}")
        self.code = code


class _Config(ctypes.Structure):
    _fields_ = [("n0", ctypes.c_uint32), ("n1", ctypes.c_uint32), ("temperature", ctypes.c_float),
                ("seed", ctypes.c_uint32), ("spin_seed", ctypes.c_uint32), ("tile_rows", ctypes.c_uint32),
                ("flip_rule", ctypes.c_uint32), ("max_samples", ctypes.c_uint32)]


class _DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 32), ("compute_units", ctypes.c_uint32),
                ("wavefront_size", ctypes.c_uint32), ("total_memory", ctypes.c_uint64), ("l2_bytes", ctypes.c_uint32),
                ("clock_khz", ctypes.c_uint32)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("memory_budget", ctypes.c_uint64), ("max_jobs", ctypes.c_uint32), ("chunk_sweeps", ctypes.c_uint32),
                ("n_sep", ctypes.c_uint32), ("n_lanes", ctypes.c_uint32)]


class _Job(ctypes.Structure):
    _fields_ = [("n0", ctypes.c_uint32), ("n1", ctypes.c_uint32), ("temperature", ctypes.c_float),
                ("seed", ctypes.c_uint32), ("spin_seed", ctypes.c_uint32), ("warmup", ctypes.c_uint32),
                ("sweeps", ctypes.c_uint32), ("flip_rule", ctypes.c_uint32), ("tag", ctypes.c_uint64)]


class _JobResult(ctypes.Structure):
    _fields_ = [("job_id", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("n0", ctypes.c_uint32), ("n1", ctypes.c_uint32),
                ("temperature", ctypes.c_float), ("sweeps_done", ctypes.c_uint32), ("error_flags", ctypes.c_uint32),
                ("status", ctypes.c_uint32), ("moments", ctypes.c_double * 6), ("energy", ctypes.c_int64),
                ("magnetization", ctypes.c_int64), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_finish", ctypes.c_double)]


class _ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("admitted", ctypes.c_uint64), ("finished", ctypes.c_uint64),
                ("queued", ctypes.c_uint64), ("active", ctypes.c_uint64), ("bytes_in_use", ctypes.c_uint64),
                ("bytes_pooled", ctypes.c_uint64), ("spin_updates", ctypes.c_uint64), ("clock", ctypes.c_double)]


def library_path() -> Path:
    return _LIB_PATH


def load_library() -> ctypes.CDLL:
    """Load libswmi.so (built by computational_physics_amd.build). Raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise SwmiError(-1, f"{_LIB_PATH} is missing: build it with `python -m computational_physics_amd.build`")
    lib = ctypes.CDLL(str(_LIB_PATH))
    P = ctypes.POINTER
    u32, u64, i64, f32, f64, vp = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64, ctypes.c_float, ctypes.c_double, ctypes.c_void_p
    sigs = {
        "swmi_last_error": (ctypes.c_char_p, []),
        "swmi_version": (ctypes.c_char_p, []),
        "swmi_get_device_info": (ctypes.c_int, [P(_DeviceInfo)]),
        "swmi_create": (ctypes.c_int, [P(_Config), u32, P(vp)]),
        "swmi_destroy": (ctypes.c_int, [vp]),
        "swmi_num_lattices": (ctypes.c_int, [vp, P(u32)]),
        "swmi_synchronize": (ctypes.c_int, [vp]),
        "swmi_upload_spins": (ctypes.c_int, [vp, u32, vp]),
        "swmi_download_spins": (ctypes.c_int, [vp, u32, vp]),
        "swmi_set_temperature": (ctypes.c_int, [vp, u32, f32]),
        "swmi_get_threshold": (ctypes.c_int, [vp, u32, P(f32), P(u64)]),
        "swmi_sweep": (ctypes.c_int, [vp, u32]),
        "swmi_sample": (ctypes.c_int, [vp]),
        "swmi_run": (ctypes.c_int, [vp, u32, u32]),
        "swmi_observables": (ctypes.c_int, [vp, u32, P(i64), P(i64)]),
        "swmi_get_moments": (ctypes.c_int, [vp, u32, P(f64)]),
        "swmi_get_series": (ctypes.c_int, [vp, u32, vp, vp, u32, P(u32)]),
        "swmi_reset_measurements": (ctypes.c_int, [vp]),
        "swmi_record_bonds": (ctypes.c_int, [vp, ctypes.c_int]),
        "swmi_download_bonds": (ctypes.c_int, [vp, u32, vp]),
        "swmi_label_bonds": (ctypes.c_int, [vp, u32, vp]),
        "swmi_download_labels": (ctypes.c_int, [vp, u32, vp]),
        "swmi_label_bonds_noalloc": (ctypes.c_int, [vp, u32, u32, vp, vp]),
        "swmi_rng_export": (ctypes.c_int, [vp, u32, vp, vp, vp, P(u32)]),
        "swmi_get_error_flags": (ctypes.c_int, [vp, u32, P(u32)]),
        "swmi_time_kernels": (ctypes.c_int, [vp, u32, P(f32)]),
        "swmi_copy_calibration": (ctypes.c_int, [ctypes.c_size_t, u32, P(f32)]),
        "swmi_set_graphs": (ctypes.c_int, [vp, ctypes.c_int]),
        "swmi_checkpoint_size": (ctypes.c_int, [vp, u32, P(u64)]),
        "swmi_checkpoint": (ctypes.c_int, [vp, u32, vp, u64]),
        "swmi_restore": (ctypes.c_int, [vp, u32, vp, u64]),
        "swmi_error_bars": (ctypes.c_int, [vp, u32, u32, P(f64)]),
        "swmi_get_tile_rows": (ctypes.c_int, [vp, u32, P(u32)]),
        "swmi_get_layout": (ctypes.c_int, [vp, u32, P(u32), P(u32)]),
        "swmi_autotune": (ctypes.c_int, [P(_Config), u32, P(u32), P(f32)]),
        "swmi_autotune_batch": (ctypes.c_int, [P(_Config), u32, u32, P(u32), P(f32)]),
        "swmi_set_schedule": (ctypes.c_int, [vp, u32, u32, u32, u32]),
        "swmi_run_scheduled": (ctypes.c_int, [vp, u32, u32]),
        "swmi_get_progress": (ctypes.c_int, [vp, u32, P(u32), P(u32)]),
        "swmi_server_create": (ctypes.c_int, [P(_ServerConfig), P(vp)]),
        "swmi_server_destroy": (ctypes.c_int, [vp]),
        "swmi_server_submit": (ctypes.c_int, [vp, P(_Job), P(u64)]),
        "swmi_server_poll": (ctypes.c_int, [vp, P(_JobResult), u32, P(u32)]),
        "swmi_server_wait": (ctypes.c_int, [vp, u32, P(u32)]),
        "swmi_server_get_stats": (ctypes.c_int, [vp, P(_ServerStats)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != 0:
        msg = load_library().swmi_last_error().decode(errors="replace")
        raise SwmiError(status, msg)


def device_info() -> dict:
    lib = load_library()
    info = _DeviceInfo()
    _check(lib.swmi_get_device_info(ctypes.byref(info)))
    return {"name": info.name.decode(), "arch": info.arch.decode(), "compute_units": info.compute_units,
            "wavefront_size": info.wavefront_size, "total_memory": info.total_memory, "l2_bytes": info.l2_bytes,
            "clock_khz": info.clock_khz}


def label_bonds_noalloc(bonds_ptr: int, n0: int, n1: int, labels_ptr: int, stream: int = 0) -> None:
    """Allocation-free labelling (swmi_label_bonds_noalloc): device bond bytes at bonds_ptr -> canonical uint32
    labels at labels_ptr, enqueued on the HIP stream `stream` (0: default stream); the caller synchronises."""
    _check(load_library().swmi_label_bonds_noalloc(ctypes.c_void_p(bonds_ptr), n0, n1, ctypes.c_void_p(labels_ptr),
                                                   ctypes.c_void_p(stream)))


def copy_calibration(nbytes: int, repeats: int = 10) -> float:
    """Milliseconds per device-to-device copy of nbytes (calibrates HBM byte counters)."""
    ms = ctypes.c_float()
    _check(load_library().swmi_copy_calibration(nbytes, repeats, ctypes.byref(ms)))
    return ms.value


def autotune(config: "LatticeConfig", sweeps: int = 20) -> tuple[int, list[float]]:
    """Time plain sweeps for tile heights 1..64 (wave rows); the fastest is used from now on in this process
    for lattices of this extent created with tile_rows=0. Returns (best, ms per sweep per height; -1 = skipped)."""
    best = ctypes.c_uint32()
    ms = (ctypes.c_float * 7)()
    _check(load_library().swmi_autotune(ctypes.byref(config._c()), sweeps, ctypes.byref(best), ms))
    return best.value, list(ms)


def autotune_batch(configs, sweeps: int = 20) -> tuple[int, list[float]]:
    """Time plain sweeps of the whole batch with one tile height for all lattices (1..64); returns the fastest
    height and the ms per sweep of each."""
    arr = (_Config * len(configs))(*[c._c() for c in configs])
    best = ctypes.c_uint32()
    ms = (ctypes.c_float * 7)()
    _check(load_library().swmi_autotune_batch(arr, len(configs), sweeps, ctypes.byref(best), ms))
    return best.value, list(ms)


@dataclass
class LatticeConfig:
    """One lattice of an engine (fields of swmi_lattice_config)."""
    n0: int
    n1: int
    temperature: float = REFERENCE_TEMPERATURE
    seed: int = 0
    spin_seed: int = 1
    tile_rows: int = 0
    flip_rule: int = 0
    max_samples: int = 0

    def _c(self) -> _Config:
        return _Config(self.n0, self.n1, self.temperature, self.seed & 0xFFFFFFFF, self.spin_seed & 0xFFFFFFFF,
                       self.tile_rows, self.flip_rule, self.max_samples)


class Engine:
    """A set of lattices resident on the GPU, advanced together by one launch sequence per sweep."""

    def __init__(self, configs):
        if isinstance(configs, LatticeConfig):
            configs = [configs]
        self.configs = list(configs)
        self._lib = load_library()
        arr = (_Config * len(self.configs))(*[c._c() for c in self.configs])
        handle = ctypes.c_void_p()
        _check(self._lib.swmi_create(arr, len(self.configs), ctypes.byref(handle)))
        self._h = handle

    # -- lifecycle --
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.swmi_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self) -> None:
        _check(self._lib.swmi_synchronize(self._h))

    def _n(self, i: int) -> int:
        c = self.configs[i]
        return c.n0 * c.n1

    # -- state --
    def spins(self, i: int = 0) -> np.ndarray:
        out = np.empty(self._n(i), dtype=np.uint8)
        _check(self._lib.swmi_download_spins(self._h, i, out.ctypes.data))
        return out

    def set_spins(self, spins: np.ndarray, i: int = 0) -> None:
        s = np.ascontiguousarray(spins, dtype=np.uint8).ravel()
        if s.size != self._n(i):
            raise ValueError("spin array has the wrong size")
        _check(self._lib.swmi_upload_spins(self._h, i, s.ctypes.data))

    def set_temperature(self, temperature: float, i: int = 0) -> None:
        _check(self._lib.swmi_set_temperature(self._h, i, temperature))
        self.configs[i].temperature = temperature

    def threshold(self, i: int = 0) -> tuple[float, int]:
        p = ctypes.c_float()
        t = ctypes.c_uint64()
        _check(self._lib.swmi_get_threshold(self._h, i, ctypes.byref(p), ctypes.byref(t)))
        return p.value, t.value

    # -- Monte Carlo --
    def sweep(self, n: int = 1) -> None:
        _check(self._lib.swmi_sweep(self._h, n))

    def set_graphs(self, enable: bool) -> None:
        _check(self._lib.swmi_set_graphs(self._h, 1 if enable else 0))

    def tile_rows(self, i: int = 0) -> int:
        t = ctypes.c_uint32()
        _check(self._lib.swmi_get_tile_rows(self._h, i, ctypes.byref(t)))
        return t.value

    def set_schedule(self, i: int, sweeps: int, warmup: int = 0, n_sep: int = 10) -> None:
        _check(self._lib.swmi_set_schedule(self._h, i, sweeps, warmup, n_sep))

    def run_scheduled(self, n_sweeps: int, n_sep: int = 10) -> None:
        _check(self._lib.swmi_run_scheduled(self._h, n_sweeps, n_sep))

    def progress(self, i: int = 0) -> tuple[int, int]:
        """(sweeps done since creation, samples taken)."""
        a, b = ctypes.c_uint32(), ctypes.c_uint32()
        _check(self._lib.swmi_get_progress(self._h, i, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def sample(self) -> None:
        _check(self._lib.swmi_sample(self._h))

    def run(self, n_sweeps: int, n_sep: int = 10) -> None:
        _check(self._lib.swmi_run(self._h, n_sweeps, n_sep))

    def energy_magnetization(self, i: int = 0) -> tuple[int, int]:
        """(E, M) = (sum s(s_right+s_down), sum s) of the current configuration."""
        e, m = ctypes.c_int64(), ctypes.c_int64()
        _check(self._lib.swmi_observables(self._h, i, ctypes.byref(e), ctypes.byref(m)))
        return e.value, m.value

    def moments(self, i: int = 0) -> np.ndarray:
        out = (ctypes.c_double * 6)()
        _check(self._lib.swmi_get_moments(self._h, i, out))
        return np.array(out[:], dtype=np.float64)

    def series(self, i: int = 0) -> tuple[np.ndarray, np.ndarray]:
        cap = self.configs[i].max_samples
        e = np.empty(cap, dtype=np.int64)
        m = np.empty(cap, dtype=np.int64)
        count = ctypes.c_uint32()
        _check(self._lib.swmi_get_series(self._h, i, e.ctypes.data, m.ctypes.data, cap, ctypes.byref(count)))
        k = min(count.value, cap)
        return e[:k].copy(), m[:k].copy()

    def reset_measurements(self) -> None:
        _check(self._lib.swmi_reset_measurements(self._h))

    # -- inspection --
    def record_bonds(self, enable: bool = True) -> None:
        _check(self._lib.swmi_record_bonds(self._h, 1 if enable else 0))

    def bonds(self, i: int = 0) -> np.ndarray:
        out = np.empty(self._n(i), dtype=np.uint8)
        _check(self._lib.swmi_download_bonds(self._h, i, out.ctypes.data))
        return out

    def label_bonds(self, bonds: np.ndarray, i: int = 0) -> np.ndarray:
        b = np.ascontiguousarray(bonds, dtype=np.uint8).ravel()
        if b.size != self._n(i):
            raise ValueError("bond array has the wrong size")
        _check(self._lib.swmi_label_bonds(self._h, i, b.ctypes.data))
        return self.labels(i)

    def labels(self, i: int = 0) -> np.ndarray:
        out = np.empty(self._n(i), dtype=np.uint32)
        _check(self._lib.swmi_download_labels(self._h, i, out.ctypes.data))
        return out

    def layout(self, i: int = 0) -> tuple[int, int]:
        """(row pitch, lcg32 groups): the lattice is stored with rows padded to `pitch` sites, in pitch*n1/64 groups."""
        pitch, groups = ctypes.c_uint32(), ctypes.c_uint32()
        _check(self._lib.swmi_get_layout(self._h, i, ctypes.byref(pitch), ctypes.byref(groups)))
        return pitch.value, groups.value

    def rng_state(self, i: int = 0) -> dict:
        """lcg32 state of every group of the padded layout: states (64 per group), selector planes, rotations."""
        _, g = self.layout(i)
        states = np.empty(64 * g, dtype=np.uint32)
        masks = np.empty(3 * g, dtype=np.uint64)
        rot = np.empty(g, dtype=np.uint32)
        it = ctypes.c_uint32()
        _check(self._lib.swmi_rng_export(self._h, i, states.ctypes.data, masks.ctypes.data, rot.ctypes.data, ctypes.byref(it)))
        return {"states": states, "masks": masks, "rot": rot, "iteration": it.value}

    # -- checkpoint / resume --
    def checkpoint(self, i: int = 0) -> bytes:
        """The lattice's complete chain state (spins, lcg32 states, counters, moments, series)."""
        size = ctypes.c_uint64()
        _check(self._lib.swmi_checkpoint_size(self._h, i, ctypes.byref(size)))
        buf = ctypes.create_string_buffer(size.value)
        _check(self._lib.swmi_checkpoint(self._h, i, buf, size.value))
        return buf.raw

    def restore(self, data: bytes, i: int = 0) -> None:
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        _check(self._lib.swmi_restore(self._h, i, buf, len(data)))

    def error_bars(self, i: int = 0, n_bins: int = 50) -> dict:
        """Binning / jackknife error bars from the on-device series (needs max_samples > 0)."""
        out = (ctypes.c_double * 8)()
        _check(self._lib.swmi_error_bars(self._h, i, n_bins, out))
        keys = ("e", "e_err", "abs_m", "abs_m_err", "c", "c_err", "tau_int_e", "samples")
        return dict(zip(keys, out[:]))

    def error_flags(self, i: int = 0) -> int:
        f = ctypes.c_uint32()
        _check(self._lib.swmi_get_error_flags(self._h, i, ctypes.byref(f)))
        return f.value

    def time_kernels(self, n_sweeps: int) -> np.ndarray:
        ms = (ctypes.c_float * 3)()
        _check(self._lib.swmi_time_kernels(self._h, n_sweeps, ms))
        return np.array(ms[:], dtype=np.float64)


@dataclass
class Job:
    """One served job (fields of swmi_job)."""
    n0: int
    n1: int
    temperature: float
    seed: int = 0
    spin_seed: int = 1
    warmup: int = 0
    sweeps: int = 1000
    tag: int = 0
    flip_rule: int = 0

    def _c(self) -> _Job:
        return _Job(self.n0, self.n1, self.temperature, self.seed & 0xFFFFFFFF, self.spin_seed & 0xFFFFFFFF, self.warmup,
                    self.sweeps, self.flip_rule, self.tag)


def _result_dict(r: _JobResult) -> dict:
    return {"job_id": r.job_id, "tag": r.tag, "n0": r.n0, "n1": r.n1, "temperature": r.temperature,
            "sweeps_done": r.sweeps_done, "error_flags": r.error_flags, "status": r.status,
            "moments": np.array(r.moments[:], dtype=np.float64), "energy": r.energy, "magnetization": r.magnetization,
            "t_submit": r.t_submit, "t_admit": r.t_admit, "t_finish": r.t_finish}


class Server:
    """Continuous-batching request loop (swmi_server_*): submit jobs at any time, poll finished results."""

    def __init__(self, memory_budget: int = 0, max_jobs: int = 0, chunk_sweeps: int = 0, n_sep: int = 0, n_lanes: int = 0):
        self._lib = load_library()
        cfg = _ServerConfig(memory_budget, max_jobs, chunk_sweeps, n_sep, n_lanes)
        h = ctypes.c_void_p()
        _check(self._lib.swmi_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.swmi_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, job: Job) -> int:
        jid = ctypes.c_uint64()
        _check(self._lib.swmi_server_submit(self._h, ctypes.byref(job._c()), ctypes.byref(jid)))
        return jid.value

    def poll(self, max_results: int = 256) -> list[dict]:
        buf = (_JobResult * max_results)()
        n = ctypes.c_uint32()
        _check(self._lib.swmi_server_poll(self._h, buf, max_results, ctypes.byref(n)))
        return [_result_dict(buf[i]) for i in range(n.value)]

    def wait(self, timeout_ms: int = 60000) -> int:
        """Block until all submitted jobs are finished (or the timeout); returns the number unfinished."""
        left = ctypes.c_uint32()
        _check(self._lib.swmi_server_wait(self._h, timeout_ms, ctypes.byref(left)))
        return left.value

    def stats(self) -> dict:
        st = _ServerStats()
        _check(self._lib.swmi_server_get_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}
