#!/usr/bin/env python3
"""Exact internal energy and specific heat per spin of the finite periodic 2D Ising model.

Ferdinand & Fisher, Phys. Rev. 185, 832 (1969) (Kaufman's solution): on an M x N torus with coupling
K = J/T (J = 1, k_B = 1)

    Z = 1/2 (2 sinh 2K)^(MN/2) (Z1 + Z2 + Z3 + Z4),
    Z1 = prod_{r=0}^{N-1} 2 cosh(M/2 g_{2r+1}),   Z2 = prod 2 sinh(M/2 g_{2r+1}),
    Z3 = prod_{r=0}^{N-1} 2 cosh(M/2 g_{2r}),     Z4 = prod 2 sinh(M/2 g_{2r}),
    cosh g_l = cosh 2K coth 2K - cos(l pi / N),   g_0 = 2K + ln tanh K,

and U/N = -(1/MN) d ln Z / dK, C/N = (K^2/MN) d^2 ln Z / dK^2. This script evaluates ln Z in mpmath at
60 digits and differentiates it numerically (mpmath.diff), an independent route from the closed-form
derivatives of ising/verification/FF_internalEnergy_specificHeat.c; tests compare the two where the
reference's long-double tool is valid.

K is 1/float32(T), the temperature as the reference simulation holds it (ising/src/ising_2d.cpp:42).

    python tests/golden/make_exact.py        # writes tests/golden/exact_values.json
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import mpmath as mp
import numpy as np

HERE = Path(__file__).resolve().parent
mp.mp.dps = 60


def log_z(K, M: int, N: int):
    K = mp.mpf(K)
    c = mp.cosh(2 * K) / mp.tanh(2 * K)

    def gamma(l: int):
        if l == 0:
            return 2 * K + mp.log(mp.tanh(K))
        return mp.acosh(c - mp.cos(mp.pi * l / N))

    half = mp.mpf(M) / 2
    lz = [mp.mpf(0)] * 4
    sign = [1, 1, 1, 1]
    for r in range(N):
        g_odd = half * gamma(2 * r + 1)
        g_even = half * gamma(2 * r)
        lz[0] += mp.log(2 * mp.cosh(g_odd))
        lz[1] += mp.log(2 * mp.sinh(g_odd))
        lz[2] += mp.log(2 * mp.cosh(g_even))
        s = 2 * mp.sinh(g_even)   # negative for r = 0 below the critical coupling
        if s < 0:
            sign[3] = -sign[3]
        lz[3] += mp.log(abs(s))
    top = max(lz)
    total = sum(sign[i] * mp.exp(lz[i] - top) for i in range(4))
    return mp.log(mp.mpf(1) / 2) + (mp.mpf(M) * N / 2) * mp.log(2 * mp.sinh(2 * K)) + top + mp.log(total)


def exact(K, M: int, N: int) -> tuple[float, float]:
    f = lambda k: log_z(k, M, N)
    d1 = mp.diff(f, mp.mpf(K), 1)
    d2 = mp.diff(f, mp.mpf(K), 2)
    u = -d1 / (M * N)
    cv = mp.mpf(K) ** 2 * d2 / (M * N)
    return float(u), float(cv)


def main() -> int:
    temps = {"2.0": 2.0, "2.2691853": 2.2691853, "2.6": 2.6}
    sizes = [16, 32, 64, 128, 1024, 4096, 8192]
    out = {"method": "mpmath 60 digits, numerical d/dK of ln Z (Ferdinand-Fisher)", "values": []}
    for L in sizes:
        for name, t in temps.items():
            K = mp.mpf(1) / mp.mpf(float(np.float32(t)))
            u, cv = exact(K, L, L)
            out["values"].append({"L": L, "T": name, "T_float": float(np.float32(t)), "K": float(K), "U_per_spin": u,
                                  "C_per_spin": cv})
            print(L, name, u, cv, flush=True)
    # the critical coupling itself, for the README check (README.md:98: -1.419076272084983 at 128^2)
    Kc = mp.log(1 + mp.sqrt(2)) / 2
    u, cv = exact(Kc, 128, 128)
    out["critical_128"] = {"K": float(Kc), "U_per_spin": u, "C_per_spin": cv}
    (HERE / "exact_values.json").write_text(json.dumps(out, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
