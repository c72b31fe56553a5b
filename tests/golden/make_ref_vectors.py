"""Writes tests/golden/ref_vectors.npz from the reference's own headers.

    python oracle/ref_build.py            # oracle/_ref/libref.so from the reference tree
    python tests/golden/make_ref_vectors.py

Inputs are seeded; outputs come from the reference's code compiled unmodified
(sobol.h, saka.h, interval.h through oracle/ref_driver.cpp).  The file holds
data only: inputs and the reference's outputs.

  sobol_*   shuffled_scrambled_sobol_2d (sobol.h:118-128) at the photon
            tracer's seed patterns (main_gpu.cu:84, :88, :143: index = thread,
            seeds (iteration, triangle, 789), (iteration, triangle, 178),
            (depth, triangle, iteration)) and at random 32-bit seeds
  saka_*    saka::reflection / saka::refraction_norm_free on one-tangent
            duals (saka.h:192-208), total internal reflection included
  iv_*      interval.h's +, -, *, /, sqrt, square, dot, cross, lengthSquared
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "ref_vectors.npz")
N_IV = {0: 256, 1: 256, 2: 256, 3: 256, 4: 256, 5: 256, 6: 256, 7: 256, 8: 256}


def sobol_inputs(rng: np.random.Generator) -> np.ndarray:
    rows = []
    for it in (0, 1, 7):
        for tri in (0, 1, 511, 131075):
            for t in range(0, 256, 5):
                rows.append((t, it, tri, 789))
                rows.append((t, it, tri, 178))
            for d in range(5):
                for t in range(0, 256, 17):
                    rows.append((t, d, tri, it))
    rnd = rng.integers(0, 1 << 32, size=(1024, 4), dtype=np.uint64)
    rows.extend(map(tuple, rnd.tolist()))
    return np.array(rows, dtype=np.uint32)


def saka_inputs(rng: np.random.Generator):
    n = 1024
    wi = rng.normal(size=(n, 6)).astype(np.float32)
    nn = rng.normal(size=(n, 6)).astype(np.float32)
    wi[:, :3] *= rng.uniform(0.2, 3.0, size=(n, 1)).astype(np.float32)
    eta = rng.uniform(0.55, 1.9, size=n).astype(np.float32)  # eta < 1 gives TIR cases
    return wi, nn, eta


def iv_inputs(rng: np.random.Generator, op: int, n: int):
    def ivs(m, lo=-4.0, hi=4.0, positive=False):
        a = rng.uniform(0.05 if positive else lo, hi, size=(m, 2)).astype(np.float32)
        a.sort(axis=1)
        point = rng.random(m) < 0.1
        a[point, 1] = a[point, 0]
        return a
    if op in (6, 7, 8):
        a = ivs(3 * n).reshape(n, 6)
        b = ivs(3 * n).reshape(n, 6)
    elif op == 3:
        a, b = ivs(n), ivs(n, 0.1, 5.0, positive=True)
        b[: n // 2] = -b[: n // 2][:, ::-1]  # divisors of both signs, never straddling zero
    elif op == 4:
        a, b = ivs(n, 0.0, 9.0, positive=True), np.zeros((n, 2), np.float32)
    else:
        a, b = ivs(n), ivs(n)
    return np.ascontiguousarray(a), np.ascontiguousarray(b)


def main() -> None:
    import ctypes as C

    from oracle import ref_build
    path = ref_build.build()
    if path is None:
        raise SystemExit("the reference tree is not here: nothing to generate")
    import oracle
    ref = oracle.ref_lib()
    ptr = oracle._ptr
    rng = np.random.default_rng(20261204)
    out: dict[str, np.ndarray] = {}

    s_in = sobol_inputs(rng)
    s_out = np.zeros((len(s_in), 2), np.float32)
    for i, (idx, p0, p1, p2) in enumerate(s_in.tolist()):
        ref.ref_sobol2d(idx, p0, p1, p2, ptr(s_out[i]))
    out["sobol_in"], out["sobol_out"] = s_in, s_out

    wi, nn, eta = saka_inputs(rng)
    for op, name in ((0, "refl"), (1, "refr")):
        res = np.zeros((len(wi), 6), np.float32)
        for i in range(len(wi)):
            ref.ref_saka_optics(op, ptr(wi[i]), ptr(nn[i]), C.c_float(eta[i]), ptr(res[i]))
        out[f"saka_{name}_out"] = res
    out["saka_wi"], out["saka_n"], out["saka_eta"] = wi, nn, eta

    for op, n in N_IV.items():
        a, b = iv_inputs(rng, op, n)
        width = 6 if op == 7 else 2
        res = np.zeros((n, 6), np.float32)
        a6 = np.zeros((n, 6), np.float32)
        b6 = np.zeros((n, 6), np.float32)
        a6[:, : a.shape[1]] = a
        b6[:, : b.shape[1]] = b
        for i in range(n):
            ref.ref_interval(op, ptr(a6[i]), ptr(b6[i]), ptr(res[i]))
        out[f"iv{op}_a"], out[f"iv{op}_b"], out[f"iv{op}_out"] = a6, b6, res[:, :width].copy()

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {len(s_in)} Sobol points, {len(wi)} dual pairs, {sum(N_IV.values())} interval ops")


if __name__ == "__main__":
    main()
