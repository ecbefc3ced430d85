"""Mixed-precision path (split 16-bit operands on the matrix cores) against the
FP64 path on the GPU.  Error definitions: normwise relative error per
coefficient set, ||c_mixed - c_fp64|| / ||c_fp64||, and the largest elementwise
relative error over coefficients with |c| >= 1e-3 max |c| of their set.
Bars (north star): <= 1e-5 on expansion coefficients, <= 1e-6 on cross
sections (the cross sections never touch the angular path and agree exactly)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from conftest import load_golden
from lmie_amd import lmie, workloads as W

pytestmark = pytest.mark.gpu

SETS = ("alpha1", "alpha2", "alpha3", "alpha4", "beta1", "beta2")


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("bf16", [0, 1])
def test_mfma16_lane_layout(product, bf16):
    say("case: mfma 16x16x32 %s lane layout" % ("bf16" if bf16 else "f16"))
    rng = np.random.Generator(np.random.PCG64(5))
    A = rng.integers(-8, 9, (16, 32)).astype(np.float32)
    B = rng.integers(-8, 9, (32, 16)).astype(np.float32)
    D = np.zeros((16, 16), np.float32)
    f = product.lmie_gpu_selftest_mfma16
    f.argtypes = [C.c_int] + [C.POINTER(C.c_float)] * 3 + [C.c_int]
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    assert f(bf16, fp(np.ascontiguousarray(A)), fp(np.ascontiguousarray(B)), fp(D), 0) == 0
    E = A.astype(np.float64) @ B.astype(np.float64)
    if not np.array_equal(D, E):
        # report which row map the hardware uses
        alt = np.zeros_like(D)
        for l in range(64):
            for v in range(4):
                alt[(l >> 4) + 4 * v, l & 15] = D[4 * (l >> 4) + v, l & 15]
        say("  rows as (l>>4)+4v match: %s" % np.array_equal(alt, E))
    assert np.array_equal(D, E)


def test_mfma32_lane_layout(product):
    say("case: mfma 32x32x16 f16 lane layout")
    rng = np.random.Generator(np.random.PCG64(6))
    A = rng.integers(-8, 9, (32, 16)).astype(np.float32)
    B = rng.integers(-8, 9, (16, 32)).astype(np.float32)
    D = np.zeros((32, 32), np.float32)
    f = product.lmie_gpu_selftest_mfma32
    f.argtypes = [C.POINTER(C.c_float)] * 3 + [C.c_int]
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    assert f(fp(np.ascontiguousarray(A)), fp(np.ascontiguousarray(B)), fp(D), 0) == 0
    E = A.astype(np.float64) @ B.astype(np.float64)
    if not np.array_equal(D, E):
        # report which row of the result each (lane group, value) holds
        for r0 in range(32):
            hits = [r for r in range(32) if np.array_equal(D[r0], E[r])]
            say("  stored row %d holds rows %s" % (r0, hits))
    assert np.array_equal(D, E)


def errors(mix, ref):
    out = {}
    for k in range(6):
        c, r = mix["gc"][k], ref["gc"][k]
        n = np.linalg.norm(r)
        big = np.abs(r) >= 1e-3 * np.abs(r).max() if n > 0 else np.zeros_like(r, bool)
        out[SETS[k]] = {"normwise": float(np.linalg.norm(c - r) / n) if n > 0 else 0.0,
                        "elem_max": float(np.max(np.abs(c[big] - r[big]) / np.abs(r[big]))) if big.any() else 0.0}
    for q in ("cext", "csca", "cbak", "g"):
        out[q] = float(abs(mix[q] - ref[q]) / abs(ref[q]))
    return out


CASES = {
    "cfg0": lambda lib, prec: lmie.solve(lib, W.CFG0, None, prec),
    "cfg1": lambda lib, prec: lmie.solve(lib, W.CFG1, None, prec),
    "cfg2": lambda lib, prec: lmie.solve(lib, W.CFG2, None, prec),
    "cfg4": lambda lib, prec: lmie.solve_mix(lib, [(W.cfg4_mode_params(k), None) for k in range(2)],
                                             [m["weight"] for m in W.CFG4_MODES], prec),
}


@pytest.mark.parametrize("case", list(CASES))
def test_mixed_precision_error(product, case):
    say("case: mixed vs fp64", case)
    ref = CASES[case](product, "fp64")
    table = {}
    for prec in ("mixed", "mixed-bf16"):
        table[prec] = errors(CASES[case](product, prec), ref)
        say("  %-10s " % prec + " ".join("%s %.1e/%.1e" % (k, v["normwise"], v["elem_max"])
                                         for k, v in table[prec].items() if isinstance(v, dict)))
    os.makedirs("gpurun_out", exist_ok=True)
    with open("gpurun_out/mixed_error_%s.json" % case, "w") as fh:
        json.dump(table, fh, indent=1)
    e = table["mixed"]
    for q in ("cext", "csca", "cbak"):
        assert e[q] <= 1e-6, q
    for k in SETS:
        assert e[k]["normwise"] <= 1e-5, (k, e[k])
