"""Shared fixtures.  Tests marked ``gpu`` need an MI355X; everything else runs on CPU."""
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")

from lmie_amd import abi  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and the built HIP library")


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    meta = json.loads(str(z["inputs_json"]))
    return z, meta["params"], (meta["seeds"] or None)


def relnorm(a, b):
    """Normwise relative difference ||a - b|| / ||b|| (absolute when b = 0)."""
    a = np.asarray(a, dtype=float).ravel()
    b = np.asarray(b, dtype=float).ravel()
    n = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / n if n > 0 else d


OUTPUT_KEYS = ("cext", "csca", "cbak", "g", "reff", "veff", "gavg", "vavg", "ravg", "rvw", "norm",
               "r1", "r2", "gc", "lc", "pf", "theta", "save1", "save2")


def compare_all(res, z, tol, skip=()):
    """Normwise comparison of every output present in fixture z; returns the
    worst (error, key) and asserts it is within tol."""
    worst = (0.0, None)
    for k in z.files:
        if k in ("inputs_json", "loss", "loss_grad", "max_coef") or k in skip:
            continue
        if k == "n_coef":
            assert int(res[k]) == int(z[k]), "n_coef %s vs %s" % (res[k], z[k])
            continue
        assert k in res, "missing output " + k
        e = relnorm(res[k], z[k])
        if e > worst[0]:
            worst = (e, k)
    assert worst[0] <= tol, "worst %s rel err %.3e > %.1e" % (worst[1], worst[0], tol)
    return worst


def _build_oracle():
    so = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    src = os.path.join(ROOT, "oracle", "lmie_oracle.c")
    if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return so


@pytest.fixture(scope="session")
def oracle():
    """The CPU restatement (the checker)."""
    lib = abi.bind(C.CDLL(_build_oracle()))
    lib.oracle_mie_ab.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.oracle_mie_ab.restype = C.c_int
    lib.oracle_pitau.argtypes = [C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.oracle_pitau.restype = None
    lib.oracle_gsf.argtypes = [C.c_int, C.c_double] + [C.POINTER(C.c_double)] * 4
    lib.oracle_gsf.restype = None
    lib.oracle_gauss_leg.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
    lib.oracle_gauss_leg.restype = None
    lib.oracle_order_n1.argtypes = [C.c_double]
    lib.oracle_order_n1.restype = C.c_int
    return lib


@pytest.fixture(scope="session")
def reflib():
    """The reference build in oracle/_ref (skips where it was not built)."""
    so = os.path.join(ROOT, "oracle", "_ref", "liblmie_ref.so")
    if not os.path.exists(so):
        pytest.skip("reference build oracle/_ref/liblmie_ref.so not present")
    return abi.bind(C.CDLL(so))


@pytest.fixture(scope="session")
def product():
    """The product library; GPU tests call it, CPU tests only inspect it."""
    import lmie_amd
    return lmie_amd.load()
