"""ctypes mirror of the drop-in C ABI declared in include/lmie_interface.h.

The structures follow the reference layout field for field
(src/lmie_interface.h:20-101), so the same Python driver can call any library
exporting that ABI: the product library liblmie_amd.so, or a CPU build used as
a checker.  ``solve`` fills an ``lmie_in_data`` from keyword parameters, calls
``lmie_solution`` with ``alloc_out=1`` and copies every output into numpy
arrays before releasing the C-side blocks with ``lmie_out_free``.
"""
import ctypes as C

import numpy as np

DP = C.POINTER(C.c_double)
DPP = C.POINTER(DP)
DPPP = C.POINTER(DPP)

# misc/rtutil_size_dist.h:22-34
SIZE_DIST = {
    "none": 0, "modified_gamma": 1, "log_normal": 2, "power_law": 3,
    "gamma": 4, "modified_power_law": 5, "modified_bimodal_log_normal": 6,
    "exponential": 7, "mono": 8,
}

SEED_NAMES = ("lambda", "mr", "mi", "a1", "a2", "a3", "a4", "a5", "r1", "r2")
SCALAR_OUT = ("r1", "r2", "norm", "reff", "veff", "gavg", "vavg", "ravg",
              "rvw", "cext", "csca", "cbak", "g")


class LmieIn(C.Structure):
    """lmie_in_data, src/lmie_interface.h:20-58."""
    _fields_ = [
        ("save_control", C.c_int), ("calc_gc", C.c_int), ("calc_lc", C.c_int),
        ("calc_pf", C.c_int), ("dist_type", C.c_int), ("n_int1", C.c_int),
        ("n_int2", C.c_int), ("n_quad", C.c_int), ("n_angles", C.c_int),
        ("n_derivs", C.c_int),
    ] + [(n, C.c_double) for n in ("lambda_", "mr", "mi", "a1", "a2", "a3", "a4",
                                   "a5", "r1", "r2")] \
      + [(n + "_l", DP) for n in SEED_NAMES] + [("accuracy", C.c_double)]


class LmieOut(C.Structure):
    """lmie_out_data, src/lmie_interface.h:61-101."""
    _fields_ = [("n_coef", C.c_int)] \
        + [(n, C.c_double) for n in SCALAR_OUT] \
        + [("gc", DPP), ("lc", DPP), ("theta", DP), ("pf", DPP)] \
        + [(n + "_l", DP) for n in SCALAR_OUT] \
        + [("gc_l", DPPP), ("lc_l", DPPP), ("pf_l", DPPP),
           ("save1", DPP), ("save2", DPPP)]


def bind(lib):
    """Declare argument/return types of the ABI functions on a loaded CDLL."""
    lib.lmie_calc_max_coef.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_double]
    lib.lmie_calc_max_coef.restype = C.c_int
    lib.lmie_in_alloc.argtypes = [C.POINTER(LmieIn), C.c_int]
    lib.lmie_in_alloc.restype = C.c_int
    lib.lmie_in_free.argtypes = [C.POINTER(LmieIn), C.c_int]
    lib.lmie_in_free.restype = None
    lib.lmie_out_alloc.argtypes = [C.POINTER(LmieIn), C.POINTER(LmieOut), C.c_int]
    lib.lmie_out_alloc.restype = C.c_int
    lib.lmie_out_free.argtypes = [C.POINTER(LmieOut), C.c_int]
    lib.lmie_out_free.restype = None
    lib.lmie_in_zero_derivs.argtypes = [C.POINTER(LmieIn), C.c_int]
    lib.lmie_in_zero_derivs.restype = None
    lib.lmie_solution.argtypes = [C.POINTER(LmieIn), C.POINTER(LmieOut), C.c_int,
                                  C.c_int, C.c_int, C.c_int]
    lib.lmie_solution.restype = C.c_int
    return lib


def _arr(ptr, shape):
    n = int(np.prod(shape))
    if n == 0:
        return np.zeros(shape)
    return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(shape).copy()


def n_size_points(p):
    """Number of size nodes (misc/rtutil_size_dist.c:20-49)."""
    dt = p["dist_type"]
    if dt == SIZE_DIST["mono"]:
        return 1
    if dt == SIZE_DIST["modified_power_law"]:
        return (p.get("n_int1", 0) + p["n_int2"]) * p["n_quad"]
    return p["n_int2"] * p["n_quad"]


def make_in(lib, p, seeds=None):
    """Build an LmieIn from a parameter dict and optional seed vectors.

    ``seeds`` maps names in SEED_NAMES to sequences of length n_derivs.
    Returns the struct (seed vectors allocated by the library's lmie_in_alloc).
    """
    seeds = seeds or {}
    nd = 0
    for v in seeds.values():
        nd = max(nd, len(v))
    inp = LmieIn()
    inp.save_control = int(p.get("save_control", 0))
    inp.calc_gc = int(p.get("calc_gc", 1))
    inp.calc_lc = int(p.get("calc_lc", 0))
    inp.calc_pf = int(p.get("calc_pf", 0))
    dt = p["dist_type"]
    inp.dist_type = SIZE_DIST[dt] if isinstance(dt, str) else int(dt)
    inp.n_int1 = int(p.get("n_int1", 0))
    inp.n_int2 = int(p["n_int2"])
    inp.n_quad = int(p["n_quad"])
    inp.n_angles = int(p.get("n_angles", 0))
    inp.n_derivs = nd
    inp.lambda_ = p["lambda"]
    for k in ("mr", "mi", "a1", "a2", "a3", "a4", "a5", "r1", "r2"):
        setattr(inp, k, float(p.get(k, 0.0)))
    inp.accuracy = float(p.get("accuracy", 1e-7))
    if nd > 0:
        lib.lmie_in_alloc(C.byref(inp), nd)
        lib.lmie_in_zero_derivs(C.byref(inp), nd)
        for name, vec in seeds.items():
            ptr = getattr(inp, name + "_l")
            for i, v in enumerate(vec):
                ptr[i] = float(v)
    return inp


def read_out(lib, p, inp, out, max_coef):
    """Copy every output of a finished lmie_solution call into numpy."""
    nd = inp.n_derivs
    r = {"n_coef": out.n_coef, "max_coef": max_coef}
    for k in SCALAR_OUT:
        r[k] = getattr(out, k)
    if inp.calc_gc:
        r["gc"] = _arr(out.gc[0], (6, max_coef))
    if inp.calc_lc:
        r["lc"] = _arr(out.lc[0], (6, max_coef))
    if inp.calc_pf:
        r["theta"] = _arr(out.theta, (inp.n_angles,))
        r["pf"] = _arr(out.pf[0], (6, inp.n_angles))
    if nd > 0:
        for k in SCALAR_OUT:
            r[k + "_l"] = _arr(getattr(out, k + "_l"), (nd,))
        if inp.calc_gc:
            r["gc_l"] = _arr(out.gc_l[0][0], (nd, 6, max_coef))
        if inp.calc_lc:
            r["lc_l"] = _arr(out.lc_l[0][0], (nd, 6, max_coef))
        if inp.calc_pf:
            r["pf_l"] = _arr(out.pf_l[0][0], (nd, 6, inp.n_angles))
    if inp.save_control == 1 and bool(out.save1):
        n = n_size_points(p)
        q = max_coef if (inp.calc_gc or inp.calc_lc) else 0
        r["save1"] = _arr(out.save1[0], (4, n))
        if q > 0:
            r["save2"] = _arr(out.save2[0][0], (4, n, q))
    return r


def solve_resaved(lib, p1, p2, seeds2=None, n_threads=1):
    """The save/re-weight pair of lmie_solution (src/lmie_schedule.c:1087-1124):
    a save_control=1 call with p1 keeps the per-size arrays, then a
    save_control=2 call with p2 (same size grid and angles, new distribution
    parameters and weight-only seeds) re-weights them.  Returns the outputs of
    the second call.  The saved blocks are single allocations of the library,
    released with the C library's free (lmie_out_free leaves them, as the
    reference's does)."""
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    in1 = make_in(lib, dict(p1, save_control=1), None)
    out1 = LmieOut()
    rc = lib.lmie_solution(C.byref(in1), C.byref(out1), 1, 0, int(n_threads), 0)
    if rc != 0:
        raise RuntimeError("lmie_solution (save_control 1) returned %d" % rc)
    s1, s2 = out1.save1, out1.save2
    lib.lmie_out_free(C.byref(out1), 0)
    try:
        in2 = make_in(lib, dict(p2, save_control=2), seeds2)
        out2 = LmieOut()
        out2.save1, out2.save2 = s1, s2
        max_coef = lib.lmie_calc_max_coef(in2.lambda_, in2.dist_type, in2.a1, in2.a2,
                                          in2.a3, in2.r1, in2.r2)
        rc = lib.lmie_solution(C.byref(in2), C.byref(out2), 1, 0, int(n_threads), 0)
        if rc != 0:
            if in2.n_derivs > 0:
                lib.lmie_in_free(C.byref(in2), 1)
            raise RuntimeError("lmie_solution (save_control 2) returned %d" % rc)
        try:
            return read_out(lib, p2, in2, out2, max_coef)
        finally:
            lib.lmie_out_free(C.byref(out2), 1 if in2.n_derivs > 0 else 0)
            if in2.n_derivs > 0:
                lib.lmie_in_free(C.byref(in2), 1)
    finally:
        libc.free(C.cast(s1, C.c_void_p))
        libc.free(C.cast(s2, C.c_void_p))


def solve(lib, p, seeds=None, n_threads=1):
    """Run lmie_solution(alloc_out=1) and return the outputs as numpy arrays.

    Raises RuntimeError when the call returns -1.
    """
    inp = make_in(lib, p, seeds)
    out = LmieOut()
    max_coef = lib.lmie_calc_max_coef(inp.lambda_, inp.dist_type, inp.a1, inp.a2,
                                      inp.a3, inp.r1, inp.r2)
    rc = lib.lmie_solution(C.byref(inp), C.byref(out), 1, 0, int(n_threads), 0)
    if rc != 0:
        if inp.n_derivs > 0:
            lib.lmie_in_free(C.byref(inp), 1)
        raise RuntimeError("lmie_solution returned %d" % rc)
    try:
        return read_out(lib, p, inp, out, max_coef)
    finally:
        lib.lmie_out_free(C.byref(out), 1 if inp.n_derivs > 0 else 0)
        if inp.n_derivs > 0:
            lib.lmie_in_free(C.byref(inp), 1)
