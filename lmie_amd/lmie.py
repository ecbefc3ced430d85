"""ctypes wrappers of the product library (include/lmie_hip.h extensions).

``solve`` mirrors the reference call ``lmie_solution`` (src/lmie_interface.c:199-250)
with a precision option; ``solve_mix`` evaluates several distributions on one
size grid; ``adjoint`` runs the forward and one reverse pass and returns the
gradient of a scalar loss.  Inputs are the same parameter dictionaries as
lmie_amd.abi and lmie_amd.workloads.
"""
import ctypes as C
import math
import time

import numpy as np

from . import abi
from .abi import LmieIn, LmieOut

PRECISION = {"fp64": 0, "mixed": 1, "mixed-bf16": 2}

TIMERS = ("tables", "coefs", "angular", "reduce", "expand", "adj_angular", "adj_coefs", "total", "adj_gemm",
          "host_grid", "host_prep", "host_wait", "call")


class LmieOpts(C.Structure):
    """lmie_opts, include/lmie_hip.h."""
    _fields_ = [("precision", C.c_int), ("device", C.c_int), ("reserved", C.c_int * 6)]


class LmieAdjIn(C.Structure):
    """lmie_adj_in, include/lmie_hip.h."""
    _fields_ = [("cext", C.c_double), ("csca", C.c_double), ("cbak", C.c_double),
                ("g", C.c_double), ("gc", C.POINTER(C.c_double)), ("lc", C.POINTER(C.c_double))]


class LmieGrad(C.Structure):
    """lmie_grad, include/lmie_hip.h: dL/d of each scalar input."""
    _fields_ = [(n, C.c_double) for n in ("lambda_", "mr", "mi", "a1", "a2", "a3", "a4", "a5",
                                          "r1", "r2")]


def bind_all(lib):
    abi.bind(lib)
    PIn, POut = C.POINTER(LmieIn), C.POINTER(LmieOut)
    lib.lmie_solution_ex.argtypes = [PIn, POut, C.c_int, C.POINTER(LmieOpts)]
    lib.lmie_solution_ex.restype = C.c_int
    lib.lmie_solution_mix.argtypes = [C.POINTER(PIn), C.POINTER(C.c_double), C.c_int, POut,
                                      C.c_int, C.POINTER(LmieOpts)]
    lib.lmie_solution_mix.restype = C.c_int
    lib.lmie_adjoint_forward.argtypes = [C.POINTER(PIn), C.POINTER(C.c_double), C.c_int, POut,
                                         C.c_int, C.POINTER(LmieOpts), C.POINTER(C.c_void_p)]
    lib.lmie_adjoint_forward.restype = C.c_int
    lib.lmie_adjoint_reverse.argtypes = [C.c_void_p, C.POINTER(LmieAdjIn), C.POINTER(LmieGrad),
                                         C.POINTER(C.c_double)]
    lib.lmie_adjoint_reverse.restype = C.c_int
    lib.lmie_adjoint_free.argtypes = [C.c_void_p]
    lib.lmie_adjoint_free.restype = None
    PPOut = C.POINTER(POut)
    lib.lmie_solution_adjoint_batch.argtypes = [C.c_int, C.POINTER(PIn), PPOut, C.c_int,
                                                C.POINTER(LmieOpts), C.c_void_p,
                                                C.POINTER(C.c_void_p), C.POINTER(LmieGrad),
                                                C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int]
    lib.lmie_solution_adjoint_batch.restype = C.c_int
    lib.lmie_last_batch_fused.argtypes = [C.c_int]
    lib.lmie_last_batch_fused.restype = C.c_int
    lib.lmie_last_error.restype = C.c_char_p
    lib.lmie_last_precision.restype = C.c_int
    lib.lmie_gpu_last_error.restype = C.c_char_p
    lib.lmie_set_timing.argtypes = [C.c_int]
    lib.lmie_get_timing.argtypes = [C.POINTER(C.c_double), C.c_int]
    return lib


def _opts(precision, device=0):
    o = LmieOpts()
    o.precision = PRECISION[precision] if isinstance(precision, str) else int(precision)
    o.device = device
    return o


def last_error(lib):
    return (lib.lmie_last_error() or b"").decode()


def timings(lib):
    """Per-stage device times (ms) of this thread's last call, if timing is on."""
    buf = (C.c_double * len(TIMERS))()
    lib.lmie_get_timing(buf, len(TIMERS))
    return dict(zip(TIMERS, list(buf)))


def solve(lib, p, seeds=None, precision="fp64", device=0):
    """lmie_solution_ex with alloc_out = 1; returns numpy copies of every output."""
    inp = abi.make_in(lib, p, seeds)
    out = LmieOut()
    max_coef = lib.lmie_calc_max_coef(inp.lambda_, inp.dist_type, inp.a1, inp.a2, inp.a3,
                                      inp.r1, inp.r2)
    o = _opts(precision, device)
    rc = lib.lmie_solution_ex(C.byref(inp), C.byref(out), 1, C.byref(o))
    try:
        if rc != 0:
            raise RuntimeError("lmie_solution_ex failed: " + last_error(lib))
        return abi.read_out(lib, p, inp, out, max_coef)
    finally:
        # a failed call leaves what it allocated to the caller, as the reference
        # (the pointers of a failed allocation are NULL)
        lib.lmie_out_free(C.byref(out), 1 if inp.n_derivs > 0 else 0)
        if inp.n_derivs > 0:
            lib.lmie_in_free(C.byref(inp), 1)


def _modes(lib, modes):
    ins = [abi.make_in(lib, p, s) for p, s in modes]
    arr = (C.POINTER(LmieIn) * len(ins))(*[C.pointer(x) for x in ins])
    return ins, arr


def solve_mix(lib, modes, weights, precision="fp64", device=0):
    """lmie_solution_mix: modes is a list of (params, seeds) on one grid."""
    ins, arr = _modes(lib, modes)
    nd = sum(x.n_derivs for x in ins)
    w = (C.c_double * len(ins))(*weights)
    out = LmieOut()
    p0 = modes[0][0]
    max_coef = lib.lmie_calc_max_coef(ins[0].lambda_, ins[0].dist_type, ins[0].a1, ins[0].a2,
                                      ins[0].a3, ins[0].r1, ins[0].r2)
    o = _opts(precision, device)
    rc = lib.lmie_solution_mix(arr, w, len(ins), C.byref(out), 1, C.byref(o))
    try:
        if rc != 0:
            raise RuntimeError("lmie_solution_mix failed: " + last_error(lib))
        fake = LmieIn.from_buffer_copy(ins[0])
        fake.n_derivs = nd
        return abi.read_out(lib, p0, fake, out, max_coef)
    finally:
        lib.lmie_out_free(C.byref(out), 1 if nd > 0 else 0)
        for x in ins:
            if x.n_derivs > 0:
                lib.lmie_in_free(C.byref(x), 1)


class Adjoint:
    """Forward pass that keeps its device state, then reverse passes.

    modes: list of parameter dicts on one grid (one entry for a single
    distribution); weights: number weights (None for one mode).
    ``forward()`` returns the outputs; ``reverse(cot)`` takes the output
    cotangents (keys cext, csca, cbak, g: floats; gc, lc: [6][max_coef]) and
    returns dL/d input: keys lambda, mr, mi, a1..a5, r1, r2 (mode 0 for a..),
    and 'modes' = [n_modes][5] gradients of a1..a5 per mode.
    """

    def __init__(self, lib, modes, weights=None, precision="fp64", device=0):
        self.lib, self.modes = lib, modes
        self.weights = weights if weights is not None else [1.0] * len(modes)
        self.opts = _opts(precision, device)
        self.ctx = C.c_void_p()
        self.out = None

    def forward(self):
        lib = self.lib
        ins, arr = _modes(lib, [(p, None) for p in self.modes])
        n = len(ins)
        w = (C.c_double * n)(*self.weights)
        out = LmieOut()
        self.max_coef = lib.lmie_calc_max_coef(ins[0].lambda_, ins[0].dist_type, ins[0].a1, ins[0].a2,
                                               ins[0].a3, ins[0].r1, ins[0].r2)
        rc = lib.lmie_adjoint_forward(arr, w, n, C.byref(out), 1, C.byref(self.opts), C.byref(self.ctx))
        if rc != 0:
            raise RuntimeError("lmie_adjoint_forward failed: " + last_error(lib))
        try:
            fake = LmieIn.from_buffer_copy(ins[0])
            fake.n_derivs = 0
            self.out = abi.read_out(lib, self.modes[0], fake, out, self.max_coef)
        finally:
            lib.lmie_out_free(C.byref(out), 0)
        return self.out

    def reverse(self, cot):
        keep = []

        def arrp(key):
            v = cot.get(key)
            if v is None:
                return None
            a = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(6, self.max_coef))
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_double))

        adj = LmieAdjIn(float(cot.get("cext", 0.0)), float(cot.get("csca", 0.0)),
                        float(cot.get("cbak", 0.0)), float(cot.get("g", 0.0)), arrp("gc"), arrp("lc"))
        g = LmieGrad()
        n = len(self.modes)
        gm = (C.c_double * (5 * n))()
        rc = self.lib.lmie_adjoint_reverse(self.ctx, C.byref(adj), C.byref(g), gm)
        if rc != 0:
            raise RuntimeError("lmie_adjoint_reverse failed: " + last_error(self.lib))
        return {"lambda": g.lambda_, "mr": g.mr, "mi": g.mi, "a1": g.a1, "a2": g.a2, "a3": g.a3,
                "a4": g.a4, "a5": g.a5, "r1": g.r1, "r2": g.r2,
                "modes": np.array(list(gm)).reshape(n, 5)}

    def free(self):
        if self.ctx:
            self.lib.lmie_adjoint_free(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def adjoint(lib, modes, weights, cot_fn, precision="fp64", device=0):
    """Forward, cotangents from cot_fn(outputs) -> dict, one reverse pass."""
    a = Adjoint(lib, modes, weights, precision, device)
    try:
        out = a.forward()
        grad = a.reverse(cot_fn(out))
    finally:
        a.free()
    return out, grad


class LmieGcL2(C.Structure):
    """lmie_gc_l2, include/lmie_hip.h."""
    _fields_ = [("target", C.POINTER(C.c_double)), ("weight", C.POINTER(C.c_double)),
                ("gc_bar", C.POINTER(C.c_double)), ("loss", C.c_double)]


def adjoint_batch(lib, cases, targets, weights, precision="mixed", device=0, n_workers=0):
    """lmie_solution_adjoint_batch with the library's weighted L2 loss on gc
    (lmie_cotangent_gc_l2): case c is the single distribution cases[c], with
    L_c = sum_kl weights[c]_kl (gc_kl - targets[c]_kl)^2 over [6][max_coef].
    n_workers 0: one launch per kernel for all cases; > 0: that many cases at a
    time on their own streams.  Returns one (outputs, gradient, loss) triple per
    case, as ``adjoint``."""
    n = len(cases)
    ins = [abi.make_in(lib, p, None) for p in cases]
    outs = [LmieOut() for _ in range(n)]
    pin = (C.POINTER(LmieIn) * n)(*[C.pointer(x) for x in ins])
    pout = (C.POINTER(LmieOut) * n)(*[C.pointer(o) for o in outs])
    mc = [lib.lmie_calc_max_coef(x.lambda_, x.dist_type, x.a1, x.a2, x.a3, x.r1, x.r2) for x in ins]
    keep, l2 = [], []

    def full(a, m):
        # (6, m) contiguous doubles; arrays already of that form are used as they are
        a = np.asarray(a, float)
        return np.ascontiguousarray(a if a.shape == (6, m) else np.broadcast_to(a, (6, m)))

    for c in range(n):
        t, w = full(targets[c], mc[c]), full(weights[c], mc[c])
        b = np.zeros((6, mc[c]))
        keep += [t, w, b]
        dp = C.POINTER(C.c_double)
        l2.append(LmieGcL2(t.ctypes.data_as(dp), w.ctypes.data_as(dp), b.ctypes.data_as(dp), 0.0))
    users = (C.c_void_p * n)(*[C.cast(C.pointer(x), C.c_void_p) for x in l2])
    grads = (LmieGrad * n)()
    gm = (C.c_double * (5 * n))()
    status = (C.c_int * n)()
    o = _opts(precision, device)
    cot = C.cast(lib.lmie_cotangent_gc_l2, C.c_void_p)
    t0 = time.perf_counter()
    rc = lib.lmie_solution_adjoint_batch(n, pin, pout, 1, C.byref(o), cot, users, grads, gm,
                                         status, n_workers)
    adjoint_batch.call_seconds = time.perf_counter() - t0
    res, err = [], None if rc == 0 else last_error(lib)
    try:
        for c in range(n):
            if status[c] != 0:
                res.append(None)
                continue
            # the outputs carry no derivatives (read_out reads n_derivs from the input)
            fake = ins[c]
            if fake.n_derivs:
                fake = LmieIn.from_buffer_copy(ins[c])
                fake.n_derivs = 0
            out = abi.read_out(lib, cases[c], fake, outs[c], mc[c])
            g = grads[c]
            grad = {"lambda": g.lambda_, "mr": g.mr, "mi": g.mi, "a1": g.a1, "a2": g.a2,
                    "a3": g.a3, "a4": g.a4, "a5": g.a5, "r1": g.r1, "r2": g.r2,
                    "modes": np.array(list(gm[5 * c:5 * c + 5])).reshape(1, 5)}
            res.append((out, grad, l2[c].loss))
    finally:
        for c in range(n):
            if status[c] == 0:
                lib.lmie_out_free(C.byref(outs[c]), 0)
    if rc != 0:
        raise RuntimeError("lmie_solution_adjoint_batch failed: " + err)
    return res


def nan_to_none(v):
    return None if (v is None or (isinstance(v, float) and math.isnan(v))) else v
