"""CPU-side checks of the product library: it loads, exports every function the
headers in include/ declare, and its host-only paths (allocation, argument
checks that fail before any device work) behave like the reference's
(src/lmie_interface.c:24-169, :404-408)."""
import ctypes as C
import glob
import os
import re
import subprocess

import numpy as np

import pytest

from conftest import ROOT
from lmie_amd import abi, lmie

LIB = os.path.join(ROOT, "lmie_amd", "lib", "liblmie_amd.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "lmie_amd", "csrc")])
    return lmie.bind_all(C.CDLL(LIB))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(lmie_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_reference_interface():
    names = declared_functions()
    for f in ("lmie_calc_max_coef", "lmie_in_alloc", "lmie_in_free", "lmie_out_alloc", "lmie_out_free",
              "lmie_in_zero_derivs", "lmie_solution", "lmie_solution2", "lmie_solution2_l",
              "lmie_solution_slave", "lmie_core_solution"):
        assert f in names


def test_library_exports_every_declared_function(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layout_matches_the_reference():
    # field order and types of src/lmie_interface.h:20-101 on LP64
    assert C.sizeof(abi.LmieIn) == 10 * 4 + 10 * 8 + 10 * 8 + 8
    assert abi.LmieIn.lambda_.offset == 40
    assert abi.LmieIn.accuracy.offset == 200
    assert abi.LmieOut.gc.offset == 8 + 13 * 8
    assert C.sizeof(abi.LmieOut) == 8 + 13 * 8 + 4 * 8 + 13 * 8 + 3 * 8 + 2 * 8


def test_max_coef_matches_reference_formula(lib):
    from lmie_amd import workloads as W
    for p in (W.CFG0, W.CFG1, W.CFG2, W.cfg4_mode_params(0)):
        x = 2 * 3.141592653589793 * p["r2"] / p["lambda"]
        assert lib.lmie_calc_max_coef(p["lambda"], abi.SIZE_DIST[p["dist_type"]], p["a1"], p["a2"],
                                      0.0, p["r1"], p["r2"]) == 2 * W.calc_n1(x)


def test_alloc_and_free_roundtrip(lib):
    p = dict(dist_type="log_normal", n_int2=2, n_quad=4, calc_gc=1, calc_lc=1, calc_pf=1, n_angles=7,
             **{"lambda": 0.5}, mr=1.5, mi=0.01, a1=0.3, a2=0.2, r1=0.01, r2=2.0)
    inp = abi.make_in(lib, p, {"mr": [1, 0, 0]})
    assert inp.n_derivs == 3 and inp.mr_l[0] == 1.0 and inp.mi_l[2] == 0.0
    out = abi.LmieOut()
    assert lib.lmie_out_alloc(C.byref(inp), C.byref(out), 40) == 0
    # pointer-table blocks: rows contiguous
    a0 = C.cast(out.gc[0], C.c_void_p).value
    a1 = C.cast(out.gc[1], C.c_void_p).value
    assert a1 - a0 == 40 * 8
    b0 = C.cast(out.gc_l[0][0], C.c_void_p).value
    b1 = C.cast(out.gc_l[1][0], C.c_void_p).value
    assert b1 - b0 == 6 * 40 * 8
    lib.lmie_out_free(C.byref(out), 1)
    lib.lmie_in_free(C.byref(inp), 1)


def test_argument_errors_fail_before_device_work(lib):
    p = dict(dist_type="log_normal", n_int2=2, n_quad=4, calc_gc=0, calc_lc=0, calc_pf=1, n_angles=7,
             **{"lambda": 0.5}, mr=1.5, mi=0.01, a1=0.3, a2=0.2, r1=0.01, r2=2.0)
    with pytest.raises(RuntimeError):
        abi.solve(lib, p)                                  # calc_pf without gc or lc
    inp = abi.make_in(lib, dict(p, calc_gc=1))
    out = abi.LmieOut()
    assert lib.lmie_solution(C.byref(inp), C.byref(out), 1, 0, 1, 1) == -1   # use_mpi
    assert "MPI" in lmie.last_error(lib)
    lib.lmie_out_free(C.byref(out), 0)
    bad = abi.make_in(lib, dict(p, calc_gc=1, dist_type=42))
    out = abi.LmieOut()
    assert lib.lmie_solution(C.byref(bad), C.byref(out), 1, 0, 1, 0) == -1   # invalid distribution
    assert lib.lmie_solution_slave() == -1


def test_batch_reports_each_failing_case(lib):
    """lmie_solution_adjoint_batch on cases that fail in host validation: every
    case runs, each gets status -1, nothing stays allocated and the error names
    the first failing case."""
    lmie.bind_all(lib)
    good = dict(dist_type="log_normal", n_int2=2, n_quad=4, calc_gc=1, calc_lc=0, n_angles=0,
                **{"lambda": 0.5}, mr=1.5, mi=0.01, a1=0.3, a2=0.2, r1=0.01, r2=2.0)
    cases = [dict(good, dist_type=42), dict(good, dist_type=43), dict(good, dist_type=44)]
    for workers in (0, 2):
        with pytest.raises(RuntimeError, match="case 0"):
            lmie.adjoint_batch(lib, cases, [0.0] * 3, [1.0] * 3, "mixed", n_workers=workers)
    ins = [abi.make_in(lib, p) for p in cases]
    outs = [abi.LmieOut() for _ in cases]
    pin = (C.POINTER(abi.LmieIn) * 3)(*[C.pointer(x) for x in ins])
    pout = (C.POINTER(abi.LmieOut) * 3)(*[C.pointer(o) for o in outs])
    grads = (lmie.LmieGrad * 3)()
    status = (C.c_int * 3)(7, 7, 7)
    cot = C.cast(lib.lmie_cotangent_gc_l2, C.c_void_p)
    rc = lib.lmie_solution_adjoint_batch(3, pin, pout, 1, None, cot, None, grads, None, status, 3)
    assert rc == -1 and list(status) == [-1, -1, -1]
    assert lmie.last_error(lib).startswith("lmie_solution_adjoint_batch: case 0:")
    assert all(not bool(o.gc) for o in outs)
    assert lib.lmie_solution_adjoint_batch(3, pin, pout, 1, None, None, None, grads, None, status, 3) == -1
    assert "cotangent" in lmie.last_error(lib)
    assert lib.lmie_solution_adjoint_batch(0, None, None, 1, None, cot, None, None, None, None, 0) == 0
    status = (C.c_int * 3)(7, 7, 7)
    assert lib.lmie_solution_adjoint_batch(3, pin, pout, 1, None, cot, None, grads, None, status, 0) == -1
    assert list(status) == [-1, -1, -1] and all(not bool(o.gc) for o in outs)


def _build_dropin(tmpdir):
    """Compiles tests/c/dropin_caller.c against include/ and links it to the
    product library (as an existing LMie C caller would be built)."""
    exe = os.path.join(str(tmpdir), "dropin_caller")
    libdir = os.path.dirname(LIB)
    subprocess.check_call(["gcc", "-O1", "-std=gnu99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                           "-o", exe, os.path.join(ROOT, "tests", "c", "dropin_caller.c"),
                           "-L", libdir, "-llmie_amd", "-Wl,-rpath," + libdir])
    return exe


def test_c_caller_builds_against_the_drop_in_abi(lib, tmp_path):
    """A C program using only the reference's interface compiles and links
    against the product library (it runs in tests/test_gpu_fp64.py)."""
    assert os.path.exists(_build_dropin(tmp_path))


def test_scattering_files_match_the_reference_writers(lib, reflib, tmp_path):
    """The calllmie text files written by the library (header, gc and lc
    coefficients, phase function; the values and a tangent) equal byte for byte
    those of the reference's own writers (misc/rtutil_scat_io.c, built in
    oracle/_ref) on the same data, with values over the whole exponent range."""
    libc = C.CDLL(None)
    libc.fopen.restype, libc.fopen.argtypes = C.c_void_p, [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    DPP = C.POINTER(C.POINTER(C.c_double))
    for name, args in (("lmie_scat_coefs_write", [C.c_int, C.c_int]), ("lmie_phase_func_write", [C.c_int])):
        getattr(lib, name).argtypes = [C.c_void_p, C.POINTER(abi.LmieIn), C.POINTER(abi.LmieOut)] + args
        getattr(lib, name).restype = C.c_int
    reflib.scat_header_import.argtypes = [C.c_void_p] + [C.c_double] * 3 + [C.c_int] + [C.c_double] * 7 + \
        [C.c_int] * 3 + [C.c_double] * 12
    reflib.scat_header_import.restype = None
    reflib.scat_coefs_write_fp.argtypes = [C.c_void_p, C.c_int, C.c_void_p, DPP, C.c_int]
    reflib.phase_func_write_fp.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_double), DPP, C.c_int, C.c_int]
    p = dict(dist_type="power_law", n_int2=2, n_quad=8, calc_gc=1, calc_lc=1, calc_pf=1, n_angles=7,
             **{"lambda": 0.63}, mr=1.5, mi=0.01, a1=0.8, a2=0.15, r1=0.0, r2=0.0, accuracy=1e-7)
    inp = abi.make_in(lib, p, {"mr": [1.0, 0.0], "a1": [0.0, 1.0]})
    out = abi.LmieOut()
    max_coef, n_coef = 40, 37
    assert lib.lmie_out_alloc(C.byref(inp), C.byref(out), max_coef) == 0
    rng = np.random.Generator(np.random.PCG64(5))

    def val():
        return float(rng.standard_normal() * 10.0 ** int(rng.integers(-300, 300)))

    out.n_coef = n_coef
    for k in ("r1", "r2", "norm", "reff", "veff", "gavg", "vavg", "ravg", "rvw", "cext", "csca", "cbak", "g"):
        setattr(out, k, val())
    for i in range(p["n_angles"]):
        out.theta[i] = float(rng.uniform(0, 180))
    for k in range(6):
        for i in range(max_coef):
            out.gc[k][i] = val(); out.lc[k][i] = val()
            for d in range(2):
                out.gc_l[d][k][i] = val(); out.lc_l[d][k][i] = val()
        for i in range(p["n_angles"]):
            out.pf[k][i] = val()
            for d in range(2):
                out.pf_l[d][k][i] = val()
    h = C.create_string_buffer(4096)
    reflib.scat_header_import(h, inp.lambda_, inp.mr, inp.mi, inp.dist_type, inp.a1, inp.a2, inp.a3, inp.a4,
                              inp.a5, out.r1, out.r2, inp.n_int1, inp.n_int2, inp.n_quad, inp.accuracy,
                              out.norm, out.reff, out.veff, out.gavg, out.vavg, out.ravg, out.rvw,
                              out.cext, out.csca, out.cbak, out.g)
    cases = []
    for deriv in (-1, 0, 1):
        for leg in (0, 1):
            src = (out.lc if leg else out.gc) if deriv < 0 else (out.lc_l if leg else out.gc_l)[deriv]
            cases.append(("coefs_%d_%d" % (deriv, leg),
                          lambda fp, leg=leg, deriv=deriv: lib.lmie_scat_coefs_write(fp, C.byref(inp), C.byref(out),
                                                                                     leg, deriv),
                          lambda fp, src=src: reflib.scat_coefs_write_fp(fp, n_coef, h, src, 1)))
        pf = out.pf if deriv < 0 else out.pf_l[deriv]
        cases.append(("pf_%d" % deriv,
                      lambda fp, deriv=deriv: lib.lmie_phase_func_write(fp, C.byref(inp), C.byref(out), deriv),
                      lambda fp, pf=pf: reflib.phase_func_write_fp(fp, h, out.theta, pf, p["n_angles"], 1)))
    for name, ours, theirs in cases:
        texts = []
        for tag, fn in (("ours", ours), ("ref", theirs)):
            path = str(tmp_path / ("%s_%s.txt" % (name, tag)))
            fp = libc.fopen(path.encode(), b"w")
            assert fp
            assert fn(fp) == 0, (name, tag)
            libc.fclose(fp)
            texts.append(open(path, "rb").read())
        assert texts[0] == texts[1], name
        # 25 header lines, the row count, the rows
        assert texts[0].count(b"\n") == 25 + 1 + (n_coef if name.startswith("coefs") else p["n_angles"]), name
    # errors: a derivative that does not exist
    fp = libc.fopen(str(tmp_path / "bad.txt").encode(), b"w")
    assert lib.lmie_scat_coefs_write(fp, C.byref(inp), C.byref(out), 0, 2) == -1
    libc.fclose(fp)
    lib.lmie_out_free(C.byref(out), 1)
    lib.lmie_in_free(C.byref(inp), 1)


def test_out_alloc_failure_frees_everything(lib):
    """An allocation that cannot be satisfied returns -1 and leaves every array
    pointer NULL (src/lmie_interface.c:79-125 returns -1 on the first failing
    allocation); lmie_out_free on the result is then a no-op."""
    p = dict(dist_type="log_normal", n_int2=2, n_quad=4, calc_gc=1, calc_lc=1, calc_pf=1, n_angles=7,
             **{"lambda": 0.5}, mr=1.5, mi=0.01, a1=0.3, a2=0.2, r1=0.01, r2=2.0)
    inp = abi.make_in(lib, p, {"mr": [1.0] * 4096})
    out = abi.LmieOut()
    # 4096 x 6 x 2^31 doubles cannot be allocated
    assert lib.lmie_out_alloc(C.byref(inp), C.byref(out), 2 ** 31 - 1) == -1
    assert "out of memory" in lmie.last_error(lib)
    for k in ("gc", "lc", "pf", "theta", "gc_l", "lc_l", "pf_l", "cext_l", "g_l"):
        assert not bool(getattr(out, k)), k
    lib.lmie_out_free(C.byref(out), 1)
    lib.lmie_in_free(C.byref(inp), 1)


def test_writers_refuse_outputs_that_were_not_computed(lib, tmp_path):
    """The tangent file of an output that was not requested (calc_lc = 0 or
    calc_pf = 0 with seeds) is an error, not a dereference of a NULL block."""
    libc = C.CDLL(None)
    libc.fopen.restype, libc.fopen.argtypes = C.c_void_p, [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    lib.lmie_scat_coefs_write.argtypes = [C.c_void_p, C.POINTER(abi.LmieIn), C.POINTER(abi.LmieOut), C.c_int,
                                          C.c_int]
    lib.lmie_phase_func_write.argtypes = [C.c_void_p, C.POINTER(abi.LmieIn), C.POINTER(abi.LmieOut), C.c_int]
    p = dict(dist_type="log_normal", n_int2=2, n_quad=4, calc_gc=1, calc_lc=0, calc_pf=0, n_angles=7,
             **{"lambda": 0.5}, mr=1.5, mi=0.01, a1=0.3, a2=0.2, r1=0.01, r2=2.0)
    inp = abi.make_in(lib, p, {"mr": [1.0, 0.0]})
    out = abi.LmieOut()
    assert lib.lmie_out_alloc(C.byref(inp), C.byref(out), 12) == 0
    assert not bool(out.lc_l) and not bool(out.pf_l) and bool(out.gc_l)
    out.n_coef = 12
    fp = libc.fopen(str(tmp_path / "x.txt").encode(), b"w")
    for deriv in (-1, 0, 1):
        assert lib.lmie_scat_coefs_write(fp, C.byref(inp), C.byref(out), 1, deriv) == -1
        assert lib.lmie_phase_func_write(fp, C.byref(inp), C.byref(out), deriv) == -1
        assert lib.lmie_scat_coefs_write(fp, C.byref(inp), C.byref(out), 0, deriv) == 0
    libc.fclose(fp)
    lib.lmie_out_free(C.byref(out), 1)
    lib.lmie_in_free(C.byref(inp), 1)


@pytest.mark.parametrize("kind", ["power_law", "mono"])
@pytest.mark.parametrize("position", [0, 1])
def test_mixture_rejects_modes_with_their_own_grid(lib, kind, position):
    """A power-law or mono mode places its nodes from its own parameters, so it
    cannot share a mixture's grid, in any position of the mode list."""
    base = dict(dist_type="log_normal", n_int2=2, n_quad=4, calc_gc=1, calc_lc=0, n_angles=0,
                **{"lambda": 0.5}, mr=1.5, mi=0.01, a1=0.3, a2=0.2, r1=0.01, r2=2.0)
    other = dict(base, dist_type=kind, a1=0.5, a2=0.1)
    modes = [(base, None), (base, None)]
    modes[position] = (other, None)
    with pytest.raises(RuntimeError, match="cannot be mixed"):
        lmie.solve_mix(lib, modes, [0.5, 0.5], "fp64")
