"""GPU parity of the FP64 path against the reference's outputs (fixtures made
from the reference build) and against the oracle.  Normwise relative error per
output array, bar 1e-12 (north star: FP64 path within 1e-12 of the reference)."""
import ctypes as C

import numpy as np
import pytest

from conftest import compare_all, load_golden, relnorm
from lmie_amd import abi, lmie, workloads as W

pytestmark = pytest.mark.gpu

SINGLE = ["small_lognormal", "small_gamma", "small_modgamma", "small_powerlaw", "small_modpowerlaw",
          "small_bimodal", "small_exponential", "small_mono", "cfg0", "example", "cfg1"]


def say(*a):
    print(*a, flush=True)


def test_mfma_f64_lane_layout(product):
    say("case: mfma f64 16x16x4 lane layout")
    rng = np.random.Generator(np.random.PCG64(3))
    A = rng.standard_normal((16, 4))
    B = rng.standard_normal((4, 16))
    D = np.zeros((16, 16))
    f = product.lmie_gpu_selftest_mfma_f64
    f.argtypes = [C.POINTER(C.c_double)] * 3 + [C.c_int]
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    assert f(dp(np.ascontiguousarray(A)), dp(np.ascontiguousarray(B)), dp(D), 0) == 0
    assert np.abs(D - A @ B).max() <= 1e-14


@pytest.mark.parametrize("n", [84, 522, 1922, 3810])
def test_gauss_rule_is_bitwise_the_reference_rule(product, oracle, n):
    say("case: gauss rule n=%d" % n)
    x, w = np.zeros(n), np.zeros(n)
    xo, wo = (C.c_double * n)(), (C.c_double * n)()
    f = product.lmie_gpu_gauss_rule
    f.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
    assert f(n, x.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)), 0) == 0
    oracle.oracle_gauss_leg(n, 0, -1.0, 1.0, None, None, xo, wo, None, None)
    xo, wo = np.array(xo), np.array(wo)
    if not (np.array_equal(x, xo) and np.array_equal(w, wo)):
        import os
        out = os.environ.get("LMIE_TEST_OUT", "test_out")
        os.makedirs(out, exist_ok=True)
        np.savez(os.path.join(out, "gauss_rule_%d.npz" % n), x=x, w=w, xo=xo, wo=wo)
        say("  nodes differ at %d, weights at %d; max |dx| %.3e, max |dw|/w %.3e" % (
            np.sum(x != xo), np.sum(w != wo), np.abs(x - xo).max(), (np.abs(w - wo) / wo).max()))
    assert np.array_equal(x, xo)
    assert np.array_equal(w, wo)


def test_gsf_tables_are_bitwise_the_oracle_functions(product, oracle):
    """The expansion tables of the angle kernels (gsf_table: each division a
    corrected product) equal the oracle's generalized spherical functions,
    computed with plain IEEE divisions in the reference's order, bit for bit,
    and P_l its Legendre recurrence; at the nodes of a 3810-point rule, at
    mu = 0 and +-1 and at random angles, to l = 3809."""
    n_l = 3810
    rng = np.random.Generator(np.random.PCG64(31))
    x, w = np.zeros(n_l), np.zeros(n_l)
    g = product.lmie_gpu_gauss_rule
    g.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
    assert g(n_l, x.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)), 0) == 0
    mu = np.concatenate([x[n_l // 2:][::37], [0.0, 1.0, -1.0], rng.uniform(-1.0, 1.0, 40)])
    say("case: gsf tables at %d angles, l < %d" % (mu.size, n_l))
    out = np.zeros((5, n_l, mu.size))
    f = product.lmie_gpu_gsf_table
    f.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double), C.c_int]
    m_c = np.ascontiguousarray(mu)
    assert f(mu.size, m_c.ctypes.data_as(C.POINTER(C.c_double)), n_l, out.ctypes.data_as(C.POINTER(C.c_double)), 0) == 0
    bad = []
    for k, m in enumerate(mu):
        P = [(C.c_double * n_l)() for _ in range(4)]
        oracle.oracle_gsf(n_l, float(m), *P)
        for q in range(4):
            if not np.array_equal(out[q, :, k], np.array(P[q])):
                bad.append((q, float(m)))
        # P_{l+1} = l / (l + 1) (mu P_l - P_{l-1}) + mu P_l
        leg = np.zeros(n_l)
        leg[0], leg[1] = 1.0, m
        for l in range(1, n_l - 1):
            u = m * leg[l]
            leg[l + 1] = l / (l + 1.0) * (u - leg[l - 1]) + u
        if not np.array_equal(out[4, :, k], leg):
            bad.append((4, float(m)))
    assert not bad, bad[:8]


def test_pitau_tables_are_bitwise_the_oracle_functions(product, oracle):
    """The FP64 angle tables (pitau_table: pi_n + tau_n, pi_n - tau_n) equal
    the oracle's pi_n, tau_n recurrence in the reference's order, added and
    subtracted, bit for bit; at 95 angles (nodes, 0, +-1, random), n <= 1905."""
    n, q = 1905, 3810
    rng = np.random.Generator(np.random.PCG64(32))
    x, w = np.zeros(q), np.zeros(q)
    g = product.lmie_gpu_gauss_rule
    g.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
    assert g(q, x.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)), 0) == 0
    mu = np.ascontiguousarray(np.concatenate([x[q // 2:][::37], [0.0, 1.0, -1.0], rng.uniform(-1.0, 1.0, 40)]))
    say("case: pi/tau tables at %d angles, n <= %d" % (mu.size, n))
    bp, bm = np.zeros((n, mu.size)), np.zeros((n, mu.size))
    f = product.lmie_gpu_pitau_table
    f.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    assert f(mu.size, dp(mu), n, dp(bp), dp(bm), 0) == 0
    bad = []
    for k, m in enumerate(mu):
        pi, tau = (C.c_double * n)(), (C.c_double * n)()
        oracle.oracle_pitau(float(m), n, pi, tau)
        pi, tau = np.array(pi), np.array(tau)
        if not (np.array_equal(bp[:, k], pi + tau) and np.array_equal(bm[:, k], pi - tau)):
            bad.append(float(m))
    assert not bad, bad[:8]


@pytest.mark.parametrize("name", SINGLE)
def test_fp64_matches_reference(product, name):
    say("case: fp64 parity", name)
    z, p, seeds = load_golden(name)
    r = lmie.solve(product, p, seeds, "fp64")
    worst = compare_all(r, z, 1e-12)
    say("  worst %s %.2e" % (worst[1], worst[0]))


def test_fp64_cfg2_matches_reference(product):
    say("case: fp64 parity cfg2 (16384 sizes, x 914, 4 seeds)")
    z, p, seeds = load_golden("cfg2")
    r = lmie.solve(product, p, seeds, "fp64")
    worst = compare_all(r, z, 1e-12)
    say("  worst %s %.2e" % (worst[1], worst[0]))


def test_fp64_cfg3_cases_match_reference(product):
    say("case: fp64 parity cfg3 (32 retrieval cases)")
    z = np.load(load_golden.__globals__["GOLDEN"] + "/cfg3.npz")
    for c, (p, seeds) in enumerate(W.cfg3_cases()):
        r = lmie.solve(product, p, None, "fp64")
        for k in ("cext", "csca", "cbak", "g"):
            assert abs(r[k] - z[k][c]) <= 1e-12 * abs(z[k][c]), (c, k)
        if "gc_%d" % c in z.files:
            assert relnorm(r["gc"], z["gc_%d" % c]) <= 1e-12, c


def test_fp64_cfg4_two_modes_match_reference(product):
    say("case: fp64 parity cfg4 (two modes, 32768 sizes, x 1848)")
    z = np.load(load_golden.__globals__["GOLDEN"] + "/cfg4.npz")
    modes = [(W.cfg4_mode_params(k), None) for k in range(2)]
    r = lmie.solve_mix(product, modes, [m["weight"] for m in W.CFG4_MODES], "fp64")
    for k in ("cext", "csca", "cbak", "g"):
        assert abs(r[k] - float(z[k])) <= 1e-12 * abs(float(z[k])), k
    e = relnorm(r["gc"], z["gc"])
    say("  gc rel err %.2e" % e)
    assert e <= 1e-12


def test_save_control_per_size_arrays(product):
    say("case: save_control 1 per-size arrays")
    z, p, seeds = load_golden("small_lognormal")
    r = lmie.solve(product, p, seeds, "fp64")
    assert relnorm(r["save1"], z["save1"]) <= 1e-12
    assert relnorm(r["save2"], z["save2"]) <= 1e-12


def test_resaved_per_size_arrays_match_reference(product, oracle):
    say("case: save_control 2 re-weighting")
    z = np.load(load_golden.__globals__["GOLDEN"] + "/resaved.npz")
    meta = __import__("json").loads(str(z["inputs_json"]))
    p1, p2, seeds = meta["params"]["p1"], meta["params"]["p2"], meta["seeds"]
    r = abi.solve_resaved(product, p1, p2, seeds, n_threads=1)
    compare_all(r, z, 1e-12)
    # and equal to solving anew with the second distribution
    f = abi.solve(product, dict(p2, save_control=0), seeds, n_threads=1)
    for k in ("cext", "g", "gc", "gc_l", "lc", "pf"):
        assert relnorm(r[k], f[k]) <= 1e-12, k


def test_repeat_calls_are_deterministic(product):
    say("case: determinism")
    z, p, seeds = load_golden("cfg1")
    a = lmie.solve(product, p, seeds, "fp64")
    b = lmie.solve(product, p, seeds, "fp64")
    for k in ("gc", "gc_l", "cext", "g_l"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_cross_sections_only_no_angles(product, oracle, prec):
    """calc_gc = calc_lc = 0: no Gauss angles, no contraction, no expansion;
    the series and their tangents only."""
    say("case: cross sections only (%s)" % prec)
    p = dict(dist_type="gamma", n_int2=3, n_quad=7, calc_gc=0, calc_lc=0, **{"lambda": 0.63},
             mr=1.41, mi=3e-3, a1=0.7, a2=0.25, r1=0.02, r2=5.0, accuracy=-1.0)
    seeds = {"a1": [1, 0, 0], "mr": [0, 1, 0], "lambda": [0, 0, 1]}
    r = lmie.solve(product, p, seeds, prec)
    o = abi.solve(oracle, p, seeds, 1)
    for k in ("cext", "csca", "cbak", "g", "reff", "veff"):
        assert abs(r[k] - o[k]) <= 1e-12 * abs(o[k]), k
    for k in ("cext_l", "csca_l", "cbak_l", "g_l"):
        assert relnorm(r[k], o[k]) <= 1e-12, k


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_three_size_points(product, oracle, prec):
    """A grid of 3 points (one padded tile, 61 padding lanes)."""
    say("case: 3 size points (%s)" % prec)
    p = dict(dist_type="log_normal", n_int2=1, n_quad=3, calc_gc=1, calc_lc=1, **{"lambda": 0.5},
             mr=1.55, mi=0.01, a1=0.6, a2=0.2, r1=0.1, r2=2.0, accuracy=-1.0)
    r = lmie.solve(product, p, None, prec)
    o = abi.solve(oracle, p, None, 1)
    tol = 1e-12 if prec == "fp64" else 1e-5
    assert abs(r["csca"] - o["csca"]) <= 1e-12 * abs(o["csca"])
    assert relnorm(r["gc"], o["gc"]) <= tol
    assert relnorm(r["lc"], o["lc"]) <= tol


def test_compiled_c_caller_matches_reference(product, tmp_path):
    """tests/c/dropin_caller.c, built against the reference's interface and
    linked to the product library, run as a separate process on the cfg0
    inputs: every scalar and the expansion equal the reference's fixture at the
    FP64 bar."""
    import subprocess
    from test_library import _build_dropin
    say("case: compiled C caller of lmie_solution, cfg0")
    z, p, _ = load_golden("cfg0")
    exe = _build_dropin(tmp_path)
    args = [str(abi.SIZE_DIST[p["dist_type"]]), str(p["n_int2"]), str(p["n_quad"])] + \
           ["%.17g" % p[k] for k in ("lambda", "mr", "mi", "a1", "a2", "r1", "r2")]
    txt = subprocess.run([exe] + args, check=True, capture_output=True, text=True, timeout=120).stdout
    rows = {ln.split()[0]: ln.split()[1:] for ln in txt.splitlines()}
    assert int(rows["n_coef"][0]) == int(z["n_coef"])
    sc = [float(v) for v in rows["scalars"]]
    for k, v in zip(("r1", "r2", "norm", "reff", "veff", "gavg", "vavg", "ravg", "rvw", "cext", "csca",
                     "cbak", "g"), sc):
        assert relnorm(v, z[k]) <= 1e-12, (k, v, float(z[k]))
    gc = np.array([[float(v) for v in rows["gc%d" % k]] for k in range(6)])
    n = gc.shape[1]
    assert relnorm(gc, z["gc"][:, :n]) <= 1e-12


def test_reference_example_program_runs_unchanged(product, tmp_path):
    """The reference's own example program (examples/example_c.c), unmodified,
    built against include/lmie_interface.h with the reference's platform
    flags and linked to the product library (oracle/ref.mk, where the
    reference sources are present): run on the GPU, its printed output equals
    the reference's examples/example.out to the printed 7 digits."""
    import os
    import subprocess
    from conftest import ROOT, GOLDEN
    from test_oracle import _parse_example_out
    exe = os.path.join(ROOT, "oracle", "_ref", "example_c_amd")
    if not os.path.exists(exe):
        pytest.skip("the reference's example was not built here (oracle/ref.mk needs the reference sources)")
    say("case: the reference's example_c linked to the product library")
    txt = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout
    out = tmp_path / "example.out"
    out.write_text(txt)
    ours, ref = _parse_example_out(str(out)), _parse_example_out(os.path.join(GOLDEN, "example.out"))
    assert len(ours) == len(ref) == 5
    for b, (o, r) in enumerate(zip(ours, ref)):
        assert o["scal"].keys() == r["scal"].keys(), b
        for k, v in r["scal"].items():
            assert abs(o["scal"][k] - v) <= 6e-7 * abs(v) + 1e-300, (b, k, o["scal"][k], v)
        for key in ("gc", "pf"):
            a, e = np.array(o[key]), np.array(r[key])
            assert a.shape == e.shape, (b, key)
            scale = np.abs(e).max()
            assert np.all(np.abs(a - e) <= 6e-7 * np.abs(e) + 1e-9 * scale), (b, key)
