"""Reverse-mode gradients of the GPU library against
  - the reference's forward tangents contracted with dL/d gc (the fixtures'
    loss_grad: sum_kl dL/dc_kl gc_l[seed][k][l] from the reference build), and
  - central differences of the library's own FP64 forward.
Loss (workloads.expansion_loss_targets): L = sum_k w_k sum_l (c_kl - t_kl)^2,
t = c_ref (1 + z), z ~ N(0,1), w_k = 1 / ||c_ref_k||^2.
Per-component bar: |g - g_ref| <= tol * max(|g_ref|, 1e-3 ||g_ref||)."""
import os

import numpy as np
import pytest

from conftest import load_golden, relnorm
from lmie_amd import lmie, workloads as W

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-8, "mixed": 1e-5}


def say(*a):
    print(*a, flush=True)


def loss_fn(c_ref, seed):
    t, w = W.expansion_loss_targets(c_ref, seed)

    def cot(out):
        return {"gc": 2.0 * w * (out["gc"] - t)}

    return cot


def check(g, ref, tol, label, floor_frac=1e-3):
    g, ref = np.asarray(g, float), np.asarray(ref, float)
    floor = floor_frac * np.linalg.norm(ref)
    err = np.abs(g - ref) / np.maximum(np.abs(ref), floor)
    say("  %s max rel err %.2e  (%s)" % (label, err.max(), " ".join("%.3e" % v for v in g)))
    assert err.max() <= tol, (label, g, ref)
    return err


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_adjoint_cfg2_matches_reference_tangents(product, prec):
    say("case: adjoint cfg2 (gamma, 16384 sizes, x 914)", prec)
    z, p, _ = load_golden("cfg2")
    out, g = lmie.adjoint(product, [p], None, loss_fn(z["gc"], 0), prec)
    check([g["a1"], g["a2"], g["mr"], g["mi"]], z["loss_grad"], TOL[prec], prec)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_adjoint_cfg4_two_modes_match_reference_tangents(product, prec):
    say("case: adjoint cfg4 (two modes, 32768 sizes, x 1848)", prec)
    z = np.load(load_golden.__globals__["GOLDEN"] + "/cfg4.npz")
    modes = [W.cfg4_mode_params(k) for k in range(2)]
    wts = [m["weight"] for m in W.CFG4_MODES]
    out, g = lmie.adjoint(product, modes, wts, loss_fn(z["gc"], 0), prec)
    gm = g["modes"]
    ds = [W.lognormal_sigma_seed(m["sigma_g"]) for m in W.CFG4_MODES]
    ours = [gm[0][0], gm[0][1] * ds[0], gm[1][0], gm[1][1] * ds[1], g["mr"], g["mi"]]
    check(ours, z["loss_grad"], TOL[prec], prec)


def test_adjoint_cfg3_cases_match_reference_tangents(product):
    say("case: adjoint cfg3 (32 retrieval cases)")
    z = np.load(load_golden.__globals__["GOLDEN"] + "/cfg3.npz")
    for c, (p, seeds) in enumerate(W.cfg3_cases()):
        c_ref = z["gc_%d" % c] if "gc_%d" % c in z.files else lmie.solve(product, p, None, "fp64")["gc"]
        out, g = lmie.adjoint(product, [p], None, loss_fn(c_ref, 100 + c), "fp64")
        ours = [g["a1"], g["a2"] * seeds["a2"][1], g["mr"], g["mi"]]
        check(ours, z["loss_grad"][c], TOL["fp64"], "case %d" % c)


FD_CASE = dict(dist_type="log_normal", n_int2=8, n_quad=64, calc_gc=1, calc_lc=1, **{"lambda": 0.61},
               mr=1.46, mi=0.004, a1=0.9, a2=0.3, r1=0.05, r2=6.0, accuracy=-1.0)


@pytest.mark.parametrize("part", ["cext", "csca", "cbak", "g", "gc", "lc"])
def test_adjoint_against_central_differences(product, part):
    say("case: adjoint vs central differences (log-normal, 512 sizes, x_max 62), cotangent on", part)
    p = FD_CASE
    rng = np.random.Generator(np.random.PCG64(11))
    base = lmie.solve(product, p, None, "fp64")
    cot = {"cext": 0.0, "csca": 0.0, "cbak": 0.0, "g": 0.0}
    if part in ("gc", "lc"):
        cot[part] = rng.standard_normal(base[part].shape) / np.abs(base[part]).max(axis=1, keepdims=True)
    else:
        cot[part] = 1.0

    def L(r):
        if part in ("gc", "lc"):
            return float(np.sum(cot[part] * r[part]))
        return float(r[part])

    out, g = lmie.adjoint(product, [p], None, lambda o: cot, "fp64")
    bad = []
    def central(key, h):
        return (L(lmie.solve(product, dict(p, **{key: p[key] + h}), None, "fp64")) -
                L(lmie.solve(product, dict(p, **{key: p[key] - h}), None, "fp64"))) / (2 * h)

    for key in ("mr", "mi", "a1", "a2", "lambda", "r1", "r2"):
        # Richardson on two steps cancels the h^2 term: the Mie resonances make
        # the third derivative in mr large enough that a plain central
        # difference at h = 1e-5 mr is off by ~1e-5 relative
        h = 1e-5 * abs(p[key])
        fd = (4.0 * central(key, 0.5 * h) - central(key, h)) / 3.0
        err = abs(g[key] - fd) / max(abs(fd), 1e-12)
        say("  %s adjoint %.10e  central difference %.10e  rel %.1e" % (key, g[key], fd, err))
        if err > 1e-6:
            bad.append(key)
    assert not bad, bad


PL_CASE = dict(FD_CASE, dist_type="power_law", n_int2=4, n_quad=32, a1=0.8, a2=0.15, r1=0.0, r2=0.0)
MONO_CASE = dict(FD_CASE, dist_type="mono", n_int2=1, n_quad=1, a1=1.3, a2=0.0, r1=1.3, r2=1.3)


@pytest.mark.parametrize("case", ["log_normal", "power_law", "mono"])
def test_adjoint_every_input_matches_forward_tangents(product, case):
    """The reverse pass against the FP64 forward tangents of all ten scalar
    inputs (lambda, m, a1..a5, r1, r2) contracted with one random cotangent on
    every output.  The power law's grid (r1, r2, nodes, weights) is set by a1
    and a2, so there a1 and a2 also move the size parameters.  The mono
    distribution's one size is a1, but a1 does not move its range, so its
    tangent acts on the weight only, as in the reference's forward tangents
    (src/lmie_interface.c:327-331): the reverse pass follows that."""
    p = {"log_normal": FD_CASE, "power_law": PL_CASE, "mono": MONO_CASE}[case]
    say("case: adjoint vs forward tangents of every input,", case)
    names = ("lambda", "mr", "mi", "a1", "a2", "a3", "a4", "a5", "r1", "r2")
    eye = np.eye(len(names))
    seeds = {k: list(eye[i]) for i, k in enumerate(names)}
    tan = lmie.solve(product, p, seeds, "fp64")
    rng = np.random.Generator(np.random.PCG64(13))
    cot = {k: float(rng.standard_normal()) for k in ("cext", "csca", "cbak", "g")}
    for k in ("gc", "lc"):
        cot[k] = rng.standard_normal(tan[k].shape) / np.abs(tan[k]).max(axis=1, keepdims=True)
    ref = np.zeros(len(names))
    for k in ("cext", "csca", "cbak", "g"):
        ref += cot[k] * tan[k + "_l"]
    for k in ("gc", "lc"):
        ref += np.einsum("dkl,kl->d", tan[k + "_l"], cot[k])
    for prec, tol in (("fp64", TOL["fp64"]), ("mixed", TOL["mixed"])):
        _, g = lmie.adjoint(product, [p], None, lambda o: cot, prec)
        ours = np.array([g[k] for k in names])
        say("  %s lambda %.6e r1 %.6e r2 %.6e (tangents %.6e %.6e %.6e)" % (prec, g["lambda"], g["r1"], g["r2"],
                                                                             ref[0], ref[8], ref[9]))
        check(ours, ref, tol, "%s %s" % (case, prec))


@pytest.mark.parametrize("part", ["cext", "csca", "cbak", "g", "gc", "lc"])
def test_mixed_adjoint_matches_fp64_adjoint(product, part):
    """The mixed reverse pass (split 16-bit kernel matrix) against the FP64
    reverse pass (stored sums, FP64 transposed contraction), per cotangent."""
    say("case: mixed vs fp64 adjoint (log-normal, 512 sizes, x_max 62), cotangent on", part)
    p = FD_CASE
    rng = np.random.Generator(np.random.PCG64(12))
    base = lmie.solve(product, p, None, "fp64")
    cot = {"cext": 0.0, "csca": 0.0, "cbak": 0.0, "g": 0.0}
    if part in ("gc", "lc"):
        cot[part] = rng.standard_normal(base[part].shape) / np.abs(base[part]).max(axis=1, keepdims=True)
    else:
        cot[part] = 1.0
    _, g64 = lmie.adjoint(product, [p], None, lambda o: cot, "fp64")
    _, gmx = lmie.adjoint(product, [p], None, lambda o: cot, "mixed")
    keys = ("mr", "mi", "a1", "a2", "lambda", "r1", "r2")
    check([gmx[k] for k in keys], [g64[k] for k in keys], TOL["mixed"], "mixed " + part)


def test_autograd_wrapper(product):
    torch = pytest.importorskip("torch")
    from lmie_amd import autograd
    say("case: torch.autograd wrapper (cfg0-like grid)")
    p = dict(W.CFG0)
    theta = torch.tensor([p["mr"], p["mi"], p["a1"], p["a2"]], dtype=torch.float64, requires_grad=True)
    gc = autograd.expansion(theta, product, [p])
    loss = (gc ** 2).sum()
    loss.backward()
    out, g = lmie.adjoint(product, [p], None, lambda o: {"gc": 2.0 * o["gc"]}, "fp64")
    ref = np.array([g["mr"], g["mi"], g["a1"], g["a2"]])
    assert np.allclose(theta.grad.numpy(), ref, rtol=1e-12, atol=0)
    # with the wavelength and the radius range
    theta = torch.tensor([p["mr"], p["mi"], p["a1"], p["a2"], p["lambda"], p["r1"], p["r2"]], dtype=torch.float64,
                         requires_grad=True)
    (autograd.expansion(theta, product, [p], grid=True) ** 2).sum().backward()
    ref = np.array([g["mr"], g["mi"], g["a1"], g["a2"], g["lambda"], g["r1"], g["r2"]])
    assert np.allclose(theta.grad.numpy(), ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("env", [("LMIE_MIX_PRODUCTS", "3"),
                                 ("LMIE_UGEMM_WAVES", "4"), ("LMIE_UGEMM_WAVES", "2"),
                                 ("LMIE_TILE_GROUP", "4")])
def test_contraction_shapes_agree(product, env):
    """Every selectable shape of the mixed forward and reverse contractions
    gives the default shape's outputs and gradients (cfg2: 16384 sizes,
    x_max 914)."""
    name, value = env
    say("case: contraction shape %s=%s (cfg2)" % env)
    z, p, _ = load_golden("cfg2")
    fn = loss_fn(z["gc"], 0)
    out0, g0 = lmie.adjoint(product, [p], None, fn, "mixed")
    old = os.environ.get(name)
    os.environ[name] = value
    try:
        out1, g1 = lmie.adjoint(product, [p], None, fn, "mixed")
    finally:
        if old is None:
            del os.environ[name]
        else:
            os.environ[name] = old
    assert relnorm(out1["gc"], out0["gc"]) <= 1e-6
    keys = ("mr", "mi", "a1", "a2")
    check([g1[k] for k in keys], [g0[k] for k in keys], 1e-5, "shape %s=%s" % env)


@pytest.mark.parametrize("workers", [0, 4])
def test_batch_of_cfg3_cases_in_one_call(product, workers):
    """The 32 retrieval cases in one lmie_solution_adjoint_batch call (mixed;
    workers 0: every kernel launched once for all cases, 4: four cases at a
    time on their own streams): every case's outputs and gradient equal those of its own
    single call bit for bit, the library's L2 loss equals the loss formed
    here, and the gradients meet the reference tangents at the mixed bar.
    The mixed path's gradient errors are normwise (up to about 4e-6 of
    ||g||): they come from the forward coefficients' errors (up to 2.5e-5
    elementwise), which enter the loss cotangent 2 w (c - t).  In some of these
    cases one component cancels to 1e-3 of ||g||, and such a component carries
    an absolute error of a few 1e-7 ||g||, so the bars are
    ||g - g_ref|| <= 5e-6 ||g_ref|| per case and, per component,
    |g - g_ref| <= 5e-5 max(|g_ref|, 1e-2 ||g_ref||).  Every case's errors
    are printed before the bars are checked."""
    say("case: cfg3 batch, 32 cases in one call, mixed, workers", workers)
    z = np.load(load_golden.__globals__["GOLDEN"] + "/cfg3.npz")
    cases = W.cfg3_cases()
    params = [p for p, _ in cases]
    c_refs = [z["gc_%d" % c] if "gc_%d" % c in z.files else lmie.solve(product, p, None, "fp64")["gc"]
              for c, p in enumerate(params)]
    tw = [W.expansion_loss_targets(c_refs[c], 100 + c) for c in range(len(params))]
    res = lmie.adjoint_batch(product, params, [t for t, _ in tw], [w for _, w in tw], "mixed",
                             n_workers=workers)
    fused = [product.lmie_last_batch_fused(p) for p in (0, 1)]
    say("  cases fused: forward %d, reverse %d" % tuple(fused))
    assert fused == ([len(params)] * 2 if workers == 0 else [0, 0])
    errs = []
    for c, (p, seeds) in enumerate(cases):
        out_b, g_b, loss_b = res[c]
        out, g = lmie.adjoint(product, [p], None, loss_fn(c_refs[c], 100 + c), "mixed")
        for k in ("cext", "csca", "cbak", "g"):
            assert out_b[k] == out[k], (c, k)
        assert np.array_equal(out_b["gc"], out["gc"]), c
        for k in ("lambda", "mr", "mi", "a1", "a2", "r1", "r2"):
            assert g_b[k] == g[k], (c, k, g_b[k], g[k])
        t, w = tw[c]
        loss = float(np.sum(w * (out["gc"] - t) ** 2))
        assert abs(loss_b - loss) <= 1e-12 * abs(loss), (c, loss_b, loss)
        ours = [g_b["a1"], g_b["a2"] * seeds["a2"][1], g_b["mr"], g_b["mi"]]
        ref = np.asarray(z["loss_grad"][c], float)
        nw = np.linalg.norm(np.asarray(ours) - ref) / np.linalg.norm(ref)
        comp = np.abs(np.asarray(ours) - ref) / np.maximum(np.abs(ref), 1e-2 * np.linalg.norm(ref))
        say("  case %d normwise %.2e, per component (floor 1e-2) %.2e" % (c, nw, comp.max()))
        errs.append((c, nw, comp.max(), ours, ref))
    worst_nw = max(errs, key=lambda e: e[1])
    worst_c = max(errs, key=lambda e: e[2])
    say("  worst normwise %.2e (case %d), worst per component %.2e (case %d)"
        % (worst_nw[1], worst_nw[0], worst_c[2], worst_c[0]))
    assert worst_nw[1] <= 5e-6, worst_nw
    assert worst_c[2] <= 5e-5, worst_c


@pytest.mark.parametrize("prec", ["fp64", "mixed-bf16"])
def test_batch_other_precisions_equal_single_calls(product, prec):
    """A batch of six small cases in FP64 (run one by one inside the batch's
    passes) and in the bf16 split (fused with the bf16 kernels): each case equals
    its single call bit for bit."""
    say("case: batch of 6 small log-normal cases,", prec)
    cases = [dict(FD_CASE, n_int2=4 + k, mr=1.40 + 0.03 * k, mi=0.001 * (k + 1), a1=0.5 + 0.2 * k,
                  **{"lambda": 0.5 + 0.07 * k}) for k in range(6)]
    refs = [lmie.solve(product, p, None, "fp64")["gc"] for p in cases]
    tw = [W.expansion_loss_targets(refs[k], 50 + k) for k in range(6)]
    res = lmie.adjoint_batch(product, cases, [t for t, _ in tw], [w for _, w in tw], prec)
    fused = [product.lmie_last_batch_fused(p) for p in (0, 1)]
    say("  cases fused: forward %d, reverse %d" % tuple(fused))
    assert fused == ([0, 0] if prec == "fp64" else [6, 6])
    for k, p in enumerate(cases):
        out_b, g_b, _ = res[k]
        out, g = lmie.adjoint(product, [p], None, loss_fn(refs[k], 50 + k), prec)
        assert np.array_equal(out_b["gc"], out["gc"]) and np.array_equal(out_b["lc"], out["lc"]), k
        for key in ("cext", "csca", "cbak", "g"):
            assert out_b[key] == out[key], (k, key, out_b[key], out[key])
        for key in ("lambda", "mr", "mi", "a1", "a2", "r1", "r2"):
            assert g_b[key] == g[key], (k, key, g_b[key], g[key])


def test_batch_with_a_failing_case(product):
    """A fused batch whose middle case fails host validation: the other cases
    still run fused and equal their single calls; the failing case is
    reported and leaves nothing allocated."""
    say("case: fused batch of 3, the middle case invalid")
    cases = [dict(FD_CASE, n_int2=6), dict(FD_CASE, dist_type=42), dict(FD_CASE, n_int2=7, mr=1.5)]
    refs = [lmie.solve(product, p, None, "fp64")["gc"] for p in (cases[0], cases[2])]
    tw = [W.expansion_loss_targets(refs[0], 7), (0.0, 1.0), W.expansion_loss_targets(refs[1], 8)]
    with pytest.raises(RuntimeError, match="case 1"):
        lmie.adjoint_batch(product, cases, [t for t, _ in tw], [w for _, w in tw], "mixed")
    assert [product.lmie_last_batch_fused(p) for p in (0, 1)] == [2, 2]
    ins = [lmie.abi.make_in(product, p) for p in cases]
    outs = [lmie.LmieOut() for _ in cases]
    import ctypes as C
    pin = (C.POINTER(lmie.LmieIn) * 3)(*[C.pointer(x) for x in ins])
    pout = (C.POINTER(lmie.LmieOut) * 3)(*[C.pointer(o) for o in outs])
    mc = [product.lmie_calc_max_coef(x.lambda_, x.dist_type, x.a1, x.a2, x.a3, x.r1, x.r2) for x in ins]
    keep, l2 = [], []
    for k in range(3):
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(tw[k][0], float), (6, max(mc[k], 1))))
        w = np.ascontiguousarray(np.broadcast_to(np.asarray(tw[k][1], float), (6, max(mc[k], 1))))
        b = np.zeros((6, max(mc[k], 1)))
        keep += [t, w, b]
        dp = C.POINTER(C.c_double)
        l2.append(lmie.LmieGcL2(t.ctypes.data_as(dp), w.ctypes.data_as(dp), b.ctypes.data_as(dp), 0.0))
    users = (C.c_void_p * 3)(*[C.cast(C.pointer(x), C.c_void_p) for x in l2])
    grads = (lmie.LmieGrad * 3)()
    status = (C.c_int * 3)(7, 7, 7)
    o = lmie._opts("mixed")
    rc = product.lmie_solution_adjoint_batch(3, pin, pout, 1, C.byref(o), C.cast(product.lmie_cotangent_gc_l2, C.c_void_p),
                                             users, grads, None, status, 0)
    assert rc == -1 and list(status) == [0, -1, 0] and not bool(outs[1].gc)
    for k, ref_k in ((0, 0), (2, 1)):
        fake = lmie.LmieIn.from_buffer_copy(ins[k])
        gc_b = lmie.abi.read_out(product, cases[k], fake, outs[k], mc[k])["gc"]
        out, g = lmie.adjoint(product, [cases[k]], None, loss_fn(refs[ref_k], 7 if k == 0 else 8), "mixed")
        assert np.array_equal(gc_b, out["gc"]), k
        assert grads[k].mr == g["mr"] and grads[k].a1 == g["a1"], k
        product.lmie_out_free(C.byref(outs[k]), 0)


def test_batch_larger_than_one_group(product):
    """70 tiny cases: the fused mode runs them as two device batches (64 + 6);
    every case equals its single call bit for bit."""
    say("case: fused batch of 70 tiny cases (two groups)")
    cases = [dict(FD_CASE, n_int2=2, n_quad=16, calc_lc=0, a1=0.3 + 0.01 * k, mr=1.35 + 0.002 * k,
                  **{"lambda": 0.4 + 0.005 * k}) for k in range(70)]
    refs = [lmie.solve(product, p, None, "fp64")["gc"] for p in cases]
    tw = [W.expansion_loss_targets(refs[k], 200 + k) for k in range(70)]
    res = lmie.adjoint_batch(product, cases, [t for t, _ in tw], [w for _, w in tw], "mixed")
    assert [product.lmie_last_batch_fused(p) for p in (0, 1)] == [70, 70]
    for k in (0, 31, 63, 64, 69):
        out_b, g_b, _ = res[k]
        out, g = lmie.adjoint(product, [cases[k]], None, loss_fn(refs[k], 200 + k), "mixed")
        assert np.array_equal(out_b["gc"], out["gc"]), k
        assert g_b["mr"] == g["mr"] and g_b["a1"] == g["a1"] and g_b["a2"] == g["a2"], k
