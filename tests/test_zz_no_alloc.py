"""lmie_gpu_phase_func_static (include/lmie_gpu.h): the phase function from
expansion coefficients with its staging in a static device array instead of
per-call allocations.  Its results must equal lmie_gpu_phase_func's bit for bit
(the same kernel on the same inputs), on two calls in a row whose second is
smaller than the first (the array keeps the first call's contents), for
generalized-spherical-function and Legendre coefficients; sizes beyond the
array are refused before any device work.

Pytest collects this file last; it passes alone and after the rest of the
suite."""
import ctypes as C

import numpy as np
import pytest

DP = C.POINTER(C.c_double)
ARGS = [C.c_int, C.c_int, C.c_int, DP, C.c_int, DP, DP, C.c_int, C.c_int]
CAP = 1 << 21          # doubles in the static array


def say(*a):
    print(*a, flush=True)


def _fns(lib):
    for name in ("lmie_gpu_phase_func", "lmie_gpu_phase_func_static"):
        getattr(lib, name).argtypes = ARGS
        getattr(lib, name).restype = C.c_int
    lib.lmie_gpu_last_error.restype = C.c_char_p
    return lib.lmie_gpu_phase_func, lib.lmie_gpu_phase_func_static


def _coefs(rng, n_set, ld, n_coef):
    """Decaying coefficient sets [n_set][6][ld] (beta_0 = 1 for the first row),
    zero beyond n_coef."""
    l = np.arange(ld)
    x = rng.standard_normal((n_set, 6, ld)) * np.exp(-l / (0.3 * max(n_coef, 1)))
    x[:, 0, 0] = 1.0
    x[:, :, n_coef:] = 0.0
    return np.ascontiguousarray(x)


def _call(fn, n_set, n_coef, ld, coefs, theta, gsf):
    pf = np.full((n_set, 6, theta.size), np.nan)
    rc = fn(n_set, n_coef, ld, coefs.ctypes.data_as(DP), theta.size, theta.ctypes.data_as(DP),
            pf.ctypes.data_as(DP), gsf, 0)
    return rc, pf


@pytest.mark.gpu
@pytest.mark.parametrize("gsf", [1, 0])
def test_static_staging_equals_allocating_entry_point(product, gsf):
    alloc, static = _fns(product)
    rng = np.random.Generator(np.random.PCG64(70 + gsf))
    # the second call is smaller in every dimension than the first
    for n_set, ld, n_coef, n_theta in ((5, 700, 640, 1801), (2, 90, 77, 181)):
        say("case: phase_func_static gsf %d, n_set %d, ld %d, n_coef %d, n_theta %d" % (gsf, n_set, ld, n_coef, n_theta))
        coefs = _coefs(rng, n_set, ld, n_coef)
        theta = np.linspace(0.0, 180.0, n_theta)
        rc_a, pf_a = _call(alloc, n_set, n_coef, ld, coefs, theta, gsf)
        assert rc_a == 0, product.lmie_gpu_last_error()
        rc_s, pf_s = _call(static, n_set, n_coef, ld, coefs, theta, gsf)
        assert rc_s == 0, product.lmie_gpu_last_error()
        assert np.all(np.isfinite(pf_s))
        assert np.array_equal(pf_s, pf_a)
        # the forward direction of a normalised set: P11(0) = sum of beta_l
        b = coefs[0, 0, :n_coef]
        assert abs(pf_s[0, 0, 0] - float(np.sum(b))) <= 1e-12 * float(np.sum(np.abs(b)))


def _staging(n_set, ld, n_theta):
    """Doubles of the array one call uses: coefficients, angles and phase
    function, each part starting on a multiple of 32 doubles."""
    ot = (6 * n_set * ld + 31) // 32 * 32
    op = (ot + n_theta + 31) // 32 * 32
    return op + 6 * n_set * n_theta


@pytest.mark.gpu
def test_static_staging_at_its_limit(product):
    """The largest n_theta that fits beside 42 sets of 5376 coefficients, then
    a small call; one angle more is refused."""
    alloc, static = _fns(product)
    rng = np.random.Generator(np.random.PCG64(77))
    n_set, ld = 42, 5376
    n_theta = 1
    while _staging(n_set, ld, n_theta + 1) <= CAP:
        n_theta += 1
    say("case: phase_func_static at the limit, n_theta %d, %d of %d doubles"
        % (n_theta, _staging(n_set, ld, n_theta), CAP))
    for n_set_, ld_, n_coef, nth in ((n_set, ld, 300, n_theta), (1, 8, 8, 3)):
        coefs = _coefs(rng, n_set_, ld_, n_coef)
        theta = np.linspace(0.0, 180.0, nth)
        rc_a, pf_a = _call(alloc, n_set_, n_coef, ld_, coefs, theta, 1)
        rc_s, pf_s = _call(static, n_set_, n_coef, ld_, coefs, theta, 1)
        assert rc_a == 0 and rc_s == 0, product.lmie_gpu_last_error()
        assert np.array_equal(pf_s, pf_a)
    coefs = _coefs(rng, n_set, ld, 300)
    theta = np.linspace(0.0, 180.0, n_theta + 1)
    rc, _ = _call(static, n_set, 300, ld, coefs, theta, 1)
    assert rc == -1 and b"static scratch" in product.lmie_gpu_last_error()


def test_static_staging_refuses_oversize_calls(product):
    """Sizes beyond the array fail with -1 and an error naming the limit,
    before any device call (runs without a GPU)."""
    _, static = _fns(product)
    one = np.zeros(1)
    for n_set, ld, n_theta in ((1, CAP, 1), (CAP + 1, 1, 1), (1, 1, CAP + 1), (100, 2000, 2000),
                               (2 ** 31 - 1, 2 ** 31 - 1, 2 ** 31 - 1)):
        rc = static(n_set, 1, ld, one.ctypes.data_as(DP), n_theta, one.ctypes.data_as(DP),
                    one.ctypes.data_as(DP), 1, 0)
        assert rc == -1
        assert b"static scratch" in product.lmie_gpu_last_error()
