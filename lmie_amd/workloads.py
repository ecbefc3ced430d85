"""The five size-integrated Mie workloads, as lmie_in_data parameter sets.

Each workload is a dict of lmie_in_data fields (names as in
include/lmie_interface.h; ``lambda`` for the wavelength) plus, where the
workload differentiates, the seed vectors of its linearised inputs.  The
two-mode workload is a list of log-normal components sharing one size grid,
mixed by number weights; the reference interface cannot express it in one
call (its bimodal type is r^-4 weighted, misc/rtutil_size_dist.c:547-561), so
it is evaluated as the exact linear combination of its components.

Sizes quoted beside each workload follow the reference's order formula
n1 = floor(x + 4.05 x^(1/3) + 8) (src/lmie_core.c:18-26), n_qang = 2 n1.
"""
import math

import numpy as np

LN = math.log


def lognormal_sigma_seed(sigma_g):
    """d a2 / d sigma_g for a2 = (ln sigma_g)^2, so a seed on a2 differentiates
    with respect to sigma_g (doc/lmie.tex: log-normal a2 is the variance of ln r)."""
    return 2.0 * LN(sigma_g) / sigma_g


# Plain FP64, no derivatives: x_max 22.8, n1 42, 1000 size points.
CFG0 = dict(dist_type="log_normal", n_int2=10, n_quad=100, calc_gc=1,
            **{"lambda": 0.55}, mr=1.53, mi=0.008, a1=0.1, a2=LN(1.6) ** 2,
            r1=0.001, r2=2.0, accuracy=-1.0)

# FP64 with forward tangents wrt (r_g, sigma_g, m_r, m_i): x_max 228.5, n1 261, 4000 points.
CFG1 = dict(dist_type="log_normal", n_int2=40, n_quad=100, calc_gc=1,
            **{"lambda": 0.55}, mr=1.45, mi=0.001, a1=0.5, a2=LN(1.8) ** 2,
            r1=0.005, r2=20.0, accuracy=-1.0)
CFG1_SEEDS = {"a1": [1, 0, 0, 0], "a2": [0, lognormal_sigma_seed(1.8), 0, 0],
              "mr": [0, 0, 1, 0], "mi": [0, 0, 0, 1]}

# Mixed precision plus adjoint, gamma distribution: x_max 913.9, n1 961, 16384 points.
CFG2 = dict(dist_type="gamma", n_int2=256, n_quad=64, calc_gc=1,
            **{"lambda": 0.55}, mr=1.50, mi=0.01, a1=1.0, a2=0.1,
            r1=0.01, r2=80.0, accuracy=-1.0)
CFG2_SEEDS = {"a1": [1, 0, 0, 0], "a2": [0, 1, 0, 0], "mr": [0, 0, 1, 0], "mi": [0, 0, 0, 1]}


def cfg3_cases(n_cases=32, seed=1, r1=0.001, r2=30.0):
    """32 retrieval cases drawn from numpy PCG64(seed): wavelength U[0.34, 2.1],
    m_r U[1.33, 1.60], log10 m_i U[-4, -1], r_g U[0.05, 2], sigma_g U[1.3, 2.2];
    r in [r1, r2] with r1 = 0.001 um (the lower bound is a choice of this suite)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lam = rng.uniform(0.34, 2.1, n_cases)
    mr = rng.uniform(1.33, 1.60, n_cases)
    mi = 10.0 ** rng.uniform(-4.0, -1.0, n_cases)
    rg = rng.uniform(0.05, 2.0, n_cases)
    sg = rng.uniform(1.3, 2.2, n_cases)
    cases = []
    for k in range(n_cases):
        p = dict(dist_type="log_normal", n_int2=40, n_quad=100, calc_gc=1,
                 **{"lambda": float(lam[k])}, mr=float(mr[k]), mi=float(mi[k]),
                 a1=float(rg[k]), a2=LN(float(sg[k])) ** 2, r1=r1, r2=r2, accuracy=-1.0)
        seeds = {"a1": [1, 0, 0, 0], "a2": [0, lognormal_sigma_seed(float(sg[k])), 0, 0],
                 "mr": [0, 0, 1, 0], "mi": [0, 0, 0, 1]}
        cases.append((p, seeds))
    return cases


# Two log-normal modes on one grid, number ratio 1000:1: x_max 1848.0, n1 1905, 32768 points.
CFG4_GRID = dict(n_int2=512, n_quad=64, calc_gc=1, **{"lambda": 0.34}, mr=1.53, mi=0.005,
                 r1=0.001, r2=100.0, accuracy=-1.0)
CFG4_MODES = [dict(a1=0.08, a2=LN(1.5) ** 2, weight=1000.0 / 1001.0, sigma_g=1.5),
              dict(a1=2.5, a2=LN(2.0) ** 2, weight=1.0 / 1001.0, sigma_g=2.0)]


def cfg4_mode_params(k):
    """lmie_in_data parameters of mode k of the two-mode workload."""
    m = CFG4_MODES[k]
    p = dict(CFG4_GRID)
    p.update(dist_type="log_normal", a1=m["a1"], a2=m["a2"])
    return p


def cfg4_mode_seeds(k):
    """Seeds (r_g, sigma_g, m_r, m_i) of mode k."""
    return {"a1": [1, 0, 0, 0], "a2": [0, lognormal_sigma_seed(CFG4_MODES[k]["sigma_g"]), 0, 0],
            "mr": [0, 0, 1, 0], "mi": [0, 0, 0, 1]}


# The example program's scenario (examples/example_c.c:45-98): 4096 points,
# 181 output angles, accuracy 1e-7, seeds mr, mi, a1, a2.
EXAMPLE = dict(dist_type="log_normal", n_int2=64, n_quad=64, n_angles=181, calc_gc=1,
               calc_pf=1, **{"lambda": 2.25}, mr=1.53, mi=0.0054, a1=0.39, a2=0.4805,
               r1=0.005, r2=50.0, accuracy=1e-7)
EXAMPLE_SEEDS = {"mr": [1, 0, 0, 0], "mi": [0, 1, 0, 0], "a1": [0, 0, 1, 0], "a2": [0, 0, 0, 1]}


def calc_n1(x):
    """n1 = floor(x + 4.05 x^(1/3) + 8), src/lmie_core.c:18-26."""
    return int(x + 4.05 * x ** (1.0 / 3.0) + 8)


def expansion_loss_targets(c_ref, seed=0):
    """Targets and per-set weights of the expansion-coefficient loss.

    L = sum_k w_k sum_l (c_kl - t_kl)^2 with t_kl = c_ref_kl (1 + z_kl),
    z ~ N(0,1) from numpy PCG64(seed), and w_k = 1 / sum_l c_ref_kl^2, so each
    of the six coefficient sets weighs alike and the target deviation of every
    coefficient is N(0,1) times its magnitude.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    z = rng.standard_normal(c_ref.shape)
    t = c_ref * (1.0 + z)
    nrm = np.sum(c_ref * c_ref, axis=-1, keepdims=True)
    w = np.where(nrm > 0, 1.0 / np.where(nrm > 0, nrm, 1.0), 0.0)
    return t, np.broadcast_to(w, c_ref.shape).copy()
