"""Generate the golden fixtures of tests/golden/ from the reference CPU library.

Runs only where the reference sources are present: it builds them with
oracle/ref.mk into oracle/_ref/liblmie_ref.so (a checker, git-ignored) and
calls lmie_solution through the ABI mirror lmie_amd/abi.py.  Every fixture is
data only: the inputs (in the workload definitions and stored beside the
outputs) and the reference's outputs at full double precision.

    python tests/golden/gen_golden.py [names...]

Fixtures written (numpy .npz, allow_pickle never needed):
  example.npz     examples/example_c.c scenario, 4 seeds, gc + pf at 181 angles
  small_*.npz     tiny grids with save_control=1 (per-size save1/save2), one per
                  size-distribution kind, with Mie and weight-only tangents
  resaved.npz     save_control=2: the small log-normal grid saved with one
                  (r_g, sigma_g), re-weighted with another, weight-only seeds
  cfg0.npz        1000 points, x_max 22.8, no derivatives
  cfg1.npz        4000 points, x_max 228.5, 4 seeds
  cfg2.npz        16384 points, x_max 913.9, 4 seeds -> expansion-loss gradient
  cfg3.npz        32 retrieval cases -> scalars, gc of 8 cases, loss gradients
  cfg4.npz        two log-normal modes, 32768 points, x_max 1848 -> mixed gc and
                  expansion-loss gradient wrt (r_g1, sigma_g1, r_g2, sigma_g2, m_r, m_i)
"""
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from lmie_amd import abi, workloads as W  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
NT = int(os.environ.get("GOLDEN_THREADS", os.cpu_count() or 1))


def ref_lib():
    subprocess.check_call(["make", "-s", "-f", "oracle/ref.mk", "-j8"], cwd=ROOT)
    return abi.bind(C.CDLL(os.path.join(ROOT, "oracle", "_ref", "liblmie_ref.so")))


def params_json(p, seeds=None):
    return json.dumps({"params": p, "seeds": seeds or {}})


def save(name, p, seeds, r, **extra):
    arrs = {k: np.asarray(v) for k, v in r.items() if v is not None}
    arrs.update({k: np.asarray(v) for k, v in extra.items()})
    arrs["inputs_json"] = np.asarray(params_json(p, seeds))
    path = os.path.join(OUT, name + ".npz")
    np.savez_compressed(path, **arrs)
    print("wrote %s (%.1f KB)" % (path, os.path.getsize(path) / 1024.0))


def expansion_loss_grad(gc, gc_l, seed=0):
    """Loss value and its gradient along each seed, by contracting the
    reference's forward tangents with dL/dgc (loss of workloads.expansion_loss_targets)."""
    t, w = W.expansion_loss_targets(gc, seed)
    res = gc - t
    L = float(np.sum(w * res * res))
    dL = 2.0 * w * res
    grad = np.einsum("kl,dkl->d", dL, gc_l)
    return L, grad


def gen_example(lib):
    r = abi.solve(lib, W.EXAMPLE, W.EXAMPLE_SEEDS, NT)
    save("example", W.EXAMPLE, W.EXAMPLE_SEEDS, r)


SMALL = {
    # per-size fixtures: 2 x 8 = 16 points, save_control = 1
    "small_lognormal": (dict(dist_type="log_normal", n_int2=2, n_quad=8, calc_gc=1, calc_lc=1,
                             calc_pf=1, n_angles=37, **{"lambda": 0.6}, mr=1.5, mi=0.02,
                             a1=0.3, a2=0.25, r1=0.05, r2=3.0, accuracy=-1.0, save_control=1),
                        {"mr": [1, 0, 0, 0, 0], "mi": [0, 1, 0, 0, 0], "a1": [0, 0, 1, 0, 0],
                         "a2": [0, 0, 0, 1, 0], "lambda": [0, 0, 0, 0, 1]}),
    "small_gamma": (dict(dist_type="gamma", n_int2=2, n_quad=8, calc_gc=1, calc_pf=1, n_angles=19,
                         **{"lambda": 0.55}, mr=1.33, mi=1e-3, a1=0.8, a2=0.2, r1=0.02,
                         r2=4.0, accuracy=-1.0, save_control=1),
                    {"a1": [1, 0, 0, 0], "a2": [0, 1, 0, 0], "mr": [0, 0, 1, 0], "r2": [0, 0, 0, 1]}),
    "small_modgamma": (dict(dist_type="modified_gamma", n_int2=3, n_quad=6, calc_gc=1,
                            **{"lambda": 0.8}, mr=1.44, mi=0.0, a1=2.0, a2=3.0, a3=1.0,
                            r1=0.01, r2=2.0, accuracy=1e-7),
                       {"a1": [1, 0, 0], "a2": [0, 1, 0], "a3": [0, 0, 1]}),
    "small_powerlaw": (dict(dist_type="power_law", n_int2=3, n_quad=6, calc_gc=1,
                            **{"lambda": 0.65}, mr=1.5, mi=0.01, a1=0.4, a2=0.3,
                            r1=0.0, r2=0.0, accuracy=1e-9),
                       {"a1": [1, 0], "a2": [0, 1]}),
    "small_modpowerlaw": (dict(dist_type="modified_power_law", n_int1=2, n_int2=3, n_quad=6,
                               calc_gc=1, **{"lambda": 0.5}, mr=1.38, mi=2e-3, a1=-3.0,
                               r1=0.1, r2=1.2, accuracy=-1.0),
                          {"a1": [1, 0], "r1": [0, 1]}),
    "small_bimodal": (dict(dist_type="modified_bimodal_log_normal", n_int2=3, n_quad=6, calc_gc=1,
                           **{"lambda": 0.55}, mr=1.52, mi=0.008, a1=0.1, a2=0.3, a3=0.6,
                           a4=0.4, a5=0.05, r1=0.02, r2=2.5, accuracy=-1.0),
                      {"a1": [1, 0, 0], "a4": [0, 1, 0], "a5": [0, 0, 1]}),
    "small_exponential": (dict(dist_type="exponential", n_int2=3, n_quad=6, calc_gc=1,
                               **{"lambda": 0.7}, mr=1.31, mi=1e-4, a1=1.5, r1=0.01,
                               r2=1.5, accuracy=-1.0), None),
    "small_mono": (dict(dist_type="mono", n_int2=1, n_quad=1, calc_gc=1, calc_lc=1, calc_pf=1,
                        n_angles=13, **{"lambda": 0.5}, mr=1.6, mi=0.05, a1=1.1, r1=1.1,
                        r2=1.1, accuracy=-1.0),
                   {"mr": [1, 0], "mi": [0, 1]}),
}


def gen_small(lib):
    for name, (p, seeds) in SMALL.items():
        r = abi.solve(lib, p, seeds, 1)
        save(name, p, seeds, r)


# save_control = 2: saved with RESAVED_P1, re-weighted with RESAVED_P2
RESAVED_P1 = dict(SMALL["small_lognormal"][0], save_control=1)
RESAVED_P2 = dict(RESAVED_P1, a1=0.45, a2=0.18)
RESAVED_SEEDS = {"a1": [1, 0], "a2": [0, 1]}


def gen_resaved(lib):
    r = abi.solve_resaved(lib, RESAVED_P1, RESAVED_P2, RESAVED_SEEDS, 1)
    save("resaved", {"p1": RESAVED_P1, "p2": RESAVED_P2}, RESAVED_SEEDS, r)


def gen_cfg0(lib):
    r = abi.solve(lib, W.CFG0, None, NT)
    save("cfg0", W.CFG0, None, r)


def gen_cfg1(lib):
    r = abi.solve(lib, W.CFG1, W.CFG1_SEEDS, NT)
    save("cfg1", W.CFG1, W.CFG1_SEEDS, r)


def gen_cfg2(lib):
    r = abi.solve(lib, W.CFG2, W.CFG2_SEEDS, NT)
    L, grad = expansion_loss_grad(r["gc"], r["gc_l"])
    save("cfg2", W.CFG2, W.CFG2_SEEDS, r, loss=L, loss_grad=grad)


def gen_cfg3(lib, n_full=8):
    cases = W.cfg3_cases()
    keep = {}
    scal = {k: [] for k in ("cext", "csca", "cbak", "g", "n_coef", "max_coef")}
    losses, grads = [], []
    for c, (p, seeds) in enumerate(cases):
        r = abi.solve(lib, p, seeds, NT)
        for k in scal:
            scal[k].append(r[k])
        L, grad = expansion_loss_grad(r["gc"], r["gc_l"], seed=100 + c)
        losses.append(L)
        grads.append(grad)
        if c < n_full:
            keep["gc_%d" % c] = r["gc"]
    arrs = {k: np.asarray(v) for k, v in scal.items()}
    arrs.update(keep)
    save("cfg3", {"cases": "workloads.cfg3_cases()"}, None, arrs,
         loss=np.asarray(losses), loss_grad=np.asarray(grads))


def gen_cfg4(lib):
    modes = []
    for k in range(2):
        t0 = time.time()
        r = abi.solve(lib, W.cfg4_mode_params(k), W.cfg4_mode_seeds(k), NT)
        print("  mode %d: %.1f s" % (k, time.time() - t0))
        modes.append(r)
    w = [m["weight"] for m in W.CFG4_MODES]
    csca = sum(w[k] * modes[k]["csca"] for k in range(2))
    cext = sum(w[k] * modes[k]["cext"] for k in range(2))
    cbak = sum(w[k] * modes[k]["cbak"] for k in range(2))
    g = sum(w[k] * modes[k]["csca"] * modes[k]["g"] for k in range(2)) / csca
    gc = sum(w[k] * modes[k]["csca"] * modes[k]["gc"] for k in range(2)) / csca
    # tangents of the mixture along (r_g1, sigma_g1, r_g2, sigma_g2, m_r, m_i)
    gc_l = np.zeros((6,) + gc.shape)
    for k in range(2):
        m = modes[k]
        # mode-k tangent columns: 0 r_g, 1 sigma_g, 2 m_r, 3 m_i
        for d, col in ((0, 2 * k), (1, 2 * k + 1), (2, 4), (3, 5)):
            dnum = w[k] * (m["csca_l"][d] * m["gc"] + m["csca"] * m["gc_l"][d])
            gc_l[col] += (dnum - gc * w[k] * m["csca_l"][d]) / csca
    L, grad = expansion_loss_grad(gc, gc_l)
    arrs = dict(cext=cext, csca=csca, cbak=cbak, g=g, gc=gc, loss=L, loss_grad=grad)
    for k in range(2):
        for q in ("cext", "csca", "cbak", "g", "n_coef", "max_coef"):
            arrs["mode%d_%s" % (k, q)] = modes[k][q]
        for q in ("cext_l", "csca_l", "cbak_l", "g_l"):
            arrs["mode%d_%s" % (k, q)] = modes[k][q]
    save("cfg4", {"grid": W.CFG4_GRID, "modes": W.CFG4_MODES}, None, arrs)


GENS = {"example": gen_example, "small": gen_small, "resaved": gen_resaved, "cfg0": gen_cfg0, "cfg1": gen_cfg1,
        "cfg2": gen_cfg2, "cfg3": gen_cfg3, "cfg4": gen_cfg4}


def main(argv):
    lib = ref_lib()
    names = argv or list(GENS)
    for n in names:
        t0 = time.time()
        GENS[n](lib)
        print("%s: %.1f s" % (n, time.time() - t0))


if __name__ == "__main__":
    main(sys.argv[1:])
