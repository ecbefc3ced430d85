"""Benchmark of the size-integrated Mie hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--path mixed-adjoint|fp64-adjoint|mixed|fp64] [--no-cpu-baseline]

Workload: two log-normal modes (r_g 0.08 um / sigma_g 1.5 and r_g 2.5 um /
sigma_g 2.0, number ratio 1000:1) on one grid r in [0.001, 100] um with 512 x 64
= 32768 size points, wavelength 0.34 um, m = 1.53 + 0.005i: x_max 1848,
n1 1905, 3810 Gauss angles, 6 x 3810 expansion coefficients.  Inputs are a few
scalars; the size grid, the Mie coefficients and every table are generated on
the device inside the timed step.

A step of the default path (mixed-adjoint) is one forward solution in mixed
precision that keeps its reverse-pass state, the loss cotangent of the six
expansion-coefficient sets (weighted L2 against seeded targets), and one
reverse pass giving dL/d(m_r, m_i, r_g and sigma_g of both modes).  The plain
FP64 GPU forward of the same workload is timed beside it.  The JSON line reports:
  value        sum_i n1(x_i) over the size points per second (size-pt * n-term / s)
  roofline     the dominant kernel of the step, algorithmic FLOP per launch over
               its device time (HIP events on the library's stream) against the
               MI355X peak of its arithmetic (executed_frac: the matrix-core work
               actually issued, five split products per algorithmic product in the
               forward contraction and three in the reverse one): the forward angular contraction
               (8 n_qang sum_i n1(x_i)) or the reverse contraction U = G A of the
               mixed path (24 sum_i n1(x_i)^2); traffic = HBM bytes per launch
               from the committed counter summary profiles/traffic.json
               (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy), if present
  config.wall_speedup_vs_fp64  whole calls (host preparation and PCIe included)
               of the plain FP64 forward over the default path's whole calls
  config.fp64_angular  the FP64 forward's angular contraction: device time, the
               time its algorithmic FLOP take at the FP64 matrix peak, their ratio,
               and the step speedup against an FP64 forward whose contraction
               ran at that peak
  config.accuracy  the timed run's expansion coefficients against the FP64
               solution of the same workload: per set, normwise and the largest
               elementwise relative error over |c| >= 1e-3 max |c| of the set
  cpu_baseline the reference's pthreads CPU path (oracle/_ref) on the same
               workload (forward only; the reference has no reverse pass), on
               this host's cores (up to 16 threads: one GPU's share of a node),
               with the CPU model and core counts
  config.retrieval_batch  the 32-case retrieval workload in one fused batch call
               against one call per case (wall clock; --no-batch omits it)

    python bench.py --calibrate-copy BYTES   one plain device copy of BYTES
               (calibration point of the traffic counters), then exit

    python bench.py --workload retrieval-batch [--steps K] [--warmup W]
               the 32 seeded retrieval cases (log-normal, 4000 sizes each, x_max
               89-554): mixed forward + reverse of each case's L2 loss, all cases
               in one lmie_solution_adjoint_batch call (one launch per kernel for
               all cases), against the same call with concurrent workers, one call
               per case, and the plain FP64 forward one call per case; wall clock of
               whole calls (host preparation included; the cases' inputs are scalars)
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "size-pts*n-terms/s + adjoint grads/s vs plain FP64 GPU at rel err<=1e-6"
PEAK = {"fp64": 78.6e12, "bf16": 2.5e15}


def grid_n1(p):
    """n1(x_i) over the size grid of p (Gauss nodes per interval)."""
    t, _ = np.polynomial.legendre.leggauss(p["n_quad"])
    h = (p["r2"] - p["r1"]) / p["n_int2"]
    r = (p["r1"] + h * (np.arange(p["n_int2"])[:, None] + 0.5) + 0.5 * h * t[None, :]).ravel()
    x = 2 * math.pi * r / p["lambda"]
    return np.floor(x + 4.05 * x ** (1.0 / 3.0) + 8)


def sum_n1(p):
    """sum_i n1(x_i) over the grid of p."""
    return int(np.sum(grid_n1(p)))


def cpu_model():
    """CPU model name of this host (/proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_threads():
    """Worker threads for the CPU baseline: every core this process may run on,
    capped at 16, the host-core share of one GPU on a shared multi-GPU node."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return max(1, min(16, avail)), avail


def cpu_baseline():
    """The reference build (pthreads) on the whole workload: both modes on the
    full 512 x 64 grid (the two-mode solution is the weighted sum of the two
    single-mode solutions on that grid)."""
    from lmie_amd import abi, workloads as W
    so = os.path.join(ROOT, "oracle", "_ref", "liblmie_ref.so")
    if not os.path.exists(so):
        return None
    threads, avail = host_threads()
    ref = abi.bind(C.CDLL(so))
    tot_units, t0 = 0, time.perf_counter()
    for k in range(2):
        p = W.cfg4_mode_params(k)
        abi.solve(ref, p, None, threads)
        tot_units += sum_n1(p)
    dt = time.perf_counter() - t0
    return {"value": tot_units / dt, "unit": "size-pt*n-term/s", "cores": threads,
            "kind": "reference", "host_cpus": os.cpu_count(), "cpus_available": avail,
            "cpu_model": cpu_model(),
            "sample": "the whole workload: reference lmie_solution (pthreads, %d threads) on both modes "
                      "at 512x64 size points each, 3810 angles, the forward only: %.1f s" % (threads, dt)}


def retrieval_batch(lib, steps, warmup, quick=False):
    """JSON line of the 32-case retrieval workload (--workload retrieval-batch);
    quick: only the fused batch and one call per case (a summary for the main line)."""
    from lmie_amd import lmie, workloads as W
    cases = [p for p, _ in W.cfg3_cases()]
    c_ref = [lmie.solve(lib, p, None, "fp64")["gc"] for p in cases]
    tw = [W.expansion_loss_targets(c_ref[c], 100 + c) for c in range(len(cases))]
    units = sum(sum_n1(p) for p in cases)

    def timed(fn):
        for _ in range(warmup):
            fn()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        return (time.perf_counter() - t0) / steps

    def batch(nw):
        return lambda: lmie.adjoint_batch(lib, cases, [t for t, _ in tw], [w for _, w in tw],
                                          "mixed", n_workers=nw)

    def single():
        for c, p in enumerate(cases):
            t, w = tw[c]
            lmie.adjoint(lib, [p], None, lambda o: {"gc": 2.0 * w * (o["gc"] - t)}, "mixed")

    def fp64():
        for p in cases:
            lmie.solve(lib, p, None, "fp64")

    t_batch = {nw: timed(batch(nw)) for nw in ((0,) if quick else (0, 1, 4, 8))}
    t_single = timed(single)
    if quick:
        return {"workload": "32 retrieval cases, mixed forward + reverse, wall clock of whole calls",
                "fused_batch_ms": t_batch[0] * 1e3, "one_call_per_case_ms": t_single * 1e3,
                "grads_per_s": len(cases) / t_batch[0]}
    t_fp64 = timed(fp64)
    dt = t_batch[0]
    return {
        "metric": METRIC, "value": units / dt, "unit": "size-pt*n-term/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f16-split/f32-acc+f64",
        "data": "synthetic: 32 seeded retrieval cases",
        "config": {"workload": "32 retrieval cases, log-normal, 40x100 sizes each, x_max 89-554, "
                               "mixed forward + reverse of a per-case L2 loss on gc, one batch call, "
                               "one launch per kernel for all cases",
                   "sum_n1": units,
                   "batch_ms_by_workers": {str(k): v * 1e3 for k, v in t_batch.items()},
                   "workers": "0: one launch per kernel for all cases; k > 0: k cases at a time on their own streams",
                   "one_call_per_case_ms": t_single * 1e3,
                   "fp64_forward_one_call_per_case_ms": t_fp64 * 1e3,
                   "grads_per_s": len(cases) / dt,
                   "timing": "wall clock of whole calls, host preparation and transfers included"},
        "roofline": None, "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--path", default="mixed-adjoint",
                    choices=["mixed-adjoint", "fp64-adjoint", "fp64", "mixed", "mixed-bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-compare", action="store_true",
                    help="omit the separate plain-FP64 run (counter passes)")
    ap.add_argument("--no-batch", action="store_true",
                    help="omit the summary of the 32-case retrieval batch")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES")
    ap.add_argument("--workload", default="two-mode", choices=["two-mode", "retrieval-batch"])
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (the size points "
              "would shard across GPUs with one all-reduce of the raw sums); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import lmie_amd
    from lmie_amd import lmie, workloads as W
    lib = lmie_amd.load()
    if a.calibrate_copy:
        ms = C.c_double()
        f = lib.lmie_gpu_selftest_copy
        f.argtypes = [C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_double)]
        if f(a.calibrate_copy, 5, 0, C.byref(ms)) != 0:
            raise SystemExit("copy failed: " + lmie.last_error(lib))
        print(json.dumps({"copy_bytes": a.calibrate_copy, "ms": ms.value,
                          "GB/s": 2 * a.calibrate_copy / ms.value / 1e6}), flush=True)
        return
    if a.workload == "retrieval-batch":
        print(json.dumps(retrieval_batch(lib, a.steps, a.warmup)), flush=True)
        return
    modes = [(W.cfg4_mode_params(k), None) for k in range(2)]
    weights = [m["weight"] for m in W.CFG4_MODES]
    n1s = grid_n1(W.cfg4_mode_params(0))
    units = int(np.sum(n1s))
    nq = 2 * W.calc_n1(2 * math.pi * W.CFG4_GRID["r2"] / W.CFG4_GRID["lambda"])

    # loss targets from one FP64 solution (outside the timed region)
    c_ref = lmie.solve_mix(lib, modes, weights, "fp64")["gc"]
    targets, wts = W.expansion_loss_targets(c_ref, 0)

    def step(path):
        if path.endswith("-adjoint"):
            adj = lmie.Adjoint(lib, [m for m, _ in modes], weights, path.split("-")[0])
            try:
                out = adj.forward()
                t_fwd = lmie.timings(lib)
                grad = adj.reverse({"gc": 2.0 * wts * (out["gc"] - targets)})
                t_rev = lmie.timings(lib)
            finally:
                adj.free()
            tm = dict(t_fwd)
            for k in ("adj_angular", "adj_coefs", "adj_gemm"):
                tm[k] = t_rev[k]
            tm["total"] = t_fwd["total"] + t_rev["total"]
            tm["fwd_call"], tm["rev_call"] = t_fwd["call"], t_rev["call"]
            tm["call"] = t_fwd["call"] + t_rev["call"]
            return tm, grad, out
        out = lmie.solve_mix(lib, modes, weights, path)
        return lmie.timings(lib), None, out

    def run(path, steps, warmup):
        """Mean device time per step (inputs resident in HBM: the library's own
        HIP events around its device pipeline, forward + reverse) and mean wall
        time per step (host preparation and PCIe transfers included)."""
        lib.lmie_set_timing(1)
        for _ in range(warmup):
            step(path)
        t_all = []
        t0 = time.perf_counter()
        for _ in range(steps):
            tm, grad, out = step(path)
            t_all.append(tm)
        wall = (time.perf_counter() - t0) / steps
        lib.lmie_set_timing(0)
        stages = {k: float(np.mean([t[k] for t in t_all])) for k in t_all[0]}
        return stages["total"] * 1e-3, wall, stages["angular"], stages, grad, out

    dt, wall, ang_ms, stages, grad, out_last = run(a.path, a.steps, a.warmup)
    ref_dt = ref_wall = None
    if a.path != "fp64" and not a.no_fp64_compare:
        ref_dt, ref_wall, _, ref_stages, _, _ = run("fp64", max(2, a.steps // 3), 1)

    mixed = a.path.startswith("mixed")
    peak = PEAK["bf16"] if mixed else PEAK["fp64"]
    # dominant kernel: forward contraction, or the reverse contraction U = G A
    kern = {"angular_mix" if mixed else "angular_f64_dma": (8.0 * nq * units, ang_ms)}
    # matrix products issued per algorithmic product: the forward contraction
    # splits the size operand in two parts and the angle operand in three and
    # takes five of their products (three or four with LMIE_MIX_PRODUCTS=3 or
    # 4); the reverse contraction takes three (two parts of each operand)
    issued = {"angular_mix": {"3": 3.0, "4": 4.0}.get(os.environ.get("LMIE_MIX_PRODUCTS", ""), 5.0),
              "adj_ugemm": 3.0, "angular_f64_dma": 1.0}
    if mixed and a.path.endswith("-adjoint") and stages.get("adj_gemm", 0) > 0:
        kern["adj_ugemm"] = (24.0 * float(np.sum(n1s * n1s)), stages["adj_gemm"])
    kname = max(kern, key=lambda k: kern[k][1])
    flop, kms = kern[kname]
    achieved = flop / (kms * 1e-3)
    # the FP64 forward's angular contraction against the FP64 matrix peak
    fp64_angular = None
    if ref_dt is not None:
        f64_ms = ref_stages["angular"]
        f64_roof_ms = 8.0 * nq * units / PEAK["fp64"] * 1e3
        fp64_angular = {"kernel": "angular_f64_dma", "ms": f64_ms, "roofline_ms": f64_roof_ms,
                        "frac": f64_roof_ms / f64_ms,
                        # an FP64 forward whose contraction ran at the FP64 matrix
                        # peak (its other stages as measured), against this step
                        "fp64_step_at_roofline_ms": ref_dt * 1e3 - f64_ms + f64_roof_ms,
                        "speedup_vs_fp64_at_roofline": (ref_dt * 1e3 - f64_ms + f64_roof_ms) / (dt * 1e3)}
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get(kname, {}).get("bytes_per_launch")
    # error of the last timed step's coefficients against the FP64 solution of
    # the same workload (c_ref): per set, normwise and the largest elementwise
    # relative error over |c| >= 1e-3 max |c| of the set
    acc = {}
    for k, name in enumerate(("alpha1", "alpha2", "alpha3", "alpha4", "beta1", "beta2")):
        r = np.asarray(c_ref[k], float)
        c = np.asarray(out_last["gc"][k], float)[:r.size]
        big = np.abs(r) >= 1e-3 * np.abs(r).max()
        acc[name] = {"normwise": float(np.linalg.norm(c - r) / np.linalg.norm(r)),
                     "elem_max": float(np.max(np.abs(c[big] - r[big]) / np.abs(r[big])))}
    accuracy = {"vs": "FP64 GPU solution of the same workload", "sets": acc,
                "normwise_max": max(v["normwise"] for v in acc.values()),
                "elem_max": max(v["elem_max"] for v in acc.values()),
                "elem_max_without_beta1": max(v["elem_max"] for k, v in acc.items() if k != "beta1")}
    line = {
        "metric": METRIC, "value": units / dt, "unit": "size-pt*n-term/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": ("bf16-split/f32-acc+f64" if "bf16" in a.path else "f16-split/f32-acc+f64") if mixed else "f64",
        "data": "synthetic: generated size grid of two log-normal modes",
        "config": {"workload": "two-mode log-normal, 32768 sizes, x_max 1848, 3810 angles, "
                               "gc 6x3810", "path": a.path, "sum_n1": units, "n_qang": nq,
                   "stage_ms": stages,
                   "grads_per_s": (1.0 / dt) if a.path.endswith("-adjoint") else None,
                   "grad": None if grad is None else {k: (float(v) if np.ndim(v) == 0 else np.asarray(v).tolist())
                                                      for k, v in grad.items()},
                   "fp64_ms_per_step": None if ref_dt is None else ref_dt * 1e3,
                   "speedup_vs_fp64": None if ref_dt is None else ref_dt / dt,
                   "wall_ms_per_step": wall * 1e3,
                   "fp64_wall_ms_per_step": None if ref_wall is None else ref_wall * 1e3,
                   "wall_speedup_vs_fp64": None if ref_wall is None else ref_wall / wall,
                   "fp64_angular": fp64_angular,
                   "accuracy": accuracy,
                   "timing": "ms_per_step, value and fp64_ms_per_step: device pipeline per step with "
                             "inputs in HBM (HIP events on the library stream, forward + reverse); "
                             "wall_*: whole calls including host preparation and PCIe transfers"},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                     "kernel": kname, "kernel_ms": kms, "flop_per_launch": flop,
                     "executed_frac": issued[kname] * achieved / peak,
                     "kernels": {k: {"ms": v[1], "TFLOP/s": v[0] / (v[1] * 1e-3) / 1e12,
                                     "executed_frac": issued[k] * v[0] / (v[1] * 1e-3) / peak}
                                 for k, v in kern.items()}},
        "cpu_baseline": None,
    }
    if a.path == "mixed-adjoint" and not a.no_batch and not a.no_fp64_compare:
        line["config"]["retrieval_batch"] = retrieval_batch(lib, 5, 2, quick=True)
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
