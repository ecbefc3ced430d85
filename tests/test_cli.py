"""The command line program lmie_amd/bin/calllmie against the reference's own
calllmie (utils/calllmie.c, built from its sources by oracle/ref.mk into
oracle/_ref/calllmie_ref) on the same command lines: the same exit status, and
the same output files (header, gc, lc and phase-function files and their
derivative files).

Every file must have the same lines, with every printed number equal or,
where two double-precision solutions round the last printed digit differently
(the smallest coefficients carry cancellation noise at 1e-7 of their own size
in both), within one unit of that digit; numbers that vanish mathematically
and print as rounding noise may differ by 1e-10 of the file's largest number.
The count of lines that are not byte-equal is printed per file.

On the CPU the program is built against the oracle (the CPU restatement of the
solution) and this file's writers; on the GPU (marked gpu) the built program
runs against the product library."""
import os
import re
import subprocess

import pytest

from conftest import ROOT

REF = os.path.join(ROOT, "oracle", "_ref", "calllmie_ref")
OURS = os.path.join(ROOT, "lmie_amd", "bin", "calllmie")

EXAMPLE = "-dist log_normal 0.39 0.4805 0.005 50 -lambda 2.25 -m 1.53 0.0054 -n_int2 64 -n_quad 64 -n_angles 181"
# command lines: (name, options, output files as option -> (value file, derivative files))
CASES = [
    ("example_derivs", EXAMPLE + " -accuracy 1e-7 -derivs 4 -m_l 1,0,0,0 0,1,0,0 -dist_l 0,0,1,0 0,0,0,1 0,0,0,0 0,0,0,0"
     " -check_g 1e-5 -check_cbak 1e-5", ("total", "gc", "lc", "pf")),
    ("gamma_save_dist", "-dist gamma 0.5 0.1 0.01 8 -lambda 0.63 -m 1.45 0.001 -n_int2 20 -n_quad 32 -n_angles 91"
     " -save_dist -derivs 2 -lambda_l 1,0 -dist_l 0,1 0,0 0,0 0,0", ("total", "gc", "pf")),
    ("modified_power_law", "-dist modified_power_law 3.5 0.1 2.0 -lambda 0.55 -m 1.5 0.01 -n_int1 4 -n_int2 8"
     " -n_quad 16 -n_angles 37", ("total", "gc", "lc")),
    ("power_law", "-dist power_law 0.5 0.1 -lambda 0.55 -m 1.33 0.0 -n_int2 8 -n_quad 16 -derivs 1 -dist_l 1 0",
     ("total", "gc")),
    ("mono", "-dist mono 0.7 -lambda 0.55 -m 1.53 0.008 -n_int2 1 -n_quad 1 -n_angles 19", ("total", "gc", "pf")),
    ("exponential", "-dist exponential 2.0 0.01 3.0 -lambda 0.55 -m 1.53 0.008 -n_int2 10 -n_quad 16",
     ("total", "lc")),
    ("modified_gamma", "-dist modified_gamma 0.5 2.0 1.0 0.01 3.0 -lambda 0.55 -m 1.53 0.008 -n_int2 10 -n_quad 16",
     ("total", "gc")),
    ("modified_bimodal", "-dist modified_bimodal_log_normal 0.1 0.2 0.8 0.3 0.5 0.01 4.0 -lambda 0.55 -m 1.53 0.008"
     " -n_int2 16 -n_quad 16", ("total", "gc")),
]

# argument errors (checked before any solution): the reference's exit status
ERRORS = [
    "-lambda 0.5 -m 1.5 0.01 -n_int2 4 -n_quad 8",                          # no -dist
    "-dist log_normal 0.3 0.2 0.01 2 -m 1.5 0.01 -n_int2 4 -n_quad 8",      # no -lambda
    "-dist log_normal 0.3 0.2 0.01 2 -lambda 0.5 -n_int2 4 -n_quad 8",      # no -m
    "-dist log_normal 0.3 0.2 0.01 2 -lambda 0.5 -m 1.5 0.01 -n_quad 8",    # no -n_int2
    "-dist cubic 1 2 -lambda 0.5 -m 1.5 0.01 -n_int2 4 -n_quad 8",          # unknown kind
    "-dist log_normal 0.3 0.2 -lambda",                                     # too few values
    "-dist log_normal 0.3 0.2 0.01 2 -lambda 0.5 -m 1.5 0.01 -n_int2 4 -n_quad 8 -derivs 2 -m_l 1 0,1",  # list count
    "-dist log_normal 0.3 0.2 0.01 2 -lambda 0.5 -m 1.5 0.01 -n_int2 4 -n_quad 8 -check_g 1e-5",          # no -pf
    "-dist log_normal 0.3 0.2 0.01 2 -lambda 0.5 -m 1.5 0.01 -n_int2 4 -n_quad 8 -bogus",                 # option
    "-dist log_normal 0.3 0.2 0.01 2 -lambda 0.5 -m 1.5 0.01 -n_int2 4 -n_quad 8 extra",                  # argument
    "-dist modified_power_law 3.5 0.1 2.0 -lambda 0.5 -m 1.5 0.01 -n_int2 4 -n_quad 8",                   # no -n_int1
]


def _need_ref():
    if not os.path.exists(REF):
        pytest.skip("reference calllmie (oracle/_ref/calllmie_ref) not built")


def _files(tmp, tag, name, outs, n_derivs):
    """Command-line file options and the list of files they name."""
    args, files = [], []
    for o in outs:
        v = os.path.join(tmp, "%s_%s_%s.txt" % (tag, name, o))
        files.append(v)
        args += ["-" + o, v]
        if n_derivs:
            dl = [os.path.join(tmp, "%s_%s_%s_d%d.txt" % (tag, name, o, d)) for d in range(n_derivs)]
            files += dl
            args.append(",".join(dl))
    return args, files


def _n_derivs(opts):
    m = re.search(r"-derivs (\d+)", opts)
    return int(m.group(1)) if m else 0


NUM = re.compile(r"[-+]?\d\.\d{6}e[-+]\d+")


def _same_within_last_digit(a, b):
    """Lines equal, or equal up to numbers that differ by at most one unit of
    their last printed digit, or by at most 1e-10 of the largest number of the
    file (values that vanish mathematically, printed as rounding noise, such as
    the tangent of a distribution statistic along the wavelength).  Returns
    (ok, number of lines not byte-equal)."""
    la, lb = a.splitlines(), b.splitlines()
    if len(la) != len(lb):
        return False, -1
    scale = max([abs(float(v)) for v in NUM.findall(a)] + [0.0])
    differing = 0
    for x, y in zip(la, lb):
        if x == y:
            continue
        differing += 1
        nx, ny = NUM.findall(x), NUM.findall(y)
        if len(nx) != len(ny) or NUM.sub("#", x) != NUM.sub("#", y):
            return False, differing
        for p, q in zip(nx, ny):
            if p == q:
                continue
            fp, fq = float(p), float(q)
            unit = 10.0 ** (int(p.split("e")[1]) - 6)
            if abs(fp - fq) > max(1.01 * unit, 1e-10 * scale):
                return False, differing
    return True, differing


def _compare(exe, tmp_path, name, opts, outs):
    nd = _n_derivs(opts)
    ra, rf = _files(str(tmp_path), "ref", name, outs, nd)
    oa, of = _files(str(tmp_path), "ours", name, outs, nd)
    r = subprocess.run([REF] + opts.split() + ra, capture_output=True, text=True, timeout=600)
    o = subprocess.run([exe] + opts.split() + oa, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    assert o.returncode == 0, o.stderr
    report = []
    for fr, fo in zip(rf, of):
        a, b = open(fr).read(), open(fo).read()
        ok, differing = _same_within_last_digit(a, b)
        assert ok, (name, fr, differing)
        report.append((os.path.basename(fo), differing, len(a.splitlines())))
    return report


def test_cli_argument_errors_match_the_reference():
    """Each argument error ends both programs with status 1 before any
    solution (no GPU needed)."""
    _need_ref()
    if not os.path.exists(OURS):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "lmie_amd", "csrc")])
    for opts in ERRORS:
        r = subprocess.run([REF] + opts.split(), capture_output=True, text=True, timeout=60)
        o = subprocess.run([OURS] + opts.split(), capture_output=True, text=True, timeout=60)
        assert r.returncode == 1, (opts, r.stdout, r.stderr)
        assert o.returncode == 1, (opts, o.stdout, o.stderr)
        assert ("ERROR" in o.stderr) == ("ERROR" in r.stderr), (opts, o.stderr, r.stderr)


@pytest.fixture(scope="module")
def oracle_cli(tmp_path_factory):
    """The program built against the oracle and the library's writers."""
    exe = str(tmp_path_factory.mktemp("cli") / "calllmie_oracle")
    subprocess.check_call(["gcc", "-O2", "-std=gnu99", "-Wall", "-I", os.path.join(ROOT, "include"),
                           "-o", exe, os.path.join(ROOT, "lmie_amd", "csrc", "calllmie.c"),
                           os.path.join(ROOT, "lmie_amd", "csrc", "lmie_scat_io.c"),
                           os.path.join(ROOT, "oracle", "lmie_oracle.c"), "-lm", "-lpthread"])
    return exe


@pytest.mark.parametrize("name,opts,outs", CASES, ids=[c[0] for c in CASES])
def test_cli_files_match_the_reference_on_cpu(oracle_cli, tmp_path, name, opts, outs):
    _need_ref()
    for f, differing, total in _compare(oracle_cli, tmp_path, name, opts, outs):
        print("%s: %d of %d lines differ in the last digit" % (f, differing, total))


@pytest.mark.gpu
@pytest.mark.parametrize("name,opts,outs", CASES, ids=[c[0] for c in CASES])
def test_cli_files_match_the_reference_on_gpu(tmp_path, name, opts, outs):
    _need_ref()
    print("case: calllmie", name, flush=True)
    for f, differing, total in _compare(OURS, tmp_path, name, opts, outs):
        print("  %s: %d of %d lines differ in the last digit" % (f, differing, total), flush=True)
