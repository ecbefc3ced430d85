"""The host C layer and the oracle under AddressSanitizer, UndefinedBehavior-
Sanitizer and the leak checker, on the CPU.

The product's host layer (lmie_amd/csrc/lmie_host.c, lmie_scat_io.c) is linked with a test
double of the device layer (tests/c/gpu_double.c), so every entry point of the
drop-in interface and of the extensions runs without a GPU; then each device
call is forced to fail from the host, and the -1 return, the error text
(the reference's error chain, src/lmie_interface.c:212-214, 244-246) and the
release of every host allocation are checked.  The oracle is driven through
the reference interface on every size distribution.  The driver is
tests/c/host_checks.c."""
import os
import subprocess

import pytest

from conftest import ROOT

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:allocator_may_return_null=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _build(tmp_path, name, sources, defines=()):
    exe = str(tmp_path / name)
    cmd = ["gcc", "-std=gnu99", "-Wall", "-Wextra", "-Wno-unused-parameter"] + SAN + list(defines) + \
        ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "oracle")] + \
        [os.path.join(ROOT, s) for s in sources] + ["-o", exe, "-lm", "-lpthread"]
    subprocess.check_call(cmd)
    return exe


def _run(exe):
    r = subprocess.run([exe], env=ENV, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "host checks: ok" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]


def test_host_layer_under_sanitizers_with_forced_device_failures(tmp_path):
    exe = _build(tmp_path, "host_checks", ["lmie_amd/csrc/lmie_host.c", "lmie_amd/csrc/lmie_scat_io.c",
                                           "tests/c/gpu_double.c", "tests/c/host_checks.c"], ["-DWITH_EXTENSIONS"])
    _run(exe)


def test_oracle_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "oracle_checks", ["oracle/lmie_oracle.c", "tests/c/host_checks.c"])
    _run(exe)
