"""CPU checks of the arithmetic identities the gfx950 kernels rely on
(tests/c/kernel_arith.c): the Markstein-corrected quotients of the expansion
tables equal the IEEE division bit for bit; the per-size psi recurrence in
Miller's normalised form equals the reference's ratio chain and product
(src/lmie_core.c:298-330) to rounding; the reciprocal refinement of the D
chain is within an ulp; the D chain's w form equals the reference's
(src/lmie_core.c:386-405) to rounding."""
import os
import subprocess

from conftest import ROOT


def test_kernel_arithmetic_identities(tmp_path):
    exe = str(tmp_path / "kernel_arith")
    subprocess.check_call(["gcc", "-std=gnu99", "-O2", "-ffp-contract=off", "-Wall",
                           os.path.join(ROOT, "tests/c/kernel_arith.c"), "-o", exe, "-lm"])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "kernel arithmetic: ok" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
