"""The fill_invisible kernel divides each weighted sum by its weight in FP64:
q = s r, e = fma(-q, w, s), q' = fma(e, r, q) with r = 1 / w rounded, then
rounded to float.  This checks, on the CPU with the same IEEE operations,
that the result equals IEEE float division for every weight 1 .. 13 over the
subnormal, near-one and largest exponents of both signs (JXG_FULL_DIVISION=1:
every finite float, about a minute on 8 cores)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_fp64_reciprocal_division_is_ieee_float_division(tmp_path):
    exe = tmp_path / "fill_division_check"
    subprocess.check_call(["gcc", "-O2", "-march=native", "-fopenmp", "-o", str(exe),
                           os.path.join(HERE, "native", "fill_division_check.c"), "-lm"])
    args = [str(exe)] + (["full"] if os.environ.get("JXG_FULL_DIVISION") == "1" else [])
    out = subprocess.run(args, capture_output=True, text=True, timeout=900)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.split()[-1] == "0", out.stdout
