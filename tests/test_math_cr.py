"""wam_math.h against the correctly rounded result and against glibc's float
functions, over strided sweeps of whole float ranges (tests/math/math_sweep.c,
built here with gcc).  The exhaustive sweep (stride 1) is the same program:

    gcc -O2 -mfma -ffp-contract=off -fopenmp -Iwam_amd/csrc tests/math/math_sweep.c -lm -o sweep && ./sweep 1

The correctly rounded value is the x87 long-double libm result rounded to
float.  Every function must equal it on every swept argument; disagreements
with glibc are counted and only bounded (glibc's tanhf, sinhf, coshf, log10f
and atan2f are not correctly rounded for a few per cent of arguments)."""
import json
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def sweep(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("ms") / "sweep")
    subprocess.check_call(["gcc", "-O2", "-mfma", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                           "-I" + os.path.join(ROOT, "wam_amd", "csrc"), os.path.join(HERE, "math", "math_sweep.c"),
                           "-o", exe, "-lm"])
    out = subprocess.check_output([exe, "1009", "300000"], text=True)
    return json.loads(out)


@pytest.mark.parametrize("fn", ["exp", "log", "log10", "tanh", "cosh", "sinh", "sin", "cos", "pow", "atan2"])
def test_correctly_rounded(sweep, fn):
    r = sweep[fn]
    assert r["n"] > 250000
    assert r["ne_correctly_rounded"] == 0, r
    # every disagreement with glibc is one where glibc is not correctly rounded
    assert r["ne_glibc"] == r["glibc_ne_correctly_rounded"], r


def test_division_by_1p1_is_ieee(tmp_path):
    """wam_div_1p1 (PROPAGS' SIG1 weight, -MIN(0, DOP)/1.1) equals IEEE division
    by 1.1f bit for bit on every float."""
    exe = str(tmp_path / "div_sweep")
    subprocess.check_call(["gcc", "-O2", "-mfma", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                           "-I" + os.path.join(ROOT, "wam_amd", "csrc"), os.path.join(HERE, "math", "div_sweep.c"),
                           "-o", exe])
    r = json.loads(subprocess.check_output([exe, "1"], text=True))
    assert r["n"] == 2 ** 32 and r["ne"] == 0, r
