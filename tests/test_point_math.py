"""Select-form point helpers of the kernels (kah_internal.h) equal the reference formulas
they replace: minmod (src/Flux/Slope.jl:20-24) on special values and 2e6 random operands;
div_by (reciprocal, product, one FMA correction) equal to IEEE division on 2e7 random
quotients. Host build of the shared header with g++; no GPU."""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def test_minmod_select_form_equals_formula(tmp_path):
    exe = tmp_path / "point_math_check"
    src = HERE / "native" / "point_math_check.cpp"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
