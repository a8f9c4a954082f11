"""Accuracy of the FP64 Box-Muller the FP64 step kernels use
(pincosimulator_amd/csrc/twmc_f64math.h), built for the host with g++ and
checked against x87 long double references over ~1M arguments spread over the
whole 32-bit range plus the reductions' boundaries (every row switch of the ln
and sin / cos tables): ln within 2 ulp, sin / cos of 2 pi b 2^-32 within 2 ulp
(absolute, in ulps of 1/2, near their zeros), the complex sample within 4 ulp
of its radius.  A dropped series term, a wrong table row or entry, or a wrong
ln 2 split moves these errors by orders of magnitude."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pincosimulator_amd", "csrc")


def test_fp64_box_muller_accuracy(tmp_path):
    exe = str(tmp_path / "f64math_check")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-Wextra", "-Werror", "-I", CSRC,
           os.path.join(ROOT, "tests", "native", "f64math_check.cpp"), "-o", exe, "-lm"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    err = {l.split()[0]: float(l.split()[1]) for l in r.stdout.splitlines()}
    assert set(err) == {"log_u32", "sin_turn32", "cos_turn32", "box_muller"}
