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


def _table_rows(name):
    """The doubles of one table macro of twmc_f64tab.h, in order."""
    import re
    text = open(os.path.join(CSRC, "twmc_f64tab.h")).read()
    body = text.split(f"#define {name} \\\n", 1)[1].split("\n\n", 1)[0]
    return [float(v) for v in re.findall(r"[-+]?[0-9][0-9.e+-]*", body.replace("\\", " "))]


def test_tables_match_their_definition():
    """The committed tables against their definition at 40 digits: ln rows
    (stored as 2 c_j; c_j with at most 19 significant bits, |m c_j - 1| <= 2^-7 over the row's
    interval, L_j = -ln c_j - [j >= 53] ln 2 within 0.01 of its ulp, exact
    rows c = 1, 1/2 at the ends), and sin / cos rows the doubles nearest
    cos, sin (2 pi j / 256)."""
    import math

    import mpmath as mp

    mp.mp.dps = 40
    log_rows = _table_rows("TWMC_F64_LOG_ROWS")
    sc_rows = _table_rows("TWMC_F64_SINCOS_ROWS")
    assert len(log_rows) == 256 and len(sc_rows) == 512
    for j in range(128):
        c, L = log_rows[2 * j] / 2, log_rows[2 * j + 1]  # stored as 2 c_j
        m_, _ = math.frexp(c)
        assert m_ * 2 ** 19 == int(m_ * 2 ** 19), j
        lo, hi = 1 + mp.mpf(j) / 128, 1 + mp.mpf(j + 1) / 128
        assert max(abs(c * lo - 1), abs(c * hi - 1)) <= mp.mpf(2) ** -7, j
        want = -mp.log(c) - (mp.log(2) if j >= 53 else 0)
        if L == 0.0:
            assert want == 0, j
        else:
            assert abs(mp.mpf(L) - want) <= 0.01 * math.ulp(abs(L)), j
    assert log_rows[:2] == [2.0, 0.0] and log_rows[-2:] == [1.0, 0.0]
    for j in range(256):
        t = mp.mpf(j) / 128  # the angle in units of pi
        assert sc_rows[2 * j] == float(mp.cospi(t)) and sc_rows[2 * j + 1] == float(mp.sinpi(t)), j
    assert sc_rows[128:130] == [0.0, 1.0] and sc_rows[256:258] == [-1.0, 0.0] and sc_rows[384:386] == [0.0, -1.0]
