"""Coefficients of the series in twmc_f64math.h (kSeries), fitted at 50
significant digits with mpmath's Chebyshev approximation over the ranges the
functions reduce to, and printed as the doubles the header holds:

  ps(z) = (x - sin x) / x^3,       z = x^2 in [0, (pi/4)^2]:  sin x = x - x z ps(z)
  pc(z) = (cos x - 1) / z,          z = x^2 in [0, (pi/4)^2]:  cos x = 1 + z pc(z)
  pl(w) = (ln((1+f)/(1-f)) - 2f) / f^3, w = f^2 in [0, ((sqrt2-1)/(sqrt2+1))^2]:
                                    ln m = 2 f + f w pl(w),  f = (m-1)/(m+1)

Seven coefficients each keep the weighted fit error (its contribution to the
relative error of sin, cos and ln) below 2e-17; the accuracy of the compiled
functions is checked by tests/test_f64math.py.  Run: python gen_f64_series.py
"""
import mpmath as mp

mp.mp.dps = 50
Z = (mp.pi / 4) ** 2
W = ((mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)) ** 2


def ps(z):
    if z == 0:
        return mp.mpf(1) / 6
    x = mp.sqrt(z)
    return (x - mp.sin(x)) / x ** 3


def pc(z):
    if z == 0:
        return mp.mpf(-1) / 2
    x = mp.sqrt(z)
    return (mp.cos(x) - 1) / z


def pl(w):
    if w == 0:
        return mp.mpf(2) / 3
    f = mp.sqrt(w)
    return (mp.log((1 + f) / (1 - f)) - 2 * f) / f ** 3


if __name__ == "__main__":
    for name, fn, hi in (("ps", ps, Z), ("pc", pc, Z), ("pl", pl, W)):
        coef = mp.chebyfit(fn, [0, hi], 7)  # highest degree first
        print(name, ", ".join(repr(float(c)) for c in coef))
