"""Coefficients of log_pos's series (eob_fastmath.hpp): Chebyshev economization
of R(z)/z = sum_k 2 z^(k-1)/(2k+1) on [0, a], a = 0.0295 >= (3 - 2 sqrt 2)^2.

The series to z^40 in exact rationals is rewritten in Chebyshev polynomials of
t = 2z/a - 1, cut after T_D, rewritten in powers of z and rounded to double;
the error of z P(z) against R(z) is then measured on a grid in 60-digit
decimal arithmetic.  D = 6 is the degree log_pos uses
(eobgpu_amd/csrc/eob_fastmath.hpp); tests/test_derivations.py checks the
library's coefficients against ``economized(6)``.

    python tests/log_series_coefficients.py
"""
from fractions import Fraction as Fr
from decimal import Decimal, getcontext
from math import comb

getcontext().prec = 60
A = Fr(295, 10000)   # 0.0295 > (3 - 2 sqrt 2)^2 = 0.029437
K = 40


def _chebyshev_of_monomials(p):
    """Chebyshev coefficients c (p = sum c_k T_k) and the T_k as monomial lists."""
    n = len(p)
    T = [[Fr(1)], [Fr(0), Fr(1)]]
    for k in range(2, n):
        nxt = [Fr(0)] * (k + 1)
        for i, c in enumerate(T[k - 1]):
            nxt[i + 1] += 2 * c
        for i, c in enumerate(T[k - 2]):
            nxt[i] -= c
        T.append(nxt)
    rem = list(p)
    c = [Fr(0)] * n
    for k in range(n - 1, -1, -1):
        c[k] = rem[k] / T[k][k]
        for i, v in enumerate(T[k]):
            rem[i] -= c[k] * v
    return c, T


def economized(D):
    """(coefficients of P(z) = R(z)/z of degree D as doubles, lowest power first,
    max |R - z P| on [0, (3 - 2 sqrt 2)^2])."""
    f = [Fr(2, 2 * k + 1) for k in range(1, K + 1)]   # f(z) = R(z)/z = sum f[i] z^i
    g = [Fr(0)] * K                                   # g(t) = f(a (1 + t) / 2)
    for i, c in enumerate(f):
        s = c * (A / 2) ** i
        for j in range(i + 1):
            g[j] += s * comb(i, j)
    c, T = _chebyshev_of_monomials(g)
    pt = [Fr(0)] * (D + 1)
    for k in range(D + 1):
        for i, v in enumerate(T[k]):
            pt[i] += c[k] * v
    pz = [Fr(0)] * (D + 1)                            # t = 2 z / a - 1
    for i, v in enumerate(pt):
        for j in range(i + 1):
            pz[j] += v * comb(i, j) * (Fr(2) / A) ** j * (-1) ** (i - j)
    coef = [float(x) for x in pz]
    worst = Decimal(0)
    for m in range(0, 2001):
        z = Fr(m, 2000) * Fr(29437, 1000000)
        zd = Decimal(z.numerator) / Decimal(z.denominator)
        Rt = sum(Decimal(2) / (2 * k + 1) * zd ** k for k in range(1, 30))
        P = Decimal(0)
        for cc in reversed(coef):
            P = P * zd + Decimal(cc)
        worst = max(worst, abs(Rt - zd * P))
    return coef, float(worst)


if __name__ == "__main__":
    for D in (5, 6, 7):
        coef, err = economized(D)
        print(f"degree {D} in f (R degree {D + 1}): max |R - zP| = {err:.3e}; coefficients (z^1..):", [c.hex() for c in coef])
