"""Coefficients of log_pos's series (eob_fastmath.hpp): Chebyshev economization
of R(z)/z = sum_k 2 z^(k-1)/(2k+1) on [0, a], a = 0.0295 >= (3 - 2 sqrt 2)^2.

The series to z^40 in exact rationals is rewritten in Chebyshev polynomials of
t = 2z/a - 1, cut after T_D, rewritten in powers of z and rounded to double;
the error of z P(z) against R(z) is then measured on a grid in 60-digit
decimal arithmetic.  D = 6 is the degree log_pos uses.

    python eobgpu_amd/csrc/log_series_coefficients.py
"""
from fractions import Fraction as Fr
from decimal import Decimal, getcontext
import math
getcontext().prec = 60
a = Fr(295, 10000)   # 0.0295 > (3 - 2 sqrt 2)^2 = 0.029437
K = 40
f = [Fr(2, 2 * k + 1) for k in range(1, K + 1)]   # f(z) = sum f[i] z^i
# g(t) = f(a (1 + t) / 2): expand powers of (a/2)(1 + t)
from math import comb
g = [Fr(0)] * K
for i, c in enumerate(f):
    s = c * (a / 2) ** i
    for j in range(i + 1):
        g[j] += s * comb(i, j)
# monomial -> Chebyshev: t^n = sum over T_k
def cheb_of_monomials(p):
    n = len(p)
    # build T_k as monomial lists
    T = [[Fr(1)], [Fr(0), Fr(1)]]
    for k in range(2, n):
        nxt = [Fr(0)] * (k + 1)
        for i, c in enumerate(T[k - 1]): nxt[i + 1] += 2 * c
        for i, c in enumerate(T[k - 2]): nxt[i] -= c
        T.append(nxt)
    # solve triangular: p = sum c_k T_k, from the top degree down
    rem = list(p) + [Fr(0)] * 0
    c = [Fr(0)] * n
    for k in range(n - 1, -1, -1):
        lead = T[k][k]
        c[k] = rem[k] / lead
        for i, v in enumerate(T[k]): rem[i] -= c[k] * v
    return c, T
c, T = cheb_of_monomials(g)
for D in (5, 6, 7):
    # truncate to degree D in t, back to monomials in t, then in z
    pt = [Fr(0)] * (D + 1)
    for k in range(D + 1):
        for i, v in enumerate(T[k]): pt[i] += c[k] * v
    # t = 2 z / a - 1
    pz = [Fr(0)] * (D + 1)
    for i, v in enumerate(pt):
        for j in range(i + 1):   # (2z/a - 1)^i
            pz[j] += v * comb(i, j) * (Fr(2) / a) ** j * (-1) ** (i - j)
    coef = [float(x) for x in pz]
    # error of z * P(z) against R(z) on a fine grid (high precision)
    worst = Decimal(0)
    for m in range(0, 2001):
        z = Fr(m, 2000) * Fr(29437, 1000000)
        zd = Decimal(z.numerator) / Decimal(z.denominator)
        Rt = sum(Decimal(2) / (2 * k + 1) * zd ** k for k in range(1, 30))
        P = Decimal(0)
        for cc in reversed(coef): P = P * zd + Decimal(cc)
        worst = max(worst, abs(Rt - zd * P))
    print(f"degree {D} in f (R degree {D + 1}): max |R - zP| = {float(worst):.3e}; coefficients (z^1..):", [c.hex() for c in coef])
