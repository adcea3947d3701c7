"""Writes tests/golden/fixtures.json: golden vectors that pin the CPU oracle.

Sources (no Julia is available, so nothing here runs the reference):
  * values hard-coded in the reference's own tests and docs
      README 4x4            test/runtests.jl:10-27
      issue #49 (Pf=-119000) test/runtests.jl:423
      analytic 4x4 Pf       test/runtests.jl:426-430
      issue #121 (Pf=1)     test/runtests.jl:433
      issue #116 ev lists   test/runtests.jl:490,499
      issue #118 ev lists   test/runtests.jl:521
      issue #130 matrices   test/runtests.jl:512,516 (sparse I,J,V data)
      logabspfaffian(A)     docs/src/pfaffian.md:55-57
  * LAPACK gehrd via scipy.linalg.hessenberg -- the oracle the reference's hessenberg test
    uses (test/runtests.jl:152-164); generic Gaussian inputs, where the conventions agree
  * exact Pfaffians of integer matrices in rational arithmetic (fractions.Fraction)
  * constructed matrices Q*blockdiag([0 s;-s 0])*Q^T with known log|Pf| and sign

Run:  python tests/golden/gen_golden.py
"""
import json
import os
from fractions import Fraction

import numpy as np
import scipy.linalg as sl

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")
SEED = 314159


def pf_exact(M):
    """Exact Pfaffian by skew Gaussian elimination with row/column pivoting (rationals).

    Pf(A) = a_01 * Pf(A'), A'_ij = A_ij + (A_i0 A_1j - A_i1 A_0j) / a_01 on the trailing block;
    a symmetric swap of rows/columns 1 and p flips the sign.
    """
    A = [[Fraction(int(x)) for x in row] for row in M]
    n = len(A)
    if n % 2 == 1:
        return Fraction(0)
    pf = Fraction(1)
    while n > 0:
        p = next((j for j in range(1, n) if A[0][j] != 0), None)
        if p is None:
            return Fraction(0)
        if p != 1:
            A[1], A[p] = A[p], A[1]
            for row in A:
                row[1], row[p] = row[p], row[1]
            pf = -pf
        a01 = A[0][1]
        pf *= a01
        B = [[A[i][j] + (A[i][0] * A[1][j] - A[i][1] * A[0][j]) / a01 for j in range(2, n)]
             for i in range(2, n)]
        A = B
        n -= 2
    return pf


def pf_expand(M):
    """Pfaffian by cofactor expansion along the first row (brute force, tiny n only)."""
    n = len(M)
    if n == 0:
        return Fraction(1)
    if n % 2 == 1:
        return Fraction(0)
    total = Fraction(0)
    for j in range(1, n):
        if M[0][j] == 0:
            continue
        keep = [k for k in range(n) if k not in (0, j)]
        sub = [[M[a][b] for b in keep] for a in keep]
        total += (-1) ** (j + 1) * Fraction(int(M[0][j])) * pf_expand(sub)
    return total


class GQ:
    """Exact Gaussian rational re + im*i (fractions.Fraction parts)."""

    def __init__(self, re, im=0):
        self.re, self.im = Fraction(re), Fraction(im)

    def __add__(self, o):
        return GQ(self.re + o.re, self.im + o.im)

    def __sub__(self, o):
        return GQ(self.re - o.re, self.im - o.im)

    def __neg__(self):
        return GQ(-self.re, -self.im)

    def __mul__(self, o):
        return GQ(self.re * o.re - self.im * o.im, self.re * o.im + self.im * o.re)

    def __truediv__(self, o):
        d = o.re * o.re + o.im * o.im
        return GQ((self.re * o.re + self.im * o.im) / d, (self.im * o.re - self.re * o.im) / d)

    def iszero(self):
        return self.re == 0 and self.im == 0


def pf_exact_complex(Are, Aim):
    """Exact Pfaffian of a complex skew-symmetric matrix given as exact real and imaginary parts
    (skew Gaussian elimination in Gaussian rationals; the same recursion as pf_exact)."""
    A = [[GQ(Fraction(str(a)), Fraction(str(b))) for a, b in zip(ra, rb)] for ra, rb in zip(Are, Aim)]
    n = len(A)
    if n % 2 == 1:
        return GQ(0)
    pf = GQ(1)
    while n > 0:
        p = next((j for j in range(1, n) if not A[0][j].iszero()), None)
        if p is None:
            return GQ(0)
        if p != 1:
            A[1], A[p] = A[p], A[1]
            for row in A:
                row[1], row[p] = row[p], row[1]
            pf = -pf
        a01 = A[0][1]
        pf = pf * a01
        A = [[A[i][j] + (A[i][0] * A[1][j] - A[i][1] * A[0][j]) / a01 for j in range(2, n)] for i in range(2, n)]
        n -= 2
    return pf


def haar(n, rng):
    Z = rng.standard_normal((n, n))
    Q, R = np.linalg.qr(Z)
    return Q * np.sign(np.diag(R))


def main():
    rng = np.random.default_rng(SEED)
    fx = {}

    # README 4x4 (test/runtests.jl:10-27, docs/src/pfaffian.md:55-57)
    fx["readme"] = {
        "A": [[0, 2, -7, 4], [-2, 0, -8, 3], [7, 8, 0, 1], [-4, -3, -1, 0]],
        "eta": [8.306623862918073, 8.53382018538718, -1.083472677771923],
        "Q": [[1.0, 0.0, 0.0, 0.0],
              [0.0, -0.2407717061715382, -0.9592700375676934, -0.14774972261267352],
              [0.0, 0.8427009716003843, -0.2821382463434394, 0.4585336219014009],
              [0.0, -0.48154341234307674, -0.014106912317171805, 0.8763086996337883]],
        "eig_pos": [11.93445871397423, 0.7541188264752862],
        "det": 81.0,
        "inv_times_9": [[0, 1, -3, -8], [-1, 0, 4, 7], [3, -4, 0, 2], [8, -7, -2, 0]],
        "logabspfaffian": [2.1972245773362196, -1.0],
        "pfaffian_exact": -9,
    }

    # issue #49 (test/runtests.jl:423)
    fx["issue49"] = {
        "A": [[0, 14, 7, -10, 0, 10, 0, -11], [-14, 0, -10, 7, 13, -9, -12, -13],
              [-7, 10, 0, -4, 6, -17, -1, 18], [10, -7, 4, 0, -2, -4, 0, 11],
              [0, -13, -6, 2, 0, -8, -18, 17], [-10, 9, 17, 4, 8, 0, -8, 12],
              [0, 12, 1, 0, 18, 8, 0, 0], [11, 13, -18, -11, -17, -12, 0, 0]],
        "pfaffian": -119000,
    }
    assert pf_exact(fx["issue49"]["A"]) == -119000
    assert pf_expand(fx["issue49"]["A"]) == -119000

    # issue #121 (test/runtests.jl:433): exercises the tau=2 reflector of an already-zero column
    fx["issue121"] = {"A": [[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1], [0, 0, -1, 0]],
                      "pfaffian": 1}

    # analytic 4x4 Pf = a f - b e + c d (test/runtests.jl:426-430)
    cases = []
    for _ in range(10):
        a, b, c, d, e, f = (int(x) for x in rng.integers(-10, 11, size=6))
        A = [[0, a, b, c], [-a, 0, d, e], [-b, -d, 0, f], [-c, -e, -f, 0]]
        pf = c * d - b * e + a * f
        assert pf_exact(A) == pf == pf_expand(A)
        cases.append({"A": A, "pfaffian": pf})
    fx["analytic4"] = cases

    # integer skew matrices with exact Pfaffians (test/runtests.jl:390-397 style: entries in
    # 2*(-10:10), then the skew part, so integers)
    cases = []
    for n in [1, 2, 3, 4, 6, 8, 10, 11, 12, 16]:
        R = rng.integers(-10, 11, size=(n, n)) * 2
        A = (R - R.T) // 2
        pf = pf_exact(A.tolist())
        if n <= 8:
            assert pf == pf_expand(A.tolist())
        det = int(round(np.linalg.det(A))) if n <= 12 else None
        cases.append({"A": A.tolist(), "pfaffian": str(pf), "det": det})
    fx["integer_pf"] = cases

    # issue #116 (test/runtests.jl:490,499) and issue #118 (test/runtests.jl:521)
    fx["issue116_ev"] = [[0, 0, 0], [1, 2, 3, 2, 1], [0, 1, 0], [1, 0, 0], [1, 0, 1], [1, 1, 0],
                         [0, 1, 1], [0, 0, 1], [1, 1, 1], [1, 1, 0, 1, 1], [0, 0, 0, 0, 1, 1, 1],
                         [0, 1, 0, 1, 0, 1, 1, 1, 0, 0, 0, 1], [1, 1, 0, 1],
                         [0, 1, 0, 1, 1, 1, 1, 1]]
    fx["issue118_ev"] = [[1.0, 0.001, 1.0, 0.0001, 1.0], [2.0, 1e-11, 2.0, 1e-11, 2.0]]

    # issue #130 matrices (test/runtests.jl:512,516), sparse triplets, 1-based
    fx["issue130"] = _issue130()

    # LAPACK gehrd fixtures on generic Gaussian inputs
    cases = []
    for n in [2, 3, 4, 5, 10, 11, 12, 13, 31, 32, 33, 64]:
        B = rng.standard_normal((n, n))
        A = B - B.T
        H, Q = sl.hessenberg(A, calc_q=True)
        cases.append({"A": A.tolist(), "eta": np.diag(H, -1).tolist(), "Q": Q.tolist()})
    fx["gehrd"] = cases

    # constructed: A = Q blockdiag([0 s; -s 0]) Q^T, log|Pf| = sum(log s), sign = det(Q)
    cases = []
    for n in [2, 4, 8, 16, 64]:
        Q = haar(n, rng)
        s = np.exp(rng.uniform(np.log(0.1), np.log(10.0), size=n // 2))
        Bm = np.zeros((n, n))
        for k in range(n // 2):
            Bm[2 * k, 2 * k + 1] = s[k]
            Bm[2 * k + 1, 2 * k] = -s[k]
        A = Q @ Bm @ Q.T
        A = 0.5 * (A - A.T)
        sgn = float(np.sign(np.linalg.det(Q)))
        Ainv = Q @ np.linalg.inv(Bm) @ Q.T
        cases.append({"A": A.tolist(), "logpf": float(np.sum(np.log(s))), "sign": sgn,
                      "inv": (0.5 * (Ainv - Ainv.T)).tolist()})
    fx["constructed"] = cases

    # complex skew-symmetric matrices as in the reference's Pfaffian tests
    # (test/runtests.jl:407-420: A = (R - R^T)/2 with R of Gaussian integers in [-10, 10]), plus
    # cases that force a pivot swap (A[1,0] = 0) or have a zero Pfaffian (a zero row); exact
    # Pfaffians in Gaussian rationals
    crng = np.random.default_rng(SEED + 7)
    cases = []
    for n in [1, 2, 3, 4, 6, 10, 20]:
        R = crng.integers(-10, 11, (n, n)) + 1j * crng.integers(-10, 11, (n, n))
        cases.append((R - R.T) / 2)
    Sw = cases[4].copy()            # n = 6 with A[1,0] = 0: the first step swaps rows/columns
    Sw[1, 0] = Sw[0, 1] = 0
    cases.append(Sw)
    Sw2 = cases[5].copy()           # n = 10 with A[1:3, 0] = 0
    Sw2[1:3, 0] = 0
    Sw2[0, 1:3] = 0
    cases.append(Sw2)
    Z = cases[5].copy()             # zero row and column 3: Pf = 0
    Z[3, :] = 0
    Z[:, 3] = 0
    cases.append(Z)
    out = []
    for A in cases:
        pf = pf_exact_complex(A.real.tolist(), A.imag.tolist())
        out.append({"re": A.real.tolist(), "im": A.imag.tolist(),
                    "pf": [str(pf.re), str(pf.im)], "pf_float": [float(pf.re), float(pf.im)]})
    assert tuple(out[3]["pf"]) == _pf4(cases[3]), "4x4 exact Pfaffian disagrees with the analytic formula"
    fx["complex_pf"] = out

    with open(OUT, "w") as f:
        json.dump(fx, f, indent=None, separators=(",", ":"))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


def _pf4(A):
    """a f - b e + c d for A = [0 a b c; -a 0 d e; -b -d 0 f; -c -e -f 0] in exact Gaussian rationals
    (test/runtests.jl:426-430 states it for integers)."""
    g = lambda z: GQ(Fraction(str(z.real)), Fraction(str(z.imag)))
    a, b, c, d, e, f = g(A[0, 1]), g(A[0, 2]), g(A[0, 3]), g(A[1, 2]), g(A[1, 3]), g(A[2, 3])
    r = a * f - b * e + c * d
    return str(r.re), str(r.im)


def _issue130():
    I10 = [2, 7, 1, 3, 6, 8, 2, 4, 7, 9, 3, 5, 8, 10, 4, 9, 2, 7, 1, 3, 6, 8, 2, 4, 7, 9, 3, 5, 8, 10,
           4, 9]
    J10 = [1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 6, 6, 7, 7, 7, 7, 8, 8, 8, 8, 9, 9, 9, 9,
           10, 10]
    V10 = [-0.8414709848, 1.5403023059, 0.8414709848, -0.8414709848, 0.4596976941, 1.5403023059,
           0.8414709848, -0.8414709848, 0.4596976941, 1.5403023059, 0.8414709848, -0.8414709848,
           0.4596976941, 1.5403023059, 0.8414709848, 0.4596976941, -0.4596976941, 0.8414709848,
           -1.5403023059, -0.4596976941, -0.8414709848, 0.8414709848, -1.5403023059,
           -0.4596976941, -0.8414709848, 0.8414709848, -1.5403023059, -0.4596976941,
           -0.8414709848, 0.8414709848, -1.5403023059, -0.8414709848]
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "issue130_66.json")
    with open(src) as f:
        m66 = json.load(f)
    return {"m10": {"n": 10, "I": I10, "J": J10, "V": V10}, "m66": m66}


if __name__ == "__main__":
    main()
