"""Exact angular tables used by the kernel generator.

Everything here is computed in exact rational arithmetic (``fractions.Fraction``) with at most one
square-root factor per row, and only rounded to float64 when the generator prints a literal.  That
makes structural zeros exact: e.g. the Laplacian terms of a solid harmonic's Hermite expansion cancel
identically and are never emitted.

Conventions (fixed by the reference, restated here):
* Cartesian components of a shell with angular momentum l are ordered x-major descending:
  (l,0,0), (l-1,1,0), (l-1,0,1), ... (reference indexing.hpp:98-112).
* Real solid harmonics are ordered S_l, ..., S_1, C_0, ..., C_l and built from the generating
  formulas of the reference's generator (gen/cartesian_spherical_transformation.py:8-49); p shells are
  left as x, y, z (reference cartesian_spherical_transformation.hpp:38-39).
* Each Cartesian component is normalised to unit self-overlap relative to the pure-x component,
  N = sqrt(ix! iy! iz! (2l)! / ((2ix)! (2iy)! (2iz)! l!)) (reference kernel_cartesian_normalization.hpp:22-26);
  the spherical transform is applied to normalised Cartesians with the normalisation divided back out
  (gen/cartesian_spherical_transformation.py:53-67), so the net spherical function is the raw solid
  harmonic polynomial times the contracted radial part.
"""

from fractions import Fraction
from functools import lru_cache
from math import comb, factorial, isqrt

import mpmath


def cart_list(l):
    """Cartesian exponent triples of shell l in x-major descending order."""
    out = []
    for x in range(l, -1, -1):
        for y in range(l - x, -1, -1):
            out.append((x, y, l - x - y))
    return out


def ncart(l):
    return (l + 1) * (l + 2) // 2


def herm_list(L):
    """Hermite index triples (t, u, v) with t + u + v <= L, ordered by total then x-major."""
    out = []
    for n in range(L + 1):
        out.extend(cart_list(n))
    return out


class SqrtRational:
    """A value sqrt(r) * q with r, q rational; used for table entries."""

    __slots__ = ("r", "q")

    def __init__(self, r, q):
        self.r = Fraction(r)
        self.q = Fraction(q)

    def is_zero(self):
        return self.q == 0 or self.r == 0

    def to_float(self):
        if self.is_zero():
            return 0.0
        with mpmath.workdps(50):
            v = mpmath.sqrt(mpmath.mpf(self.r.numerator) / self.r.denominator) * \
                (mpmath.mpf(self.q.numerator) / self.q.denominator)
            return float(v)


def _sqrt_of_fraction_if_exact(fr):
    n, d = fr.numerator, fr.denominator
    sn, sd = isqrt(n), isqrt(d)
    if sn * sn == n and sd * sd == d:
        return Fraction(sn, sd)
    return None


@lru_cache(maxsize=None)
def cart_norm(l, c):
    """N for Cartesian component index c of shell l, as SqrtRational (value sqrt(r))."""
    x, y, z = cart_list(l)[c]
    r = Fraction(factorial(x) * factorial(y) * factorial(z) * factorial(2 * l),
                 factorial(2 * x) * factorial(2 * y) * factorial(2 * z) * factorial(l))
    return SqrtRational(r, 1)


@lru_cache(maxsize=None)
def solid_harmonic_rows(l):
    """Raw solid-harmonic coefficients: list over rows m of (r_m, {cart_index: Fraction}).

    Row value for monomial c is sqrt(r_m) * coef[c].  For l <= 1 the identity is returned.
    """
    if l <= 1:
        return [(Fraction(1), {c: Fraction(1)}) for c in range(ncart(l))]
    index = {xyz: i for i, xyz in enumerate(cart_list(l))}
    rows = []
    for row in range(2 * l + 1):
        sine = row < l
        m = l - row if sine else row - l
        poly = {}
        for k in range((l - m) // 2 + 1):
            gamma = Fraction((-1) ** k * comb(l, k) * comb(2 * l - 2 * k, l) * factorial(l - 2 * k),
                             factorial(l - 2 * k - m) * 2 ** l)
            zpow = l - 2 * k - m
            for i in range(k + 1):
                for j in range(k - i + 1):
                    h = k - i - j
                    multinom = Fraction(factorial(k), factorial(i) * factorial(j) * factorial(h))
                    for p in range(m + 1):
                        q = m - p
                        if sine:
                            trig = {1: 1, 3: -1}.get(q % 4, 0)
                        else:
                            trig = {0: 1, 2: -1}.get(q % 4, 0)
                        if trig == 0:
                            continue
                        key = (2 * i + p, 2 * j + q, 2 * h + zpow)
                        poly[key] = poly.get(key, 0) + gamma * multinom * comb(m, p) * trig
        r = Fraction((2 if (sine or m > 0) else 1) * factorial(l - m), factorial(l + m))
        coef = {index[key]: v for key, v in poly.items() if v != 0}
        rows.append((r, coef))
    return rows


def out_basis_rows(l, spherical):
    """Rows of the output basis of shell l in terms of raw Cartesian monomials.

    Spherical: the raw solid harmonics.  Cartesian: the normalised components N_c * monomial_c.
    Each row is (r, {c: q}) with value sqrt(r) * q.
    """
    if spherical:
        return solid_harmonic_rows(l)
    rows = []
    for c in range(ncart(l)):
        n = cart_norm(l, c)
        rows.append((n.r, {c: Fraction(1)}))
    return rows


def n_out(l, spherical):
    return 2 * l + 1 if spherical else ncart(l)


@lru_cache(maxsize=None)
def single_centre_e(i, t):
    """Rational part e(i, t) of the single-centre coefficient E^{i,0}_t = e(i,t) (1/2q)^((i+t)/2).

    e(i,t) = i! / (((i-t)/2)! t! 2^((i-t)/2)) for i+t even, else 0 (reference
    mcmurchie_davidson_E_term.hpp:132-156).
    """
    if t < 0 or t > i or (i + t) % 2:
        return Fraction(0)
    h = (i - t) // 2
    return Fraction(factorial(i), factorial(h) * factorial(t) * 2 ** h)


@lru_cache(maxsize=None)
def aux_transform(lc, spherical):
    """Hermite-to-output transform of a single-centre (auxiliary) shell.

    Returns a list over output rows m of (r_m, {s_triple: (Fraction coef, j)}) meaning
      out_m = sum_s sqrt(r_m) * coef * (-1)^|s| * (1/2q)^lc * (2q)^j * R_{t+s}
    with j = (lc - |s|) / 2 an integer >= 0.  The common factor (-1)^lc (1/2q)^lc is applied by the
    caller, so the per-term sign is (-1)^(|s| - lc) = (-1)^(2j) = +1 and is dropped.
    For spherical output only |s| = lc (j = 0) survives: Hobson's theorem, exact here.
    """
    carts = cart_list(lc)
    rows = []
    for r, coef in out_basis_rows(lc, spherical):
        terms = {}
        for c, qc in coef.items():
            cx, cy, cz = carts[c]
            for sx in range(cx % 2, cx + 1, 2):
                for sy in range(cy % 2, cy + 1, 2):
                    for sz in range(cz % 2, cz + 1, 2):
                        e = single_centre_e(cx, sx) * single_centre_e(cy, sy) * single_centre_e(cz, sz)
                        j = (lc - (sx + sy + sz)) // 2
                        key = (sx, sy, sz)
                        prev = terms.get(key, (Fraction(0), j))[0]
                        terms[key] = (prev + qc * e, j)
        terms = {k: v for k, v in terms.items() if v[0] != 0}
        rows.append((r, terms))
    return rows


@lru_cache(maxsize=None)
def pair_out_rows(la, lb, spherical):
    """Output rows of a shell pair: list over (mi, mj) of (r, {(ci, cj): Fraction}).

    value(mi, mj) = sqrt(r) * sum_{ci,cj} q * monomial_i(ci) monomial_j(cj)
    """
    ra = out_basis_rows(la, spherical)
    rb = out_basis_rows(lb, spherical)
    rows = []
    for r1, c1 in ra:
        for r2, c2 in rb:
            terms = {}
            for ci, qi in c1.items():
                for cj, qj in c2.items():
                    terms[(ci, cj)] = qi * qj
            rows.append((r1 * r2, terms))
    return rows


@lru_cache(maxsize=None)
def pair_support(la, lb, spherical):
    """Structural Hermite support of each pair output row: list over (mi,mj) of sorted t triples."""
    ca, cb = cart_list(la), cart_list(lb)
    out = []
    for _, terms in pair_out_rows(la, lb, spherical):
        sup = set()
        for (ci, cj) in terms:
            ix, iy, iz = ca[ci]
            jx, jy, jz = cb[cj]
            for tx in range(ix + jx + 1):
                for ty in range(iy + jy + 1):
                    for tz in range(iz + jz + 1):
                        sup.add((tx, ty, tz))
        out.append(sorted(sup, key=lambda t: (sum(t), -t[0], -t[1])))
    return out


def fmt(v):
    """float64 literal with round-trip precision."""
    if v == 0.0:
        return "0.0"
    s = repr(float(v))
    if "e" not in s and "." not in s:
        s += ".0"
    return s
