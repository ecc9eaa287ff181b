"""The Joe-Kuo D(6) direction-number table (vitrum_amd/csrc/joe_kuo_d6.inc),
which the product and the oracle both compile, checked against the
mathematics that fixes it rather than against itself: rows are dimensions
2..513 in order; every row's polynomial x^s + a_1 x^(s-1) + ... + a_(s-1) x + 1
is primitive over GF(2) (x has multiplicative order 2^s - 1 modulo it);
within a degree the polynomials are distinct and listed by increasing a, and
every fully used degree lists all phi(2^s - 1) / s primitive polynomials;
each initial direction number m_k is odd and below 2^k (Joe and Kuo 2008,
section 2).  A mistyped a or m in the shared table fails one of these (the
reference's KATs cover the first dimensions only)."""
import math
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "vitrum_amd", "csrc", "joe_kuo_d6.inc")


def rows():
    text = open(INC).read()
    body = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = []
    # rows {dim, s, a, {m_1, ..., m_s, 0 padding}}
    for m in re.finditer(r"\{\s*(\d+)\s*,\s*(\d+)\s*,\s*(\d+)\s*,\s*\{([^{}]*)\}\s*\}", body):
        dim, s, a = int(m.group(1)), int(m.group(2)), int(m.group(3))
        ms = [int(v, 0) for v in m.group(4).split(",") if v.strip()]
        assert not any(ms[s:]), dim  # padding after the s initial numbers
        out.append([dim, s, a] + ms[:s])
    return out


def order_of_x(poly, s):
    """Multiplicative order of x modulo poly (bit i = coefficient of x^i)."""
    r, k = 1, 0
    while True:
        r <<= 1
        if r >> s & 1:
            r ^= poly
        k += 1
        if r == 1 or k > (1 << s):
            return k


def phi(n):
    res, m, p = n, n, 2
    while p * p <= m:
        if m % p == 0:
            while m % p == 0:
                m //= p
            res -= res // p
        p += 1
    if m > 1:
        res -= res // m
    return res


def test_table_is_joe_kuo_by_construction():
    rs = rows()
    assert [r[0] for r in rs] == list(range(2, 514))
    by_degree = {}
    for dim, s, a, *m in rs:
        assert len(m) == s, dim
        poly = (1 << s) | (a << 1) | 1
        assert order_of_x(poly, s) == (1 << s) - 1, f"dimension {dim}: polynomial not primitive"
        for k, mk in enumerate(m, start=1):
            assert mk % 2 == 1 and mk < (1 << k), (dim, k, mk)
        by_degree.setdefault(s, []).append(a)
    degrees = [r[1] for r in rs]
    assert degrees == sorted(degrees)
    last = max(by_degree)
    for s, alist in by_degree.items():
        assert alist == sorted(set(alist)), s
        if s < last:
            assert len(alist) == phi((1 << s) - 1) // s, s
