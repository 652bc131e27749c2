"""Independent pure-Python restatement of GF(2^128) and GF(2^8) arithmetic for
pinning the oracle: products by shift-and-add on Python integers and
reduction by polynomial long division (a different algorithm from the
oracle's word-wise fold).  Small cases only.
"""

P128 = (1 << 128) | 0x87  # x^128 + x^7 + x^2 + x + 1 (CF/field/gf2_128.rs:10)
P8 = 0x11B  # x^8 + x^4 + x^3 + x + 1 (CF/field/gf2_8.rs:12)


def clmul(a, b):
    r = 0
    while b:
        if b & 1:
            r ^= a
        a <<= 1
        b >>= 1
    return r


def polymod(x, p):
    dp = p.bit_length() - 1
    while x.bit_length() - 1 >= dp:
        x ^= p << (x.bit_length() - 1 - dp)
    return x


def mul(a, b):
    return polymod(clmul(a, b), P128)


def mul8(a, b):
    return polymod(clmul(a, b), P8)


def pow_(a, e):
    r = 1
    while e:
        if e & 1:
            r = mul(r, a)
        a = mul(a, a)
        e >>= 1
    return r


def inv(a):
    return pow_(a, (1 << 128) - 2)


def eq_entry(point, i):
    """prod_j (bit_j(i) ? r_j : 1 + r_j), computed per entry."""
    acc = 1
    for j, r in enumerate(point):
        acc = mul(acc, r if (i >> j) & 1 else r ^ 1)
    return acc
