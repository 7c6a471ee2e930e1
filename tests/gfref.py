"""Textbook GF(2^n) arithmetic in Python integers for pinning the oracle.

Carry-less schoolbook product followed by long division by the full
irreducible polynomial (x^n + low part).  This is written from the
mathematics, not from the reference's bit-serial loops, so a dropped term,
wrong shift or wrong polynomial in the oracle shows up as a mismatch.
"""

POLY = {16: (1 << 16) | 0x2D, 32: (1 << 32) | 0x8299, 128: (1 << 128) | 0x87}


def clmul(a: int, b: int) -> int:
    r = 0
    i = 0
    while b >> i:
        if (b >> i) & 1:
            r ^= a << i
        i += 1
    return r


def reduce(v: int, bits: int) -> int:
    p = POLY[bits]
    while v.bit_length() > bits:
        v ^= p << (v.bit_length() - 1 - bits)
    return v


def mul(a: int, b: int, bits: int) -> int:
    return reduce(clmul(a, b), bits)


def subspace_poly(i: int, x: int, bits: int) -> int:
    """W_i(x) = prod over u in span{1, 2, 4, ..., 2^(i-1)} of (x + u)."""
    r = 1
    for u in range(1 << i):
        r = mul(r, x ^ u, bits)
    return r
