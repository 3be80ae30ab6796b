"""The bit-sliced Game of Life arithmetic of gol_bits_kernel (b2r_gol.hip),
restated bit for bit on the CPU: full adders as xor3 / majority, the x-sum of
three 2-bit z-sums and the rule on the carry-save y-sum, checked against the
count and the rule they stand for (gol-3D.cu:93-99: born at n == 5, survives
2 <= n <= 7, with box = n + self) over every input they can receive, and the
integer arithmetic of the byte <-> bit packing (gb_pack's byte dot products,
gb_unpack's nibble spread).  The GPU
parity tests check the kernel itself; these pin the formulation, so that a
wrong truth table or a dropped carry fails here with the input that breaks it."""
import itertools


def xor3(a, b, c):
    return a ^ b ^ c


def maj3(a, b, c):
    return (a & b) | (a & c) | (b & c)


def bit(v, i):
    return (v >> i) & 1


def xsum(l, s, r):
    """gb_xsum: three 2-bit numbers (bit planes 0, 1) -> 4-bit sum."""
    l0, l1, s0, s1, r0, r1 = bit(l, 0), bit(l, 1), bit(s, 0), bit(s, 1), bit(r, 0), bit(r, 1)
    x0b, k1 = xor3(l0, s0, r0), maj3(l0, s0, r0)
    u, v = xor3(l1, s1, r1), maj3(l1, s1, r1)
    x1b, w = u ^ k1, u & k1
    return x0b | (x1b << 1) | ((v ^ w) << 2) | ((v & w) << 3)


def rule(U, X, D, self):
    """gb_rule on one cell: U, X, D are 4-bit x-sums of rows y-1, y, y+1."""
    b0, c1 = xor3(bit(U, 0), bit(X, 0), bit(D, 0)), maj3(bit(U, 0), bit(X, 0), bit(D, 0))
    e1, c2a = xor3(bit(U, 1), bit(X, 1), bit(D, 1)), maj3(bit(U, 1), bit(X, 1), bit(D, 1))
    b1, c2b = e1 ^ c1, e1 & c1
    e2, c3a = xor3(bit(U, 2), bit(X, 2), bit(D, 2)), maj3(bit(U, 2), bit(X, 2), bit(D, 2))
    b2, c3b = xor3(e2, c2a, c2b), maj3(e2, c2a, c2b)
    e3, c4a = xor3(bit(U, 3), bit(X, 3), bit(D, 3)), maj3(bit(U, 3), bit(X, 3), bit(D, 3))
    b3, c4b = xor3(e3, c3a, c3b), maj3(e3, c3a, c3b)
    lo = b2 | b1 | b0
    ge3 = b2 | (b1 & b0)
    eq5 = b2 & (1 - b1) & b0
    m = eq5 | (self & ge3)
    z = b3 & self & (1 - lo)
    tt = z | ((1 - b3) & m)
    return tt & (1 - c4a) & (1 - c4b)


def test_xsum_of_three_zsums():
    for l, s, r in itertools.product(range(4), repeat=3):
        assert xsum(l, s, r) == l + s + r, (l, s, r)


def test_rule_on_every_box_sum():
    # x-sums are 0..9 (three z-sums of 0..3 each)
    for U, X, D, self in itertools.product(range(10), range(10), range(10), (0, 1)):
        box = U + X + D
        n = box - self
        want = int((not self and n == 5) or (self and 2 <= n <= 7))
        assert rule(U, X, D, self) == want, (U, X, D, self)


def test_rule_on_every_neighbourhood_through_the_adders():
    # all 2^27 boxes are too many; every count of each z column, through the
    # x-sum, covers every reachable (U, X, D) with its multiplicity structure
    zs = range(4)
    for zl, zc, zr in itertools.product(zs, repeat=3):
        X = xsum(zl, zc, zr)
        for U, D in itertools.product(range(10), repeat=2):
            for self in (0, 1):
                if self and zc == 0:
                    continue  # the cell itself lies in the middle z column
                box = U + X + D
                want = int(box == 5 or (self and 3 <= box <= 8))
                assert rule(U, X, D, self) == want


def test_unpack_nibble_spread():
    # gb_spread4: nibble * 0x204081 & 0x01010101 puts nibble bit i in byte i
    for n in range(16):
        v = (n * 0x204081) & 0x01010101
        assert v == sum(((n >> i) & 1) << (8 * i) for i in range(4)), n


def test_pack_byte_dot_products():
    # gb_pack's dot products: bytes (bit 0 each) of an even and an odd dword
    # against weights 1, 2, 4, 8 and 16, 32, 64, 128 give cells 0..7 as bits
    def udot4(a, b, c):
        return c + sum(((a >> 8 * i) & 255) * ((b >> 8 * i) & 255) for i in range(4))
    m = 0x01010101
    for bits in range(256):
        cells = [(bits >> i) & 1 for i in range(8)]
        # odd byte values: bit 0 is the cell, the mask drops the rest
        e = sum((c | 6) << (8 * i) for i, c in enumerate(cells[:4]))
        o = sum((c | 2) << (8 * i) for i, c in enumerate(cells[4:]))
        got = udot4(o & m, 0x80402010, udot4(e & m, 0x08040201, 0))
        assert got == bits


def gb_pack(a, b):
    """gb_pack on 8 dwords (a: cells 0..15, b: 16..31): byte dot products of
    even/odd dword pairs, then the two byte permutes and the shift-or."""
    def udot4(x, w, c):
        return c + sum(((x >> 8 * i) & 255) * ((w >> 8 * i) & 255) for i in range(4))
    m = 0x01010101

    def two(e, o):
        return udot4(o & m, 0x80402010, udot4(e & m, 0x08040201, 0))
    n = [two(a[0], a[1]), two(a[2], a[3]), two(b[0], b[1]), two(b[2], b[3])]
    lo = (n[0] & 255) | ((n[1] & 255) << 8)
    hi = (n[2] & 255) | ((n[3] & 255) << 8)
    return (lo | (hi << 16)) & 0xFFFFFFFF


def test_pack_mask_on_the_word_equals_mask_on_the_bytes():
    # the kernel masks the packed word (bits 0-15 from the first 16 bytes,
    # 16-31 from the second) instead of the 8 input dwords
    import random
    rng = random.Random(7)
    for _ in range(2000):
        a = [rng.getrandbits(32) for _ in range(4)]
        b = [rng.getrandbits(32) for _ in range(4)]
        for ma, mb in itertools.product((0, 0xFFFFFFFF), repeat=2):
            want = gb_pack([x & ma for x in a], [x & mb for x in b])
            got = gb_pack(a, b) & ((ma & 0xFFFF) | (mb & 0xFFFF0000))
            assert got == want
    # and on cells: bit i of the word is cell i
    cells = [rng.getrandbits(1) for _ in range(32)]
    dw = [sum(cells[4 * k + j] << (8 * j) for j in range(4)) for k in range(8)]
    assert gb_pack(dw[:4], dw[4:]) == sum(c << i for i, c in enumerate(cells))
