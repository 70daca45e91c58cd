"""CPU: the Peano-Hilbert state table the FP32 walk's wavefront order is built from.

`c_hilbert` in csg_device.h maps (state, octant digit) to (next state, position of
that child in Hilbert order).  Walking a 16-level path of octant digits through it
must give the index Skilling's transform ("Programming the Hilbert curve", 2004)
gives for the coordinates whose bits those digits are (digit bit 2 = first axis).
The transform here is written from the paper's AxesToTranspose steps, independent
of the table.  Two properties that single out a Hilbert order among all orders of
the children are checked on the table alone: every state's children form a
permutation, and children consecutive in Hilbert order are face neighbours (their
digits differ in one bit), also across the boundary of two sibling cells.
"""
import os
import re

import numpy as np

HEADER = os.path.join(os.path.dirname(__file__), "..", "cosmosim_amd", "csrc", "csg_device.h")


def table():
    src = open(HEADER).read()
    body = re.search(r"c_hilbert\[24 \* 8\] = \{(.*?)\};", src, re.S).group(1)
    body = re.sub(r"//[^\n]*", "", body)
    vals = [int(v) for v in body.replace("\n", " ").split(",") if v.strip()]
    assert len(vals) == 192
    return np.array(vals, dtype=np.int64)


def skilling_index(X, bits):
    """Hilbert index of the 3 coordinates X (ints of `bits` bits), Skilling 2004:
    inverse undo, Gray encode, then interleave (X[0] the most significant)."""
    X = list(X)
    M = 1 << (bits - 1)
    Q = M
    while Q > 1:
        P = Q - 1
        for i in range(3):
            if X[i] & Q:
                X[0] ^= P
            else:
                t = (X[0] ^ X[i]) & P
                X[0] ^= t
                X[i] ^= t
        Q >>= 1
    for i in range(1, 3):
        X[i] ^= X[i - 1]
    t = 0
    Q = M
    while Q > 1:
        if X[2] & Q:
            t ^= Q - 1
        Q >>= 1
    for i in range(3):
        X[i] ^= t
    h = 0
    for k in range(bits - 1, -1, -1):
        for i in range(3):
            h = (h << 1) | ((X[i] >> k) & 1)
    return h


def walk(T, digits):
    st, h = 0, 0
    for m in digits:
        e = int(T[8 * st + m])
        h = (h << 3) | (e & 7)
        st = e >> 3
    return h


def test_table_matches_skilling_transform():
    T = table()
    rng = np.random.default_rng(11)
    for _ in range(4000):
        digits = rng.integers(0, 8, 16)
        X = [0, 0, 0]
        for m in digits:
            for a in range(3):
                X[a] = (X[a] << 1) | ((int(m) >> (2 - a)) & 1)
        assert walk(T, digits) == skilling_index(X, 16), digits


def test_table_is_a_face_neighbour_order():
    T = table()
    for st in range(24):
        row = T[8 * st: 8 * st + 8]
        assert sorted(row & 7) == list(range(8))
        assert all(0 <= (e >> 3) < 24 for e in row)
        by_h = [int(np.nonzero((row & 7) == h)[0][0]) for h in range(8)]
        for a, b in zip(by_h, by_h[1:]):
            assert bin(a ^ b).count("1") == 1, (st, by_h)
    # across cells: the last grandchild of child h and the first of child h + 1 are
    # face neighbours too (two levels of coordinates)
    for st in range(24):
        row = T[8 * st: 8 * st + 8]
        by_h = [int(np.nonzero((row & 7) == h)[0][0]) for h in range(8)]
        for a, b in zip(by_h, by_h[1:]):
            sa, sb = row[a] >> 3, row[b] >> 3
            last = [m for m in range(8) if (T[8 * sa + m] & 7) == 7][0]
            first = [m for m in range(8) if (T[8 * sb + m] & 7) == 0][0]
            pa = [((a >> (2 - i)) & 1) * 2 + ((last >> (2 - i)) & 1) for i in range(3)]
            pb = [((b >> (2 - i)) & 1) * 2 + ((first >> (2 - i)) & 1) for i in range(3)]
            assert sum(abs(x - y) for x, y in zip(pa, pb)) == 1, (st, a, b)
