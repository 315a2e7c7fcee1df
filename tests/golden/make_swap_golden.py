#!/usr/bin/env python3
"""Write tests/golden/swap_hashes.json from the reference's swap-matrix files.

Reads taichi_with_jerks/multipole_rotation_matrix/rotate_order_{3..30}.txt and
the inline orders 0-2 of taichi_with_jerks/kernel_rotate_two_arrays.h:19-32 as
data (text parsing only; nothing of the reference is compiled or run), turns
each order into its integer form (entry * 2^(n-1)), and stores per order the
number of non-zeros and the SHA-256 of the canonical 'row col value' lines.
The test suite compares these hashes with the tables derived independently by
taichi_collisional_amd/csrc/gen_swap_tables.py.

Only needs /root/reference at generation time; the JSON is the fixture.
"""
import hashlib
import json
import os
import re

REF = "/root/reference/taichi_with_jerks/multipole_rotation_matrix"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "swap_hashes.json")


def parse_order(n):
    text = open(os.path.join(REF, "rotate_order_%d.txt" % n)).read()
    m = re.search(r"coeff\s*=\s*1\s*/\s*([0-9.]+)", text)
    denom = int(float(m.group(1)))
    assert denom == 1 << (n - 1), (n, denom)
    entries = {}
    for blk in re.finditer(r"arr\[(\d+)\]\s*=\s*\((.*?)\)\s*\*\s*coeff", text, re.S):
        row = int(blk.group(1))
        for t in re.finditer(r"([+-]?)\s*(-?\d+)\s*\*\s*arrold\[(\d+)\]", blk.group(2)):
            v = int(t.group(2)) * (-1 if t.group(1) == "-" else 1)
            col = int(t.group(3))
            if v != 0:
                entries[(row, col)] = entries.get((row, col), 0) + v
    return entries


def low_orders():
    # orders 0..2 as written inline (kernel_rotate_two_arrays.h:19-32), integer form
    o0 = {(0, 0): 1}
    o1 = {(1, 1): 1, (2, 3): 1, (3, 2): 1}
    # order 2: coeff = 1/2 -> integer form is the numerator
    o2 = {(4, 5): 4, (5, 4): 1, (6, 6): -1, (6, 8): 1, (7, 7): 2, (8, 6): 3, (8, 8): 1}
    return [o0, o1, o2]


def digest(entries):
    lines = ["%d %d %d" % (r, c, v) for (r, c), v in sorted(entries.items())]
    return hashlib.sha256("\n".join(lines).encode()).hexdigest()


def main():
    orders = low_orders() + [parse_order(n) for n in range(3, 31)]
    out = {"source": "taichi_with_jerks/multipole_rotation_matrix/rotate_order_{3..30}.txt "
                     "+ kernel_rotate_two_arrays.h:19-32 (orders 0-2)",
           "form": "entry * 2^(n-1) as integer; lines 'row col value' sorted, global indices",
           "orders": [{"n": n, "nnz": len(e), "sha256": digest(e)} for n, e in enumerate(orders)]}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
