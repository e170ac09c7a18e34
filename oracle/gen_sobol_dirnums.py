"""Writes pg_sobol_dirnums.h: the Sobol direction numbers pine's Sobol sampler
uses, in Joe & Kuo's compact form (degree s, primitive-polynomial bits a,
initial direction numbers m_1..m_s) for dimensions 1..1023.

Pine ships them expanded into 32-bit generator matrices
(src/util/sobolmetrices.cpp:5-6662, 1024 dimensions x 52 columns; they are the
Joe-Kuo "new-joe-kuo-6.21201" numbers that pbrt also uses).  Bit parity with
pine's SobolSampler needs exactly these numbers, so this script reads the table
as text where the reference sources lie, recovers (s, a, m) per dimension and
checks that the library's own recurrence (pg_sobol.h, sobol_expand) rebuilds
every column pine's sampler can reach (columns 0..31: pine's Sobol index is a
32-bit value, lowdiscrepancy.h:139-158).  Dimension 0 is the van der Corput
sequence and needs no numbers.

Run once, where /root/reference exists:  python oracle/gen_sobol_dirnums.py
"""
import os
import re
import sys

import numpy as np

REF = "/root/reference/src/util/sobolmetrices.cpp"
OUT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pine_amd", "csrc",
                   "pg_sobol_dirnums.h")
NDIM, NCOL = 1024, 52


def read_table():
    text = open(REF).read()
    start = text.index("SobolMatrices32[")
    body = text[text.index("{", start) + 1:text.index("};", start)]
    vals = [int(x, 16) for x in re.findall(r"0x[0-9a-fA-F]+", body)]
    assert len(vals) == NDIM * NCOL, len(vals)
    return np.array(vals, dtype=np.uint64).reshape(NDIM, NCOL)


def expand(s, a, m, ncol=32):
    """Bratley-Fox recurrence in 32-bit fixed point: column k-1 holds v_k * 2^32;
    v_k = v_{k-s} ^ (v_{k-s} >> s) ^ XOR_j a_j v_{k-j}, a_1 the top bit of a."""
    c = [0] * ncol
    for k in range(1, ncol + 1):
        if k <= s:
            c[k - 1] = (m[k - 1] << (32 - k)) & 0xFFFFFFFF if k <= 32 else m[k - 1] >> (k - 32)
        else:
            v = c[k - s - 1] ^ (c[k - s - 1] >> s)
            for j in range(1, s):
                if (a >> (s - 1 - j)) & 1:
                    v ^= c[k - j - 1]
            c[k - 1] = v
    return c


def solve_gf2(vectors, rhs):
    """Bits x_j with XOR_j x_j vectors[j] == rhs (Python ints as GF(2) vectors), or None."""
    pivots = []  # (pivot bit, vector, combination mask)
    for j, v in enumerate(vectors):
        mask = 1 << j
        for pb, pv, pm in pivots:
            if (v >> pb) & 1:
                v ^= pv
                mask ^= pm
        if v:
            pivots.append((v.bit_length() - 1, v, mask))
    combo = 0
    for pb, pv, pm in pivots:
        if (rhs >> pb) & 1:
            rhs ^= pv
            combo ^= pm
    return None if rhs else combo


def recover(col):
    """Smallest (s, a, m) whose recurrence reproduces columns 0..31.  For a fixed
    degree the recurrence is linear in a_1..a_{s-1}, so a is solved over GF(2)."""
    c = [int(x) for x in col[:32]]
    for s in range(1, 19):
        m = [c[k - 1] >> (32 - k) for k in range(1, s + 1)]
        pack = lambda words: sum(w << (32 * i) for i, w in enumerate(words))
        ks = range(s + 1, 33)
        rhs = pack([c[k - 1] ^ c[k - s - 1] ^ (c[k - s - 1] >> s) for k in ks])
        vecs = [pack([c[k - j - 1] for k in ks]) for j in range(1, s)]
        combo = solve_gf2(vecs, rhs)
        if combo is None:
            continue
        a = 0
        for j in range(1, s):
            if (combo >> (j - 1)) & 1:
                a |= 1 << (s - 1 - j)
        if expand(s, a, m) == c:
            return s, a, m
    raise ValueError("no recurrence reproduces this dimension")


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} not found: the direction numbers are read from pine's table")
    tab = read_table()
    rows = []
    for d in range(1, NDIM):
        s, a, m = recover(tab[d])
        rows.append((s, a, m))
    # check: every dimension rebuilt bit for bit over the 32 reachable columns
    for d, (s, a, m) in enumerate(rows, start=1):
        assert expand(s, a, m) == [int(x) for x in tab[d, :32]], d
    stream = []
    for s, a, m in rows:
        stream += [s, a] + m
    with open(OUT, "w") as f:
        f.write("// Generated by oracle/gen_sobol_dirnums.py: Joe & Kuo (2008) Sobol direction numbers\n")
        f.write("// (new-joe-kuo-6.21201, as pine's src/util/sobolmetrices.cpp holds them expanded)\n")
        f.write("// for dimensions 1..1023, each as {s, a, m_1..m_s}: degree, the polynomial's\n")
        f.write("// middle coefficients a_1..a_{s-1} (a_1 the top bit) and the initial odd m_k < 2^k.\n")
        f.write("#pragma once\n#include <cstdint>\n\n")
        f.write(f"static const int kSobolDirnumCount = {len(stream)};\n")
        f.write("static const uint16_t kSobolDirnums[] = {\n")
        for i in range(0, len(stream), 16):
            f.write("    " + ", ".join(str(x) for x in stream[i:i + 16]) + ",\n")
        f.write("};\n")
    print(f"wrote {OUT}: {len(rows)} dimensions, {len(stream)} numbers, max degree "
          f"{max(r[0] for r in rows)}")


if __name__ == "__main__":
    main()
