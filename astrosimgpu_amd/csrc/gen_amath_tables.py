"""Generate amath_tables.h: the lookup tables of the FP64 exp, log, sin and cos
in amath.h.

Both functions follow the table-driven method of glibc 2.35's exp and log
(Szabolcs Nagy's algorithms, shipped in glibc since 2.28 and in Arm's
optimized-routines). The tables are not copied from anywhere: they are
recomputed here from the construction rules those algorithms publish.

exp, N = 128 entries, for k in [0, N):
    H[k] = 2^(k/N) rounded to nearest double,
    T[k] = (2^(k/N) - H[k]) / H[k] rounded to nearest double,
    tab[2k] = bits(T[k]), tab[2k+1] = bits(H[k]) - k * 2^52 / N.

log, N = 128 subintervals of z in [0x1.6p-1, 0x1.6p0): entry i holds
(invc, logc), where invc is a double within 2^29 ulps of 1/center(i),
logc = RN(log(1/invc)) is a multiple of 2^-43 (so that k*ln2hi + logc is
exact) and |log(1/invc) - logc| < 2^-66. Several doubles meet that rule per
subinterval; gen_log_table.c enumerates them in quad precision and keeps the
one whose results equal the host C library's log on that subinterval (about a
minute on 8 cores). Its output file is read here.

sin/cos, 110 entries for k in [0, 110) (glibc 2.35's table-driven sin/cos,
the IBM Accurate Mathematical Library method, looks up the nearest multiple
k/128 of the reduced argument, which stays below 109.5/128):
    tab[4k] = RN(sin(k/128)), tab[4k+1] = RN(sin(k/128) - tab[4k]),
    tab[4k+2] = RN(cos(k/128)), tab[4k+3] = RN(cos(k/128) - tab[4k+2]),
i.e. each value as an unevaluated double-double sum, from 300-bit mpmath.

Usage:  python gen_amath_tables.py LOGTAB_FILE > amath_tables.h
"""

import struct
import sys

import mpmath

mpmath.mp.prec = 300


def bits(d: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", d))[0]


def exp_table(n: int = 128):
    out = []
    for k in range(n):
        v = mpmath.power(2, mpmath.mpf(k) / n)
        h = float(v)
        t = float((v - mpmath.mpf(h)) / mpmath.mpf(h))
        out.append((bits(t), (bits(h) - (k << 45)) & 0xFFFFFFFFFFFFFFFF))
    return out


def sincos_table(n: int = 110):
    out = []
    for k in range(n):
        x = mpmath.mpf(k) / 128
        row = []
        for v in (mpmath.sin(x), mpmath.cos(x)):
            hi = float(v)
            row += [hi, float(v - mpmath.mpf(hi))]
        out.append(row)
    return out


def log_table(path: str):
    entries = {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            i = int(parts[0])
            invc = float.fromhex(parts[1])
            logc = float.fromhex(parts[2])
            entries[i] = (invc, logc)
    if sorted(entries) != list(range(128)):
        raise SystemExit("log table search output is incomplete")
    return [entries[i] for i in range(128)]


def main() -> None:
    if len(sys.argv) != 2:
        raise SystemExit(__doc__)
    print("// Generated by gen_amath_tables.py. Do not edit by hand.")
    print("#pragma once")
    print("#include <cstdint>")
    print()
    print("// exp: {bits(T[k]), bits(H[k]) - (k << 45)}, k = 0..127")
    print("ASG_TABLE std::uint64_t ASG_EXP_TAB[256] = {")
    for t, h in exp_table():
        print(f"    0x{t:016x}ULL, 0x{h:016x}ULL,")
    print("};")
    print()
    print("// log: {invc, logc}, i = 0..127")
    print("ASG_TABLE double ASG_LOG_TAB[256] = {")
    for invc, logc in log_table(sys.argv[1]):
        print(f"    {invc.hex()}, {logc.hex()},")
    print("};")
    print()
    print("// sin/cos: {sin hi, sin lo, cos hi, cos lo} of k/128, k = 0..109")
    print("ASG_TABLE double ASG_SINCOS_TAB[440] = {")
    for row in sincos_table():
        print("    " + ", ".join(v.hex() for v in row) + ",")
    print("};")


if __name__ == "__main__":
    main()
