#!/usr/bin/env python3
"""Writes gx_pow5.h: 128-bit truncations of 5^q for the decimal -> binary64 conversion
of gx_decimal.h (q in [GX_POW5_MIN_Q, GX_POW5_MAX_Q]).

Entry q holds t_q = floor(5^q * 2^s_q) with s_q chosen so that 2^127 <= t_q < 2^128,
as two 64-bit words (high, low), and s_q.  Exact integer arithmetic throughout.

    python3 gen_pow5.py > gx_pow5.h
"""

QMIN, QMAX = -342, 308


def entry(q):
    if q >= 0:
        p = 5 ** q
        s = 128 - p.bit_length()
        t = p << s if s >= 0 else p >> -s
    else:
        p = 5 ** -q
        s = 127 + p.bit_length()
        t = (1 << s) // p
    assert (1 << 127) <= t < (1 << 128), q
    return t >> 64, t & ((1 << 64) - 1), s


def main():
    rows = [entry(q) for q in range(QMIN, QMAX + 1)]
    out = ["// gx_pow5.h -- generated by gen_pow5.py; do not edit.",
           "// t_q = floor(5^q * 2^s_q), 2^127 <= t_q < 2^128, for q in [GX_POW5_MIN_Q, GX_POW5_MAX_Q]:",
           "// GX_POW5_WORDS lists (high, low) per q, GX_POW5_SHIFTS lists s_q.",
           "#pragma once",
           f"#define GX_POW5_MIN_Q ({QMIN})",
           f"#define GX_POW5_MAX_Q ({QMAX})",
           "#define GX_POW5_WORDS \\"]
    for i, (hi, lo, _) in enumerate(rows):
        sep = "," if i + 1 < len(rows) else ""
        out.append(f"  0x{hi:016x}ull, 0x{lo:016x}ull{sep} \\")
    out.append("")
    out.append("#define GX_POW5_SHIFTS \\")
    shifts = [r[2] for r in rows]
    for i in range(0, len(shifts), 16):
        part = ", ".join(str(v) for v in shifts[i:i + 16])
        sep = "," if i + 16 < len(shifts) else ""
        out.append(f"  {part}{sep} \\")
    out.append("")
    print("\n".join(out))


if __name__ == "__main__":
    main()
