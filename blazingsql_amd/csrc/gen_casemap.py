"""Writes casemap.h: the Unicode case data the string kernels use for
UPPER/LOWER/INITCAP, for the Basic Multilingual Plane (code points above
U+FFFF are left unchanged, as libcudf's 64K-entry character tables do).

Source: Python's unicodedata (its Unicode version is recorded in the header).
  UPPER_MAP / LOWER_MAP: code point -> full case mapping (1 to 3 code points,
      the unconditional mappings of SpecialCasing included), sorted by code
      point, one uint64 each: cp << 48 | m0 << 32 | m1 << 16 | m2 (unused 0).
  ALPHA_RANGES: [start, end) ranges of alphabetic code points (category L*).
Run: python3 gen_casemap.py  (rewrites casemap.h beside this file)
"""
import os
import unicodedata


def mapping(cp, fn):
    c = chr(cp)
    m = fn(c)
    if m == c or len(m) > 3 or any(ord(x) > 0xFFFF for x in m):
        return None
    # context-dependent lowering (final sigma) is not part of the mapping
    return [ord(x) for x in m]


def pack(cp, m):
    m = m + [0] * (3 - len(m))
    return (cp << 48) | (m[0] << 32) | (m[1] << 16) | m[2]


def main():
    upper, lower, alpha = [], [], []
    start = None
    for cp in range(0x10000):
        if 0xD800 <= cp <= 0xDFFF:
            is_alpha = False
        else:
            u = mapping(cp, str.upper)
            l = mapping(cp, str.lower)
            if u:
                upper.append(pack(cp, u))
            if l:
                lower.append(pack(cp, l))
            is_alpha = chr(cp).isalpha()
        if is_alpha and start is None:
            start = cp
        if not is_alpha and start is not None:
            alpha.append((start, cp))
            start = None
    if start is not None:
        alpha.append((start, 0x10000))

    text = [
        "// casemap.h -- generated by gen_casemap.py from Python's unicodedata "
        f"(Unicode {unicodedata.unidata_version}); do not edit.",
        "// Full case mappings and alphabetic ranges of the Basic Multilingual Plane",
        "// for UPPER / LOWER / INITCAP (see gen_casemap.py for the layout).",
        "#pragma once",
        "#include <stdint.h>",
        "",
        f"#define BSQL_CASE_UPPER_N {len(upper)}",
        f"#define BSQL_CASE_LOWER_N {len(lower)}",
        f"#define BSQL_ALPHA_RANGES_N {len(alpha)}",
        "",
        "#define BSQL_CASE_UPPER_VALUES \\",
    ]
    def macro(vals, per):
        lines = []
        for i in range(0, len(vals), per):
            lines.append("    " + ", ".join(vals[i:i + per]) + ", \\")
        lines[-1] = lines[-1][:-4]
        return lines
    text += macro([f"0x{v:016x}ULL" for v in upper], 4)
    text += ["", "#define BSQL_CASE_LOWER_VALUES \\"]
    text += macro([f"0x{v:016x}ULL" for v in lower], 4)
    text += ["", "#define BSQL_ALPHA_RANGES_VALUES \\"]
    text += macro([f"0x{a:04x}u, 0x{b:05x}u" for a, b in alpha], 6)
    text.append("")
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "casemap.h"), "w") as f:
        f.write("\n".join(text))


if __name__ == "__main__":
    main()
