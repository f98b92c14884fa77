#!/usr/bin/env python3
"""Recipe that writes ``maroontree_amd/csrc/av1_tables.h`` from the reference's table text.

Runs only in the build container (it reads the reference tree at /root/reference); the
generated header is committed, so nothing on a GPU box needs the reference.

What it extracts, and why these must be identical rather than re-derived:

* ``Dc_Qlookup`` / ``Ac_Qlookup`` for 8/10/12-bit -- AV1-normative dequantizer tables
  (reference ``src/coef_q.rs:44,59,2616,2634,2652,2671``).
* The inverse quantizer matrices, 15 levels x {luma, chroma} x 3344 weights in the decoder's
  transposed (x-major, ``rc = x*h + y``) order -- AV1-normative (reference
  ``src/qm_tables.rs:33``).
* The encoder-defined forward ADST-4/8/16 matrices in Q12 (reference ``src/dct.rs:3178``,
  ``src/dct.rs:304``, ``src/dct.rs:445``). They are not normative, but the levels and
  targets the encoder produces depend on every entry, so bit-exact parity needs the same
  integers.

The header also records a FNV-1a checksum of each table so that a test can pin the committed
data against the values read here (``tests/golden/tables_checksum.json``).
"""
import json
import os
import re
import sys

REF = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT_H = os.path.join(HERE, "..", "maroontree_amd", "csrc", "av1_tables.h")
OUT_J = os.path.join(HERE, "..", "tests", "golden", "tables_checksum.json")


def _ints_after(text, name, count):
    start = text.index(name)
    body = text[text.index("=", start):]
    vals = [int(v) for v in re.findall(r"-?\b\d+\b", body)]
    # the first numbers after '=' may include array dims in the type; skip none: the
    # type precedes '=' so the first integers are the data.
    return vals[:count]


def fnv1a(vals, width):
    h = 0xCBF29CE484222325
    for v in vals:
        v &= (1 << (8 * width)) - 1
        for b in range(width):
            h ^= (v >> (8 * b)) & 0xFF
            h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def main():
    coef_q = open(os.path.join(REF, "coef_q.rs")).read()
    qm = open(os.path.join(REF, "qm_tables.rs")).read()
    dct = open(os.path.join(REF, "dct.rs")).read()

    dc = [_ints_after(coef_q, f"static DC_QLOOKUP_{b}:", 256) for b in (8, 10, 12)]
    ac = [_ints_after(coef_q, f"static AC_QLOOKUP_{b}:", 256) for b in (8, 10, 12)]
    iqm = _ints_after(qm, "static AV1_IQM:", 15 * 2 * 3344)
    adst4 = _ints_after(dct, "static ADST4_FWD_Q12:", 16)
    adst8 = _ints_after(dct, "static ADST8_FWD_Q12:", 64)
    adst16 = _ints_after(dct, "static ADST16_FWD_Q12:", 256)
    assert len(iqm) == 100320 and all(1 <= v <= 255 for v in iqm)
    assert all(len(t) == 256 for t in dc + ac)

    def arr(vals, per_line=20):
        lines = []
        for i in range(0, len(vals), per_line):
            lines.append("  " + ", ".join(str(v) for v in vals[i:i + per_line]) + ",")
        return "\n".join(lines)

    sums = {
        "dc_qlookup": fnv1a(sum(dc, []), 2),
        "ac_qlookup": fnv1a(sum(ac, []), 2),
        "iqm": fnv1a(iqm, 1),
        "adst4": fnv1a(adst4, 4),
        "adst8": fnv1a(adst8, 4),
        "adst16": fnv1a(adst16, 4),
    }
    out = []
    out.append("// Generated by oracle/gen_tables.py -- do not edit by hand.")
    out.append("// AV1-normative dequantizer and quantizer-matrix tables, and the encoder's forward")
    out.append("// ADST matrices (Q12), with the values of the maroontree encoder (reference")
    out.append("// src/coef_q.rs:44-2690, src/qm_tables.rs:33, src/dct.rs:304,445,3178).")
    out.append("#pragma once")
    out.append("#include <stdint.h>")
    out.append("")
    out.append("#define MT_QM_USED_SIZE 3344")
    out.append("#define MT_QM_LEVELS 15")
    for k, v in sums.items():
        out.append(f"#define MT_TABLE_FNV_{k.upper()} 0x{v:016x}ull")
    out.append("")
    out.append("// [bit-depth index: 0 = 8-bit, 1 = 10-bit, 2 = 12-bit][qindex]")
    out.append("static const uint16_t MT_DC_QLOOKUP[3][256] = {")
    out.append(arr(sum(dc, [])))
    out.append("};")
    out.append("static const uint16_t MT_AC_QLOOKUP[3][256] = {")
    out.append(arr(sum(ac, [])))
    out.append("};")
    out.append("")
    out.append("// [level][0 = luma, 1 = chroma][offset of the size + rc], rc = x*h + y")
    out.append("static const uint8_t MT_IQM[MT_QM_LEVELS][2][MT_QM_USED_SIZE] = {")
    out.append(arr(iqm, 32))
    out.append("};")
    out.append("")
    out.append("// Forward ADST matrices, row = output frequency, column = input sample, Q12.")
    out.append("static const int32_t MT_FADST4_Q12[4][4] = {")
    out.append(arr(adst4, 4))
    out.append("};")
    out.append("static const int32_t MT_FADST8_Q12[8][8] = {")
    out.append(arr(adst8, 8))
    out.append("};")
    out.append("static const int32_t MT_FADST16_Q12[16][16] = {")
    out.append(arr(adst16, 16))
    out.append("};")
    with open(OUT_H, "w") as f:
        f.write("\n".join(out) + "\n")
    with open(OUT_J, "w") as f:
        json.dump({k: f"0x{v:016x}" for k, v in sums.items()}, f, indent=1)
        f.write("\n")
    print("wrote", OUT_H, OUT_J)


if __name__ == "__main__":
    sys.exit(main())
