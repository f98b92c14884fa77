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
OUT_H2 = os.path.join(HERE, "..", "maroontree_amd", "csrc", "av1_cdf_tables.h")

# Default coefficient-coding CDFs (AV1 spec defaults, per qctx) in the raw (non-inverted) form
# the reference stores (src/coef_q.rs), the AV1 scans (src/coef_q.rs, src/tables.rs:348-455)
# and the coeff_base context offsets (src/tables.rs:354, :456, :466), plus the intra mode /
# tx-type CDFs of the mode rate, and the partition CDFs of the decision snapshot Cdfs::new
# (src/tables.rs:377, :403; src/coder/cdfs.rs:544-555).
CDF_TABLES_COEF_Q = [
    "SKIP_TX4", "SKIP_TX8", "SKIP_TX16", "SKIP_TX32", "SKIP_TX64",
    "BASE_TOK_TX4_LUMA_Q", "BASE_TOK_TX4_CHROMA_Q", "BASE_TOK_TX8_LUMA_Q", "BASE_TOK_TX8_CHROMA_Q",
    "BASE_TOK_TX16_LUMA_Q", "BASE_TOK_TX16_CHROMA_Q", "BASE_TOK_TX32_LUMA_Q", "BASE_TOK_TX32_CHROMA_Q",
    "BASE_TOK_TX64_LUMA_Q",
    "BR_TOK_TX4_LUMA_Q", "BR_TOK_TX4_CHROMA_Q", "BR_TOK_TX8_LUMA_Q", "BR_TOK_TX8_CHROMA_Q",
    "BR_TOK_TX16_LUMA_Q", "BR_TOK_TX16_CHROMA_Q", "BR_TOK_TX32_LUMA_Q", "BR_TOK_TX32_CHROMA_Q",
    "EOB_BASE_TX4_LUMA_Q", "EOB_BASE_TX4_CHROMA_Q", "EOB_BASE_TX8_LUMA_Q", "EOB_BASE_TX8_CHROMA_Q",
    "EOB_BASE_TX16_LUMA_Q", "EOB_BASE_TX16_CHROMA_Q", "EOB_BASE_TX32_LUMA_Q", "EOB_BASE_TX32_CHROMA_Q",
    "EOB_BASE_TX64_LUMA_Q",
    "EOB_HI_TX4_LUMA", "EOB_HI_TX4_CHROMA", "EOB_HI_TX8_LUMA", "EOB_HI_TX8_CHROMA",
    "EOB_HI_TX16_LUMA", "EOB_HI_TX16_CHROMA", "EOB_HI_TX32_LUMA", "EOB_HI_TX32_CHROMA",
    "EOB_HI_TX64_LUMA",
    "DC_SIGN_Q",
    "EOB_BIN_16_LUMA", "EOB_BIN_16_CHROMA", "EOB_BIN_16_LUMA_1D",
    "EOB_BIN_32_LUMA", "EOB_BIN_32_CHROMA",
    "EOB_BIN_64_LUMA", "EOB_BIN_64_CHROMA", "EOB_BIN_64_LUMA_1D",
    "EOB_BIN_128_LUMA", "EOB_BIN_128_CHROMA", "EOB_BIN_128_LUMA_1D",
    "EOB_BIN_256_LUMA", "EOB_BIN_256_CHROMA",
    "EOB_BIN_512_LUMA", "EOB_BIN_512_CHROMA",
    "EOB_BIN_1024_LUMA", "EOB_BIN_1024_CHROMA",
    "SCAN_16X16", "SCAN_16X4", "SCAN_4X16", "SCAN_8X16", "SCAN_16X8", "SCAN_32X32",
    "SCAN_16X32", "SCAN_32X16",
]
CDF_TABLES_TABLES = [
    "SCAN_4X4", "SCAN_4X8", "SCAN_8X4", "SCAN_8X8",
    "LO_CTX_OFF", "LO_CTX_OFF_WLH", "LO_CTX_OFF_WGH",
    "KF_Y_MODE_CDF", "UV_MODE_NOCFL_CDF", "UV_MODE_CFL_CDF", "TXTP_INTRA1_TX8", "TXTP_INTRA2_TX16",
    "TXTP_INTRA1_TX4", "ANGLE_DELTA_CDF", "INTRA_FILTER_TAPS", "FILTER_INTRA_CDF",
    "FILTER_INTRA_MODE_CDF", "PART_SPLIT_CDF", "PART_BL8_CDF", "SKIP_CDF",
]


def _typed_table(text, name):
    """(ctype, dims, values) of `static NAME: <rust array type> = [...]`."""
    m = re.search(r"static " + name + r":\s*([^=]+)=", text)
    ty = m.group(1).strip()
    dims = [int(d) for d in re.findall(r";\s*(\d+)\s*\]", ty)]
    base = re.search(r"(u8|u16|u32|i8|i16|i32)", ty).group(1)
    count = 1
    for d in dims:
        count *= d
    body = text[m.end():]
    vals = [int(v) for v in re.findall(r"-?\b\d+\b", body)][:count]
    assert len(vals) == count, name
    ctype = {"u8": "uint8_t", "u16": "uint16_t", "u32": "uint32_t", "i8": "int8_t",
             "i16": "int16_t", "i32": "int32_t"}[base]
    return ctype, list(reversed(dims)), vals


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
    tables_rs = open(os.path.join(REF, "tables.rs")).read()
    out2 = ["// Generated by oracle/gen_tables.py -- do not edit by hand.",
            "// AV1 default coefficient CDFs (raw form, per qctx), scans, coeff_base context",
            "// offsets and intra mode / tx-type CDFs, with the values of the maroontree encoder",
            "// (reference src/coef_q.rs, src/tables.rs).",
            "#pragma once", "#include <stdint.h>", ""]
    for src_text, names in ((coef_q, CDF_TABLES_COEF_Q), (tables_rs, CDF_TABLES_TABLES)):
        for name in names:
            ctype, dims, vals = _typed_table(src_text, name)
            sums["cdf_" + name.lower()] = fnv1a(vals, 4)
            dim_s = "".join(f"[{d}]" for d in dims)
            out2.append(f"static const {ctype} MT_{name}{dim_s} = {{")
            out2.append(arr(vals, 16))
            out2.append("};")
    # Intra prediction tables (reference src/intrapred.rs): the directional derivatives
    # (:1081), the SMOOTH weights of sm_weights() (:1091) laid out so that the weights of
    # size n start at index n, and the CfL CDFs (:1048, :1051).
    intra = open(os.path.join(REF, "intrapred.rs")).read()
    for name in ("DR_INTRA_DERIVATIVE", "CFL_SIGN_CDF", "CFL_ALPHA_CDF"):
        ctype, dims, vals = _typed_table(intra, name)
        sums["cdf_" + name.lower()] = fnv1a(vals, 4)
        dim_s = "".join(f"[{d}]" for d in dims)
        out2.append(f"static const {ctype} MT_{name}{dim_s} = {{")
        out2.append(arr(vals, 16))
        out2.append("};")
    body = intra[intra.index("fn sm_weights("):]
    body = body[:body.index("_ => unreachable")]
    sm = [0] * 128
    for n in (4, 8, 16, 32, 64):
        m = re.search(r"\b%d => &\[([^\]]*)\]" % n, body)
        w = [int(v) for v in re.findall(r"\d+", m.group(1))]
        assert len(w) == n, n
        sm[n:2 * n] = w
    sums["cdf_sm_weights"] = fnv1a(sm, 4)
    out2.append("static const int32_t MT_SM_WEIGHTS[128] = {")
    out2.append(arr(sm, 16))
    out2.append("};")
    with open(OUT_H2, "w") as f:
        f.write("\n".join(out2) + "\n")
    with open(OUT_J, "w") as f:
        json.dump({k: f"0x{v:016x}" for k, v in sums.items()}, f, indent=1)
        f.write("\n")
    print("wrote", OUT_H, OUT_J)


if __name__ == "__main__":
    sys.exit(main())
