"""Writes the numeric tables of ITU-T H.264 that the analysis needs as C headers.

The values must equal the ones nervosys/IronAV codes with bit for bit (CABAC context initialisation,
significance/level context offsets, the LPS state transition, scans, dequantisation scales and the
encoder's forward quantisation multipliers), so they are read from its sources rather than retyped:

    crates/ironav-codec/src/h264/tables.rs   CABAC_INIT_I, CABAC_INIT_PB, SIG_OFF, LAST_OFF, ABS_OFF,
                                             SIG_OFF_8X8, LAST_OFF_8X8, LPS_RANGE, DEQUANT8_COEFF_INIT,
                                             ZIGZAG_DIRECT, ALPHA, BETA, TC0
    crates/ironav-codec/src/h264/cabac.rs    TRANS_LPS
    crates/ironav-codec/src/h264/ps.rs       ZIGZAG4
    crates/ironav-codec/src/h264/mb.rs       CHROMA_QP
    crates/ironav-codec/src/h264/enc/frame.rs MF, V, MF8

With a fourth argument `decoder` it writes the tables of the oracle's decoder instead (all three
CABAC_INIT_PB sets, the 4x4 dequantisation scales, the default scaling lists and the scans):

    crates/ironav-codec/src/h264/tables.rs   CABAC_INIT_I, CABAC_INIT_PB, SIG_OFF, LAST_OFF, ABS_OFF,
                                             SIG_OFF_8X8, LAST_OFF_8X8, LPS_RANGE, DEQUANT4_COEFF_INIT,
                                             DEQUANT8_COEFF_INIT, DEQUANT8_COEFF_INIT_SCAN, DEFAULT_SCALING4,
                                             DEFAULT_SCALING8, ZIGZAG_DIRECT, ALPHA, BETA, TC0
    crates/ironav-codec/src/h264/cabac.rs    TRANS_LPS
    crates/ironav-codec/src/h264/ps.rs       ZIGZAG4

Usage: python gen_tables.py <reference-root> <out.h> <PREFIX> [decoder]
The generated headers are committed; this script only reruns when the tables must be regenerated.
"""
import ast
import re
import sys

WANT = {
    "crates/ironav-codec/src/h264/tables.rs": [
        "CABAC_INIT_I", "CABAC_INIT_PB", "SIG_OFF", "LAST_OFF", "ABS_OFF", "SIG_OFF_8X8", "LAST_OFF_8X8",
        "LPS_RANGE", "DEQUANT8_COEFF_INIT", "ZIGZAG_DIRECT", "ALPHA", "BETA", "TC0",
    ],
    "crates/ironav-codec/src/h264/cabac.rs": ["TRANS_LPS"],
    "crates/ironav-codec/src/h264/ps.rs": ["ZIGZAG4"],
    "crates/ironav-codec/src/h264/mb.rs": ["CHROMA_QP"],
    "crates/ironav-codec/src/h264/enc/frame.rs": ["MF", "V", "MF8"],
}

WANT_DECODER = {
    "crates/ironav-codec/src/h264/tables.rs": [
        "CABAC_INIT_I", "CABAC_INIT_PB", "SIG_OFF", "LAST_OFF", "ABS_OFF", "SIG_OFF_8X8", "LAST_OFF_8X8",
        "LPS_RANGE", "DEQUANT4_COEFF_INIT", "DEQUANT8_COEFF_INIT", "DEQUANT8_COEFF_INIT_SCAN", "DEFAULT_SCALING4",
        "DEFAULT_SCALING8", "ZIGZAG_DIRECT", "ALPHA", "BETA", "TC0",
    ],
    "crates/ironav-codec/src/h264/cabac.rs": ["TRANS_LPS"],
    "crates/ironav-codec/src/h264/ps.rs": ["ZIGZAG4"],
}

CTYPE = {"u8": "uint8_t", "i8": "int8_t", "u16": "uint16_t", "i32": "int32_t"}


def parse(text, name):
    m = re.search(r"(?:pub(?:\(crate\))?\s+)?const\s+" + name + r"\s*:\s*([^=]+)=", text)
    if not m:
        raise KeyError(name)
    ty = m.group(1).strip()
    i = text.index("[", m.end())
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "[":
            depth += 1
        elif text[j] == "]":
            depth -= 1
            if depth == 0:
                break
    value = ast.literal_eval(re.sub(r"\s+", "", text[i:j + 1]))
    base = re.search(r"\b(u8|i8|u16|i32)\b", ty).group(1)
    return base, value


def dims(v):
    d = []
    while isinstance(v, list):
        d.append(len(v))
        v = v[0]
    return d


def emit(name, base, value, prefix):
    def init(x, depth=0):
        if not isinstance(x, list):
            return str(x)
        if not isinstance(x[0], list):
            vals = [str(v) for v in x]
            return "{" + ",".join(vals) + "}"
        return "{\n  " + ",\n  ".join(init(y, depth + 1) for y in x) + "}"

    ds = "".join(f"[{n}]" for n in dims(value))
    return f"{prefix}TABLE {CTYPE[base]} {prefix}{name}{ds} = " + init(value) + ";\n"


def main():
    root, out, prefix = sys.argv[1], sys.argv[2], sys.argv[3]
    decoder = len(sys.argv) > 4 and sys.argv[4] == "decoder"
    parts = [
        "/* Generated by ironav_amd/csrc/gen_tables.py from nervosys/IronAV's h264 sources: numeric tables of",
        " * ITU-T H.264 (and the encoder's forward quantisation multipliers) that must match bit for bit. */",
        f"#ifndef {prefix}TABLES_H",
        f"#define {prefix}TABLES_H",
        "#include <stdint.h>",
        "/* constant expressions in C++, so that code may select among entries without reading memory */",
        "#ifdef __cplusplus",
        f"#define {prefix}TABLE static constexpr",
        "#else",
        f"#define {prefix}TABLE static const",
        "#endif",
        "",
    ]
    for rel, names in (WANT_DECODER if decoder else WANT).items():
        text = open(f"{root}/{rel}").read()
        for n in names:
            base, value = parse(text, n)
            if n == "CABAC_INIT_PB" and not decoder:
                value = value[0]  # cabac_init_idc 0 is the only one the encoder uses
                n = "CABAC_INIT_P"
            parts.append(emit(n, base, value, prefix))
    parts.append(f"#endif\n")
    open(out, "w").write("\n".join(parts))


if __name__ == "__main__":
    main()
