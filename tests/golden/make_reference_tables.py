"""Extract the reference's constant tables into a JSON fixture.

Reads the reference's Rust sources as text (nothing is compiled or run) and
records the numeric tables the tests pin the oracle and the product's host
tables against:

* EXP2FT (src/adaptive_quant.rs:737): the 64 (value, correction) bit pairs of
  fast_exp2's table.
* K_COEFF_ORDER_8X8 / 16X8 / 16X16 / 32X16 / 32X32 (src/ac_context.rs:164-340):
  the natural coefficient orders.
* DEQUANT_MATRIX_8X8 / DEQUANT_MATRIX_16X8 (src/quant_weights.rs:246,455):
  the stored DCT8 and 16x8 step tables.
* WC8 / WC16 / WC32 / WC64 (src/dct.rs): butterfly constants.
* RESAMPLE_SCALE_64_TO_8 (src/dct.rs:2005): DCT64 DC resampling scales.
* WC128 (src/dct128.rs:76): the 128-point butterfly constants.
* AFV_BASIS_TRANSPOSE (src/afv.rs:48): the 16x16 AFV basis, row j = pixel j.
* AQ (src/adaptive_quant.rs:41-206,410-433,547-729): the adaptive
  quantization constants, by name: module constants, the numbers of every
  `let` in the masking, modulation and field functions ("fn.var": [numbers
  of the initializer]), the knot schedules, and the literals of the
  per-pixel clamp and the exponent's log2(e).

Usage: python tests/golden/make_reference_tables.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json")

NUM = r"-?\d[\d_]*\.?[\d_]*(?:[eE][-+]?\d+)?"


def read(rel):
    with open(os.path.join(REF, rel), encoding="utf-8") as f:
        return f.read()


def block_after(text, marker):
    """Text of the bracketed initializer that follows `marker`."""
    i = text.index(marker)
    i = text.index("=", i)
    start = text.index("[", i)
    depth = 0
    for j in range(start, len(text)):
        if text[j] == "[":
            depth += 1
        elif text[j] == "]":
            depth -= 1
            if depth == 0:
                return text[start : j + 1]
    raise ValueError(marker)


def strip_comments(s):
    return re.sub(r"//[^\n]*", "", s)


def numbers(s):
    return [float(x.replace("_", "")) for x in re.findall(NUM, strip_comments(s))]


def literals(s):
    """Float literals of an expression (digits glued to a name, such as the
    32 of an f32 suffix or the 2 of exp2, are not literals)."""
    found = re.findall(r"(?<![\w.])(" + NUM + ")", strip_comments(s))
    return [float(x.replace("_", "")) for x in found if re.search(r"[.eE]", x)]


def fn_body(text, name):
    """Body of `fn name` (from its opening brace to the matching one)."""
    m = re.search(r"fn " + name + r"\b", text)
    start = text.index("{", m.end())
    depth = 0
    for j in range(start, len(text)):
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
            if depth == 0:
                return text[start : j + 1]
    raise ValueError(name)


def aq_constants(aq):
    out = {}
    for m in re.finditer(r"(?:pub\(crate\) )?const (\w+): f32 = ([^;]+);", aq):
        out[m.group(1)] = literals(m.group(2))
    for name in ("AQ_KNOTS", "AQ_STRENGTH", "AQ_BASE_LEVEL"):
        out[name] = numbers(block_after(aq, "static " + name))[-7:]
    for fn in ("ratio_cubic_to_simple_gamma", "compute_mask", "masking_sqrt",
               "hf_modulation", "blue_modulation", "gamma_modulation",
               "fill_quant_field_scalar"):
        body = strip_comments(fn_body(aq, fn))
        for m in re.finditer(r"let (?:mut )?(\w+)(?:: [^=]+)? = ([^;]+);", body):
            nums = literals(m.group(2))
            if nums:
                out[fn + "." + m.group(1)] = nums
    body = strip_comments(fn_body(aq, "hf_modulation_strength"))
    out["hf_modulation_strength"] = literals(body)
    body = strip_comments(fn_body(aq, "fill_quant_field_scalar"))
    out["fill_quant_field_scalar.diff_clamp"] = literals(
        re.search(r"if diff >= (" + NUM + ")", body).group(1))
    out["fill_quant_field_scalar.exp_mul"] = literals(
        re.search(r"fast_exp2\(out_val \* (" + NUM + r")\)", body).group(1))
    return out


def main():
    out = {}
    aq = read("src/adaptive_quant.rs")
    exp2 = block_after(aq, "static EXP2FT")
    pairs = re.findall(r"\(0x([0-9A-Fa-f]+),\s*0x([0-9A-Fa-f]+)\)", exp2)
    out["EXP2FT"] = [[int(a, 16), int(b, 16)] for a, b in pairs]
    assert len(out["EXP2FT"]) == 64
    out["AQ"] = aq_constants(aq)

    ctx = read("src/ac_context.rs")
    for name, n in [("K_COEFF_ORDER_8X8", 64), ("K_COEFF_ORDER_16X8", 128),
                    ("K_COEFF_ORDER_16X16", 256), ("K_COEFF_ORDER_32X16", 512),
                    ("K_COEFF_ORDER_32X32", 1024)]:
        m = re.search(r"static\s+" + name + r"\s*:", ctx)
        vals = [int(v) for v in numbers(block_after(ctx[m.start():], name))]
        # the type annotation's length is parsed too; drop it
        vals = vals[-n:] if len(vals) > n else vals
        assert len(vals) == n, (name, len(vals))
        out[name] = vals

    qw = read("src/quant_weights.rs")
    for name, n in [("DEQUANT_MATRIX_8X8", 64), ("DEQUANT_MATRIX_16X8", 128)]:
        vals = numbers(block_after(qw, "static " + name))
        vals = vals[-3 * n:]
        assert len(vals) == 3 * n, (name, len(vals))
        out[name] = [vals[c * n:(c + 1) * n] for c in range(3)]

    dct = read("src/dct.rs")
    for name, n in [("WC8", 4), ("WC16", 8), ("WC32", 16), ("WC64", 32)]:
        m = re.search(r"const " + name + r": \[f32; \d+\] = ", dct)
        vals = numbers(block_after(dct[m.start():], name))[-n:]
        out[name] = vals

    m = re.search(r"const RESAMPLE_SCALE_64_TO_8: \[f32; 8\] = ", dct)
    out["RESAMPLE_SCALE_64_TO_8"] = numbers(block_after(dct[m.start():], "RESAMPLE_SCALE_64_TO_8"))[-8:]

    d128 = read("src/dct128.rs")
    m = re.search(r"static WC128: \[f32; 64\] = ", d128)
    vals = numbers(block_after(d128[m.start():], "WC128"))[-64:]
    assert len(vals) == 64
    out["WC128"] = vals

    afv = read("src/afv.rs")
    vals = numbers(block_after(afv, "static AFV_BASIS_TRANSPOSE"))[-256:]
    assert len(vals) == 256
    out["AFV_BASIS_TRANSPOSE"] = [vals[j * 16:(j + 1) * 16] for j in range(16)]

    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
