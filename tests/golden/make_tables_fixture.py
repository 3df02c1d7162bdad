"""Extract the reference decoder's numeric tables into tests/golden/ref_tables.json.

Run in the build container (where the reference source is mounted read-only):

    python tests/golden/make_tables_fixture.py /root/reference

Only numbers are kept (coefficient matrices and lookup tables); they pin the
oracle's and the kernels' tables against the reference's own.  Sources:
  src/transform.rs  idct_raw_4 (:406), idct_odd_8/16/32 (:67-404), DST4 (:30),
                    DEQUANT_SCALE (:37)
  src/intra.rs      INTRA_PRED_ANGLE (:31), INV_ANGLE (:37)
  src/deblock.rs    BETA (:35), TC (:44)
  src/decode.rs     qpc table (:7599)
  src/mc.rs         LUMA_FILTER (:30), CHROMA_FILTER (:37)
  src/config.rs     QUANT_INTRA_DEFAULT_8X8 (:360), QUANT_INTER_DEFAULT_8X8 (:366)
"""
import json
import os
import re
import sys


def odd_matrix(src: str, n: int):
    m = re.search(r"fn idct_odd_%d\(c: \[i32; %d\]\) -> \[i32; %d\] \{(.*?)\n\}" % (n, n, n // 2), src, re.S)
    body = m.group(1)
    body = body[body.index("[") + 1: body.rindex("]")]
    rows = [r for r in re.split(r",\s*\n", body) if r.strip()]
    out = []
    for r in rows:
        terms = re.findall(r"(-?\s*\d+)\s*\*\s*c\[(\d+)\]", r.replace("+ -", "-").replace("+", " "))
        coefs = {int(j): int(v.replace(" ", "")) for v, j in terms}
        out.append([coefs[j] for j in range(1, n, 2)])
    return out


def int_table(src: str, name: str):
    m = re.search(r"static %s: \[[^\]]*\] = \[(.*?)\];" % re.escape(name), src, re.S)
    return [int(v) for v in re.findall(r"-?\d+", m.group(1))]


def int_rows(src: str, name: str, width: int):
    """A static 2-D table [[T; width]; rows] as a list of rows."""
    m = re.search(r"static %s: \[\[[^\]]*\]; \d+\] = \[(.*?)\];" % re.escape(name), src, re.S)
    vals = [int(v) for v in re.findall(r"-?\d+", m.group(1))]
    return [vals[i:i + width] for i in range(0, len(vals), width)]


def main(ref: str) -> None:
    tr = open(os.path.join(ref, "src/transform.rs")).read()
    intra = open(os.path.join(ref, "src/intra.rs")).read()
    deb = open(os.path.join(ref, "src/deblock.rs")).read()
    dec = open(os.path.join(ref, "src/decode.rs")).read()
    mc = open(os.path.join(ref, "src/mc.rs")).read()
    cfg = open(os.path.join(ref, "src/config.rs")).read()
    dst = re.search(r"static DST4: \[\[i32; 4\]; 4\] = \[(.*?)\];", tr, re.S).group(1)
    dst_rows = [[int(v) for v in re.findall(r"-?\d+", r)] for r in re.findall(r"\[([^\]]*)\]", dst)]
    ls = re.search(r"const DEQUANT_SCALE: \[i64; 6\] = \[(.*?)\];", tr).group(1)
    qpc = re.search(r"const T: \[i32; 14\] = \[(.*?)\];", dec).group(1)
    raw4 = re.search(r"fn idct_raw_4\(c: \[i32; 4\]\) -> \[i32; 4\] \{(.*?)\n\}", tr, re.S).group(1)
    fixture = {
        "idct_odd": {str(n): odd_matrix(tr, n) for n in (8, 16, 32)},
        "idct_raw_4_constants": sorted(set(int(v) for v in re.findall(r"(\d+)\s*\*", raw4))),
        "dst4": dst_rows,
        "dequant_scale": [int(v) for v in re.findall(r"\d+", ls)],
        "intra_pred_angle": int_table(intra, "INTRA_PRED_ANGLE"),
        "inv_angle": int_table(intra, "INV_ANGLE"),
        "beta": int_table(deb, "BETA"),
        "tc": int_table(deb, "TC"),
        "qpc_30_43": [int(v) for v in re.findall(r"\d+", qpc)],
        "luma_filter": int_rows(mc, "LUMA_FILTER", 8),
        "chroma_filter": int_rows(mc, "CHROMA_FILTER", 4),
        "scaling_default_intra_8x8": int_table(cfg, "QUANT_INTRA_DEFAULT_8X8"),
        "scaling_default_inter_8x8": int_table(cfg, "QUANT_INTER_DEFAULT_8X8"),
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_tables.json")
    with open(out, "w") as f:
        json.dump(fixture, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
