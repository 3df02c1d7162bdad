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
  src/cabac/contexts.rs  initValues of the I-slice context set (:139 init_islice),
                    the P/B rows (:234 init, LAST_SIG_PREFIX / SIG_COEFF_FLAG /
                    COEFF_G1 / COEFF_G2), the intra-mode contexts (:370)
  src/cabac/engine.rs    RANGE_TAB_LPS (:35), TRANS_IDX_MPS (:55), TRANS_IDX_LPS (:63)
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


def fn_body(src: str, header: str) -> str:
    """The text of the function whose signature starts with `header`."""
    i = src.index(header)
    j = src.index("{", i)
    depth, k = 0, j
    while True:
        if src[k] == "{":
            depth += 1
        elif src[k] == "}":
            depth -= 1
            if depth == 0:
                return src[j:k]
        k += 1


def ctx_islice(ctxs: str):
    """initValues of the I-slice context set: field -> list."""
    body = fn_body(ctxs, "pub(crate) fn init_islice(qp: u8) -> Self")
    out = {}
    for m in re.finditer(r"(\w+): arr\(\s*\[(.*?)\],\s*qp\s*,?\s*\)", body, re.S):
        vals = m.group(2)
        rep = re.match(r"\s*(\d+)\s*;\s*(\d+)\s*$", vals)
        out[m.group(1)] = [int(rep.group(1))] * int(rep.group(2)) if rep else [int(v) for v in re.findall(r"\d+", vals)]
    for m in re.finditer(r"(\w+): c\((\d+), qp\)", body):
        out[m.group(1)] = [int(m.group(2))]
    return out


def ctx_intra_modes(ctxs: str):
    i = ctxs.index("impl IntraModeContexts")
    body = fn_body(ctxs[i:], "pub(crate) fn init_islice(qp: u8) -> Self")
    return {m.group(1): [int(m.group(2))] for m in re.finditer(r"(\w+): CtxModel::init\((\d+), qp\)", body)}


def ctx_typed(ctxs: str):
    """The multi-initType rows the P/B constructor selects from."""
    body = fn_body(ctxs, "pub(crate) fn init(init_type: u8, qp: u8) -> Self")
    rows = {}
    for name in ("split_cu", "split_tf", "cbf_luma"):
        m = re.search(r"let %s = \[(.*?)\]\[it\];" % name, body, re.S)
        rows[name] = [[int(v) for v in re.findall(r"\d+", r)] for r in re.findall(r"\[([^\[\]]*)\]", m.group(1))]
    m = re.search(r"let cbf_chroma_all: \[\[u8; 4\]; 3\] = \[(.*?)\];", body, re.S)
    rows["cbf_chroma"] = [[int(v) for v in re.findall(r"\d+", r)] for r in re.findall(r"\[([^\[\]]*)\]", m.group(1))]
    m = re.search(r"let csb = \[(.*?)\]\[it\];", body, re.S)
    rows["coded_sub_block_flag"] = [[int(v) for v in re.findall(r"\d+", r)] for r in re.findall(r"\[([^\[\]]*)\]", m.group(1))]
    m = re.search(r"sao_type_idx: c\(\[(.*?)\]\[it\], qp\)", body)
    rows["sao_type_idx"] = [int(v) for v in re.findall(r"\d+", m.group(1))]
    m = re.search(r"let pm = \[(.*?)\];", body)
    rows["part_mode_pm"] = [int(v) for v in re.findall(r"\d+", m.group(1))]
    rows["last_sig_prefix"] = int_table(ctxs, "LAST_SIG_PREFIX")
    rows["sig_coeff_flag"] = int_rows(ctxs, "SIG_COEFF_FLAG", 42)
    rows["coeff_g1"] = int_table(ctxs, "COEFF_G1")
    rows["coeff_g2"] = int_table(ctxs, "COEFF_G2")
    i = ctxs.index("impl IntraModeContexts")
    ib = fn_body(ctxs[i:], "pub(crate) fn init(init_type: u8, qp: u8) -> Self")
    for name in ("prev_intra_luma_pred_flag", "intra_chroma_pred_mode"):
        m = re.search(r"%s: CtxModel::init\(\[(.*?)\]\[it\], qp\)" % name, ib)
        rows[name] = [int(v) for v in re.findall(r"\d+", m.group(1))]
    return rows


def main(ref: str) -> None:
    tr = open(os.path.join(ref, "src/transform.rs")).read()
    intra = open(os.path.join(ref, "src/intra.rs")).read()
    deb = open(os.path.join(ref, "src/deblock.rs")).read()
    dec = open(os.path.join(ref, "src/decode.rs")).read()
    mc = open(os.path.join(ref, "src/mc.rs")).read()
    cfg = open(os.path.join(ref, "src/config.rs")).read()
    # integer literal suffixes (184u8) dropped so that only the values remain
    ctxs = re.sub(r"(\d)(u8|u16|u32|usize)\b", r"\1", open(os.path.join(ref, "src/cabac/contexts.rs")).read())
    eng = open(os.path.join(ref, "src/cabac/engine.rs")).read()
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
        "cabac_init_islice": {**ctx_islice(ctxs), **ctx_intra_modes(ctxs)},
        "cabac_init_typed": ctx_typed(ctxs),
        "cabac_range_lps": int_rows(eng, "RANGE_TAB_LPS", 4),
        "cabac_trans_mps": int_table(eng, "TRANS_IDX_MPS"),
        "cabac_trans_lps": int_table(eng, "TRANS_IDX_LPS"),
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_tables.json")
    with open(out, "w") as f:
        json.dump(fixture, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
