// cabac_kat.cpp — known-answer and inverse-property checks of the CABAC
// engine, the Exp-Golomb bit I/O and the parameter-set parsers, built from
// the library's host sources by tests/test_cabac_engine.py.  Each check
// prints "name ok" or "name FAIL ..."; the exit status counts failures.
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../hpvcd_amd/csrc/bitio.h"
#include "../hpvcd_amd/csrc/cabac.h"
#include "../hpvcd_amd/csrc/hevc_ps.h"
#include "../hpvcd_amd/csrc/parse.h"

namespace hpvcd {
static std::string g_err;
hpvcd_status set_error(hpvcd_status st, const char* msg) {
    g_err = msg ? msg : "";
    return st;
}
}  // namespace hpvcd

extern "C" const char* hpvcd_last_error(void) { return hpvcd::g_err.c_str(); }

using namespace hpvcd;

static int failures = 0;
static void check(const char* name, bool ok, const std::string& why = "") {
    std::printf("%s %s%s%s\n", name, ok ? "ok" : "FAIL", ok ? "" : " ", ok ? "" : why.c_str());
    if (!ok) failures++;
}

// The reference engine test pcm_reinit_preserves_prefetched_cabac_bytes
// (src/cabac/engine.rs tests): after the terminated segment the stream aligns
// to byte 2, 16 raw PCM bits read 0x1234 and the engine restarts at byte 4.
static void kat_pcm_realign() {
    const uint8_t data[8] = {0x80, 0x00, 0x12, 0x34, 0xab, 0xcd, 0xef, 0x42};
    CabacDec d;
    d.start(data, sizeof(data));
    d.byte_align();
    const uint32_t pcm = d.read_bits(16);
    d.init_engine();
    const uint32_t want = ((uint32_t)data[4] << 1) | (uint32_t)(data[5] >> 7);
    check("pcm_realign", pcm == 0x1234 && d.offset() == want && d.range == 510,
          "pcm=" + std::to_string(pcm) + " offset=" + std::to_string(d.offset()));
}

// The reference engine test bypass_alignment_changes_only_range: alignment
// sets ivlCurrRange to 256 and touches neither the offset nor the position.
static void kat_bypass_alignment() {
    const uint8_t data[8] = {0x96, 0x35, 0xca, 0x71, 0x0f, 0xe4, 0x58, 0xa2};
    CabacDec d;
    d.start(data, sizeof(data));
    uint8_t ctx = (uint8_t)(23 | (1 << 6));
    for (int i = 0; i < 5; i++) d.decode(ctx);
    const uint32_t off = d.offset();
    const size_t pos = d.pos();
    d.range = 319;
    d.align_bypass();
    check("bypass_alignment", d.range == 256 && d.offset() == off && d.pos() == pos);
}

// Decoding what the encoder wrote gives back every bin, for random context
// bins, bypass runs, aligned bypass runs, and terminating bins followed by a
// PCM-style raw field and an engine restart (9.3.2.5, 9.3.4.3.5).
static void prop_engine_inverse() {
    std::mt19937 rng(12345);
    bool ok = true;
    std::string why;
    for (int trial = 0; trial < 200 && ok; trial++) {
        struct Op {
            int kind, ctx, bin;
            uint32_t raw;
            int nraw;
        };
        std::vector<Op> ops;
        const int n = 50 + (int)(rng() % 400);
        for (int i = 0; i < n; i++) {
            Op o{};
            const int r = (int)(rng() % 20);
            o.kind = r < 13 ? 0 : (r < 17 ? 1 : (r < 19 ? 2 : 3));
            o.ctx = (int)(rng() % 8);
            // skewed bins so contexts adapt towards both ends
            o.bin = (int)((rng() % 100) < (uint32_t)(o.ctx * 12 + 5));
            if (o.kind == 3) {
                o.nraw = 1 + (int)(rng() % 16);
                o.raw = rng() & ((1u << o.nraw) - 1);
            }
            ops.push_back(o);
        }
        BitWriter w;
        CabacEnc e;
        e.start(&w);
        uint8_t ectx[8], dctx[8];
        for (int c = 0; c < 8; c++) ectx[c] = dctx[c] = (uint8_t)((c * 7) | ((c & 1) << 6));
        for (const Op& o : ops) {
            if (o.kind == 0) e.encode(ectx[o.ctx], o.bin);
            else if (o.kind == 1) e.bypass(o.bin);
            else if (o.kind == 2) e.terminate(0);
            else {  // pcm_flag = 1: flush, align, raw bits, restart
                e.terminate(1);
                w.align_zero();
                w.bits(o.raw, o.nraw);
                e.start(&w);
            }
        }
        e.terminate(1);
        w.align_zero();
        CabacDec d;
        d.start(w.buf.data(), w.buf.size());
        for (size_t i = 0; i < ops.size() && ok; i++) {
            const Op& o = ops[i];
            int got;
            if (o.kind == 0) got = d.decode(dctx[o.ctx]);
            else if (o.kind == 1) got = d.bypass();
            else if (o.kind == 2) got = d.terminate();
            else {
                const int t = d.terminate();
                d.byte_align();
                const uint32_t raw = d.read_bits(o.nraw);
                d.init_engine();
                if (t != 1 || raw != o.raw) {
                    ok = false;
                    why = "terminated raw field at op " + std::to_string(i);
                }
                continue;
            }
            const int want = o.kind == 2 ? 0 : o.bin;
            if (got != want) {
                ok = false;
                why = "trial " + std::to_string(trial) + " op " + std::to_string(i);
            }
        }
        if (ok && (d.terminate() != 1 || d.failed())) {
            ok = false;
            why = "final terminating bin";
        }
        // the decoder reads exactly the bits the encoder wrote, stop bit included
        if (ok) {
            d.byte_align();
            if (d.pos() != w.buf.size() * 8) {
                ok = false;
                why = "decoder position " + std::to_string(d.pos()) + " vs " + std::to_string(w.buf.size() * 8);
            }
        }
    }
    check("engine_inverse", ok, why);
}

// ue(v) / se(v) known codes (H.265 Table 9-2 / 9-3) and NAL escaping
static void kat_expgolomb() {
    BitWriter w;
    w.ue(0);   // 1
    w.ue(1);   // 010
    w.ue(2);   // 011
    w.ue(7);   // 0001000
    w.se(-1);  // 011
    w.se(2);   // 00100
    w.align_zero();
    // 1 010 011 0001000 011 00100 -> 1010 0110 0010 0001 1001 00(00)
    const std::vector<uint8_t> want = {0xa6, 0x21, 0x90};
    bool ok = w.buf == want;
    BitReader r(w.buf.data(), w.buf.size());
    ok = ok && r.ue() == 0 && r.ue() == 1 && r.ue() == 2 && r.ue() == 7 && r.se() == -1 && r.se() == 2;
    check("expgolomb", ok);
    const std::vector<uint8_t> rb = {0x00, 0x00, 0x01, 0x00, 0x00, 0x00, 0x00, 0x03, 0x05};
    const std::vector<uint8_t> nal = nal_from_rbsp(NAL_SPS, rb);
    const std::vector<uint8_t> esc = {0x42, 0x01, 0x00, 0x00, 0x03, 0x01, 0x00, 0x00,
                                      0x03, 0x00, 0x00, 0x03, 0x03, 0x05};
    std::vector<uint8_t> back = rbsp_from_nal(nal.data(), nal.size());
    back.erase(back.begin(), back.begin() + 2);
    check("emulation_prevention", nal == esc && back == rb);
}

// Parameter sets written and parsed back; a 4:2:2 SPS is refused by the
// picture parser with the chroma status.
static void kat_param_sets() {
    Sps s;
    s.chroma_format_idc = 3;
    s.width = 1920;
    s.height = 1088;
    s.conf_win[3] = 8;
    s.bit_depth_luma = s.bit_depth_chroma = 10;
    s.scaling_list_enabled = s.scaling_list_data_present = true;
    s.scaling.set_default();
    s.scaling.e[2][1][5] = 77;
    s.scaling.e[2][1][64] = 40;
    s.scaling.refresh_flat(2, 1);
    std::memcpy(s.scaling.e[3][1], s.scaling.e[2][1], 66);
    s.pcm = true;
    s.pcm_bit_depth_luma = 9;
    s.pcm_bit_depth_chroma = 7;
    s.log2_min_pcm_cb = 3;
    s.log2_max_pcm_cb = 5;
    s.vui = s.colour_description = true;
    s.matrix = 9;
    s.full_range = true;
    s.implicit_rdpcm = s.persistent_rice = true;
    std::vector<uint8_t> nal = write_sps(s);
    Sps t;
    const int st = parse_sps(rbsp_from_nal(nal.data(), nal.size()), t);
    const bool ok = st == 0 && t.chroma_format_idc == 3 && t.width == 1920 && t.height == 1088 &&
                    t.conf_win[3] == 8 && t.bit_depth_luma == 10 && t.pcm && t.pcm_bit_depth_luma == 9 &&
                    t.pcm_bit_depth_chroma == 7 && t.matrix == 9 && t.full_range && t.implicit_rdpcm &&
                    t.persistent_rice && !t.explicit_rdpcm &&
                    std::memcmp(&t.scaling.e[0][0][0], &s.scaling.e[0][0][0], sizeof(s.scaling.e)) == 0;
    check("sps_roundtrip", ok, "status " + std::to_string(st));

    Pps p;
    p.tiles = true;
    p.tile_cols = 3;
    p.tile_rows = 2;
    p.uniform_spacing = false;
    p.col_widths = {5, 9};
    p.row_heights = {4};
    p.cu_qp_delta = true;
    p.diff_cu_qp_delta_depth = 1;
    p.cb_qp_offset = -3;
    p.transform_skip = true;
    p.log2_max_ts = 3;
    p.cross_component = true;
    p.act = true;
    p.deblocking_control = p.deblocking_override = true;
    p.beta_offset_div2 = -2;
    std::vector<uint8_t> pn = write_pps(p, s);
    Pps q;
    const int st2 = parse_pps(rbsp_from_nal(pn.data(), pn.size()), t, q);
    const bool ok2 = st2 == 0 && q.tiles && q.tile_cols == 3 && q.tile_rows == 2 && !q.uniform_spacing &&
                     q.col_widths == p.col_widths && q.row_heights == p.row_heights && q.cu_qp_delta &&
                     q.diff_cu_qp_delta_depth == 1 && q.cb_qp_offset == -3 && q.log2_max_ts == 3 &&
                     q.cross_component && q.act && q.act_cr_qp_offset == -3 && q.deblocking_override &&
                     q.beta_offset_div2 == -2;
    check("pps_roundtrip", ok2, "status " + std::to_string(st2));

    // uniform tile spacing (6-3 / 6-4): 30 CTB columns in 4 tiles
    Pps u;
    u.tiles = true;
    u.tile_cols = 4;
    u.tile_rows = 1;
    std::vector<int> cb, rb;
    tile_boundaries(u, 30, 17, cb, rb);
    check("tile_uniform_spacing", cb == std::vector<int>({0, 7, 15, 22, 30}) && rb == std::vector<int>({0, 17}));

    Sps s422 = s;
    s422.chroma_format_idc = 2;
    s422.scaling_list_enabled = false;
    std::vector<uint8_t> n422 = write_sps(s422);
    Sps t422;
    Pps p0;
    const int st3 = parse_sps(rbsp_from_nal(n422.data(), n422.size()), t422);
    CodedPicture cp;
    cp.slices.push_back({0x28, 0x01, 0xac});
    ParsedPicture pp;
    const hpvcd_status st4 = parse_picture(t422, p0, cp, pp);
    check("chroma_422_refused", st3 == 0 && t422.chroma_format_idc == 2 && st4 == HPVCD_E_UNSUPPORTED_CHROMA,
          "status " + std::to_string((int)st4));
}

// A run of n bypass bins read at once (one long division per 16 bins)
// equals n bins read one at a time, for random streams, ranges after a
// decision and run lengths 0..32, interleaved with context-coded bins.
static void prop_bypass_runs() {
    std::mt19937 rng(777);
    std::vector<uint8_t> data(4096);
    bool ok = true;
    std::string why;
    for (int trial = 0; trial < 200 && ok; trial++) {
        for (auto& b : data) b = (uint8_t)rng();
        CabacDec a, b;
        a.start(data.data(), data.size());
        b.start(data.data(), data.size());
        if (a.offset() >= a.range) continue;  // not a conforming start (9.3.2.5)
        uint8_t ca = (uint8_t)(rng() & 127), cb = ca;
        for (int op = 0; op < 300 && ok; op++) {
            if (rng() % 3 == 0) {
                const int x = a.decode(ca), y = b.decode(cb);
                if (x != y) { ok = false; why = "decision"; }
                continue;
            }
            const int n = (int)(rng() % 33);
            const uint32_t x = a.bypass_bits(n);
            uint32_t y = 0;
            for (int i = 0; i < n; i++) y = (y << 1) | (uint32_t)b.bypass();
            if (x != y || a.offset() != b.offset() || a.pos() != b.pos()) {
                ok = false;
                why = "trial " + std::to_string(trial) + " op " + std::to_string(op) + " n " + std::to_string(n);
            }
        }
    }
    check("bypass_runs", ok, why);
}

int main() {
    prop_bypass_runs();
    kat_pcm_realign();
    kat_bypass_alignment();
    prop_engine_inverse();
    kat_expgolomb();
    kat_param_sets();
    return failures;
}
