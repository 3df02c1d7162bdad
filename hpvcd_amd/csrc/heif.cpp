// heif.cpp — the HEIF container: parse of a HEIC file into reconstruction
// records (the container half of decode_heic), and the generator's HEIC
// writer.
//
// The parse follows the reference's container reading and its decode-side
// rules (src/heif.rs:242 parse, :573 parse_iloc, :670 parse_iprp, :882
// parse_grid_item, :972 build_item, :1116 parse_colr_into, :1140
// irot/imir; src/lib.rs:1497 decode_heic_impl, :1610-1653 grid placement,
// :1724 decode_grid, :1210 hvcC NAL length size, :1349-1382 the item's
// slices; src/config.rs:1284 parse_hvcc_full), including its quirks: only
// the last iloc extent of an item is kept, ipma entries are read with 16-bit
// item ids and 7-bit property indices, a grid's tiles take the first tile's
// conformance window and colour, tiles that fail to parse are left black,
// and a second irot/imir after a first non-identity one is ignored.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hpvcd.h"
#include "gen.h"
#include "gen_hevc.h"
#include "heif.h"
#include "hevc_ps.h"
#include "internal.h"
#include "parse.h"

namespace hpvcd {

namespace {

uint32_t rd16(const uint8_t* p) { return (uint32_t)p[0] << 8 | p[1]; }
uint32_t rd32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

struct Span {
    const uint8_t* p = nullptr;
    size_t n = 0;
};

struct Box {
    char type[5];
    Span body;
};

// boxes at one level: stop at a size below 8, past the end or above the cap
std::vector<Box> boxes(Span s, uint64_t max_box) {
    std::vector<Box> out;
    size_t pos = 0;
    while (pos + 8 <= s.n) {
        const size_t size = rd32(s.p + pos);
        if (size < 8 || pos + size > s.n || size > max_box) break;
        Box b;
        std::memcpy(b.type, s.p + pos + 4, 4);
        b.type[4] = 0;
        b.body = {s.p + pos + 8, size - 8};
        out.push_back(b);
        pos += size;
    }
    return out;
}

bool is(const Box& b, const char* t) { return std::memcmp(b.type, t, 4) == 0; }

uint64_t rdn(Span s, size_t off, int n) {
    if (off + (size_t)n > s.n) return 0;
    uint64_t v = 0;
    if (n != 1 && n != 2 && n != 4 && n != 8) return 0;
    for (int i = 0; i < n; i++) v = (v << 8) | s.p[off + i];
    return v;
}

struct Prop {
    char type[5] = {0, 0, 0, 0, 0};
    Span data;
};

// irot / imir onto an orientation (src/heif.rs:1140-1160): only from the
// identity; a later transform is dropped
uint8_t apply_irot(uint8_t steps, uint8_t prev) {
    if (prev != HPVCD_ORIENT_NORMAL) return prev;
    static const uint8_t m[4] = {HPVCD_ORIENT_NORMAL, HPVCD_ORIENT_ROTATE_270, HPVCD_ORIENT_ROTATE_180,
                                 HPVCD_ORIENT_ROTATE_90};
    return m[steps & 3];
}
uint8_t apply_imir(uint8_t axis, uint8_t prev) {
    if (prev != HPVCD_ORIENT_NORMAL) return prev;
    return (axis & 1) ? HPVCD_ORIENT_FLIP_V : HPVCD_ORIENT_FLIP_H;
}

}  // namespace

// CICP matrix_coefficients as the reference's MatrixCoefficients::from_u8
// reads them: values without a variant are unspecified (2)
uint8_t canonical_matrix(int m) {
    if (m == 0 || m == 1 || (m >= 4 && m <= 17)) return (uint8_t)m;
    return 2;
}

hpvcd_status heif_parse(const uint8_t* file, size_t n, const HeifLimits& L, HeifFile& out) {
    out = HeifFile();
    const Span all{file, n};
    bool brand = false;
    for (const Box& b : boxes(all, L.max_box_size))
        if (is(b, "ftyp")) {
            for (size_t i = 0; i + 4 <= b.body.n; i += 4) {
                const uint8_t* c = b.body.p + i;
                if (!std::memcmp(c, "heic", 4) || !std::memcmp(c, "mif1", 4) || !std::memcmp(c, "miaf", 4) ||
                    !std::memcmp(c, "heif", 4))
                    brand = true;
            }
            break;
        }
    if (!brand) return set_error(HPVCD_E_NOT_HEIF, "not a HEIF file (no heic/mif1/miaf/heif brand)");
    Span meta;
    bool have_meta = false;
    for (const Box& b : boxes(all, L.max_box_size))
        if (is(b, "meta")) {
            meta = b.body;
            have_meta = true;
            break;
        }
    if (!have_meta) return set_error(HPVCD_E_MISSING_BOX, "missing box: meta");
    if (meta.n < 4) return set_error(HPVCD_E_MISSING_BOX, "missing box: meta body");
    meta = {meta.p + 4, meta.n - 4};

    uint32_t pitm = 1;
    Span iloc, iinf, iref, iprp;
    uint64_t idat_off = 0;
    for (const Box& b : boxes(meta, L.max_box_size)) {
        if (is(b, "pitm") && b.body.n >= 4) {
            const Span h{b.body.p + 4, b.body.n - 4};
            if (b.body.p[0] == 1) {
                if (h.n >= 4) pitm = (uint16_t)rd32(h.p);
            } else if (h.n >= 2) {
                pitm = rd16(h.p);
            }
        } else if (is(b, "iloc")) iloc = b.body;
        else if (is(b, "iinf")) iinf = b.body;
        else if (is(b, "iref")) iref = b.body;
        else if (is(b, "iprp")) iprp = b.body;
        else if (is(b, "idat")) idat_off = (uint64_t)(b.body.p - file);
    }
    pitm &= 0xffff;

    // iloc: the last extent of each item (src/heif.rs:573)
    std::map<uint32_t, std::pair<uint64_t, uint64_t>> ext;
    if (iloc.n >= 4) {
        const int version = iloc.p[0];
        const Span d{iloc.p + 4, iloc.n - 4};
        if (d.n >= 4) {
            const int off_size = d.p[0] >> 4, len_size = d.p[0] & 15, base_size = d.p[1] >> 4;
            const int index_size = (version == 1 || version == 2) ? (d.p[1] & 15) : 0;
            size_t count, pos;
            if (version == 2) {
                count = (size_t)rdn(d, 2, 4);
                pos = 6;
            } else {
                count = (size_t)rdn(d, 2, 2);
                pos = 4;
            }
            count = std::min<size_t>(count, L.max_items);
            for (size_t i = 0; i < count && pos < d.n; i++) {
                const size_t before = pos;
                uint32_t id;
                if (version == 2) {
                    id = (uint16_t)rdn(d, pos, 4);
                    pos += 4;
                } else {
                    id = (uint32_t)rdn(d, pos, 2);
                    pos += 2;
                }
                int cm = 0;
                if (version == 1 || version == 2) {
                    cm = (int)(rdn(d, pos, 2) & 15);
                    pos += 2;
                }
                pos += 2;  // data_reference_index
                const uint64_t base = rdn(d, pos, base_size);
                pos += base_size;
                size_t ne = (size_t)rdn(d, pos, 2);
                pos += 2;
                ne = std::min<size_t>(ne, L.max_extents_per_item);
                for (size_t e = 0; e < ne && pos < d.n; e++) {
                    pos += index_size;
                    const uint64_t eo = rdn(d, pos, off_size);
                    pos += off_size;
                    const uint64_t el = rdn(d, pos, len_size);
                    pos += len_size;
                    const uint64_t abs = cm == 1 ? idat_off + base + eo : base + eo;
                    ext[id] = {abs, el};
                }
                if (pos == before) break;
            }
        }
    }
    // iinf: item types (infe v2: id at +12, type at +16..20 of each entry box)
    std::map<uint32_t, std::string> types;
    if (iinf.n >= 4) {
        const Span r{iinf.p + 4, iinf.n - 4};
        size_t count = r.n >= 2 ? rd16(r.p) : 0;
        count = std::min<size_t>(count, L.max_items);
        size_t pos = 2;
        for (size_t i = 0; i < count; i++) {
            if (pos + 8 > r.n) break;
            const size_t bsz = rd32(r.p + pos);
            if (bsz < 12 || pos + bsz > r.n) break;
            if (bsz >= 20) types[rd16(r.p + pos + 12)] = std::string((const char*)r.p + pos + 16, 4);
            pos += bsz;
        }
    }
    // iref: dimg lists (grid tiles)
    std::map<uint32_t, std::vector<uint32_t>> dimg;
    if (iref.n >= 4) {
        const Span r{iref.p + 4, iref.n - 4};
        size_t pos = 0;
        while (pos + 8 <= r.n) {
            const size_t bsz = rd32(r.p + pos);
            if (bsz < 12 || pos + bsz > r.n) break;
            if (!std::memcmp(r.p + pos + 4, "dimg", 4)) {
                const uint32_t from = rd16(r.p + pos + 8);
                const size_t cnt = pos + 12 <= r.n ? rd16(r.p + pos + 10) : 0;
                std::vector<uint32_t> to;
                for (size_t k = 0; k < cnt && pos + 14 + 2 * k <= r.n; k++) to.push_back(rd16(r.p + pos + 12 + 2 * k));
                dimg[from] = to;
            }
            pos += bsz;
        }
    }
    // iprp: ipco properties (1-based), ipma associations
    std::vector<Prop> props(1);
    std::map<uint32_t, std::vector<uint8_t>> assoc;
    for (const Box& b : boxes(iprp, L.max_box_size)) {
        if (is(b, "ipco")) {
            for (const Box& pb : boxes(b.body, L.max_box_size)) {
                if (props.size() >= std::max<size_t>((size_t)L.max_items * 4, 64)) break;
                Prop p;
                std::memcpy(p.type, pb.type, 5);
                p.data = pb.body;
                props.push_back(p);
            }
        }
        if (is(b, "ipma") && b.body.n >= 4) {
            const Span r{b.body.p + 4, b.body.n - 4};
            const uint32_t entries = r.n >= 4 ? rd32(r.p) : 0;
            size_t pos = 4;
            for (uint32_t e = 0; e < entries; e++) {
                if (pos + 3 > r.n) break;
                const uint32_t id = rd16(r.p + pos);
                pos += 2;
                const size_t cnt = r.p[pos++];
                std::vector<uint8_t> idx;
                for (size_t k = 0; k < cnt && pos < r.n; k++) idx.push_back(r.p[pos++] & 0x7f);
                assoc[id] = idx;
            }
        }
    }
    auto item_props = [&](uint32_t id, HeifItem& it, bool need_hvcc) -> hpvcd_status {
        auto a = assoc.find(id);
        if (a != assoc.end())
            for (uint8_t k : a->second) {
                if (k == 0 || k >= props.size()) continue;
                const Prop& p = props[k];
                const std::string t(p.type, 4);
                if (t == "hvcC" && need_hvcc) {
                    if (p.data.n > L.max_hvcc_size) return set_error(HPVCD_E_LIMIT, "hvcC size beyond the limit");
                    it.hvcc.assign(p.data.p, p.data.p + p.data.n);
                } else if (t == "ispe" && need_hvcc) {
                    if (p.data.n >= 12) {
                        it.display_w = rd32(p.data.p + 4);
                        it.display_h = rd32(p.data.p + 8);
                        if (hpvcd_status st = check_image(L, it.display_w, it.display_h)) return st;
                    }
                } else if (t == "colr" && p.data.n >= 4) {
                    if (!std::memcmp(p.data.p, "nclx", 4) && p.data.n >= 11) {
                        it.has_cicp = true;
                        it.cicp_matrix = canonical_matrix(rd16(p.data.p + 8) & 0xff);
                        it.cicp_full_range = (p.data.p[10] & 0x80) != 0;
                    }
                } else if (t == "irot" && p.data.n >= 1) {
                    it.orientation = apply_irot(p.data.p[0] & 3, it.orientation);
                } else if (t == "imir" && p.data.n >= 1) {
                    it.orientation = apply_imir(p.data.p[0] & 1, it.orientation);
                }
            }
        return HPVCD_OK;
    };
    auto build_item = [&](uint32_t id, HeifItem& it) -> hpvcd_status {
        it = HeifItem();
        it.id = id;
        auto e = ext.find(id);
        if (e == ext.end()) return set_error(HPVCD_E_MISSING_BOX, "missing box: iloc entry for item");
        it.offset = e->second.first;
        it.length = e->second.second;
        if (it.length > L.max_item_size) return set_error(HPVCD_E_LIMIT, "item data size beyond the limit");
        if (hpvcd_status st = item_props(id, it, true)) return st;
        if (it.hvcc.empty()) return set_error(HPVCD_E_MISSING_BOX, "missing box: hvcC property");
        return HPVCD_OK;
    };

    auto t = types.find(pitm);
    const bool grid = t != types.end() && t->second == "grid";
    out.is_grid = grid;
    if (!grid) return build_item(pitm, out.primary);
    // grid descriptor (ISO/IEC 23008-12 6.6.2.3): version, flags (bit 0: 32-bit
    // sizes), rows - 1, columns - 1, output width, height
    out.rows = out.cols = 1;
    auto ge = ext.find(pitm);
    if (ge != ext.end()) {
        const uint64_t a = ge->second.first, len = ge->second.second;
        if (a + len >= a && a + len <= n && len >= 4) {
            const uint8_t* d = file + a;
            out.rows = d[2] + 1u;
            out.cols = d[3] + 1u;
            if (d[1] & 1) {
                out.out_w = len >= 8 ? rd32(d + 4) : 0;
                out.out_h = len >= 12 ? rd32(d + 8) : 0;
            } else {
                out.out_w = len >= 6 ? rd16(d + 4) : 0;
                out.out_h = len >= 8 ? rd16(d + 6) : 0;
            }
        }
    }
    if (hpvcd_status st = check_image(L, out.out_w, out.out_h)) return st;
    std::vector<uint32_t> ids = dimg.count(pitm) ? dimg[pitm] : std::vector<uint32_t>();
    if (ids.size() > L.max_tiles) return set_error(HPVCD_E_LIMIT, "grid tiles beyond the limit");
    ids.resize(std::min<size_t>(ids.size(), (size_t)out.rows * out.cols));
    for (uint32_t id : ids) {
        HeifItem it;
        if (build_item(id, it) == HPVCD_OK) out.tiles.push_back(std::move(it));
    }
    HeifItem g;
    item_props(pitm, g, false);
    out.grid_orientation = g.orientation;
    out.primary = out.tiles.empty() ? HeifItem() : out.tiles[0];
    return HPVCD_OK;
}

hpvcd_status check_image(const HeifLimits& L, uint64_t w, uint64_t h) {
    if (w == 0 || h == 0) return set_error(HPVCD_E_BAD_DIMENSIONS, "zero image dimensions");
    if (w > L.max_dimension || h > L.max_dimension) return set_error(HPVCD_E_LIMIT, "image dimension beyond the limit");
    if (w * h > L.max_pixels) return set_error(HPVCD_E_LIMIT, "image pixel count beyond the limit");
    return HPVCD_OK;
}

hpvcd_status parse_hvcc(const std::vector<uint8_t>& hvcc, Sps& sps, Pps& pps) {
    if (hvcc.size() < 23) return set_error(HPVCD_E_PARAMSET, "hvcC too short");
    const size_t arrays = hvcc[22];
    size_t pos = 23;
    std::vector<uint8_t> s_rb, p_rb;
    bool have_s = false, have_p = false;
    for (size_t a = 0; a < arrays; a++) {
        if (pos + 3 > hvcc.size()) break;
        const int type = hvcc[pos] & 0x3f;
        const size_t cnt = rd16(&hvcc[pos + 1]);
        pos += 3;
        for (size_t k = 0; k < cnt; k++) {
            if (pos + 2 > hvcc.size()) break;
            const size_t len = rd16(&hvcc[pos]);
            pos += 2;
            if (pos + len > hvcc.size()) break;
            if (len >= 2) {
                if (type == NAL_SPS) {
                    s_rb = rbsp_from_nal(&hvcc[pos], len);
                    have_s = true;
                } else if (type == NAL_PPS) {
                    p_rb = rbsp_from_nal(&hvcc[pos], len);
                    have_p = true;
                }
            }
            pos += len;
        }
    }
    if (!have_s) return set_error(HPVCD_E_PARAMSET, "hvcC without SPS");
    if (!have_p) return set_error(HPVCD_E_PARAMSET, "hvcC without PPS");
    if (int st = parse_sps(s_rb, sps)) return (hpvcd_status)st;
    if (int st = parse_pps(p_rb, sps, pps)) return (hpvcd_status)st;
    return HPVCD_OK;
}

// One item's coded picture: the slices of its length-prefixed sample
// (src/lib.rs:1349-1382)
hpvcd_status item_picture(const uint8_t* file, size_t n, const HeifItem& it, const std::vector<uint8_t>& hvcc,
                          ParsedPicture& out, Sps& sps) {
    if (it.offset > n || it.length > n - it.offset) return set_error(HPVCD_E_BITSTREAM, "image data extends past file end");
    Pps pps;
    if (hpvcd_status st = parse_hvcc(hvcc, sps, pps)) return st;
    const int len_size = (hvcc[21] & 3) + 1;
    std::vector<std::pair<const uint8_t*, size_t>> nals;
    if (len_size == 3) {  // split_nals takes 1, 2 and 4; 3-byte lengths here
        size_t pos = 0;
        const uint8_t* s = file + it.offset;
        while (pos < it.length) {
            if (it.length - pos < 3) return set_error(HPVCD_E_BITSTREAM, "truncated HEIF NAL length prefix");
            const size_t m = (size_t)s[pos] << 16 | (size_t)s[pos + 1] << 8 | s[pos + 2];
            pos += 3;
            if (m > it.length - pos) return set_error(HPVCD_E_BITSTREAM, "truncated HEIF NAL payload");
            nals.emplace_back(s + pos, m);
            pos += m;
        }
    } else if (hpvcd_status st = split_nals(file + it.offset, (size_t)it.length, len_size, nals)) {
        return st;
    }
    CodedPicture cp;
    for (const auto& nal : nals) {
        if (nal.second < 2) return set_error(HPVCD_E_BITSTREAM, "HEIF sample contains an undersized NAL unit");
        const int type = (nal.first[0] >> 1) & 63;
        if (type > 31) continue;
        cp.removed.emplace_back();
        cp.slices.push_back(rbsp_from_nal(nal.first, nal.second, &cp.removed.back()));
    }
    if (cp.slices.empty()) return set_error(HPVCD_E_BITSTREAM, "no VCL slice NAL found");
    if (cp.slices[0].size() < 3 || !(cp.slices[0][2] & 0x80))
        return set_error(HPVCD_E_BITSTREAM, "HEIF item does not begin with an independent first slice");
    return parse_picture(sps, pps, cp, out);
}

namespace {

int default_kind(int chroma_idc, bool deep, int want) {
    const bool mono = chroma_idc == 0;
    if (want == HEIC_KIND_RGB8) return deep ? HPVCD_OUT_RGB8_FROM16 : HPVCD_OUT_RGB8;
    if (want == HEIC_KIND_YUV)
        return (mono ? HPVCD_OUT_YUV400_8 : (chroma_idc == 3 ? HPVCD_OUT_YUV444_8 : HPVCD_OUT_YUV420_8)) + (deep ? 1 : 0);
    if (want >= 0) return want;
    return mono ? (deep ? HPVCD_OUT_LUMA16 : HPVCD_OUT_LUMA8) : (deep ? HPVCD_OUT_RGB16 : HPVCD_OUT_RGB8);
}

hpvcd_output make_output(int w, int h, int kind, uint8_t orientation) {
    hpvcd_output o;
    std::memset(&o, 0, sizeof(o));
    o.width = w;
    o.height = h;
    o.kind = (uint8_t)kind;
    if (HPVCD_OUT_IS_YUV(kind)) {
        o.channels = HPVCD_OUT_YUV_CHROMA(kind) ? 3 : 1;
        o.bytes_per_sample = HPVCD_OUT_YUV_WIDE(kind) ? 2 : 1;
    } else {
        o.channels = (kind == HPVCD_OUT_LUMA8 || kind == HPVCD_OUT_LUMA16) ? 1 : 3;
        o.bytes_per_sample = (kind == HPVCD_OUT_RGB16 || kind == HPVCD_OUT_LUMA16) ? 2 : 1;
    }
    o.orientation = orientation == HPVCD_ORIENT_NORMAL ? 0 : orientation;
    return o;
}

void set_colour(hpvcd_picture& p, uint8_t matrix, bool full) {
    p.matrix = matrix;
    p.flags = (p.flags & ~HPVCD_PF_FULL_RANGE) | (full ? HPVCD_PF_FULL_RANGE : 0u);
}

}  // namespace

hpvcd_status heic_records(const uint8_t* file, size_t n, const HeifLimits& L, int threads, int want_kind,
                          hpvcd_records** out, HeicInfo* info) {
    *out = nullptr;
    HeifFile hf;
    if (hpvcd_status st = heif_parse(file, n, L, hf)) return st;
    std::vector<ParsedPicture> pics;
    std::vector<hpvcd_output> outs;
    HeicInfo inf;
    if (!hf.is_grid) {
        // single item (src/lib.rs:1497-1560): display size from ispe, the
        // window at the conformance-window offset, VUI colour unless its
        // matrix is unspecified, then colr, then sRGB
        if (hpvcd_status st = check_image(L, hf.primary.display_w, hf.primary.display_h)) return st;
        pics.resize(1);
        Sps sps;
        if (hpvcd_status st = item_picture(file, n, hf.primary, hf.primary.hvcc, pics[0], sps)) return st;
        hpvcd_picture& p = pics[0].pic;
        const int dw = (int)hf.primary.display_w, dh = (int)hf.primary.display_h;
        p.win_w = p.copy_w = dw;
        p.win_h = p.copy_h = dh;
        p.dst_x = p.dst_y = 0;
        p.out_index = 0;
        uint8_t m = canonical_matrix(sps.matrix);
        bool full = sps.full_range;
        if (m == 2) {
            if (hf.primary.has_cicp) {
                m = hf.primary.cicp_matrix;
                full = hf.primary.cicp_full_range;
            } else {
                m = 1;
                full = true;
            }
        }
        set_colour(p, m, full);
        const int kind = default_kind(p.chroma_idc, p.bit_depth_luma > 8, want_kind);
        p.out_kind = (uint8_t)kind;
        outs.push_back(make_output(dw, dh, kind, hf.primary.orientation));
        inf.orientation = hf.primary.orientation;
        inf.tiles = 1;
    } else {
        // grid (src/lib.rs:1724-1870)
        if (hf.out_w == 0 || hf.out_h == 0 || hf.tiles.empty())
            return set_error(HPVCD_E_BITSTREAM, "grid has zero size or no tiles");
        const std::vector<uint8_t>* fallback = nullptr;
        for (const HeifItem& t : hf.tiles)
            if (!t.hvcc.empty()) {
                fallback = &t.hvcc;
                break;
            }
        const std::vector<uint8_t>& h0 = !hf.tiles[0].hvcc.empty() ? hf.tiles[0].hvcc : *fallback;
        Sps s0;
        Pps p0;
        const bool have0 = parse_hvcc(h0, s0, p0) == HPVCD_OK;
        int tile_w = 0, tile_h = 0, crop_l = 0, crop_t = 0;
        if (have0) {
            tile_w = std::max(0, s0.width - s0.conf_win[0] - s0.conf_win[1]);
            tile_h = std::max(0, s0.height - s0.conf_win[2] - s0.conf_win[3]);
            crop_l = s0.conf_win[0];
            crop_t = s0.conf_win[2];
        }
        if (tile_w <= 0 || tile_h <= 0) {
            const HeifItem& t = hf.tiles[0];
            const bool same = t.display_w == hf.out_w && t.display_h == hf.out_h && (hf.cols > 1 || hf.rows > 1);
            // raw planes take the grid's share without the display fallback
            // (src/lib.rs:1121-1127)
            if (t.display_w && t.display_h && !same && want_kind != HEIC_KIND_YUV) {
                tile_w = (int)t.display_w;
                tile_h = (int)t.display_h;
            } else {
                tile_w = (int)((hf.out_w + hf.cols - 1) / hf.cols);
                tile_h = (int)((hf.out_h + hf.rows - 1) / hf.rows);
            }
        }
        const uint8_t matrix = have0 ? canonical_matrix(s0.matrix) : 1;
        const bool full = have0 ? s0.full_range : true;
        const bool mono = have0 && s0.chroma_format_idc == 0;
        const bool deep = have0 && s0.bit_depth_luma > 8;
        const int chroma0 = have0 ? s0.chroma_format_idc : 1;
        const int kind = default_kind(chroma0, deep, want_kind);
        const size_t nt = hf.tiles.size();
        std::vector<ParsedPicture> all(nt);
        std::vector<hpvcd_status> sts(nt, HPVCD_OK);
        // tiles parse on the decoder's worker threads
        // (each thread takes the next unparsed tile: tiles of unequal size
        // balance out)
        const unsigned nth = (unsigned)std::max(1, std::min<int>(threads, (int)nt));
        std::atomic<size_t> next{0};
        std::vector<std::thread> th;
        for (unsigned w = 0; w < nth; w++)
            th.emplace_back([&]() {
                for (size_t i = next++; i < nt; i = next++) {
                    const HeifItem& t = hf.tiles[i];
                    const std::vector<uint8_t>& hv = !t.hvcc.empty() ? t.hvcc : *fallback;
                    Sps s;
                    sts[i] = item_picture(file, n, t, hv, all[i], s);
                }
            });
        for (auto& x : th) x.join();
        for (size_t i = 0; i < nt; i++) {
            if (sts[i] == HPVCD_E_ALLOC) return set_error(HPVCD_E_ALLOC, "out of memory parsing a grid tile");
            if (sts[i] != HPVCD_OK) {
                inf.failed_tiles++;
                continue;
            }
            hpvcd_picture& p = all[i].pic;
            // a tile of another format than the first is not composed
            if ((p.chroma_idc == 0) != mono || (p.bit_depth_luma > 8) != deep ||
                (HPVCD_OUT_IS_YUV(kind) && p.chroma_idc != chroma0)) {
                inf.failed_tiles++;
                continue;
            }
            const int row = (int)(i / hf.cols), col = (int)(i % hf.cols);
            const int dx = col * tile_w, dy = row * tile_h;
            const int cw = std::min(tile_w, std::max(0, (int)hf.out_w - dx));
            const int ch = std::min(tile_h, std::max(0, (int)hf.out_h - dy));
            if (cw <= 0 || ch <= 0) continue;
            p.crop_left = crop_l;
            p.crop_top = crop_t;
            p.win_w = tile_w;
            p.win_h = tile_h;
            p.dst_x = dx;
            p.dst_y = dy;
            p.copy_w = cw;
            p.copy_h = ch;
            p.out_index = 0;
            p.out_kind = (uint8_t)kind;
            set_colour(p, matrix, full);
            pics.push_back(std::move(all[i]));
        }
        outs.push_back(make_output((int)hf.out_w, (int)hf.out_h, kind, hf.grid_orientation));
        inf.orientation = hf.grid_orientation;
        inf.tiles = (int)nt;
        inf.rows = (int)hf.rows;
        inf.cols = (int)hf.cols;
    }
    if (info) *info = inf;
    return assemble_records(pics, outs, out);
}

// ---------------------------------------------------------------------------
// writer
// ---------------------------------------------------------------------------
namespace {

struct Out {
    std::vector<uint8_t> b;
    void u8(uint32_t v) { b.push_back((uint8_t)v); }
    void u16(uint32_t v) {
        u8(v >> 8);
        u8(v);
    }
    void u32(uint32_t v) {
        u16(v >> 16);
        u16(v & 0xffff);
    }
    void tag(const char* t) { b.insert(b.end(), t, t + 4); }
    void bytes(const std::vector<uint8_t>& v) { b.insert(b.end(), v.begin(), v.end()); }
    // box with the payload `body`
    void box(const char* t, const std::vector<uint8_t>& body) {
        u32((uint32_t)(8 + body.size()));
        tag(t);
        bytes(body);
    }
    void fullbox(const char* t, int version, uint32_t flags, const std::vector<uint8_t>& body) {
        u32((uint32_t)(12 + body.size()));
        tag(t);
        u8(version);
        u8(flags >> 16);
        u16(flags & 0xffff);
        bytes(body);
    }
};

// HEVCDecoderConfigurationRecord (ISO/IEC 14496-15 8.3.3.1) with one array
// each of VPS, SPS and PPS, 4-byte NAL lengths
std::vector<uint8_t> hvcc_record(const Sps& s, const std::vector<std::vector<uint8_t>>& ps) {
    Out o;
    o.u8(1);  // configurationVersion
    // general_profile_space .. general_level_idc as the SPS codes them
    // (RBSP bytes 3..14: after the NAL header and the sub-layer byte)
    const std::vector<uint8_t> sps = rbsp_from_nal(ps[1].data(), ps[1].size());
    for (int i = 0; i < 12; i++) o.u8(sps.size() > 3 + (size_t)i ? sps[3 + i] : 0);
    o.u16(0xf000);                               // min_spatial_segmentation_idc 0
    o.u8(0xfc);                                  // parallelismType 0
    o.u8(0xfc | (uint32_t)s.chroma_format_idc);  // chromaFormat
    o.u8(0xf8 | (uint32_t)(s.bit_depth_luma - 8));
    o.u8(0xf8 | (uint32_t)(s.bit_depth_chroma - 8));
    o.u16(0);                                    // avgFrameRate
    o.u8(0x0f);                                  // constantFrameRate 0, numTemporalLayers 1, nested 1, lengthSizeMinusOne 3
    o.u8(3);                                     // numOfArrays
    const uint8_t types[3] = {NAL_VPS, NAL_SPS, NAL_PPS};
    for (int a = 0; a < 3; a++) {
        o.u8(0x80 | types[a]);  // array_completeness
        o.u16(1);
        o.u16((uint32_t)ps[a].size());
        o.bytes(ps[a]);
    }
    return o.b;
}

}  // namespace

hpvcd_status gen_heic(const hpvcd_gen_config& c, std::vector<uint8_t>& file) {
    if (c.orientation == HPVCD_ORIENT_TRANSPOSE || c.orientation == HPVCD_ORIENT_TRANSVERSE)
        return set_error(HPVCD_E_UNSUPPORTED, "HEIC writer: transpose and transverse have no irot/imir form");
    if (c.num_outputs > 1) return set_error(HPVCD_E_UNSUPPORTED, "HEIC writer: one image per file");
    if (c.inter) return set_error(HPVCD_E_UNSUPPORTED, "HEIC writer: image items are intra pictures");
    std::vector<PicRecords> recs;
    std::vector<PicStream> streams;
    if (hpvcd_status st = gen_coded(c, recs, streams)) return st;
    Sps sps;
    Pps pps;
    gen_parameter_sets(c, sps, pps);
    const std::vector<std::vector<uint8_t>> ps = {write_vps(sps), write_sps(sps), write_pps(pps, sps)};
    const std::vector<uint8_t> hvcc = hvcc_record(sps, ps);
    const uint32_t ntiles = (uint32_t)streams.size();
    const int win_w = c.tile_width - c.conf_win[0] - c.conf_win[1];
    const int win_h = c.tile_height - c.conf_win[2] - c.conf_win[3];
    // one picture whose output is its window: a plain hvc1 item; otherwise a grid
    const bool grid = ntiles > 1 || c.out_width != win_w || c.out_height != win_h;
    // samples: length-prefixed slice NAL units
    std::vector<std::vector<uint8_t>> samples(ntiles);
    for (uint32_t i = 0; i < ntiles; i++)
        for (const auto& nal : gen_slice_nals(c, sps, pps, i, streams[i])) {
            Out o;
            o.u32((uint32_t)nal.size());
            o.bytes(nal);
            samples[i].insert(samples[i].end(), o.b.begin(), o.b.end());
        }
    const uint32_t grid_id = ntiles + 1, pitm = grid ? grid_id : 1;
    // the grid descriptor (in idat)
    Out gd;
    const bool wide = c.out_width > 65535 || c.out_height > 65535;
    gd.u8(0);
    gd.u8(wide ? 1 : 0);
    gd.u8((uint32_t)c.grid_rows - 1);
    gd.u8((uint32_t)c.grid_cols - 1);
    if (wide) {
        gd.u32((uint32_t)c.out_width);
        gd.u32((uint32_t)c.out_height);
    } else {
        gd.u16((uint32_t)c.out_width);
        gd.u16((uint32_t)c.out_height);
    }
    if (grid && (c.grid_rows > 256 || c.grid_cols > 256))
        return set_error(HPVCD_E_UNSUPPORTED, "HEIC writer: more than 256 grid rows or columns");

    auto build_meta = [&](uint32_t mdat_payload) -> std::vector<uint8_t> {
        Out meta;
        {
            Out h;
            h.u32(0);
            h.tag("pict");
            for (int i = 0; i < 12; i++) h.u8(0);
            h.u8(0);  // empty name
            meta.fullbox("hdlr", 0, 0, h.b);
        }
        {
            Out p;
            p.u16(pitm);
            meta.fullbox("pitm", 0, 0, p.b);
        }
        {
            Out inf;
            inf.u16(ntiles + (grid ? 1 : 0));
            for (uint32_t i = 1; i <= ntiles + (grid ? 1u : 0u); i++) {
                Out e;
                e.u16(i);
                e.u16(0);
                e.tag(i == grid_id && grid ? "grid" : "hvc1");
                e.u8(0);
                inf.fullbox("infe", 2, i <= ntiles && grid ? 1 : 0, e.b);  // tiles are hidden items
            }
            meta.fullbox("iinf", 0, 0, inf.b);
        }
        if (grid) {
            Out r;
            Out d;
            d.u16(grid_id);
            d.u16(ntiles);
            for (uint32_t i = 1; i <= ntiles; i++) d.u16(i);
            r.box("dimg", d.b);
            meta.fullbox("iref", 0, 0, r.b);
        }
        {
            // properties: 1 hvcC, 2 ispe (tile window), 3 colr nclx, [4 ispe (grid)], [irot / imir]
            Out ipco;
            ipco.box("hvcC", hvcc);
            Out is;
            is.u32(0);
            is.u32((uint32_t)win_w);
            is.u32((uint32_t)win_h);
            ipco.box("ispe", is.b);
            Out colr;
            colr.tag("nclx");
            colr.u16(2);
            colr.u16(2);
            colr.u16(c.matrix);
            colr.u8(c.full_range ? 0x80 : 0);
            ipco.box("colr", colr.b);
            int next = 4, ispe_grid = 0, orient = 0;
            if (grid) {
                Out ig;
                ig.u32(0);
                ig.u32((uint32_t)c.out_width);
                ig.u32((uint32_t)c.out_height);
                ipco.box("ispe", ig.b);
                ispe_grid = next++;
            }
            const int o = c.orientation;
            if (o == HPVCD_ORIENT_FLIP_H || o == HPVCD_ORIENT_FLIP_V) {
                Out m;
                m.u8(o == HPVCD_ORIENT_FLIP_V ? 1 : 0);
                ipco.box("imir", m.b);
                orient = next++;
            } else if (o == HPVCD_ORIENT_ROTATE_90 || o == HPVCD_ORIENT_ROTATE_180 || o == HPVCD_ORIENT_ROTATE_270) {
                Out r;
                r.u8(o == HPVCD_ORIENT_ROTATE_270 ? 1 : (o == HPVCD_ORIENT_ROTATE_180 ? 2 : 3));
                ipco.box("irot", r.b);
                orient = next++;
            }
            Out ipma;
            const uint32_t entries = ntiles + (grid ? 1 : 0);
            ipma.u32(entries);
            for (uint32_t i = 1; i <= ntiles; i++) {
                ipma.u16(i);
                std::vector<uint8_t> a = {0x81, 2};
                if (!grid) {
                    a.push_back(3);
                    if (orient) a.push_back((uint8_t)(0x80 | orient));
                }
                ipma.u8((uint32_t)a.size());
                ipma.bytes(a);
            }
            if (grid) {
                ipma.u16(grid_id);
                std::vector<uint8_t> a = {(uint8_t)ispe_grid, 3};
                if (orient) a.push_back((uint8_t)(0x80 | orient));
                ipma.u8((uint32_t)a.size());
                ipma.bytes(a);
            }
            Out iprp;
            iprp.box("ipco", ipco.b);
            iprp.fullbox("ipma", 0, 0, ipma.b);
            meta.box("iprp", iprp.b);
        }
        {
            // iloc v1: 4-byte offsets and lengths, no base offset; tiles in
            // mdat (construction method 0), the grid descriptor in idat (1)
            Out l;
            l.u8(0x44);
            l.u8(0x00);
            l.u16(ntiles + (grid ? 1 : 0));
            uint32_t off = mdat_payload;
            for (uint32_t i = 1; i <= ntiles; i++) {
                l.u16(i);
                l.u16(0);  // construction_method 0
                l.u16(0);  // data_reference_index
                l.u16(1);
                l.u32(off);
                l.u32((uint32_t)samples[i - 1].size());
                off += (uint32_t)samples[i - 1].size();
            }
            if (grid) {
                l.u16(grid_id);
                l.u16(1);
                l.u16(0);
                l.u16(1);
                l.u32(0);
                l.u32((uint32_t)gd.b.size());
            }
            meta.fullbox("iloc", 1, 0, l.b);
        }
        if (grid) meta.box("idat", gd.b);
        Out m;
        m.fullbox("meta", 0, 0, meta.b);
        return m.b;
    };
    Out ftyp;
    ftyp.tag("heic");
    ftyp.u32(0);
    ftyp.tag("mif1");
    ftyp.tag("heic");
    Out head;
    head.box("ftyp", ftyp.b);
    size_t total = 0;
    for (const auto& s : samples) total += s.size();
    if (total > 0xffffffffull - 65536) return set_error(HPVCD_E_UNSUPPORTED, "HEIC writer: file beyond 4 GiB");
    // the meta size does not depend on the offsets it holds (fixed-width fields)
    const size_t meta_size = build_meta(0).size();
    const uint32_t mdat_payload = (uint32_t)(head.b.size() + meta_size + 8);
    const std::vector<uint8_t> meta = build_meta(mdat_payload);
    file = head.b;
    file.insert(file.end(), meta.begin(), meta.end());
    Out mh;
    mh.u32((uint32_t)(8 + total));
    mh.tag("mdat");
    file.insert(file.end(), mh.b.begin(), mh.b.end());
    for (const auto& s : samples) file.insert(file.end(), s.begin(), s.end());
    return HPVCD_OK;
}

HeifLimits limits_of(const hpvcd_settings* s) {
    HeifLimits L;
    if (s) {
        L.max_dimension = s->max_dimension;
        L.max_pixels = s->max_pixels;
        L.max_box_size = s->max_box_size;
        L.max_item_size = s->max_item_size;
        L.max_items = s->max_items;
        L.max_tiles = s->max_tiles;
    }
    return L;
}

}  // namespace hpvcd

using namespace hpvcd;

extern "C" hpvcd_status hpvcd_gen_heic(const hpvcd_gen_config* cfg, uint8_t** data, size_t* size) {
    if (!cfg || !data || !size) return set_error(HPVCD_E_INVALID, "null argument");
    *data = nullptr;
    *size = 0;
    std::vector<uint8_t> f;
    if (hpvcd_status st = gen_heic(*cfg, f)) return st;
    *data = copy_out(f, size);
    return *data ? HPVCD_OK : set_error(HPVCD_E_ALLOC, "HEIC writer: out of memory");
}

extern "C" hpvcd_status hpvcd_parse_heic(const uint8_t* data, size_t size, const hpvcd_settings* settings,
                                         int32_t out_kind, hpvcd_records** out) {
    if (!data || !out) return set_error(HPVCD_E_INVALID, "null argument");
    const int threads = settings && settings->threads > 0 ? settings->threads
                                                          : (int)std::max(1u, std::thread::hardware_concurrency());
    return heic_records(data, size, limits_of(settings), threads, out_kind, out, nullptr);
}
