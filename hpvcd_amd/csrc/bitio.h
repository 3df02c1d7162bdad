// bitio.h — RBSP bit writer/reader, Exp-Golomb codes and NAL emulation
// prevention (H.265 7.2, 7.3.1.1, 9.2).  The reader mirrors the reference's
// BitReader (src/bitreader.rs) over an unescaped RBSP.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace hpvcd {

struct BitWriter {
    std::vector<uint8_t> buf;
    uint32_t acc = 0;  // pending bits, left-aligned in the low `n` bits
    int n = 0;

    void bit(uint32_t b) {
        acc = (acc << 1) | (b & 1u);
        if (++n == 8) {
            buf.push_back((uint8_t)acc);
            acc = 0;
            n = 0;
        }
    }
    void bits(uint32_t v, int count) {
        for (int i = count - 1; i >= 0; i--) bit((v >> i) & 1u);
    }
    void flag(bool f) { bit(f ? 1u : 0u); }
    void ue(uint32_t v) {  // 9.2 ue(v)
        const uint64_t x = (uint64_t)v + 1;
        int len = 0;
        while ((x >> (len + 1)) != 0) len++;
        bits(0, len);
        bits((uint32_t)x, len + 1);
    }
    void se(int32_t v) { ue(v > 0 ? (uint32_t)(2 * v - 1) : (uint32_t)(-2 * (int64_t)v)); }
    bool aligned() const { return n == 0; }
    void align_zero() {
        while (n) bit(0);
    }
    // byte_alignment() / rbsp_trailing_bits(): a one bit, then zeros
    void align_one_zero() {
        bit(1);
        align_zero();
    }
    size_t bit_count() const { return buf.size() * 8 + (size_t)n; }
};

struct BitReader {
    const uint8_t* p = nullptr;
    size_t len = 0;    // bytes
    size_t pos = 0;    // bits consumed
    bool overrun = false;

    BitReader() = default;
    BitReader(const uint8_t* data, size_t bytes) : p(data), len(bytes) {}

    uint32_t bit() {
        if (pos >= len * 8) {
            overrun = true;
            pos++;
            return 0;
        }
        uint32_t b = (p[pos >> 3] >> (7 - (pos & 7))) & 1u;
        pos++;
        return b;
    }
    uint32_t bits(int count) {
        uint32_t v = 0;
        for (int i = 0; i < count; i++) v = (v << 1) | bit();
        return v;
    }
    bool flag() { return bit() != 0; }
    uint32_t ue() {
        int zeros = 0;
        while (!bit()) {
            if (++zeros > 31 || overrun) {
                overrun = true;
                return 0;
            }
        }
        if (!zeros) return 0;
        const uint64_t v = ((uint64_t)1 << zeros) - 1 + bits(zeros);
        return v > 0xffffffffull ? 0xffffffffu : (uint32_t)v;
    }
    int32_t se() {
        const uint32_t k = ue();
        return (k & 1) ? (int32_t)((k + 1) / 2) : -(int32_t)(k / 2);
    }
    void skip_to_byte() { pos = (pos + 7) & ~(size_t)7; }
    size_t byte_pos() const { return (pos + 7) >> 3; }
    size_t bits_left() const { return pos >= len * 8 ? 0 : len * 8 - pos; }
};

// RBSP -> NAL payload with emulation_prevention_three_byte after every
// 0x0000 followed by a byte <= 3 (7.4.2).  `hdr` is the two-byte NAL header.
inline std::vector<uint8_t> nal_from_rbsp(uint8_t type, const std::vector<uint8_t>& rbsp) {
    std::vector<uint8_t> out;
    out.reserve(rbsp.size() + rbsp.size() / 64 + 4);
    out.push_back((uint8_t)(type << 1));  // forbidden_zero_bit 0, nal_unit_type, nuh_layer_id 0
    out.push_back(1);                     // nuh_temporal_id_plus1
    int zeros = 0;
    for (uint8_t b : rbsp) {
        if (zeros >= 2 && b <= 3) {
            out.push_back(3);
            zeros = 0;
        }
        out.push_back(b);
        zeros = b == 0 ? zeros + 1 : 0;
    }
    return out;
}

// NAL payload (header included) -> RBSP without the emulation prevention
// bytes; the two header bytes are kept at the front.  `removed`, when given,
// receives the NAL positions of the dropped bytes.
inline std::vector<uint8_t> rbsp_from_nal(const uint8_t* nal, size_t n, std::vector<uint32_t>* removed = nullptr) {
    std::vector<uint8_t> out;
    out.reserve(n);
    int zeros = 0;
    for (size_t i = 0; i < n; i++) {
        const uint8_t b = nal[i];
        if (i >= 2 && zeros >= 2 && b == 3) {
            zeros = 0;
            if (removed) removed->push_back((uint32_t)i);
            continue;
        }
        out.push_back(b);
        zeros = (i >= 2 && b == 0) ? zeros + 1 : 0;
    }
    return out;
}

}  // namespace hpvcd
