// hash.cpp — TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// The decoded picture hashes of hash::verify (src/hevc/hash.rs:109-158),
// restated over one component's message (its samples in raster order as
// bytes, low byte first above 8 bits): the bit-serial CRC-16 (:160-174),
// the position-masked checksum (:176-189) and MD5 (:191-250). The MD5
// constants are computed as the reference computes them, from sin() in
// double precision.
#include <cmath>
#include <cstring>

#include "oracle.h"

extern "C" uint16_t orc_crc16(const uint8_t* data, size_t n) {
    uint32_t crc = 0xFFFF;
    for (size_t i = 0; i < n; ++i)
        for (int bit = 0; bit < 8; ++bit) {
            const uint32_t msb = (crc >> 15) & 1, v = (data[i] >> (7 - bit)) & 1;
            crc = (((crc << 1) + v) & 0xFFFF) ^ (msb * 0x1021);
        }
    for (int k = 0; k < 16; ++k) {
        const uint32_t msb = (crc >> 15) & 1;
        crc = ((crc << 1) & 0xFFFF) ^ (msb * 0x1021);
    }
    return (uint16_t)crc;
}

extern "C" uint32_t orc_checksum(const uint8_t* data, size_t n, int width, int wide) {
    const size_t per = wide ? 2 : 1;
    uint32_t sum = 0;
    for (size_t i = 0; i * per < n; ++i) {
        const size_t x = i % (size_t)width, y = i / (size_t)width;
        const uint32_t mask = (uint32_t)((x & 0xFF) ^ (y & 0xFF) ^ (x >> 8) ^ (y >> 8));
        sum += data[i * per] ^ mask;
        if (wide) sum += data[i * per + 1] ^ mask;
    }
    return sum;
}

extern "C" void orc_md5(const uint8_t* data, size_t n, uint8_t out[16]) {
    static const uint32_t S[64] = {7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22,
                                   5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20,
                                   4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23,
                                   6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21};
    uint32_t K[64];
    for (int i = 0; i < 64; ++i) K[i] = (uint32_t)(std::fabs(std::sin((double)i + 1.0)) * 4294967296.0);
    uint32_t a0 = 0x67452301, b0 = 0xefcdab89, c0 = 0x98badcfe, d0 = 0x10325476;
    // Padded message: 0x80, zeros to 56 mod 64, the bit length (LE u64).
    size_t total = n + 1;
    while (total % 64 != 56) ++total;
    total += 8;
    const uint64_t bits = (uint64_t)n * 8;
    auto byte_at = [&](size_t j) -> uint8_t {
        if (j < n) return data[j];
        if (j == n) return 0x80;
        if (j >= total - 8) return (uint8_t)(bits >> (8 * (j - (total - 8))));
        return 0;
    };
    for (size_t blk = 0; blk < total; blk += 64) {
        uint32_t m[16];
        for (int i = 0; i < 16; ++i)
            m[i] = (uint32_t)byte_at(blk + 4 * i) | (uint32_t)byte_at(blk + 4 * i + 1) << 8 |
                   (uint32_t)byte_at(blk + 4 * i + 2) << 16 | (uint32_t)byte_at(blk + 4 * i + 3) << 24;
        uint32_t a = a0, b = b0, c = c0, d = d0;
        for (int i = 0; i < 64; ++i) {
            uint32_t f;
            int g;
            switch (i / 16) {
                case 0: f = (b & c) | (~b & d), g = i; break;
                case 1: f = (d & b) | (~d & c), g = (5 * i + 1) % 16; break;
                case 2: f = b ^ c ^ d, g = (3 * i + 5) % 16; break;
                default: f = c ^ (b | ~d), g = (7 * i) % 16; break;
            }
            const uint32_t f2 = f + a + K[i] + m[g];
            a = d, d = c, c = b;
            b = b + ((f2 << S[i]) | (f2 >> (32 - S[i])));
        }
        a0 += a, b0 += b, c0 += c, d0 += d;
    }
    const uint32_t st[4] = {a0, b0, c0, d0};
    for (int i = 0; i < 16; ++i) out[i] = (uint8_t)(st[i / 4] >> (8 * (i % 4)));
}
