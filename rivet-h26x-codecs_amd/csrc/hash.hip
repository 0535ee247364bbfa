// Decoded picture hash (H.265 D.3.20) of one picture's planes on the
// device: the hashes of hash::verify (src/hevc/hash.rs:109-158). Each
// component's message is its plane's samples in raster order as bytes, the
// low byte first when a sample takes two (depth above 8), without the row
// padding of the device planes.
//
// - Checksum (:176-189): a sum of independent terms, one workgroup per
//   row and an atomic add per workgroup.
// - CRC-16 (:160-174): the reference shifts the message bits into a 16-bit
//   register seeded 0xFFFF and flushes it with 16 zero bits. Over GF(2)
//   that register is x^16 (0xFFFF x^(8n) + M(x)) mod P, with M the message
//   as a polynomial, n its length in bytes and P = x^16 + x^12 + x^5 + 1.
//   M(x) mod P is linear in the message, so each lane reduces a piece of a
//   row, shifts it by x^(8 k) mod P (k = bytes after the piece) and the
//   pieces meet by XOR; the rows likewise.
// - MD5 (:191-250): serial within a message. One workgroup per component
//   stages 4 KiB of the padded message in LDS at a time and one lane runs
//   the 64-step compression over its blocks.
#include <hip/hip_runtime.h>

#include "rv_kernels.h"

namespace rv {
namespace {

// ---- checksum --------------------------------------------------------

__global__ void __launch_bounds__(64) hash_zero(uint32_t* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* red) {
    for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// One workgroup per (row, component): sum over the row of each byte XOR
// its sample's position mask (x & 0xFF) ^ (y & 0xFF) ^ (x >> 8) ^ (y >> 8),
// wrapping at 2^32.
__global__ void __launch_bounds__(256) hash_checksum(HashPlanes H, uint32_t* out) {
    __shared__ uint32_t red[4];
    const int c = blockIdx.y, y = blockIdx.x;
    if (y >= H.h[c]) return;
    const uint8_t* row = H.p[c] + (size_t)y * H.pitch[c];
    const uint32_t my = (uint32_t)((y & 0xFF) ^ (y >> 8));
    uint32_t sum = 0;
    for (int x = threadIdx.x; x < H.w[c]; x += 256) {
        const uint32_t m = my ^ (uint32_t)((x & 0xFF) ^ (x >> 8));
        if (H.ss == 1) sum += row[x] ^ m;
        else sum += (row[2 * x] ^ m) + (row[2 * x + 1] ^ m);
    }
    sum = block_sum(sum, red);
    if (threadIdx.x == 0) atomicAdd(out + c, sum);
}

// ---- CRC-16 ----------------------------------------------------------

// a b mod P for a, b of degree < 16 (Horner over b's bits).
__device__ uint32_t gf_mul(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int i = 15; i >= 0; --i) {
        r <<= 1;
        if (r & 0x10000u) r ^= 0x11021u;
        if ((b >> i) & 1u) r ^= a;
    }
    return r;
}

// x^(8 n) mod P.
__device__ uint32_t gf_xpow8(uint64_t n) {
    uint32_t r = 1u, b = 0x100u;  // x^8
    while (n) {
        if (n & 1u) r = gf_mul(r, b);
        b = gf_mul(b, b);
        n >>= 1;
    }
    return r;
}

__device__ __forceinline__ uint32_t block_xor(uint32_t v, uint32_t* red) {
    for (int o = 32; o > 0; o >>= 1) v ^= (uint32_t)__shfl_xor((int)v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] ^ red[1] ^ red[2] ^ red[3];
}

// One workgroup per (row, component): R_y = the row's bytes as a
// polynomial mod P. Lane t reduces bytes [t k, t k + k) of the row a byte
// at a time (r x^8 + b mod P through the table T[h] = h x^16 mod P) and
// shifts the result past the row's later bytes.
__global__ void __launch_bounds__(256) hash_crc_rows(HashPlanes H, uint32_t* rows, int hmax) {
    __shared__ uint32_t T[256];
    __shared__ uint32_t red[4];
    const int c = blockIdx.y, y = blockIdx.x;
    T[threadIdx.x] = gf_mul(threadIdx.x, 0x1021u);  // x^16 mod P = 0x1021
    __syncthreads();
    if (y >= H.h[c]) return;
    const int L = H.w[c] * H.ss;
    const int k = (L + 255) / 256, lo = min((int)threadIdx.x * k, L), hi = min(lo + k, L);
    const uint8_t* row = H.p[c] + (size_t)y * H.pitch[c];
    uint32_t r = 0;
    for (int i = lo; i < hi; ++i) r = (((r << 8) & 0xFFFFu) | row[i]) ^ T[r >> 8];
    if (hi < L) r = gf_mul(r, gf_xpow8((uint64_t)(L - hi)));
    r = block_xor(r, red);
    if (threadIdx.x == 0) rows[(size_t)c * hmax + y] = r;
}

// One workgroup per component: M = XOR over rows of R_y x^(8 L (h-1-y)),
// lane t taking a run of rows by Horner; then the seed and the flush.
__global__ void __launch_bounds__(256) hash_crc_final(HashPlanes H, const uint32_t* rows, int hmax, uint16_t* out) {
    __shared__ uint32_t red[4];
    const int c = blockIdx.x, h = H.h[c];
    const uint64_t L = (uint64_t)H.w[c] * H.ss;
    const int k = (h + 255) / 256, lo = min((int)threadIdx.x * k, h), hi = min(lo + k, h);
    const uint32_t xl = gf_xpow8(L);
    uint32_t m = 0;
    for (int y = lo; y < hi; ++y) m = gf_mul(m, xl) ^ rows[(size_t)c * hmax + y];
    if (hi > lo && hi < h) m = gf_mul(m, gf_xpow8(L * (uint64_t)(h - hi)));
    m = block_xor(m, red);
    if (threadIdx.x == 0) out[c] = (uint16_t)gf_mul(gf_mul(0xFFFFu, gf_xpow8(L * (uint64_t)h)) ^ m, 0x1021u);
}

// ---- MD5 -------------------------------------------------------------

// RFC 1321's sine table, floor(|sin(i + 1)| 2^32), and rotations.
__constant__ uint32_t kMd5K[64] = {
    0xd76aa478u, 0xe8c7b756u, 0x242070dbu, 0xc1bdceeeu, 0xf57c0fafu, 0x4787c62au, 0xa8304613u, 0xfd469501u,
    0x698098d8u, 0x8b44f7afu, 0xffff5bb1u, 0x895cd7beu, 0x6b901122u, 0xfd987193u, 0xa679438eu, 0x49b40821u,
    0xf61e2562u, 0xc040b340u, 0x265e5a51u, 0xe9b6c7aau, 0xd62f105du, 0x02441453u, 0xd8a1e681u, 0xe7d3fbc8u,
    0x21e1cde6u, 0xc33707d6u, 0xf4d50d87u, 0x455a14edu, 0xa9e3e905u, 0xfcefa3f8u, 0x676f02d9u, 0x8d2a4c8au,
    0xfffa3942u, 0x8771f681u, 0x6d9d6122u, 0xfde5380cu, 0xa4beea44u, 0x4bdecfa9u, 0xf6bb4b60u, 0xbebfbc70u,
    0x289b7ec6u, 0xeaa127fau, 0xd4ef3085u, 0x04881d05u, 0xd9d4d039u, 0xe6db99e5u, 0x1fa27cf8u, 0xc4ac5665u,
    0xf4292244u, 0x432aff97u, 0xab9423a7u, 0xfc93a039u, 0x655b59c3u, 0x8f0ccc92u, 0xffeff47du, 0x85845dd1u,
    0x6fa87e4fu, 0xfe2ce6e0u, 0xa3014314u, 0x4e0811a1u, 0xf7537e82u, 0xbd3af235u, 0x2ad7d2bbu, 0xeb86d391u,
};
constexpr int kMd5S[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
constexpr int kMd5Chunk = 4096;  // padded-message bytes staged per round

// One 64-byte block (16 little-endian words) into the state.
__device__ __forceinline__ void md5_block(const uint32_t* m, uint32_t (&st)[4]) {
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
#pragma unroll
    for (int i = 0; i < 64; ++i) {
        uint32_t f;
        int g;
        if (i < 16) f = (b & c) | (~b & d), g = i;
        else if (i < 32) f = (d & b) | (~d & c), g = (5 * i + 1) & 15;
        else if (i < 48) f = b ^ c ^ d, g = (3 * i + 5) & 15;
        else f = c ^ (b | ~d), g = (7 * i) & 15;
        f += a + kMd5K[i] + m[g];
        a = d, d = c, c = b;
        b += __builtin_rotateleft32(f, kMd5S[i >> 4][i & 3]);
    }
    st[0] += a, st[1] += b, st[2] += c, st[3] += d;
}

// One workgroup per component. Byte j of the padded message (length N, a
// multiple of 64): message byte j below n, then 0x80, zeros, and the bit
// length 8 n as a little-endian u64 in the last 8 bytes.
__global__ void __launch_bounds__(256) hash_md5(HashPlanes H, uint8_t* out) {
    __shared__ uint32_t msg[kMd5Chunk / 4];
    const int c = blockIdx.x, t = threadIdx.x;
    const uint64_t L = (uint64_t)H.w[c] * H.ss, n = L * (uint64_t)H.h[c];
    const uint64_t N = (n + 8) / 64 * 64 + 64, bits = 8 * n;
    const uint8_t* pl = H.p[c];
    const size_t pitch = (size_t)H.pitch[c];
    uint32_t st[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
    for (uint64_t base = 0; base < N; base += kMd5Chunk) {
        // Lane t forms bytes base + 16 t .. + 15.
        uint64_t j = base + 16 * (uint64_t)t;
        uint64_t y = j < n ? j / L : 0, x = j < n ? j - y * L : 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint32_t w = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k, ++j) {
                uint32_t v;
                if (j < n) {
                    v = pl[y * pitch + x];
                    if (++x == L) x = 0, ++y;
                } else if (j == n) {
                    v = 0x80u;
                } else if (j >= N - 8 && j < N) {
                    v = (uint32_t)(bits >> (8 * (j - (N - 8)))) & 0xFFu;
                } else {
                    v = 0u;
                }
                w |= v << (8 * k);
            }
            msg[4 * t + q] = w;
        }
        __syncthreads();
        if (t == 0) {
            const uint64_t left = (N - base) / 64;
            const int nb = left < kMd5Chunk / 64 ? (int)left : kMd5Chunk / 64;
            for (int b = 0; b < nb; ++b) md5_block(msg + 16 * b, st);
        }
        __syncthreads();
    }
    if (t == 0)  // the lane that ran the compression holds the state
        for (int i = 0; i < 16; ++i) out[16 * c + i] = (uint8_t)(st[i >> 2] >> (8 * (i & 3)));
}

// One workgroup per picture: the MD5 of the picture packed planar (the
// Y rows, then Cb, then Cr, each cropped to the output size, samples as
// bytes low byte first), the per-frame line of examples/h26xdec.rs:
// 120-131. The message runs across the three planes; otherwise as
// hash_md5.
__global__ void __launch_bounds__(256) frame_md5(const FrameMd5* F, uint8_t* out) {
    __shared__ uint32_t msg[kMd5Chunk / 4];
    const FrameMd5 f = F[blockIdx.x];
    const int t = threadIdx.x;
    uint64_t len[3], n = 0;
    for (int c = 0; c < 3; ++c) n += len[c] = (uint64_t)f.w[c] * f.ss * (uint64_t)f.h[c];
    const uint64_t N = (n + 8) / 64 * 64 + 64, bits = 8 * n;
    uint32_t st[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
    for (uint64_t base = 0; base < N; base += kMd5Chunk) {
        // Lane t forms bytes base + 16 t .. + 15: plane c, row y, byte x.
        uint64_t j = base + 16 * (uint64_t)t;
        int c = 0;
        uint64_t r = j;
        while (c < 2 && r >= len[c]) r -= len[c++];
        uint64_t L = (uint64_t)f.w[c] * f.ss;
        uint64_t y = j < n ? r / L : 0, x = j < n ? r - y * L : 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint32_t w = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k, ++j) {
                uint32_t v;
                if (j < n) {
                    v = f.p[c][y * (uint64_t)f.pitch[c] + x];
                    if (++x == L) {
                        x = 0;
                        if (++y == (uint64_t)f.h[c] && c < 2) {
                            y = 0;
                            ++c;
                            L = (uint64_t)f.w[c] * f.ss;
                        }
                    }
                } else if (j == n) {
                    v = 0x80u;
                } else if (j >= N - 8 && j < N) {
                    v = (uint32_t)(bits >> (8 * (j - (N - 8)))) & 0xFFu;
                } else {
                    v = 0u;
                }
                w |= v << (8 * k);
            }
            msg[4 * t + q] = w;
        }
        __syncthreads();
        if (t == 0) {
            const uint64_t left = (N - base) / 64;
            const int nb = left < kMd5Chunk / 64 ? (int)left : kMd5Chunk / 64;
            for (int b = 0; b < nb; ++b) md5_block(msg + 16 * b, st);
        }
        __syncthreads();
    }
    if (t == 0)
        for (int i = 0; i < 16; ++i) out[16 * blockIdx.x + i] = (uint8_t)(st[i >> 2] >> (8 * (i & 3)));
}

}  // namespace

hipError_t frame_md5_launch(const FrameMd5* frames, int n, uint8_t* out, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(frame_md5, dim3(n), dim3(256), 0, st, frames, out);
    return hipGetLastError();
}

size_t hash_scratch_bytes(int type, const HashPlanes& H) {
    const int hmax = max(H.h[0], max(H.h[1], H.h[2]));
    return 256 + (type == 1 ? (size_t)3 * hmax * sizeof(uint32_t) : 0);
}

hipError_t hash_launch(int type, const HashPlanes& H, void* scratch, hipStream_t st) {
    const int hmax = max(H.h[0], max(H.h[1], H.h[2]));
    if (type == 0) {
        hipLaunchKernelGGL(hash_md5, dim3(3), dim3(256), 0, st, H, (uint8_t*)scratch);
    } else if (type == 1) {
        uint32_t* rows = (uint32_t*)((char*)scratch + 256);
        hipLaunchKernelGGL(hash_crc_rows, dim3(hmax, 3), dim3(256), 0, st, H, rows, hmax);
        hipLaunchKernelGGL(hash_crc_final, dim3(3), dim3(256), 0, st, H, (const uint32_t*)rows, hmax,
                           (uint16_t*)scratch);
    } else {
        hipLaunchKernelGGL(hash_zero, dim3(1), dim3(64), 0, st, (uint32_t*)scratch, 3);
        hipLaunchKernelGGL(hash_checksum, dim3(hmax, 3), dim3(256), 0, st, H, (uint32_t*)scratch);
    }
    return hipGetLastError();
}

}  // namespace rv
