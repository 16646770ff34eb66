// Baseline JPEG encoder on the GPU: the JPEG stage of the render pipeline,
// behind the interface of rfkt::jpeg_encoder
// (src/librefrakt/include/librefrakt/interface/jpeg_encoder.h:11-24), whose
// backends are nvJPEG (src/librefrakt/src/impl/jpeg_encoder/nvjpeg.cpp:124-196:
// interleaved RGB in, 4:4:4 sampling, quality 1..100) and stb on the host
// (cpu_jpeg_encoder.cpp:14-41). Output: a JFIF file, baseline sequential
// Huffman coding, 4:4:4 YCbCr, the quantisation and Huffman tables of the
// JPEG standard's Annex K with the IJG quality scaling, and a restart marker
// every kRestart MCUs so the entropy-coded segments are independent:
//
//   1. jpeg_dct_kernel   one wave per 8x8 block, one lane per coefficient:
//                        colour conversion, integer 2-D DCT, quantisation,
//                        zigzag order, and the block's non-zero masks (ballot)
//   2. jpeg_huff_kernel  one wave per restart segment, one lane per block:
//                        DC prediction, run-length + Huffman code lengths,
//                        a wave prefix sum for the blocks' bit offsets, the
//                        bits ORed into LDS, then byte stuffing, padding and
//                        the RSTn marker into a fixed slot per segment
//   3. jpeg_scan_kernel  exclusive scan of the segment lengths (one group)
//   4. jpeg_gather_kernel  segments packed end to end
//
// All arithmetic is integer, so the bytes are fixed by the inputs: the host
// adds the headers and the EOI marker when the caller retrieves the file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/refrakt.h"

namespace rfk {
int set_error(int code, const std::string& msg);
}

namespace {

int fail(int code, const std::string& m) { return rfk::set_error(code, m); }

#define RFK_JHIP(expr)                                                                                  \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail(RFK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr uint32_t kRestart = 8;   // MCUs per restart interval (entropy-coded segment)
constexpr uint32_t kBlockCap = 420;  // bytes of one coded block, worst case with stuffing
constexpr uint32_t kSegCap = kRestart * 3 * kBlockCap + 8;

// orthonormal 8-point DCT-II rows, round(2^14 * a(u) * cos((2x + 1) u pi / 16))
__constant__ int16_t c_dct[64] = {
    5793, 5793,  5793,  5793,  5793,  5793,  5793,  5793,  8035, 6811,  4551,  1598,  -1598, -4551, -6811, -8035,
    7568, 3135,  -3135, -7568, -7568, -3135, 3135,  7568,  6811, -1598, -8035, -4551, 4551,  8035,  1598,  -6811,
    5793, -5793, -5793, 5793,  5793,  -5793, -5793, 5793,  4551, -8035, 1598,  6811,  -6811, -1598, 8035,  -4551,
    3135, -7568, 7568,  -3135, -3135, 7568,  -7568, 3135,  1598, -4551, 6811,  -8035, 8035,  -6811, 4551,  -1598};

// natural index of zigzag position k
__constant__ uint8_t c_zigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                     12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                     35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                     58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct quant_tables {
    uint16_t q[2][64];      // luminance, chrominance; natural order
    uint32_t magic[2][64];  // ceil(2^20 / q): floor(a / q) = (a * magic) >> 20 for a < 2^12
};

// Huffman code of each symbol: code | length << 16 (DC: 2 x 16, AC: 2 x 256)
struct huff_codes {
    uint32_t dc[2][16];
    uint32_t ac[2][256];
};

constexpr uint32_t kDctWaves = 4;  // blocks per workgroup

// LDS writes of a wave's lanes visible to its other lanes
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// One wave per 8x8 block (kDctWaves blocks per group, bx fastest), all three
// components (Y, Cb, Cr) per lane. coef: [block][3][64] int16 in zigzag
// order; nz: [block][3] bit k set when coefficient k is non-zero.
__global__ void __launch_bounds__(64 * kDctWaves)
    jpeg_dct_kernel(const uint8_t* __restrict__ img, size_t pitch, uint32_t bpp, uint32_t width, uint32_t height,
                    uint32_t blocks_x, uint32_t blocks, quant_tables qt, int16_t* __restrict__ coef,
                    uint64_t* __restrict__ nz) {
    __shared__ int32_t s_px[kDctWaves][3][64];
    __shared__ int32_t s_t[kDctWaves][3][64];
    const uint32_t wv = threadIdx.x >> 6, k = __lane_id();
    const uint32_t b = blockIdx.x * kDctWaves + wv;
    if (b >= blocks) return;  // uniform per wave; only wave barriers follow
    const uint32_t bx = b % blocks_x, by = b / blocks_x;
    int32_t(*px)[64] = s_px[wv];
    int32_t(*tt)[64] = s_t[wv];
    // pixel k of the block, edges replicated; JFIF full-range YCbCr with 16
    // fractional bits, chroma rounded with 2^15 - 1 so it stays below 256
    {
        const uint32_t x = std::min(bx * 8u + (k & 7u), width - 1u), y = std::min(by * 8u + (k >> 3), height - 1u);
        const uint8_t* p = img + size_t(y) * pitch + size_t(x) * bpp;
        const int32_t r = p[0], g = p[1], bl = p[2];
        px[0][k] = ((19595 * r + 38470 * g + 7471 * bl + 32768) >> 16) - 128;
        px[1][k] = ((-11059 * r - 21709 * g + 32768 * bl + (128 << 16) + 32767) >> 16) - 128;
        px[2][k] = ((32768 * r - 27439 * g - 5329 * bl + (128 << 16) + 32767) >> 16) - 128;
    }
    wave_sync();
    // rows: t[v][u] = sum_x D[u][x] s[v][x], kept with 3 fractional bits
    {
        const uint32_t v = k >> 3, u = k & 7u;
        int32_t t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
        for (uint32_t x = 0; x < 8; x++) {
            const int32_t dx = c_dct[u * 8 + x];
            t0 += dx * px[0][v * 8 + x];
            t1 += dx * px[1][v * 8 + x];
            t2 += dx * px[2][v * 8 + x];
        }
        tt[0][k] = (t0 + 1024) >> 11;
        tt[1][k] = (t1 + 1024) >> 11;
        tt[2][k] = (t2 + 1024) >> 11;
    }
    wave_sync();
    // columns at natural position p = zigzag[k] (the result scaled by 2^17),
    // quantised with rounding half away from zero:
    // floor((|f| + q 2^16) / (q 2^17)) = floor(((|f| + q 2^16) >> 17) / q)
    const uint32_t p = c_zigzag[k], v2 = p >> 3, u = p & 7u;
    int32_t f[3] = {0, 0, 0};
#pragma unroll
    for (uint32_t v = 0; v < 8; v++) {
        const int32_t dv = c_dct[v2 * 8 + v];
        f[0] += dv * tt[0][v * 8 + u];
        f[1] += dv * tt[1][v * 8 + u];
        f[2] += dv * tt[2][v * 8 + u];
    }
#pragma unroll
    for (uint32_t c = 0; c < 3; c++) {
        const uint32_t t = c ? 1u : 0u;
        const uint32_t a = (uint32_t(f[c] < 0 ? -f[c] : f[c]) + (uint32_t(qt.q[t][p]) << 16)) >> 17;
        const int32_t mag = int32_t((a * qt.magic[t][p]) >> 20);
        int32_t q = f[c] < 0 ? -mag : mag;
        // baseline AC magnitudes have at most 10 bits (8-bit input stays below 1000)
        if (k) q = std::clamp(q, -1023, 1023);
        coef[(size_t(b) * 3 + c) * 64 + k] = int16_t(q);
        const uint64_t m = __ballot(q != 0);
        if (k == 0) nz[size_t(b) * 3 + c] = m;
    }
}

__device__ __forceinline__ uint32_t magnitude_bits(int32_t v, uint32_t* size) {
    const uint32_t a = uint32_t(v < 0 ? -v : v);
    const uint32_t s = a ? 32u - __clz(a) : 0u;
    *size = s;
    return uint32_t(v < 0 ? v - 1 : v) & ((1u << s) - 1u);
}

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(v, d);
        if (__lane_id() >= d) v += u;
    }
    return v;
}

constexpr uint32_t kSegBlocks = kRestart * 3;
constexpr uint32_t kSegWords = (kSegBlocks * 1664 + 31) / 32;  // unstuffed bits of a segment, worst case
constexpr uint32_t kHuffWaves = 4;

// Bits of a segment, MSB first, in LDS words (big-endian byte order);
// blocks own disjoint bit ranges, so only boundary words are shared.
struct bit_sink {
    uint32_t* words;
    uint32_t pos;
    __device__ void put(uint32_t code, uint32_t len) {
        const uint32_t w = pos >> 5, sh = pos & 31u, room = 32u - sh;
        if (len <= room) {
            atomicOr(&words[w], code << (room - len));
        } else {
            atomicOr(&words[w], code >> (len - room));
            atomicOr(&words[w + 1], code << (32u - (len - room)));
        }
        pos += len;
    }
};

// One wave per restart segment (kHuffWaves per group): lane j codes block j
// of the segment (MCU j / 3, component j % 3). Pass 1 sums each block's code
// length, a wave prefix sum places the blocks, pass 2 ORs the bits into LDS;
// then the lanes copy the bytes out with 0x00 stuffed after every 0xFF, pad
// the last byte with 1-bits and end every segment but the last with RSTn.
__global__ void __launch_bounds__(64 * kHuffWaves)
    jpeg_huff_kernel(const int16_t* __restrict__ coef, const uint64_t* __restrict__ nz, uint32_t mcus,
                     uint32_t segments, const huff_codes* __restrict__ hc, uint8_t* __restrict__ seg_buf,
                     uint32_t* __restrict__ seg_len) {
    __shared__ huff_codes s_hc;
    __shared__ __align__(16) int16_t s_coef[kHuffWaves][kSegBlocks * 64];
    __shared__ uint32_t s_bits[kHuffWaves][kSegWords];
    const uint32_t* src = reinterpret_cast<const uint32_t*>(hc);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&s_hc);
    for (uint32_t i = threadIdx.x; i < sizeof(huff_codes) / 4; i += blockDim.x) dst[i] = src[i];
    const uint32_t wv = threadIdx.x >> 6, lane = __lane_id();
    const uint32_t s = blockIdx.x * kHuffWaves + wv;
    const bool live = s < segments;  // uniform per wave
    const uint32_t m0 = s * kRestart, nblk = live ? (std::min(mcus, m0 + kRestart) - m0) * 3u : 0u;
    int16_t* cf = s_coef[wv];
    uint32_t* bits = s_bits[wv];
    if (live) {
        // the segment's blocks are contiguous: [m0 * 3, m0 * 3 + nblk) x 64 coefficients
        const uint4* g = reinterpret_cast<const uint4*>(coef + size_t(m0) * 3 * 64);
        uint4* l = reinterpret_cast<uint4*>(cf);
        for (uint32_t i = lane; i < nblk * 8u; i += 64) l[i] = g[i];
        for (uint32_t i = lane; i < kSegWords; i += 64) bits[i] = 0u;
    }
    __syncthreads();
    if (!live) return;  // no barrier follows
    const bool mine = lane < nblk;
    const uint32_t c = lane % 3u, t = c ? 1u : 0u;
    const int16_t* blk = cf + lane * 64u;
    const int32_t dc = mine ? blk[0] : 0;
    const int32_t prev = (mine && lane >= 3u) ? cf[(lane - 3u) * 64u] : 0;
    const uint64_t mask = mine ? nz[size_t(m0) * 3 + lane] & ~1ull : 0ull;
    uint32_t dsize;
    const uint32_t dbits = magnitude_bits(dc - prev, &dsize);
    const uint32_t zrl = s_hc.ac[t][0xF0], eob = s_hc.ac[t][0x00];
    // pass 1: coded length of the block
    uint32_t len = 0;
    if (mine) {
        len = (s_hc.dc[t][dsize] >> 16) + dsize;
        uint32_t last = 0;
        for (uint64_t m = mask; m; m &= m - 1ull) {
            const uint32_t k = uint32_t(__builtin_ctzll(m));
            const uint32_t run = k - last - 1u;
            uint32_t size;
            magnitude_bits(blk[k], &size);
            len += (run >> 4) * (zrl >> 16) + (s_hc.ac[t][((run & 15u) << 4) | size] >> 16) + size;
            last = k;
        }
        if (last != 63u) len += eob >> 16;
    }
    const uint32_t incl = wave_inclusive_sum(len);
    const uint32_t total = __shfl(incl, 63);
    // pass 2: the bits at the block's offset
    if (mine) {
        bit_sink w{bits, incl - len};
        const uint32_t hdc = s_hc.dc[t][dsize];
        w.put(hdc & 0xFFFFu, hdc >> 16);
        if (dsize) w.put(dbits, dsize);
        uint32_t last = 0;
        for (uint64_t m = mask; m; m &= m - 1ull) {
            const uint32_t k = uint32_t(__builtin_ctzll(m));
            uint32_t run = k - last - 1u;
            for (; run >= 16u; run -= 16u) w.put(zrl & 0xFFFFu, zrl >> 16);
            uint32_t size;
            const uint32_t vb = magnitude_bits(blk[k], &size);
            const uint32_t hac = s_hc.ac[t][(run << 4) | size];
            w.put(hac & 0xFFFFu, hac >> 16);
            w.put(vb, size);
            last = k;
        }
        if (last != 63u) w.put(eob & 0xFFFFu, eob >> 16);
    }
    // pad the last byte with 1-bits
    if (lane == 0 && (total & 7u)) {
        bit_sink w{bits, total};
        const uint32_t padn = 8u - (total & 7u);
        w.put((1u << padn) - 1u, padn);
    }
    wave_sync();
    // bytes out, 0x00 after every 0xFF: lane L copies bytes [L * per, ...)
    const uint32_t nbytes = (total + 7u) >> 3, per = (nbytes + 63u) >> 6;
    const uint32_t lo = std::min(nbytes, lane * per), hi = std::min(nbytes, lo + per);
    auto byte_at = [&](uint32_t i) { return uint8_t(bits[i >> 2] >> (24u - 8u * (i & 3u))); };
    uint32_t ff = 0;
    for (uint32_t i = lo; i < hi; i++) ff += byte_at(i) == 0xFFu;
    const uint32_t ffincl = wave_inclusive_sum(ff);
    uint8_t* out = seg_buf + size_t(s) * kSegCap;
    uint32_t o = lo + ffincl - ff;
    for (uint32_t i = lo; i < hi; i++) {
        const uint8_t b = byte_at(i);
        out[o++] = b;
        if (b == 0xFFu) out[o++] = 0;
    }
    if (lane == 63) {
        uint32_t n = nbytes + ffincl;
        if (s + 1u < segments) {
            out[n++] = 0xFF;
            out[n++] = uint8_t(0xD0u + (s & 7u));
        }
        seg_len[s] = n;
    }
}

// one group of 1024: seg_off = exclusive scan of seg_len, *total = the sum
__global__ void __launch_bounds__(1024)
    jpeg_scan_kernel(const uint32_t* __restrict__ seg_len, uint32_t n, uint32_t* __restrict__ seg_off,
                     uint32_t* __restrict__ total) {
    __shared__ uint32_t s_wave[16];
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t lo = threadIdx.x * per, hi = std::min(n, lo + per);
    uint32_t sum = 0;
    for (uint32_t i = lo; i < hi; i++) sum += seg_len[i];
    uint32_t incl = sum;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(incl, d);
        if (__lane_id() >= d) incl += u;
    }
    if (__lane_id() == 63) s_wave[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (uint32_t w = 0; w < 16; w++) {
        if (w < (threadIdx.x >> 6)) base += s_wave[w];
        all += s_wave[w];
    }
    uint32_t run = base + incl - sum;
    for (uint32_t i = lo; i < hi; i++) {
        seg_off[i] = run;
        run += seg_len[i];
    }
    if (threadIdx.x == 0) *total = all;
}

// grid = segments: segment s's bytes to out + seg_off[s]
__global__ void __launch_bounds__(256)
    jpeg_gather_kernel(const uint8_t* __restrict__ seg_buf, const uint32_t* __restrict__ seg_len,
                       const uint32_t* __restrict__ seg_off, uint8_t* __restrict__ out) {
    const uint32_t s = blockIdx.x, n = seg_len[s];
    const uint8_t* src = seg_buf + size_t(s) * kSegCap;
    uint8_t* dst = out + seg_off[s];
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// ---------------------------------------------------------------- tables (host)

// JPEG standard Annex K.1: luminance and chrominance quantisation, natural order
const uint8_t kStdQuant[2][64] = {
    {16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57,
     69, 56, 14, 17, 22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64,
     81, 104, 113, 92, 49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99},
    {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99,
     99, 99, 47, 66, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
     99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99}};

const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                             41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                             30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Annex K.3: code counts per length 1..16 and symbols, DC then AC, luminance then chrominance
const uint8_t kBitsDc[2][16] = {{0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0},
                                {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0}};
const uint8_t kValDc[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kBitsAc[2][16] = {{0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d},
                                {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77}};
const uint8_t kValAc[2][162] = {
    {0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71,
     0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
     0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37,
     0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
     0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
     0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
     0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
     0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
     0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa},
    {0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22,
     0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1,
     0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36,
     0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58,
     0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
     0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a,
     0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba,
     0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda,
     0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa}};

// canonical codes in symbol order of the table (Annex C)
void make_codes(const uint8_t bits[16], const uint8_t* vals, uint32_t* out) {
    uint32_t code = 0, i = 0;
    for (uint32_t len = 1; len <= 16; len++) {
        for (uint32_t n = 0; n < bits[len - 1]; n++, i++) out[vals[i]] = code++ | (len << 16);
        code <<= 1;
    }
}

// IJG quality scaling of the Annex K tables, baseline range 1..255
quant_tables make_quant(int quality) {
    const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
    quant_tables qt;
    for (int t = 0; t < 2; t++)
        for (int i = 0; i < 64; i++) {
            qt.q[t][i] = uint16_t(std::clamp((int(kStdQuant[t][i]) * scale + 50) / 100, 1, 255));
            qt.magic[t][i] = ((1u << 20) + qt.q[t][i] - 1u) / qt.q[t][i];
        }
    return qt;
}

void put16(std::vector<uint8_t>& h, uint32_t v) {
    h.push_back(uint8_t(v >> 8));
    h.push_back(uint8_t(v));
}

// SOI, JFIF APP0, DQT, SOF0, DHT, DRI, SOS
std::vector<uint8_t> make_header(const quant_tables& qt, uint32_t width, uint32_t height) {
    std::vector<uint8_t> h = {0xFF, 0xD8, 0xFF, 0xE0, 0, 16, 'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0};
    h.insert(h.end(), {0xFF, 0xDB});
    put16(h, 2 + 2 * 65);
    for (uint32_t t = 0; t < 2; t++) {
        h.push_back(uint8_t(t));
        for (uint32_t k = 0; k < 64; k++) h.push_back(uint8_t(qt.q[t][kZigzag[k]]));
    }
    h.insert(h.end(), {0xFF, 0xC0, 0, 17, 8});
    put16(h, height);
    put16(h, width);
    h.insert(h.end(), {3, 1, 0x11, 0, 2, 0x11, 1, 3, 0x11, 1});
    h.insert(h.end(), {0xFF, 0xC4});
    put16(h, 2 + 2 * (17 + 12) + 2 * (17 + 162));
    for (uint32_t t = 0; t < 2; t++) {
        h.push_back(uint8_t(0x00 | t));
        h.insert(h.end(), kBitsDc[t], kBitsDc[t] + 16);
        h.insert(h.end(), kValDc, kValDc + 12);
        h.push_back(uint8_t(0x10 | t));
        h.insert(h.end(), kBitsAc[t], kBitsAc[t] + 16);
        h.insert(h.end(), kValAc[t], kValAc[t] + 162);
    }
    h.insert(h.end(), {0xFF, 0xDD, 0, 4});
    put16(h, kRestart);
    h.insert(h.end(), {0xFF, 0xDA, 0, 12, 3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0});
    return h;
}

}  // namespace

struct rfk_jpeg {
    huff_codes* d_codes = nullptr;
    int16_t* coef = nullptr;
    uint64_t* nz = nullptr;
    uint8_t *seg_buf = nullptr, *out = nullptr;
    uint32_t *seg_len = nullptr, *seg_off = nullptr, *total_dev = nullptr;
    uint32_t* total_host = nullptr;  // pinned
    uint8_t* stage = nullptr;        // pinned copy of the coded segments
    size_t cap_blocks = 0, cap_segments = 0, cap_stage = 0;
    hipEvent_t done = nullptr;
    std::vector<uint8_t> header;
    bool pending = false;
};

static void free_jpeg_buffers(rfk_jpeg* j) {
    auto f = [](void* p) { if (p) (void)hipFree(p); };
    f(j->coef); f(j->nz); f(j->seg_buf); f(j->out); f(j->seg_len); f(j->seg_off);
    j->coef = nullptr; j->nz = nullptr; j->seg_buf = nullptr; j->out = nullptr; j->seg_len = nullptr;
    j->seg_off = nullptr;
    j->cap_blocks = j->cap_segments = 0;
}

extern "C" {

int rfk_jpeg_create(rfk_jpeg** out) {
    if (!out) return fail(RFK_ERR_INVALID, "null argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(RFK_ERR_NODEVICE, "no GPU visible");
    auto j = new rfk_jpeg;
    huff_codes hc;
    std::memset(&hc, 0, sizeof hc);
    for (int t = 0; t < 2; t++) {
        make_codes(kBitsDc[t], kValDc, hc.dc[t]);
        make_codes(kBitsAc[t], kValAc[t], hc.ac[t]);
    }
    hipError_t e = hipMalloc(&j->d_codes, sizeof hc);
    if (e == hipSuccess) e = hipMemcpy(j->d_codes, &hc, sizeof hc, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&j->total_dev, 16);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&j->total_host), 16, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&j->done, hipEventDisableTiming);
    if (e != hipSuccess) {
        rfk_jpeg_free(j);
        return fail(RFK_ERR_HIP, std::string("jpeg encoder set-up: ") + hipGetErrorString(e));
    }
    *out = j;
    return RFK_OK;
}

int rfk_jpeg_encode(rfk_jpeg* j, const void* image, int in_format, size_t pitch_bytes, uint32_t width,
                    uint32_t height, int quality, void* stream) {
    if (!j || !image) return fail(RFK_ERR_INVALID, "null argument");
    if (in_format != RFK_PIXEL_RGB8 && in_format != RFK_PIXEL_RGBA8)
        return fail(RFK_ERR_INVALID, "jpeg input must be RFK_PIXEL_RGB8 or RFK_PIXEL_RGBA8");
    if (width == 0 || height == 0 || width > 65535 || height > 65535)
        return fail(RFK_ERR_INVALID, "jpeg image size must be 1..65535 pixels per side");
    if (quality < 1) return fail(RFK_ERR_INVALID, "jpeg quality must be 1..100");
    quality = std::min(quality, 100);
    const uint32_t bpp = in_format == RFK_PIXEL_RGB8 ? 3u : 4u;
    if (pitch_bytes < size_t(width) * bpp) return fail(RFK_ERR_INVALID, "pitch below the row size");
    if (j->pending) RFK_JHIP(hipEventSynchronize(j->done));
    const uint32_t bx = (width + 7) / 8, by = (height + 7) / 8;
    const size_t blocks = size_t(bx) * by;
    const uint32_t segments = uint32_t((blocks + kRestart - 1) / kRestart);
    if (blocks > j->cap_blocks || segments > j->cap_segments) {
        free_jpeg_buffers(j);
        RFK_JHIP(hipMalloc(&j->coef, blocks * 3 * 64 * sizeof(int16_t)));
        RFK_JHIP(hipMalloc(&j->nz, blocks * 3 * sizeof(uint64_t)));
        RFK_JHIP(hipMalloc(&j->seg_buf, size_t(segments) * kSegCap));
        RFK_JHIP(hipMalloc(&j->out, size_t(segments) * kSegCap));
        RFK_JHIP(hipMalloc(&j->seg_len, size_t(segments) * 4));
        RFK_JHIP(hipMalloc(&j->seg_off, size_t(segments) * 4));
        j->cap_blocks = blocks;
        j->cap_segments = segments;
    }
    const quant_tables qt = make_quant(quality);
    hipStream_t s = static_cast<hipStream_t>(stream);
    jpeg_dct_kernel<<<uint32_t((blocks + kDctWaves - 1) / kDctWaves), 64 * kDctWaves, 0, s>>>(
        static_cast<const uint8_t*>(image), pitch_bytes, bpp, width, height, bx, uint32_t(blocks), qt, j->coef, j->nz);
    jpeg_huff_kernel<<<(segments + kHuffWaves - 1) / kHuffWaves, 64 * kHuffWaves, 0, s>>>(
        j->coef, j->nz, uint32_t(blocks), segments, j->d_codes, j->seg_buf, j->seg_len);
    jpeg_scan_kernel<<<1, 1024, 0, s>>>(j->seg_len, segments, j->seg_off, j->total_dev);
    jpeg_gather_kernel<<<segments, 256, 0, s>>>(j->seg_buf, j->seg_len, j->seg_off, j->out);
    RFK_JHIP(hipGetLastError());
    RFK_JHIP(hipMemcpyAsync(j->total_host, j->total_dev, 4, hipMemcpyDeviceToHost, s));
    RFK_JHIP(hipEventRecord(j->done, s));
    j->header = make_header(qt, width, height);
    j->pending = true;
    return RFK_OK;
}

int rfk_jpeg_retrieve(rfk_jpeg* j, uint8_t* out, size_t cap, size_t* size) {
    if (!j || !size) return fail(RFK_ERR_INVALID, "null argument");
    if (!j->pending) return fail(RFK_ERR_INVALID, "no encoded image: call rfk_jpeg_encode first");
    RFK_JHIP(hipEventSynchronize(j->done));
    const size_t coded = *j->total_host, total = j->header.size() + coded + 2;
    *size = total;
    if (!out) return RFK_OK;
    if (cap < total) return fail(RFK_ERR_INVALID, "output buffer too small for the JPEG file");
    if (coded > j->cap_stage) {
        if (j->stage) (void)hipHostFree(j->stage);
        j->stage = nullptr;
        j->cap_stage = 0;
        RFK_JHIP(hipHostMalloc(reinterpret_cast<void**>(&j->stage), coded * 2, hipHostMallocDefault));
        j->cap_stage = coded * 2;
    }
    std::memcpy(out, j->header.data(), j->header.size());
    RFK_JHIP(hipMemcpy(j->stage, j->out, coded, hipMemcpyDeviceToHost));
    std::memcpy(out + j->header.size(), j->stage, coded);
    out[total - 2] = 0xFF;
    out[total - 1] = 0xD9;
    return RFK_OK;
}

int rfk_jpeg_free(rfk_jpeg* j) {
    if (!j) return RFK_OK;
    if (j->pending) (void)hipEventSynchronize(j->done);
    free_jpeg_buffers(j);
    if (j->d_codes) (void)hipFree(j->d_codes);
    if (j->total_dev) (void)hipFree(j->total_dev);
    if (j->total_host) (void)hipHostFree(j->total_host);
    if (j->stage) (void)hipHostFree(j->stage);
    if (j->done) (void)hipEventDestroy(j->done);
    delete j;
    return RFK_OK;
}

}  // extern "C"
