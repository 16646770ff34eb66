/*
 * ORACLE — test infrastructure only. Nothing in the product imports, links
 * or calls this file; tests/ use it as the checker of the GPU JPEG encoder.
 *
 * A scalar CPU restatement of the baseline JPEG encoder that stands behind
 * rfkt::jpeg_encoder::encode_image (interface/jpeg_encoder.h:11-24; nvJPEG
 * backend impl/jpeg_encoder/nvjpeg.cpp:124-196 with 4:4:4 sampling and the
 * quality parameter): JFIF, 4:4:4 YCbCr, ISO/IEC 10918-1 Annex K tables with
 * the IJG quality scaling, restart marker every 8 MCUs. The integer colour
 * conversion, DCT and quantisation are the encoder's own fixed-point
 * definition (stated below); the DCT matrix and the zigzag order are derived
 * here from their definitions rather than taken from the GPU code. Pinned by
 * an independent decoder (PIL / libjpeg in tests/test_cpu_oracle.py):
 * decoded images match the input to the quality's error level.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define RESTART 8

/* Annex K.1 */
static const uint8_t std_quant[2][64] = {
    {16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57,
     69, 56, 14, 17, 22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64,
     81, 104, 113, 92, 49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99},
    {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99,
     99, 99, 47, 66, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
     99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99}};

/* Annex K.3 */
static const uint8_t bits_dc[2][16] = {{0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0},
                                       {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0}};
static const uint8_t bits_ac[2][16] = {{0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d},
                                       {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77}};
static const uint8_t vals_ac[2][162] = {
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

typedef struct {
    uint8_t* p;
    size_t n, cap;
    int overflow;
    uint32_t acc;
    int nbits;
} sink;

static void byte_out(sink* s, uint8_t b) {
    if (s->n < s->cap) s->p[s->n] = b;
    else s->overflow = 1;
    s->n++;
}

static void bits_out(sink* s, uint32_t code, int len) {
    for (int i = len - 1; i >= 0; i--) {
        s->acc = (s->acc << 1) | ((code >> i) & 1u);
        if (++s->nbits == 8) {
            uint8_t b = (uint8_t)s->acc;
            byte_out(s, b);
            if (b == 0xFF) byte_out(s, 0);
            s->acc = 0;
            s->nbits = 0;
        }
    }
}

/* canonical Huffman codes (ISO/IEC 10918-1 Annex C) */
static void codes(const uint8_t bits[16], const uint8_t* vals, uint16_t code[256], uint8_t len[256]) {
    int k = 0;
    uint32_t c = 0;
    for (int l = 1; l <= 16; l++) {
        for (int i = 0; i < bits[l - 1]; i++, k++) {
            code[vals[k]] = (uint16_t)c++;
            len[vals[k]] = (uint8_t)l;
        }
        c <<= 1;
    }
}

static int category(int v) {
    int a = v < 0 ? -v : v, s = 0;
    while (a) {
        s++;
        a >>= 1;
    }
    return s;
}

static void put16(sink* s, int v) {
    byte_out(s, (uint8_t)(v >> 8));
    byte_out(s, (uint8_t)v);
}

long orc_jpeg_encode(const uint8_t* img, size_t pitch, int bpp, int w, int h, int quality, uint8_t* out, size_t cap) {
    if (w < 1 || h < 1 || quality < 1 || (bpp != 3 && bpp != 4)) return -1;
    if (quality > 100) quality = 100;
    /* DCT matrix from its definition, zigzag by walking the anti-diagonals */
    int D[8][8], zz[64];
    for (int u = 0; u < 8; u++)
        for (int x = 0; x < 8; x++)
            D[u][x] = (int)lround(16384.0 * (u ? 0.5 : sqrt(0.125)) * cos((2 * x + 1) * u * M_PI / 16.0));
    for (int d = 0, k = 0; d < 15; d++)
        for (int i = 0; i < 8; i++) {
            int r = (d & 1) ? i : d - i, c = d - r; /* odd diagonals walk down, even ones up */
            if (r >= 0 && r < 8 && c >= 0 && c < 8) zz[k++] = r * 8 + c;
        }
    int Q[2][64];
    int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
    for (int t = 0; t < 2; t++)
        for (int i = 0; i < 64; i++) {
            int q = (std_quant[t][i] * scale + 50) / 100;
            Q[t][i] = q < 1 ? 1 : q > 255 ? 255 : q;
        }
    uint16_t dcc[2][256], acc[2][256];
    uint8_t dcl[2][256], acl[2][256];
    static const uint8_t dcvals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
    for (int t = 0; t < 2; t++) {
        codes(bits_dc[t], dcvals, dcc[t], dcl[t]);
        codes(bits_ac[t], vals_ac[t], acc[t], acl[t]);
    }

    sink s = {out, 0, cap, 0, 0, 0};
    /* SOI + JFIF APP0 */
    static const uint8_t app0[20] = {0xFF, 0xD8, 0xFF, 0xE0, 0, 16, 'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0};
    for (int i = 0; i < 20; i++) byte_out(&s, app0[i]);
    byte_out(&s, 0xFF); byte_out(&s, 0xDB); put16(&s, 2 + 2 * 65);
    for (int t = 0; t < 2; t++) {
        byte_out(&s, (uint8_t)t);
        for (int k = 0; k < 64; k++) byte_out(&s, (uint8_t)Q[t][zz[k]]);
    }
    byte_out(&s, 0xFF); byte_out(&s, 0xC0); put16(&s, 17); byte_out(&s, 8); put16(&s, h); put16(&s, w);
    byte_out(&s, 3);
    for (int c = 0; c < 3; c++) {
        byte_out(&s, (uint8_t)(c + 1));
        byte_out(&s, 0x11);
        byte_out(&s, (uint8_t)(c ? 1 : 0));
    }
    byte_out(&s, 0xFF); byte_out(&s, 0xC4); put16(&s, 2 + 2 * (17 + 12) + 2 * (17 + 162));
    for (int t = 0; t < 2; t++) {
        byte_out(&s, (uint8_t)t);
        for (int i = 0; i < 16; i++) byte_out(&s, bits_dc[t][i]);
        for (int i = 0; i < 12; i++) byte_out(&s, dcvals[i]);
        byte_out(&s, (uint8_t)(0x10 | t));
        for (int i = 0; i < 16; i++) byte_out(&s, bits_ac[t][i]);
        for (int i = 0; i < 162; i++) byte_out(&s, vals_ac[t][i]);
    }
    byte_out(&s, 0xFF); byte_out(&s, 0xDD); put16(&s, 4); put16(&s, RESTART);
    static const uint8_t sos[14] = {0xFF, 0xDA, 0, 12, 3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0};
    for (int i = 0; i < 14; i++) byte_out(&s, sos[i]);

    const int bw = (w + 7) / 8, bh = (h + 7) / 8, mcus = bw * bh;
    int pred[3] = {0, 0, 0};
    for (int m = 0; m < mcus; m++) {
        if (m && m % RESTART == 0) {
            if (s.nbits) bits_out(&s, (1u << (8 - s.nbits)) - 1u, 8 - s.nbits);
            byte_out(&s, 0xFF);
            byte_out(&s, (uint8_t)(0xD0 + ((m / RESTART - 1) & 7)));
            pred[0] = pred[1] = pred[2] = 0;
        }
        const int bx = m % bw, by = m / bw;
        for (int c = 0; c < 3; c++) {
            int sp[64], t1[64], q[64];
            for (int k = 0; k < 64; k++) {
                int x = bx * 8 + (k & 7), y = by * 8 + (k >> 3);
                if (x > w - 1) x = w - 1;
                if (y > h - 1) y = h - 1;
                const uint8_t* p = img + (size_t)y * pitch + (size_t)x * bpp;
                int R = p[0], G = p[1], B = p[2], v;
                /* JFIF full-range YCbCr, 16 fractional bits; chroma rounds with
                   2^15 - 1 so B = 255 (Cb) and R = 255 (Cr) stay at 255 */
                if (c == 0) v = (19595 * R + 38470 * G + 7471 * B + 32768) >> 16;
                else if (c == 1) v = (-11059 * R - 21709 * G + 32768 * B + (128 << 16) + 32767) >> 16;
                else v = (32768 * R - 27439 * G - 5329 * B + (128 << 16) + 32767) >> 16;
                sp[k] = v - 128;
            }
            /* rows then columns; 3 fractional bits between the passes, the
               result scaled by 2^17; quantised rounding half away from zero */
            for (int v = 0; v < 8; v++)
                for (int u = 0; u < 8; u++) {
                    int acc_ = 0;
                    for (int x = 0; x < 8; x++) acc_ += D[u][x] * sp[v * 8 + x];
                    t1[v * 8 + u] = (acc_ + 1024) >> 11;
                }
            for (int k = 0; k < 64; k++) {
                const int p = zz[k], v2 = p >> 3, u = p & 7;
                int f = 0;
                for (int v = 0; v < 8; v++) f += D[v2][v] * t1[v * 8 + u];
                const int d = Q[c ? 1 : 0][p] << 17;
                int qv = f >= 0 ? (f + d / 2) / d : -((-f + d / 2) / d);
                if (k && qv > 1023) qv = 1023;
                if (k && qv < -1023) qv = -1023;
                q[k] = qv;
            }
            const int tb = c ? 1 : 0;
            int diff = q[0] - pred[c], cat = category(diff);
            pred[c] = q[0];
            bits_out(&s, dcc[tb][cat], dcl[tb][cat]);
            if (cat) bits_out(&s, (uint32_t)(diff < 0 ? diff - 1 : diff) & ((1u << cat) - 1u), cat);
            int run = 0;
            for (int k = 1; k < 64; k++) {
                if (q[k] == 0) {
                    run++;
                    continue;
                }
                while (run > 15) {
                    bits_out(&s, acc[tb][0xF0], acl[tb][0xF0]);
                    run -= 16;
                }
                int cv = category(q[k]), sym = (run << 4) | cv;
                bits_out(&s, acc[tb][sym], acl[tb][sym]);
                bits_out(&s, (uint32_t)(q[k] < 0 ? q[k] - 1 : q[k]) & ((1u << cv) - 1u), cv);
                run = 0;
            }
            if (run) bits_out(&s, acc[tb][0x00], acl[tb][0x00]);
        }
    }
    if (s.nbits) bits_out(&s, (1u << (8 - s.nbits)) - 1u, 8 - s.nbits);
    byte_out(&s, 0xFF);
    byte_out(&s, 0xD9);
    return s.overflow ? -(long)s.n : (long)s.n;
}
