/*
 * turbo_oracle.c -- CPU restatement of Wiseung/Turbo_CPP src/lte_turbo.cpp.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for turbo_cpp_amd, never part of the
 * product.  Built with gcc (-O2, no FMA contraction, no -march) so that its
 * float arithmetic follows the same operation order as the reference's
 * Release build on x86-64.
 *
 * Citations are file:line into the reference (lte_turbo.cpp unless named).
 */
#include "turbo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ QPP ---
 * Table of (K, f1, f2), 36.212 Table 5.1.3-3; lte_turbo.cpp:54-102.  The
 * values are the standard's and have to be identical, so they are data. */
static const uint16_t kQpp[188][3] = {
    {40, 3, 10},     {48, 7, 12},     {56, 19, 42},    {64, 7, 16},     {72, 7, 18},
    {80, 11, 20},    {88, 5, 22},     {96, 11, 24},    {104, 7, 26},    {112, 41, 84},
    {120, 103, 90},  {128, 15, 32},   {136, 9, 34},    {144, 17, 108},  {152, 9, 38},
    {160, 21, 120},  {168, 101, 84},  {176, 21, 44},   {184, 57, 46},   {192, 23, 48},
    {200, 13, 50},   {208, 27, 52},   {216, 11, 36},   {224, 27, 56},   {232, 85, 58},
    {240, 29, 60},   {248, 33, 62},   {256, 15, 32},   {264, 17, 198},  {272, 33, 68},
    {280, 103, 210}, {288, 19, 36},   {296, 19, 74},   {304, 37, 76},   {312, 19, 78},
    {320, 21, 120},  {328, 21, 82},   {336, 115, 84},  {344, 193, 86},  {352, 21, 44},
    {360, 133, 90},  {368, 81, 46},   {376, 45, 94},   {384, 23, 48},   {392, 243, 98},
    {400, 151, 40},  {408, 155, 102}, {416, 25, 52},   {424, 51, 106},  {432, 47, 72},
    {440, 91, 110},  {448, 29, 168},  {456, 29, 114},  {464, 247, 58},  {472, 29, 118},
    {480, 89, 180},  {488, 91, 122},  {496, 157, 62},  {504, 55, 84},   {512, 31, 64},
    {528, 17, 66},   {544, 35, 68},   {560, 227, 420}, {576, 65, 96},   {592, 19, 74},
    {608, 37, 76},   {624, 41, 234},  {640, 39, 80},   {656, 185, 82},  {672, 43, 252},
    {688, 21, 86},   {704, 155, 44},  {720, 79, 120},  {736, 139, 92},  {752, 23, 94},
    {768, 217, 48},  {784, 25, 98},   {800, 17, 80},   {816, 127, 102}, {832, 25, 52},
    {848, 239, 106}, {864, 17, 48},   {880, 137, 110}, {896, 215, 112}, {912, 29, 114},
    {928, 15, 58},   {944, 147, 118}, {960, 29, 60},   {976, 59, 122},  {992, 65, 124},
    {1008, 55, 84},  {1024, 31, 64},  {1056, 17, 66},  {1088, 171, 204}, {1120, 67, 140},
    {1152, 35, 72},  {1184, 19, 74},  {1216, 39, 76},  {1248, 19, 78},  {1280, 199, 240},
    {1312, 21, 82},  {1344, 211, 252}, {1376, 21, 86}, {1408, 43, 88},  {1440, 149, 60},
    {1472, 45, 92},  {1504, 49, 846}, {1536, 71, 48},  {1568, 13, 28},  {1600, 17, 80},
    {1632, 25, 102}, {1664, 183, 104}, {1696, 55, 954}, {1728, 127, 96}, {1760, 27, 110},
    {1792, 29, 112}, {1824, 29, 114}, {1856, 57, 116}, {1888, 45, 354}, {1920, 31, 120},
    {1952, 59, 610}, {1984, 185, 124}, {2016, 113, 420}, {2048, 31, 64}, {2112, 17, 66},
    {2176, 171, 136}, {2240, 209, 420}, {2304, 253, 216}, {2368, 367, 444}, {2432, 265, 456},
    {2496, 181, 468}, {2560, 39, 80},  {2624, 27, 164}, {2688, 127, 504}, {2752, 143, 172},
    {2816, 43, 88},  {2880, 29, 300}, {2944, 45, 92},  {3008, 157, 188}, {3072, 47, 96},
    {3136, 13, 28},  {3200, 111, 240}, {3264, 443, 204}, {3328, 51, 104}, {3392, 51, 212},
    {3456, 451, 192}, {3520, 257, 220}, {3584, 57, 336}, {3648, 313, 228}, {3712, 271, 232},
    {3776, 179, 236}, {3840, 331, 120}, {3904, 363, 244}, {3968, 375, 248}, {4032, 127, 168},
    {4096, 31, 64},  {4160, 33, 130}, {4224, 43, 264}, {4288, 33, 134}, {4352, 477, 408},
    {4416, 35, 138}, {4480, 233, 280}, {4544, 357, 142}, {4608, 337, 480}, {4672, 37, 146},
    {4736, 71, 444}, {4800, 71, 120}, {4864, 37, 152}, {4928, 39, 462}, {4992, 127, 234},
    {5056, 39, 158}, {5120, 39, 80},  {5184, 31, 96},  {5248, 113, 902}, {5312, 41, 166},
    {5376, 251, 336}, {5440, 43, 170}, {5504, 21, 86}, {5568, 43, 174}, {5632, 45, 176},
    {5696, 45, 178}, {5760, 161, 120}, {5824, 89, 182}, {5888, 323, 184}, {5952, 47, 186},
    {6016, 23, 94},  {6080, 47, 190}, {6144, 263, 480},
};

size_t or_qpp_count(void) { return 188; }

void or_qpp_entry(size_t idx, uint32_t* k, uint32_t* f1, uint32_t* f2) {
  *k = kQpp[idx][0];
  *f1 = kQpp[idx][1];
  *f2 = kQpp[idx][2];
}

/* lte_turbo.cpp:157-161 (linear search), :500-506 */
int or_qpp_index(uint32_t k) {
  for (int i = 0; i < 188; ++i) {
    if (kQpp[i][0] == k) return i;
  }
  return -1;
}

/* Pi(i) = (f1*i + f2*i^2) mod K in 64-bit; lte_turbo.cpp:606-613 */
void or_qpp_perm(uint32_t k, uint32_t* perm) {
  int idx = or_qpp_index(k);
  uint64_t f1 = kQpp[idx][1], f2 = kQpp[idx][2];
  for (uint64_t i = 0; i < k; ++i) perm[i] = (uint32_t)((f1 * i + f2 * i * i) % k);
}

/* ------------------------------------------------------------------ CRC ---
 * Generator polynomials exactly as the reference's tap arrays
 * (lte_turbo.cpp:104-116), written as the low N bits of G(D) with the D^N
 * term implicit.  They are NOT the 36.212 polynomials. */
static const uint32_t kCrcPoly[4] = {0xC199BFu, 0xC00063u, 0x8811u, 0xD9u};
static const uint32_t kCrcLen[4] = {24, 24, 16, 8};

uint32_t or_crc_len(int kind) { return kCrcLen[kind]; }

/* Remainder of bits(D) * D^N mod G(D), one bit at a time (an LFSR form of the
 * long division at lte_turbo.cpp:118-135). */
static uint32_t crc_remainder(int kind, const uint8_t* bits, size_t n) {
  const uint32_t len = kCrcLen[kind];
  const uint32_t top = 1u << (len - 1);
  const uint32_t mask = (len == 32) ? 0xFFFFFFFFu : ((1u << len) - 1u);
  uint32_t r = 0;
  for (size_t i = 0; i < n; ++i) {
    uint32_t fb = ((r & top) ? 1u : 0u) ^ (bits[i] & 1u);
    r = (r << 1) & mask;
    if (fb) r ^= kCrcPoly[kind];
  }
  return r;
}

void or_crc_append(int kind, const uint8_t* bits, size_t n, uint8_t* out) {
  const uint32_t len = kCrcLen[kind];
  uint32_t r = crc_remainder(kind, bits, n);
  if (out != bits) memmove(out, bits, n);
  for (uint32_t j = 0; j < len; ++j) out[n + j] = (uint8_t)((r >> (len - 1 - j)) & 1u);
}

/* lte_turbo.cpp:137-155: false when shorter than N, else the remainder of
 * the leading n-N bits must equal the trailing N bits. */
int or_crc_check(int kind, const uint8_t* bits, size_t n) {
  const uint32_t len = kCrcLen[kind];
  if (n < len) return 0;
  uint32_t r = crc_remainder(kind, bits, n - len);
  for (uint32_t j = 0; j < len; ++j) {
    if (((r >> (len - 1 - j)) & 1u) != bits[n - len + j]) return 0;
  }
  return 1;
}

/* --------------------------------------------------------- segmentation ---
 * 36.212 5.1.2 as implemented at lte_turbo.cpp:540-604. */
int or_segment(const uint8_t* bits, size_t b, size_t* seg, uint8_t* blocks) {
  if (b == 0) return -1;
  size_t c, l, bp;
  if (b <= 6144) {
    l = 0;
    c = 1;
    bp = b;
  } else {
    l = 24;
    c = (b + (6144 - 24) - 1) / (6144 - 24);
    bp = b + c * l;
  }
  int ip = -1;
  for (int i = 0; i < 188; ++i) {
    if (c * kQpp[i][0] >= bp) {
      ip = i;
      break;
    }
  }
  if (ip < 0) return -2;
  size_t kp = kQpp[ip][0], km = 0, cm = 0, cp;
  if (c == 1) {
    cp = 1;
  } else {
    int im = -1;
    for (int i = 187; i >= 0; --i) {
      if (kQpp[i][0] < kp) {
        im = i;
        break;
      }
    }
    if (im < 0) return -2;
    km = kQpp[im][0];
    cm = (c * kp - bp) / (kp - km);
    cp = c - cm;
  }
  size_t f = cp * kp + cm * km - bp;
  seg[0] = c;
  seg[1] = cp;
  seg[2] = cm;
  seg[3] = kp;
  seg[4] = km;
  seg[5] = f;
  seg[6] = l;
  if (blocks) {
    size_t cursor = 0, out = 0;
    for (size_t r = 0; r < c; ++r) {
      size_t kr = (r < cm) ? km : kp;
      size_t off = (r == 0) ? f : 0;
      size_t pay = kr - off - l;
      memset(blocks + out, 0, kr);
      memcpy(blocks + out + off, bits + cursor, pay);
      cursor += pay;
      if (l == 24) or_crc_append(OR_CRC24B, blocks + out + off, pay, blocks + out + off);
      out += kr;
    }
  }
  return 0;
}

/* -------------------------------------------------------------- trellis ---
 * RSC step of lte_turbo.cpp:333-341.  State bit 2 is the newest delay. */
static inline void rsc_step(unsigned s, unsigned u, unsigned* parity, unsigned* next) {
  unsigned s0 = s & 1u, s1 = (s >> 1) & 1u, s2 = (s >> 2) & 1u;
  unsigned fb = u ^ s1 ^ s2;
  *parity = fb ^ s0 ^ s2;
  *next = (fb << 2) | (s2 << 1) | s1;
}

/* Inputs that drive a state to zero in three steps (lte_turbo.cpp:343-353). */
static void term_inputs(unsigned s, unsigned in[3]) {
  for (int i = 0; i < 3; ++i) {
    unsigned p, n;
    in[i] = ((s >> 1) ^ (s >> 2)) & 1u;
    rsc_step(s, in[i], &p, &n);
    s = n;
  }
}

/* -------------------------------------------------------------- encoder ---
 * PCCC encoder with the reference's tail multiplexing (lte_turbo.cpp:624-688). */
int or_encode_cb(const uint8_t* cb, uint32_t k, uint32_t filler, uint8_t* d0, uint8_t* d1,
                 uint8_t* d2) {
  if (or_qpp_index(k) < 0) return -1;
  uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * k);
  or_qpp_perm(k, perm);
  unsigned su = 0, sl = 0, p, n;
  for (uint32_t i = 0; i < k; ++i) {
    rsc_step(su, cb[i] & 1u, &p, &n);
    su = n;
    d0[i] = (i < filler) ? 0 : (cb[i] & 1u);
    d1[i] = (i < filler) ? 0 : (uint8_t)p;
    rsc_step(sl, cb[perm[i]] & 1u, &p, &n);
    sl = n;
    d2[i] = (uint8_t)p;
  }
  free(perm);
  unsigned iu[3], il[3], xu[3], zu[3], xl[3], zl[3];
  term_inputs(su, iu);
  term_inputs(sl, il);
  for (int i = 0; i < 3; ++i) {
    xu[i] = iu[i];
    rsc_step(su, iu[i], &zu[i], &n);
    su = n;
    xl[i] = il[i];
    rsc_step(sl, il[i], &zl[i], &n);
    sl = n;
  }
  /* lte_turbo.cpp:664-677 */
  d0[k + 0] = (uint8_t)xu[0];
  d0[k + 1] = (uint8_t)zu[1];
  d0[k + 2] = (uint8_t)xu[1];
  d0[k + 3] = (uint8_t)zu[2];
  d1[k + 0] = (uint8_t)zu[0];
  d1[k + 1] = (uint8_t)xu[1];
  d1[k + 2] = (uint8_t)zu[2];
  d1[k + 3] = (uint8_t)xl[0];
  d2[k + 0] = (uint8_t)zl[0];
  d2[k + 1] = (uint8_t)xl[1];
  d2[k + 2] = (uint8_t)zl[1];
  d2[k + 3] = (uint8_t)xl[2];
  return 0;
}

/* -------------------------------------------------------- rate matching ---
 * E per code block, lte_turbo.cpp:167-185.  Returns -1 on the reference's
 * invalid_argument cases. */
int or_compute_e(size_t r, size_t g_total, size_t c, size_t nl, size_t qm, size_t* e) {
  if (c == 0 || nl == 0 || qm == 0) return -1;
  size_t nq = nl * qm;
  if (g_total % nq != 0) return -1;
  size_t gp = g_total / nq;
  size_t gam = gp % c;
  *e = (r + gam + 1 <= c) ? nq * (gp / c) : nq * ((gp + c - 1) / c);
  return 0;
}

/* 36.212 inter-column permutation (lte_turbo.cpp:49-52). */
static const uint8_t kColPerm[32] = {0, 16, 8, 24, 4, 20, 12, 28, 2, 18, 10, 26, 6, 22, 14, 30,
                                     1, 17, 9, 25, 5, 21, 13, 29, 3, 19, 11, 27, 7, 23, 15, 31};

/* Sub-block interleaver read-out maps: entry j is the stream index placed at
 * interleaver output j, or -1 for a dummy/NULL bit.  d(0)/d(1) follow
 * lte_turbo.cpp:195-219 (filler bits NULL), d(2) follows :221-236. */
static void subblock_maps(size_t d, size_t filler, int* m01, int* m2, size_t* rows_out) {
  size_t rows = (d + 31) / 32, kpi = rows * 32, dummy = kpi - d;
  for (size_t col = 0; col < 32; ++col) {
    for (size_t row = 0; row < rows; ++row) {
      size_t src = row * 32 + kColPerm[col]; /* position in the padded matrix */
      int v = -1;
      if (src >= dummy) {
        size_t i = src - dummy;
        v = (i < filler) ? -1 : (int)i;
      }
      m01[col * rows + row] = v;
    }
  }
  for (size_t j = 0; j < kpi; ++j) {
    size_t pos = (kColPerm[j / rows] + 32 * (j % rows) + 1) % kpi;
    m2[j] = (pos >= dummy) ? (int)(pos - dummy) : -1;
  }
  *rows_out = rows;
}

/* Circular buffer entry idx -> (stream, index) or -1 (lte_turbo.cpp:255-262). */
static int circ_entry(size_t idx, size_t kpi, const int* m01, const int* m2, int* stream) {
  if (idx < kpi) {
    *stream = 0;
    return m01[idx];
  }
  size_t local = idx - kpi;
  if ((local & 1u) == 0) {
    *stream = 1;
    return m01[local / 2];
  }
  *stream = 2;
  return m2[local / 2];
}

static size_t k0_of(size_t rows, size_t ncb, uint32_t rv) {
  return rows * (2 * ((ncb + 8 * rows - 1) / (8 * rows)) * rv + 2);
}

/* lte_turbo.cpp:250-278 */
void or_rate_match(const uint8_t* d0, const uint8_t* d1, const uint8_t* d2, size_t d,
                   size_t filler, size_t e, uint32_t rv, uint8_t* out) {
  size_t rows, kpi = ((d + 31) / 32) * 32, ncb = 3 * kpi;
  int* m01 = (int*)malloc(sizeof(int) * kpi);
  int* m2 = (int*)malloc(sizeof(int) * kpi);
  subblock_maps(d, filler, m01, m2, &rows);
  const uint8_t* streams[3] = {d0, d1, d2};
  size_t k0 = k0_of(rows, ncb, rv), k = 0, j = 0;
  while (k < e) {
    int st;
    int v = circ_entry((k0 + j) % ncb, kpi, m01, m2, &st);
    if (v >= 0) out[k++] = streams[st][v];
    ++j;
  }
  free(m01);
  free(m2);
}

/* Soft inverse: every received bit adds +-8 to its circular-buffer entry,
 * then entries fold back into d0/d1/d2 (lte_turbo.cpp:280-325, 722-734). */
void or_derate_match(const uint8_t* bits, size_t e, size_t d, size_t filler, uint32_t rv,
                     float* out) {
  size_t rows, kpi = ((d + 31) / 32) * 32, ncb = 3 * kpi;
  int* m01 = (int*)malloc(sizeof(int) * kpi);
  int* m2 = (int*)malloc(sizeof(int) * kpi);
  float* circ = (float*)calloc(ncb, sizeof(float));
  subblock_maps(d, filler, m01, m2, &rows);
  size_t k0 = k0_of(rows, ncb, rv), k = 0, j = 0;
  while (k < e) {
    size_t idx = (k0 + j) % ncb;
    int st;
    if (circ_entry(idx, kpi, m01, m2, &st) >= 0) {
      circ[idx] += bits[k] ? 8.0f : -8.0f;
      ++k;
    }
    ++j;
  }
  memset(out, 0, sizeof(float) * 3 * d);
  for (size_t i = 0; i < kpi; ++i) {
    /* entries never visited hold 0.0f, so adding them is a no-op */
    if (m01[i] >= 0) out[m01[i]] += circ[i];
    if (m01[i] >= 0) out[d + m01[i]] += circ[kpi + 2 * i];
    if (m2[i] >= 0) out[2 * d + m2[i]] += circ[kpi + 2 * i + 1];
  }
  free(circ);
  free(m01);
  free(m2);
}

/* Soft-input variant of the de-rate matcher: received LLR k (bit 1 > 0)
 * instead of +-8 per hard bit (lte_turbo.cpp:280-325, the same circular
 * buffer walk and accumulation order).  accumulate adds this transmission's
 * contribution to out (HARQ soft combining) instead of overwriting it. */
void or_derate_match_soft(const float* llr, size_t e, size_t d, size_t filler, uint32_t rv,
                          float* out, int accumulate) {
  size_t rows, kpi = ((d + 31) / 32) * 32, ncb = 3 * kpi;
  int* m01 = (int*)malloc(sizeof(int) * kpi);
  int* m2 = (int*)malloc(sizeof(int) * kpi);
  float* circ = (float*)calloc(ncb, sizeof(float));
  char* seen = (char*)calloc(ncb, 1);
  subblock_maps(d, filler, m01, m2, &rows);
  size_t k0 = k0_of(rows, ncb, rv), k = 0, j = 0;
  while (k < e) {
    size_t idx = (k0 + j) % ncb;
    int st;
    if (circ_entry(idx, kpi, m01, m2, &st) >= 0) {
      circ[idx] += llr[k];
      seen[idx] = 1;
      ++k;
    }
    ++j;
  }
  if (!accumulate) memset(out, 0, sizeof(float) * 3 * d);
  for (size_t idx = 0; idx < ncb; ++idx) {
    int st;
    int v = circ_entry(idx, kpi, m01, m2, &st);
    if (v >= 0 && seen[idx]) out[(size_t)st * d + (size_t)v] += circ[idx];
  }
  free(seen);
  free(circ);
  free(m01);
  free(m2);
}

/* Transport-block assembly after decoding (lte_turbo.cpp:958-979): hard
 * decisions (LLR >= 0 is 1, :488-494) of each code block's posterior, filler
 * bits dropped, CRC24B checked and dropped when c > 1 (a single block reports
 * ok), CRC24A over the concatenation, payload = all but the last 24 bits.
 * post: the blocks' K_r posteriors back to back.  Returns -1 when the blocks
 * do not add up to b + 24 bits. */
int or_assemble_tb(const float* post, const uint32_t* k, const uint32_t* filler, size_t c, size_t b,
                   uint8_t* bits_out, int* cb_ok, int* tb_ok) {
  size_t total = 0;
  for (size_t r = 0; r < c; ++r) total += k[r] - filler[r] - (c > 1 ? 24 : 0);
  if (total != b + 24) return -1;
  uint8_t* all = (uint8_t*)malloc(total + 1);
  uint8_t* blk = (uint8_t*)malloc(6144 + 1);
  size_t pos = 0, off = 0;
  for (size_t r = 0; r < c; ++r) {
    size_t n = k[r] - filler[r];
    for (size_t i = 0; i < n; ++i) blk[i] = post[off + filler[r] + i] >= 0.0f ? 1 : 0;
    cb_ok[r] = 1;
    if (c > 1) {
      cb_ok[r] = or_crc_check(OR_CRC24B, blk, n);
      n -= 24;
    }
    memcpy(all + pos, blk, n);
    pos += n;
    off += k[r];
  }
  *tb_ok = or_crc_check(OR_CRC24A, all, total);
  memcpy(bits_out, all, b);
  free(blk);
  free(all);
  return 0;
}

/* ------------------------------------------------------ Log-MAP decoder ---
 * max* with the reference's infinity rules (lte_turbo.cpp:380-390). */
static float maxstar(float x, float y) {
  if (!isfinite(x)) return y;
  if (!isfinite(y)) return x;
  float hi = (x < y) ? y : x;
  float lo = (y < x) ? y : x;
  return hi + log1pf(expf(lo - hi));
}

/* Subtract the largest entry from the finite entries (lte_turbo.cpp:392-402). */
static void normalize8(float* m) {
  float mx = m[0];
  for (int i = 1; i < 8; ++i)
    if (mx < m[i]) mx = m[i];
  if (!isfinite(mx)) return;
  for (int i = 0; i < 8; ++i)
    if (isfinite(m[i])) m[i] -= mx;
}

/* Tail beta: each state's forced three-step path to state 0
 * (lte_turbo.cpp:357-378). */
static void tail_beta(const float* ts, const float* tp, float* beta) {
  for (unsigned s = 0; s < 8; ++s) {
    unsigned in[3], cur = s, p, n;
    float metric = 0.0f;
    term_inputs(s, in);
    for (int t = 0; t < 3; ++t) {
      rsc_step(cur, in[t], &p, &n);
      cur = n;
      float a = 0.5f * ((in[t] ? 1.0f : -1.0f) * ts[t]);
      float b = 0.5f * ((p ? 1.0f : -1.0f) * tp[t]);
      metric += a + b;
    }
    beta[s] = (cur == 0) ? metric : -INFINITY;
  }
  normalize8(beta);
}

static inline float branch(unsigned u, unsigned p, float sa, float par) {
  return 0.5f * ((u ? 1.0f : -1.0f) * sa + (p ? 1.0f : -1.0f) * par);
}

/* lte_turbo.cpp:410-486 */
void or_siso(const float* sys, const float* parity, const float* apriori, uint32_t k,
             const float* ts, const float* tp, float* ext) {
  float* al = (float*)malloc(sizeof(float) * 8 * (k + 1));
  float* be = (float*)malloc(sizeof(float) * 8 * (k + 1));
  for (size_t i = 0; i < 8 * (size_t)(k + 1); ++i) al[i] = be[i] = -INFINITY;
  al[0] = 0.0f;
  tail_beta(ts, tp, be + 8 * (size_t)k);
  for (uint32_t t = 0; t < k; ++t) {
    const float sa = sys[t] + apriori[t];
    float* a0 = al + 8 * (size_t)t;
    float* a1 = a0 + 8;
    for (unsigned s = 0; s < 8; ++s) {
      if (!isfinite(a0[s])) continue;
      for (unsigned u = 0; u < 2; ++u) {
        unsigned p, n;
        rsc_step(s, u, &p, &n);
        float c = a0[s] + branch(u, p, sa, parity[t]);
        a1[n] = isfinite(a1[n]) ? maxstar(a1[n], c) : c;
      }
    }
    normalize8(a1);
  }
  for (uint32_t t = k; t-- > 0;) {
    const float sa = sys[t] + apriori[t];
    float* b0 = be + 8 * (size_t)t;
    float* b1 = b0 + 8;
    for (unsigned s = 0; s < 8; ++s) {
      for (unsigned u = 0; u < 2; ++u) {
        unsigned p, n;
        rsc_step(s, u, &p, &n);
        if (!isfinite(b1[n])) continue;
        float c = b1[n] + branch(u, p, sa, parity[t]);
        b0[s] = isfinite(b0[s]) ? maxstar(b0[s], c) : c;
      }
    }
    normalize8(b0);
  }
  for (uint32_t t = 0; t < k; ++t) {
    const float sa = sys[t] + apriori[t];
    float l[2] = {-INFINITY, -INFINITY};
    for (unsigned s = 0; s < 8; ++s) {
      for (unsigned u = 0; u < 2; ++u) {
        unsigned p, n;
        rsc_step(s, u, &p, &n);
        float m = al[8 * (size_t)t + s] + branch(u, p, sa, parity[t]) + be[8 * (size_t)(t + 1) + n];
        l[u] = isfinite(l[u]) ? maxstar(l[u], m) : m;
      }
    }
    ext[t] = (l[1] - l[0]) - sys[t] - apriori[t];
  }
  free(al);
  free(be);
}

/* Decoder-side tail read-out, reproduced as the reference does it
 * (lte_turbo.cpp:758-773); the lower-code values do not match the encoder's
 * multiplexing at :664-677 and parity requires keeping that. */
static void tail_llrs(const float* d0, const float* d1, const float* d2, uint32_t k,
                      float us[4], float up[4], float ls[4], float lp[4]) {
  us[0] = d0[k + 0];
  us[1] = d0[k + 2];
  us[2] = 0.0f;
  us[3] = 0.0f;
  up[0] = d1[k + 0];
  up[1] = d0[k + 1];
  up[2] = d1[k + 2];
  up[3] = 0.0f;
  ls[0] = d1[k + 3];
  ls[1] = d2[k + 0];
  ls[2] = d2[k + 2];
  ls[3] = 0.0f;
  lp[0] = d2[k + 1];
  lp[1] = d0[k + 3];
  lp[2] = d2[k + 3];
  lp[3] = 0.0f;
}

/* CRC24B early-stop test on hard decisions (lte_turbo.cpp:792-797). */
static int cb_crc_passes(const float* post, uint32_t k, uint32_t filler, uint8_t* scratch) {
  for (uint32_t i = 0; i < k; ++i) scratch[i] = (i < filler) ? 0 : (post[i] >= 0.0f);
  return or_crc_check(OR_CRC24B, scratch + filler, k - filler);
}

/* Turbo iteration loop, lte_turbo.cpp:736-813. */
int or_decode_cb(const float* d0, const float* d1, const float* d2, uint32_t k,
                 uint32_t filler, uint32_t max_iterations, int early_stop, float* post,
                 int* crc_ok, uint32_t* iterations_used) {
  if (or_qpp_index(k) < 0) return -1;
  uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * k);
  or_qpp_perm(k, perm);
  float* buf = (float*)calloc(6 * (size_t)k, sizeof(float));
  float *ap = buf, *e1 = buf + k, *e2 = buf + 2 * k, *si = buf + 3 * k, *ia = buf + 4 * k,
        *e2i = buf + 5 * k;
  uint8_t* bits = (uint8_t*)malloc(k);
  float us[4], up[4], ls[4], lp[4];
  tail_llrs(d0, d1, d2, k, us, up, ls, lp);
  for (uint32_t i = 0; i < k; ++i) si[i] = d0[perm[i]];
  for (uint32_t i = 0; i < k; ++i) post[i] = 0.0f;
  int ok = 0;
  for (uint32_t it = 0; it < max_iterations; ++it) {
    or_siso(d0, d1, ap, k, us, up, e1);
    for (uint32_t i = 0; i < k; ++i) ia[i] = e1[perm[i]];
    or_siso(si, d2, ia, k, ls, lp, e2i);
    for (uint32_t j = 0; j < k; ++j) e2[perm[j]] = e2i[j];
    for (uint32_t i = 0; i < k; ++i) {
      ap[i] = e2[i];
      post[i] = d0[i] + e1[i] + e2[i];
    }
    if (iterations_used) *iterations_used = it + 1;
    if (early_stop) {
      ok = cb_crc_passes(post, k, filler, bits);
      if (ok) break;
    }
  }
  if (crc_ok) *crc_ok = ok;
  free(bits);
  free(buf);
  free(perm);
  return 0;
}

/* ------------------------------- exact segment-parallel decoder (GPU path) ---
 * Not a reference function: the schedule the library's exact GPU decoder
 * runs, stated serially.  Every SISO half computes the reference's recursions
 * (lte_turbo.cpp:410-486, same operations in the same order) split into
 * M = K / seg_len segments that run side by side:
 *   1. speculation: each segment runs beta down and alpha up its own steps
 *      from a guess of its boundary metrics (the exact values of the same
 *      boundary in the previous half of the same kind, uniform at first),
 *      keeping the metrics at every kChunk-th step (checkpoints);
 *   2. fix-up rounds: each segment takes its neighbour's boundary value as it
 *      now stands (segment 0's alpha_0 and the last segment's tail beta are
 *      exact) and, if it differs from the value it started from, recomputes
 *      chunk by chunk until the recomputed metrics equal the stored
 *      checkpoint bit for bit: from there on the recursion is the same
 *      deterministic function of the same inputs, so nothing further changes.
 *      A segment that never meets its checkpoints hands a new boundary value
 *      on; rounds repeat until no boundary value changes;
 *   3. the extrinsics, per chunk, from the checkpoints: beta of the chunk
 *      recomputed downwards, alpha stepped upwards, the reference's 16-term
 *      LLR accumulation per step.
 * The metrics every step ends with are therefore those of the full-length
 * serial recursion, and the posteriors equal or_decode_cb's bit for bit
 * (test_exact_segmented_equals_serial).  `work` (may be NULL) accumulates
 * [0] halves, [1] fix-up rounds, [2] chunks recomputed counting the slowest
 * segment of each round (the cost on a GPU whose lanes run in lockstep),
 * [3] chunks of one full pass (for scale). */
enum { kChunk = 8 };

static void exact_alpha_step(const float* a0, float sa, float par, float* a1) {
  for (int i = 0; i < 8; ++i) a1[i] = -INFINITY;
  for (unsigned s = 0; s < 8; ++s) {
    if (!isfinite(a0[s])) continue;
    for (unsigned u = 0; u < 2; ++u) {
      unsigned p, n;
      rsc_step(s, u, &p, &n);
      float c = a0[s] + branch(u, p, sa, par);
      a1[n] = isfinite(a1[n]) ? maxstar(a1[n], c) : c;
    }
  }
  normalize8(a1);
}

static void exact_beta_step(const float* b1, float sa, float par, float* b0) {
  for (int i = 0; i < 8; ++i) b0[i] = -INFINITY;
  for (unsigned s = 0; s < 8; ++s) {
    for (unsigned u = 0; u < 2; ++u) {
      unsigned p, n;
      rsc_step(s, u, &p, &n);
      if (!isfinite(b1[n])) continue;
      float c = b1[n] + branch(u, p, sa, par);
      b0[s] = isfinite(b0[s]) ? maxstar(b0[s], c) : c;
    }
  }
  normalize8(b0);
}

/* beta over chunk w of segment q (steps [lo, hi)), from b (beta at hi) down to beta at lo */
static void exact_beta_chunk(const float* sys, const float* par, const float* ap, uint32_t lo, uint32_t hi,
                             float* b) {
  float nb[8];
  for (uint32_t t = hi; t-- > lo;) {
    exact_beta_step(b, sys[t] + ap[t], par[t], nb);
    memcpy(b, nb, sizeof nb);
  }
}

static void exact_alpha_chunk(const float* sys, const float* par, const float* ap, uint32_t lo, uint32_t hi,
                              float* a) {
  float na[8];
  for (uint32_t t = lo; t < hi; ++t) {
    exact_alpha_step(a, sys[t] + ap[t], par[t], na);
    memcpy(a, na, sizeof na);
  }
}

/* One SISO half, segment-parallel with exact boundaries.  nii: per segment
 * [alpha at its first step | beta after its last step], in/out (the guesses
 * in, the exact values out). */
static void exact_siso_seg(const float* sys, const float* par, const float* ap, uint32_t k, uint32_t seg,
                           const float* ts, const float* tp, float* nii, float* ext, uint64_t* work) {
  const uint32_t m = k / seg, nw = (seg + kChunk - 1) / kChunk;
  float* ack = (float*)malloc(sizeof(float) * 8 * (size_t)m * nw);
  float* bck = (float*)malloc(sizeof(float) * 8 * (size_t)m * nw);
  float* edge = (float*)malloc(sizeof(float) * 16 * (size_t)m);  /* [alpha at end | beta at start] */
  float* prev = (float*)malloc(sizeof(float) * 16 * (size_t)m);
  float tail[8], a0[8];
  tail_beta(ts, tp, tail);
  for (int i = 0; i < 8; ++i) a0[i] = i ? -INFINITY : 0.0f;
#define CLO(w) ((w) * kChunk)
#define CHI(w) (((w) + 1) * kChunk < seg ? ((w) + 1) * kChunk : seg)
  /* 1. speculation */
  for (uint32_t q = 0; q < m; ++q) {
    const uint32_t base = q * seg;
    float b[8], a[8];
    memcpy(b, q == m - 1 ? tail : nii + 16 * q + 8, sizeof b);
    for (uint32_t w = nw; w-- > 0;) {
      memcpy(bck + 8 * ((size_t)q * nw + w), b, sizeof b);
      exact_beta_chunk(sys, par, ap, base + CLO(w), base + CHI(w), b);
    }
    memcpy(edge + 16 * q + 8, b, sizeof b);
    memcpy(a, q == 0 ? a0 : nii + 16 * q, sizeof a);
    for (uint32_t w = 0; w < nw; ++w) {
      memcpy(ack + 8 * ((size_t)q * nw + w), a, sizeof a);
      exact_alpha_chunk(sys, par, ap, base + CLO(w), base + CHI(w), a);
    }
    memcpy(edge + 16 * q, a, sizeof a);
  }
  if (work) {
    work[0] += 1;
    work[3] += nw;
  }
  /* 2. fix-up rounds: every segment reads its neighbours' edges as the
   * previous round left them (lanes in lockstep), beta and alpha alike */
  for (;;) {
    memcpy(prev, edge, sizeof(float) * 16 * (size_t)m);
    int changed = 0;
    uint32_t slowest = 0;
    for (uint32_t q = 0; q < m; ++q) {
      const uint32_t base = q * seg;
      uint32_t done_b = 0, done_a = 0;
      /* beta: new value after the last step = next segment's beta at its start */
      float b[8], a[8];
      memcpy(b, q == m - 1 ? tail : prev + 16 * (q + 1) + 8, sizeof b);
      int met = 0;
      for (uint32_t w = nw; w-- > 0;) {
        float* ck = bck + 8 * ((size_t)q * nw + w);
        if (!memcmp(ck, b, sizeof b)) {
          met = 1;
          break;
        }
        memcpy(ck, b, sizeof b);
        exact_beta_chunk(sys, par, ap, base + CLO(w), base + CHI(w), b);
        ++done_b;
      }
      if (!met && memcmp(edge + 16 * q + 8, b, sizeof b)) {
        memcpy(edge + 16 * q + 8, b, sizeof b);
        changed = 1;
      }
      memcpy(a, q == 0 ? a0 : prev + 16 * (q - 1), sizeof a);
      met = 0;
      for (uint32_t w = 0; w < nw; ++w) {
        float* ck = ack + 8 * ((size_t)q * nw + w);
        if (!memcmp(ck, a, sizeof a)) {
          met = 1;
          break;
        }
        memcpy(ck, a, sizeof a);
        exact_alpha_chunk(sys, par, ap, base + CLO(w), base + CHI(w), a);
        ++done_a;
      }
      if (!met && memcmp(edge + 16 * q, a, sizeof a)) {
        memcpy(edge + 16 * q, a, sizeof a);
        changed = 1;
      }
      const uint32_t d = done_a > done_b ? done_a : done_b;
      if (d > slowest) slowest = d;
    }
    if (work) {
      work[1] += 1;
      work[2] += slowest;
    }
    if (!changed) break;
  }
  /* 3. extrinsics from the checkpoints */
  for (uint32_t q = 0; q < m; ++q) {
    const uint32_t base = q * seg;
    for (uint32_t w = 0; w < nw; ++w) {
      const uint32_t lo = base + CLO(w), hi = base + CHI(w);
      float bs[kChunk + 1][8], a[8], na[8];
      memcpy(bs[hi - lo], bck + 8 * ((size_t)q * nw + w), sizeof a);
      for (uint32_t t = hi; t-- > lo;) exact_beta_step(bs[t - lo + 1], sys[t] + ap[t], par[t], bs[t - lo]);
      memcpy(a, ack + 8 * ((size_t)q * nw + w), sizeof a);
      for (uint32_t t = lo; t < hi; ++t) {
        const float sa = sys[t] + ap[t];
        float l[2] = {-INFINITY, -INFINITY};
        for (unsigned s = 0; s < 8; ++s)
          for (unsigned u = 0; u < 2; ++u) {
            unsigned p, n;
            rsc_step(s, u, &p, &n);
            float mm = a[s] + branch(u, p, sa, par[t]) + bs[t - lo + 1][n];
            l[u] = isfinite(l[u]) ? maxstar(l[u], mm) : mm;
          }
        ext[t] = (l[1] - l[0]) - sys[t] - ap[t];
        exact_alpha_step(a, sa, par[t], na);
        memcpy(a, na, sizeof a);
      }
    }
    memcpy(nii + 16 * q, ack + 8 * ((size_t)q * nw), sizeof(float) * 8);
    memcpy(nii + 16 * q + 8, bck + 8 * ((size_t)q * nw + nw - 1), sizeof(float) * 8);
  }
#undef CLO
#undef CHI
  free(prev);
  free(edge);
  free(bck);
  free(ack);
}

int or_decode_exact_seg(const float* d0, const float* d1, const float* d2, uint32_t k, uint32_t filler,
                        uint32_t max_iterations, int early_stop, uint32_t seg_len, float* post, int* crc_ok,
                        uint32_t* iterations_used, uint64_t* work) {
  if (or_qpp_index(k) < 0 || seg_len == 0 || k % seg_len != 0) return -1;
  const uint32_t m = k / seg_len;
  uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * k);
  or_qpp_perm(k, perm);
  float* buf = (float*)calloc(6 * (size_t)k, sizeof(float));
  float *ap = buf, *e1 = buf + k, *e2 = buf + 2 * k, *si = buf + 3 * k, *ia = buf + 4 * k, *e2i = buf + 5 * k;
  float* nii = (float*)calloc(32 * (size_t)m, sizeof(float)); /* upper | lower halves; uniform guesses */
  uint8_t* bits = (uint8_t*)malloc(k);
  float us[4], up[4], ls[4], lp[4];
  tail_llrs(d0, d1, d2, k, us, up, ls, lp);
  for (uint32_t i = 0; i < k; ++i) si[i] = d0[perm[i]];
  for (uint32_t i = 0; i < k; ++i) post[i] = 0.0f;
  int ok = 0;
  for (uint32_t it = 0; it < max_iterations; ++it) {
    exact_siso_seg(d0, d1, ap, k, seg_len, us, up, nii, e1, work);
    for (uint32_t i = 0; i < k; ++i) ia[i] = e1[perm[i]];
    exact_siso_seg(si, d2, ia, k, seg_len, ls, lp, nii + 16 * (size_t)m, e2i, work);
    for (uint32_t j = 0; j < k; ++j) e2[perm[j]] = e2i[j];
    for (uint32_t i = 0; i < k; ++i) {
      ap[i] = e2[i];
      post[i] = d0[i] + e1[i] + e2[i];
    }
    if (iterations_used) *iterations_used = it + 1;
    if (early_stop) {
      ok = cb_crc_passes(post, k, filler, bits);
      if (ok) break;
    }
  }
  if (crc_ok) *crc_ok = ok;
  free(bits);
  free(nii);
  free(buf);
  free(perm);
  return 0;
}

/* ------------------------------------------- windowed decoder (GPU path) ---
 * The algorithm the batched HIP kernel runs; not a reference function.  The
 * code block is split into M = K / seg_len segments.  Each SISO half runs an
 * independent forward/backward pass per segment; a segment's initial alpha
 * (beta) is the alpha (beta) the neighbouring segment produced at the shared
 * boundary in the previous iteration (next-iteration initialisation), uniform
 * in the first iteration.  Segment 0 starts from alpha_0 = state 0 and the
 * last segment from the reference's tail beta, so seg_len == K reduces to an
 * ordinary full-length Log-MAP pass.
 *
 * The arithmetic is the GPU kernel's, written out on the CPU:
 *   log domain, metrics in base-2 units (natural-log metric * log2e):
 *                x = sa*(log2e/2), y = par*(log2e/2), g1 = x + y, g2 = x - y
 *                (branch metrics +-g1, +-g2),
 *                max*(a,b) = max(a,b) + log2(1 + 2^-|a-b|),
 *                metrics renormalised to state 0's metric once per chunk of 8
 *                steps of a pass (chunks aligned at the pass's first step:
 *                beta after the chunk's lowest step, alpha after its 8th),
 *                extrinsic = ln2 * (LSE{alpha~(n)+beta(n) : u(n)=1}
 *                          - LSE{...: u(n)=0}) - sa,
 *                LSE(v) = m + log2(sum 2^(v_i-m)).
 *   linear domain: x = sa/2, y = par/2; branch weights
 *                exp(gamma - |x| - |y|) = U_u * Q_c with 2^(-|sa| log2e),
 *                metrics rescaled by a power of two to a maximum in [0.5,1)
 *                each step, extrinsic = ln(sum_{u=1} B(n) beta(n)) -
 *                ln(sum_{u=0} B(n) beta(n)) with B(n) the alpha update
 *                without its U factor; sums below FLT_MIN are clamped.
 */
#define OR_NEG (-1.0e30f)
static const float kLog2e = 1.44269504088896340736f;
static const float kLn2 = 0.693147180559945309417f;

/* For next state n: predecessors 2(n&3) and 2(n&3)+1, input u(n), parity of
 * the branch from each predecessor. */
static unsigned tr_u(unsigned n) { return ((n >> 2) ^ (n >> 1) ^ n) & 1u; }
static unsigned tr_pa(unsigned n) { return (n & 3u) << 1; }
static unsigned tr_cpa(unsigned n) { return ((n >> 2) ^ (n >> 1)) & 1u; }

static const float kHalfLog2e = 0.721347520444481703680f;

static inline float lg_maxstar(float a, float b) {
  float m = fmaxf(a, b);
  return m + log2f(1.0f + exp2f(-fabsf(a - b)));
}

/* gamma for (u, c) from g1 = x+y, g2 = x-y */
static inline float gsel(unsigned u, unsigned c, float g1, float g2) {
  if (u) return c ? g1 : g2;
  return c ? -g2 : -g1;
}

/* metrics relative to state 0 */
static void lg_norm(float* v) {
  const float m = v[0];
  for (int i = 0; i < 8; ++i) v[i] = v[i] - m;
}

static inline void lin_weights(float x, float y, float* u1, float* u0, float* q1, float* q0) {
  float ex = exp2f(-fabsf(x + x) * kLog2e);
  float ey = exp2f(-fabsf(y + y) * kLog2e);
  *u1 = (x >= 0.0f) ? 1.0f : ex;
  *u0 = (x >= 0.0f) ? ex : 1.0f;
  *q1 = (y >= 0.0f) ? 1.0f : ey;
  *q0 = (y >= 0.0f) ? ey : 1.0f;
}

static void lin_norm(float* v) {
  float m = v[0];
  for (int i = 1; i < 8; ++i) m = fmaxf(m, v[i]);
  if (m > 0.0f) {
    int e;
    (void)frexpf(m, &e);
    for (int i = 0; i < 8; ++i) v[i] = ldexpf(v[i], -e);
  }
}

/* One segment of one SISO half.  in_sa/in_par point at the segment's first
 * step and are valid on [-ga, len + gb): ga warm-up steps before the segment
 * and gb after it.  ain is alpha at -ga, bin is beta at len + gb.  Outputs:
 * ext[0..len), aexp = alpha at len - g (next segment's warm-up start) and
 * bexp = beta at g (previous segment's warm-up start). */
static void seg_siso(const float* in_sa, const float* in_par, int len, int ga, int gb, int g,
                     int lin, const float* ain, const float* bin, float* ext, float* aexp,
                     float* bexp, float* bstore /* 8*(len+1) */) {
  float b[8];
  memcpy(b, bin, sizeof b);
  if (gb == 0) memcpy(bstore + 8 * (size_t)len, b, sizeof b);
  for (int t = len + gb; t-- > 0;) {
    float x = 0.5f * in_sa[t], y = 0.5f * in_par[t];
    float nb[8];
    if (!lin) {
      const float x2 = in_sa[t] * kHalfLog2e, y2 = in_par[t] * kHalfLog2e;
      float g1 = x2 + y2, g2 = x2 - y2;
      for (unsigned p = 0; p < 8; ++p) {
        unsigned c0, n0, c1, n1;
        rsc_step(p, 0, &c0, &n0);
        rsc_step(p, 1, &c1, &n1);
        nb[p] = lg_maxstar(b[n0] + gsel(0, c0, g1, g2), b[n1] + gsel(1, c1, g1, g2));
      }
      /* chunks of the own pass start at local step 0, those of the warm-up
       * over the next segment at local step len */
      if ((t >= len ? t - len : t) % 8 == 0) lg_norm(nb);
    } else {
      float u1, u0, q1, q0;
      lin_weights(x, y, &u1, &u0, &q1, &q0);
      float w[2][2] = {{u0 * q0, u0 * q1}, {u1 * q0, u1 * q1}};
      for (unsigned p = 0; p < 8; ++p) {
        unsigned c0, n0, c1, n1;
        rsc_step(p, 0, &c0, &n0);
        rsc_step(p, 1, &c1, &n1);
        nb[p] = fmaf(w[1][c1], b[n1], w[0][c0] * b[n0]);
      }
      lin_norm(nb);
    }
    memcpy(b, nb, sizeof b);
    if (t <= len) memcpy(bstore + 8 * (size_t)t, b, sizeof b);
    if (t == g) memcpy(bexp, b, sizeof b);
  }
  float a[8];
  memcpy(a, ain, sizeof a);
  for (int t = -ga; t < len; ++t) {
    if (t == len - g) memcpy(aexp, a, sizeof a);
    float x = 0.5f * in_sa[t], y = 0.5f * in_par[t];
    const float* bn = bstore + 8 * (size_t)(t + 1);
    float na[8];
    if (!lin) {
      const float x2 = in_sa[t] * kHalfLog2e, y2 = in_par[t] * kHalfLog2e;
      float g1 = x2 + y2, g2 = x2 - y2;
      float tv[8];
      for (unsigned n = 0; n < 8; ++n) {
        unsigned u = tr_u(n), pa = tr_pa(n), ca = tr_cpa(n);
        na[n] = lg_maxstar(a[pa] + gsel(u, ca, g1, g2), a[pa + 1] + gsel(u, ca ^ 1u, g1, g2));
        tv[n] = (t >= 0) ? na[n] + bn[n] : 0.0f;
      }
      if (t >= 0) {
        float m1 = OR_NEG, m0 = OR_NEG;
        for (unsigned n = 0; n < 8; ++n) {
          if (tr_u(n)) m1 = fmaxf(m1, tv[n]);
          else m0 = fmaxf(m0, tv[n]);
        }
        float s1 = 0.0f, s0 = 0.0f;
        for (unsigned n = 0; n < 8; ++n) {
          if (tr_u(n)) s1 += exp2f(tv[n] - m1);
          else s0 += exp2f(tv[n] - m0);
        }
        float l1 = m1 + log2f(s1), l0 = m0 + log2f(s0);
        ext[t] = (l1 - l0) * kLn2 - in_sa[t];
      }
      /* chunks of the warm-up start at local step -ga, the own pass's at 0 */
      if ((t < 0 ? t + ga : t) % 8 == 7) lg_norm(na);
    } else {
      float u1, u0, q1, q0;
      lin_weights(x, y, &u1, &u0, &q1, &q0);
      float q[2] = {q0, q1}, uu[2] = {u0, u1};
      float s1 = 0.0f, s0 = 0.0f;
      for (unsigned n = 0; n < 8; ++n) {
        unsigned u = tr_u(n), pa = tr_pa(n), ca = tr_cpa(n);
        float bb = fmaf(a[pa + 1], q[ca ^ 1u], a[pa] * q[ca]);
        if (t >= 0) {
          if (u) s1 = fmaf(bb, bn[n], s1);
          else s0 = fmaf(bb, bn[n], s0);
        }
        na[n] = uu[u] * bb;
      }
      if (t >= 0) {
        s1 = fmaxf(s1, 1.17549435e-38f);
        s0 = fmaxf(s0, 1.17549435e-38f);
        ext[t] = kLn2 * (log2f(s1) - log2f(s0));
      }
      lin_norm(na);
    }
    memcpy(a, na, sizeof a);
  }
}

float or_round_half(float x) {
  if (x > 65504.0f) x = 65504.0f;
  if (x < -65504.0f) x = -65504.0f;
  const float ax = fabsf(x);
  if (ax < 6.103515625e-05f) { /* fp16 subnormals: multiples of 2^-24 */
    const float q = rintf(ax * 16777216.0f) * 5.9604644775390625e-08f;
    return copysignf(q, x);
  }
  /* keep 10 of the 23 fraction bits, ties to even; a carry moves into the
   * exponent as it should */
  uint32_t u;
  memcpy(&u, &x, sizeof u);
  u += 0x0FFFu + ((u >> 13) & 1u);
  u &= ~0x1FFFu;
  float r;
  memcpy(&r, &u, sizeof r);
  return r;
}

int or_decode_window(const float* d0, const float* d1, const float* d2, uint32_t k,
                     uint32_t filler, uint32_t max_iterations, int early_stop,
                     uint32_t seg_len, uint32_t warmup, int lin, uint32_t ex_bits, float* post,
                     int* crc_ok, uint32_t* iterations_used) {
  if (or_qpp_index(k) < 0 || seg_len == 0 || k % seg_len != 0) return -1;
  if (ex_bits != 32 && ex_bits != 16) return -1;
  const uint32_t nseg = k / seg_len;
  /* warmup >= k: every segment's warm-ups reach the block ends (alpha from
   * alpha_0 at step 0, beta from the tail beta at step K), so each segment
   * computes the full-length recursions' metrics and the result equals the
   * one-segment decode (seg_len == k) bit for bit -- a check of the segment
   * plumbing; otherwise the warm-up stays inside the neighbouring segment */
  const int full = nseg > 1 && warmup >= k;
  if (nseg > 1 && !full && warmup >= seg_len) return -1;
  const int g = (nseg > 1 && !full) ? (int)warmup : 0;
  uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * k);
  or_qpp_perm(k, perm);
  /* E1 | E2 as held between halves | ext | E2 as computed */
  float* e = (float*)calloc(4 * (size_t)k, sizeof(float));
  float *e1 = e, *e2 = e + k, *ex = e + 2 * k, *e2raw = e + 3 * (size_t)k;
  float* sa = (float*)malloc(sizeof(float) * k);
  float* si = (float*)malloc(sizeof(float) * k);
  float* bst = (float*)malloc(sizeof(float) * 8 * (seg_len + 1));
  /* boundary metrics [half][alpha|beta][segment][8]: alpha at mS-g (m>0),
   * beta at (m+1)S+g (m<M-1) */
  const size_t nb = 2 * 2 * (size_t)nseg * 8;
  float* ab = (float*)malloc(sizeof(float) * nb);
  float* nab = (float*)malloc(sizeof(float) * nb);
  uint8_t* bits = (uint8_t*)malloc(k);
  float us[4], up[4], ls[4], lp[4], tb[2][8];
  tail_llrs(d0, d1, d2, k, us, up, ls, lp);
  tail_beta(us, up, tb[0]);
  tail_beta(ls, lp, tb[1]);
  for (int h = 0; h < 2; ++h)
    for (int s = 0; s < 8; ++s) {
      float v = isfinite(tb[h][s]) ? tb[h][s] : OR_NEG;
      tb[h][s] = lin ? exp2f(v * kLog2e) : v * kLog2e;
    }
  const float uni = lin ? 1.0f : 0.0f, zero = lin ? 0.0f : OR_NEG;
  for (size_t i = 0; i < nb; ++i) ab[i] = uni;
  for (int h = 0; h < 2; ++h) {
    float* a0 = ab + ((size_t)(h * 2 + 0) * nseg) * 8;
    float* bl = ab + ((size_t)(h * 2 + 1) * nseg + nseg - 1) * 8;
    for (int s = 0; s < 8; ++s) {
      a0[s] = (s == 0) ? uni : zero;
      bl[s] = tb[h][s];
    }
  }
  for (uint32_t i = 0; i < k; ++i) si[i] = d0[perm[i]];
  for (uint32_t i = 0; i < k; ++i) post[i] = 0.0f;
  int ok = 0;
  for (uint32_t it = 0; it < max_iterations; ++it) {
    for (int h = 0; h < 2; ++h) {
      const float* par = h ? d2 : d1;
      for (uint32_t i = 0; i < k; ++i) sa[i] = h ? (si[i] + e1[perm[i]]) : (d0[i] + e2[i]);
      memcpy(nab, ab, sizeof(float) * nb);
      for (uint32_t m = 0; m < nseg; ++m) {
        const int ga = full ? (int)(m * seg_len) : ((m > 0) ? g : 0);
        const int gb = full ? (int)((nseg - 1 - m) * seg_len) : ((m + 1 < nseg) ? g : 0);
        float* ain = ab + ((size_t)(h * 2 + 0) * nseg + (full ? 0 : m)) * 8;
        float* bin = ab + ((size_t)(h * 2 + 1) * nseg + (full ? nseg - 1 : m)) * 8;
        float aexp[8], bexp[8];
        seg_siso(sa + (size_t)m * seg_len, par + (size_t)m * seg_len, (int)seg_len, ga, gb, g, lin,
                 ain, bin, ex + (size_t)m * seg_len, aexp, bexp, bst);
        if (m + 1 < nseg) memcpy(nab + ((size_t)(h * 2 + 0) * nseg + m + 1) * 8, aexp, sizeof aexp);
        if (m > 0) memcpy(nab + ((size_t)(h * 2 + 1) * nseg + m - 1) * 8, bexp, sizeof bexp);
      }
      memcpy(ab, nab, sizeof(float) * nb);
      /* the held extrinsics feed the next half; the posterior takes the
       * fresh lower extrinsic */
      if (h == 0) {
        for (uint32_t i = 0; i < k; ++i) e1[i] = (ex_bits == 16) ? or_round_half(ex[i]) : ex[i];
      } else {
        for (uint32_t j = 0; j < k; ++j) {
          e2[perm[j]] = (ex_bits == 16) ? or_round_half(ex[j]) : ex[j];
          e2raw[perm[j]] = ex[j];
        }
      }
    }
    for (uint32_t i = 0; i < k; ++i) post[i] = d0[i] + e1[i] + e2raw[i];
    if (iterations_used) *iterations_used = it + 1;
    if (early_stop) {
      ok = cb_crc_passes(post, k, filler, bits);
      if (ok) break;
    }
  }
  if (crc_ok) *crc_ok = ok;
  free(bits);
  free(nab);
  free(ab);
  free(bst);
  free(si);
  free(sa);
  free(e);
  free(perm);
  return 0;
}

/* ------------------------------------------------------------ brute force ---
 * Exact APP LLR of the upper code: sum over all 2^K inputs of
 * exp(sum of branch metrics), each path continued by its forced termination
 * and weighted by the tail branch metrics, in float64.  The terminated
 * trellis ends in state 0, which is what the reference's tail beta encodes. */
void or_brute_app(const float* sys, const float* parity, const float* apriori, uint32_t k,
                  const float* ts, const float* tp, double* app) {
  double* num = (double*)calloc(k, sizeof(double));
  double* den = (double*)calloc(k, sizeof(double));
  double* lw = (double*)malloc(sizeof(double) << k);
  double mx = -INFINITY;
  for (uint64_t w = 0; w < (1ull << k); ++w) {
    unsigned s = 0, p, n;
    double metric = 0.0;
    for (uint32_t t = 0; t < k; ++t) {
      unsigned u = (w >> t) & 1u;
      rsc_step(s, u, &p, &n);
      metric += 0.5 * ((u ? 1.0 : -1.0) * ((double)sys[t] + apriori[t]) + (p ? 1.0 : -1.0) * parity[t]);
      s = n;
    }
    unsigned in[3];
    term_inputs(s, in);
    for (int t = 0; t < 3; ++t) {
      rsc_step(s, in[t], &p, &n);
      metric += 0.5 * ((in[t] ? 1.0 : -1.0) * ts[t]) + 0.5 * ((p ? 1.0 : -1.0) * tp[t]);
      s = n;
    }
    lw[w] = metric;
    if (metric > mx) mx = metric;
  }
  for (uint64_t w = 0; w < (1ull << k); ++w) {
    double pr = exp(lw[w] - mx);
    for (uint32_t t = 0; t < k; ++t) {
      if ((w >> t) & 1u) num[t] += pr;
      else den[t] += pr;
    }
  }
  for (uint32_t t = 0; t < k; ++t) app[t] = log(num[t]) - log(den[t]);
  free(lw);
  free(num);
  free(den);
}
