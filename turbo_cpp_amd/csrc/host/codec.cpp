// Host side of the turbo codec chain: QPP, CRC, segmentation, encoder, rate
// matching and the CPU Log-MAP decoder, with the conventions of
// Wiseung/Turbo_CPP src/lte_turbo.cpp (cited per function) reproduced
// bit-exactly, including those that differ from 36.212.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>

#include "codec_internal.h"
#include "turbo_cpp/c_api.h"

namespace turbo_cpp {
namespace detail {

const QppEntry kQppTable[188] = {
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

int qpp_index(std::size_t k) {
  // The table is sorted by K: binary search.
  const auto* end = kQppTable + 188;
  const auto* it = std::lower_bound(kQppTable, end, k,
                                    [](const QppEntry& e, std::size_t v) { return e.k < v; });
  return (it != end && it->k == k) ? static_cast<int>(it - kQppTable) : -1;
}

namespace {

// ---------------------------------------------------------------- CRC ------
// Generator polynomials of the reference's tap arrays (lte_turbo.cpp:104-116)
// as the low N bits of G(D); D^N implicit.  Not the 36.212 polynomials.
struct CrcSpec {
  unsigned len;
  std::uint32_t poly;
};
constexpr CrcSpec kCrc[4] = {{24, 0xC199BFu}, {24, 0xC00063u}, {16, 0x8811u}, {8, 0xD9u}};

// Byte-at-a-time table for the MSB-first shift register of a CRC of length
// len: entry v is the register after shifting in the 8 message bits of v
// from a zero register.
struct CrcTable {
  std::array<std::uint32_t, 256> t{};
  constexpr explicit CrcTable(CrcSpec s) {
    const std::uint32_t top = 1u << (s.len - 1), mask = (s.len == 32) ? ~0u : ((1u << s.len) - 1u);
    for (unsigned v = 0; v < 256; ++v) {
      std::uint32_t r = 0;
      for (int b = 7; b >= 0; --b) {
        const std::uint32_t fb = ((r & top) ? 1u : 0u) ^ ((v >> b) & 1u);
        r = (r << 1) & mask;
        if (fb) r ^= s.poly;
      }
      t[v] = r;
    }
  }
};
constexpr CrcTable kCrcTables[4] = {CrcTable(kCrc[0]), CrcTable(kCrc[1]), CrcTable(kCrc[2]),
                                    CrcTable(kCrc[3])};

// Remainder of m(D) * D^N mod G(D) for the 0/1 bits m (the long division of
// lte_turbo.cpp:118-135), eight bits per table step.
std::uint32_t crc_register(int kind, std::span<const std::uint8_t> bits) {
  const CrcSpec s = kCrc[kind];
  const auto& tab = kCrcTables[kind].t;
  const std::uint32_t mask = (1u << s.len) - 1u, top = 1u << (s.len - 1);
  std::uint32_t r = 0;
  std::size_t i = 0;
  for (; i + 8 <= bits.size(); i += 8) {
    unsigned byte = 0;
    for (int b = 0; b < 8; ++b) byte = (byte << 1) | (bits[i + b] & 1u);
    const unsigned idx = ((r >> (s.len - 8)) ^ byte) & 0xFFu;
    r = ((r << 8) & mask) ^ tab[idx];
  }
  for (; i < bits.size(); ++i) {
    const std::uint32_t fb = ((r & top) ? 1u : 0u) ^ (bits[i] & 1u);
    r = (r << 1) & mask;
    if (fb) r ^= s.poly;
  }
  return r;
}

std::vector<std::uint8_t> crc_append(int kind, std::span<const std::uint8_t> bits) {
  const unsigned len = kCrc[kind].len;
  const std::uint32_t r = crc_register(kind, bits);
  std::vector<std::uint8_t> out(bits.begin(), bits.end());
  for (unsigned j = 0; j < len; ++j) out.push_back(static_cast<std::uint8_t>((r >> (len - 1 - j)) & 1u));
  return out;
}

// False when shorter than N (lte_turbo.cpp:140-142); otherwise the trailing N
// bits must equal the remainder of the leading bits.
bool crc_check(int kind, std::span<const std::uint8_t> bits) {
  const unsigned len = kCrc[kind].len;
  if (bits.size() < len) return false;
  const std::uint32_t r = crc_register(kind, bits.first(bits.size() - len));
  for (unsigned j = 0; j < len; ++j) {
    if (bits[bits.size() - len + j] != ((r >> (len - 1 - j)) & 1u)) return false;
  }
  return true;
}

std::size_t ceil_div(std::size_t a, std::size_t b) { return (a + b - 1) / b; }

// -------------------------------------------------------- rate matching ---
constexpr std::array<std::uint8_t, 32> kColumnPermutation = {
    0, 16, 8, 24, 4, 20, 12, 28, 2, 18, 10, 26, 6, 22, 14, 30,
    1, 17, 9, 25, 5, 21, 13, 29, 3, 19, 11, 27, 7, 23, 15, 31};

// Circular buffer of one code block as a table: entry j holds the source
// (stream << 16 | index) feeding position j, or -1 for <NULL>/dummy bits.
// Systematic part: sub-block interleaver with filler bits NULL in d(0) and
// d(1) (lte_turbo.cpp:195-219); d(2) interleaver of :221-236; parity bits
// interlaced (:255-262); N_cb = 3 K_pi since n_soft is not used (:243).
struct CircularBuffer {
  std::size_t rows = 0, kpi = 0, ncb = 0;
  std::vector<std::int32_t> src;
};

CircularBuffer make_circular_buffer(std::size_t d, std::size_t filler) {
  CircularBuffer cb;
  cb.rows = ceil_div(d, 32);
  cb.kpi = cb.rows * 32;
  cb.ncb = 3 * cb.kpi;
  cb.src.assign(cb.ncb, -1);
  const std::size_t dummy = cb.kpi - d;
  for (std::size_t j = 0; j < cb.kpi; ++j) {
    // d(0), d(1): column-wise read of the column-permuted matrix
    const std::size_t col = j / cb.rows, row = j % cb.rows;
    const std::size_t cell = row * 32 + kColumnPermutation[col];
    std::int32_t i01 = -1;
    if (cell >= dummy && cell - dummy >= filler) i01 = static_cast<std::int32_t>(cell - dummy);
    // d(2): pi(j) = (P(j / R) + 32 (j mod R) + 1) mod K_pi
    const std::size_t cell2 = (kColumnPermutation[j / cb.rows] + 32 * (j % cb.rows) + 1) % cb.kpi;
    const std::int32_t i2 = (cell2 >= dummy) ? static_cast<std::int32_t>(cell2 - dummy) : -1;
    if (i01 >= 0) {
      cb.src[j] = i01;
      cb.src[cb.kpi + 2 * j] = (1 << 16) | i01;
    }
    if (i2 >= 0) cb.src[cb.kpi + 2 * j + 1] = (2 << 16) | i2;
  }
  return cb;
}

// k0 = R (2 ceil(N_cb / (8 R)) rv + 2) (lte_turbo.cpp:263-264).
std::size_t start_position(const CircularBuffer& cb, std::uint32_t rv) {
  return cb.rows * (2 * ceil_div(cb.ncb, 8 * cb.rows) * rv + 2);
}

std::vector<std::uint8_t> rate_match_bits(std::span<const std::uint8_t> d0, std::span<const std::uint8_t> d1,
                                          std::span<const std::uint8_t> d2, std::size_t filler,
                                          std::size_t e, std::uint32_t rv) {
  // lte_turbo.cpp:250-278
  const auto cb = make_circular_buffer(d0.size(), filler);
  const std::span<const std::uint8_t> streams[3] = {d0, d1, d2};
  std::vector<std::uint8_t> out;
  out.reserve(e);
  std::size_t j = start_position(cb, rv) % cb.ncb;
  while (out.size() < e) {
    const std::int32_t s = cb.src[j];
    if (s >= 0) out.push_back(streams[s >> 16][s & 0xFFFF]);
    if (++j == cb.ncb) j = 0;
  }
  return out;
}

std::vector<float> derate_match(std::span<const std::uint8_t> bits, std::size_t d, std::size_t filler,
                                std::uint32_t rv) {
  // lte_turbo.cpp:280-325: +-8 per received bit accumulated on its circular
  // buffer entry, then each entry lands on its stream position.
  const auto cb = make_circular_buffer(d, filler);
  std::vector<float> acc(cb.ncb, 0.0F);
  std::size_t j = start_position(cb, rv) % cb.ncb;
  for (std::size_t k = 0; k < bits.size();) {
    if (cb.src[j] >= 0) {
      acc[j] += bits[k] ? 8.0F : -8.0F;
      ++k;
    }
    if (++j == cb.ncb) j = 0;
  }
  std::vector<float> out(3 * d, 0.0F);
  for (std::size_t p = 0; p < cb.ncb; ++p) {
    const std::int32_t s = cb.src[p];
    if (s >= 0) out[(s >> 16) * d + (s & 0xFFFF)] += acc[p];
  }
  return out;
}

// ------------------------------------------------------ Log-MAP decoder ---
// Jacobian logarithm with the reference's infinity rules (lte_turbo.cpp:380-390).
inline float jacobian(float x, float y) {
  if (!std::isfinite(x)) return y;
  if (!std::isfinite(y)) return x;
  const float hi = std::max(x, y), lo = std::min(x, y);
  return hi + std::log1pf(std::exp(lo - hi));
}

// lte_turbo.cpp:392-402
inline void renormalise(float* m) {
  const float mx = *std::max_element(m, m + 8);
  if (!std::isfinite(mx)) return;
  for (int s = 0; s < 8; ++s)
    if (std::isfinite(m[s])) m[s] -= mx;
}

constexpr float kNegInf = -std::numeric_limits<float>::infinity();

// Branch metric 1/2[(2u-1)(Ls+La) + (2p-1)Lp] (lte_turbo.cpp:437-438).
inline float branch_metric(unsigned u, unsigned p, float sa, float lp) {
  return 0.5F * ((u ? 1.0F : -1.0F) * sa + (p ? 1.0F : -1.0F) * lp);
}

struct SisoWork {
  std::vector<float> alpha, beta;
};

// One SISO half (lte_turbo.cpp:410-486).  sys/apriori K, parity at least K.
void siso(std::span<const float> sys, std::span<const float> parity, std::span<const float> apriori,
          const float* tail_sys, const float* tail_par, std::span<float> ext, SisoWork& w) {
  const std::size_t k = sys.size();
  w.alpha.assign(8 * (k + 1), kNegInf);
  w.beta.assign(8 * (k + 1), kNegInf);
  w.alpha[0] = 0.0F;
  tail_beta_log(tail_sys, tail_par, &w.beta[8 * k]);
  for (std::size_t t = 0; t < k; ++t) {
    const float sa = sys[t] + apriori[t];
    const float* a0 = &w.alpha[8 * t];
    float* a1 = &w.alpha[8 * (t + 1)];
    for (unsigned s = 0; s < 8; ++s) {
      if (!std::isfinite(a0[s])) continue;
      for (unsigned u = 0; u < 2; ++u) {
        const auto tr = rsc(s, u);
        const float c = a0[s] + branch_metric(u, tr.parity, sa, parity[t]);
        a1[tr.next] = std::isfinite(a1[tr.next]) ? jacobian(a1[tr.next], c) : c;
      }
    }
    renormalise(a1);
  }
  for (std::size_t t = k; t-- > 0;) {
    const float sa = sys[t] + apriori[t];
    const float* b1 = &w.beta[8 * (t + 1)];
    float* b0 = &w.beta[8 * t];
    for (unsigned s = 0; s < 8; ++s) {
      for (unsigned u = 0; u < 2; ++u) {
        const auto tr = rsc(s, u);
        if (!std::isfinite(b1[tr.next])) continue;
        const float c = b1[tr.next] + branch_metric(u, tr.parity, sa, parity[t]);
        b0[s] = std::isfinite(b0[s]) ? jacobian(b0[s], c) : c;
      }
    }
    renormalise(b0);
  }
  for (std::size_t t = 0; t < k; ++t) {
    const float sa = sys[t] + apriori[t];
    float l[2] = {kNegInf, kNegInf};
    for (unsigned s = 0; s < 8; ++s) {
      for (unsigned u = 0; u < 2; ++u) {
        const auto tr = rsc(s, u);
        const float m = w.alpha[8 * t + s] + branch_metric(u, tr.parity, sa, parity[t]) +
                        w.beta[8 * (t + 1) + tr.next];
        l[u] = std::isfinite(l[u]) ? jacobian(l[u], m) : m;
      }
    }
    ext[t] = (l[1] - l[0]) - sys[t] - apriori[t];
  }
}

}  // namespace

std::size_t compute_e(std::size_t r, std::size_t g_total, std::size_t c, std::size_t num_layers,
                      std::size_t mod_order) {
  // lte_turbo.cpp:167-185
  if (c == 0) throw std::invalid_argument("code block count must be positive");
  if (num_layers == 0 || mod_order == 0)
    throw std::invalid_argument("num_layers and mod_order must be positive");
  const std::size_t nq = num_layers * mod_order;
  if (g_total % nq != 0) throw std::invalid_argument("g_total must be divisible by num_layers * mod_order");
  const std::size_t gp = g_total / nq, gamma = gp % c;
  return (r + gamma + 1 <= c) ? nq * (gp / c) : nq * ceil_div(gp, c);
}

DerateGeometry derate_geometry(std::size_t d, std::size_t filler) {
  const auto cb = make_circular_buffer(d, filler);
  DerateGeometry g;
  g.rank.assign(3 * d, 0xFFFFU);
  std::vector<std::uint32_t> before(cb.ncb + 1, 0);  // non-NULL entries in [0, j)
  std::uint32_t n = 0;
  for (std::size_t j = 0; j < cb.ncb; ++j) {
    before[j] = n;
    const std::int32_t s = cb.src[j];
    if (s < 0) continue;
    if (n > 0xFFFEU) throw std::invalid_argument("circular buffer too long for soft de-rate matching");
    g.rank[(s >> 16) * d + (s & 0xFFFF)] = static_cast<std::uint16_t>(n++);
  }
  before[cb.ncb] = n;
  if (n == 0) throw std::invalid_argument("circular buffer holds no bits");
  g.n_valid = n;
  for (std::uint32_t rv = 0; rv < 4; ++rv) g.start[rv] = before[start_position(cb, rv) % cb.ncb] % n;
  return g;
}

void tail_beta_log(const float tail_sys[3], const float tail_par[3], float beta[8]) {
  for (unsigned s = 0; s < 8; ++s) {
    unsigned cur = s;
    float metric = 0.0F;
    for (int t = 0; t < 3; ++t) {
      const unsigned u = termination_input(cur);
      const auto tr = rsc(cur, u);
      metric += 0.5F * ((u ? 1.0F : -1.0F) * tail_sys[t]) + 0.5F * ((tr.parity ? 1.0F : -1.0F) * tail_par[t]);
      cur = tr.next;
    }
    beta[s] = (cur == 0) ? metric : kNegInf;
  }
  renormalise(beta);
}

void decoder_tail_llrs(const float* d0, const float* d1, const float* d2, std::size_t k, float o[12]) {
  // lte_turbo.cpp:758-773 (third entries of each group are unused there: 0)
  o[0] = d0[k + 0];
  o[1] = d0[k + 2];
  o[2] = 0.0F;
  o[3] = d1[k + 0];
  o[4] = d0[k + 1];
  o[5] = d1[k + 2];
  o[6] = d1[k + 3];
  o[7] = d2[k + 0];
  o[8] = d2[k + 2];
  o[9] = d2[k + 1];
  o[10] = d0[k + 3];
  o[11] = d2[k + 3];
}

namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

void throw_status(int status, const char* what) {
  switch (status) {
    case TC_ERR_INVALID_ARGUMENT: throw std::invalid_argument(what);
    case TC_ERR_OUT_OF_RANGE: throw std::out_of_range(what);
    default: throw std::runtime_error(what);
  }
}

std::vector<float> decode_cpu(const CodeBlockDescriptor& desc, std::span<const float> d0,
                              std::span<const float> d1, std::span<const float> d2,
                              std::uint32_t max_iterations, bool early_stop, bool* crc_ok,
                              std::uint32_t* iterations_used) {
  // lte_turbo.cpp:736-813
  if (d0.size() != desc.d_r || d1.size() != desc.d_r || d2.size() != desc.d_r)
    throw std::invalid_argument("invalid LLR sizes for turbo decode");
  const std::size_t k = desc.k_r;
  const auto perm = qpp_permutation(k);
  std::vector<float> buf(5 * k, 0.0F);
  std::span<float> ap(buf.data(), k), e1(buf.data() + k, k), si(buf.data() + 2 * k, k),
      ia(buf.data() + 3 * k, k), e2i(buf.data() + 4 * k, k);
  for (std::size_t i = 0; i < k; ++i) si[i] = d0[perm[i]];
  float tails[12];
  decoder_tail_llrs(d0.data(), d1.data(), d2.data(), k, tails);
  const std::span<const float> sys = d0.first(k);
  std::vector<float> post(k, 0.0F);
  std::vector<std::uint8_t> bits(k);
  SisoWork work;
  bool ok = false;
  for (std::uint32_t it = 0; it < max_iterations; ++it) {
    siso(sys, d1, ap, tails + 0, tails + 3, e1, work);
    for (std::size_t i = 0; i < k; ++i) ia[i] = e1[perm[i]];
    siso(si, d2, ia, tails + 6, tails + 9, e2i, work);
    for (std::size_t j = 0; j < k; ++j) ap[perm[j]] = e2i[j];  // a-priori of the next pass = ext2
    for (std::size_t i = 0; i < k; ++i) post[i] = sys[i] + e1[i] + ap[i];
    if (iterations_used) *iterations_used = it + 1;
    if (early_stop) {
      for (std::size_t i = 0; i < k; ++i) bits[i] = (i < desc.filler_count) ? 0 : (post[i] >= 0.0F);
      ok = crc_check(1, std::span<const std::uint8_t>(bits).subspan(desc.filler_count));
      if (ok) break;
    }
  }
  if (crc_ok) *crc_ok = ok;
  return post;
}

}  // namespace detail

// ---------------------------------------------------------- public API ----
std::span<const QppEntry> qpp_interleaver_table() { return {detail::kQppTable, 188}; }

const QppEntry& qpp_entry_for_k(std::size_t k) {
  const int i = detail::qpp_index(k);
  if (i < 0) throw std::out_of_range("unsupported QPP block size");
  return detail::kQppTable[i];
}

std::vector<std::uint8_t> append_crc24a(std::span<const std::uint8_t> b) { return detail::crc_append(0, b); }
std::vector<std::uint8_t> append_crc24b(std::span<const std::uint8_t> b) { return detail::crc_append(1, b); }
std::vector<std::uint8_t> append_crc16(std::span<const std::uint8_t> b) { return detail::crc_append(2, b); }
std::vector<std::uint8_t> append_crc8(std::span<const std::uint8_t> b) { return detail::crc_append(3, b); }
bool check_crc24a(std::span<const std::uint8_t> b) { return detail::crc_check(0, b); }
bool check_crc24b(std::span<const std::uint8_t> b) { return detail::crc_check(1, b); }
bool check_crc16(std::span<const std::uint8_t> b) { return detail::crc_check(2, b); }
bool check_crc8(std::span<const std::uint8_t> b) { return detail::crc_check(3, b); }

namespace detail {
SegmentationResult segmentation_sizes(std::size_t b) {
  // 36.212 5.1.2 as at lte_turbo.cpp:540-604 (sizes only)
  if (b == 0) throw std::invalid_argument("transport block cannot be empty");
  constexpr std::size_t Z = 6144;
  SegmentationResult r;
  std::size_t bp = b;
  if (b > Z) {
    r.l = 24;
    r.c = detail::ceil_div(b, Z - 24);
    bp = b + r.c * 24;
  } else {
    r.c = 1;
  }
  const auto table = qpp_interleaver_table();
  const auto up = std::find_if(table.begin(), table.end(),
                               [&](const QppEntry& e) { return r.c * e.k >= bp; });
  if (up == table.end()) throw std::runtime_error("failed to select K+");
  r.k_plus = up->k;
  if (r.c == 1) {
    r.c_plus = 1;
  } else {
    if (up == table.begin()) throw std::runtime_error("failed to select K-");
    r.k_minus = std::prev(up)->k;
    r.c_minus = (r.c * r.k_plus - bp) / (r.k_plus - r.k_minus);
    r.c_plus = r.c - r.c_minus;
  }
  r.filler_bits = r.c_plus * r.k_plus + r.c_minus * r.k_minus - bp;
  return r;
}
}  // namespace detail

SegmentationResult segment_transport_block(std::span<const std::uint8_t> bits) {
  // 36.212 5.1.2 as at lte_turbo.cpp:540-604
  SegmentationResult r = detail::segmentation_sizes(bits.size());
  std::size_t cursor = 0;
  for (std::size_t i = 0; i < r.c; ++i) {
    const std::size_t kr = (i < r.c_minus) ? r.k_minus : r.k_plus;
    const std::size_t off = (i == 0) ? r.filler_bits : 0;
    const std::size_t payload = kr - off - r.l;
    std::vector<std::uint8_t> blk(kr, 0);
    std::copy_n(bits.begin() + static_cast<std::ptrdiff_t>(cursor), payload, blk.begin() + static_cast<std::ptrdiff_t>(off));
    cursor += payload;
    if (r.l == 24) {
      const std::uint32_t crc = detail::crc_register(1, std::span<const std::uint8_t>(blk).subspan(off, payload));
      for (unsigned j = 0; j < 24; ++j) blk[off + payload + j] = static_cast<std::uint8_t>((crc >> (23 - j)) & 1u);
    }
    r.code_blocks.push_back(std::move(blk));
  }
  return r;
}

std::vector<std::size_t> qpp_permutation(std::size_t k) {
  // Pi(i) = (f1 i + f2 i^2) mod K (lte_turbo.cpp:606-613), by the recurrence
  // Pi(i+1) = Pi(i) + (f1 + f2 (2i+1)) mod K.
  const auto& e = qpp_entry_for_k(k);
  std::vector<std::size_t> p(k);
  std::size_t v = 0, step = (e.f1 + e.f2) % k;
  const std::size_t inc = (2 * e.f2) % k;
  for (std::size_t i = 0; i < k; ++i) {
    p[i] = v;
    v = (v + step) % k;
    step = (step + inc) % k;
  }
  return p;
}

std::vector<std::size_t> qpp_inverse_permutation(std::size_t k) {
  const auto p = qpp_permutation(k);
  std::vector<std::size_t> inv(k);
  for (std::size_t i = 0; i < k; ++i) inv[p[i]] = i;
  return inv;
}

EncodedCodeBlock turbo_encode_code_block(const CodeBlockDescriptor& desc, std::span<const std::uint8_t> cb) {
  // lte_turbo.cpp:624-688
  if (cb.size() != desc.k_r) throw std::invalid_argument("code block length mismatch");
  const std::size_t k = desc.k_r;
  const auto perm = qpp_permutation(k);
  EncodedCodeBlock out;
  out.descriptor = desc;
  out.cb_bits.assign(cb.begin(), cb.end());
  out.systematic.resize(k + 4);
  out.parity0.resize(k + 4);
  out.parity1.resize(k + 4);
  unsigned su = 0, sl = 0;
  for (std::size_t i = 0; i < k; ++i) {
    const auto tu = detail::rsc(su, cb[i]);
    const auto tl = detail::rsc(sl, cb[perm[i]]);
    const bool fill = i < desc.filler_count;
    out.systematic[i] = fill ? 0 : cb[i];
    out.parity0[i] = fill ? 0 : tu.parity;
    out.parity1[i] = tl.parity;
    su = tu.next;
    sl = tl.next;
  }
  std::uint8_t xu[3], zu[3], xl[3], zl[3];
  for (int i = 0; i < 3; ++i) {
    xu[i] = static_cast<std::uint8_t>(detail::termination_input(su));
    xl[i] = static_cast<std::uint8_t>(detail::termination_input(sl));
    const auto tu = detail::rsc(su, xu[i]);
    const auto tl = detail::rsc(sl, xl[i]);
    zu[i] = tu.parity;
    zl[i] = tl.parity;
    su = tu.next;
    sl = tl.next;
  }
  // tail multiplexing of lte_turbo.cpp:664-677 (not the 36.212 order)
  const std::uint8_t t0[4] = {xu[0], zu[1], xu[1], zu[2]};
  const std::uint8_t t1[4] = {zu[0], xu[1], zu[2], xl[0]};
  const std::uint8_t t2[4] = {zl[0], xl[1], zl[1], xl[2]};
  for (int i = 0; i < 4; ++i) {
    out.systematic[k + i] = t0[i];
    out.parity0[k + i] = t1[i];
    out.parity1[k + i] = t2[i];
  }
  for (int i = 0; i < 3; ++i) {
    out.upper_tail_systematic[i] = xu[i];
    out.upper_tail_parity[i] = zu[i];
    out.lower_tail_systematic[i] = xl[i];
    out.lower_tail_parity[i] = zl[i];
  }
  out.descriptor.d_r = k + 4;
  return out;
}

EncodedTransportBlock encode_transport_block(const LteTurboCodecParams& params,
                                             std::span<const std::uint8_t> transport_bits) {
  // lte_turbo.cpp:690-720
  if (params.transport_block_bits != transport_bits.size())
    throw std::invalid_argument("transport block size does not match params");
  EncodedTransportBlock out;
  out.params = params;
  out.transport_bits.assign(transport_bits.begin(), transport_bits.end());
  out.tb_bits_with_crc = append_crc24a(transport_bits);
  const auto seg = segment_transport_block(out.tb_bits_with_crc);
  for (std::size_t r = 0; r < seg.code_blocks.size(); ++r) {
    const auto& bits = seg.code_blocks[r];
    CodeBlockDescriptor d;
    d.block_index = r;
    d.k_r = bits.size();
    d.d_r = bits.size() + 4;
    d.e_r = detail::compute_e(r, params.g_total, seg.c, params.num_layers, params.mod_order);
    d.filler_count = (r == 0) ? seg.filler_bits : 0;
    d.qpp_index = static_cast<std::size_t>(detail::qpp_index(d.k_r));
    auto enc = turbo_encode_code_block(d, bits);
    enc.rate_matched_bits = detail::rate_match_bits(enc.systematic, enc.parity0, enc.parity1,
                                                    d.filler_count, d.e_r, params.rv);
    out.code_blocks.push_back(std::move(enc));
  }
  return out;
}

std::vector<float> derate_match_code_block(const LteTurboCodecParams& params, const CodeBlockDescriptor& desc,
                                           std::span<const std::uint8_t> rm) {
  return detail::derate_match(rm, desc.d_r, desc.filler_count, params.rv);
}

std::vector<float> turbo_decode_code_block_cpu(const CodeBlockDescriptor& desc, std::span<const float> d0,
                                               std::span<const float> d1, std::span<const float> d2,
                                               std::uint32_t max_iterations, bool early_stop, bool* crc_ok,
                                               std::uint32_t* iterations_used) {
  return detail::decode_cpu(desc, d0, d1, d2, max_iterations, early_stop, crc_ok, iterations_used);
}

std::string to_string(LteChannelType t) {
  switch (t) {
    case LteChannelType::UL_SCH: return "UL_SCH";
    case LteChannelType::DL_SCH: return "DL_SCH";
    case LteChannelType::PCH: return "PCH";
    case LteChannelType::MCH: return "MCH";
  }
  throw std::invalid_argument("unknown LTE channel type");
}

}  // namespace turbo_cpp

// ------------------------------------------------------ C ABI: host codec --
namespace {
template <typename F>
tc_status guarded(F&& f) {
  try {
    f();
    return TC_OK;
  } catch (const std::invalid_argument& e) {
    turbo_cpp::detail::set_last_error(e.what());
    return TC_ERR_INVALID_ARGUMENT;
  } catch (const std::out_of_range& e) {
    turbo_cpp::detail::set_last_error(e.what());
    return TC_ERR_OUT_OF_RANGE;
  } catch (const std::exception& e) {
    turbo_cpp::detail::set_last_error(e.what());
    return TC_ERR_RUNTIME;
  }
}
}  // namespace

extern "C" {

const char* tc_last_error(void) { return turbo_cpp::detail::last_error(); }

tc_status tc_crc_append(int32_t kind, const uint8_t* bits, size_t n, uint8_t* out) {
  return guarded([&] {
    if (kind < 0 || kind > 3) throw std::invalid_argument("unknown CRC kind");
    const auto v = turbo_cpp::detail::crc_append(kind, std::span<const uint8_t>(bits, n));
    std::memcpy(out, v.data(), v.size());
  });
}

tc_status tc_crc_check(int32_t kind, const uint8_t* bits, size_t n, int32_t* ok) {
  return guarded([&] {
    if (kind < 0 || kind > 3) throw std::invalid_argument("unknown CRC kind");
    *ok = turbo_cpp::detail::crc_check(kind, std::span<const uint8_t>(bits, n));
  });
}

tc_status tc_segment(const uint8_t* bits, size_t b, uint64_t* seg, uint8_t* blocks) {
  return guarded([&] {
    const auto r = turbo_cpp::segment_transport_block(std::span<const uint8_t>(bits, b));
    const uint64_t v[7] = {r.c, r.c_plus, r.c_minus, r.k_plus, r.k_minus, r.filler_bits, r.l};
    std::memcpy(seg, v, sizeof v);
    if (blocks) {
      for (const auto& blk : r.code_blocks) {
        std::memcpy(blocks, blk.data(), blk.size());
        blocks += blk.size();
      }
    }
  });
}

tc_status tc_qpp_entry(uint32_t k, uint32_t* f1, uint32_t* f2, uint32_t* index) {
  return guarded([&] {
    const auto& e = turbo_cpp::qpp_entry_for_k(k);
    if (f1) *f1 = static_cast<uint32_t>(e.f1);
    if (f2) *f2 = static_cast<uint32_t>(e.f2);
    if (index) *index = static_cast<uint32_t>(turbo_cpp::detail::qpp_index(k));
  });
}

tc_status tc_encode_code_block(const uint8_t* cb, uint32_t k, uint32_t filler, uint8_t* d0, uint8_t* d1,
                               uint8_t* d2) {
  return guarded([&] {
    turbo_cpp::CodeBlockDescriptor d;
    d.k_r = k;
    d.d_r = k + 4;
    d.filler_count = filler;
    (void)turbo_cpp::qpp_entry_for_k(k);
    const auto e = turbo_cpp::turbo_encode_code_block(d, std::span<const uint8_t>(cb, k));
    std::memcpy(d0, e.systematic.data(), k + 4);
    std::memcpy(d1, e.parity0.data(), k + 4);
    std::memcpy(d2, e.parity1.data(), k + 4);
  });
}

tc_status tc_compute_e(uint64_t r, uint64_t g_total, uint64_t c, uint32_t nl, uint32_t qm, uint64_t* e) {
  return guarded([&] { *e = turbo_cpp::detail::compute_e(r, g_total, c, nl, qm); });
}

tc_status tc_rate_match(const uint8_t* d0, const uint8_t* d1, const uint8_t* d2, uint64_t d, uint64_t filler,
                        uint64_t e, uint32_t rv, uint8_t* out) {
  return guarded([&] {
    const auto v = turbo_cpp::detail::rate_match_bits(std::span<const uint8_t>(d0, d), std::span<const uint8_t>(d1, d),
                                                      std::span<const uint8_t>(d2, d), filler, e, rv);
    std::memcpy(out, v.data(), v.size());
  });
}

tc_status tc_derate_match(const uint8_t* bits, uint64_t e, uint64_t d, uint64_t filler, uint32_t rv, float* out) {
  return guarded([&] {
    const auto v = turbo_cpp::detail::derate_match(std::span<const uint8_t>(bits, e), d, filler, rv);
    std::memcpy(out, v.data(), v.size() * sizeof(float));
  });
}

tc_status tc_decode_code_block_cpu(const tc_cb_desc* desc, const float* d0, const float* d1, const float* d2,
                                   uint32_t max_iterations, int32_t early_stop, float* posterior, int32_t* crc_ok,
                                   uint32_t* iterations_used) {
  return guarded([&] {
    turbo_cpp::CodeBlockDescriptor d;
    d.block_index = desc->block_index;
    d.k_r = desc->k_r;
    d.d_r = desc->d_r;
    d.e_r = desc->e_r;
    d.filler_count = desc->filler_count;
    d.qpp_index = desc->qpp_index;
    (void)turbo_cpp::qpp_entry_for_k(d.k_r);
    bool ok = false;
    uint32_t it = 0;
    const auto p = turbo_cpp::detail::decode_cpu(d, std::span<const float>(d0, d.d_r), std::span<const float>(d1, d.d_r),
                                                 std::span<const float>(d2, d.d_r), max_iterations, early_stop != 0, &ok,
                                                 &it);
    std::memcpy(posterior, p.data(), p.size() * sizeof(float));
    if (crc_ok) *crc_ok = ok;
    if (iterations_used) *iterations_used = it;
  });
}

}  // extern "C"
