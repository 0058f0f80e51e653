// ref_shim.cpp -- extern "C" access to the compiled reference CPU codec.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together
// with the reference's own src/lte_turbo.cpp (read in place under
// /root/reference, CPU-only: TURBO_CPP_HAS_CUDA=0) into oracle/_ref/.  It is
// used to validate the C restatement (turbo_oracle.c), to write the golden
// fixtures under tests/golden/ and as the timed CPU baseline of bench.py.
// Nothing in the product loads it.
#include "turbo_cpp/lte_turbo.h"

#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <thread>
#include <vector>

namespace tc = turbo_cpp;

namespace {
// 0 ok, 1 invalid_argument, 2 out_of_range, 3 runtime_error / other
template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::invalid_argument&) {
    return 1;
  } catch (const std::out_of_range&) {
    return 2;
  } catch (const std::exception&) {
    return 3;
  }
}

tc::CodeBlockDescriptor make_desc(size_t k, size_t filler, size_t e) {
  tc::CodeBlockDescriptor d;
  d.k_r = k;
  d.d_r = k + 4;
  d.e_r = e;
  d.filler_count = filler;
  return d;
}
}  // namespace

extern "C" {

size_t ref_qpp_table(uint32_t* out /* 3*188 */) {
  auto t = tc::qpp_interleaver_table();
  for (size_t i = 0; i < t.size(); ++i) {
    out[3 * i] = static_cast<uint32_t>(t[i].k);
    out[3 * i + 1] = static_cast<uint32_t>(t[i].f1);
    out[3 * i + 2] = static_cast<uint32_t>(t[i].f2);
  }
  return t.size();
}

int ref_qpp_perm(uint32_t k, uint32_t* perm) {
  return guarded([&] {
    auto p = tc::qpp_permutation(k);
    for (size_t i = 0; i < p.size(); ++i) perm[i] = static_cast<uint32_t>(p[i]);
  });
}

int ref_crc_append(int kind, const uint8_t* bits, size_t n, uint8_t* out) {
  std::span<const uint8_t> s(bits, n);
  std::vector<uint8_t> r;
  switch (kind) {
    case 0: r = tc::append_crc24a(s); break;
    case 1: r = tc::append_crc24b(s); break;
    case 2: r = tc::append_crc16(s); break;
    default: r = tc::append_crc8(s); break;
  }
  std::memcpy(out, r.data(), r.size());
  return static_cast<int>(r.size() - n);
}

int ref_crc_check(int kind, const uint8_t* bits, size_t n) {
  std::span<const uint8_t> s(bits, n);
  switch (kind) {
    case 0: return tc::check_crc24a(s);
    case 1: return tc::check_crc24b(s);
    case 2: return tc::check_crc16(s);
    default: return tc::check_crc8(s);
  }
}

int ref_segment(const uint8_t* bits, size_t b, size_t* seg, uint8_t* blocks) {
  return guarded([&] {
    auto r = tc::segment_transport_block(std::span<const uint8_t>(bits, b));
    seg[0] = r.c;
    seg[1] = r.c_plus;
    seg[2] = r.c_minus;
    seg[3] = r.k_plus;
    seg[4] = r.k_minus;
    seg[5] = r.filler_bits;
    seg[6] = r.l;
    if (blocks) {
      size_t off = 0;
      for (auto& blk : r.code_blocks) {
        std::memcpy(blocks + off, blk.data(), blk.size());
        off += blk.size();
      }
    }
  });
}

int ref_encode_cb(const uint8_t* cb, size_t k, size_t filler, uint8_t* d0, uint8_t* d1,
                  uint8_t* d2) {
  return guarded([&] {
    auto e = tc::turbo_encode_code_block(make_desc(k, filler, 0),
                                         std::span<const uint8_t>(cb, k));
    std::memcpy(d0, e.systematic.data(), k + 4);
    std::memcpy(d1, e.parity0.data(), k + 4);
    std::memcpy(d2, e.parity1.data(), k + 4);
  });
}

// Encode a transport block.  desc_out[6*r..] = descriptor of CB r;
// rm_out receives the concatenated rate-matched bits.  Returns the status;
// *n_cb is the number of code blocks.
int ref_encode_tb(size_t tbs, size_t g_total, uint32_t rv, uint32_t nl, uint32_t qm,
                  const uint8_t* bits, size_t* n_cb, size_t* desc_out, uint8_t* rm_out,
                  uint8_t* cbbits_out) {
  return guarded([&] {
    tc::LteTurboCodecParams p;
    p.transport_block_bits = tbs;
    p.g_total = g_total;
    p.rv = rv;
    p.num_layers = nl;
    p.mod_order = qm;
    auto enc = tc::encode_transport_block(p, std::span<const uint8_t>(bits, tbs));
    *n_cb = enc.code_blocks.size();
    size_t off = 0, coff = 0;
    for (size_t r = 0; r < enc.code_blocks.size(); ++r) {
      const auto& cb = enc.code_blocks[r];
      const auto& d = cb.descriptor;
      size_t* o = desc_out + 6 * r;
      o[0] = d.block_index;
      o[1] = d.k_r;
      o[2] = d.d_r;
      o[3] = d.e_r;
      o[4] = d.filler_count;
      o[5] = d.qpp_index;
      if (rm_out) std::memcpy(rm_out + off, cb.rate_matched_bits.data(), cb.rate_matched_bits.size());
      off += cb.rate_matched_bits.size();
      if (cbbits_out) std::memcpy(cbbits_out + coff, cb.cb_bits.data(), cb.cb_bits.size());
      coff += cb.cb_bits.size();
    }
  });
}

int ref_derate(size_t k, size_t filler, size_t e, uint32_t rv, const uint8_t* bits,
               float* out /* 3*(k+4) */) {
  return guarded([&] {
    tc::LteTurboCodecParams p;
    p.rv = rv;
    auto v = tc::derate_match_code_block(p, make_desc(k, filler, e),
                                         std::span<const uint8_t>(bits, e));
    std::memcpy(out, v.data(), v.size() * sizeof(float));
  });
}

int ref_decode_cb(size_t k, size_t filler, const float* d0, const float* d1, const float* d2,
                  uint32_t max_it, int early, float* post, int* crc_ok, uint32_t* iters) {
  return guarded([&] {
    auto d = make_desc(k, filler, 0);
    bool ok = false;
    uint32_t it = 0;
    auto r = tc::turbo_decode_code_block_cpu(d, std::span<const float>(d0, k + 4),
                                             std::span<const float>(d1, k + 4),
                                             std::span<const float>(d2, k + 4), max_it,
                                             early != 0, &ok, &it);
    std::memcpy(post, r.data(), r.size() * sizeof(float));
    if (crc_ok) *crc_ok = ok;
    if (iters) *iters = it;
  });
}

// Decode n code blocks of length k (llr: n x 3(k+4), d0|d1|d2 per block) on
// `threads` std::threads, one block per call of the reference decoder.
int ref_decode_many(size_t n, size_t k, const float* llr, uint32_t max_it, float* post,
                    int threads) {
  if (threads < 1) threads = 1;
  std::vector<std::thread> pool;
  std::vector<int> status(threads, 0);
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      for (size_t i = t; i < n; i += threads) {
        const float* b = llr + i * 3 * (k + 4);
        int s = ref_decode_cb(k, 0, b, b + (k + 4), b + 2 * (k + 4), max_it, 0, post + i * k,
                              nullptr, nullptr);
        if (s) status[t] = s;
      }
    });
  }
  for (auto& th : pool) th.join();
  for (int s : status)
    if (s) return s;
  return 0;
}

// Full transport-block round trip on the CPU backend: encode with the given
// params, decode, report TB CRC and the decoded bits.
int ref_tb_roundtrip(size_t tbs, size_t g_total, uint32_t rv, uint32_t max_it, int early,
                     const uint8_t* bits, uint8_t* out_bits, int* tb_crc_ok, uint32_t* iters,
                     float* post_out, size_t* post_len) {
  return guarded([&] {
    tc::LteTurboCodecParams p;
    p.transport_block_bits = tbs;
    p.g_total = g_total;
    p.rv = rv;
    p.max_iterations = max_it;
    p.enable_early_stop = early != 0;
    auto enc = tc::encode_transport_block(p, std::span<const uint8_t>(bits, tbs));
    auto dec = tc::decode_transport_block(p, enc, tc::DecoderBackend::CpuReference);
    std::memcpy(out_bits, dec.tb_bits.data(), dec.tb_bits.size());
    *tb_crc_ok = dec.transport_block_crc_ok;
    *iters = dec.iterations_used;
    if (post_out) std::memcpy(post_out, dec.posterior_llr->data(), dec.posterior_llr->size() * sizeof(float));
    if (post_len) *post_len = dec.posterior_llr->size();
  });
}

// Decode a transport block with the CPU backend from given rate-matched bits
// (bits of all code blocks back to back, the encoder's lengths): encodes
// `payload` to obtain the descriptors, replaces the rate-matched bits with
// `rm`, then decodes.  Outputs payload bits, per-block CRC flags, TB CRC flag,
// iterations used and the posterior LLRs.
int ref_decode_tb(size_t tbs, size_t g_total, uint32_t rv, uint32_t max_it, int early,
                  const uint8_t* payload, const uint8_t* rm, uint8_t* out_bits, int* cb_ok, int* tb_ok,
                  uint32_t* iters, float* post_out) {
  return guarded([&] {
    tc::LteTurboCodecParams p;
    p.transport_block_bits = tbs;
    p.g_total = g_total;
    p.rv = rv;
    p.max_iterations = max_it;
    p.enable_early_stop = early != 0;
    auto enc = tc::encode_transport_block(p, std::span<const uint8_t>(payload, tbs));
    size_t off = 0;
    for (auto& cb : enc.code_blocks) {
      std::memcpy(cb.rate_matched_bits.data(), rm + off, cb.rate_matched_bits.size());
      off += cb.rate_matched_bits.size();
    }
    auto dec = tc::decode_transport_block(p, enc, tc::DecoderBackend::CpuReference);
    std::memcpy(out_bits, dec.tb_bits.data(), dec.tb_bits.size());
    for (size_t r = 0; r < dec.code_block_crc_ok.size(); ++r) cb_ok[r] = dec.code_block_crc_ok[r];
    *tb_ok = dec.transport_block_crc_ok;
    *iters = dec.iterations_used;
    std::memcpy(post_out, dec.posterior_llr->data(), dec.posterior_llr->size() * sizeof(float));
  });
}

}  // extern "C"
