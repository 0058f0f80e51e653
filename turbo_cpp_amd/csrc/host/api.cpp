// C++ API entry points that reach the GPU: the per-code-block decoders
// (lte_turbo.h:129-155) and decode_transport_block (lte_turbo.h:157-159).
// They call the C ABI of the HIP library and turn its status codes back
// into the exception types the reference throws.
#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "codec_internal.h"
#include "turbo_cpp/c_api.h"

namespace turbo_cpp {
namespace detail {
std::vector<float> decode_cpu(const CodeBlockDescriptor& desc, std::span<const float> d0, std::span<const float> d1,
                              std::span<const float> d2, std::uint32_t max_iterations, bool early_stop, bool* crc_ok,
                              std::uint32_t* iterations_used);
}  // namespace detail

namespace {

void check(tc_status s) {
  if (s != TC_OK) detail::throw_status(s, tc_last_error());
}

tc_cb_desc to_c(const CodeBlockDescriptor& d) {
  return {d.block_index, d.k_r, d.d_r, d.e_r, d.filler_count, d.qpp_index};
}

std::vector<float> gpu_decode(int decoder, const CodeBlockDescriptor& desc, std::span<const float> d0,
                              std::span<const float> d1, std::span<const float> d2, std::uint32_t max_iterations,
                              bool early_stop, bool* crc_ok, std::uint32_t* iterations_used) {
  tc_options o{};
  o.max_iterations = max_iterations;
  o.early_stop = early_stop ? 1 : 0;
  o.decoder = decoder;
  const tc_cb_desc cd = to_c(desc);
  std::vector<float> post(desc.k_r, 0.0F);
  int32_t ok = 0;
  uint32_t it = 0;
  check(tc_decode_code_block(&cd, d0.data(), d0.size(), d1.data(), d1.size(), d2.data(), d2.size(), &o, post.data(),
                             &ok, &it));
  if (crc_ok) *crc_ok = ok != 0;
  if (iterations_used && max_iterations > 0) *iterations_used = it;
  return post;
}

}  // namespace

// GpuStage1Exact: exact full-length Log-MAP on the GPU (lte_turbo.cpp:815-838).
std::vector<float> turbo_decode_code_block_gpu_stage1(const CodeBlockDescriptor& desc, std::span<const float> d0,
                                                      std::span<const float> d1, std::span<const float> d2,
                                                      std::uint32_t max_iterations, bool early_stop, bool* crc_ok,
                                                      std::uint32_t* iterations_used) {
  return gpu_decode(TC_DECODER_EXACT, desc, d0, d1, d2, max_iterations, early_stop, crc_ok, iterations_used);
}

// GpuStage2Windowed: segment-parallel decoder (lte_turbo.cpp:840-863).
std::vector<float> turbo_decode_code_block_gpu_stage2(const CodeBlockDescriptor& desc, std::span<const float> d0,
                                                      std::span<const float> d1, std::span<const float> d2,
                                                      std::uint32_t max_iterations, bool early_stop, bool* crc_ok,
                                                      std::uint32_t* iterations_used) {
  return gpu_decode(TC_DECODER_WINDOW_LOG, desc, d0, d1, d2, max_iterations, early_stop, crc_ok, iterations_used);
}

// The reference's tile8 variant computes the exact decoder's maths with one
// lane per trellis state (lte_turbo_cuda.cu:298-401); here it is the serial
// exact decoder, which maps lanes to states the same way (same numbers as
// Stage1, bit for bit).
std::vector<float> turbo_decode_code_block_gpu_stage1_tile8_experimental(
    const CodeBlockDescriptor& desc, std::span<const float> d0, std::span<const float> d1, std::span<const float> d2,
    std::uint32_t max_iterations, bool early_stop, bool* crc_ok, std::uint32_t* iterations_used) {
  return gpu_decode(TC_DECODER_EXACT_SERIAL, desc, d0, d1, d2, max_iterations, early_stop, crc_ok, iterations_used);
}

namespace {
std::vector<std::uint8_t> hard_bits(const std::vector<float>& post) {
  std::vector<std::uint8_t> b(post.size());
  for (std::size_t i = 0; i < post.size(); ++i) b[i] = post[i] >= 0.0F ? 1U : 0U;
  return b;
}
}  // namespace

namespace {
// True when the code blocks are those encode_transport_block makes for
// `params` (lte_turbo.cpp:690-720), so the device pipeline, which derives the
// geometry from the parameters, sees the same blocks.
bool standard_geometry(const LteTurboCodecParams& params, const EncodedTransportBlock& encoded) {
  try {
    if (params.transport_block_bits == 0) return false;
    // the device de-rate tables cover rv 0..3; the host path takes k0 of any rv
    // (lte_turbo.cpp:263-264)
    if (params.rv > 3) return false;
    const auto seg = detail::segmentation_sizes(params.transport_block_bits + 24);
    if (seg.c != encoded.code_blocks.size()) return false;
    for (std::size_t r = 0; r < seg.c; ++r) {
      const auto& d = encoded.code_blocks[r].descriptor;
      const std::size_t k = r < seg.c_minus ? seg.k_minus : seg.k_plus;
      const std::size_t f = r == 0 ? seg.filler_bits : 0;
      const std::size_t e = detail::compute_e(r, params.g_total, seg.c, params.num_layers, params.mod_order);
      if (d.k_r != k || d.d_r != k + 4 || d.filler_count != f || d.e_r != e ||
          encoded.code_blocks[r].rate_matched_bits.size() != e)
        return false;
    }
    return true;
  } catch (const std::exception&) {
    return false;
  }
}

// The whole receive chain on the device: +-8 per received bit (the
// reference's de-rate input, lte_turbo.cpp:280-325), soft de-rate matching,
// batched decode and TB assembly.
DecodedTransportBlock decode_on_device(const LteTurboCodecParams& params, const EncodedTransportBlock& encoded,
                                       int decoder) {
  std::vector<float> rx;
  std::size_t nk = 0;
  for (const auto& blk : encoded.code_blocks) {
    for (std::uint8_t b : blk.rate_matched_bits) rx.push_back(b ? 8.0F : -8.0F);
    nk += blk.descriptor.k_r;
  }
  const std::size_t c = encoded.code_blocks.size();
  tc_tb_desc t{params.transport_block_bits, params.g_total, params.rv, params.num_layers, params.mod_order};
  tc_options o{};
  o.max_iterations = params.max_iterations;
  o.early_stop = params.enable_early_stop ? 1 : 0;
  o.decoder = decoder;
  DecodedTransportBlock out;
  out.tb_bits.resize(params.transport_block_bits);
  std::vector<int32_t> cb_ok(c, 0);
  std::vector<float> post(nk);
  int32_t tb_ok = 0;
  uint32_t iters = 0;
  check(tc_decode_transport_blocks_host(&t, 1, &o, rx.data(), out.tb_bits.data(), &tb_ok, cb_ok.data(), &iters,
                                        post.data()));
  for (int32_t v : cb_ok) out.code_block_crc_ok.push_back(v != 0);
  out.transport_block_crc_ok = tb_ok != 0;
  out.iterations_used = iters;
  out.posterior_llr = std::move(post);
  return out;
}
}  // namespace

// lte_turbo.cpp:891-980.  CpuReference decodes block by block.  Both GPU
// backends decode all code blocks of the transport block in one batched
// launch (the reference used one std::async thread per block instead); for
// transport blocks laid out as encode_transport_block makes them, de-rate
// matching and TB assembly run on the device too.
DecodedTransportBlock decode_transport_block(const LteTurboCodecParams& params, const EncodedTransportBlock& encoded,
                                             DecoderBackend backend) {
  if (backend != DecoderBackend::CpuReference && standard_geometry(params, encoded))
    return decode_on_device(params, encoded,
                            backend == DecoderBackend::GpuStage1Exact ? TC_DECODER_EXACT : TC_DECODER_WINDOW_LOG);
  DecodedTransportBlock out;
  const std::size_t c = encoded.code_blocks.size();
  out.posterior_llr = std::vector<float>{};
  std::vector<std::vector<float>> posts(c);
  std::vector<std::uint32_t> iters(c, 0);
  std::vector<bool> ok(c, false);
  std::vector<std::vector<float>> llrs(c);
  for (std::size_t i = 0; i < c; ++i) {
    const auto& blk = encoded.code_blocks[i];
    llrs[i] = derate_match_code_block(params, blk.descriptor, blk.rate_matched_bits);
  }
  if (backend == DecoderBackend::CpuReference) {
    for (std::size_t i = 0; i < c; ++i) {
      const auto& d = encoded.code_blocks[i].descriptor;
      const std::size_t n = d.d_r;
      bool cok = false;
      std::uint32_t it = 0;
      posts[i] = detail::decode_cpu(d, std::span<const float>(llrs[i].data(), n),
                                    std::span<const float>(llrs[i].data() + n, n),
                                    std::span<const float>(llrs[i].data() + 2 * n, n), params.max_iterations,
                                    params.enable_early_stop, &cok, &it);
      ok[i] = cok;
      iters[i] = it;
    }
  } else {
    for (std::size_t i = 0; i < c; ++i) {
      const auto& d = encoded.code_blocks[i].descriptor;
      if (d.d_r != d.k_r + 4 || llrs[i].size() != 3 * d.d_r)
        throw std::invalid_argument("invalid LLR sizes for turbo decode");
    }
    tc_options o{};
    o.max_iterations = params.max_iterations;
    o.early_stop = params.enable_early_stop ? 1 : 0;
    o.decoder = (backend == DecoderBackend::GpuStage1Exact) ? TC_DECODER_EXACT : TC_DECODER_WINDOW_LOG;
    std::vector<std::uint32_t> ks(c), fs(c);
    std::vector<float> packed;
    std::size_t post_total = 0;
    for (std::size_t i = 0; i < c; ++i) {
      ks[i] = static_cast<std::uint32_t>(encoded.code_blocks[i].descriptor.k_r);
      fs[i] = static_cast<std::uint32_t>(encoded.code_blocks[i].descriptor.filler_count);
      packed.insert(packed.end(), llrs[i].begin(), llrs[i].end());
      post_total += ks[i];
    }
    std::vector<float> post(post_total);
    std::vector<int32_t> cok(c, 0);
    check(tc_decode_batch_host(ks.data(), fs.data(), c, &o, packed.data(), post.data(), cok.data(), iters.data()));
    std::size_t off = 0;
    for (std::size_t i = 0; i < c; ++i) {
      posts[i].assign(post.begin() + static_cast<std::ptrdiff_t>(off),
                      post.begin() + static_cast<std::ptrdiff_t>(off + ks[i]));
      off += ks[i];
      ok[i] = cok[i] != 0;
      if (params.max_iterations == 0) iters[i] = 0;
    }
  }
  std::vector<std::uint8_t> tb_bits_with_crc;
  for (std::size_t i = 0; i < c; ++i) {
    const auto& blk = encoded.code_blocks[i];
    out.iterations_used = std::max(out.iterations_used, iters[i]);
    auto bits = hard_bits(posts[i]);
    bits.erase(bits.begin(), bits.begin() + static_cast<std::ptrdiff_t>(blk.descriptor.filler_count));
    bool block_ok = true;
    if (c > 1) {
      block_ok = check_crc24b(bits);
      bits.resize(bits.size() - 24);
    }
    tb_bits_with_crc.insert(tb_bits_with_crc.end(), bits.begin(), bits.end());
    out.code_block_crc_ok.push_back(block_ok);
    out.posterior_llr->insert(out.posterior_llr->end(), posts[i].begin(), posts[i].end());
  }
  out.transport_block_crc_ok = check_crc24a(tb_bits_with_crc);
  out.tb_bits.assign(tb_bits_with_crc.begin(), tb_bits_with_crc.end() - 24);
  return out;
}

}  // namespace turbo_cpp
