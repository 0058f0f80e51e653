// C++ API test of libturbo_cpp_amd through include/turbo_cpp/lte_turbo.h,
// mirroring the assertions of the reference's tests/test_main.cpp (cited per
// test).  Usage: api_test cpu | gpu.  The gpu mode runs the GPU backends
// (exact and windowed) where the reference compiled them under
// TURBO_CPP_HAS_CUDA.
#include <cmath>
#include <cstring>
#include <future>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "turbo_cpp/lte_turbo.h"

namespace tc = turbo_cpp;

namespace {

bool g_gpu = false;

void expect(bool ok, const char* what) {
  if (!ok) throw std::runtime_error(what);
}

std::vector<std::uint8_t> hard(const std::vector<float>& p) {
  std::vector<std::uint8_t> b(p.size());
  for (std::size_t i = 0; i < p.size(); ++i) b[i] = p[i] >= 0.0F;
  return b;
}

std::vector<std::uint8_t> pattern(std::size_t n, std::size_t a, std::size_t b) {
  std::vector<std::uint8_t> v(n);
  for (std::size_t i = 0; i < n; ++i) v[i] = static_cast<std::uint8_t>((i * a + b) & 1U);
  return v;
}

std::size_t full_rate_g(const std::vector<std::uint8_t>& bits) {
  const auto seg = tc::segment_transport_block(tc::append_crc24a(bits));
  std::size_t g = 0;
  for (const auto& c : seg.code_blocks) g += 3 * (c.size() + 4);
  return g;
}

void test_crc() {  // test_main.cpp:25-35
  const std::vector<std::uint8_t> bits = {1, 0, 1, 1, 0, 1, 0, 0, 1, 1, 0, 1};
  expect(tc::check_crc24a(tc::append_crc24a(bits)), "CRC24A failed");
  expect(tc::check_crc24b(tc::append_crc24b(bits)), "CRC24B failed");
  expect(tc::check_crc16(tc::append_crc16(bits)), "CRC16 failed");
  expect(tc::check_crc8(tc::append_crc8(bits)), "CRC8 failed");
}

void test_qpp_table() {  // test_main.cpp:37-44
  const auto t = tc::qpp_interleaver_table();
  expect(t.size() == 188, "QPP table size mismatch");
  expect(t.front().k == 40 && t.front().f1 == 3 && t.front().f2 == 10, "QPP first entry mismatch");
  expect(t.back().k == 6144 && t.back().f1 == 263 && t.back().f2 == 480, "QPP last entry mismatch");
  bool threw = false;
  try {
    (void)tc::qpp_entry_for_k(41);
  } catch (const std::out_of_range&) {
    threw = true;
  }
  expect(threw, "K=41 must throw out_of_range");
}

void test_segmentation() {  // test_main.cpp:46-56
  const auto s40 = tc::segment_transport_block(std::vector<std::uint8_t>(40, 1));
  expect(s40.c == 1 && s40.code_blocks.front().size() == 40, "B=40 segmentation");
  const auto s6145 = tc::segment_transport_block(std::vector<std::uint8_t>(6145, 0));
  expect(s6145.c > 1 && s6145.filler_bits > 0, "B=6145 segmentation");
}

void roundtrip(std::size_t tbs, std::size_t a, std::size_t b, bool compare_stage1) {
  tc::LteTurboCodecParams p;
  p.transport_block_bits = tbs;
  p.max_iterations = 6;
  const auto bits = pattern(tbs, a, b);
  p.g_total = full_rate_g(bits);
  const auto enc = tc::encode_transport_block(p, bits);
  const auto cpu = tc::decode_transport_block(p, enc, tc::DecoderBackend::CpuReference);
  expect(cpu.tb_bits == bits, "CPU roundtrip mismatch");
  expect(cpu.transport_block_crc_ok, "CPU TB CRC");
  if (!g_gpu) return;
  const auto s1 = tc::decode_transport_block(p, enc, tc::DecoderBackend::GpuStage1Exact);
  expect(s1.tb_bits == bits, "GPU Stage1 roundtrip mismatch");
  expect(s1.transport_block_crc_ok, "GPU Stage1 CRC");
  expect(s1.posterior_llr.has_value() && s1.posterior_llr->size() == cpu.posterior_llr->size(), "Stage1 size");
  if (compare_stage1) {
    for (std::size_t i = 0; i < cpu.posterior_llr->size(); ++i)  // test_main.cpp:141-151
      expect(std::fabs(cpu.posterior_llr->at(i) - s1.posterior_llr->at(i)) <= 2e-5F * std::max(1.0F, std::fabs(cpu.posterior_llr->at(i))),
             "CPU vs GPU Stage1 posterior mismatch");
  }
  const auto s2 = tc::decode_transport_block(p, enc, tc::DecoderBackend::GpuStage2Windowed);
  expect(s2.tb_bits == bits, "GPU Stage2 roundtrip mismatch");
  expect(s2.transport_block_crc_ok, "GPU Stage2 CRC");
  for (float v : *s2.posterior_llr) expect(std::isfinite(v), "Stage2 posterior non-finite");
  // tile8 hard decisions (test_main.cpp:153-175)
  const auto& blk = enc.code_blocks.front();
  const auto llr = tc::derate_match_code_block(p, blk.descriptor, blk.rate_matched_bits);
  const std::size_t d = blk.descriptor.d_r;
  bool ok = false;
  std::uint32_t it = 0;
  const auto t8 = tc::turbo_decode_code_block_gpu_stage1_tile8_experimental(
      blk.descriptor, std::span<const float>(llr.data(), d), std::span<const float>(llr.data() + d, d),
      std::span<const float>(llr.data() + 2 * d, d), p.max_iterations, false, &ok, &it);
  expect(t8.size() == blk.cb_bits.size() && hard(t8) == blk.cb_bits, "tile8 hard decisions");
}

void test_errors() {
  tc::CodeBlockDescriptor d;
  d.k_r = 40;
  d.d_r = 44;
  std::vector<float> z(44, 0.0F), z43(43, 0.0F);
  bool threw = false;
  try {
    (void)tc::turbo_decode_code_block_cpu(d, z43, z, z, 8, false);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  expect(threw, "CPU decode with wrong span size must throw invalid_argument");
  threw = false;
  try {
    (void)tc::turbo_decode_code_block_gpu_stage2(d, z43, z, z, 8, false);
  } catch (const std::invalid_argument&) {
    threw = true;
  } catch (const std::runtime_error&) {
    threw = !g_gpu;  // no device: runtime_error (lte_turbo.cpp:861)
  }
  expect(threw, "GPU decode with wrong span size must throw");
}

// The reference's decode_transport_block runs one GPU decode per code block
// on std::async threads (lte_turbo.cpp:941-951): eight host threads decode
// different blocks at once through both GPU entry points; every result equals
// the same call made alone, bit for bit.
void test_concurrent_calls() {
  if (!g_gpu) return;
  constexpr int kBlocks = 8;
  const std::size_t k = 1024, d = k + 4;
  std::vector<std::vector<float>> llr(kBlocks);
  std::uint32_t lcg = 12345;
  for (int b = 0; b < kBlocks; ++b) {
    std::vector<std::uint8_t> cb(k);
    for (auto& x : cb) x = (lcg = lcg * 1664525U + 1013904223U) >> 31;
    tc::CodeBlockDescriptor cd;
    cd.k_r = k;
    cd.d_r = d;
    const auto enc = tc::turbo_encode_code_block(cd, cb);
    expect(enc.systematic.size() == d && enc.parity0.size() == d && enc.parity1.size() == d, "encoder stream size");
    llr[b].resize(3 * d);
    for (std::size_t i = 0; i < d; ++i) {
      const std::uint8_t bits[3] = {enc.systematic[i], enc.parity0[i], enc.parity1[i]};
      for (int s = 0; s < 3; ++s) {
        const float noise = static_cast<float>((lcg = lcg * 1664525U + 1013904223U) >> 8) / 16777216.0F - 0.5F;
        llr[b][s * d + i] = (bits[s] ? 1.0F : -1.0F) * 1.5F + 3.0F * noise;  // bit 1 > 0
      }
    }
  }
  tc::CodeBlockDescriptor desc;
  desc.k_r = k;
  desc.d_r = d;
  auto decode = [&](int b, bool windowed) {
    const std::span<const float> d0(llr[b].data(), d), d1(llr[b].data() + d, d), d2(llr[b].data() + 2 * d, d);
    return windowed ? tc::turbo_decode_code_block_gpu_stage2(desc, d0, d1, d2, 6, false)
                    : tc::turbo_decode_code_block_gpu_stage1(desc, d0, d1, d2, 6, false);
  };
  std::vector<std::vector<float>> alone[2];
  for (int w = 0; w < 2; ++w)
    for (int b = 0; b < kBlocks; ++b) alone[w].push_back(decode(b, w == 1));
  for (int round = 0; round < 3; ++round) {
    std::vector<std::future<std::pair<std::vector<float>, std::vector<float>>>> jobs;
    for (int b = 0; b < kBlocks; ++b)
      jobs.push_back(std::async(std::launch::async, [&, b] { return std::make_pair(decode(b, false), decode(b, true)); }));
    for (int b = 0; b < kBlocks; ++b) {
      const auto r = jobs[b].get();
      expect(std::memcmp(r.first.data(), alone[0][b].data(), k * sizeof(float)) == 0, "concurrent Stage1 decode differs");
      expect(std::memcmp(r.second.data(), alone[1][b].data(), k * sizeof(float)) == 0, "concurrent Stage2 decode differs");
    }
  }
}

}  // namespace

int main(int argc, char** argv) {
  g_gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
  try {
    test_crc();
    test_qpp_table();
    test_segmentation();
    roundtrip(40, 7, 3, true);       // test_main.cpp:58-187
    roundtrip(512, 11, 5, false);    // :189-255
    roundtrip(6112, 19, 11, false);  // :257-303
    roundtrip(12000, 17, 9, false);  // :305-344
    test_errors();
    test_concurrent_calls();
    std::cout << "All tests passed (" << (g_gpu ? "gpu" : "cpu") << ")\n";
    return 0;
  } catch (const std::exception& e) {
    std::cerr << "Test failure: " << e.what() << '\n';
    return 1;
  }
}
