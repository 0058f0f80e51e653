// C ABI of the transport-block chain on the GPU: soft de-rate matching with
// HARQ combining, batched transport-block decoding (de-rate matching, turbo
// decoding and TB assembly chained on one stream) and batched
// transport-block encoding (CRC24A, segmentation with CRC24B, turbo encoder,
// rate matching).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <stdexcept>
#include <memory>
#include <utility>
#include <vector>

#include "../host/codec_internal.h"
#include "hip_util.h"
#include "decoder_types.h"
#include "derate_plan.h"
#include "rx_types.h"
#include "turbo_cpp/c_api.h"

namespace tcdev {
__global__ void derate_soft_kernel(const DerateCb*, const uint16_t*, const float*, float*, int32_t, int32_t);
__global__ void derate_soft2_kernel(const DerateCb*, const uint16_t*, const float*, int32_t, const float*, int32_t,
                                    float*, int32_t);
__global__ void tb_assemble_kernel(const TbInfo*, const TbCb*, const float*, const uint32_t*, uint32_t,
                                   const uint32_t*, const uint32_t*, uint8_t*, int32_t*, int32_t*, uint32_t*);
__global__ void tb_acc_zero_kernel(uint32_t*, uint32_t);
__global__ void tb_finalize_kernel(const uint32_t*, uint32_t, uint32_t, int32_t*, uint32_t*);
__global__ void tb_crc24a_kernel(const TbInfo*, const uint8_t*, const uint32_t*, uint32_t*);
__global__ void cb_encode_kernel(const TxCb*, const TbInfo*, const uint8_t*, const uint32_t*, const uint32_t*,
                                 uint8_t*, uint8_t*);
__global__ void rate_match_kernel(const TxCb*, const uint32_t*, const uint8_t*, int32_t, uint8_t*);
}  // namespace tcdev

using namespace tchost;


namespace {

void build_derate_plan(tc_derate_plan& p, const uint32_t* k, const uint32_t* filler, const uint64_t* e,
                       const uint32_t* rv, size_t n) {
  if (n && (!k || !e)) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
  std::map<std::pair<uint32_t, uint32_t>, std::pair<uint32_t, turbo_cpp::detail::DerateGeometry>> tables;
  std::vector<uint16_t> rank;
  std::vector<tcdev::DerateCb> cbs(n);
  uint64_t in = 0, out = 0;
  for (size_t i = 0; i < n; ++i) {
    const uint32_t f = filler ? filler[i] : 0;
    if (turbo_cpp::detail::qpp_index(k[i]) < 0) throw Failure{TC_ERR_OUT_OF_RANGE, "unsupported QPP block size"};
    if (f >= k[i]) throw Failure{TC_ERR_INVALID_ARGUMENT, "filler count must be below K"};
    if (e[i] == 0 || e[i] > 0xFFFFFFFFull) throw Failure{TC_ERR_INVALID_ARGUMENT, "E must be in [1, 2^32)"};
    if (rv && rv[i] > 3) throw Failure{TC_ERR_INVALID_ARGUMENT, "rv must be 0..3"};
    const uint32_t d = k[i] + 4;
    auto it = tables.find({d, f});
    if (it == tables.end()) {
      auto g = turbo_cpp::detail::derate_geometry(d, f);
      const uint32_t off = static_cast<uint32_t>(rank.size());
      rank.insert(rank.end(), g.rank.begin(), g.rank.end());
      it = tables.emplace(std::make_pair(d, f), std::make_pair(off, std::move(g))).first;
    }
    const auto& g = it->second.second;
    tcdev::DerateCb& c = cbs[i];
    c.in_off = in;
    c.out_off = out;
    c.e = static_cast<uint32_t>(e[i]);
    c.d = d;
    c.n_valid = g.n_valid;
    c.rank_off = it->second.first;
    std::copy(g.start, g.start + 4, c.start);
    c.rv = rv ? rv[i] : 0;
    c.pad = 0;
    p.max_held = std::max(p.max_held, std::min(c.e, g.n_valid));
    p.ks.push_back(k[i]);
    p.fs.push_back(f);
    in += e[i];
    out += 3ull * d;
  }
  p.n_cb = n;
  p.rx_floats = in;
  p.llr_floats = out;
  if (n == 0) return;
  require_device();
  p.d_cbs = dev_alloc<tcdev::DerateCb>(n);
  dev_upload(p.d_cbs, cbs);
  p.d_rank = dev_alloc<uint16_t>(rank.size());
  dev_upload(p.d_rank, rank);
  uploads_done();
}

void run_derate(const tc_derate_plan& p, const float* rx, int32_t rv, int32_t accumulate, float* llr, hipStream_t s) {
  if (rv < -1 || rv > 3) throw Failure{TC_ERR_INVALID_ARGUMENT, "rv must be -1 (per block) or 0..3"};
  if (p.n_cb == 0) return;
  if (!rx || !llr) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
  if (p.n_cb > 0x7FFFFFFFull) throw Failure{TC_ERR_INVALID_ARGUMENT, "too many code blocks"};
  const size_t lds = static_cast<size_t>(p.max_held) * sizeof(float);
  hipLaunchKernelGGL(tcdev::derate_soft_kernel, dim3(static_cast<uint32_t>(p.n_cb)), dim3(tcdev::kDerateThreads), lds,
                     s, p.d_cbs, p.d_rank, rx, llr, rv, accumulate ? 1 : 0);
  hip_check(hipGetLastError(), "de-rate matching launch");
}

// Two transmissions in one pass (HARQ combining): the same result as
// run_derate(rx_a, rv_a, accumulate) then run_derate(rx_b, rv_b, 1).
void run_derate2(const tc_derate_plan& p, const float* rx_a, int32_t rv_a, const float* rx_b, int32_t rv_b,
                 int32_t accumulate, float* llr, hipStream_t s) {
  if (rv_a < -1 || rv_a > 3 || rv_b < -1 || rv_b > 3)
    throw Failure{TC_ERR_INVALID_ARGUMENT, "rv must be -1 (per block) or 0..3"};
  if (p.n_cb == 0) return;
  if (!rx_a || !rx_b || !llr) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
  if (p.n_cb > 0x7FFFFFFFull) throw Failure{TC_ERR_INVALID_ARGUMENT, "too many code blocks"};
  const size_t lds = 2 * static_cast<size_t>(p.max_held) * sizeof(float);
  hipLaunchKernelGGL(tcdev::derate_soft2_kernel, dim3(static_cast<uint32_t>(p.n_cb)), dim3(tcdev::kDerateThreads),
                     lds, s, p.d_cbs, p.d_rank, rx_a, rv_a, rx_b, rv_b, llr, accumulate ? 1 : 0);
  hip_check(hipGetLastError(), "de-rate matching launch");
}

}  // namespace

struct tc_tb_plan {
  size_t n_tb = 0, n_cb = 0;
  uint64_t bits_total = 0;
  uint32_t max_iterations = 0;
  bool exact = false;  // exact decoder: de-rate matching goes through d_soft
  std::vector<uint32_t> ks, fs;
  std::vector<uint64_t> es;
  tc_derate_plan* derate = nullptr;
  tc_plan* decoder = nullptr;
  tcdev::TbInfo* d_tbs = nullptr;
  tcdev::TbCb* d_cbs = nullptr;
  uint32_t* d_pow24a = nullptr;
  uint32_t* d_pow24b = nullptr;
  float* d_soft = nullptr;  // decoder input; the HARQ soft buffer between calls
  bool hold = true;         // decodes from received LLRs refresh d_soft
  int32_t staging = TC_STAGING_AUTO;  // window decoders: where de-rate matching runs
  float* d_post = nullptr;  // posteriors when the caller passes none
  uint32_t* d_cb_iters = nullptr;
  // transmit side
  tcdev::TxCb* d_tx = nullptr;
  uint32_t* d_valid = nullptr;   // per (K, F): non-NULL circular-buffer entries in order
  uint8_t* d_streams = nullptr;  // encoder output [d0|d1|d2] per block
  uint32_t* d_tb_crc = nullptr;  // CRC24A per TB
  // TB assembly in the window decoder's epilogue
  tcdev::TbCbMap* d_map = nullptr;
  uint32_t* d_acc = nullptr;  // 3 words per TB
  ~tc_tb_plan() {
    tc_derate_plan_destroy(derate);
    tc_plan_destroy(decoder);
    for (void* p : {(void*)d_tbs, (void*)d_cbs, (void*)d_pow24a, (void*)d_pow24b, (void*)d_soft, (void*)d_post,
                    (void*)d_cb_iters, (void*)d_tx, (void*)d_valid, (void*)d_streams, (void*)d_tb_crc,
                    (void*)d_map, (void*)d_acc})
      if (p) (void)hipFree(p);
  }
};

namespace {

// D^m mod G(D) for m < n, G = D^24 + poly (24 low coefficients).
std::vector<uint32_t> crc_pow_table(uint32_t poly, size_t n) {
  std::vector<uint32_t> t(n);
  uint32_t r = 1;
  for (size_t m = 0; m < n; ++m) {
    t[m] = r;
    r = ((r << 1) & 0xFFFFFFu) ^ (((r >> 23) & 1u) ? poly : 0u);
  }
  return t;
}

void check_status(tc_status s) {
  if (s != TC_OK) throw Failure{s, turbo_cpp::detail::last_error()};
}

void build_tb_plan(tc_tb_plan& p, const tc_tb_desc* tbs, size_t n_tb, const tc_options* opts) {
  if (n_tb && !tbs) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
  std::vector<tcdev::TbInfo> info(n_tb);
  std::vector<tcdev::TbCb> cbs;
  std::vector<uint32_t> rvs;
  std::vector<tcdev::TxCb> tx;
  std::vector<tcdev::TbCbMap> map;
  std::vector<uint32_t> valid;
  struct ValidInfo {
    uint32_t off, n, start[4];
  };
  std::map<std::pair<uint32_t, uint32_t>, ValidInfo> valid_of;
  uint64_t bits = 0, post = 0, max_len = 24, d_bytes = 0, rm_bits = 0;
  for (size_t t = 0; t < n_tb; ++t) {
    const tc_tb_desc& d = tbs[t];
    if (d.transport_block_bits == 0) throw Failure{TC_ERR_INVALID_ARGUMENT, "transport block cannot be empty"};
    if (d.rv > 3) throw Failure{TC_ERR_INVALID_ARGUMENT, "rv must be 0..3"};
    if (d.transport_block_bits > 0xFFFFFFFFull - 24) throw Failure{TC_ERR_INVALID_ARGUMENT, "transport block too long"};
    turbo_cpp::SegmentationResult seg;
    std::vector<uint64_t> e(0);
    try {
      if (d.code_block_crc) {
        if (turbo_cpp::detail::qpp_index(d.transport_block_bits + 24) < 0)
          throw Failure{TC_ERR_OUT_OF_RANGE, "code_block_crc needs B + 24 to be a QPP block size"};
        seg.c = 1;
        seg.c_plus = 1;
        seg.k_plus = d.transport_block_bits + 24;
        seg.filler_bits = 0;
        seg.l = 24;
      } else {
        seg = turbo_cpp::detail::segmentation_sizes(d.transport_block_bits + 24);
      }
      for (size_t r = 0; r < seg.c; ++r)
        e.push_back(turbo_cpp::detail::compute_e(r, d.g_total, seg.c, d.num_layers, d.mod_order));
    } catch (const std::invalid_argument& ex) {
      throw Failure{TC_ERR_INVALID_ARGUMENT, ex.what()};
    } catch (const std::runtime_error& ex) {
      throw Failure{TC_ERR_INVALID_ARGUMENT, ex.what()};
    }
    info[t] = {bits, static_cast<uint32_t>(d.transport_block_bits), static_cast<uint32_t>(p.ks.size()),
               static_cast<uint32_t>(seg.c), d.code_block_crc ? 1u : 0u};
    uint32_t cursor = 0;
    for (size_t r = 0; r < seg.c; ++r) {
      const uint32_t k = static_cast<uint32_t>(r < seg.c_minus ? seg.k_minus : seg.k_plus);
      const uint32_t f = static_cast<uint32_t>(r == 0 ? seg.filler_bits : 0);
      if (e[r] == 0) throw Failure{TC_ERR_INVALID_ARGUMENT, "g_total leaves a code block without coded bits"};
      p.ks.push_back(k);
      p.fs.push_back(f);
      p.es.push_back(e[r]);
      rvs.push_back(d.rv);
      cbs.push_back({post, k, f});
      tcdev::TxCb x{};
      x.cb_off = post;
      x.d_off = d_bytes;
      x.rm_off = rm_bits;
      x.tb = static_cast<uint32_t>(t);
      x.k = k;
      x.filler = f;
      x.cursor = cursor;
      x.seg = (seg.c > 1 || d.code_block_crc) ? 1u : 0u;
      x.payload = k - f - 24 * x.seg;
      const auto& qe = turbo_cpp::detail::kQppTable[turbo_cpp::detail::qpp_index(k)];
      x.f1 = static_cast<uint32_t>(qe.f1);
      x.f2 = static_cast<uint32_t>(qe.f2);
      x.e = static_cast<uint32_t>(e[r]);
      x.rv = d.rv;
      auto it = valid_of.find({k, f});
      if (it == valid_of.end()) {
        const auto g = turbo_cpp::detail::derate_geometry(k + 4, f);
        std::vector<uint32_t> v(g.n_valid);
        for (uint32_t s2 = 0; s2 < 3; ++s2)
          for (uint32_t i = 0; i < k + 4; ++i) {
            const uint16_t rk = g.rank[s2 * (k + 4) + i];
            if (rk != tcdev::kNullRank) v[rk] = (s2 << 16) | i;
          }
        ValidInfo vi{static_cast<uint32_t>(valid.size()), g.n_valid, {g.start[0], g.start[1], g.start[2], g.start[3]}};
        valid.insert(valid.end(), v.begin(), v.end());
        it = valid_of.emplace(std::make_pair(k, f), vi).first;
      }
      x.valid_off = it->second.off;
      x.n_valid = it->second.n;
      std::copy(it->second.start, it->second.start + 4, x.start);
      tx.push_back(x);
      map.push_back({static_cast<uint32_t>(t), cursor, x.payload, x.seg});
      cursor += x.payload;
      d_bytes += 3ull * (k + 4);
      rm_bits += e[r];
      post += k;
    }
    bits += d.transport_block_bits;
    max_len = std::max<uint64_t>(max_len, d.transport_block_bits + 24);
  }
  p.n_tb = n_tb;
  p.n_cb = p.ks.size();
  p.bits_total = bits;
  const tc_options o = opts ? *opts : tc_options{8, 0, TC_DECODER_WINDOW_LOG, 0, 0, 0};
  p.max_iterations = o.max_iterations;
  p.exact = o.decoder == TC_DECODER_EXACT_SERIAL;
  check_status(tc_derate_plan_create(p.ks.data(), p.fs.data(), p.es.data(), rvs.data(), p.n_cb, &p.derate));
  check_status(tc_plan_create(p.ks.data(), p.fs.data(), p.n_cb, &o, &p.decoder));
  if (n_tb == 0) return;
  p.d_tbs = dev_alloc<tcdev::TbInfo>(n_tb);
  dev_upload(p.d_tbs, info);
  p.d_cbs = dev_alloc<tcdev::TbCb>(cbs.size());
  dev_upload(p.d_cbs, cbs);
  const auto pa = crc_pow_table(tcdev::kCrc24APoly, max_len);
  const auto pb = crc_pow_table(tcdev::kCrc24BPoly, 6144);
  p.d_pow24a = dev_alloc<uint32_t>(pa.size());
  dev_upload(p.d_pow24a, pa);
  p.d_pow24b = dev_alloc<uint32_t>(pb.size());
  dev_upload(p.d_pow24b, pb);
  p.d_soft = dev_alloc<float>(tc_derate_plan_llr_floats(p.derate));
  p.d_post = dev_alloc<float>(post);
  p.d_cb_iters = dev_alloc<uint32_t>(p.n_cb);
  p.d_tx = dev_alloc<tcdev::TxCb>(tx.size());
  dev_upload(p.d_tx, tx);
  p.d_valid = dev_alloc<uint32_t>(valid.size());
  dev_upload(p.d_valid, valid);
  p.d_streams = dev_alloc<uint8_t>(d_bytes);
  p.d_tb_crc = dev_alloc<uint32_t>(n_tb);
  p.d_map = dev_alloc<tcdev::TbCbMap>(map.size());
  dev_upload(p.d_map, map);
  p.d_acc = dev_alloc<uint32_t>(3 * n_tb);
  uploads_done();
}

void run_encode(tc_tb_plan& p, const uint8_t* payload, int32_t rv, uint8_t* rm, uint8_t* cb_bits, hipStream_t s) {
  if (rv < -1 || rv > 3) throw Failure{TC_ERR_INVALID_ARGUMENT, "rv must be -1 (per block) or 0..3"};
  if (p.n_tb == 0) return;
  if (!payload || !rm) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
  if (p.n_tb > 0x7FFFFFFFull || p.n_cb > 0x7FFFFFFFull) throw Failure{TC_ERR_INVALID_ARGUMENT, "batch too large"};
  hipLaunchKernelGGL(tcdev::tb_crc24a_kernel, dim3(static_cast<uint32_t>(p.n_tb)), dim3(256), 0, s, p.d_tbs, payload,
                     p.d_pow24a, p.d_tb_crc);
  hip_check(hipGetLastError(), "CRC24A launch");
  hipLaunchKernelGGL(tcdev::cb_encode_kernel, dim3(static_cast<uint32_t>(p.n_cb)), dim3(64), 0, s, p.d_tx, p.d_tbs,
                     payload, p.d_tb_crc, p.d_pow24b, cb_bits, p.d_streams);
  hip_check(hipGetLastError(), "encoder launch");
  hipLaunchKernelGGL(tcdev::rate_match_kernel, dim3(static_cast<uint32_t>(p.n_cb)), dim3(256), 0, s, p.d_tx,
                     p.d_valid, p.d_streams, rv, rm);
  hip_check(hipGetLastError(), "rate matching launch");
}

// n_tx received transmissions (0: decode the held input as it stands): the
// decoder de-rate matches them in its input staging and refreshes the held
// input (the HARQ soft buffer), then the TB assembly runs.
void run_tb(tc_tb_plan& p, uint32_t n_tx, const float* const* rx, const int32_t* rv, int32_t accumulate,
            uint8_t* bits, int32_t* tb_ok, int32_t* cb_ok, uint32_t* iters, float* post, hipStream_t s) {
  for (uint32_t t = 0; rv && t < n_tx; ++t)
    if (rv[t] < -1 || rv[t] > 3) throw Failure{TC_ERR_INVALID_ARGUMENT, "rv must be -1 (per block) or 0..3"};
  if (p.n_tb == 0) return;
  float* post_buf = post ? post : p.d_post;
  if (accumulate && !p.hold)
    throw Failure{TC_ERR_INVALID_ARGUMENT, "the plan does not hold its decoder input (tc_tb_plan_set_hold)"};
  const bool exact = p.exact;
  const bool separate = n_tx && p.staging == TC_STAGING_SEPARATE;
  if (!exact && separate) {
    // window decoders, de-rate kernels first: each transmission into d_soft (added in order onto
    // the held input or the transmissions before it, the fused staging's order), then the decode
    // of d_soft with the TB assembly in its epilogue
    // (transmissions taken two at a time in one pass each, the same left-to-right sums)
    for (uint32_t t = 0; t < n_tx;) {
      if (!rx[t]) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
      const int32_t acc = (t > 0 || accumulate) ? 1 : 0;
      if (t + 1 < n_tx) {
        if (!rx[t + 1]) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
        run_derate2(*p.derate, rx[t], rv ? rv[t] : -1, rx[t + 1], rv ? rv[t + 1] : -1, acc, p.d_soft, s);
        t += 2;
      } else {
        run_derate(*p.derate, rx[t], rv ? rv[t] : -1, acc, p.d_soft, s);
        t += 1;
      }
    }
    n_tx = 0;
  }
  if (!exact) {
    // window decoders: de-rate matching in the input staging, TB assembly in the epilogue
    if (p.n_tb > 0x7FFFFFFFull / 3) throw Failure{TC_ERR_INVALID_ARGUMENT, "too many transport blocks"};
    const uint32_t nacc = static_cast<uint32_t>(3 * p.n_tb);
    hipLaunchKernelGGL(tcdev::tb_acc_zero_kernel, dim3((nacc + 255) / 256), dim3(256), 0, s, p.d_acc, nacc);
    hip_check(hipGetLastError(), "accumulator clear launch");
    tcdev::TbOut tb{1, p.max_iterations, p.d_tbs, p.d_map, p.d_pow24a, p.d_pow24b, bits, tb_ok, cb_ok, iters,
                    p.d_acc};
    tcdev::RxIn x{};
    if (n_tx) {
      if (n_tx > static_cast<uint32_t>(tcdev::kMaxTx)) throw Failure{TC_ERR_INVALID_ARGUMENT, "n_tx must be 1..4"};
      for (uint32_t t = 0; t < static_cast<uint32_t>(tcdev::kMaxTx); ++t) {
        const int32_t r = (rv && t < n_tx) ? rv[t] : -1;
        x.rv[t] = r < 0 ? tcdev::kBlockRv : static_cast<uint32_t>(r);
        x.rx[t] = t < n_tx ? rx[t] : nullptr;
        if (t < n_tx && !x.rx[t]) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
      }
      x.n_tx = n_tx;
      x.base = accumulate ? 1 : 0;
      x.keep = p.hold ? p.d_soft : nullptr;
      x.rank = p.derate->d_rank;
      x.dcb = p.derate->d_cbs;
    }
    if (!tchost::launch_window(p.decoder, n_tx ? nullptr : p.d_soft, post_buf, nullptr, p.d_cb_iters, s,
                               n_tx ? &x : nullptr, &tb))
      throw Failure{TC_ERR_RUNTIME, "window decoder expected"};
    if (tb_ok || iters) {
      hipLaunchKernelGGL(tcdev::tb_finalize_kernel, dim3(static_cast<uint32_t>((p.n_tb + 255) / 256)), dim3(256), 0, s,
                         p.d_acc, static_cast<uint32_t>(p.n_tb), p.max_iterations, tb_ok, iters);
      hip_check(hipGetLastError(), "transport block verdict launch");
    }
    return;
  }
  if (n_tx)
    check_status(tc_plan_decode_rx(p.decoder, p.derate, n_tx, rx, rv, accumulate,
                                   (p.hold || exact) ? p.d_soft : nullptr, post_buf, nullptr, p.d_cb_iters, s));
  else
    check_status(tc_plan_decode(p.decoder, p.d_soft, post_buf, nullptr, p.d_cb_iters, s));
  if (p.n_tb > 0x7FFFFFFFull) throw Failure{TC_ERR_INVALID_ARGUMENT, "too many transport blocks"};
  hipLaunchKernelGGL(tcdev::tb_assemble_kernel, dim3(static_cast<uint32_t>(p.n_tb)), dim3(tcdev::kTbThreads), 0, s,
                     p.d_tbs, p.d_cbs, post_buf, p.d_cb_iters, p.max_iterations, p.d_pow24a, p.d_pow24b, bits, tb_ok,
                     cb_ok, iters);
  hip_check(hipGetLastError(), "transport block assembly launch");
}

}  // namespace

extern "C" {

tc_status tc_derate_plan_create(const uint32_t* k, const uint32_t* filler, const uint64_t* e, const uint32_t* rv,
                                size_t n_cb, tc_derate_plan** plan) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    auto p = std::make_unique<tc_derate_plan>();
    build_derate_plan(*p, k, filler, e, rv, n_cb);
    uploads_done();
    *plan = p.release();
  });
}

void tc_derate_plan_destroy(tc_derate_plan* plan) { delete plan; }
size_t tc_derate_plan_rx_floats(const tc_derate_plan* plan) { return plan ? plan->rx_floats : 0; }
size_t tc_derate_plan_llr_floats(const tc_derate_plan* plan) { return plan ? plan->llr_floats : 0; }

tc_status tc_derate_plan_run(tc_derate_plan* plan, const float* rx_dev, int32_t rv, int32_t accumulate,
                             float* llr_dev, void* hip_stream) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    run_derate(*plan, rx_dev, rv, accumulate, llr_dev, static_cast<hipStream_t>(hip_stream));
  });
}

tc_status tc_derate_match_soft(const float* rx, uint64_t e, uint32_t k, uint32_t filler, uint32_t rv,
                               int32_t accumulate, float* out) {
  return guarded([&] {
    if (!rx || !out) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    tc_derate_plan p;
    build_derate_plan(p, &k, &filler, &e, &rv, 1);
    const hipStream_t s = tchost::host_stream();
    DevBuf<float> d_rx(e), d_out(p.llr_floats);
    hip_check(hipMemcpyAsync(d_rx.p, rx, e * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    if (accumulate)
      hip_check(hipMemcpyAsync(d_out.p, out, p.llr_floats * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    run_derate(p, d_rx.p, -1, accumulate, d_out.p, s);
    hip_check(hipMemcpyAsync(out, d_out.p, p.llr_floats * sizeof(float), hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(s), "de-rate matching");
  });
}

tc_status tc_tb_plan_create(const tc_tb_desc* tbs, size_t n_tb, const tc_options* opts, tc_tb_plan** plan) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    auto p = std::make_unique<tc_tb_plan>();
    build_tb_plan(*p, tbs, n_tb, opts);
    uploads_done();
    *plan = p.release();
  });
}

void tc_tb_plan_destroy(tc_tb_plan* plan) { delete plan; }

tc_status tc_tb_plan_sizes(const tc_tb_plan* plan, uint64_t sizes[4]) {
  return guarded([&] {
    if (!plan || !sizes) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    sizes[0] = plan->n_cb;
    sizes[1] = tc_derate_plan_rx_floats(plan->derate);
    sizes[2] = plan->bits_total;
    sizes[3] = tc_plan_post_floats(plan->decoder);
  });
}

tc_status tc_tb_plan_code_blocks(const tc_tb_plan* plan, uint32_t* k, uint32_t* filler, uint64_t* e) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    if (k) std::copy(plan->ks.begin(), plan->ks.end(), k);
    if (filler) std::copy(plan->fs.begin(), plan->fs.end(), filler);
    if (e) std::copy(plan->es.begin(), plan->es.end(), e);
  });
}

tc_status tc_tb_plan_decode(tc_tb_plan* plan, const float* rx_dev, int32_t rv, int32_t accumulate, uint8_t* bits_dev,
                            int32_t* tb_crc_ok_dev, int32_t* cb_crc_ok_dev, uint32_t* iterations_dev,
                            float* posterior_dev, void* hip_stream) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    if (rv < -1 || rv > 3) throw Failure{TC_ERR_INVALID_ARGUMENT, "rv must be -1 (per block) or 0..3"};
    const float* rxs[1] = {rx_dev};
    const int32_t rvs[1] = {rv};
    run_tb(*plan, rx_dev ? 1u : 0u, rxs, rvs, accumulate, bits_dev, tb_crc_ok_dev, cb_crc_ok_dev, iterations_dev,
           posterior_dev, static_cast<hipStream_t>(hip_stream));
  });
}

tc_status tc_tb_plan_set_hold(tc_tb_plan* plan, int32_t hold) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    plan->hold = hold != 0;
  });
}

tc_status tc_tb_plan_set_staging(tc_tb_plan* plan, int32_t mode) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    if (mode < TC_STAGING_AUTO || mode > TC_STAGING_SEPARATE)
      throw Failure{TC_ERR_INVALID_ARGUMENT, "staging must be TC_STAGING_AUTO, _FUSED or _SEPARATE"};
    plan->staging = mode;
  });
}

tc_status tc_tb_plan_decode_multi(tc_tb_plan* plan, uint32_t n_tx, const float* const* rx_dev, const int32_t* rv,
                                  uint8_t* bits_dev, int32_t* tb_crc_ok_dev, int32_t* cb_crc_ok_dev,
                                  uint32_t* iterations_dev, float* posterior_dev, void* hip_stream) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    if (n_tx == 0 || n_tx > 4) throw Failure{TC_ERR_INVALID_ARGUMENT, "n_tx must be 1..4"};
    if (!rx_dev) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    run_tb(*plan, n_tx, rx_dev, rv, 0, bits_dev, tb_crc_ok_dev, cb_crc_ok_dev, iterations_dev, posterior_dev,
           static_cast<hipStream_t>(hip_stream));
  });
}

tc_status tc_tb_plan_soft_combine(tc_tb_plan* plan, const float* rx_dev, int32_t rv, int32_t accumulate,
                                  void* hip_stream) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    if (plan->n_tb == 0) return;
    if (!rx_dev) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    run_derate(*plan->derate, rx_dev, rv, accumulate, plan->d_soft, static_cast<hipStream_t>(hip_stream));
  });
}

tc_status tc_tb_plan_encode(tc_tb_plan* plan, const uint8_t* payload_dev, int32_t rv, uint8_t* rm_bits_dev,
                            uint8_t* cb_bits_dev, void* hip_stream) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    run_encode(*plan, payload_dev, rv, rm_bits_dev, cb_bits_dev, static_cast<hipStream_t>(hip_stream));
  });
}

tc_status tc_encode_transport_blocks_host(const tc_tb_desc* tbs, size_t n_tb, const uint8_t* payload,
                                          uint8_t* rm_bits, uint8_t* cb_bits) {
  return guarded([&] {
    tc_tb_plan p;
    build_tb_plan(p, tbs, n_tb, nullptr);
    if (n_tb == 0) return;
    if (!payload || !rm_bits) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    const size_t nrm = tc_derate_plan_rx_floats(p.derate), ncb = tc_plan_post_floats(p.decoder);
    const hipStream_t s = tchost::host_stream();
    DevBuf<uint8_t> d_pay(p.bits_total), d_rm(nrm), d_cb(ncb);
    hip_check(hipMemcpyAsync(d_pay.p, payload, p.bits_total, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    run_encode(p, d_pay.p, -1, d_rm.p, cb_bits ? d_cb.p : nullptr, s);
    hip_check(hipMemcpyAsync(rm_bits, d_rm.p, nrm, hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    if (cb_bits) hip_check(hipMemcpyAsync(cb_bits, d_cb.p, ncb, hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(s), "transport block encode");
  });
}

tc_status tc_decode_transport_blocks_host(const tc_tb_desc* tbs, size_t n_tb, const tc_options* opts, const float* rx,
                                          uint8_t* bits, int32_t* tb_crc_ok, int32_t* cb_crc_ok,
                                          uint32_t* iterations, float* posterior) {
  return guarded([&] {
    tc_tb_plan p;
    build_tb_plan(p, tbs, n_tb, opts);
    if (n_tb == 0) return;
    if (!rx || !bits || !tb_crc_ok) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    const size_t nrx = tc_derate_plan_rx_floats(p.derate), npost = tc_plan_post_floats(p.decoder);
    const hipStream_t s = tchost::host_stream();
    DevBuf<float> d_rx(nrx), d_post(npost);
    DevBuf<uint8_t> d_bits(p.bits_total);
    DevBuf<int32_t> d_tb(n_tb), d_cb(p.n_cb);
    DevBuf<uint32_t> d_it(n_tb);
    hip_check(hipMemcpyAsync(d_rx.p, rx, nrx * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    const float* rxs[1] = {d_rx.p};
    run_tb(p, 1, rxs, nullptr, 0, d_bits.p, d_tb.p, d_cb.p, d_it.p, d_post.p, s);
    hip_check(hipMemcpyAsync(bits, d_bits.p, p.bits_total, hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    hip_check(hipMemcpyAsync(tb_crc_ok, d_tb.p, n_tb * sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    if (cb_crc_ok) hip_check(hipMemcpyAsync(cb_crc_ok, d_cb.p, p.n_cb * sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    if (iterations) hip_check(hipMemcpyAsync(iterations, d_it.p, n_tb * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    if (posterior) hip_check(hipMemcpyAsync(posterior, d_post.p, npost * sizeof(float), hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(s), "transport block decode");
  });
}

}  // extern "C"
