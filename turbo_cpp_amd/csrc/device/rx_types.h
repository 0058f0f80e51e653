// Plain structures shared by the receive-chain plan builders and kernels:
// soft de-rate matching with HARQ combining, and transport-block assembly.
#pragma once

#include <cstdint>

namespace tcdev {

// One code block of a soft de-rate matching batch.  Received LLR k of the
// block (in_off + k, k < e) lands on circular-buffer rank
// (start[rv] + k) mod n_valid; rank_off locates the (K, F) table that gives
// the rank of each [d0|d1|d2] output position (0xFFFF = NULL entry).
struct DerateCb {
  uint64_t in_off;
  uint64_t out_off;
  uint32_t e;
  uint32_t d;
  uint32_t n_valid;
  uint32_t rank_off;
  uint32_t start[4];
  uint32_t rv;
  uint32_t pad;
};
static_assert(sizeof(DerateCb) == 56, "DerateCb layout");

// Transport-block assembly: one TB of B payload bits carried by code blocks
// cb_first .. cb_first + c - 1; its payload lands at bits_off of the output.
// cb_crc: a single code block of B + 24 bits ending in a CRC24B (no CRC24A).
struct TbInfo {
  uint64_t bits_off;
  uint32_t b;
  uint32_t cb_first;
  uint32_t c;
  uint32_t cb_crc;  // 1: one code block carrying a CRC24B instead of a CRC24A
};
static_assert(sizeof(TbInfo) == 24, "TbInfo layout");

// A code block as the assembly sees it: its K posteriors at post_off.
struct TbCb {
  uint64_t post_off;
  uint32_t k;
  uint32_t filler;
};
static_assert(sizeof(TbCb) == 16, "TbCb layout");

constexpr int kTbThreads = 256;

// Transport-block assembly done by the window decoder after a block's last
// iteration (the TB plan's fast path): where the block's payload goes in its
// TB and which CRCs it carries.
struct TbCbMap {
  uint32_t tb;     // transport block index
  uint32_t pos;    // first payload bit of the block within its TB (with CRC24A)
  uint32_t np;     // bits the block contributes to the TB (K - F - 24 * seg)
  uint32_t seg;    // 1: the block ends in a CRC24B
};
struct TbOut {
  int32_t on;
  uint32_t max_iterations;
  const TbInfo* tbs;
  const TbCbMap* map;
  const uint32_t* pow24a;
  const uint32_t* pow24b;
  uint8_t* bits;
  int32_t* tb_ok;
  int32_t* cb_ok;
  uint32_t* tb_iters;
  uint32_t* acc;  // per TB: CRC24A remainder (XOR), unused, max iterations (tb_finalize_kernel reads them)
};
// D^m mod G(D) tables: CRC24B for m < 6144 (a code block), CRC24A up to the
// longest transport block of a plan.
constexpr uint32_t kCrc24APoly = 0xC199BFu;  // lte_turbo.cpp:104-107, D^24 implicit
constexpr uint32_t kCrc24BPoly = 0xC00063u;  // lte_turbo.cpp:108-111

// Transmit side of a code block (GPU encoder and rate matcher).  Its bits
// are filler zeros, then payload bits [cursor, cursor + payload) of the TB
// with its CRC24A, then (seg) a CRC24B over those payload bits.
struct TxCb {
  uint64_t cb_off;   // first code-block bit in the code-block bit buffer (sum of K)
  uint64_t d_off;    // first byte of [d0|d1|d2] (3 (K + 4) bytes)
  uint64_t rm_off;   // first rate-matched bit (sum of E)
  uint32_t tb, k, filler, cursor, payload, seg;
  uint32_t f1, f2;
  uint32_t e, n_valid, valid_off, rv;
  uint32_t start[4];
};

constexpr uint16_t kNullRank = 0xFFFF;
constexpr int kDerateThreads = 1024;

}  // namespace tcdev
