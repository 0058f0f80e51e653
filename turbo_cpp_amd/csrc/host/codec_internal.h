// Internal helpers shared by the host codec, the C ABI and the HIP host code.
#pragma once

#include <cstddef>
#include <cstdint>
#include <span>
#include <string>
#include <vector>

#include "turbo_cpp/lte_turbo.h"

namespace turbo_cpp::detail {

// 36.212 Table 5.1.3-3 (K, f1, f2); same order as the reference's kQppTable
// (lte_turbo.cpp:54-102), so qpp_index values agree.
extern const QppEntry kQppTable[188];

// Index into kQppTable, or -1 when K is not a table size.
int qpp_index(std::size_t k);

// One step of the 8-state RSC constituent encoder (lte_turbo.cpp:333-341):
// state bit 2 is the newest delay, feedback = u ^ s1 ^ s2,
// parity = feedback ^ s0 ^ s2, next = (feedback << 2) | (s2 << 1) | s1.
struct RscTransition {
  std::uint8_t parity;
  std::uint8_t next;
};
constexpr RscTransition rsc(unsigned state, unsigned u) {
  const unsigned s0 = state & 1U, s1 = (state >> 1) & 1U, s2 = (state >> 2) & 1U;
  const unsigned fb = (u ^ s1 ^ s2) & 1U;
  return {static_cast<std::uint8_t>(fb ^ s0 ^ s2), static_cast<std::uint8_t>((fb << 2) | (s2 << 1) | s1)};
}
// Input that moves `state` one step towards state 0 (lte_turbo.cpp:343-353).
constexpr unsigned termination_input(unsigned state) { return ((state >> 1) ^ (state >> 2)) & 1U; }

// Tail beta in the log domain with the reference's -inf/normalisation rules
// (lte_turbo.cpp:357-378, 392-402).
void tail_beta_log(const float tail_sys[3], const float tail_par[3], float beta[8]);

// The decoder's read-out of the tail LLRs (lte_turbo.cpp:758-773), kept as
// the reference has it.  Layout: [upper sys 3][upper par 3][lower sys 3][lower par 3].
void decoder_tail_llrs(const float* d0, const float* d1, const float* d2, std::size_t k,
                       float out[12]);

// E_r of code block r (lte_turbo.cpp:167-185); invalid_argument on bad input.
std::size_t compute_e(std::size_t r, std::size_t g_total, std::size_t c, std::size_t num_layers,
                      std::size_t mod_order);

// Segmentation sizes (C, C+/-, K+/-, F, L) of a transport block of b bits
// including its CRC24A (lte_turbo.cpp:540-604); code_blocks stays empty.
SegmentationResult segmentation_sizes(std::size_t b);

// Soft de-rate matching geometry of one code block (d = K + 4 per stream,
// filler bits F), from the circular buffer of lte_turbo.cpp:195-264:
// rank[s * d + i] is the position of stream s entry i among the non-NULL
// circular-buffer entries (0xFFFF for the NULL filler entries of d(0)/d(1)),
// n_valid their count, start[rv] the rank of the first entry read from k0(rv).
// Received bit k of a transmission lands on rank (start[rv] + k) mod n_valid.
struct DerateGeometry {
  std::vector<std::uint16_t> rank;
  std::uint32_t n_valid = 0;
  std::uint32_t start[4] = {0, 0, 0, 0};
};
DerateGeometry derate_geometry(std::size_t d, std::size_t filler);

// Thread-local last-error message used by the C ABI.
void set_last_error(const std::string& msg);
const char* last_error();

// C ABI status codes (c_api.h) and their exception mapping.
[[noreturn]] void throw_status(int status, const char* what);

}  // namespace turbo_cpp::detail
