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

// Thread-local last-error message used by the C ABI.
void set_last_error(const std::string& msg);
const char* last_error();

// C ABI status codes (c_api.h) and their exception mapping.
[[noreturn]] void throw_status(int status, const char* what);

}  // namespace turbo_cpp::detail
