// Plain structures of the differentiable decoder (sliding-window Log-MAP with
// per-half-iteration extrinsic scales, forward and adjoint kernels).
#pragma once

#include <cstdint>

namespace tcdev {

// One wave of work: nblk code blocks of the same K, each split into M
// segments of S steps (lane = block * M + segment); warm-up g <= S steps.
struct DiffJob {
  uint32_t k, seg, nseg, nblk, warm, f1, f2, cb_first;
};

// A code block: offsets (in elements) of its channel LLRs [d0|d1|d2]
// (3(K+4)), its posteriors (K), its saved extrinsics (2I x K) and its two
// extrinsic-adjoint buffers (2 x K).
struct DiffCb {
  uint64_t llr_off, post_off, save_off, ebar_off;
};

template <typename R>
struct DiffArgs {
  const DiffJob* jobs;
  uint32_t njobs;
  uint32_t halves;  // 2I
  const DiffCb* cbs;
  const R* llr;
  const R* scales;  // [2I]
  R* post;
  R* saved;
  // backward only
  const R* grad_post;
  R* grad_llr;
  R* ebar;
  R* dscale_part;  // [njobs][2I]
  R* scratch;      // per lane, slot_elems each
  uint32_t slot_elems;
  // the forward's alpha / beta checkpoints, kept for the adjoint: job jb's
  // half h at ckpt + (jb * halves + h) * ckpt_job_half, per lane
  // diff_ck_lane_elems elements in [element][lane] order
  R* ckpt;
  uint64_t ckpt_job_half;
  uint32_t max_seg, max_warm;
  // 1: every job's blocks fit their LDS copies (diff_lds_elems) in the
  // dynamic LDS of its workgroup (staging and the adjoint combine then go
  // through LDS with coalesced global accesses, and the recursion passes
  // read their inputs from LDS)
  int32_t use_lds;
};

// Checkpoints of one lane and half: alpha at the start of each of its chunks
// (at most (g + S + 14) / 8 of them) and beta at the top of each chunk (at
// most (g + S + 7) / 8 + 1 slots), 8 states each.
__host__ __device__ constexpr uint32_t diff_ck_alpha(uint32_t s, uint32_t g) { return (g + s + 14) / 8; }
__host__ __device__ constexpr uint32_t diff_ck_beta(uint32_t s, uint32_t g) { return (g + s + 7) / 8 + 1; }
__host__ __device__ constexpr uint32_t diff_ck_lane_elems(uint32_t s, uint32_t g) {
  return 8 * (diff_ck_alpha(s, g) + diff_ck_beta(s, g));
}

// LDS copies per block: kDiffLdsArrays arrays of diff_lds_stride(K)
// elements, element i at i + i / 32 (one pad per 32 elements, so lanes that
// read positions a multiple of 32 apart use different banks).
constexpr uint32_t kDiffLdsArrays = 4;
__host__ __device__ constexpr uint32_t diff_lds_stride(uint32_t k) { return k + (k >> 5) + 1; }
// Staged inputs of one half per block, read by the recursion passes: pairs
// (sys + la, parity) by position p = t + g in the half's order (the lower
// half's sys + la already gathered through the interleaver), g zero pairs
// either side of the block, pair p at p + p / 32; the blocks' pair arrays
// follow arrays 0 and 1 (they overlap arrays 2 and 3, which only the
// adjoint's combine uses, after the passes).
__host__ __device__ constexpr uint32_t diff_stage_pairs(uint32_t k, uint32_t g) { return diff_lds_stride(k + 2 * g); }
// LDS elements of a job of nblk blocks of size k with warm-up g
__host__ __device__ constexpr uint32_t diff_lds_elems(uint32_t k, uint32_t g, uint32_t nblk) {
  return nblk * (2 * diff_lds_stride(k) + 2 * diff_stage_pairs(k, g));
}

}  // namespace tcdev
