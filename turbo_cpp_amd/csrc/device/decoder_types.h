// Plain structures shared by the host plan builder and the decoder kernels.
#pragma once

#include <cstdint>

#include "rx_types.h"

namespace tcdev {

// One code block of a batch.  llr_off: start of [d0|d1|d2] (3*(K+4) floats);
// post_off: start of its K posterior LLRs; work_off: start of its private
// work space (exact decoder only: 16 (K+1) + 4K floats).
struct CbInfo {
  uint64_t llr_off;
  uint64_t post_off;
  uint64_t work_off;
  uint32_t k;
  uint32_t filler;
  uint32_t f1;
  uint32_t f2;
};

// Exact decoder: kExactBlocksPerWave code blocks of one K per wave, listed
// that many indices per group; kNoBlock fills a short group.
constexpr uint32_t kNoBlock = 0xFFFFFFFFu;
constexpr uint32_t kExactBlocksPerWave = 2;

// Window decoder: a job is what one work-group of 64-lane wavefronts (one,
// or two or three for blocks of more than 64 segments) decodes at a time:
// nblk code blocks of the same K, each split into M segments of length S
// (lane = block * M + segment, lanes numbered across the waves).
constexpr int kWaveLanes = 64;
constexpr int kSubWin = 8;       // W: sub-window of the beta recompute
constexpr int kMaxSubWins = 24;  // nW <= 24, so S <= 192
constexpr int kMaxSeg = kSubWin * kMaxSubWins;
constexpr int kExRow = kWaveLanes;  // LDS extrinsic row: one entry per lane
// Per-wave scratch slot, in float2: staged channel LLRs for both halves
// ([t][lane], S <= kMaxSeg), then in floats the boundary metrics
// [lane][half][16], the beta checkpoints [lane][w][8], the posteriors
// [t][lane] plus one dummy slot per lane, the alpha checkpoints [lane][w][8]
// and the extrinsics held between halves [t][lane] plus one dummy slot per
// lane (the last two: exact segment boundaries only, whose fp32 extrinsics
// live here rather than in LDS).
constexpr int kSlotFloat2 = 2 * kMaxSeg * kWaveLanes + 16 * kWaveLanes + 4 * kMaxSubWins * kWaveLanes +
                            (kMaxSeg + 1) * kWaveLanes / 2 + 4 * kMaxSubWins * kWaveLanes +
                            (kMaxSeg + 1) * kWaveLanes / 2;

struct WinJob {
  uint32_t k;
  uint32_t seg;      // S
  uint32_t nseg;     // M
  uint32_t nblk;     // code blocks in this job
  uint32_t warm;     // warm-up steps g (0 when M == 1)
  uint32_t f1, f2;
  uint32_t cb_off;   // first entry in the job's code-block index list
  uint32_t qtab_off; // first entry of the per-K table of Pi(t) for t < S
  uint32_t div_magic;// ceil(2^32 / S) for exact i / S with i < 2^16
  uint32_t pad0, pad1;
};

// Received LLRs de-rate matched inside the window decoder's input staging
// (n_tx > 0).  The decoder input of block c at stream s, index i is
//   (base ? keep[c][s][i] : 0) + sum over transmissions t of the soft de-rate
//   value of rx[t] with redundancy version rv[t] (kBlockRv: the block's own),
// in that order, i.e. what tc_derate_plan_run would leave in keep after one
// run per transmission (the first without accumulation unless base).  keep,
// if set, receives the result in the decoder input layout.
constexpr int kMaxTx = 4;
constexpr uint32_t kBlockRv = 0xFFFFFFFFu;
struct RxIn {
  const float* rx[kMaxTx];
  uint32_t rv[kMaxTx];
  uint32_t n_tx;
  int32_t base;
  float* keep;
  const uint16_t* rank;
  const DerateCb* dcb;
};

// Arguments of the window kernels.
struct WinArgs {
  const WinJob* jobs;
  uint32_t njobs;
  const uint32_t* cb_index;
  const CbInfo* cbs;
  const uint32_t* qtab;
  const uint32_t* crc_pow;  // D^m mod G(D) of CRC24B for m < kCrcPowLen
  const float* llr;
  float* post;
  float2* scratch;
  int32_t* crc_ok;
  uint32_t* iters;
  uint32_t max_iterations;
  int32_t early_stop;
  RxIn rx;  // n_tx == 0: the decoder input is llr
  TbOut tb;  // on == 0: posteriors only
  // Job queue: [0] next job, [1] work-groups that found it empty.  Both read
  // zero at launch; the last work-group to leave puts them back to zero, so the kernel
  // needs no clearing launch (and no memset node under graph capture).
  uint32_t* sched;
};
constexpr int kCrcPowLen = 6144;
// LDS of one window job: the extrinsics held between half-iterations
// [S][64 waves] plus one dummy slot per lane (ex_bytes each: 2 for fp16, 4 for
// fp32) and one CRC partial per lane.  At S = 96 with fp16 this is 12.4 KiB,
// so LDS does not keep a CU below the eight waves (two per SIMD) that the
// register budget allows; fp32 (24.5 KiB) allows six (the probability-domain
// decoder's fp32 variant; the others hold fp32 in the scratch slot,
// tile_job_lds_bytes).
constexpr int job_lds_bytes(int s, int ex_bytes, int waves = 1) {
  return (s * kExRow + kWaveLanes) * ex_bytes * waves + kWaveLanes * 4 * waves;
}
// LDS of a one-wave job whose fp32-held extrinsics sit in the wave's scratch
// slot (the exact segment-parallel decoder; the log window decoder with
// extrinsic_bits = 32): the staging / copy-out tile (ceil(S/2) rows of 64
// floats), which also holds the CRC partials during the iterations.  At
// S = 96 this is 12 KiB, so eight waves (two per SIMD, what the registers
// allow) fit a CU, where 24.8 KiB of fp32 extrinsics in LDS allowed six.
constexpr int tile_job_lds_bytes(int s) {
  return ((s + 1) / 2) * kWaveLanes * 4 > kWaveLanes * 4 ? ((s + 1) / 2) * kWaveLanes * 4 : kWaveLanes * 4;
}
constexpr int kExtrinsicBits = 16;  // window decoders' default held width

}  // namespace tcdev
