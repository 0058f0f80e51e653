/*
 * turbo_cpp/c_api.h -- C ABI of the MI355X turbo decoder (libturbo_cpp_amd.so).
 *
 * Plain pointers and sizes only; no C++ or HIP types cross this boundary.
 * Every entry point returns a tc_status; on failure tc_last_error() holds a
 * thread-local message.  The C++ API in lte_turbo.h is implemented on top of
 * these entry points and maps the status codes back to the reference's
 * exception types:
 *   TC_ERR_INVALID_ARGUMENT -> std::invalid_argument  (lte_turbo.cpp:744-747)
 *   TC_ERR_OUT_OF_RANGE     -> std::out_of_range      (lte_turbo.cpp:502-504)
 *   TC_ERR_RUNTIME          -> std::runtime_error     (lte_turbo_cuda.cu:29-35,
 *                              lte_turbo.cpp:836,861,887)
 *
 * Soft-input layout used by every decoder: per code block r the reference's
 * de-rate-matcher output [d0 | d1 | d2], each d_r = K_r + 4 floats
 * (lte_turbo.cpp:728-733), LLR > 0 meaning bit 1.  Batched calls pack code
 * blocks back to back: block r starts at sum_{s<r} 3*(K_s+4) floats; its
 * posterior (K_r floats, natural order) at sum_{s<r} K_s.
 */
#ifndef TURBO_CPP_C_API_H
#define TURBO_CPP_C_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t tc_status;
#define TC_OK 0
#define TC_ERR_INVALID_ARGUMENT 1
#define TC_ERR_OUT_OF_RANGE 2
#define TC_ERR_RUNTIME 3

/* Decoder algorithm. */
#define TC_DECODER_EXACT 0        /* the reference's Log-MAP arithmetic, segment-parallel with exact
                                     segment boundaries (bit-identical to TC_DECODER_EXACT_SERIAL) */
#define TC_DECODER_WINDOW_LOG 1   /* segment-parallel Log-MAP (max* in the log domain), warm-up +
                                     next-iteration boundaries */
#define TC_DECODER_WINDOW_LIN 2   /* segment-parallel MAP in the probability domain */
#define TC_DECODER_EXACT_SERIAL 3 /* the reference's arithmetic, one full-length recursion per block */

/* Mirrors turbo_cpp::CodeBlockDescriptor (lte_turbo.h:38-45). */
typedef struct tc_cb_desc {
  uint64_t block_index;
  uint64_t k_r;
  uint64_t d_r;
  uint64_t e_r;
  uint64_t filler_count;
  uint64_t qpp_index;
} tc_cb_desc;

typedef struct tc_options {
  uint32_t max_iterations; /* turbo iterations (reference default 8) */
  int32_t early_stop;      /* stop a code block once its CRC24B passes */
  int32_t decoder;         /* TC_DECODER_* */
  uint32_t min_segment;    /* window decoders: lower bound on the segment length, 0 = 32 */
  uint32_t warmup;         /* window decoders: warm-up steps per boundary, 0 = 32 */
  uint32_t waves_per_cu;   /* window decoders: resident waves per CU, 0 = as many as fit */
  uint32_t extrinsic_bits; /* width of the extrinsics held between half-iterations: 16
                              (fp16) or 32 (fp32); 0 = default (window decoders 16, exact
                              decoders 32, which accept only 32) */
  uint32_t max_waves_per_block; /* window decoders (log, lin): a code block may span up to
                              this many waves of one work-group, 1-3 (0 = 1), so the
                              segment length may go down to ceil(K / (64 n)): window
                              lengths 32 and 64 at K = 6144 need 3 and 2 */
} tc_options;

/* Geometry the window decoders chose for one K (for tests and reports). */
typedef struct tc_segment_info {
  uint32_t k;
  uint32_t segment_length; /* S, divides K */
  uint32_t segments;       /* M = K / S: lanes per code block */
  uint32_t blocks_per_wave;/* code blocks per job of waves_per_block waves (one job: 64 x
                              waves_per_block lanes) */
  uint32_t warmup;         /* effective warm-up steps (0 when M == 1) */
  uint32_t extrinsic_bits; /* width of the extrinsics held between
                              half-iterations: 16 (fp16, saturated at
                              +-65504, round to nearest even) or 32 */
  uint32_t waves_per_block;/* waves one code block spans: ceil(M / 64) */
} tc_segment_info;

const char* tc_last_error(void);
tc_status tc_device_count(int32_t* count);
tc_status tc_segment_geometry(uint32_t k, const tc_options* opts, tc_segment_info* out);

/* ---- per-code-block decode of host buffers --------------------------------
 * Replaces turbo_decode_code_block_gpu_stage1 / _stage2 /
 * _stage1_tile8_experimental (lte_turbo.h:129-155; lte_turbo.cpp:815-889 ->
 * lte_turbo_cuda.cu:1391-1899).  n0/n1/n2 must equal desc->d_r.  crc_ok and
 * iterations_used may be NULL. */
tc_status tc_decode_code_block(const tc_cb_desc* desc, const float* d0, size_t n0,
                               const float* d1, size_t n1, const float* d2, size_t n2,
                               const tc_options* opts, float* posterior, int32_t* crc_ok,
                               uint32_t* iterations_used);

/* The same decode without allocating device memory.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (serving loops that must
 * not allocate, harnesses that forbid it).  Same arguments, checks, kernels
 * and results as tc_decode_code_block; the inputs, outputs and work space
 * live in a static device array of 640 KiB that each GPU the library runs on
 * holds until the process ends, sized for the largest code block (K = 6144).
 * Calls on one GPU share that array: calls that overlap in time on one GPU
 * (e.g. from several host threads) race and must be serialised by the
 * caller. */
tc_status tc_decode_code_block_noalloc(const tc_cb_desc* desc, const float* d0, size_t n0,
                                       const float* d1, size_t n1, const float* d2, size_t n2,
                                       const tc_options* opts, float* posterior, int32_t* crc_ok,
                                       uint32_t* iterations_used);

/* ---- batched decode, device-resident buffers ------------------------------
 * New: no counterpart in the reference, which decodes one code block per call
 * (lte_turbo.cpp:941-951 runs those calls on std::async threads).  A plan
 * fixes the K and filler of every code block and owns the device work
 * space; tc_plan_decode runs one persistent kernel over all blocks on the
 * given hipStream_t (NULL = default stream).  llr_dev / post_dev use the
 * packed layout above; crc_ok_dev / iters_dev (n_cb entries) may be NULL. */
typedef struct tc_plan tc_plan;
tc_status tc_plan_create(const uint32_t* k, const uint32_t* filler, size_t n_cb,
                         const tc_options* opts, tc_plan** plan);
void tc_plan_destroy(tc_plan* plan);
size_t tc_plan_llr_floats(const tc_plan* plan);  /* total floats of llr_dev */
size_t tc_plan_post_floats(const tc_plan* plan); /* total floats of post_dev */
tc_status tc_plan_decode(tc_plan* plan, const float* llr_dev, float* post_dev, int32_t* crc_ok_dev,
                         uint32_t* iters_dev, void* hip_stream);

/* Batched decode of host buffers (copies in, decodes, copies out). */
tc_status tc_decode_batch_host(const uint32_t* k, const uint32_t* filler, size_t n_cb,
                               const tc_options* opts, const float* llr, float* post,
                               int32_t* crc_ok, uint32_t* iterations_used);

/* ---- soft de-rate matching with HARQ combining (GPU) ----------------------
 * Float-LLR form of the reference's de-rate matcher (lte_turbo.cpp:280-325,
 * called per code block from lte_turbo.cpp:722-734, which takes hard bits and
 * adds +-8 per bit): received LLR k of a code block (bit 1 > 0) is added onto
 * its circular-buffer entry, walking from k0(rv) and skipping NULL entries
 * (repetitions summed in k order), and every entry lands on its stream
 * position; NULL (filler) positions get 0.  +-8 inputs reproduce the
 * reference's output bit for bit.  With accumulate != 0 the result is added
 * to what the output already holds (HARQ soft combining of retransmissions,
 * e.g. rv 0 then rv 2); positions the transmission does not reach keep their
 * value.
 * Input: the E_r LLRs of every code block back to back.  Output: the decoder
 * plan's input layout ([d0|d1|d2] of K+4 floats per block, blocks back to
 * back), so a run feeds tc_plan_decode on the same stream directly.
 * rv (run): -1 uses the per-block rv given at creation (NULL there = 0),
 * 0..3 applies to every block.  A run allocates nothing; runs on one plan
 * from several streams do not conflict (the plan is read-only). */
typedef struct tc_derate_plan tc_derate_plan;
tc_status tc_derate_plan_create(const uint32_t* k, const uint32_t* filler, const uint64_t* e,
                                const uint32_t* rv, size_t n_cb, tc_derate_plan** plan);
void tc_derate_plan_destroy(tc_derate_plan* plan);
size_t tc_derate_plan_rx_floats(const tc_derate_plan* plan);
size_t tc_derate_plan_llr_floats(const tc_derate_plan* plan);
tc_status tc_derate_plan_run(tc_derate_plan* plan, const float* rx_dev, int32_t rv, int32_t accumulate,
                             float* llr_dev, void* hip_stream);
/* One code block from host buffers (out: 3 (K+4) floats; read first when
 * accumulate != 0). */
tc_status tc_derate_match_soft(const float* rx, uint64_t e, uint32_t k, uint32_t filler, uint32_t rv,
                               int32_t accumulate, float* out);

/* Decode straight from received LLRs: the window decoders de-rate match
 * (tc_derate_plan semantics) inside their input staging, so no decoder input
 * is written before the decode.  `derate` must cover the same blocks (K,
 * filler) as `plan`.  The decoder input is, per position,
 *   (accumulate ? keep : 0) + de-rate(rx_dev[0], rv[0]) + ... + de-rate(rx_dev[n_tx-1], rv[n_tx-1])
 * (HARQ combining of up to 4 transmissions; rv NULL or -1 = the blocks' own),
 * bit-identical to running tc_derate_plan_run per transmission into keep and
 * decoding keep.  keep (may be NULL unless accumulate) receives that input.
 * The exact decoder has no fused staging: it de-rate matches into keep
 * (required) and decodes it. */
tc_status tc_plan_decode_rx(tc_plan* plan, const tc_derate_plan* derate, uint32_t n_tx,
                            const float* const* rx_dev, const int32_t* rv, int32_t accumulate,
                            float* keep_dev, float* post_dev, int32_t* crc_ok_dev,
                            uint32_t* iters_dev, void* hip_stream);

/* ---- batched transport-block decoding (GPU) --------------------------------
 * The reference's decode_transport_block (lte_turbo.cpp:891-980) for many
 * transport blocks at once, from received LLRs: soft de-rate matching (with
 * HARQ combining), turbo decoding of every code block in one batched launch,
 * and TB assembly on the device (hard decisions, filler strip, CRC24B per code
 * block when C > 1, CRC24A per TB, iterations = max over the TB's blocks).
 * Segmentation and E_r follow encode_transport_block (lte_turbo.cpp:540-604,
 * 167-185, 690-720) from each descriptor.
 * Device buffers of tc_tb_plan_decode: rx = sum E_r floats (blocks of all TBs
 * back to back, TB order); bits = sum B bytes (0/1, TB payloads back to back);
 * tb_crc_ok / iterations: one per TB; cb_crc_ok: one per code block;
 * posterior: sum K_r floats (NULL: kept inside the plan).  Any output may be
 * NULL.  The plan holds the decoder input between calls: accumulate != 0 adds
 * this transmission onto it (HARQ), rv as in tc_derate_plan_run; rx NULL
 * decodes the held input as it stands (e.g. after tc_tb_plan_soft_combine).
 * A call allocates nothing; calls on one plan must not overlap (shared
 * buffers). */
typedef struct {
  uint64_t transport_block_bits; /* B: payload bits without the CRC24A */
  uint64_t g_total;              /* coded bits for the TB */
  uint32_t rv;
  uint32_t num_layers;
  uint32_t mod_order;
  /* 1: no CRC24A and no segmentation -- one code block of K = B + 24 bits
   * whose last 24 bits are its CRC24B, the layout of the reference's
   * segmented code blocks (early stop applies); for code-block-level batches.
   * tb_crc_ok then reports that CRC24B. */
  uint32_t code_block_crc;
} tc_tb_desc;
typedef struct tc_tb_plan tc_tb_plan;
tc_status tc_tb_plan_create(const tc_tb_desc* tbs, size_t n_tb, const tc_options* opts,
                            tc_tb_plan** plan);
void tc_tb_plan_destroy(tc_tb_plan* plan);
/* sizes[0] = code blocks, [1] = rx floats, [2] = payload bits, [3] = posterior floats. */
tc_status tc_tb_plan_sizes(const tc_tb_plan* plan, uint64_t sizes[4]);
/* K_r, F_r and E_r of every code block (each array may be NULL). */
tc_status tc_tb_plan_code_blocks(const tc_tb_plan* plan, uint32_t* k, uint32_t* filler, uint64_t* e);
tc_status tc_tb_plan_decode(tc_tb_plan* plan, const float* rx_dev, int32_t rv, int32_t accumulate,
                            uint8_t* bits_dev, int32_t* tb_crc_ok_dev, int32_t* cb_crc_ok_dev,
                            uint32_t* iterations_dev, float* posterior_dev, void* hip_stream);
/* The transmit side of the same plan, bit-exact with encode_transport_block
 * (lte_turbo.cpp:690-720): CRC24A per TB, segmentation with filler and
 * CRC24B, turbo encoder with the reference's tail multiplexing, rate
 * matching.  payload = sum B bytes (0/1, TB order); rm_bits = sum E_r bytes
 * (the reference's rate_matched_bits of every block back to back); cb_bits
 * (may be NULL) = sum K_r bytes (the reference's cb_bits).  rv as in
 * tc_derate_plan_run. */
tc_status tc_tb_plan_encode(tc_tb_plan* plan, const uint8_t* payload_dev, int32_t rv,
                            uint8_t* rm_bits_dev, uint8_t* cb_bits_dev, void* hip_stream);
tc_status tc_encode_transport_blocks_host(const tc_tb_desc* tbs, size_t n_tb, const uint8_t* payload,
                                          uint8_t* rm_bits, uint8_t* cb_bits);
/* Soft de-rate matching of one transmission into the plan's held decoder
 * input without decoding (HARQ: combine rv 0 and rv 2, then decode with
 * rx NULL). */
tc_status tc_tb_plan_soft_combine(tc_tb_plan* plan, const float* rx_dev, int32_t rv, int32_t accumulate,
                                  void* hip_stream);
/* hold != 0 (the default): decodes from received LLRs leave the combined
 * decoder input in the plan for later accumulation.  hold == 0 is for
 * receivers that do not combine retransmissions: accumulate != 0 is then
 * rejected, and with TC_STAGING_FUSED the decode skips writing the decoder
 * input (one pass over it fewer; the default staging writes it either way).
 * rx NULL and tc_tb_plan_soft_combine keep working either way. */
tc_status tc_tb_plan_set_hold(tc_tb_plan* plan, int32_t hold);
/* Where the window decoders de-rate match received transmissions (the exact
 * decoder always uses the second form):
 *   TC_STAGING_FUSED    - inside the decoder's input staging: one launch, the
 *                         decoder gathers each LLR through the sub-block
 *                         interleaver;
 *   TC_STAGING_SEPARATE - one de-rate kernel per transmission into the plan's
 *                         decoder input (coalesced, LDS-staged), then the
 *                         decode of that input;
 *   TC_STAGING_AUTO     - the default, the fused form (measured faster on
 *                         every workload tried: one full-rate transmission,
 *                         punctured ones, HARQ pairs).
 * Both give the same results bit for bit.  No counterpart in the reference
 * (whose de-rate matcher is a host function, lte_turbo.cpp:280-325). */
#define TC_STAGING_AUTO 0
#define TC_STAGING_FUSED 1
#define TC_STAGING_SEPARATE 2
tc_status tc_tb_plan_set_staging(tc_tb_plan* plan, int32_t mode);
/* HARQ in one call: the held decoder input becomes the combination of n_tx
 * (1..4) transmissions (rx_dev[t] with rv[t], rv NULL = the descriptors'),
 * then decode and assembly as tc_tb_plan_decode. */
tc_status tc_tb_plan_decode_multi(tc_tb_plan* plan, uint32_t n_tx, const float* const* rx_dev,
                                  const int32_t* rv, uint8_t* bits_dev, int32_t* tb_crc_ok_dev,
                                  int32_t* cb_crc_ok_dev, uint32_t* iterations_dev, float* posterior_dev,
                                  void* hip_stream);
/* Same from host buffers (one transmission, no HARQ). */
tc_status tc_decode_transport_blocks_host(const tc_tb_desc* tbs, size_t n_tb, const tc_options* opts,
                                          const float* rx, uint8_t* bits, int32_t* tb_crc_ok,
                                          int32_t* cb_crc_ok, uint32_t* iterations, float* posterior);

/* ---- channel front end (GPU) -----------------------------------------------
 * No counterpart in the reference (it has no channel code): Gray QAM mapping
 * of 3GPP TS 36.211 7.1 (qm 2/4/6, unit average energy; bits b0,b2,b4 on I,
 * b1,b3,b5 on Q), AWGN or i.i.d. Rayleigh fading known at the receiver
 * (h ~ CN(0,1) per symbol), noise CN(0, n0), and the max-log demapper
 * LLR = (min_{b=0} |y-hs|^2 - min_{b=1} |y-hs|^2) / n0 (bit 1 > 0).
 * Segments lie back to back: bits (and LLRs) sum n_bits, symbols sum
 * n_bits/qm; y and h hold (re, im) per symbol.  tc_channel_llr maps, fades,
 * adds noise from (seed, symbol index) and demaps in one pass (y, h may be
 * NULL); tc_qam_demap demaps given y (h NULL = 1).  Both only enqueue. */
typedef struct {
  uint64_t n_bits;   /* multiple of qm */
  uint32_t qm;       /* 2, 4 or 6 */
  uint32_t rayleigh; /* 0: AWGN, 1: Rayleigh */
  float n0;          /* noise variance per complex symbol */
  uint32_t pad;
} tc_channel_seg;
typedef struct tc_channel_plan tc_channel_plan;
tc_status tc_channel_plan_create(const tc_channel_seg* segs, size_t n_seg, tc_channel_plan** plan);
void tc_channel_plan_destroy(tc_channel_plan* plan);
/* sizes[0] = bits, sizes[1] = symbols */
tc_status tc_channel_plan_sizes(const tc_channel_plan* plan, uint64_t sizes[2]);
tc_status tc_channel_llr(tc_channel_plan* plan, uint64_t seed, const uint8_t* bits_dev, float* llr_dev,
                         float* y_dev, float* h_dev, void* hip_stream);
tc_status tc_qam_demap(tc_channel_plan* plan, const float* y_dev, const float* h_dev, float* llr_dev,
                       void* hip_stream);

/* ---- differentiable decoder (GPU) ------------------------------------------
 * No counterpart in the reference: a sliding-window Log-MAP turbo decoder
 * whose half-iteration h multiplies its extrinsic output by a learnable scale
 * w[h] wherever it is used (next half's a-priori input; the last two halves
 * also in the posterior post = d0 + w[2I-2] e1 + w[2I-1] e2), with the
 * hand-written adjoint of its alpha/beta/LLR recursions.  With w = 1 and one
 * segment per block (K <= min_segment) it is the reference decoder
 * (lte_turbo.cpp:736-813, no early stop).  Segments of S steps
 * (S = smallest divisor of K >= max(min_segment, K/64)) start their alpha and
 * beta recursions `warmup` (<= S) steps outside the segment from zero metrics.
 * fp64 != 0 runs every buffer in double (for finite-difference checks),
 * else float.  Layouts: llr = [d0|d1|d2] of K+4 per block, post K per block,
 * scales 2I, saved = the forward's extrinsics and its alpha/beta checkpoints
 * (every 8 steps of each segment's window, every half-iteration) for the
 * backward, which recomputes the metrics between checkpoints (sizes from
 * tc_diff_plan_sizes: [0] llr, [1] post, [2] scales, [3] saved elements).
 * grad_llr is overwritten (same layout as llr); grad_scales gets 2I values.
 * Calls enqueue on the given stream and allocate nothing; calls on one plan
 * must not overlap (shared work space). */
typedef struct tc_diff_plan tc_diff_plan;
tc_status tc_diff_plan_create(const uint32_t* k, size_t n_cb, uint32_t iterations,
                              uint32_t min_segment, uint32_t warmup, int32_t fp64,
                              tc_diff_plan** plan);
void tc_diff_plan_destroy(tc_diff_plan* plan);
tc_status tc_diff_plan_sizes(const tc_diff_plan* plan, uint64_t sizes[4]);
tc_status tc_diff_forward(tc_diff_plan* plan, const void* llr, const void* scales, void* post,
                          void* saved, void* hip_stream);
tc_status tc_diff_backward(tc_diff_plan* plan, const void* llr, const void* scales,
                           const void* saved, const void* grad_post, void* grad_llr,
                           void* grad_scales, void* hip_stream);

/* ---- host codec (the reference's CPU chain, bit-exact) ---------------------
 * CRC kinds: 0 = CRC24A, 1 = CRC24B, 2 = CRC16, 3 = CRC8 with the
 * reference's generator polynomials (lte_turbo.cpp:104-116). */
tc_status tc_crc_append(int32_t kind, const uint8_t* bits, size_t n, uint8_t* out);
tc_status tc_crc_check(int32_t kind, const uint8_t* bits, size_t n, int32_t* ok);
/* seg[0..6] = {c, c_plus, c_minus, k_plus, k_minus, filler_bits, l}; blocks
 * (may be NULL) receives all code blocks back to back (lte_turbo.cpp:540-604). */
tc_status tc_segment(const uint8_t* bits, size_t b, uint64_t* seg, uint8_t* blocks);
tc_status tc_qpp_entry(uint32_t k, uint32_t* f1, uint32_t* f2, uint32_t* index);
/* d0/d1/d2: K+4 bytes each (lte_turbo.cpp:624-688). */
tc_status tc_encode_code_block(const uint8_t* cb_bits, uint32_t k, uint32_t filler, uint8_t* d0,
                               uint8_t* d1, uint8_t* d2);
tc_status tc_compute_e(uint64_t r, uint64_t g_total, uint64_t c, uint32_t num_layers,
                       uint32_t mod_order, uint64_t* e);
tc_status tc_rate_match(const uint8_t* d0, const uint8_t* d1, const uint8_t* d2, uint64_t d,
                        uint64_t filler, uint64_t e, uint32_t rv, uint8_t* out);
/* out: 3*d floats d0|d1|d2 (lte_turbo.cpp:722-734). */
tc_status tc_derate_match(const uint8_t* bits, uint64_t e, uint64_t d, uint64_t filler,
                          uint32_t rv, float* out);
/* Host Log-MAP decode (turbo_decode_code_block_cpu, lte_turbo.cpp:736-813). */
tc_status tc_decode_code_block_cpu(const tc_cb_desc* desc, const float* d0, const float* d1,
                                   const float* d2, uint32_t max_iterations, int32_t early_stop,
                                   float* posterior, int32_t* crc_ok, uint32_t* iterations_used);

#ifdef __cplusplus
}
#endif
#endif
