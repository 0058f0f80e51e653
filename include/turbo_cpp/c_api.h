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
#define TC_DECODER_EXACT 0      /* full-length Log-MAP with the reference's arithmetic */
#define TC_DECODER_WINDOW_LOG 1 /* segment-parallel Log-MAP (max* in the log domain) */
#define TC_DECODER_WINDOW_LIN 2 /* segment-parallel MAP in the probability domain */

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
} tc_options;

/* Geometry the window decoders chose for one K (for tests and reports). */
typedef struct tc_segment_info {
  uint32_t k;
  uint32_t segment_length; /* S, divides K */
  uint32_t segments;       /* M = K / S: lanes per code block */
  uint32_t blocks_per_wave;/* code blocks packed into one 64-lane wavefront */
  uint32_t warmup;         /* effective warm-up steps (0 when M == 1) */
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
