/*
 * turbo_oracle.h -- CPU restatement of the Turbo_CPP LTE codec chain.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP decoder
 * and the host codec of turbo_cpp_amd.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it.  The product never links,
 * imports or calls anything here.
 *
 * Every function restates the behaviour of Wiseung/Turbo_CPP
 * src/lte_turbo.cpp (cited per function in turbo_oracle.c) including its
 * non-36.212 conventions: CRC polynomials, RSC feedback 1+D+D^2, tail
 * multiplexing and the decoder's lower-code tail read-out.
 *
 * Pinning (what checks each function besides the compiled reference build in
 * oracle/_ref when it exists):
 *   or_qpp_*            reference known-answer test (188 entries, first/last
 *                       entry), bijection of every entry, closed form.
 *   or_crc_*            append/check round trip, linearity over GF(2),
 *                       golden vectors written from oracle/_ref.
 *   or_segment          36.212 5.1.2 closed form on the reference's spot sizes.
 *   or_encode_cb        golden vectors; brute-force trellis termination
 *                       (both constituent encoders end in state 0).
 *   or_rate_match /
 *   or_derate_match     golden vectors; full-rate rv=0 visits every non-NULL
 *                       circular-buffer entry exactly once (permutation).
 *   or_siso / or_decode golden vectors; brute-force enumeration of all
 *                       2^K paths for K<=12 in float64 (exact APP).
 *   or_derate_match_soft  +-8 inputs equal the reference's de-rate matcher
 *                       (golden vectors, every rv, E above and below the
 *                       circular buffer); additivity of HARQ combining.
 *   or_assemble_tb      equality with the reference's decode_transport_block
 *                       (golden vectors with corrupted code blocks: bits,
 *                       per-block and TB CRC flags).
 *   or_decode_window    the decoder the library runs on the GPU in windowed
 *                       mode; pinned by equality with or_decode when there is
 *                       one window per code block, and by BER.
 *   or_round_half       equality with numpy's float32 -> float16 -> float32
 *                       conversion (IEEE round to nearest even) on normal,
 *                       subnormal, tie and saturating inputs.
 */
#ifndef TURBO_ORACLE_H
#define TURBO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_CRC24A = 0, OR_CRC24B = 1, OR_CRC16 = 2, OR_CRC8 = 3 };

/* QPP interleaver table (188 entries). */
size_t or_qpp_count(void);
void or_qpp_entry(size_t idx, uint32_t* k, uint32_t* f1, uint32_t* f2);
int or_qpp_index(uint32_t k); /* -1 when k is not a table size */
void or_qpp_perm(uint32_t k, uint32_t* perm);

/* CRC.  bits are 0/1 bytes. */
uint32_t or_crc_len(int kind);
void or_crc_append(int kind, const uint8_t* bits, size_t n, uint8_t* out);
int or_crc_check(int kind, const uint8_t* bits, size_t n);

/* Segmentation: fills seg[0..6] = {c, c_plus, c_minus, k_plus, k_minus, filler, l}.
 * blocks (may be NULL) receives the concatenated code blocks (sum of K).
 * Returns 0 on success, -1 for an empty input, -2 if no K+ / K- exists. */
int or_segment(const uint8_t* bits, size_t b, size_t* seg, uint8_t* blocks);

/* Turbo encoder: d0,d1,d2 each K+4 bytes. */
int or_encode_cb(const uint8_t* cb_bits, uint32_t k, uint32_t filler,
                 uint8_t* d0, uint8_t* d1, uint8_t* d2);

/* Rate matching. */
int or_compute_e(size_t r, size_t g_total, size_t c, size_t nl, size_t qm, size_t* e);
void or_rate_match(const uint8_t* d0, const uint8_t* d1, const uint8_t* d2,
                   size_t d, size_t filler, size_t e, uint32_t rv, uint8_t* out);
void or_derate_match(const uint8_t* bits, size_t e, size_t d, size_t filler,
                     uint32_t rv, float* out /* 3*d: d0|d1|d2 */);

/* Soft de-rate matching of E received LLRs (HARQ: accumulate into out). */
void or_derate_match_soft(const float* llr, size_t e, size_t d, size_t filler, uint32_t rv,
                          float* out, int accumulate);

/* Transport-block assembly after decoding: hard decisions, filler strip,
 * CRC24B per block (c > 1), CRC24A per TB, payload of b bits. */
int or_assemble_tb(const float* post, const uint32_t* k, const uint32_t* filler, size_t c, size_t b,
                   uint8_t* bits_out, int* cb_ok, int* tb_ok);

/* One Log-MAP SISO half: sys, apriori K; parity K+4 (first K used). */
void or_siso(const float* sys, const float* parity, const float* apriori, uint32_t k,
             const float* tail_sys4, const float* tail_par4, float* ext);

/* Full turbo decode of one code block (posterior K floats). */
int or_decode_cb(const float* d0, const float* d1, const float* d2, uint32_t k,
                 uint32_t filler, uint32_t max_iterations, int early_stop,
                 float* posterior, int* crc_ok, uint32_t* iterations_used);

/* Windowed decoder used by the GPU batched path (see turbo_oracle.c).
 * ex_bits: 32, or 16 to hold the extrinsics between half-iterations as fp16. */
int or_decode_window(const float* d0, const float* d1, const float* d2, uint32_t k,
                     uint32_t filler, uint32_t max_iterations, int early_stop,
                     uint32_t seg_len, uint32_t warmup, int linear_domain, uint32_t ex_bits,
                     float* posterior, int* crc_ok, uint32_t* iterations_used);

/* Segment-parallel decode with exact segment boundaries (speculation and
 * bit-exact fix-up, see turbo_oracle.c): the schedule of the library's exact
 * GPU decoder.  Equals or_decode_cb bit for bit.  work (may be NULL, 4
 * entries, accumulated): halves, fix-up rounds, chunks recomputed by the
 * slowest segment of each round, chunks of one full pass. */
int or_decode_exact_seg(const float* d0, const float* d1, const float* d2, uint32_t k, uint32_t filler,
                        uint32_t max_iterations, int early_stop, uint32_t seg_len, float* posterior,
                        int* crc_ok, uint32_t* iterations_used, uint64_t* work);

/* x saturated at +-65504 and rounded to the nearest fp16 value (ties to
 * even, fp16 subnormals kept), returned as a float. */
float or_round_half(float x);

/* Exact APP of the upper constituent code by enumerating 2^K inputs (K<=16). */
void or_brute_app(const float* sys, const float* parity, const float* apriori, uint32_t k,
                  const float* tail_sys4, const float* tail_par4, double* app);

#ifdef __cplusplus
}
#endif
#endif
