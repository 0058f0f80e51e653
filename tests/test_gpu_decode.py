"""HIP decoders against the oracle, the golden vectors of the reference and
the reference's BER.  Marked gpu: they need an MI355X.

Tolerances (stated per test): exact decoder posterior LLRs within 1e-4
absolute of the reference (lte_turbo.cpp:736-813) with identical hard
decisions where |LLR| >= 1e-3; window decoders against the oracle's
restatement of the same algorithm (or_decode_window) and, for BER, inside
the reference decoder's 95% Clopper-Pearson interval.
"""
import os

import numpy as np
import pytest

import oracle as O
import turbo_cpp_amd as T
from tests.helpers import assert_llr_parity, bpsk_awgn_llr, clopper_pearson, golden, random_blocks

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _split(llr, k):
    d = k + 4
    return llr[:d], llr[d:2 * d], llr[2 * d:]


GOLDEN_DECODES = [(40, 100, 6), (64, 0, 8), (544, 50, 6), (1024, 50, 8), (6144, 100, 8), (6144, 50, 8),
                  (6144, 25, 8)]


@pytest.mark.parametrize("k,eb,it", GOLDEN_DECODES)
def test_exact_decoder_matches_reference_golden(k, eb, it):
    g = golden()
    tag = f"dec_k{k}_eb{eb}_it{it}"
    _say(tag)
    llr = g[tag + "_llr"]
    post, ok, used = T.turbo_decode_code_block_gpu_stage1(T.CodeBlockDescriptor(0, k, k + 4), llr[0], llr[1],
                                                          llr[2], it, False)
    assert used == it and not ok
    assert_llr_parity(post, g[tag + "_post"], atol=1e-4)


def test_exact_decoder_early_stop_matches_reference():
    g = golden()
    llr = g["es_k1024_llr"]
    post, ok, used = T.turbo_decode_code_block_gpu_stage1(T.CodeBlockDescriptor(0, 1024, 1028), *llr, 8, True)
    assert [int(ok), used] == list(g["es_k1024_meta"])
    assert_llr_parity(post, g["es_k1024_post"], atol=1e-4)


def test_exact_batched_equals_single_calls():
    """Uneven K in one launch: every block equals its single-block decode."""
    rng = np.random.default_rng(11)
    ks = [40, 6144, 104, 1024, 544, 40]
    llrs, singles = [], []
    for k in ks:
        cb = rng.integers(0, 2, k).astype(np.uint8)
        l = np.concatenate(bpsk_awgn_llr(O.encode_cb(cb), k, 0.5, rng))
        llrs.append(l)
        singles.append(T.turbo_decode_code_block_gpu_stage1(T.CodeBlockDescriptor(0, k, k + 4), *_split(l, k), 6,
                                                            False)[0])
    post, _, its = T.decode_batch_host(ks, np.concatenate(llrs), T.DECODER_EXACT, 6)
    off = 0
    for k, s in zip(ks, singles):
        assert np.array_equal(post[off:off + k].view(np.uint32), s.view(np.uint32))
        off += k
    assert list(its) == [6] * len(ks)


# -------------------------------------- exact decoders: segmented == serial
def _table_ks():
    return [int(e[0]) for e in T.qpp_interleaver_table()]


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def test_exact_serial_matches_reference_golden():
    """The serial form (one recursion per block, behind the tile8 entry point)
    against the reference's golden posteriors: 1e-4 absolute, hard decisions
    equal where |LLR| >= 1e-3."""
    g = golden()
    for k, eb, it in GOLDEN_DECODES:
        tag = f"dec_k{k}_eb{eb}_it{it}"
        _say("serial", tag)
        llr = g[tag + "_llr"]
        post, _, used = T.turbo_decode_code_block_gpu_stage1_tile8_experimental(T.CodeBlockDescriptor(0, k, k + 4),
                                                                                llr[0], llr[1], llr[2], it, False)
        assert used == it
        assert_llr_parity(post, g[tag + "_post"], atol=1e-4)


def test_exact_segmented_equals_serial_all_sizes():
    """TC_DECODER_EXACT (segment-parallel recursions with boundaries fixed up
    until every segment meets its checkpoints bit for bit) and
    TC_DECODER_EXACT_SERIAL (one full-length recursion per block) give the
    same posteriors bit for bit: one block of each of the 188 QPP sizes at
    Eb/N0 0..2 dB, 8 iterations, one batch each."""
    ks = _table_ks()
    rng = np.random.default_rng(188)
    llrs = []
    for k in ks:
        cb = rng.integers(0, 2, k).astype(np.uint8)
        llrs.append(np.concatenate(bpsk_awgn_llr(O.encode_cb(cb), k, float(rng.uniform(0.0, 2.0)), rng)))
    llr = np.concatenate(llrs)
    _say("exact segmented vs serial: 188 sizes")
    seg, _, its = T.decode_batch_host(ks, llr, T.DECODER_EXACT, 8)
    ser, _, its2 = T.decode_batch_host(ks, llr, T.DECODER_EXACT_SERIAL, 8)
    assert _bits_equal(seg, ser)
    assert list(its) == [8] * len(ks) and list(its2) == [8] * len(ks)


@pytest.mark.parametrize("ebn0", [-0.5, 0.0, 0.5, 1.0])
def test_exact_segmented_equals_serial_k6144(ebn0):
    """64 blocks of K=6144 (64 segments of 96 steps each) from below the
    decoding threshold to above it, where boundary fix-ups need the most
    rounds: bit for bit the serial decoder's posteriors."""
    bits, llr = random_blocks(64, 6144, ebn0, 6144 + int(100 * ebn0), O.encode_cb)
    _say(f"exact segmented vs serial K=6144 ebn0={ebn0}")
    seg, _, _ = T.decode_batch_host([6144] * 64, llr, T.DECODER_EXACT, 8)
    ser, _, _ = T.decode_batch_host([6144] * 64, llr, T.DECODER_EXACT_SERIAL, 8)
    assert _bits_equal(seg, ser)


def test_exact_segmented_equals_serial_early_stop():
    """With CRC24B early stop on mixed K (several blocks per wave for K <= 64,
    fillers): posteriors, CRC verdicts and stopping iterations equal the
    serial decoder's, and the verdicts the CPU reference decoder's."""
    rng = np.random.default_rng(96)
    table = _table_ks()
    ks = [6144] * 8 + [40] * 9 + [int(x) for x in rng.choice(table, 47)]
    fs = [8 if i % 5 == 0 and k > 64 else 0 for i, k in enumerate(ks)]
    eb = [float(x) for x in rng.uniform(-0.5, 2.5, len(ks))]
    _, llrs = _crc_blocks(ks, fs, eb, 96)
    llr = np.concatenate(llrs)
    _say("exact segmented vs serial, early stop:", len(ks), "blocks")
    seg, c1, i1 = T.decode_batch_host(ks, llr, T.DECODER_EXACT, 8, True, fillers=fs)
    ser, c2, i2 = T.decode_batch_host(ks, llr, T.DECODER_EXACT_SERIAL, 8, True, fillers=fs)
    assert _bits_equal(seg, ser)
    assert np.array_equal(c1, c2) and np.array_equal(i1, i2)
    assert 0 < c1.sum() < len(ks)
    for j in range(0, len(ks), 9):  # verdicts of the reference decoder on the CPU (a sample)
        k, f, l = ks[j], fs[j], llrs[j]
        _, ok, used = T.turbo_decode_code_block_cpu(T.CodeBlockDescriptor(0, k, k + 4, filler_count=f), *_split(l, k),
                                                    8, True)
        assert (bool(c1[j]), int(i1[j])) == (ok, used), j


# ------------------------------------------------------------ window decoders
def _window_bar(geo, ref):
    """(atol, rtol, hard_tol) of window decoder vs oracle.  exp2/log2
    rounding (v_exp_f32/v_log_f32 against libm) alone leaves ~1e-4 relative;
    with extrinsics held as fp16, such a difference can move a held value
    across an fp16 rounding boundary, i.e. by one fp16 step (2^-11 of the
    held value, not of the bit's own LLR), which the following
    half-iterations spread to neighbouring bits: the absolute bar scales
    with the largest LLR of the block."""
    assert geo["extrinsic_bits"] in (16, 32)
    if geo["extrinsic_bits"] == 32:
        return 1e-3, 1e-4, 1e-3
    a = 1e-3 * max(1.0, float(np.abs(ref).max()))
    return a, 2e-3, a


WINDOW_CASES = [(6144, 0.5), (6144, 1.0), (1024, 0.5), (40, 0.5), (1952, 0.75), (6080, 0.25)]


@pytest.mark.parametrize("decoder", [T.DECODER_WINDOW_LOG, T.DECODER_WINDOW_LIN])
@pytest.mark.parametrize("k,ebn0", WINDOW_CASES)
def test_window_decoder_matches_oracle(k, ebn0, decoder):
    """Same algorithm on both sides (segments, warm-up, NII, arithmetic);
    they differ only in exp2/log2 rounding (v_exp_f32/v_log_f32 vs libm).
    Bar (_window_bar), extrinsics held as fp16: 1e-3 max|LLR| + 2e-3 |LLR|,
    hard decisions equal where |LLR| >= 1e-3 max|LLR|."""
    geo = T.segment_geometry(k, decoder)
    _say(f"k={k} ebn0={ebn0} decoder={decoder} geometry={geo}")
    rng = np.random.default_rng(k + int(100 * ebn0))
    bits, llr = random_blocks(3, k, ebn0, int(rng.integers(1 << 30)), O.encode_cb)
    post, _, its = T.decode_batch_host([k] * 3, llr, decoder, 8)
    for i in range(3):
        ref = O.decode_window(*_split(llr[i], k), k, geo["segment_length"], geo["warmup"],
                              decoder == T.DECODER_WINDOW_LIN, 0, 8, False, geo["extrinsic_bits"])[0]
        atol, rtol, hard = _window_bar(geo, ref)
        assert_llr_parity(post[i * k:(i + 1) * k], ref, atol=atol, rtol=rtol, hard_tol=hard)
    assert list(its) == [8, 8, 8]


@pytest.mark.parametrize("early_stop", [False, True])
@pytest.mark.parametrize("decoder", [T.DECODER_WINDOW_LOG, T.DECODER_WINDOW_LIN])
def test_window_fp32_held_all_sizes_match_oracle(decoder, early_stop):
    """Extrinsics held in fp32 between half-iterations (extrinsic_bits=32):
    one block of each of the 188 QPP sizes (Eb/N0 0..2.5 dB) against the
    oracle's restatement with 32-bit held extrinsics, at the 32-bit bar of
    _window_bar: 1e-3 absolute + 1e-4 relative, hard decisions equal where
    |LLR| >= 1e-3.  With early stop (CRC24B blocks) the stopping iteration and
    the verdict equal the oracle's."""
    import concurrent.futures as cf
    ks = _table_ks()
    rng = np.random.default_rng(32 + decoder + 2 * early_stop)
    eb = [float(x) for x in rng.uniform(0.0, 2.5, len(ks))]
    _, llrs = _crc_blocks(ks, [0] * len(ks), eb, 320 + decoder)
    _say(f"fp32-held window decoder={decoder} es={early_stop}: 188 sizes")
    post, crc, its = T.decode_batch_host(ks, np.concatenate(llrs), decoder, 8, early_stop, extrinsic_bits=32)
    lin = decoder == T.DECODER_WINDOW_LIN

    def ref(i):
        k = ks[i]
        geo = T.segment_geometry(k, decoder, extrinsic_bits=32)
        assert geo["extrinsic_bits"] == 32
        return O.decode_window(*_split(llrs[i], k), k, geo["segment_length"], geo["warmup"], lin, 0, 8, early_stop, 32)

    with cf.ThreadPoolExecutor(8) as ex:
        refs = list(ex.map(ref, range(len(ks))))
    off = 0
    for i, k in enumerate(ks):
        r, ok, used = refs[i]
        assert_llr_parity(post[off:off + k], r, atol=1e-3, rtol=1e-4, hard_tol=1e-3)
        if early_stop:
            assert (bool(crc[i]), int(its[i])) == (ok, used), (i, k)
        off += k


@pytest.mark.parametrize("decoder", [T.DECODER_WINDOW_LOG, T.DECODER_WINDOW_LIN])
def test_window_single_segment_matches_reference(decoder):
    """K=40 uses one segment per block (M=1): no boundaries, so the window
    decoder is a full-length MAP decoder.  It matches the oracle's restatement
    with fp16-held extrinsics at the window bar, and the reference's
    posteriors to the fp16 rounding of the held extrinsics (the oracle with
    32-bit extrinsics matches the reference to 2e-4: test_oracle)."""
    g = golden()
    llr = g["dec_k40_eb100_it6_llr"].reshape(-1)
    geo = T.segment_geometry(40, decoder)
    assert geo["segments"] == 1 and geo["extrinsic_bits"] == 16
    post, _, _ = T.decode_batch_host([40], llr, decoder, 6)
    ref = O.decode_window(*_split(llr, 40), 40, 40, 0, decoder == T.DECODER_WINDOW_LIN, 0, 6, False, 16)[0]
    atol, rtol, hard = _window_bar(geo, ref)
    assert_llr_parity(post, ref, atol=atol, rtol=rtol, hard_tol=hard)
    gold = g["dec_k40_eb100_it6_post"]
    assert_llr_parity(post, gold, atol=2e-3 * np.abs(gold).max(), rtol=2e-3, hard_tol=2e-3 * np.abs(gold).max())


def test_window_batched_mixed_sizes_equal_single():
    """Blocks of different K packed into shared jobs give the same bits as
    decoding each alone (the packing does not change any lane's arithmetic)."""
    rng = np.random.default_rng(5)
    ks = [6144, 40, 40, 1024, 6144, 544, 3200, 40]
    llrs = []
    for k in ks:
        cb = rng.integers(0, 2, k).astype(np.uint8)
        llrs.append(np.concatenate(bpsk_awgn_llr(O.encode_cb(cb), k, 0.5, rng)))
    post, _, _ = T.decode_batch_host(ks, np.concatenate(llrs), T.DECODER_WINDOW_LOG, 8)
    off = 0
    for k, l in zip(ks, llrs):
        single, _, _ = T.decode_batch_host([k], l, T.DECODER_WINDOW_LOG, 8)
        assert np.array_equal(post[off:off + k].view(np.uint32), single.view(np.uint32))
        off += k


def test_window_decoder_is_deterministic():
    bits, llr = random_blocks(70, 6144, 0.5, 99, O.encode_cb)
    a, _, _ = T.decode_batch_host([6144] * 70, llr, T.DECODER_WINDOW_LOG, 8)
    b, _, _ = T.decode_batch_host([6144] * 70, llr, T.DECODER_WINDOW_LOG, 8)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _reference_posteriors(llr, k, iters):
    """The reference's CPU decoder (oracle/_ref, its own lte_turbo.cpp) on all
    host threads, or the oracle's restatement (bit-exact with it) where that
    build is absent."""
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    if O.ref_available():
        return O.ref_decode_many(llr, k, iters, threads)
    import concurrent.futures as cf
    with cf.ThreadPoolExecutor(threads) as ex:
        return np.stack(list(ex.map(lambda l: O.decode_cb(*_split(l, k), k, 0, iters, False)[0], llr)))


# Eb/N0 (dB) and blocks of K=6144: enough for >= 100 bit errors of the reference
# decoder where it makes them at feasible sample sizes (its failures cluster in
# a few blocks, ~1 in 100 blocks at 0.75-1.0 dB); at 1.25 and 1.5 dB the
# decoders made no error in 585 blocks each (3.6 Mbit) of the benchmark's sweep,
# so >= 100 errors would need a sample beyond a test's reach: those two points
# are checked against the interval of the reference's (zero or few) errors on
# 512 blocks and reported as unresolvable.
BER_POINTS = [(0.0, 64), (0.25, 64), (0.5, 256), (0.75, 8192), (1.0, 8192), (1.25, 512), (1.5, 512)]


@pytest.mark.parametrize("ebn0,n", BER_POINTS)
def test_ber_inside_reference_interval(ebn0, n):
    """BER and BLER of the exact decoder and of both window decoders (the
    headline geometry: fp16-held extrinsics) lie inside the 95%
    Clopper-Pearson intervals of the reference decoder's bit and block
    errors on the same noisy K=6144 blocks, 8 iterations (BASELINE config 2's
    Eb/N0 sweep, 0 to 1.5 dB in 0.25 dB steps)."""
    k = 6144
    bits, llr = random_blocks(n, k, ebn0, 1234 + int(ebn0 * 100), O.encode_cb)
    _say(f"ebn0={ebn0}: {n} blocks, reference decoder on the CPU")
    ref = _reference_posteriors(llr, k, 8)
    ref_err = ((ref >= 0) != bits).sum(1)
    lo_b, hi_b = clopper_pearson(int(ref_err.sum()), n * k)
    lo_f, hi_f = clopper_pearson(int((ref_err > 0).sum()), n)
    resolvable = int(ref_err.sum()) >= 100
    _say(f"ebn0={ebn0} reference bit errors={int(ref_err.sum())} BER={ref_err.sum() / (n * k):.3g} "
         f"BLER={(ref_err > 0).mean():.3g}" + ("" if resolvable else " (unresolvable: fewer than 100 errors)"))
    assert resolvable or ebn0 >= 1.25, "a point planned as resolvable made fewer than 100 reference errors"
    for dec in (T.DECODER_EXACT, T.DECODER_WINDOW_LOG, T.DECODER_WINDOW_LIN):
        post, _, _ = T.decode_batch_host([k] * n, llr, dec, 8, min_segment=32, warmup=16)
        err = ((post.reshape(n, k) >= 0) != bits).sum(1)
        ber, bler = err.sum() / (n * k), (err > 0).mean()
        _say(f"  decoder={dec} bit errors={int(err.sum())} BER={ber:.3g} BLER={bler:.3g}")
        assert lo_b <= ber <= hi_b, (dec, ber, lo_b, hi_b)
        assert lo_f <= bler <= hi_f, (dec, bler, lo_f, hi_f)


# ------------------------------------------- reference test mirrors (GPU) --
def _tb(tbs, pattern, iters=6):
    bits = np.array([(i * pattern[0] + pattern[1]) & 1 for i in range(tbs)], np.uint8)
    seg = T.segment_transport_block(T.append_crc24a(bits))
    p = T.LteTurboCodecParams(transport_block_bits=tbs, g_total=sum(3 * (len(c) + 4) for c in seg.code_blocks),
                              max_iterations=iters)
    return bits, p, T.encode_transport_block(p, bits)


@pytest.mark.parametrize("tbs,pattern", [(40, (7, 3)), (512, (11, 5)), (6112, (19, 11)), (12000, (17, 9))])
def test_reference_roundtrips_gpu(tbs, pattern):
    # reference tests/test_main.cpp:132-186, 211-254, 287-302, 336-343
    bits, p, enc = _tb(tbs, pattern)
    cpu = T.decode_transport_block(p, enc, T.DecoderBackend.CpuReference)
    for backend in (T.DecoderBackend.GpuStage1Exact, T.DecoderBackend.GpuStage2Windowed):
        dec = T.decode_transport_block(p, enc, backend)
        assert np.array_equal(dec.tb_bits, bits), backend
        assert dec.transport_block_crc_ok
        assert dec.posterior_llr is not None and len(dec.posterior_llr) == len(cpu.posterior_llr)
        assert np.all(np.isfinite(dec.posterior_llr))
        if backend == T.DecoderBackend.GpuStage1Exact:
            # test_main.cpp:142 (Stage1 vs CPU, 2e-5) on noise-free +-8 inputs
            np.testing.assert_allclose(dec.posterior_llr, cpu.posterior_llr, atol=2e-5, rtol=1e-6)


@pytest.mark.parametrize("backend", ["GpuStage1Exact", "GpuStage2Windowed"])
def test_gpu_roundtrip_rv_above_3(backend):
    """rv = 4: the reference takes k0 of any rv modulo the circular buffer
    (lte_turbo.cpp:263-264, 294-295); the device de-rate tables cover rv 0..3,
    so such TBs take the host de-rate + batched decode path and still decode,
    matching the CPU decoder (Stage1 within 2e-5 as test_main.cpp:142)."""
    bits = np.array([(i * 13 + 7) & 1 for i in range(12000)], np.uint8)
    seg = T.segment_transport_block(T.append_crc24a(bits))
    p = T.LteTurboCodecParams(transport_block_bits=12000, g_total=sum(3 * (len(c) + 4) for c in seg.code_blocks),
                              rv=4, max_iterations=6)
    enc = T.encode_transport_block(p, bits)
    cpu = T.decode_transport_block(p, enc, T.DecoderBackend.CpuReference)
    dec = T.decode_transport_block(p, enc, getattr(T.DecoderBackend, backend))
    assert np.array_equal(dec.tb_bits, bits) and dec.transport_block_crc_ok
    if backend == "GpuStage1Exact":
        np.testing.assert_allclose(dec.posterior_llr, cpu.posterior_llr, atol=2e-5, rtol=1e-6)


def test_tile8_alias_hard_decisions():
    # reference tests/test_main.cpp:153-175: tile8 hard decisions equal cb_bits
    bits, p, enc = _tb(40, (7, 3))
    blk = enc.code_blocks[0]
    l = T.derate_match_code_block(p, blk.descriptor, blk.rate_matched_bits)
    post, _, _ = T.turbo_decode_code_block_gpu_stage1_tile8_experimental(blk.descriptor, *_split(l, 64), 6, False)
    assert np.all(np.isfinite(post))
    assert np.array_equal((post >= 0).astype(np.uint8), blk.cb_bits)


def test_gpu_error_paths():
    d = T.CodeBlockDescriptor(0, 40, 44)
    z = np.zeros(44, np.float32)
    with pytest.raises(ValueError):
        T.turbo_decode_code_block_gpu_stage1(d, z[:43], z, z, 8, False)
    with pytest.raises(IndexError):
        T.turbo_decode_code_block_gpu_stage2(T.CodeBlockDescriptor(0, 41, 45), np.zeros(45, np.float32),
                                             np.zeros(45, np.float32), np.zeros(45, np.float32), 8, False)
    with pytest.raises(ValueError):
        T.decode_batch_host([40], np.zeros(10, np.float32))


def test_plan_on_device_buffers():
    torch = pytest.importorskip("torch")
    ks = [6144] * 5 + [1024] * 3
    rng = np.random.default_rng(3)
    llrs = []
    for k in ks:
        cb = rng.integers(0, 2, k).astype(np.uint8)
        llrs.append(np.concatenate(bpsk_awgn_llr(O.encode_cb(cb), k, 0.75, rng)))
    host = np.concatenate(llrs)
    plan = T.Plan(ks, T.DECODER_WINDOW_LOG, 8)
    x = torch.from_numpy(host).cuda()
    y = torch.empty(plan.post_floats, dtype=torch.float32, device="cuda")
    it = torch.zeros(len(ks), dtype=torch.int32, device="cuda")
    plan.decode(x.data_ptr(), y.data_ptr(), 0, it.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    want, _, _ = T.decode_batch_host(ks, host, T.DECODER_WINDOW_LOG, 8)
    assert np.array_equal(y.cpu().numpy().view(np.uint32), want.view(np.uint32))
    assert it.cpu().tolist() == [8] * len(ks)


# ------------------------------------------------------- early stop (CRC24B)
def _crc_blocks(ks, fillers, ebn0s, seed):
    """Code blocks that carry a CRC24B over bits [F, K) like segmented
    transport blocks (filler bits zero), BPSK/AWGN LLRs."""
    rng = np.random.default_rng(seed)
    bits, llrs = [], []
    for k, f, e in zip(ks, fillers, ebn0s):
        payload = rng.integers(0, 2, k - f - 24).astype(np.uint8)
        cb = np.concatenate([np.zeros(f, np.uint8), O.crc_append(O.CRC24B, payload)])
        bits.append(cb)
        llrs.append(np.concatenate(bpsk_awgn_llr(O.encode_cb(cb, f), k, e, rng)))
    return bits, llrs


def test_exact_early_stop_blocks_sharing_a_wave():
    """The exact decoder runs two blocks of one K per wave in lockstep; a block
    whose CRC24B passes must freeze (posterior and iteration count of the
    passing iteration) while its partner goes on.  Each block of the batch
    equals its own single-block decode with early stop, bit for bit, and the
    verdicts match the reference's decoder on the CPU."""
    ks = [1024, 1024, 1024, 1024, 6144, 6144, 40]
    fs = [0, 8, 0, 0, 0, 24, 0]
    eb = [2.0, -1.0, 0.6, 1.5, -1.0, 1.0, 3.0]
    _, llrs = _crc_blocks(ks, fs, eb, 91)
    post, crc, its = T.decode_batch_host(ks, np.concatenate(llrs), T.DECODER_EXACT, 8, True, fillers=fs)
    off = 0
    for k, f, l, c, i in zip(ks, fs, llrs, crc, its):
        d = T.CodeBlockDescriptor(0, k, k + 4, filler_count=f)
        single, ok, used = T.turbo_decode_code_block_gpu_stage1(d, *_split(l, k), 8, True)
        cpu, ok_cpu, used_cpu = T.turbo_decode_code_block_cpu(d, *_split(l, k), 8, True)
        _say(f"K={k} F={f}: batch it={int(i)} crc={bool(c)} single it={used} crc={ok} cpu it={used_cpu} crc={ok_cpu}")
        assert (int(i), bool(c)) == (used, ok) == (used_cpu, ok_cpu)
        assert np.array_equal(post[off:off + k].view(np.uint32), single.view(np.uint32))
        off += k
    assert crc.sum() >= 3 and not crc[1] and not crc[4]


@pytest.mark.parametrize("decoder", [T.DECODER_WINDOW_LOG, T.DECODER_WINDOW_LIN])
def test_window_early_stop_matches_oracle(decoder):
    """Stopping iteration and CRC verdict per block equal the oracle's
    restatement; posteriors agree to the window-decoder bar.  Mixed K, so
    some jobs pack several blocks per wavefront that stop at different
    iterations (and one block at -1 dB that never passes)."""
    ks = [1024, 6144, 6144, 40, 40, 40, 544, 544, 1024, 104]
    fs = [8, 0, 24, 0, 0, 0, 8, 0, 0, 0]
    eb = [1.0, 1.0, -1.0, 2.0, 0.0, 3.0, 1.5, 0.5, 1.2, 2.5]
    bits, llrs = _crc_blocks(ks, fs, eb, 77)
    post, crc, its = T.decode_batch_host(ks, np.concatenate(llrs), decoder, 8, True, fillers=fs)
    off = 0
    lin = decoder == T.DECODER_WINDOW_LIN
    for i, (k, f, l) in enumerate(zip(ks, fs, llrs)):
        geo = T.segment_geometry(k, decoder)
        ref, ok, used = O.decode_window(*_split(l, k), k, geo["segment_length"], geo["warmup"], lin, f, 8, True,
                                        geo["extrinsic_bits"])
        _say(f"block {i}: K={k} F={f} gpu it={its[i]} crc={crc[i]}  oracle it={used} crc={ok}")
        assert (int(its[i]), bool(crc[i])) == (used, ok)
        atol, rtol, hard = _window_bar(geo, ref)
        assert_llr_parity(post[off:off + k], ref, atol=atol, rtol=rtol, hard_tol=hard)
        off += k
    assert crc.sum() >= 6 and not crc[2]


@pytest.mark.parametrize("early_stop", [False, True])
@pytest.mark.parametrize("xbits", [16, 32])
@pytest.mark.parametrize("warm", [8, 16, 24])
def test_window_lean_kernel_equals_general(warm, xbits, early_stop):
    """A plan whose every job has the lean geometry (one block in 64
    segments, segment length and warm-up multiples of 8: K >= 2048 a
    multiple of 512) runs the lean log-domain kernel; one K=40 block more
    makes the same blocks run the general kernel.  Posteriors, CRC verdicts
    and stopping iterations are the same bits (fillers, blocks that stop at
    different iterations and ones that never pass included)."""
    ks = [2048, 6144, 3072, 2560, 6144, 4096, 5632, 6144, 2048, 3584]
    fs = [8, 0, 0, 24, 0, 0, 0, 16, 0, 0]
    eb = [1.5, 0.4, -1.0, 1.0, 0.7, 2.0, 0.9, 0.5, -0.5, 1.2]
    _, llrs = _crc_blocks(ks, fs, eb, 60 + warm)
    _, short = _crc_blocks([40], [0], [2.0], 61)
    lean = T.decode_batch_host(ks, np.concatenate(llrs), T.DECODER_WINDOW_LOG, 8, early_stop, fs, 0, warm, xbits)
    gen = T.decode_batch_host(ks + [40], np.concatenate(llrs + short), T.DECODER_WINDOW_LOG, 8, early_stop,
                              fs + [0], 0, warm, xbits)
    n = sum(ks)
    _say(f"warm-up {warm}, fp{xbits} held, early stop {early_stop}: iterations {list(lean[2])}")
    assert np.array_equal(lean[0].view(np.uint32), gen[0][:n].view(np.uint32))
    assert list(lean[1]) == list(gen[1][:-1]) and list(lean[2]) == list(gen[2][:-1])
    if early_stop:
        assert lean[1].sum() >= 5 and not lean[1][2] and min(lean[2]) < 8


@pytest.mark.parametrize("early_stop", [False, True])
@pytest.mark.parametrize("decoder", [T.DECODER_WINDOW_LOG, T.DECODER_WINDOW_LIN])
@pytest.mark.parametrize("k", [96, 352, 544, 768])
def test_window_spare_lane_jobs_match_oracle(k, decoder, early_stop):
    """Jobs that pack several code blocks of several segments each into one
    wave and leave lanes idle (M = 3/11/17/24 segments, 21/5/3/2 blocks per
    wave, 1/9/13/16 spare lanes), with 2 x blocks_per_wave + 1 blocks so the
    last job is partly empty too: the idle lanes' shadowing, the per-lane
    staging and the strided copy-out.  Each block equals the oracle's
    restatement at the window bar (_window_bar); with early stop, stopping
    iteration and CRC verdict equal the oracle's."""
    geo = T.segment_geometry(k, decoder)
    m, bpw = geo["segments"], geo["blocks_per_wave"]
    assert 1 < m and bpw > 1 and m * bpw < 64
    n = 2 * bpw + 1
    rng = np.random.default_rng(1000 * k + 10 * decoder + early_stop)
    ebs = [float(x) for x in rng.uniform(0.0, 3.0, n)]
    if early_stop:
        _, llrs = _crc_blocks([k] * n, [0] * n, ebs, 1000 + k)
    else:
        llrs = []
        for e in ebs:
            cb = rng.integers(0, 2, k).astype(np.uint8)
            llrs.append(np.concatenate(bpsk_awgn_llr(O.encode_cb(cb), k, e, rng)))
    _say(f"spare lanes: K={k} M={m} blocks/wave={bpw} n={n} decoder={decoder} es={early_stop}")
    post, crc, its = T.decode_batch_host([k] * n, np.concatenate(llrs), decoder, 8, early_stop)
    lin = decoder == T.DECODER_WINDOW_LIN
    for i, l in enumerate(llrs):
        ref, ok, used = O.decode_window(*_split(l, k), k, geo["segment_length"], geo["warmup"], lin, 0, 8,
                                        early_stop, geo["extrinsic_bits"])
        atol, rtol, hard = _window_bar(geo, ref)
        assert_llr_parity(post[i * k:(i + 1) * k], ref, atol=atol, rtol=rtol, hard_tol=hard)
        if early_stop:
            assert (int(its[i]), bool(crc[i])) == (used, ok), i
        else:
            assert its[i] == 8
    if early_stop:
        assert crc.sum() > 0


# ------------------------------------------------ blocks on several waves
# (K, segment length, waves per block, warm-up): 192, 96, 128, 126 and 160
# segments; the last two leave lanes of their last wave idle
WIDE_CASES = [(6144, 32, 3, 24), (6144, 64, 2, 16), (6144, 48, 2, 32), (4032, 32, 2, 16), (5120, 32, 3, 8)]


def _wide_geo(k, decoder, seg, waves, warm, xbits=0):
    geo = T.segment_geometry(k, decoder, seg, warm, xbits, waves)
    assert (geo["segment_length"], geo["waves_per_block"]) == (seg, waves), geo
    return geo


@pytest.mark.parametrize("xbits", [16, 32])
@pytest.mark.parametrize("decoder", [T.DECODER_WINDOW_LOG, T.DECODER_WINDOW_LIN])
@pytest.mark.parametrize("k,seg,waves,warm", WIDE_CASES)
def test_window_multi_wave_matches_oracle(k, seg, waves, warm, decoder, xbits):
    """Code blocks of more than 64 segments span 2 or 3 waves of one
    work-group (max_waves_per_block): window lengths 32, 48 and 64 at
    K=6144.  Each block equals the oracle's restatement with the same
    segments and warm-up at the window bar (_window_bar: fp32-held 1e-3 +
    1e-4 |LLR|, hard decisions equal where |LLR| >= 1e-3; fp16-held scaled
    by max |LLR|); a second decode of the same blocks gives the same bits
    (the waves' hand-offs through memory are ordered by barriers, not by
    timing)."""
    geo = _wide_geo(k, decoder, seg, waves, warm, xbits)
    rng = np.random.default_rng(k + seg + 7 * decoder + xbits)
    llrs = []
    for e in (0.25, 0.6, 1.0):
        cb = rng.integers(0, 2, k).astype(np.uint8)
        llrs.append(np.concatenate(bpsk_awgn_llr(O.encode_cb(cb), k, e, rng)))
    _say(f"several waves: K={k} S={seg} waves={waves} g={warm} decoder={decoder} fp{xbits}")
    args = ([k] * 3, np.concatenate(llrs), decoder, 8, False, None, seg, warm, xbits, waves)
    post, _, its = T.decode_batch_host(*args)
    again, _, _ = T.decode_batch_host(*args)
    assert np.array_equal(post.view(np.uint32), again.view(np.uint32))
    lin = decoder == T.DECODER_WINDOW_LIN
    for i, l in enumerate(llrs):
        ref = O.decode_window(*_split(l, k), k, seg, geo["warmup"], lin, 0, 8, False, geo["extrinsic_bits"])[0]
        atol, rtol, hard = _window_bar(geo, ref)
        assert_llr_parity(post[i * k:(i + 1) * k], ref, atol=atol, rtol=rtol, hard_tol=hard)
    assert list(its) == [8, 8, 8]


@pytest.mark.parametrize("decoder", [T.DECODER_WINDOW_LOG, T.DECODER_WINDOW_LIN])
def test_window_multi_wave_plan_early_stop_matches_oracle(decoder):
    """One plan of 3-wave jobs (max_waves_per_block 3, segment length from
    32): K=6144 and 5120 blocks span three waves; shorter blocks are packed
    192 lanes to a job (K=1024: 6 blocks of 32 segments, K=544: 11 of 17, so
    blocks straddle waves; K=40: one segment each), with CRC24B early stop,
    fillers and a block at -1 dB that never passes.  Stopping iteration and
    CRC verdict equal the oracle's, posteriors at the window bar."""
    ks = [6144, 1024, 6144, 40, 5120, 544, 6144, 1024, 40, 2048, 544, 544]
    fs = [0, 8, 24, 0, 0, 0, 16, 0, 0, 8, 0, 0]
    eb = [1.0, 1.2, -1.0, 2.0, 0.8, 1.5, 0.6, 0.3, 0.0, 1.1, 2.5, 0.9]
    _, llrs = _crc_blocks(ks, fs, eb, 88 + decoder)
    post, crc, its = T.decode_batch_host(ks, np.concatenate(llrs), decoder, 8, True, fs, 32, 16, 0, 3)
    lin = decoder == T.DECODER_WINDOW_LIN
    off = 0
    for i, (k, f, l) in enumerate(zip(ks, fs, llrs)):
        geo = T.segment_geometry(k, decoder, 32, 16, 0, 3)
        ref, ok, used = O.decode_window(*_split(l, k), k, geo["segment_length"], geo["warmup"], lin, f, 8, True,
                                        geo["extrinsic_bits"])
        _say(f"block {i}: K={k} M={geo['segments']} gpu it={its[i]} crc={crc[i]}  oracle it={used} crc={ok}")
        assert (int(its[i]), bool(crc[i])) == (used, ok)
        atol, rtol, hard = _window_bar(geo, ref)
        assert_llr_parity(post[off:off + k], ref, atol=atol, rtol=rtol, hard_tol=hard)
        off += k
    assert crc.sum() >= 6 and not crc[2] and min(its) < 8


@pytest.mark.parametrize("early_stop", [False, True])
def test_window_blocks_in_multi_wave_jobs_equal_one_wave(early_stop):
    """Blocks of at most 64 segments decode to the same bits in a plan of
    3-wave jobs (192 lanes; blocks packed across the waves, a K=6144 block of
    192 segments makes the plan wide) as in one-wave jobs of the same
    segment length and warm-up: posteriors, CRC verdicts, iterations."""
    ks = [1024] * 7 + [40] * 5 + [544] * 12 + [2048] * 2 + [352] * 6
    for k in set(ks):
        a, b = (T.segment_geometry(k, T.DECODER_WINDOW_LOG, 32, 16, 0, n) for n in (1, 3))
        assert (a["segment_length"], a["warmup"]) == (b["segment_length"], b["warmup"])
    rng = np.random.default_rng(5 + early_stop)
    eb = [float(x) for x in rng.uniform(-0.5, 2.5, len(ks))]
    _, llrs = _crc_blocks(ks, [0] * len(ks), eb, 70 + early_stop)
    _, big = _crc_blocks([6144], [0], [1.0], 71)
    one = T.decode_batch_host(ks, np.concatenate(llrs), T.DECODER_WINDOW_LOG, 8, early_stop, None, 32, 16, 0, 1)
    wide = T.decode_batch_host(ks + [6144], np.concatenate(llrs + big), T.DECODER_WINDOW_LOG, 8, early_stop,
                               None, 32, 16, 0, 3)
    n = sum(ks)
    _say(f"small blocks in 3-wave jobs, early stop {early_stop}: iterations {list(one[2])}")
    assert np.array_equal(one[0].view(np.uint32), wide[0][:n].view(np.uint32))
    assert list(one[1]) == list(wide[1][:-1]) and list(one[2]) == list(wide[2][:-1])


def test_window_early_stop_equals_fixed_iterations():
    """A block that stops at iteration n has bit-for-bit the posterior of the
    same decoder run for exactly n iterations without early stop."""
    ks = [6144, 1024, 40, 40]
    fs = [0, 8, 0, 0]
    bits, llrs = _crc_blocks(ks, fs, [1.0, 1.0, 2.0, 1.0], 5)
    post, crc, its = T.decode_batch_host(ks, np.concatenate(llrs), T.DECODER_WINDOW_LOG, 8, True, fillers=fs)
    off = 0
    for i, k in enumerate(ks):
        assert crc[i] and its[i] < 8
        fixed, _, _ = T.decode_batch_host([k], llrs[i], T.DECODER_WINDOW_LOG, int(its[i]))
        assert np.array_equal(post[off:off + k].view(np.uint32), fixed.view(np.uint32))
        off += k


@pytest.mark.parametrize("backend", ["GpuStage1Exact", "GpuStage2Windowed"])
def test_transport_block_early_stop(backend):
    """Segmented TB (C=2, CRC24B per block) with early stop on: decoded,
    CRC-clean, and fewer iterations than the maximum (noise-free inputs)."""
    bits, p, enc = _tb(12000, (17, 9), iters=8)
    p.enable_early_stop = True
    cpu = T.decode_transport_block(p, enc, T.DecoderBackend.CpuReference)
    dec = T.decode_transport_block(p, enc, getattr(T.DecoderBackend, backend))
    assert np.array_equal(dec.tb_bits, bits) and dec.transport_block_crc_ok
    assert dec.iterations_used < 8
    if backend == "GpuStage1Exact":
        assert dec.iterations_used == cpu.iterations_used
    assert all(dec.code_block_crc_ok)


# ------------------------------------------------ host threads at once --
def test_concurrent_host_threads_match_sequential():
    """Eight host threads decode at once through the host-buffer entry points
    (the per-block GPU calls, as the reference's decode_transport_block issues
    them from std::async threads, lte_turbo.cpp:941-951, and the batched host
    call).  Each thread runs on a stream of its own; every result equals the
    same call made alone, bit for bit."""
    import concurrent.futures as cf
    rng = np.random.default_rng(808)
    ks = [6144, 1024, 544, 40, 2048, 6144, 104, 3200]
    llrs = []
    for k in ks:
        cb = rng.integers(0, 2, k).astype(np.uint8)
        llrs.append(np.concatenate(bpsk_awgn_llr(O.encode_cb(cb), k, 0.75, rng)))

    def work(i):
        k, l = ks[i], llrs[i]
        d = T.CodeBlockDescriptor(0, k, k + 4)
        a = T.turbo_decode_code_block_gpu_stage1(d, *_split(l, k), 8, False)[0]
        b = T.turbo_decode_code_block_gpu_stage2(d, *_split(l, k), 8, False)[0]
        c = T.decode_batch_host([k, k], np.concatenate([l, l]), T.DECODER_WINDOW_LIN, 8)[0]
        return a, b, c

    alone = [work(i) for i in range(len(ks))]
    _say("8 host threads decoding at once")
    for _ in range(3):
        with cf.ThreadPoolExecutor(len(ks)) as ex:
            together = list(ex.map(work, range(len(ks))))
        for x, y in zip(together, alone):
            for u, v in zip(x, y):
                assert _bits_equal(u, v)
