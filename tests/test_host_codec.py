"""Host codec of turbo_cpp_amd (C ABI + Python mirror) against the oracle
and the golden vectors; library exports; error behaviour.  CPU only."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
import turbo_cpp_amd as T
from tests.helpers import bpsk_awgn_llr, golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_c_abi_symbol():
    """Every function declared in include/turbo_cpp/c_api.h is exported."""
    hdr = open(os.path.join(ROOT, "include", "turbo_cpp", "c_api.h")).read()
    names = set(re.findall(r"\b(tc_[a-z_0-9]+)\s*\(", hdr))
    assert len(names) >= 19
    lib = ctypes.CDLL(T.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_library_exports_cpp_api():
    """The C++ API of lte_turbo.h is exported (mangled names present)."""
    out = os.popen(f"nm -D --defined-only {T.LIB_PATH}").read()
    for fn in ("turbo_decode_code_block_cpu", "turbo_decode_code_block_gpu_stage1",
               "turbo_decode_code_block_gpu_stage2", "turbo_decode_code_block_gpu_stage1_tile8_experimental",
               "decode_transport_block", "encode_transport_block", "derate_match_code_block",
               "segment_transport_block", "qpp_interleaver_table", "append_crc24a", "check_crc24b"):
        assert fn in out, fn


def test_qpp_table_matches_reference_kat():
    t = T.qpp_interleaver_table()
    assert len(t) == 188 and t[0] == (40, 3, 10) and t[-1] == (6144, 263, 480)
    assert t == O.qpp_table()


def test_crc_golden_and_oracle():
    g = golden()
    fns = [T.append_crc24a, T.append_crc24b, T.append_crc16, T.append_crc8]
    chk = [T.check_crc24a, T.check_crc24b, T.check_crc16, T.check_crc8]
    for kind in range(4):
        for n in (1, 7, 40, 333):
            out = fns[kind](g[f"crc{kind}_n{n}_in"])
            assert np.array_equal(out, g[f"crc{kind}_n{n}_out"])
            assert chk[kind](out)
    rng = np.random.default_rng(0)
    for kind in range(4):
        for n in (0, 9, 100, 1001):  # byte-table path and bit tail
            b = rng.integers(0, 2, n).astype(np.uint8)
            assert np.array_equal(fns[kind](b), O.crc_append(kind, b))


def test_reference_test_crc():
    # reference tests/test_main.cpp:25-35
    bits = [1, 0, 1, 1, 0, 1, 0, 0, 1, 1, 0, 1]
    assert T.check_crc24a(T.append_crc24a(bits))
    assert T.check_crc24b(T.append_crc24b(bits))
    assert T.check_crc16(T.append_crc16(bits))
    assert T.check_crc8(T.append_crc8(bits))


def test_segmentation_golden_and_reference_cases():
    g = golden()
    for b, fields in zip(g["seg_sizes"], g["seg_fields"]):
        s = T.segment_transport_block(g[f"seg_b{b}_in"])
        assert [s.c, s.c_plus, s.c_minus, s.k_plus, s.k_minus, s.filler_bits, s.l] == list(fields)
        assert np.array_equal(np.concatenate(s.code_blocks), g[f"seg_b{b}_blocks"])
    # reference tests/test_main.cpp:46-56
    s = T.segment_transport_block(np.ones(40, np.uint8))
    assert s.c == 1 and len(s.code_blocks[0]) == 40
    s = T.segment_transport_block(np.zeros(6145, np.uint8))
    assert s.c > 1 and s.filler_bits > 0


def test_encoder_and_rate_matching_golden():
    g = golden()
    for k, f in [(40, 0), (64, 0), (544, 8), (1024, 0), (6144, 8), (6016, 24)]:
        d = T.encode_code_block_streams(g[f"enc_k{k}_f{f}_in"], f)
        assert np.array_equal(np.stack(d), g[f"enc_k{k}_f{f}_out"])
    for ci, tbs, g_total, rv, r, k, e, f in g["tb_meta"]:
        key = f"tb{ci}_rv{rv}_cb{r}"
        p = T.LteTurboCodecParams(transport_block_bits=int(tbs), g_total=int(g_total), rv=int(rv))
        enc = T.encode_transport_block(p, g[f"tb{ci}_bits"])
        blk = enc.code_blocks[r]
        assert (blk.descriptor.k_r, blk.descriptor.e_r, blk.descriptor.filler_count) == (k, e, f)
        assert np.array_equal(blk.cb_bits, g[f"tb{ci}_rv0_cb{r}_cbbits"])
        assert np.array_equal(blk.rate_matched_bits, g[key + "_rm"]), key
        assert np.array_equal(T.derate_match_code_block(p, blk.descriptor, blk.rate_matched_bits), g[key + "_derm"])


@pytest.mark.parametrize("k,eb,it", [(40, 100, 6), (544, 50, 6), (1024, 50, 8), (6144, 50, 8)])
def test_cpu_decoder_bit_exact_golden(k, eb, it):
    g = golden()
    tag = f"dec_k{k}_eb{eb}_it{it}"
    llr = g[tag + "_llr"]
    post, ok, used = T.turbo_decode_code_block_cpu(T.CodeBlockDescriptor(0, k, k + 4), llr[0], llr[1], llr[2], it,
                                                   False)
    assert used == it and not ok
    assert np.array_equal(post.view(np.uint32), g[tag + "_post"].view(np.uint32))


def test_cpu_decoder_early_stop_golden():
    g = golden()
    llr = g["es_k1024_llr"]
    post, ok, used = T.turbo_decode_code_block_cpu(T.CodeBlockDescriptor(0, 1024, 1028), *llr, 8, True)
    assert [int(ok), used] == list(g["es_k1024_meta"])
    assert np.array_equal(post.view(np.uint32), g["es_k1024_post"].view(np.uint32))


def _tb_roundtrip(tbs, pattern, backend, iters=6):
    bits = np.array([(i * pattern[0] + pattern[1]) & 1 for i in range(tbs)], np.uint8)
    seg = T.segment_transport_block(T.append_crc24a(bits))
    p = T.LteTurboCodecParams(transport_block_bits=tbs, g_total=sum(3 * (len(c) + 4) for c in seg.code_blocks),
                              max_iterations=iters)
    enc = T.encode_transport_block(p, bits)
    return bits, enc, T.decode_transport_block(p, enc, backend)


@pytest.mark.parametrize("tbs,pattern", [(40, (7, 3)), (512, (11, 5)), (12000, (17, 9))])
def test_reference_roundtrips_cpu(tbs, pattern):
    # reference tests/test_main.cpp:58-131, 189-209, 305-334 (CPU parts)
    bits, enc, dec = _tb_roundtrip(tbs, pattern, T.DecoderBackend.CpuReference)
    assert np.array_equal(dec.tb_bits, bits)
    assert dec.transport_block_crc_ok
    assert len(dec.posterior_llr) == sum(c.descriptor.k_r for c in enc.code_blocks)


def test_reference_large_single_code_block_layout():
    # reference tests/test_main.cpp:257-285
    bits, enc, _ = _tb_roundtrip(6112, (19, 11), T.DecoderBackend.CpuReference, iters=1)
    assert len(enc.code_blocks) == 1
    assert enc.code_blocks[0].descriptor.k_r == 6144
    assert enc.code_blocks[0].descriptor.filler_count > 0


def test_error_mapping():
    d = T.CodeBlockDescriptor(0, 40, 44)
    z = np.zeros(44, np.float32)
    with pytest.raises(ValueError):  # std::invalid_argument: span sizes
        T.turbo_decode_code_block_cpu(d, z[:43], z, z, 8, False)
    with pytest.raises(IndexError):  # std::out_of_range: K not in the QPP table
        T.qpp_entry_for_k(41)
    with pytest.raises(ValueError):  # empty transport block
        T.segment_transport_block(np.zeros(0, np.uint8))
    with pytest.raises(ValueError):  # g_total not divisible by layers*Qm
        T.compute_e(0, 101, 1, 1, 2)
    with pytest.raises(ValueError):
        T.turbo_encode_code_block(T.CodeBlockDescriptor(0, 40, 44), np.zeros(39, np.uint8))


@pytest.mark.skipif(T.device_count() > 0, reason="a GPU is present")
def test_gpu_entry_points_fail_loudly_without_device():
    # lte_turbo.cpp:836,861,887: a GPU backend without a device is a runtime_error
    d = T.CodeBlockDescriptor(0, 40, 44)
    z = np.zeros(44, np.float32)
    for fn in (T.turbo_decode_code_block_gpu_stage1, T.turbo_decode_code_block_gpu_stage2):
        with pytest.raises(RuntimeError):
            fn(d, z, z, z, 8, False)
    with pytest.raises(RuntimeError):
        T.Plan([6144])
    with pytest.raises(RuntimeError):
        T.DeratePlan([6144], [18444])
    with pytest.raises(RuntimeError):
        T.derate_match_soft(np.zeros(100, np.float32), 40)


def test_soft_derate_argument_errors_before_device():
    # argument checks come before the device is touched: same errors with or without a GPU
    with pytest.raises(ValueError):
        T.derate_match_soft(np.zeros(10, np.float32), 40, 0, 4)  # rv outside 0..3
    with pytest.raises(IndexError):
        T.derate_match_soft(np.zeros(10, np.float32), 41, 0, 0)  # K not a QPP size
    with pytest.raises(ValueError):
        T.DeratePlan([40], [0])  # E = 0
    with pytest.raises(ValueError):
        T.DeratePlan([40], [10], fillers=[40])  # F >= K


def test_segment_geometry_all_table_sizes():
    """Every K of the table gets S | K, S <= 96, M = K/S <= 64, warm-up < S."""
    for (k, _, _) in T.qpp_interleaver_table():
        g = T.segment_geometry(k)
        s, m = g["segment_length"], g["segments"]
        assert k % s == 0 and s * m == k and m <= 64 and s <= 96
        assert g["blocks_per_wave"] == 64 // m
        assert (g["warmup"] == 0) if m == 1 else (0 < g["warmup"] < s)
    assert T.segment_geometry(6144)["segment_length"] == 96


def test_segment_geometry_several_waves_per_block():
    """max_waves_per_block n (window log/lin decoders) lowers the bound on
    S to ceil(K / (64 n)): window lengths 32 and 64 at K=6144 on 3 and 2
    waves; every table K keeps S | K, M <= 64 n, waves_per_block =
    ceil(M / 64) <= n and blocks_per_wave = 64 waves_per_block // M; n above
    3, or above 1 for the exact decoders, is an argument error."""
    L, N = T.DECODER_WINDOW_LOG, T.DECODER_WINDOW_LIN
    g = T.segment_geometry(6144, L, 32, 16, 0, 3)
    assert (g["segment_length"], g["segments"], g["waves_per_block"], g["blocks_per_wave"]) == (32, 192, 3, 1)
    g = T.segment_geometry(6144, N, 64, 16, 32, 3)
    assert (g["segment_length"], g["segments"], g["waves_per_block"], g["extrinsic_bits"]) == (64, 96, 2, 32)
    assert T.segment_geometry(6144, L, 32, 16, 0, 2)["segment_length"] == 48
    for n in (1, 2, 3):
        for (k, _, _) in T.qpp_interleaver_table():
            g = T.segment_geometry(k, L, 32, 16, 0, n)
            s, m, w = g["segment_length"], g["segments"], g["waves_per_block"]
            assert k % s == 0 and s * m == k and m <= 64 * n and w == (m + 63) // 64 <= n
            assert g["blocks_per_wave"] == 64 * w // m
            assert s == T.segment_geometry(k, L, 32, 16, 0, 1)["segment_length"] or m > 64
    with pytest.raises(ValueError):
        T.segment_geometry(6144, L, 32, 16, 0, 4)
    with pytest.raises(ValueError):
        T.segment_geometry(6144, T.DECODER_EXACT, 32, 16, 0, 2)


def test_single_hip_runtime_per_process():
    """Loading the library (and torch after it) maps one HIP runtime only."""
    import subprocess
    import sys

    code = ("import turbo_cpp_amd as T; T.lib(); import torch; "
            "m=open('/proc/self/maps').read(); "
            "print(len({l.split()[-1] for l in m.splitlines() if 'amdhip64' in l}))")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().splitlines()[-1] == "1"


def test_tb_plan_argument_errors_before_device():
    # descriptor checks come before the device is touched
    with pytest.raises(ValueError):
        T.TbPlan([(0, 100, 0, 1, 2)])
    with pytest.raises(ValueError):
        T.TbPlan([(40, 101, 0, 1, 2)])  # g_total not divisible by layers * Qm
    with pytest.raises(ValueError):
        T.TbPlan([(40, 100, 4, 1, 2)])


def test_channel_plan_argument_errors_before_device():
    with pytest.raises(ValueError):
        T.ChannelPlan([(10, 3, 0, 1.0)])  # qm
    with pytest.raises(ValueError):
        T.ChannelPlan([(10, 4, 0, 1.0)])  # n_bits not a multiple of qm
    with pytest.raises(ValueError):
        T.ChannelPlan([(12, 4, 0, 0.0)])  # n0



def test_transport_block_entry_points_check_buffer_lengths():
    """The batched TB entry points take no lengths in C; the Python mirror
    checks them against the descriptors first (a short array would otherwise
    be read past its end)."""
    tbs = [(1000, 3108, 0, 1, 2), (40, 150, 1, 1, 2), (6120, 6000, 0, 1, 4, 1)]
    ks, es = T._tb_sizes(T._tb_descs(tbs), len(tbs))
    assert ks == [1024, 64, 6144] and es[1] == 150 and sum(es) == 3108 + 150 + 6000
    with pytest.raises(ValueError):
        T.decode_transport_blocks(tbs, np.zeros(sum(es) - 1, np.float32))
    with pytest.raises(ValueError):
        T.encode_transport_blocks(tbs, [np.zeros(1000, np.uint8), np.zeros(39, np.uint8), np.zeros(6120, np.uint8)])
    with pytest.raises(ValueError):
        T.encode_transport_blocks(tbs, [np.zeros(1000, np.uint8)])
