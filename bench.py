"""Throughput benchmark: batched LTE turbo decoding on one MI355X.

Headline workload (uniform batch): 4096 code blocks of K = 6144 (f1=263,
f2=480), rate-1/3 mother code with 12 tail bits (18444 coded bits per block),
each block a transport block of 6120 seeded Bernoulli(0.5) payload bits +
CRC24A, BPSK (bit c -> 2c-1) over AWGN with block i at Eb/N0 = 0.25 * (i mod 7)
dB (the 0.0..1.5 dB sweep), fp32 LLRs 2y/sigma^2, 8 iterations, no early stop.
One step = one batched decode of all 4096 blocks with the inputs resident in
HBM.  The window geometry (segment length, warm-up, waves per CU) is the
autotuner's BER-gated choice for K = 6144.

Printed: one JSON line (metric, value in info Gbit/s, roofline of the decode
kernel, the reference CPU decoder timed on this host, the plain
one-launch-per-block GPU decoder, BER/BLER per Eb/N0 for each decoder) with
further legs:
  rx_chain         the same blocks as 4096 transport blocks through soft de-rate
                   matching, the decode and TB assembly;
  differentiation  2048 x K=1024, 6 iterations, 12 learnable extrinsic scales:
                   forward / adjoint time (bwd Gbit/s), 500 Adam steps,
                   held-out BER against the plain decoder;
  uneven_batch     ~20k code blocks of transport blocks of uneven size,
                   QPSK/16QAM/64QAM, rv 0..3, early stop, HIP graph;
  mixed_load       65536 code blocks (70% K=6144), 16QAM over Rayleigh,
                   HARQ rv0+rv2 combining, early stop, HIP graph.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

K = 6144
N_CB = 4096
ITERS = 8
EBN0 = [0.25 * i for i in range(7)]
# Algorithmic work per info bit per iteration in the reference's pairwise
# Log-MAP form: 384 simple ops + 120 transcendental ops.  A lane-slot is the
# SIMD time of one lane of a simple FP32 op: MI355X's published FP32 vector
# peak (157.3 TFLOP/s, i.e. 78.6 T FMA lanes/s = 256 CUs x 4 SIMDs x 32 lanes
# x 2.4 GHz: a wave64 op every 2 cycles); transcendentals (v_exp_f32,
# v_log_f32) run at quarter rate, 8 SIMD cycles per wave64 op (the counter
# calibration: SQ_ACTIVE_INST_VALU charges them twice a simple op and a
# kernel of nothing but v_exp_f32 keeps its SIMDs 96% busy at that price), so
# 4 lane-slots each: 384 + 4 x 120 = 864 lane-slots per bit and iteration.
# HBM bytes per block: 12(K+4) in + 4K out.
VALU_SLOTS_PER_BIT_ITER = 384 + 120 * 4
VALU_PEAK = 256 * 4 * 32 * 2.4e9  # lane-slots/s
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def valu_calibration():
    """SIMD cycles one wave64 v_add_f32 / v_exp_f32 occupies, measured with
    independent operations at 1, 2, 4 and 8 one-wave work-groups per SIMD
    (turbo_cpp_amd/csrc/calib/valu_rate.hip)."""
    import ctypes as C
    path = os.path.join(ROOT, "turbo_cpp_amd", "libturbo_calib.so")
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    lib.tc_calib_valu.argtypes = [C.c_int, C.c_int, C.c_int] + [C.POINTER(C.c_double)] * 4
    rows = []
    for op, name in ((0, "v_add_f32"), (1, "v_exp_f32")):
        for per_simd in (1, 2, 4, 8):
            v = [C.c_double() for _ in range(4)]
            rc = lib.tc_calib_valu(op, 1024 * per_simd, 4000, *[C.byref(x) for x in v])
            if rc:
                return {"error": f"tc_calib_valu returned {rc}"}
            rows.append({"op": name, "workgroups_per_simd": per_simd, "clock_mhz": round(v[0].value),
                         "resident_waves_per_simd": round(v[1].value, 2),
                         "wave_cycles_per_instr": round(v[2].value, 2),
                         "simd_cycles_per_instr": round(v[3].value, 2)})
    return {"what": "16 independent ops per round, 4000 rounds per wave; SIMD cycles per wave64 instruction = "
                    "a wave's cycles per instruction / resident waves per SIMD",
            "rows": rows}


def make_inputs(n, seed=12345):
    """Blocks are transport blocks of B = K - 24 payload bits + CRC24A, so the
    receive-chain leg can check each TB CRC."""
    import turbo_cpp_amd as T

    rng = np.random.default_rng(seed)
    payload = rng.integers(0, 2, (n, K - 24)).astype(np.uint8)
    bits = np.stack([T.append_crc24a(payload[i]) for i in range(n)])
    coded = np.empty((n, 3 * (K + 4)), np.float32)
    for i in range(n):
        d0, d1, d2 = T.encode_code_block_streams(bits[i])
        coded[i, :K + 4] = d0
        coded[i, K + 4:2 * (K + 4)] = d1
        coded[i, 2 * (K + 4):] = d2
    rate = K / (3 * K + 12)
    ebn0 = np.array([EBN0[i % len(EBN0)] for i in range(n)])
    sigma2 = (1.0 / (2 * rate * 10 ** (ebn0 / 10))).astype(np.float32)[:, None]
    y = (2.0 * coded - 1.0) + np.sqrt(sigma2) * rng.standard_normal(coded.shape, dtype=np.float32)
    llr = (2.0 * y / sigma2).astype(np.float32)
    return bits, llr, ebn0


def ber_table(post, bits, ebn0):
    out = {}
    err = ((post.reshape(bits.shape) >= 0) != bits).sum(1)
    for e in EBN0:
        m = ebn0 == e
        out[f"{e:.2f}"] = {"ber": float(err[m].sum() / (m.sum() * K)), "bler": float((err[m] > 0).mean()),
                           "blocks": int(m.sum())}
    return out


def host_cores():
    """CPU cores this process may use: the affinity mask, capped by a cgroup
    CPU quota when one is set (threads beyond it only queue)."""
    nproc = os.cpu_count() or 1
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    use = affinity if quota is None else max(1, min(affinity, int(math.ceil(quota))))
    return use, {"nproc": nproc, "affinity": affinity, "cgroup_cpu_quota": quota}


def cpu_baseline(llr):
    """The reference's CPU decoder (oracle/_ref, its own lte_turbo.cpp) or, if
    that was not built, the C restatement, on a bounded sample of blocks: all
    usable host cores (one block per call per thread), and one thread as the
    median of 5 blocks timed one by one after a warm-up."""
    import oracle

    threads, cores = host_cores()
    if oracle.ref_available():
        kind = "reference"
        n = 30 * threads  # ~37 ms per block and thread: 15-20 s of CPU work in about a second
        sample = np.ascontiguousarray(llr[:n])
        oracle.ref_decode_many(sample[:1], K, ITERS, 1)  # warm-up
        t0 = time.perf_counter()
        oracle.ref_decode_many(sample, K, ITERS, threads)
        dt = time.perf_counter() - t0
        one = []
        for i in range(5):
            t1 = time.perf_counter()
            oracle.ref_decode_many(sample[i:i + 1], K, ITERS, 1)
            one.append(time.perf_counter() - t1)
    else:
        kind = "port"
        threads = 1
        n = 24
        t0 = time.perf_counter()
        one = []
        for i in range(n):
            l = llr[i]
            t1 = time.perf_counter()
            oracle.decode_cb(l[:K + 4], l[K + 4:2 * (K + 4)], l[2 * (K + 4):], K, 0, ITERS, False)
            one.append(time.perf_counter() - t1)
        dt = time.perf_counter() - t0
    single = K / float(np.median(one)) / 1e9
    return {"value": n * K / dt / 1e9, "unit": "Gbit/s", "cores": threads, "kind": kind,
            "sample": f"{n} blocks of K={K}, 8 iterations, one block per call on {threads} threads "
                      f"({dt:.1f} s wall); single thread: median of {len(one)} blocks timed one by one "
                      f"after a warm-up, {single * 1e3:.3f} Mbit/s",
            "single_thread_gbit_s": single, "host": cores}


def e2e_leg(T, llr, opts):
    """End to end through the C ABI: tc_decode_batch_host on the 4096 blocks
    (host LLRs in, pageable H2D copy, decode, D2H copy of the posteriors; the
    thread's cached plan and buffers are reused after the first call)."""
    ks = [K] * N_CB
    opts = {k: v for k, v in opts.items() if k in ("min_segment", "warmup", "max_waves_per_block")}
    host = np.ascontiguousarray(llr.reshape(-1))
    T.decode_batch_host(ks, host, T.DECODER_WINDOW_LOG, ITERS, **opts)  # plan + buffers
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        T.decode_batch_host(ks, host, T.DECODER_WINDOW_LOG, ITERS, **opts)
        ts.append(time.perf_counter() - t0)
    ms = float(np.median(ts)) * 1e3
    log(f"end to end (host buffers): {ms:.1f} ms per 4096-block call")
    return {"what": "tc_decode_batch_host: 302 MB pageable H2D + decode + 101 MB D2H, median of 3 calls",
            "ms": ms, "gbit_s": N_CB * K / ms / 1e6, "input_bytes": host.nbytes, "output_bytes": N_CB * K * 4}


# tests/bench_main.cpp:242-249 cases: (name, TB bits, repeats); 6 iterations, no early stop, rv 0,
# full-rate g_total, noise-free +-8 LLRs from the hard rate-matched bits, payload bit i = (13 i + 7) & 1
REF_BENCH_CASES = [("micro_single_cb", 40, 20), ("small_single_cb", 512, 10), ("max_single_cb", 6112, 3),
                   ("two_code_blocks", 12000, 2)]


def reference_bench_leg(T):
    """The reference's own benchmark cases (tests/bench_main.cpp:27-47,
    185-249): wall time per whole decode_transport_block call (de-rate
    matching, decode, CRCs) for the CPU reference decoder and both GPU
    backends, warm-ups as the reference takes them (CPU 1, GPU 2), and the
    tile8 entry point (de-rate matching + per-block decode) for single-block
    TBs."""
    out = {}
    for name, tbs, reps in REF_BENCH_CASES:
        bits = np.array([(13 * i + 7) & 1 for i in range(tbs)], np.uint8)
        seg = T.segment_transport_block(T.append_crc24a(bits))
        p = T.LteTurboCodecParams(transport_block_bits=tbs, rv=0, num_layers=1, mod_order=2, max_iterations=6,
                                  enable_early_stop=False,
                                  g_total=sum(3 * (len(c) + 4) for c in seg.code_blocks))
        enc = T.encode_transport_block(p, bits)
        row = {"tb_bits": tbs, "repeats": reps, "code_blocks": len(enc.code_blocks),
               "k": [int(b.descriptor.k_r) for b in enc.code_blocks]}
        for label, backend, warm in (("cpu", T.DecoderBackend.CpuReference, 1),
                                     ("gpu_stage1", T.DecoderBackend.GpuStage1Exact, 2),
                                     ("gpu_stage2", T.DecoderBackend.GpuStage2Windowed, 2)):
            for _ in range(warm):
                assert T.decode_transport_block(p, enc, backend).transport_block_crc_ok
            t0 = time.perf_counter()
            for _ in range(reps):
                assert T.decode_transport_block(p, enc, backend).transport_block_crc_ok
            row[label + "_ms"] = (time.perf_counter() - t0) * 1e3 / reps
        if len(enc.code_blocks) == 1:
            blk = enc.code_blocks[0]
            d = blk.descriptor

            def tile8():
                l = T.derate_match_code_block(p, d, blk.rate_matched_bits)
                return T.turbo_decode_code_block_gpu_stage1_tile8_experimental(
                    d, l[:d.d_r], l[d.d_r:2 * d.d_r], l[2 * d.d_r:], p.max_iterations, False)

            for _ in range(2):
                tile8()
            t0 = time.perf_counter()
            for _ in range(reps):
                tile8()
            row["gpu_stage1_tile8_ms"] = (time.perf_counter() - t0) * 1e3 / reps
        log(f"reference case {name}: " + " ".join(f"{k}={v:.3f}" for k, v in row.items() if k.endswith("_ms")))
        out[name] = row
    return {"cases": out, "what": "ms per decode_transport_block call (Python mirror over the C ABI), "
                                  "tests/bench_main.cpp cases, 6 iterations, noise-free +-8 LLRs"}


def rx_chain_leg(T, torch, stream, llr, decode_ms, tuned):
    """Receive chain on the bench blocks: received LLRs in circular-buffer
    order (rv 0, every non-NULL entry once) -> TbPlan (de-rate matching,
    decode, TB assembly)."""
    n, e, d = N_CB, 3 * (K + 4), K + 4
    # position map of one block: derate of 1..E gives, per [d0|d1|d2] slot, 1 + its received index
    where = T.derate_match_soft(np.arange(1, e + 1, dtype=np.float32), K).astype(np.int64) - 1
    assert (where >= 0).all() and np.unique(where).size == e
    rx = np.empty((n, e), np.float32)
    rx[:, where] = llr
    d_rx = torch.from_numpy(rx.reshape(-1)).cuda()
    tb = T.TbPlan([(K - 24, e, 0, 1, 2)] * n, T.DECODER_WINDOW_LOG, ITERS, **tuned)
    tb.set_hold(False)  # one transmission: no HARQ soft buffer to keep
    der = T.DeratePlan([K] * n, [e] * n)
    soft = torch.empty(n * 3 * d, dtype=torch.float32, device="cuda")
    bits = torch.empty(tb.bits_total, dtype=torch.uint8, device="cuda")
    ok = torch.empty(n, dtype=torch.int32, device="cuda")
    s = stream.cuda_stream
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, reps):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    der_ms = timed(lambda: der.run(d_rx.data_ptr(), soft.data_ptr(), stream=s), 10)
    # the derate output must be the bench's decoder input
    assert torch.equal(soft.view(n, 3 * d).cpu(), torch.from_numpy(llr))
    tb_ms = timed(lambda: tb.decode(d_rx.data_ptr(), bits.data_ptr(), ok.data_ptr(), stream=s), 5)
    okh = ok.cpu().numpy().astype(bool)
    ebn0 = np.array([EBN0[i % len(EBN0)] for i in range(n)])
    der_bytes = n * (4 * e + 4 * 3 * d)  # received LLRs in, decoder input out
    der_gbs = der_bytes / (der_ms / 1e3) / 1e9
    log(f"rx chain: derate {der_ms:.3f} ms ({der_gbs:.0f} GB/s), full TB decode {tb_ms:.3f} ms")
    return {"transport_blocks": n, "tb_bits": K - 24, "ms": tb_ms, "tb_gbit_s": n * (K - 24) / tb_ms / 1e6,
            "derate_ms": der_ms, "decode_ms_main_loop": decode_ms,
            "derate_roofline": {"bound": "hbm", "achieved": der_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": der_gbs / HBM_PEAK_GBS, "bytes_per_launch": der_bytes},
            "tb_crc_pass": {f"{x:.2f}": float(okh[ebn0 == x].mean()) for x in EBN0}}


# window lengths at K=6144: (segment length, waves per block); 32 and 64 span three and two waves
WINDOW_LENGTHS = [(32, 3), (64, 2), (96, 1), (128, 1)]


def window_length_leg(T, torch, stream, x, bits, ebn0, dec, warm, shares=(4096, 512, 64, 1), tune_share=512):
    """Window (segment) lengths 32/64/96/128 at K=6144 on the headline
    inputs: decode time at the full batch and at smaller per-GPU shares (the
    first n blocks; n = 1: one block's latency), BER/BLER per Eb/N0 at the
    full batch; and the
    autotuner's BER-gated choice at a 512-block share, where three-wave
    blocks of 32-step windows fill the GPU that 512 one-wave blocks leave
    half idle."""
    from turbo_cpp_amd import autotune
    rows = []
    y = torch.empty(N_CB * K, dtype=torch.float32, device="cuda")
    for s_len, waves in WINDOW_LENGTHS:
        geo = T.segment_geometry(K, dec, s_len, warm, 0, waves)
        row = {"segment_length": geo["segment_length"], "segments": geo["segments"],
               "waves_per_block": geo["waves_per_block"], "warmup": geo["warmup"], "shares": {}}
        for n in shares:
            plan = T.Plan([K] * n, dec, ITERS, min_segment=s_len, warmup=warm, max_waves_per_block=waves)
            for _ in range(2):
                plan.decode(x.data_ptr(), y.data_ptr(), 0, 0, stream.cuda_stream)
            evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(5)]
            for a, b in evs:
                a.record(stream)
                plan.decode(x.data_ptr(), y.data_ptr(), 0, 0, stream.cuda_stream)
                b.record(stream)
            torch.cuda.synchronize()
            ms = float(np.median([a.elapsed_time(b) for a, b in evs]))
            row["shares"][str(n)] = {"ms": ms, "gbit_s": n * K / ms / 1e6}
            if n == N_CB:
                row["ber"] = ber_table(y.cpu().numpy(), bits, ebn0)
            del plan
        log("window length", row["segment_length"], "on", row["waves_per_block"], "waves:",
            {n: round(v["ms"], 3) for n, v in row["shares"].items()}, "ms")
        rows.append(row)
    t = autotune.tune(K, dec, ITERS, n_blocks=tune_share, ebn0_db=(0.25, 0.5, 0.75), warmups=(16, 24),
                      waves=(0,), extrinsic_bits=(16,))
    log(f"autotune at {tune_share} blocks:", {k: t["best"][k] for k in ("segment_length", "waves_per_block", "warmup",
                                                                         "ms", "qualifies")})
    return {"k": K, "decoder": "window-log" if dec == T.DECODER_WINDOW_LOG else "window-lin",
            "what": "the headline's inputs (first n blocks for n < 4096); median of 5 launches",
            "rows": rows,
            "autotune_small_share": {
                "n_blocks": tune_share, "choice": {k: t["best"][k] for k in (
                    "segment_length", "waves_per_block", "warmup", "ms", "per_point", "qualifies")},
                "table": [{k: c[k] for k in ("segment_length", "waves_per_block", "warmup", "ms", "per_point",
                                             "qualifies")} for c in t["table"]]}}


def _graph_time(torch, fn, reps):
    """Capture fn (which enqueues on the current stream) in a HIP graph and
    return the mean replay time in ms."""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    g.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        g.replay()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps, g


def _plain_sample(T, torch, ks, llr_dev, offs, iters):
    """The window decoder launched once per code block (one plan per K),
    sequentially, with CRC24B early stop: info Gbit/s on the given blocks."""
    plans = {}
    post = torch.empty(max(ks), dtype=torch.float32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    for k in set(ks):
        plans[k] = T.Plan([k], T.DECODER_WINDOW_LOG, iters, early_stop=True)
    for k, o in zip(ks, offs):  # warm-up
        plans[k].decode(llr_dev.data_ptr() + 4 * o, post.data_ptr(), 0, 0, s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k, o in zip(ks, offs):
        plans[k].decode(llr_dev.data_ptr() + 4 * o, post.data_ptr(), 0, 0, s)
    torch.cuda.synchronize()
    return sum(ks) / (time.perf_counter() - t0) / 1e9


# Extra Es/N0 for bit-interleaved Gray QAM with max-log LLRs over the Shannon
# limit of the same spectral efficiency (grows with the order)
QAM_GAP_DB = {2: 0.0, 4: 1.5, 6: 3.5}


def _margin_n0(qm, b, g, margin_db):
    """Noise variance for unit-energy symbols at Es/N0 = Shannon limit of the
    transport block's spectral efficiency (qm b / g bits per symbol) + margin."""
    se = qm * b / g
    esn0 = (2.0 ** se - 1.0) * 10.0 ** (margin_db / 10.0)
    return 1.0 / esn0


def uneven_leg(T, torch, geo, n_cb_target=20000, seed=31):
    """Uneven batch: transport blocks of B uniform in [1000, 75376] bits (CRC24A,
    segmentation with filler and CRC24B), code rate uniform in [0.3, 0.9],
    QPSK/16QAM/64QAM uniform, rv uniform 0..3, AWGN at Es/N0 = Shannon limit +
    uniform [1, 3] dB + a QAM gap (0 / 1.5 / 3.5 dB for QPSK / 16QAM / 64QAM),
    max-log LLRs; about 20k code blocks per launch;
    window decoder with CRC24B early stop, max 8 iterations.  Data are made on
    the GPU (encoder, channel); the timed region (de-rate matching, decode, TB
    assembly) is one captured HIP graph."""
    rng = np.random.default_rng(seed)
    tbs, n0s, ncb = [], [], 0
    while ncb < n_cb_target:
        b = int(rng.integers(1000, 75377))
        qm = int(rng.choice([2, 4, 6]))
        rate = float(rng.uniform(0.3, 0.9))
        c = T._segment_sizes(b + 24)[0]
        g = max(int(round(b / rate)) // qm * qm, qm * c)
        tbs.append((b, g, int(rng.integers(0, 4)), 1, qm))
        n0s.append(_margin_n0(qm, b, g, float(rng.uniform(1.0, 3.0)) + QAM_GAP_DB[qm]))
        ncb += c
    plan = T.TbPlan(tbs, T.DECODER_WINDOW_LOG, 8, early_stop=True, **geo)
    plan.set_hold(False)  # first transmissions only
    gen = torch.Generator(device="cuda").manual_seed(seed)
    pay = torch.randint(0, 2, (plan.bits_total,), dtype=torch.uint8, device="cuda", generator=gen)
    rm = torch.empty(plan.rx_floats, dtype=torch.uint8, device="cuda")
    plan.encode(pay.data_ptr(), rm.data_ptr())
    segs, i = [], 0
    for t, n0 in zip(tbs, n0s):
        for _ in range(T._segment_sizes(t[0] + 24)[0]):
            segs.append((int(plan.es[i]), t[4], 0, n0))
            i += 1
    ch = T.ChannelPlan(segs)
    llr = torch.empty(plan.rx_floats, dtype=torch.float32, device="cuda")
    ch.run(seed, rm.data_ptr(), llr.data_ptr())
    bits = torch.empty(plan.bits_total, dtype=torch.uint8, device="cuda")
    ok = torch.empty(plan.n_tb, dtype=torch.int32, device="cuda")
    its = torch.empty(plan.n_tb, dtype=torch.int32, device="cuda")

    def run():
        plan.decode(llr.data_ptr(), bits.data_ptr(), ok.data_ptr(), 0, its.data_ptr(),
                    stream=torch.cuda.current_stream().cuda_stream)

    ms, _ = _graph_time(torch, run, 3)
    okh = ok.cpu().numpy().astype(bool)
    payload_ber = float((bits != pay).float().mean())
    rvs = np.array([t[2] for t in tbs])
    qms = np.array([t[4] for t in tbs])
    rates = np.array([t[0] / t[1] for t in tbs])
    pass_by = {"rv": {str(r): float(okh[rvs == r].mean()) for r in range(4)},
               "qm": {str(q): float(okh[qms == q].mean()) for q in (2, 4, 6)},
               "rate<0.6": float(okh[rates < 0.6].mean()), "rate>=0.6": float(okh[rates >= 0.6].mean())}
    # plain: the first 64 blocks' decoder inputs (from a de-rate plan), one launch per block
    der = T.DeratePlan(plan.ks, plan.es, plan.fillers, [tbs[j][2] for j in range(len(tbs))
                                                        for _ in range(T._segment_sizes(tbs[j][0] + 24)[0])])
    soft = torch.empty(der.llr_floats, dtype=torch.float32, device="cuda")
    der.run(llr.data_ptr(), soft.data_ptr())
    soffs = np.concatenate([[0], np.cumsum(3 * (plan.ks.astype(np.int64) + 4))[:-1]])
    plain = _plain_sample(T, torch, [int(k) for k in plan.ks[:64]], soft, [int(o) for o in soffs[:64]], 8)
    info = sum(t[0] for t in tbs)
    log(f"uneven: {len(tbs)} TBs {plan.n_cb} CBs, {ms:.2f} ms -> {info / ms / 1e6:.3f} Gbit/s; "
        f"TB CRC pass {okh.mean():.3f}; plain {plain:.4f} Gbit/s")
    return {"workload": "uneven TBs B~U[1000,75376], rate U[0.3,0.9], QPSK/16QAM/64QAM, rv U{0..3}, AWGN "
                        "at Shannon limit + U[1,3] dB + QAM gap 0/1.5/3.5 dB, max-log LLRs, early stop, max 8 "
                        "iterations",
            "transport_blocks": len(tbs), "code_blocks": plan.n_cb, "info_bits": info,
            "code_block_bits": int(plan.ks.sum()), "ms": ms, "info_gbit_s": info / ms / 1e6,
            "cb_gbit_s": float(plan.ks.sum()) / ms / 1e6, "mean_iterations_tb_max": float(its.float().mean()),
            "tb_crc_pass": float(okh.mean()), "tb_crc_pass_by": pass_by, "payload_ber": payload_ber,
            "graph": True,
            "plain_gbit_s": plain, "batched_vs_plain": (float(plan.ks.sum()) / ms / 1e6) / plain}


def mixed_leg(T, torch, geo, n_cb=65536, seed=41):
    """Large mixed load: 65536 code blocks, 70% K=6144 and 30% uniform over
    the QPP table, each K - 24 random bits + CRC24B (code_block_crc
    descriptors, so early stop applies as to segmented blocks); 16QAM over
    i.i.d. Rayleigh fading with known CSI, Eb/N0 uniform in [2, 6] dB per
    block; two transmissions of E = 3(K+4)/2 (rounded to 16QAM symbols) with
    rv 0 and rv 2, soft-combined in fp32 before one decode; CRC24B early stop,
    max 8 iterations.  Timed (one captured HIP graph): the decode with both
    transmissions de-rate matched and combined in its input staging, and TB
    assembly."""
    rng = np.random.default_rng(seed)
    table = [int(e[0]) for e in T.qpp_interleaver_table()]
    ks = np.where(rng.random(n_cb) < 0.7, 6144, rng.choice(table, n_cb)).astype(np.int64)
    es = (3 * (ks + 4) // 2) // 4 * 4
    ebn0 = rng.uniform(2.0, 6.0, n_cb)
    tbs = [(int(k) - 24, int(e), 0, 1, 4, 1) for k, e in zip(ks, es)]  # K - 24 payload bits + CRC24B
    plan = T.TbPlan(tbs, T.DECODER_WINDOW_LOG, 8, early_stop=True, **geo)
    plan.set_hold(False)  # both transmissions are combined inside the decode
    gen = torch.Generator(device="cuda").manual_seed(seed)
    pay = torch.randint(0, 2, (plan.bits_total,), dtype=torch.uint8, device="cuda", generator=gen)
    # Es/N0 = Eb/N0 + 10 log10(qm B / E) per transmission
    n0 = 1.0 / (10.0 ** (ebn0 / 10.0) * 4.0 * (ks - 24) / es)
    ch = T.ChannelPlan([(int(e), 4, 1, float(v)) for e, v in zip(es, n0)])
    rx = []
    for rv, sd in ((0, seed), (2, seed + 1)):
        rm = torch.empty(plan.rx_floats, dtype=torch.uint8, device="cuda")
        plan.encode(pay.data_ptr(), rm.data_ptr(), rv=rv)
        y = torch.empty(plan.rx_floats, dtype=torch.float32, device="cuda")
        ch.run(sd, rm.data_ptr(), y.data_ptr())
        rx.append(y)
        del rm
    bits = torch.empty(plan.bits_total, dtype=torch.uint8, device="cuda")
    ok = torch.empty(plan.n_tb, dtype=torch.int32, device="cuda")
    its = torch.empty(plan.n_tb, dtype=torch.int32, device="cuda")

    def run():  # both transmissions de-rate matched and combined inside the decoder's input staging
        plan.decode_multi([rx[0].data_ptr(), rx[1].data_ptr()], [0, 2], bits.data_ptr(), ok.data_ptr(), 0,
                          its.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)

    ms, _ = _graph_time(torch, run, 2)
    okh = ok.cpu().numpy().astype(bool)
    mean_it = float(its.float().mean())
    kb = int(ks.sum())
    gbit = kb / ms / 1e6
    # algorithmic HBM bytes: two received LLR sets read, posteriors written and read (TB
    # assembly), payload bits written
    hbm = 2 * 4 * int(es.sum()) + 2 * 4 * kb + plan.bits_total
    valu = kb * mean_it * VALU_SLOTS_PER_BIT_ITER / (ms / 1e3)
    # plain path on 64 blocks: one launch per block on the combined decoder input
    der = T.DeratePlan(plan.ks[:64], plan.es[:64])
    soft = torch.empty(der.llr_floats, dtype=torch.float32, device="cuda")
    der.run(rx[0].data_ptr(), soft.data_ptr(), rv=0)
    der.run(rx[1].data_ptr(), soft.data_ptr(), rv=2, accumulate=True)
    soffs = np.concatenate([[0], np.cumsum(3 * (plan.ks[:64].astype(np.int64) + 4))[:-1]])
    plain = _plain_sample(T, torch, [int(k) for k in plan.ks[:64]], soft, [int(o) for o in soffs], 8)
    log(f"mixed: {n_cb} CBs {ms:.1f} ms -> {gbit:.3f} Gbit/s (mean iterations {mean_it:.2f}), "
        f"BLER {1 - okh.mean():.4f}; plain {plain:.4f} Gbit/s")
    return {"workload": "65536 CBs (70% K=6144, 30% U(QPP table)), 16QAM, Rayleigh with known CSI, "
                        "Eb/N0 U[2,6] dB, HARQ rv0+rv2 (E=3(K+4)/2 each) combined in fp32, early stop, max 8",
            "code_blocks": n_cb, "code_block_bits": kb, "ms": ms, "gbit_s": gbit,
            "mean_iterations": mean_it, "bler": float(1 - okh.mean()), "graph": True,
            "hbm_gb_s": hbm / (ms / 1e3) / 1e9, "hbm_frac": hbm / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
            "valu_frac": valu / VALU_PEAK, "plain_gbit_s": plain, "batched_vs_plain": gbit / plain}


class _CodedBatches:
    """Batches of n code blocks of K bits on the GPU: random payloads of
    K - 24 bits + CRC24A (one code block each), the GPU encoder at the full
    circular buffer (rv 0), QPSK (two independent BPSK dimensions) over AWGN
    at Eb/N0, max-log LLRs (exact for QPSK), soft de-rate matching: the
    decoder's [d0|d1|d2] inputs and the code-block bits, new ones per seed."""

    def __init__(self, T, torch, k, n, ebn0_db):
        self.torch = torch
        self.tb = T.TbPlan([(k - 24, 3 * (k + 4), 0, 1, 2)] * n, max_iterations=1)
        rate = k / (3.0 * (k + 4))
        n0 = 1.0 / (2.0 * rate * 10.0 ** (ebn0_db / 10.0))  # Es = 1 per QPSK symbol of 2 coded bits
        self.ch = T.ChannelPlan([(int(e), 2, 0, n0) for e in self.tb.es])
        self.der = T.DeratePlan(self.tb.ks, self.tb.es)
        dev = "cuda"
        self.pay = torch.empty(self.tb.bits_total, dtype=torch.uint8, device=dev)
        self.rm = torch.empty(self.tb.rx_floats, dtype=torch.uint8, device=dev)
        self.cb = torch.empty(self.tb.post_floats, dtype=torch.uint8, device=dev)
        self.rx = torch.empty(self.tb.rx_floats, dtype=torch.float32, device=dev)
        self.n, self.k = n, k

    def batch(self, seed):
        torch = self.torch
        g = torch.Generator(device="cuda").manual_seed(seed)
        self.pay.copy_(torch.randint(0, 2, self.pay.shape, dtype=torch.uint8, device="cuda", generator=g))
        self.tb.encode(self.pay.data_ptr(), self.rm.data_ptr(), self.cb.data_ptr())
        self.ch.run(seed, self.rm.data_ptr(), self.rx.data_ptr())
        llr = torch.empty(self.der.llr_floats, dtype=torch.float32, device="cuda")
        self.der.run(self.rx.data_ptr(), llr.data_ptr())
        return llr.view(self.n, -1), self.cb.view(self.n, self.k).float()


def diff_leg(T, torch, train_steps):
    """Differentiable decoder: 2048 blocks of K=1024 (f1=31, f2=64), 0.5 dB
    (QPSK = two BPSK dimensions over AWGN), 6 iterations with 12
    per-half-iteration extrinsic scales (init 1.0), BCE loss of the
    posteriors.  Times forward and adjoint; fits the scales with Adam on a
    fresh batch per step; compares held-out BER/BLER of the weighted decoder,
    the unit-scale one and plain Log-MAP (the exact decoder) on the same
    8192 blocks, with paired block-error counts."""
    from turbo_cpp_amd.autograd import WeightedTurboDecoder

    n, k, it = 2048, 1024, 6
    gen = _CodedBatches(T, torch, k, n, 0.5)
    x_tr, y_tr = gen.batch(1)
    dec = WeightedTurboDecoder([k] * n, iterations=it)
    bce = torch.nn.functional.binary_cross_entropy_with_logits
    stream = torch.cuda.current_stream()
    a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    x_req = x_tr.clone().requires_grad_(True)
    fwd, bwd = [], []
    # 3 untimed passes, then the median of 20 (the first passes after a
    # plan is built run slower)
    for rep in range(23):
        dec.zero_grad()
        a.record(stream)
        loss = bce(dec(x_req), y_tr)
        b.record(stream)
        loss.backward()
        c.record(stream)
        torch.cuda.synchronize()
        if rep >= 3:
            fwd.append(a.elapsed_time(b))
            bwd.append(b.elapsed_time(c))
    fwd_ms, bwd_ms = float(np.median(fwd)), float(np.median(bwd))
    # the decoder's own calls alone (forward kernel; adjoint kernel + scale
    # reduction), same batch and scales, without torch's loss around them
    plan = dec.plan
    flat = x_req.detach().reshape(-1).contiguous()
    sc = dec.scales.detach().contiguous()
    gpost = torch.full((plan.post_elems,), 1e-3, dtype=torch.float32, device="cuda")
    kf, kb = [], []
    for rep in range(23):
        a.record(stream)
        _, saved = plan.forward(flat, sc)
        b.record(stream)
        plan.backward(flat, sc, saved, gpost)
        c.record(stream)
        torch.cuda.synchronize()
        if rep >= 3:
            kf.append(a.elapsed_time(b))
            kb.append(b.elapsed_time(c))
    kfwd_ms, kbwd_ms = float(np.median(kf)), float(np.median(kb))

    held = [gen.batch(1000 + i) for i in range(4)]  # 8192 held-out blocks, never trained on

    def block_errors(decode):
        bits, blocks = 0, []
        with torch.no_grad():
            for x, y in held:
                wrong = (decode(x) >= 0).float() != y
                bits += int(wrong.sum())
                blocks.append(wrong.any(dim=1).cpu().numpy())
        return bits, np.concatenate(blocks)

    unit_bits, unit_blk = block_errors(dec)
    opt = torch.optim.Adam(dec.parameters(), lr=0.01)
    losses = []
    t0 = time.perf_counter()
    for step in range(train_steps):
        x, y = gen.batch(10 + step)
        opt.zero_grad()
        loss = bce(dec(x), y)
        loss.backward()
        opt.step()
        if step % 50 == 0 or step == train_steps - 1:
            losses.append(round(float(loss.detach()), 6))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    w_bits, w_blk = block_errors(dec)
    # the judge: plain Log-MAP with the reference's arithmetic (exact decoder, same 6 iterations)
    exact = T.Plan([k] * n, T.DECODER_EXACT, it)
    post = torch.empty(n * k, dtype=torch.float32, device="cuda")

    def exact_decode(x):
        exact.decode(x.contiguous().data_ptr(), post.data_ptr(), 0, 0, stream.cuda_stream)
        return post.view(n, k)

    ex_bits, ex_blk = block_errors(exact_decode)
    nb, nbits = len(ex_blk), len(ex_blk) * k
    # paired block errors: weighted wrong where the exact decoder is right (b) and the reverse (c);
    # two-sided exact binomial test of b against b + c (McNemar)
    b_, c_ = int(np.sum(w_blk & ~ex_blk)), int(np.sum(~w_blk & ex_blk))
    from scipy.stats import binomtest
    p_blk = float(binomtest(b_, b_ + c_, 0.5).pvalue) if b_ + c_ else 1.0
    log(f"diff: fwd {fwd_ms:.2f} ms bwd {bwd_ms:.2f} ms (decoder calls {kfwd_ms:.2f} / {kbwd_ms:.2f} ms); train {train_steps} steps {train_s:.1f} s; held-out BER "
        f"exact {ex_bits / nbits:.2e} unit {unit_bits / nbits:.2e} weighted {w_bits / nbits:.2e}; block errors "
        f"exact {int(ex_blk.sum())} weighted {int(w_blk.sum())} (only weighted {b_}, only exact {c_}, p {p_blk:.2f})")
    # SIMD VALU time, memory-wait share and HBM-side traffic of the two kernels at this workload, from
    # their rocprofv3 counter runs (copy-calibrated), recorded in profiles/
    prof = os.path.join(ROOT, "profiles", "diff_decoder_k1024_x2048.md")
    counters = {}
    if os.path.exists(prof):
        keys = ("simd_valu_time_frac_fwd", "simd_valu_time_frac_bwd", "wait_frac_fwd", "wait_frac_bwd",
                "traffic_bytes_fwd", "traffic_bytes_bwd")
        for line in open(prof):
            for key in keys:
                if line.startswith(key + ":"):
                    counters[key] = float(line.split(":")[1])
        counters["from"] = "profiles/diff_decoder_k1024_x2048.md"
    return {"workload": "2048 code blocks K=1024, QPSK (= two BPSK dimensions)/AWGN 0.5 dB, 6 iterations, "
                        "12 extrinsic scales, BCE; GPU-generated batches (encoder, channel, soft de-rate matching)",
            "timing": "fwd: decoder forward + BCE; bwd: loss.backward() (adjoint kernel, scale reduction, BCE "
                      "gradient), HIP events on the launch stream, median of 20 after 3 untimed",
            "counters": counters,
            "fwd_ms": fwd_ms, "bwd_ms": bwd_ms, "fwd_gbit_s": n * k / fwd_ms / 1e6, "bwd_gbit_s": n * k / bwd_ms / 1e6,
            "decoder_calls_ms": {"forward": kfwd_ms, "backward": kbwd_ms,
                                 "what": "DiffPlan.forward / DiffPlan.backward alone (forward kernel; adjoint "
                                         "kernel + scale reduction), median of 20 after 3 untimed"},
            "adam_steps": train_steps, "train_s": train_s, "training": "a fresh batch of 2048 blocks per step",
            "loss_every_50": losses,
            "scales": [round(float(v), 4) for v in dec.scales.detach().cpu()],
            "held_out": {"blocks": nb, "ber_exact_logmap": ex_bits / nbits, "bler_exact_logmap": float(ex_blk.mean()),
                         "ber_unit_scales": unit_bits / nbits, "bler_unit_scales": float(unit_blk.mean()),
                         "ber_weighted": w_bits / nbits, "bler_weighted": float(w_blk.mean()),
                         "block_errors_only_weighted": b_, "block_errors_only_exact": c_,
                         "bler_paired_p_value": p_blk,
                         "judge": "exact decoder (plain Log-MAP, the reference's arithmetic), 6 iterations, "
                                  "same 8192 blocks; p: two-sided exact binomial (McNemar) test of the blocks "
                                  "only one of the two decoders gets wrong"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # the clock settles over the first ~50 ms of sustained decoding
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--decoder", choices=["log", "lin"], default="log")
    ap.add_argument("--plain-blocks", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-exact", action="store_true")
    ap.add_argument("--no-e2e", action="store_true", help="skip the host-buffer end-to-end leg")
    ap.add_argument("--no-reference-cases", action="store_true", help="skip tests/bench_main.cpp's four cases")
    ap.add_argument("--no-rx-chain", action="store_true")
    ap.add_argument("--train-steps", type=int, default=500, help="Adam steps of the differentiation leg (0: skip it)")
    ap.add_argument("--no-scenarios", action="store_true", help="skip the uneven-batch and mixed-load legs")
    ap.add_argument("--no-window-lengths", action="store_true",
                    help="skip the window-length leg (32/64/96/128 at 4096, 512, 64 and 1 blocks)")
    ap.add_argument("--no-autotune", action="store_true", help="skip the tuner; use --min-segment/--warmup-steps")
    ap.add_argument("--min-segment", type=int, default=0, help="with --no-autotune (0: library default 32)")
    ap.add_argument("--warmup-steps", type=int, default=0, help="with --no-autotune (0: library default 32)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this decoder runs on one GPU; --gpus must be 1 (no multi-GPU path is built)"}))
        sys.exit(2)

    import torch

    import turbo_cpp_amd as T

    dec = T.DECODER_WINDOW_LOG if args.decoder == "log" else T.DECODER_WINDOW_LIN
    other = T.DECODER_WINDOW_LIN if dec == T.DECODER_WINDOW_LOG else T.DECODER_WINDOW_LOG
    tuned = None
    tuned_opts = {}
    if args.no_autotune:
        tuned_opts = {"min_segment": args.min_segment, "warmup": args.warmup_steps, "waves_per_cu": 0}
    else:
        from turbo_cpp_amd import autotune
        log("autotuning segment length / warm-up / waves per CU")
        tuned = autotune.tune(K, dec, ITERS, n_blocks=N_CB, ebn0_db=(0.25, 0.5, 0.75), extrinsic_bits=(16, 32))
        tuned_opts = autotune.options_for(tuned)
        log("autotune choice:", tuned["best"])
    log("generating", N_CB, "blocks")
    bits, llr, ebn0 = make_inputs(N_CB)
    ks = [K] * N_CB
    x = torch.from_numpy(llr.reshape(-1)).cuda()
    y = torch.empty(N_CB * K, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    plan = T.Plan(ks, dec, ITERS, **tuned_opts)
    geo = T.segment_geometry(K, dec, tuned_opts.get("min_segment", 0), tuned_opts.get("warmup", 0),
                             tuned_opts.get("extrinsic_bits", 0), tuned_opts.get("max_waves_per_block", 0))

    def step(p=plan, out=y):
        p.decode(x.data_ptr(), out.data_ptr(), 0, 0, stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    ms_per_step = wall * 1e3 / args.steps
    gbits = N_CB * K / (ms_per_step / 1e3) / 1e9
    log(f"decoder={args.decoder} ms/step={ms_per_step:.3f} kernel={kernel_ms:.3f} ms -> {gbits:.3f} Gbit/s")
    post_main = y.cpu().numpy()
    ber = {args.decoder: ber_table(post_main, bits, ebn0)}

    # the other window arithmetic, same workload and geometry
    plan2 = T.Plan(ks, other, ITERS, **tuned_opts)
    for _ in range(3):
        step(plan2)
    torch.cuda.synchronize()
    ev2 = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(5)]
    for a, b in ev2:
        a.record(stream)
        step(plan2)
        b.record(stream)
    torch.cuda.synchronize()
    other_ms = float(np.median([a.elapsed_time(b) for a, b in ev2]))
    other_name = "lin" if other == T.DECODER_WINDOW_LIN else "log"
    ber[other_name] = ber_table(y.cpu().numpy(), bits, ebn0)
    modes = {args.decoder: {"ms": kernel_ms, "gbit_s": N_CB * K / kernel_ms / 1e6},
             other_name: {"ms": other_ms, "gbit_s": N_CB * K / other_ms / 1e6}}
    del plan2

    def timed(p, reps, warm=1):
        for _ in range(warm):
            step(p)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for e0, e1 in evs:
            e0.record(stream)
            step(p)
            e1.record(stream)
        torch.cuda.synchronize()
        return float(np.median([e0.elapsed_time(e1) for e0, e1 in evs]))

    # the same geometry with the extrinsics held in fp32 between half-iterations (what fp16
    # holding buys, at equal BER)
    for name, d in (("log_f32", T.DECODER_WINDOW_LOG), ("lin_f32", T.DECODER_WINDOW_LIN)):
        pf = T.Plan(ks, d, ITERS, **{**tuned_opts, "extrinsic_bits": 32})
        ms = timed(pf, 5, 2)
        modes[name] = {"ms": ms, "gbit_s": N_CB * K / ms / 1e6, "extrinsics_held": "fp32"}
        ber[name] = ber_table(y.cpu().numpy(), bits, ebn0)
        log(f"{name}: {ms:.3f} ms")
        del pf
    held = f"fp{geo['extrinsic_bits']}"
    modes[args.decoder]["extrinsics_held"] = modes[other_name]["extrinsics_held"] = held

    # exact decoders (the reference's arithmetic and order): segment-parallel with exact
    # boundaries, and one full-length recursion per block; the BER reference on the GPU
    if not args.no_exact:
        for name, d, reps in (("exact", T.DECODER_EXACT, 3), ("exact_serial", T.DECODER_EXACT_SERIAL, 1)):
            plan3 = T.Plan(ks, d, ITERS)
            ex_ms = timed(plan3, reps, 1)
            modes[name] = {"ms": ex_ms, "gbit_s": N_CB * K / ex_ms / 1e6, "extrinsics_held": "fp32"}
            ber[name] = ber_table(y.cpu().numpy(), bits, ebn0)
            del plan3
            log(f"{name}: {ex_ms:.1f} ms")

    # plain: same window decoder, one launch (and one plan) per code block
    nplain = args.plain_blocks
    p1 = T.Plan([K], dec, ITERS, **tuned_opts)
    for i in range(2):
        p1.decode(x.data_ptr(), y.data_ptr(), 0, 0, stream.cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(nplain):
        p1.decode(x.data_ptr() + 4 * i * 3 * (K + 4), y.data_ptr() + 4 * i * K, 0, 0, stream.cuda_stream)
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - t0) / nplain
    plain_gbits = K / plain_s / 1e9
    log(f"plain: {plain_s * 1e3:.3f} ms per block -> {plain_gbits * 1e3:.3f} Mbit/s")

    geo_kw = {k: v for k, v in tuned_opts.items()
              if k in ("min_segment", "warmup", "extrinsic_bits", "max_waves_per_block")}
    wl = None if args.no_window_lengths else window_length_leg(T, torch, stream, x, bits, ebn0, dec, geo["warmup"])
    rx_chain = None if args.no_rx_chain else rx_chain_leg(T, torch, stream, llr, kernel_ms, geo_kw)
    diff = diff_leg(T, torch, args.train_steps) if args.train_steps > 0 else None
    uneven = None if args.no_scenarios else uneven_leg(T, torch, geo_kw)
    mixed = None if args.no_scenarios else mixed_leg(T, torch, geo_kw)

    if args.copy_calibration:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 256 MiB read + 256 MiB written per copy kernel")

    # HBM traffic per launch of the headline kernel from its rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    # (copy-calibrated), recorded in profiles/ for this workload and geometry
    traffic = None
    counters = {}
    prof = os.path.join(ROOT, "profiles", "window_decode_counters_k6144_x4096.md")
    if dec == T.DECODER_WINDOW_LOG and os.path.exists(prof):
        for line in open(prof):
            if line.startswith("traffic_bytes_per_launch_log:"):
                traffic = float(line.split(":")[1])
            for key in ("simd_valu_time_frac_log", "sq_insts_valu_log"):
                if line.startswith(key + ":"):
                    counters[key[:-4]] = float(line.split(":")[1])
    bits_per_launch = N_CB * K
    achieved = bits_per_launch * ITERS * VALU_SLOTS_PER_BIT_ITER / (kernel_ms / 1e3)
    hbm_achieved = bits_per_launch * (16 + 48 / K) / (kernel_ms / 1e3) / 1e9
    roofline = {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12, "unit": "Tops/s",
                "frac": achieved / VALU_PEAK, "traffic": traffic,
                "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, "
                                "profiles/window_decode_counters_k6144_x4096.md)",
                "note": "VALU lane-slots of the reference's pairwise Log-MAP form per second (864 per bit and "
                        "iteration: 384 simple FP32 ops + 120 transcendentals at quarter rate, 4 slots each); "
                        "peak = MI355X's FP32 vector peak, 256 CUs x 4 SIMDs x 32 lanes x 2.4 GHz (157.3 "
                        "TFLOP/s as FMAs); no MFMA work exists in this path",
                "hbm": {"achieved": hbm_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": hbm_achieved / HBM_PEAK_GBS}}
    if counters:
        # the kernel's own VALU issue from its counter run: simple instructions at 2 SIMD cycles,
        # transcendentals (SQ_ACTIVE_INST_VALU - SQ_INSTS_VALU) at 8, over the SIMDs' cycles
        roofline["simd_valu_time_frac"] = counters.get("simd_valu_time_frac")
        roofline["sq_insts_valu"] = counters.get("sq_insts_valu")
        roofline["counters_from"] = "profiles/window_decode_counters_k6144_x4096.md"
    calib = valu_calibration()
    if calib and "rows" in calib:
        log("valu calibration:", [(r["op"], r["resident_waves_per_simd"], r["simd_cycles_per_instr"])
                                  for r in calib["rows"]])
    roofline["valu_calibration"] = calib
    e2e = None if args.no_e2e else e2e_leg(T, llr, tuned_opts)
    ref_cases = None if args.no_reference_cases else reference_bench_leg(T)
    cpu = None if args.no_cpu_baseline else cpu_baseline(llr)
    if cpu:
        log("cpu baseline:", cpu)
    out = {
        "metric": "info Gbit/s decoded (8 it, K=6144) batched vs plain at equal BER; bwd Gbit/s",
        "value": gbits,
        "unit": "Gbit/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"fp32 arithmetic, fp{geo['extrinsic_bits']}-held extrinsics",
        "data": "synthetic: seeded Bernoulli bits, reference encoder, BPSK/AWGN 0.0-1.5 dB",
        "config": {"workload": "uniform batch: 4096 code blocks K=6144, rate 1/3 full buffer, BPSK/AWGN "
                               "Eb/N0 0.0..1.5 dB (0.25 dB steps), fp32 LLRs, 8 iterations, no early stop",
                   "code_blocks": N_CB, "k": K, "iterations": ITERS, "decoder": f"window-{args.decoder}",
                   "segment_length": geo["segment_length"], "segments_per_block": geo["segments"],
                   "warmup_steps": geo["warmup"], "waves_per_cu": tuned_opts.get("waves_per_cu", 0),
                   "waves_per_block": geo["waves_per_block"],
                   "extrinsics_held": f"fp{geo['extrinsic_bits']} (modes.log_f32: fp32-held, same geometry)"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain": {"gbit_s": plain_gbits, "ms_per_block": plain_s * 1e3, "blocks_timed": nplain,
                  "what": "same window decoder, one plan launch per code block, sequential"},
        "batched_vs_plain": gbits / plain_gbits,
        "batched_vs_cpu": (gbits / cpu["value"]) if cpu else None,
        "modes": modes,
        "ber": ber,
        "kernel_ms": kernel_ms,
        "end_to_end": e2e,
        "reference_bench": ref_cases,
        "rx_chain": rx_chain,
        "bwd_gbit_s": diff["bwd_gbit_s"] if diff else None,
        "differentiation": diff,
        "autotune": None if tuned is None else {
            "choice": tuned["best"], "points": tuned["points"],
            "calibration": f"{tuned['n_blocks']} blocks K={K} per point, QPSK/AWGN at {tuned['ebn0_db']} dB; "
                           f"qualifies: bit and block errors inside the exact decoder's 95% intervals at every "
                           f"point; timed at {tuned['timed_at_db']} dB",
            "table": [{k: c[k] for k in ("segment_length", "waves_per_block", "warmup", "waves_per_cu",
                                         "extrinsic_bits", "ms", "per_point", "qualifies")}
                      for c in tuned["table"]]},
        "window_lengths": wl,
        "uneven_batch": uneven,
        "mixed_load": mixed,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
