"""Receive chain on the GPU: soft de-rate matching and HARQ combining.

Bar: bit-exact.  +-8 inputs must reproduce the reference's de-rate matcher
(lte_turbo.cpp:280-325) from the golden vectors; float inputs must equal the
oracle's soft restatement (or_derate_match_soft) bit for bit, including
repetition (E > N_valid), puncturing, every rv and accumulation."""
import numpy as np
import pytest

import oracle as O
import turbo_cpp_amd as T
from tests.helpers import assert_llr_parity, golden

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_soft_derate_pm8_matches_reference_golden():
    g = golden()
    for ci, tbs, g_total, rv, r, k, e, f in g["tb_meta"]:
        key = f"tb{ci}_rv{rv}_cb{r}"
        _say("derate", key)
        rx = np.where(g[key + "_rm"] > 0, 8.0, -8.0).astype(np.float32)
        out = T.derate_match_soft(rx, int(k), int(f), int(rv))
        assert np.array_equal(_bits(out), _bits(g[key + "_derm"])), key


def _mixed_batch(seed):
    rng = np.random.default_rng(seed)
    table = [int(e[0]) for e in T.qpp_interleaver_table()]
    ks = [6144, 40, 1024, 6144, 544, 2048, 6080] + [int(rng.choice(table)) for _ in range(9)]
    n = len(ks)
    fs = [int(rng.integers(0, min(k, 64))) if i % 3 == 0 else 0 for i, k in enumerate(ks)]
    es = []
    for k, f in zip(ks, fs):
        nv = 3 * (k + 4) - 2 * f
        es.append(int(rng.choice([nv, nv // 3 + 1, nv + 17, 2 * nv + 5, 100])))
    rvs = [int(x) for x in rng.integers(0, 4, n)]
    return ks, fs, es, rvs, rng


def test_soft_derate_plan_mixed_batch_matches_oracle():
    import torch
    ks, fs, es, rvs, rng = _mixed_batch(3)
    rx = rng.normal(0.0, 4.0, sum(es)).astype(np.float32)
    plan = T.DeratePlan(ks, es, fs, rvs)
    assert plan.rx_floats == sum(es) and plan.llr_floats == sum(3 * (k + 4) for k in ks)
    d_rx = torch.from_numpy(rx).cuda()
    d_llr = torch.full((plan.llr_floats,), 123.0, dtype=torch.float32, device="cuda")
    _say("derate plan", len(ks), "blocks")
    plan.run(d_rx.data_ptr(), d_llr.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    got = d_llr.cpu().numpy()
    ioff = ooff = 0
    for k, f, e, rv in zip(ks, fs, es, rvs):
        ref = O.derate_match_soft(rx[ioff:ioff + e], k + 4, f, rv)
        seg = got[ooff:ooff + 3 * (k + 4)]
        assert np.array_equal(_bits(seg), _bits(ref)), (k, f, e, rv)
        ioff += e
        ooff += 3 * (k + 4)


def test_soft_derate_harq_combining_matches_oracle():
    """rv 0 then rv 2 accumulated on the device, against the oracle."""
    import torch
    ks, fs, es, _, rng = _mixed_batch(7)
    rx0 = rng.normal(0.0, 3.0, sum(es)).astype(np.float32)
    rx2 = rng.normal(0.0, 3.0, sum(es)).astype(np.float32)
    plan = T.DeratePlan(ks, es, fs)
    d_llr = torch.empty(plan.llr_floats, dtype=torch.float32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    _say("harq rv0+rv2")
    plan.run(torch.from_numpy(rx0).cuda().data_ptr(), d_llr.data_ptr(), rv=0, stream=s)
    plan.run(torch.from_numpy(rx2).cuda().data_ptr(), d_llr.data_ptr(), rv=2, accumulate=True, stream=s)
    got = d_llr.cpu().numpy()
    ioff = ooff = 0
    for k, f, e in zip(ks, fs, es):
        ref = O.derate_match_soft(rx0[ioff:ioff + e], k + 4, f, 0)
        ref = O.derate_match_soft(rx2[ioff:ioff + e], k + 4, f, 2, out=ref)
        assert np.array_equal(_bits(got[ooff:ooff + 3 * (k + 4)]), _bits(ref)), (k, f, e)
        ioff += e
        ooff += 3 * (k + 4)
    # the host-buffer entry point accumulates the same way
    k, f, e = ks[0], fs[0], es[0]
    host = T.derate_match_soft(rx0[:e], k, f, 0)
    host = T.derate_match_soft(rx2[:e], k, f, 2, out=host)
    assert np.array_equal(_bits(host), _bits(got[:3 * (k + 4)]))


def test_soft_derate_feeds_decoder_plan():
    """De-rate matching and decoding chained on one stream on device buffers
    give the same posteriors as the host path (host de-rate + batched decode)."""
    import torch
    rng = np.random.default_rng(11)
    ks = [6144, 1024, 40]
    es = [12000, 4000, 150]
    cbs = [rng.integers(0, 2, k).astype(np.uint8) for k in ks]
    rx = []
    for cb, k, e in zip(cbs, ks, es):
        d = T.encode_code_block_streams(cb)
        bits = O.rate_match(*d, 0, e, 0)
        y = (2.0 * bits - 1.0) + rng.normal(0.0, 0.8, e)
        rx.append((2.0 * y / 0.64).astype(np.float32))
    rx = np.concatenate(rx)
    dp = T.DeratePlan(ks, es)
    pl = T.Plan(ks, decoder=T.DECODER_WINDOW_LOG, max_iterations=6)
    s = torch.cuda.current_stream().cuda_stream
    d_rx = torch.from_numpy(rx).cuda()
    d_llr = torch.empty(dp.llr_floats, dtype=torch.float32, device="cuda")
    d_post = torch.empty(pl.post_floats, dtype=torch.float32, device="cuda")
    _say("derate -> decode chain")
    dp.run(d_rx.data_ptr(), d_llr.data_ptr(), stream=s)
    pl.decode(d_llr.data_ptr(), d_post.data_ptr(), stream=s)
    got = d_post.cpu().numpy()
    llr = np.concatenate([O.derate_match_soft(rx[sum(es[:i]):sum(es[:i + 1])], k + 4, 0, 0)
                          for i, k in enumerate(ks)])
    ref, _, _ = T.decode_batch_host(ks, llr, T.DECODER_WINDOW_LOG, 6)
    assert np.array_equal(_bits(got), _bits(ref))
    hard = (got >= 0).astype(np.uint8)
    assert np.array_equal(hard, np.concatenate(cbs))


def test_soft_derate_errors():
    p = T.DeratePlan([40], [10])
    with pytest.raises(ValueError):
        p.run(0, 0, rv=4)
    with pytest.raises(ValueError):
        p.run(0, 0)  # null device pointers


# ------------------------------------------------------ transport blocks ---
@pytest.mark.parametrize("ci", range(6))
def test_tb_pipeline_matches_reference_golden(ci):
    """Device receive chain with the exact decoder against the reference's
    decode_transport_block on corrupted TBs: CRC flags and iterations equal.
    On blocks that converge (their CRC passes in the reference) posteriors
    are within 1e-4 absolute plus 1e-6 relative (+-8 inputs drive |LLR| to
    ~350, where one float32 ulp is 3e-5 and the GPU's exp/log differ from
    glibc's by a few ulps) and payload bits are equal except where the
    reference's |LLR| < 1e-3.  A block that does not converge is chaotic in
    the reference itself (a 1-ulp change of one input LLR moves its
    posteriors by ~2e-4), so there only the CRC outcome is compared."""
    g = golden()
    _, tbs, g_total, it, es, c, tb_ok_ref, used_ref = (int(x) for x in g["tbd_meta"][ci])
    _say("tb golden", ci, tbs, c)
    rx = np.where(g[f"tbd{ci}_rm"] > 0, 8.0, -8.0).astype(np.float32)
    bits, tb_ok, cb_ok, its, post = T.decode_transport_blocks([(tbs, g_total, 0, 1, 2)], rx, T.DECODER_EXACT, it,
                                                              bool(es))
    ref_post = g[f"tbd{ci}_post"]
    ref_cb_ok = g[f"tbd{ci}_cb_ok"].astype(bool)
    assert bool(tb_ok[0]) == bool(tb_ok_ref) and int(its[0]) == used_ref
    assert np.array_equal(cb_ok, ref_cb_ok)
    ks, fs = g[f"tbd{ci}_k"], g[f"tbd{ci}_f"]
    src, conv, off = [], [], 0  # payload bit j of the TB comes from posterior src[j]
    for r, (k, f) in enumerate(zip(ks, fs)):
        converged = ref_cb_ok[r] if c > 1 else bool(tb_ok_ref)
        n = int(k) - int(f) - (24 if c > 1 else 0)
        src.append(off + int(f) + np.arange(n))
        conv.append(np.full(n, converged))
        if converged:
            assert_llr_parity(post[off:off + int(k)], ref_post[off:off + int(k)], atol=1e-4, rtol=1e-6)
        off += int(k)
    src, conv = np.concatenate(src)[:tbs], np.concatenate(conv)[:tbs]
    firm = (np.abs(ref_post[src]) >= 1e-3) & conv
    assert np.array_equal(bits[0][firm], g[f"tbd{ci}_bits"][firm])


def _tb_batch(seed, n_tb=24):
    rng = np.random.default_rng(seed)
    tbs, rx_all, payloads = [], [], []
    sizes = [16, 40, 328, 1000, 6112, 6200, 12000, 20000, 30000]
    for i in range(n_tb):
        b = int(rng.choice(sizes))
        seg = T._segment_sizes(b + 24)
        c = seg[0]
        ks = [seg[4]] * seg[2] + [seg[3]] * seg[1]
        nv = sum(3 * (k + 4) for k in ks) - 2 * seg[5]
        g_total = int(nv * rng.choice([0.5, 0.8, 1.0, 1.3])) // 2 * 2
        rv = int(rng.integers(0, 4))
        payload = rng.integers(0, 2, b).astype(np.uint8)
        p = T.LteTurboCodecParams(transport_block_bits=b, g_total=g_total, rv=rv)
        enc = T.encode_transport_block(p, payload)
        snr_bad = (i % 5 == 4)
        for blk in enc.code_blocks:
            x = 2.0 * blk.rate_matched_bits.astype(np.float64) - 1.0
            sigma = 1.2 if snr_bad else 0.7
            y = x + sigma * rng.standard_normal(len(x))
            rx_all.append((2.0 * y / sigma ** 2).astype(np.float32))
        tbs.append((b, g_total, rv, 1, 2))
        payloads.append(payload)
    return tbs, np.concatenate(rx_all), payloads


@pytest.mark.parametrize("decoder,early,waves,staging", [
    (T.DECODER_WINDOW_LOG, False, 1, None), (T.DECODER_WINDOW_LIN, True, 1, None), (T.DECODER_EXACT, True, 1, None),
    (T.DECODER_WINDOW_LOG, True, 3, "fused"), (T.DECODER_WINDOW_LOG, False, 3, "separate"),
    (T.DECODER_WINDOW_LIN, True, 2, "fused")])
def test_tb_pipeline_batch_matches_oracle_assembly(decoder, early, waves, staging):
    """Many TBs of mixed sizes, rv and code rates in one plan: the device
    assembly equals the oracle's (lte_turbo.cpp:958-979) on the same
    posteriors bit for bit, and the posteriors equal the batched decoder's on
    the oracle's soft de-rate matching.  waves > 1: segment length from 32 on
    jobs of 2 or 3 waves (the K=6144 blocks span them, shorter blocks are
    packed across them; the TB assembly runs on the first wave), with de-rate
    matching fused into the decoder's staging or in separate kernels."""
    import torch
    tbs, rx, payloads = _tb_batch(21 + decoder + (10 * waves if waves > 1 else 0))
    geo = dict(min_segment=32, max_waves_per_block=waves) if waves > 1 else {}
    plan = T.TbPlan(tbs, decoder=decoder, max_iterations=6, early_stop=early, **geo)
    if staging:
        plan.set_staging(T.STAGING_FUSED if staging == "fused" else T.STAGING_SEPARATE)
    assert plan.rx_floats == rx.size
    dev = dict(rx=torch.from_numpy(rx).cuda(),
               bits=torch.zeros(plan.bits_total, dtype=torch.uint8, device="cuda"),
               tb=torch.full((plan.n_tb,), 7, dtype=torch.int32, device="cuda"),
               cb=torch.full((plan.n_cb,), 7, dtype=torch.int32, device="cuda"),
               it=torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda"),
               post=torch.zeros(plan.post_floats, dtype=torch.float32, device="cuda"))
    _say("tb batch", decoder, early, plan.n_tb, "TBs", plan.n_cb, "CBs")
    plan.decode(dev["rx"].data_ptr(), dev["bits"].data_ptr(), dev["tb"].data_ptr(), dev["cb"].data_ptr(),
                dev["it"].data_ptr(), dev["post"].data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    h = {k: v.cpu().numpy() for k, v in dev.items()}
    # decoder input from the oracle, decoded by the same batched decoder
    llr, ioff = [], 0
    for k, f, e, rv in zip(plan.ks, plan.fillers, plan.es, np.repeat([t[2] for t in tbs],
                                                                     [T._segment_sizes(t[0] + 24)[0] for t in tbs])):
        llr.append(O.derate_match_soft(rx[ioff:ioff + int(e)], int(k) + 4, int(f), int(rv)))
        ioff += int(e)
    post_ref, _, its_ref = T.decode_batch_host(plan.ks, np.concatenate(llr), decoder, 6, early,
                                               fillers=plan.fillers, **geo)
    if waves > 1:
        assert max(T.segment_geometry(int(k), decoder, 32, 0, 0, waves)["waves_per_block"] for k in plan.ks) == waves
    assert np.array_equal(h["post"].view(np.uint32), post_ref.view(np.uint32))
    boff = coff = poff = 0
    n_ok = 0
    for i, t in enumerate(tbs):
        c = T._segment_sizes(t[0] + 24)[0]
        ks, fs = plan.ks[coff:coff + c], plan.fillers[coff:coff + c]
        bits, cb_ok, tb_ok = O.assemble_tb(h["post"][poff:poff + int(ks.sum())], ks, fs, t[0])
        assert np.array_equal(h["bits"][boff:boff + t[0]], bits), i
        assert np.array_equal(h["cb"][coff:coff + c].astype(bool), cb_ok), i
        assert bool(h["tb"][i]) == tb_ok, i
        assert h["it"][i] == its_ref[coff:coff + c].max(), i
        n_ok += tb_ok and np.array_equal(bits, payloads[i])
        boff += t[0]
        coff += c
        poff += int(ks.sum())
    assert n_ok >= len(tbs) // 2  # most TBs are decodable at these SNRs


def test_tb_pipeline_harq_combining():
    """rv 0 then rv 2 accumulated in the plan's soft buffer equals decoding
    the oracle's combined de-rate matching; the second transmission rescues
    TBs the first could not decode."""
    import torch
    rng = np.random.default_rng(99)
    b, g_total = 12000, 14000  # rate ~0.86 per transmission
    payload = rng.integers(0, 2, b).astype(np.uint8)
    rxs = []
    for rv in (0, 2):
        enc = T.encode_transport_block(T.LteTurboCodecParams(transport_block_bits=b, g_total=g_total, rv=rv),
                                       payload)
        x = np.concatenate([2.0 * blk.rate_matched_bits - 1.0 for blk in enc.code_blocks])
        y = x + 0.8 * rng.standard_normal(len(x))
        rxs.append((2.0 * y / 0.64).astype(np.float32))
    plan = T.TbPlan([(b, g_total, 0, 1, 2)], max_iterations=8)
    s = torch.cuda.current_stream().cuda_stream
    bits = torch.zeros(b, dtype=torch.uint8, device="cuda")
    tb = torch.zeros(1, dtype=torch.int32, device="cuda")
    post = torch.zeros(plan.post_floats, dtype=torch.float32, device="cuda")
    _say("tb harq")
    plan.decode(torch.from_numpy(rxs[0]).cuda().data_ptr(), bits.data_ptr(), tb.data_ptr(), rv=0, stream=s)
    first_ok = bool(tb.cpu()[0])
    plan.decode(torch.from_numpy(rxs[1]).cuda().data_ptr(), bits.data_ptr(), tb.data_ptr(), post_ptr=post.data_ptr(),
                rv=2, accumulate=True, stream=s)
    torch.cuda.synchronize()
    llr, ioff = [], 0
    for k, f, e in zip(plan.ks, plan.fillers, plan.es):
        l0 = O.derate_match_soft(rxs[0][ioff:ioff + int(e)], int(k) + 4, int(f), 0)
        llr.append(O.derate_match_soft(rxs[1][ioff:ioff + int(e)], int(k) + 4, int(f), 2, out=l0))
        ioff += int(e)
    post_ref, _, _ = T.decode_batch_host(plan.ks, np.concatenate(llr), T.DECODER_WINDOW_LOG, 8,
                                         fillers=plan.fillers)
    assert np.array_equal(post.cpu().numpy().view(np.uint32), post_ref.view(np.uint32))
    assert bool(tb.cpu()[0]) and np.array_equal(bits.cpu().numpy(), payload)
    _say("first transmission alone decoded:", first_ok)
    # combine both transmissions without decoding, then decode the held input once
    post2 = torch.zeros_like(post)
    plan.soft_combine(torch.from_numpy(rxs[0]).cuda().data_ptr(), rv=0, stream=s)
    plan.soft_combine(torch.from_numpy(rxs[1]).cuda().data_ptr(), rv=2, accumulate=True, stream=s)
    plan.decode(0, bits.data_ptr(), tb.data_ptr(), post_ptr=post2.data_ptr(), stream=s)
    torch.cuda.synchronize()
    assert torch.equal(post2, post)



# ------------------------------------------------------- transmit side ---
def _random_tbs(seed, n):
    rng = np.random.default_rng(seed)
    tbs, pays = [], []
    for i in range(n):
        b = int(rng.choice([16, 40, 100, 512, 1000, 6120, 6200, 12000, 23456, 40000, 75376]))
        c = T._segment_sizes(b + 24)
        ks = [c[4]] * c[2] + [c[3]] * c[1]
        qm = int(rng.choice([2, 4, 6]))
        nl = int(rng.choice([1, 2]))
        nv = sum(3 * (k + 4) for k in ks)
        g = int(nv * rng.uniform(0.35, 1.6)) // (qm * nl) * (qm * nl)
        tbs.append((b, max(g, qm * nl * c[0]), int(rng.integers(0, 4)), nl, qm))
        pays.append(rng.integers(0, 2, b).astype(np.uint8))
    return tbs, pays


def test_gpu_encoder_matches_reference_golden():
    """tc_tb_plan_encode reproduces the reference's encode_transport_block
    (lte_turbo.cpp:690-720: CRC24A, segmentation with filler and CRC24B, the
    encoder with its tail multiplexing, rate matching) bit for bit on the
    golden transport blocks: 8 sizes (16 .. 12000 bits, fillers, C = 2,
    punctured and full-rate g_total) x rv 0..3, rate-matched bits of every code
    block and the code-block bits."""
    g = golden()
    meta = g["tb_meta"]
    cases = sorted({(int(r[0]), int(r[1]), int(r[2]), int(r[3])) for r in meta})
    tbs = [(b, gt, rv, 1, 2) for _, b, gt, rv in cases]
    pays = [g[f"tb{ci}_bits"] for ci, _, _, _ in cases]
    _say("gpu encoder vs golden:", len(tbs), "TBs")
    rm, cb = T.encode_transport_blocks(tbs, pays)
    roff = coff = 0
    for ci, b, gt, rv in cases:
        blocks = sorted(int(r[4]) for r in meta if (int(r[0]), int(r[3])) == (ci, rv))
        for r in blocks:
            want = g[f"tb{ci}_rv{rv}_cb{r}_rm"]
            assert np.array_equal(rm[roff:roff + want.size], want), (ci, rv, r)
            roff += want.size
            cbb = g[f"tb{ci}_rv0_cb{r}_cbbits"]
            assert np.array_equal(cb[coff:coff + cbb.size], cbb), (ci, rv, r)
            coff += cbb.size
    assert roff == rm.size and coff == cb.size


def test_gpu_encoder_matches_oracle_chain():
    """The same on 40 random TBs (16 .. 75376 bits, so up to 13 code blocks;
    Qm 2/4/6, 1-2 layers, rv 0-3, E above and below the circular buffer)
    against the oracle's chain: or_crc_append(CRC24A) -> or_segment ->
    or_encode_cb -> or_compute_e / or_rate_match (lte_turbo.cpp:104-325,
    540-720)."""
    tbs, pays = _random_tbs(5, 40)
    assert max(t[0] for t in tbs) == 75376
    _say("gpu encoder vs oracle:", len(tbs), "TBs")
    rm, cb = T.encode_transport_blocks(tbs, pays)
    roff = coff = 0
    for (b, gt, rv, nl, qm), pay in zip(tbs, pays):
        seg, blocks = O.segment(O.crc_append(O.CRC24A, pay))
        for r, blk in enumerate(blocks):
            f = seg["filler_bits"] if r == 0 else 0
            e = O.compute_e(r, gt, seg["c"], nl, qm)
            want = O.rate_match(*O.encode_cb(blk, f), f, e, rv)
            assert np.array_equal(rm[roff:roff + e], want), (b, gt, rv, r)
            assert np.array_equal(cb[coff:coff + blk.size], blk), (b, gt, rv, r)
            roff += e
            coff += blk.size
    assert roff == rm.size and coff == cb.size


def test_gpu_encode_decode_round_trip_on_device():
    """Encode on the device (rv from the descriptors, then rv 0 override),
    BPSK without noise, decode with the TB plan: every TB CRC passes and the
    payloads come back."""
    import torch
    tbs, pays = _random_tbs(9, 16)
    full = []  # every coded bit sent at least once, so noise-free decoding succeeds for any rv
    for (b, g, rv, nl, qm) in tbs:
        c = T._segment_sizes(b + 24)
        nv = sum(3 * (k + 4) for k in [c[4]] * c[2] + [c[3]] * c[1])
        q = qm * nl
        full.append((b, max(g, -(-int(1.1 * nv) // q) * q), rv, nl, qm))
    tbs = full
    plan = T.TbPlan(tbs, max_iterations=4, early_stop=True)
    pay = torch.from_numpy(np.concatenate(pays)).cuda()
    rm = torch.empty(plan.rx_floats, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    bits = torch.empty(plan.bits_total, dtype=torch.uint8, device="cuda")
    ok = torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda")
    for rv in (-1, 0):
        _say("device round trip rv", rv)
        plan.encode(pay.data_ptr(), rm.data_ptr(), rv=rv, stream=s)
        rx = (rm.float() * 2.0 - 1.0) * 4.0
        plan.decode(rx.data_ptr(), bits.data_ptr(), ok.data_ptr(), rv=rv, stream=s)
        torch.cuda.synchronize()
        assert bool(ok.all()) and torch.equal(bits, pay)


@pytest.mark.parametrize("waves", [1, 3])
def test_tb_pipeline_graph_capture_replays_match_eager(waves):
    """The TB pipeline (de-rate matching, decode with CRC24B early stop,
    assembly) allocates nothing and records no memset, so it can be captured
    in a HIP graph: three replays, with eager work in between, each equal
    the eager run bit for bit.  waves = 3: segment length from 32 on jobs of
    three waves (the job queue's per-work-group tickets and its reset by the
    last work-group out, replayed)."""
    import torch
    tbs, rx, _ = _tb_batch(31, n_tb=12)
    geo = dict(min_segment=32, max_waves_per_block=waves) if waves > 1 else {}
    plan = T.TbPlan(tbs, decoder=T.DECODER_WINDOW_LOG, max_iterations=8, early_stop=True, **geo)
    if waves > 1:
        assert max(T.segment_geometry(int(k), T.DECODER_WINDOW_LOG, 32, 0, 0, waves)["waves_per_block"]
                   for k in plan.ks) == waves
    d_rx = torch.from_numpy(rx).cuda()

    def outputs():
        return (torch.zeros(plan.bits_total, dtype=torch.uint8, device="cuda"),
                torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda"),
                torch.zeros(plan.n_cb, dtype=torch.int32, device="cuda"),
                torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda"),
                torch.zeros(plan.post_floats, dtype=torch.float32, device="cuda"))

    def run(o, stream):
        plan.decode(d_rx.data_ptr(), o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr(), o[3].data_ptr(),
                    o[4].data_ptr(), stream=stream)

    eager = outputs()
    _say("graph: eager")
    run(eager, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    cap = outputs()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):  # warm-up on a side stream before capture
        run(cap, side.cuda_stream)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    _say("graph: capture")
    with torch.cuda.graph(g):
        run(cap, torch.cuda.current_stream().cuda_stream)
    other = outputs()
    for rep in range(3):
        for t in cap:
            t.zero_()
        torch.cuda.synchronize()
        _say("graph: replay", rep)
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(eager, cap):
            assert torch.equal(a, b), rep
        run(other, torch.cuda.current_stream().cuda_stream)  # eager work between replays
        torch.cuda.synchronize()
        for a, b in zip(eager, other):
            assert torch.equal(a, b), rep


def test_code_block_crc_mode():
    """code_block_crc descriptors: one code block of K = B + 24 ending in a
    CRC24B.  Encoding equals the host chain (append_crc24b, encoder, rate
    matching); decoding stops early on that CRC24B and reports it as the
    TB check."""
    import torch
    rng = np.random.default_rng(17)
    ks = [6144, 1024, 40, 3200, 6144]
    tbs = [(k - 24, int(3 * (k + 4) * 0.8) // 2 * 2, i % 4, 1, 2, 1) for i, k in enumerate(ks)]
    pays = [rng.integers(0, 2, k - 24).astype(np.uint8) for k in ks]
    rm, cb = T.encode_transport_blocks(tbs, pays)
    roff = coff = 0
    for (b, g, rv, _, _, _), pay, k in zip(tbs, pays, ks):
        bits = T.append_crc24b(pay)
        d = T.encode_code_block_streams(bits)
        ref = O.rate_match(*d, 0, g, rv)
        assert np.array_equal(cb[coff:coff + k], bits)
        assert np.array_equal(rm[roff:roff + g], ref)
        roff += g
        coff += k
    plan = T.TbPlan(tbs, max_iterations=8, early_stop=True)
    rx = torch.from_numpy((2.0 * rm.astype(np.float32) - 1.0) * 3.0).cuda()
    bits = torch.empty(plan.bits_total, dtype=torch.uint8, device="cuda")
    ok = torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda")
    it = torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda")
    _say("code-block CRC mode")
    plan.decode(rx.data_ptr(), bits.data_ptr(), ok.data_ptr(), 0, it.data_ptr())
    torch.cuda.synchronize()
    assert bool(ok.all()) and np.array_equal(bits.cpu().numpy(), np.concatenate(pays))
    assert int(it.max()) < 8  # early stop fired


@pytest.mark.parametrize("decoder", [T.DECODER_WINDOW_LOG, T.DECODER_WINDOW_LIN, T.DECODER_EXACT])
def test_fused_derate_decode_equals_separate(decoder):
    """De-rate matching inside the decoder's input staging (tc_plan_decode_rx)
    is bit-identical to the de-rate kernel followed by a decode: one
    transmission, HARQ of two (rv 0 + rv 2) in one call, and accumulation
    onto a held input; the held input written by the fused path equals the
    de-rate kernel's output.  Mixed K, fillers, repetition, early stop."""
    import torch
    ks, fs, es, _, rng = _mixed_batch(23)
    n = len(ks)
    its = 3 if decoder == T.DECODER_EXACT else 6
    rx0 = torch.from_numpy(rng.normal(0, 2.5, sum(es)).astype(np.float32)).cuda()
    rx2 = torch.from_numpy(rng.normal(0, 2.5, sum(es)).astype(np.float32)).cuda()
    der = T.DeratePlan(ks, es, fs)
    plan = T.Plan(ks, decoder, its, early_stop=True, fillers=fs)
    s = torch.cuda.current_stream().cuda_stream

    def z(m, dt=torch.float32):
        return torch.zeros(m, dtype=dt, device="cuda")

    # separate: de-rate kernel, then decode
    soft = z(der.llr_floats)
    der.run(rx0.data_ptr(), soft.data_ptr(), rv=0, stream=s)
    soft1 = soft.clone()
    der.run(rx2.data_ptr(), soft.data_ptr(), rv=2, accumulate=True, stream=s)
    post_a, crc_a, it_a = z(plan.post_floats), z(n, torch.int32), z(n, torch.int32)
    plan.decode(soft1.data_ptr(), post_a.data_ptr(), crc_a.data_ptr(), it_a.data_ptr(), s)
    post_b = z(plan.post_floats)
    plan.decode(soft.data_ptr(), post_b.data_ptr(), 0, 0, s)
    _say("fused decode", decoder)
    # fused, one transmission
    keep, post_c, crc_c, it_c = z(der.llr_floats), z(plan.post_floats), z(n, torch.int32), z(n, torch.int32)
    plan.decode_rx(der, [rx0.data_ptr()], post_c.data_ptr(), rvs=[0], keep_ptr=keep.data_ptr(),
                   crc_ptr=crc_c.data_ptr(), iters_ptr=it_c.data_ptr(), stream=s)
    torch.cuda.synchronize()
    assert torch.equal(post_c, post_a) and torch.equal(crc_c, crc_a) and torch.equal(it_c, it_a)
    assert torch.equal(keep, soft1)
    # fused, HARQ of two transmissions in one call
    keep2, post_d = z(der.llr_floats), z(plan.post_floats)
    plan.decode_rx(der, [rx0.data_ptr(), rx2.data_ptr()], post_d.data_ptr(), rvs=[0, 2], keep_ptr=keep2.data_ptr(),
                   stream=s)
    torch.cuda.synchronize()
    assert torch.equal(post_d, post_b) and torch.equal(keep2, soft)
    # fused, accumulating onto a held input
    keep3, post_e = soft1.clone(), z(plan.post_floats)
    plan.decode_rx(der, [rx2.data_ptr()], post_e.data_ptr(), rvs=[2], accumulate=True, keep_ptr=keep3.data_ptr(),
                   stream=s)
    torch.cuda.synchronize()
    assert torch.equal(post_e, post_b) and torch.equal(keep3, soft)


def test_tb_decode_multi_equals_combine_then_decode():
    import torch
    tbs, rx, _ = _tb_batch(41, n_tb=10)
    plan = T.TbPlan(tbs, max_iterations=6, early_stop=True)
    rng = np.random.default_rng(5)
    r0 = torch.from_numpy(rx).cuda()
    r2 = torch.from_numpy(rx[rng.permutation(rx.size)]).cuda()
    s = torch.cuda.current_stream().cuda_stream
    outs = []
    for mode in ("combine", "multi"):
        bits = torch.zeros(plan.bits_total, dtype=torch.uint8, device="cuda")
        post = torch.zeros(plan.post_floats, dtype=torch.float32, device="cuda")
        ok = torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda")
        if mode == "combine":
            plan.soft_combine(r0.data_ptr(), rv=0, stream=s)
            plan.soft_combine(r2.data_ptr(), rv=2, accumulate=True, stream=s)
            plan.decode(0, bits.data_ptr(), ok.data_ptr(), post_ptr=post.data_ptr(), stream=s)
        else:
            plan.decode_multi([r0.data_ptr(), r2.data_ptr()], [0, 2], bits.data_ptr(), ok.data_ptr(),
                              post_ptr=post.data_ptr(), stream=s)
        torch.cuda.synchronize()
        outs.append((bits, post, ok))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_tb_plan_without_held_input():
    """set_hold(False): same decode results, no accumulation allowed; rx NULL
    after soft_combine still works."""
    import torch
    tbs, rx, _ = _tb_batch(51, n_tb=8)
    outs = []
    for hold in (True, False):
        plan = T.TbPlan(tbs, max_iterations=6, early_stop=True)
        plan.set_hold(hold)
        r = torch.from_numpy(rx).cuda()
        post = torch.zeros(plan.post_floats, dtype=torch.float32, device="cuda")
        ok = torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda")
        plan.decode(r.data_ptr(), 0, ok.data_ptr(), post_ptr=post.data_ptr())
        torch.cuda.synchronize()
        outs.append((post, ok))
        if not hold:
            with pytest.raises(ValueError):
                plan.decode(r.data_ptr(), accumulate=True)
            plan.soft_combine(r.data_ptr())
            post2 = torch.zeros_like(post)
            plan.decode(0, post_ptr=post2.data_ptr())
            torch.cuda.synchronize()
            assert torch.equal(post2, post)
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("decoder", [T.DECODER_WINDOW_LOG, T.DECODER_WINDOW_LIN])
def test_tb_staging_modes_agree(decoder):
    """De-rate matching in the window decoder's input staging
    (STAGING_FUSED), as de-rate kernels before the decode (STAGING_SEPARATE)
    and the default (STAGING_AUTO) give the same payload bits, CRC flags,
    iteration counts and posteriors, bit for bit: one transmission, two
    combined in one call, and a transmission accumulated onto the held
    input.  The batch mixes punctured and repeated blocks (E at 0.5-1.3 of
    the circular buffer)."""
    import torch
    tbs, rx, _ = _tb_batch(61, n_tb=10)
    rng = np.random.default_rng(7)
    r0 = torch.from_numpy(rx).cuda()
    r2 = torch.from_numpy(rx[rng.permutation(rx.size)]).cuda()
    s = torch.cuda.current_stream().cuda_stream
    res = []
    for mode in (T.STAGING_FUSED, T.STAGING_SEPARATE, T.STAGING_AUTO):
        plan = T.TbPlan(tbs, decoder, max_iterations=6, early_stop=True)
        plan.set_staging(mode)
        outs = []
        for case in ("one", "multi", "accumulate"):
            bits = torch.zeros(plan.bits_total, dtype=torch.uint8, device="cuda")
            post = torch.zeros(plan.post_floats, dtype=torch.float32, device="cuda")
            ok = torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda")
            cb_ok = torch.zeros(plan.n_cb, dtype=torch.int32, device="cuda")
            its = torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda")
            args = dict(bits_ptr=bits.data_ptr(), tb_crc_ptr=ok.data_ptr(), cb_crc_ptr=cb_ok.data_ptr(),
                        iters_ptr=its.data_ptr(), post_ptr=post.data_ptr(), stream=s)
            _say("staging", mode, "decoder", decoder, case)
            if case == "one":
                plan.decode(r0.data_ptr(), **args)
            elif case == "multi":
                plan.decode_multi([r0.data_ptr(), r2.data_ptr()], [0, 2], **args)
            else:
                plan.decode(r0.data_ptr(), rv=1, **args)
                plan.decode(r2.data_ptr(), rv=3, accumulate=True, **args)
            torch.cuda.synchronize()
            outs.append((bits.cpu(), ok.cpu(), cb_ok.cpu(), its.cpu(), post.cpu()))
        res.append(outs)
    for other in res[1:]:
        for a, b in zip(res[0], other):
            for x, y in zip(a, b):
                assert torch.equal(x, y)


def _tbs_rx(sizes, seed):
    """Transport blocks of the given payload sizes, full circular buffer at
    rv 0, BPSK/AWGN LLRs of the rate-matched bits (noisy enough that some
    blocks need several iterations and a few fail)."""
    rng = np.random.default_rng(seed)
    tbs, rx = [], []
    for i, b in enumerate(sizes):
        seg = T._segment_sizes(b + 24)
        ks = [seg[4]] * seg[2] + [seg[3]] * seg[1]
        g_total = sum(3 * (k + 4) for k in ks) - 2 * seg[5]
        p = T.LteTurboCodecParams(transport_block_bits=b, g_total=g_total, rv=0)
        enc = T.encode_transport_block(p, rng.integers(0, 2, b).astype(np.uint8))
        sigma = (1.45, 1.15, 1.0)[i % 3]  # Eb/N0 about -1.5, 0.5 and 1.7 dB
        for blk in enc.code_blocks:
            y = 2.0 * blk.rate_matched_bits.astype(np.float64) - 1.0 + sigma * rng.standard_normal(len(blk.rate_matched_bits))
            rx.append((2.0 * y / sigma ** 2).astype(np.float32))
        tbs.append((b, g_total, 0, 1, 2))
    return tbs, np.concatenate(rx)


@pytest.mark.parametrize("early", [False, True])
def test_lean_tb_plans_equal_general(early):
    """Transport blocks whose code blocks all have the lean geometry (K a
    multiple of 512 from 2048, warm-up 16: K = 6144, 4096, 3072, 5120 here,
    one or two blocks per TB) run the lean log kernel, with de-rate matching
    in its input staging or as separate kernels and the TB assembly in its
    epilogue.  The same TBs plus one short TB run the general kernel.
    Payload bits, TB and block CRC flags, iterations and posteriors are the
    same bits in all four runs."""
    import torch
    sizes = [6120, 4072, 3048, 12216, 8120, 5096, 6120, 12216]
    tbs, rx = _tbs_rx(sizes, 90 + early)
    short_tbs, short_rx = _tbs_rx([16], 3)
    s = torch.cuda.current_stream().cuda_stream
    res = []
    for extra in (False, True):
        all_tbs = tbs + (short_tbs if extra else [])
        r = torch.from_numpy(np.concatenate([rx, short_rx]) if extra else rx).cuda()
        for mode in (T.STAGING_FUSED, T.STAGING_SEPARATE):
            plan = T.TbPlan(all_tbs, T.DECODER_WINDOW_LOG, max_iterations=8, early_stop=early, warmup=16)
            plan.set_staging(mode)
            nb = sum(a for a, *_ in tbs)
            bits = torch.zeros(plan.bits_total, dtype=torch.uint8, device="cuda")
            post = torch.zeros(plan.post_floats, dtype=torch.float32, device="cuda")
            ok = torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda")
            cb_ok = torch.zeros(plan.n_cb, dtype=torch.int32, device="cuda")
            its = torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda")
            _say("lean TB plan: short TB", extra, "staging", mode, "early stop", early)
            plan.decode(r.data_ptr(), bits_ptr=bits.data_ptr(), tb_crc_ptr=ok.data_ptr(), cb_crc_ptr=cb_ok.data_ptr(),
                        iters_ptr=its.data_ptr(), post_ptr=post.data_ptr(), stream=s)
            torch.cuda.synchronize()
            ncb = len(plan.ks) - (1 if extra else 0)  # the short TB's block comes last
            nk = sum(int(k) for k in plan.ks[:ncb])
            res.append((bits[:nb].cpu(), ok[:len(tbs)].cpu(), cb_ok[:ncb].cpu(), its[:len(tbs)].cpu(),
                        post[:nk].cpu()))
    assert 0 < int(res[0][1].sum()) < len(tbs)
    for other in res[1:]:
        for x, y in zip(res[0], other):
            assert torch.equal(x, y)


def test_tb_staging_mode_errors():
    tbs, _, _ = _tb_batch(62, n_tb=2)
    plan = T.TbPlan(tbs)
    for bad in (-1, 3):
        with pytest.raises(ValueError):
            plan.set_staging(bad)


# -------------------------------------------------- config 5: mixed load ---
def test_mixed_load_harq_rayleigh_16qam_matches_reference_chain():
    """The large mixed load at test size: 320 code blocks, 70% K=6144 and 30%
    drawn from the QPP table, each K - 24 payload bits + CRC24B (code-block
    CRC descriptors, so early stop applies as to the reference's segmented
    blocks, lte_turbo.cpp:792-804); GPU encoder with rv 0 and rv 2, 16QAM over
    i.i.d. Rayleigh fading with known CSI at Eb/N0 U[2,6] dB, both
    transmissions soft-combined (lte_turbo.cpp:280-325 in fp32) and decoded
    with early stop in one TbPlan.decode_multi call.

    Bars: posteriors, CRC flags and iteration counts bit-exact with the
    oracle's soft de-rate matching (or_derate_match_soft, rv 0 then rv 2 added)
    followed by the batched decode of those inputs; each block's TB flag is the
    CRC24B of its hard decisions and its payload the first K - 24 of them
    (or_crc_check, the reference's polynomial); CRC verdicts and iteration
    counts equal the oracle's windowed decoder (or_decode_window, fp16-held
    extrinsics, early stop) on all but at most 1% of the blocks, where one
    fp16 rounding step of a held extrinsic (hardware exp2/log2 against libm)
    can move a marginal block's stopping iteration."""
    import concurrent.futures as cf

    import torch
    rng = np.random.default_rng(505)
    n = 320
    table = [int(e[0]) for e in T.qpp_interleaver_table()]
    ks = np.where(rng.random(n) < 0.7, 6144, rng.choice(table, n)).astype(np.int64)
    es = (3 * (ks + 4) // 2) // 4 * 4
    ebn0 = rng.uniform(2.0, 6.0, n)
    tbs = [(int(k) - 24, int(e), 0, 1, 4, 1) for k, e in zip(ks, es)]
    plan = T.TbPlan(tbs, T.DECODER_WINDOW_LOG, 8, early_stop=True)
    assert list(plan.ks) == list(ks) and list(plan.es) == list(es) and not plan.fillers.any()
    pay_h = rng.integers(0, 2, plan.bits_total).astype(np.uint8)
    pay = torch.from_numpy(pay_h).cuda()
    n0 = 1.0 / (10.0 ** (ebn0 / 10.0) * 4.0 * (ks - 24) / es)
    ch = T.ChannelPlan([(int(e), 4, 1, float(v)) for e, v in zip(es, n0)])
    s = torch.cuda.current_stream().cuda_stream
    rx = []
    for rv, seed in ((0, 71), (2, 72)):
        rm = torch.empty(plan.rx_floats, dtype=torch.uint8, device="cuda")
        plan.encode(pay.data_ptr(), rm.data_ptr(), rv=rv, stream=s)
        y = torch.empty(plan.rx_floats, dtype=torch.float32, device="cuda")
        ch.run(seed, rm.data_ptr(), y.data_ptr(), stream=s)
        rx.append(y)
    bits = torch.zeros(plan.bits_total, dtype=torch.uint8, device="cuda")
    tb_ok = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    cb_ok = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    its = torch.zeros(n, dtype=torch.int32, device="cuda")
    post = torch.zeros(plan.post_floats, dtype=torch.float32, device="cuda")
    _say("mixed load", n, "blocks: decode_multi rv0+rv2")
    plan.decode_multi([rx[0].data_ptr(), rx[1].data_ptr()], [0, 2], bits.data_ptr(), tb_ok.data_ptr(),
                      cb_ok.data_ptr(), its.data_ptr(), post.data_ptr(), stream=s)
    torch.cuda.synchronize()
    post_h, bits_h = post.cpu().numpy(), bits.cpu().numpy()
    tb_h, cb_h, its_h = tb_ok.cpu().numpy(), cb_ok.cpu().numpy(), its.cpu().numpy()
    rx_h = [r.cpu().numpy() for r in rx]
    # reference chain: the oracle's soft de-rate matching of both transmissions, added
    llr, ioff = [], 0
    for k, e in zip(ks, es):
        l0 = O.derate_match_soft(rx_h[0][ioff:ioff + int(e)], int(k) + 4, 0, 0)
        llr.append(O.derate_match_soft(rx_h[1][ioff:ioff + int(e)], int(k) + 4, 0, 2, out=l0))
        ioff += int(e)
    _say("mixed load: batched decode of the oracle-combined inputs")
    post_ref, crc_ref, its_ref = T.decode_batch_host(ks, np.concatenate(llr), T.DECODER_WINDOW_LOG, 8,
                                                     early_stop=True)
    assert np.array_equal(post_h.view(np.uint32), post_ref.view(np.uint32))
    assert np.array_equal(tb_h.astype(bool), crc_ref) and np.array_equal(cb_h, tb_h)
    assert np.array_equal(its_h, its_ref.astype(np.int32))
    poff = boff = 0
    for i, k in enumerate(ks):
        hard = (post_h[poff:poff + k] >= 0).astype(np.uint8)
        assert bool(tb_h[i]) == O.crc_check(O.CRC24B, hard), i
        assert np.array_equal(bits_h[boff:boff + k - 24], hard[:k - 24]), i
        poff += int(k)
        boff += int(k) - 24
    bler = 1.0 - tb_h.mean()
    _say(f"mixed load: BLER {bler:.4f}, mean iterations {its_h.mean():.2f}")
    assert bler < 0.1 and its_h.mean() < 4.0
    # the oracle's windowed decoder, same geometry and held-extrinsic width

    def one(i):
        k = int(ks[i])
        geo = T.segment_geometry(k)
        d = k + 4
        _, ok, used = O.decode_window(llr[i][:d], llr[i][d:2 * d], llr[i][2 * d:], k, geo["segment_length"],
                                      geo["warmup"], False, 0, 8, True, geo["extrinsic_bits"])
        return ok, used

    with cf.ThreadPoolExecutor(8) as ex:
        res = list(ex.map(one, range(n)))
    differ = [i for i, (ok, used) in enumerate(res) if ok != bool(tb_h[i]) or used != its_h[i]]
    _say(f"mixed load: oracle window decoder differs on {len(differ)} of {n} blocks")
    assert len(differ) <= n // 100, differ
