"""Pins the oracle (CPU restatement) before it is trusted as the checker.

Checks: the reference's own known-answer test (QPP table), golden vectors
written from the compiled reference (tests/golden), the compiled reference
itself where oracle/_ref exists, and what the mathematics fixes: QPP
bijection, CRC polynomial and GF(2) linearity, the segmentation closed form,
trellis termination, the rate-matching permutation, and brute-force APP
enumeration for the SISO.
"""
import numpy as np
import pytest

import oracle as O
from tests.helpers import bpsk_awgn_llr, golden


# --------------------------------------------------------------------- QPP --
def test_qpp_table_known_answers():
    # reference tests/test_main.cpp:37-44
    t = O.qpp_table()
    assert len(t) == 188
    assert t[0] == (40, 3, 10)
    assert t[-1] == (6144, 263, 480)
    ks = [e[0] for e in t]
    grid = list(range(40, 513, 8)) + list(range(528, 1025, 16)) + list(range(1056, 2049, 32)) + \
        list(range(2112, 6145, 64))
    assert ks == grid


@pytest.mark.parametrize("k", [e[0] for e in O.qpp_table()])
def test_qpp_is_bijection_closed_form(k):
    p = O.qpp_perm(k)
    _, f1, f2 = [e for e in O.qpp_table() if e[0] == k][0]
    i = np.arange(k, dtype=np.int64)
    assert np.array_equal(p, (f1 * i + f2 * i * i) % k)
    assert np.array_equal(np.sort(p), np.arange(k))


# --------------------------------------------------------------------- CRC --
POLY = {O.CRC24A: 0xC199BF, O.CRC24B: 0xC00063, O.CRC16: 0x8811, O.CRC8: 0xD9}
LEN = {O.CRC24A: 24, O.CRC24B: 24, O.CRC16: 16, O.CRC8: 8}


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_crc_single_bit_is_generator(kind):
    # remainder of D^N mod G(D) is G(D) - D^N: the reference's tap array
    # (lte_turbo.cpp:104-116), which differs from 36.212
    out = O.crc_append(kind, [1])
    n = LEN[kind]
    val = int("".join(str(b) for b in out[1:]), 2)
    assert val == POLY[kind]
    assert O.crc_check(kind, out)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_crc_linearity_and_detection(kind):
    rng = np.random.default_rng(kind)
    n = LEN[kind]
    for m in (5, 64, 300):
        a = rng.integers(0, 2, m).astype(np.uint8)
        b = rng.integers(0, 2, m).astype(np.uint8)
        ca, cb, cab = O.crc_append(kind, a), O.crc_append(kind, b), O.crc_append(kind, a ^ b)
        assert np.array_equal(ca[m:] ^ cb[m:], cab[m:])
        flipped = ca.copy()
        flipped[rng.integers(len(flipped))] ^= 1
        assert not O.crc_check(kind, flipped)  # every single-bit error is detected
    assert not O.crc_check(kind, np.zeros(n - 1, np.uint8))  # shorter than N
    assert O.crc_check(kind, np.zeros(n, np.uint8))


def test_crc_golden():
    g = golden()
    for kind in range(4):
        for n in (1, 7, 40, 333):
            assert np.array_equal(O.crc_append(kind, g[f"crc{kind}_n{n}_in"]), g[f"crc{kind}_n{n}_out"])


# ------------------------------------------------------------ segmentation --
@pytest.mark.parametrize("b,c,ks,f", [
    (64, 1, [64], 0),            # TB 40 + CRC24A
    (536, 1, [544], 8),          # TB 512
    (6136, 1, [6144], 8),        # TB 6112
    (12024, 2, [6016, 6080], 24),  # TB 12000
    (75400, 13, [5824] * 13, None),  # TB 75376
])
def test_segmentation_spot_values(b, c, ks, f):
    s, blocks = O.segment(np.zeros(b, np.uint8))
    assert s["c"] == c
    assert [len(x) for x in blocks] == ks
    if f is not None:
        assert s["filler_bits"] == f
    # 36.212 5.1.2 closed form: B' = B + C L, F = sum K_r - B'
    l = 0 if c == 1 else 24
    assert s["filler_bits"] == sum(ks) - (b + c * l)


def test_segmentation_golden():
    g = golden()
    for b, fields in zip(g["seg_sizes"], g["seg_fields"]):
        s, blocks = O.segment(g[f"seg_b{b}_in"])
        assert [s[x] for x in ("c", "c_plus", "c_minus", "k_plus", "k_minus", "filler_bits", "l")] == list(fields)
        assert np.array_equal(np.concatenate(blocks), g[f"seg_b{b}_blocks"])


def test_segmentation_crc24b_per_block():
    rng = np.random.default_rng(3)
    s, blocks = O.segment(rng.integers(0, 2, 20000).astype(np.uint8))
    for r, blk in enumerate(blocks):
        f = s["filler_bits"] if r == 0 else 0
        assert O.crc_check(O.CRC24B, blk[f:])


# ---------------------------------------------------------------- encoder --
def _rsc(s, u):
    s0, s1, s2 = s & 1, (s >> 1) & 1, (s >> 2) & 1
    fb = u ^ s1 ^ s2
    return fb ^ s0 ^ s2, (fb << 2) | (s2 << 1) | s1


def test_encoder_golden():
    g = golden()
    for k, f in [(40, 0), (64, 0), (544, 8), (1024, 0), (6144, 8), (6016, 24)]:
        d = O.encode_cb(g[f"enc_k{k}_f{f}_in"], f)
        assert np.array_equal(np.stack(d), g[f"enc_k{k}_f{f}_out"])


@pytest.mark.parametrize("k", [40, 104, 1024])
def test_encoder_parity_and_termination(k):
    """Independent re-encode of the parity streams and of the tails: both
    constituent encoders end in state 0 after their three tail steps, with the
    tail bits multiplexed as lte_turbo.cpp:664-677 does."""
    rng = np.random.default_rng(k)
    cb = rng.integers(0, 2, k).astype(np.uint8)
    d0, d1, d2 = O.encode_cb(cb)
    perm = O.qpp_perm(k)
    tails = []
    for seq, par_stream in ((cb, d1), (cb[perm], d2)):
        s = 0
        for i, u in enumerate(seq):
            p, s = _rsc(s, int(u))
            assert p == par_stream[i]
        xs, zs = [], []
        for _ in range(3):
            u = ((s >> 1) ^ (s >> 2)) & 1
            p, s = _rsc(s, u)
            xs.append(u)
            zs.append(p)
        assert s == 0
        tails.append((xs, zs))
    (xu, zu), (xl, zl) = tails
    assert list(d0[k:]) == [xu[0], zu[1], xu[1], zu[2]]
    assert list(d1[k:]) == [zu[0], xu[1], zu[2], xl[0]]
    assert list(d2[k:]) == [zl[0], xl[1], zl[1], xl[2]]
    assert np.array_equal(d0[:k], cb)


# ---------------------------------------------------------- rate matching --
def test_rate_matching_golden():
    g = golden()
    for ci, tbs, g_total, rv, r, k, e, f in g["tb_meta"]:
        key = f"tb{ci}_rv{rv}_cb{r}"
        cbb = g[f"tb{ci}_rv0_cb{r}_cbbits"]
        d = O.encode_cb(cbb, int(f))
        assert np.array_equal(O.rate_match(*d, int(f), int(e), int(rv)), g[key + "_rm"]), key
        assert np.array_equal(O.derate_match(g[key + "_rm"], int(k) + 4, int(f), int(rv)), g[key + "_derm"]), key


def test_soft_derate_of_pm8_is_reference_derate_golden():
    """+-8 per hard bit through the soft de-rate matcher reproduces the
    reference's de-rate matcher (lte_turbo.cpp:280-325) bit for bit, for every
    rv, with puncturing (E < N_valid) and repetition (E > N_valid)."""
    g = golden()
    for ci, tbs, g_total, rv, r, k, e, f in g["tb_meta"]:
        key = f"tb{ci}_rv{rv}_cb{r}"
        rx = np.where(g[key + "_rm"] > 0, 8.0, -8.0).astype(np.float32)
        soft = O.derate_match_soft(rx, int(k) + 4, int(f), int(rv))
        assert np.array_equal(soft.view(np.uint32), g[key + "_derm"].view(np.uint32)), key


@pytest.mark.parametrize("k,f,e", [(40, 0, 50), (544, 8, 3000), (1024, 0, 7000)])
def test_soft_derate_harq_combining_is_additive(k, f, e):
    """HARQ combining: accumulating rv2 onto rv0 equals the sum of the two
    separate de-rate matchings (small-integer LLRs, so float sums are exact),
    and a single LLR routed alone lands on exactly one output position."""
    rng = np.random.default_rng(k)
    d = k + 4
    x0 = rng.integers(-20, 21, e).astype(np.float32)
    x2 = rng.integers(-20, 21, e).astype(np.float32)
    a = O.derate_match_soft(x0, d, f, 0)
    b = O.derate_match_soft(x2, d, f, 2)
    comb = O.derate_match_soft(x2, d, f, 2, out=a.copy())
    assert np.array_equal(comb, a + b)
    # a one-hot received vector reaches the output once per repetition of its entry
    nv = 3 * d - 2 * f
    for j in (0, e // 2, e - 1):
        hot = np.zeros(e, np.float32)
        hot[j] = 1.0
        out = O.derate_match_soft(hot, d, f, 1)
        assert np.count_nonzero(out) == 1 and out.sum() == 1.0
        assert out[:f].sum() == 0 and out[d:d + f].sum() == 0  # filler positions stay 0
    assert nv == np.count_nonzero(O.derate_match_soft(np.ones(max(e, nv), np.float32), d, f, 0))


def _oracle_tb_receive(g, ci):
    """The oracle's receive chain for golden TB decode case ci: +-8 soft
    de-rate matching, exact decode per block, TB assembly."""
    _, tbs, g_total, it, es, c, _, _ = g["tbd_meta"][ci]
    ks, fs, es_ = g[f"tbd{ci}_k"], g[f"tbd{ci}_f"], g[f"tbd{ci}_e"]
    rm = g[f"tbd{ci}_rm"]
    posts, iters, off = [], [], 0
    for k, f, e in zip(ks, fs, es_):
        rx = np.where(rm[off:off + e] > 0, 8.0, -8.0).astype(np.float32)
        off += e
        llr = O.derate_match_soft(rx, int(k) + 4, int(f), 0)
        d = int(k) + 4
        p, ok, used = O.decode_cb(llr[:d], llr[d:2 * d], llr[2 * d:], int(k), int(f), int(it), bool(es))
        posts.append(p)
        iters.append(used)
    post = np.concatenate(posts)
    bits, cb_ok, tb_ok = O.assemble_tb(post, ks, fs, int(tbs))
    return post, bits, cb_ok, tb_ok, max(iters)


@pytest.mark.parametrize("ci", range(6))
def test_tb_receive_chain_matches_reference_golden(ci):
    """Soft de-rate matching + decode + TB assembly of the oracle equal the
    reference's decode_transport_block (lte_turbo.cpp:891-980) on corrupted
    transport blocks: payload bits, per-block and TB CRC flags, iterations
    used and the posterior LLRs bit for bit."""
    g = golden()
    post, bits, cb_ok, tb_ok, used = _oracle_tb_receive(g, ci)
    meta = g["tbd_meta"][ci]
    assert np.array_equal(post.view(np.uint32), g[f"tbd{ci}_post"].view(np.uint32))
    assert np.array_equal(bits, g[f"tbd{ci}_bits"])
    assert np.array_equal(cb_ok, g[f"tbd{ci}_cb_ok"].astype(bool))
    assert tb_ok == bool(meta[6]) and used == meta[7]


def test_non_converged_block_is_ill_conditioned():
    """Why the GPU tests compare posteriors only on blocks that converge: in
    the reference decoder (the oracle is bit-exact with it) a 1-ulp change of
    one input LLR of the non-converged block of golden TB case 4 moves its
    posteriors by more than 1e-4, while converged blocks do not move."""
    g = golden()
    ci = 4
    ks, fs, es = g[f"tbd{ci}_k"], g[f"tbd{ci}_f"], g[f"tbd{ci}_e"]
    rm = g[f"tbd{ci}_rm"]
    off = 0
    moved = []
    for k, f, e in zip(ks, fs, es):
        rx = np.where(rm[off:off + e] > 0, 8.0, -8.0).astype(np.float32)
        off += e
        llr = O.derate_match_soft(rx, int(k) + 4, int(f), 0)
        d = int(k) + 4
        a, ok, _ = O.decode_cb(llr[:d], llr[d:2 * d], llr[2 * d:], int(k), int(f), 8, True)
        llr[5] = np.nextafter(llr[5], np.float32(100))
        b, _, _ = O.decode_cb(llr[:d], llr[d:2 * d], llr[2 * d:], int(k), int(f), 8, True)
        moved.append((ok, float(np.max(np.abs(a - b)))))
    assert any(not ok and m > 1e-4 for ok, m in moved)
    assert all(m < 1e-4 for ok, m in moved if ok)


def test_assemble_tb_detects_single_bit_errors():
    """One flipped hard decision in a segmented TB fails exactly its block's
    CRC24B, and the TB CRC24A unless the bit is part of the block's own CRC24B
    (dropped before the TB check); filler LLRs are ignored."""
    rng = np.random.default_rng(5)
    payload = rng.integers(0, 2, 12000).astype(np.uint8)
    tb = O.crc_append(0, payload)
    seg, blocks = O.segment(tb)
    ks = [seg["k_minus"]] * seg["c_minus"] + [seg["k_plus"]] * seg["c_plus"]
    fs = [seg["filler_bits"]] + [0] * (len(ks) - 1)
    post = (2.0 * np.concatenate(blocks).astype(np.float32) - 1.0) * 3.0
    post[:fs[0]] = -post[:fs[0]]  # filler decisions do not matter
    bits, cb_ok, tb_ok = O.assemble_tb(post, ks, fs, len(payload))
    assert np.array_equal(bits, payload) and cb_ok.all() and tb_ok
    for pos in (fs[0], ks[0] - 1, ks[0] + 17, sum(ks) - 1):
        bad = post.copy()
        bad[pos] = -bad[pos]
        bits, cb_ok, tb_ok = O.assemble_tb(bad, ks, fs, len(payload))
        blk = 0 if pos < ks[0] else 1
        in_cb_crc = (pos - (0 if blk == 0 else ks[0])) >= ks[blk] - 24
        assert tb_ok == in_cb_crc and list(cb_ok) == [i != blk for i in range(len(ks))]


@pytest.mark.parametrize("k,f", [(40, 0), (544, 0), (6144, 0), (6144, 8)])
def test_full_rate_visits_every_bit_once(k, f):
    """At E = number of non-NULL circular-buffer entries and rv=0 each coded
    bit is read exactly once, so de-rate matching returns +-8 everywhere except
    the NULL filler positions of d0/d1."""
    rng = np.random.default_rng(k + f)
    cb = rng.integers(0, 2, k).astype(np.uint8)
    cb[:f] = 0
    d = O.encode_cb(cb, f)
    e = 3 * (k + 4) - 2 * f
    llr = O.derate_match(O.rate_match(*d, f, e, 0), k + 4, f, 0)
    want = np.concatenate([8.0 * (2.0 * s.astype(np.float32) - 1.0) for s in d])
    want[:f] = 0.0
    want[k + 4:k + 4 + f] = 0.0
    assert np.array_equal(llr, want)


# ------------------------------------------------------------- decoder -------
@pytest.mark.parametrize("k,seed", [(8, 1), (10, 2), (12, 3)])
def test_siso_matches_brute_force(k, seed):
    """Log-MAP SISO = exact APP from enumerating all 2^K input words, each
    terminated by its forced tail (float64), minus systematic and a-priori."""
    rng = np.random.default_rng(seed)
    sys_ = rng.normal(0, 2, k).astype(np.float32)
    par = rng.normal(0, 2, k + 4).astype(np.float32)
    ap = rng.normal(0, 1, k).astype(np.float32)
    ts = np.array([rng.normal(), rng.normal(), 0.0, 0.0], np.float32)
    tp = np.append(rng.normal(0, 1, 3), 0.0).astype(np.float32)
    ext = O.siso(sys_, par, ap, ts, tp)
    app = O.brute_app(sys_, par[:k], ap, ts, tp)
    np.testing.assert_allclose(ext, app - sys_ - ap, atol=2e-4)


def _golden_decodes():
    return [(40, 100, 6), (64, 0, 8), (544, 50, 6), (1024, 50, 8), (6144, 100, 8), (6144, 50, 8), (6144, 25, 8)]


@pytest.mark.parametrize("k,eb,it", _golden_decodes())
def test_decoder_golden_bit_exact(k, eb, it):
    g = golden()
    tag = f"dec_k{k}_eb{eb}_it{it}"
    llr = g[tag + "_llr"]
    post, _, used = O.decode_cb(llr[0], llr[1], llr[2], k, 0, it, False)
    assert used == it
    assert np.array_equal(post.view(np.uint32), g[tag + "_post"].view(np.uint32))


def test_decoder_early_stop_golden():
    g = golden()
    llr = g["es_k1024_llr"]
    post, ok, used = O.decode_cb(llr[0], llr[1], llr[2], 1024, 0, 8, True)
    assert [int(ok), used] == list(g["es_k1024_meta"])
    assert np.array_equal(post.view(np.uint32), g["es_k1024_post"].view(np.uint32))


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("k", [40, 56, 544, 1952, 6144])
def test_decoder_matches_compiled_reference(k):
    rng = np.random.default_rng(k)
    for ebn0 in (0.0, 0.75):
        cb = rng.integers(0, 2, k).astype(np.uint8)
        llr = bpsk_awgn_llr(O.encode_cb(cb), k, ebn0, rng)
        for es in (False, True):
            a = O.decode_cb(*llr, k, 0, 6, es)
            b = O.ref_decode_cb(*llr, k, 0, 6, es)
            assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
            assert a[1:] == b[1:]


# ------------------------------------------------- windowed decoder oracle ---
@pytest.mark.parametrize("k", [40, 544, 6144])
@pytest.mark.parametrize("lin", [False, True])
def test_window_single_segment_equals_exact(k, lin):
    """One segment per block: no boundaries, so the windowed arithmetic must
    reproduce the exact decoder up to float rounding."""
    rng = np.random.default_rng(7 * k)
    cb = rng.integers(0, 2, k).astype(np.uint8)
    llr = bpsk_awgn_llr(O.encode_cb(cb), k, 0.75, rng)
    exact = O.decode_cb(*llr, k, 0, 8, False)[0]
    win = O.decode_window(*llr, k, k, 0, lin, 0, 8, False)[0]
    np.testing.assert_allclose(win, exact, rtol=2e-5, atol=2e-4)


def test_round_half_equals_ieee_conversion():
    """or_round_half against numpy's float16 conversion (IEEE 754 round to
    nearest even): random normals across the fp16 range, fp16 subnormals,
    exact ties between neighbouring fp16 values, and saturation."""
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(10.0) ** rng.integers(-6, 5, 20000),
        rng.uniform(-6.2e-5, 6.2e-5, 5000).astype(np.float32),
        np.float32([0.0, -0.0, 65504.0, -65504.0, 65519.0, 1e9, -1e9, 6.1035156e-05, 5.96e-08, 2.98e-08]),
    ]).astype(np.float32)
    # exact midpoints between consecutive fp16 values (ties)
    h = rng.integers(1, 0x7BFF, 4000).astype(np.uint16).view(np.float16).astype(np.float32)
    nxt = np.nextafter(h.astype(np.float16), np.float16(np.inf)).astype(np.float32)
    x = np.concatenate([x, (h + nxt) / 2, -(h + nxt) / 2]).astype(np.float32)
    want = np.clip(x, -65504, 65504).astype(np.float16).astype(np.float32)
    got = O.round_half(x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_window_half_extrinsics_close_to_float():
    """Holding the extrinsics as fp16 moves posteriors by about the fp16
    rounding of what is held (relative 2^-11), not more, and leaves the hard
    decisions of a converged block unchanged; with 32 bits the argument is
    a no-op."""
    k = 6144
    rng = np.random.default_rng(3)
    cb = rng.integers(0, 2, k).astype(np.uint8)
    llr = bpsk_awgn_llr(O.encode_cb(cb), k, 1.0, rng)
    full = O.decode_window(*llr, k, 96, 16, False, 0, 8, False)[0]
    same = O.decode_window(*llr, k, 96, 16, False, 0, 8, False, 32)[0]
    half = O.decode_window(*llr, k, 96, 16, False, 0, 8, False, 16)[0]
    assert np.array_equal(full.view(np.uint32), same.view(np.uint32))
    assert not np.array_equal(full, half)
    assert np.array_equal(full >= 0, half >= 0) and np.array_equal(half >= 0, cb.astype(bool))
    np.testing.assert_allclose(half, full, rtol=2e-3, atol=2e-2)


def test_window_decoder_rejects_bad_geometry():
    z = np.zeros(6148, np.float32)
    with pytest.raises(ValueError):
        O.decode_window(z, z, z, 6144, 100)  # does not divide K
    with pytest.raises(ValueError):
        O.decode_window(z, z, z, 6144, 96, 96)  # warm-up must stay below S
    with pytest.raises(ValueError):
        O.decode_window(z, z, z, 6144, 96, 16, ex_bits=8)


# ------------------------------------------- differentiable decoder (FP64) --
def _sw_inputs(k, n, seed):
    import torch
    rng = np.random.default_rng(seed)
    llr = []
    for _ in range(n):
        cb = rng.integers(0, 2, k).astype(np.uint8)
        d = O.encode_cb(cb, 0)
        llr.append(np.concatenate([(2.0 * (2.0 * x - 1.0) + rng.normal(0, 1.3, k + 4)) for x in d]))
    return torch.tensor(np.stack(llr), dtype=torch.float64)


def test_sw_reference_single_segment_is_the_reference_decoder():
    """One segment and unit scales: the FP64 sliding-window decoder is the
    reference decoder (or_decode_cb, float32): within 1e-4 relative."""
    import torch
    from oracle import sw_reverse as R
    k = 64
    x = _sw_inputs(k, 2, 1)
    post = R.decode(x, torch.ones(12, dtype=torch.float64), k, np.array(O.qpp_perm(k), np.int64), k, 0, 6)
    for i in range(2):
        ref, _, _ = O.decode_cb(*np.split(x[i].numpy().astype(np.float32), 3), k, 0, 6)
        assert np.max(np.abs(post[i].numpy() - ref) / (1 + np.abs(ref))) < 1e-4


def test_sw_reference_full_warmup_equals_single_segment():
    """Segments whose warm-ups reach both block ends compute exact metrics:
    5 segments of 8 with warm-up 40 equal one segment of 40 (FP64, 1e-9)."""
    import torch
    from oracle import sw_reverse as R
    k = 40
    x = _sw_inputs(k, 2, 2)
    w = torch.linspace(0.7, 1.2, 12, dtype=torch.float64)
    perm = np.array(O.qpp_perm(k), np.int64)
    a = R.decode(x, w, k, perm, 40, 0, 6)
    b = R.decode(x, w, k, perm, 8, 40, 6)
    c = R.decode(x, w, k, perm, 8, 4, 6)
    assert float((a - b).abs().max()) < 1e-9
    assert float((a - c).abs().max()) > 1e-6  # short warm-ups do change the metrics


def test_sw_reference_gradcheck():
    """Central finite differences of the FP64 reference (K=40, 5 segments)."""
    import torch
    from oracle import sw_reverse as R
    k = 40
    x = (_sw_inputs(k, 1, 3) * 0.3).requires_grad_(True)
    w = torch.linspace(0.7, 1.2, 4, dtype=torch.float64).requires_grad_(True)
    perm = np.array(O.qpp_perm(k), np.int64)
    assert torch.autograd.gradcheck(lambda a, b: R.decode(a, b, k, perm, 8, 8, 2), (x, w), eps=1e-6, atol=1e-6)


# ------------------------------------------------------ channel front end --
def test_channel_oracle_constellations_and_closed_forms():
    """36.211 7.1 spot values, unit energy and the Gray property of the QAM
    tables; the brute-force max-log LLR reduces to the QPSK closed form
    -2 sqrt(2) Re(z) |h|^2 / N0 (bit 1 = negative level)."""
    from oracle import channel as CH
    p4 = CH.constellation(4) * np.sqrt(10)
    assert p4[0b0000] == 1 + 1j and p4[0b0001] == 1 + 3j and p4[0b0010] == 3 + 1j and p4[0b1111] == -3 - 3j
    p6 = CH.constellation(6) * np.sqrt(42)
    for bits, val in ((0b000000, 3 + 3j), (0b000001, 3 + 1j), (0b000011, 1 + 1j), (0b001000, 5 + 3j),
                      (0b000101, 3 + 7j), (0b111111, -7 - 7j)):
        assert abs(p6[bits] - val) < 1e-9
    for qm in (2, 4, 6):
        p = CH.constellation(qm)
        assert abs(np.mean(np.abs(p) ** 2) - 1.0) < 1e-12
        for v in range(len(p)):
            d = np.abs(p - p[v])
            d[v] = 9
            for u in np.flatnonzero(np.isclose(d, d.min())):
                assert bin(u ^ v).count("1") == 1
    rng = np.random.default_rng(0)
    y = rng.normal(size=50) + 1j * rng.normal(size=50)
    h = rng.normal(size=50) + 1j * rng.normal(size=50)
    llr = CH.maxlog_llr(y, h, 0.5, 2).reshape(-1, 2)
    z = y * np.conj(h) / np.abs(h) ** 2
    g = np.abs(h) ** 2 / 0.5
    assert np.allclose(llr[:, 0], -2 * np.sqrt(2) * z.real * g) and np.allclose(llr[:, 1], -2 * np.sqrt(2) * z.imag * g)



@pytest.mark.parametrize("k,seg", [(40, 8), (40, 40), (352, 32), (1024, 32), (6144, 96), (6144, 192)])
def test_exact_segmented_equals_serial(k, seg):
    """or_decode_exact_seg (segment-parallel recursions whose boundaries are
    fixed up until every segment meets its own checkpoints bit for bit) gives
    the serial decoder's posteriors (or_decode_cb, lte_turbo.cpp:736-813) bit
    for bit, with and without early stop, at low and high SNR; the fix-up
    work is reported in `work`."""
    rng = np.random.default_rng(k + seg)
    for ebn0, es in ((0.0, False), (0.75, True), (2.0, True)):
        payload = rng.integers(0, 2, k - 24).astype(np.uint8)
        cb = O.crc_append(O.CRC24B, payload)
        d0, d1, d2 = bpsk_awgn_llr(O.encode_cb(cb), k, ebn0, rng)
        want, ok, used = O.decode_cb(d0, d1, d2, k, 0, 8, es)
        work = np.zeros(4, np.uint64)
        got, ok2, used2 = O.decode_exact_seg(d0, d1, d2, k, seg, 0, 8, es, work=work)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (ebn0, es)
        assert (ok2, used2) == (ok, used)
        assert work[0] == 2 * used and work[1] >= work[0]


@pytest.mark.parametrize("linear", [False, True])
@pytest.mark.parametrize("k,seg", [(544, 32), (1024, 32), (6144, 96)])
def test_window_full_warmup_equals_one_segment(k, seg, linear):
    """or_decode_window with warm-ups that reach both block ends (warmup >=
    K: alpha from alpha_0 at step 0, beta from the tail beta at step K) makes
    every segment compute the full-length recursions, so the multi-segment
    decode equals the one-segment decode bit for bit, for fp32- and
    fp16-held extrinsics and with early stop: an independent check of the
    oracle's segment offsets, per-segment interleaving and extrinsic
    assembly (the one-segment decode itself equals the exact decoder to
    rounding, test_window_single_segment_equals_exact)."""
    rng = np.random.default_rng(k + seg + linear)
    payload = rng.integers(0, 2, k - 24).astype(np.uint8)
    cb = O.crc_append(O.CRC24B, payload)
    for ebn0, es in ((0.25, False), (1.0, True)):
        d0, d1, d2 = bpsk_awgn_llr(O.encode_cb(cb), k, ebn0, rng)
        for ex_bits in (32, 16):
            one = O.decode_window(d0, d1, d2, k, k, 0, linear, 0, 8, es, ex_bits)
            many = O.decode_window(d0, d1, d2, k, seg, k, linear, 0, 8, es, ex_bits)
            assert np.array_equal(many[0].view(np.uint32), one[0].view(np.uint32)), (ebn0, ex_bits)
            assert many[1:] == one[1:]
