"""Write golden vectors of the reference CPU codec chain into tests/golden/.

The vectors come from the reference's own src/lte_turbo.cpp, compiled CPU-only
into oracle/_ref/libturbo_ref.so by oracle/Makefile (this script builds it
when /root/reference exists).  Each fixture is data only: inputs and the
reference's outputs.  Re-run with `python tests/golden/make_golden.py`.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def bpsk_awgn(bits_streams, k, ebn0_db, rng):
    """LLR = 2y/sigma^2 of BPSK c -> 2c-1 over AWGN, rate K/(3K+12)."""
    rate = k / (3 * k + 12)
    sigma2 = 1.0 / (2.0 * rate * 10.0 ** (ebn0_db / 10.0))
    out = []
    for s in bits_streams:
        y = (2.0 * s.astype(np.float64) - 1.0) + np.sqrt(sigma2) * rng.standard_normal(len(s))
        out.append((2.0 * y / sigma2).astype(np.float32))
    return out


def main():
    oracle.build(with_reference=True)
    rng = np.random.default_rng(20261203)
    data = {}
    # CRC: random messages, each CRC kind
    for kind in range(4):
        for n in (1, 7, 40, 333):
            m = rng.integers(0, 2, n).astype(np.uint8)
            data[f"crc{kind}_n{n}_in"] = m
            data[f"crc{kind}_n{n}_out"] = oracle.ref_crc_append(kind, m)
    # segmentation
    sizes = np.array([40, 64, 512, 6112 + 24, 6144, 6145, 12024, 20000, 75376 + 24], np.int64)
    segs = []
    for b in sizes:
        bits = rng.integers(0, 2, int(b)).astype(np.uint8)
        s, blocks = oracle.ref_segment(bits)
        segs.append([s[x] for x in ("c", "c_plus", "c_minus", "k_plus", "k_minus", "filler_bits", "l")])
        data[f"seg_b{b}_in"] = bits
        data[f"seg_b{b}_blocks"] = blocks
    data["seg_sizes"] = sizes
    data["seg_fields"] = np.array(segs, np.int64)
    # encoder (per code block)
    enc_cases = [(40, 0), (64, 0), (544, 8), (1024, 0), (6144, 8), (6016, 24)]
    for k, f in enc_cases:
        cb = rng.integers(0, 2, k).astype(np.uint8)
        cb[:f] = 0
        d = oracle.ref_encode_cb(cb, f)
        data[f"enc_k{k}_f{f}_in"] = cb
        data[f"enc_k{k}_f{f}_out"] = np.stack(d)
    # transport-block encode (CRC24A, segmentation, encoder, rate matching of
    # encode_transport_block) and de-rate matching of each code block
    tb_cases = [(16, None), (40, None), (40, 150), (512, None), (512, 1000), (6112, None), (12000, None),
                (12000, 30000)]
    tb_meta = []
    for ci, (tbs, g) in enumerate(tb_cases):
        bits = rng.integers(0, 2, tbs).astype(np.uint8)
        seg, blocks = oracle.ref_segment(oracle.ref_crc_append(0, bits))
        ks = [seg["k_minus"]] * seg["c_minus"] + [seg["k_plus"]] * seg["c_plus"]
        g_total = g if g is not None else sum(3 * (k + 4) for k in ks)
        for rv in range(4):
            cbs = oracle.ref_encode_tb(bits, g_total, rv=rv)
            for r, c in enumerate(cbs):
                tb_meta.append([ci, tbs, g_total, rv, r, c["k"], c["e"], c["filler"]])
                key = f"tb{ci}_rv{rv}_cb{r}"
                data[key + "_rm"] = c["rm"]
                data[key + "_derm"] = oracle.ref_derate(c["rm"], c["k"], c["filler"], rv)
                if rv == 0:
                    data[key + "_cbbits"] = c["cb_bits"]
        data[f"tb{ci}_bits"] = bits
    data["tb_meta"] = np.array(tb_meta, np.int64)
    # full decodes (APP posteriors of turbo_decode_code_block_cpu)
    dec_cases = [(40, 1.0, 6), (64, 0.0, 8), (544, 0.5, 6), (1024, 0.5, 8), (6144, 1.0, 8), (6144, 0.5, 8),
                 (6144, 0.25, 8)]
    for k, ebn0, it in dec_cases:
        cb = rng.integers(0, 2, k).astype(np.uint8)
        d = oracle.ref_encode_cb(cb, 0)
        llr = bpsk_awgn(d, k, ebn0, rng)
        post, ok, used = oracle.ref_decode_cb(*llr, k, 0, it, False)
        tag = f"dec_k{k}_eb{int(ebn0 * 100)}_it{it}"
        data[tag + "_bits"] = cb
        data[tag + "_llr"] = np.stack(llr)
        data[tag + "_post"] = post
    # early stop: code block carrying its own CRC24B (segmented TB style)
    k = 1024
    payload = rng.integers(0, 2, k - 24).astype(np.uint8)
    cb = oracle.ref_crc_append(1, payload)
    d = oracle.ref_encode_cb(cb, 0)
    llr = bpsk_awgn(d, k, 1.0, rng)
    post, ok, used = oracle.ref_decode_cb(*llr, k, 0, 8, True)
    data["es_k1024_bits"] = cb
    data["es_k1024_llr"] = np.stack(llr)
    data["es_k1024_post"] = post
    data["es_k1024_meta"] = np.array([int(ok), used], np.int64)
    # transport-block decodes (decode_transport_block, CpuReference) from
    # rate-matched bits with some bits flipped, so that some code blocks and
    # transport blocks fail their CRC
    tbd_cases = [(40, None, 0.0, 8, False), (512, 1000, 0.25, 6, False), (12000, None, 0.0, 4, True),
                 (12000, 30000, 0.2, 8, False), (20000, None, 0.2, 8, True), (40000, 90000, 0.12, 3, False)]
    tbd_meta = []
    for ci, (tbs, g, flip, it, es) in enumerate(tbd_cases):
        payload = rng.integers(0, 2, tbs).astype(np.uint8)
        seg, _ = oracle.ref_segment(oracle.ref_crc_append(0, payload))
        ks = [seg["k_minus"]] * seg["c_minus"] + [seg["k_plus"]] * seg["c_plus"]
        g_total = g if g is not None else sum(3 * (k + 4) for k in ks)
        cbs = oracle.ref_encode_tb(payload, g_total)
        rm = []
        for r, c in enumerate(cbs):
            p = flip * ((r + ci) % 3) / 2.0  # blocks corrupted to different degrees, some clean
            noise = (rng.random(c["e"]) < p).astype(np.uint8)
            rm.append(c["rm"] ^ noise)
        rm = np.concatenate(rm)
        bits, cb_ok, tb_ok, used, post = oracle.ref_decode_tb(payload, g_total, rm, 0, it, es)
        nk = sum(c["k"] for c in cbs)
        tbd_meta.append([ci, tbs, g_total, it, int(es), len(cbs), int(tb_ok), used])
        data[f"tbd{ci}_payload"] = payload
        data[f"tbd{ci}_rm"] = rm
        data[f"tbd{ci}_bits"] = bits
        data[f"tbd{ci}_cb_ok"] = cb_ok[:len(cbs)].astype(np.int64)
        data[f"tbd{ci}_post"] = post[:nk]
        data[f"tbd{ci}_k"] = np.array([c["k"] for c in cbs], np.int64)
        data[f"tbd{ci}_f"] = np.array([c["filler"] for c in cbs], np.int64)
        data[f"tbd{ci}_e"] = np.array([c["e"] for c in cbs], np.int64)
    data["tbd_meta"] = np.array(tbd_meta, np.int64)
    out = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(out, **data)
    print("wrote", out, os.path.getsize(out), "bytes,", len(data), "arrays")


if __name__ == "__main__":
    main()
