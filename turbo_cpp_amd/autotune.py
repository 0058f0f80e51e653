"""Autotuner of the window decoders' launch geometry.

Knobs (tc_options): the segment length S (via min_segment and
max_waves_per_block n; S is the smallest divisor of K >= max(min_segment,
K/(64 n)), which also fixes the segments per block M = K/S, the waves one
block spans, ceil(M/64), and the blocks per job), the warm-up steps per
segment boundary, the resident waves per CU (occupancy) and the width of the
extrinsics held between half-iterations (fp16 or fp32).  Window lengths 32
and 64 at K=6144 are blocks on three and two waves.

For one K the tuner makes calibration batches on the GPU at several Eb/N0
points (random payloads of K - 24 bits with CRC24A, the GPU encoder, full
circular buffer at rv 0, QPSK over AWGN, max-log LLRs, soft de-rate
matching), decodes them with the exact decoder (the reference's algorithm)
and with every candidate, and times each candidate's decode with HIP events
on the middle point's batch.  A candidate qualifies when, at every point,
both its bit-error and its block-error counts lie inside the 95%
Clopper-Pearson intervals of the exact decoder's counts on the same blocks;
the fastest qualifying candidate is returned with the whole table.
"""
from __future__ import annotations

import itertools
import math
from dataclasses import asdict, dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import (DECODER_EXACT, DECODER_WINDOW_LOG, ChannelPlan, DeratePlan, Plan, TbPlan, segment_geometry)


@dataclass
class Candidate:
    min_segment: int
    warmup: int
    waves_per_cu: int
    extrinsic_bits: int = 16
    max_waves_per_block: int = 1
    waves_per_block: int = 1
    segment_length: int = 0
    segments: int = 0
    blocks_per_wave: int = 0
    ms: float = float("nan")
    bit_errors: int = -1  # summed over the calibration points
    block_errors: int = -1
    per_point: Optional[list] = None  # [(bit errors, block errors)] per Eb/N0 point
    qualifies: bool = False


def _beta_ppf(q, a, b):
    from scipy.stats import beta
    return float(beta.ppf(q, a, b))


def clopper_pearson(k: int, n: int, alpha: float = 0.05):
    lo = 0.0 if k == 0 else _beta_ppf(alpha / 2, k, n - k + 1)
    hi = 1.0 if k == n else _beta_ppf(1 - alpha / 2, k + 1, n - k)
    return lo, hi


def calibration_batch(k: int, n_blocks: int, ebn0_db: float, seed: int = 1):
    """Decoder inputs ([d0|d1|d2] per block, device) and the code-block bits (device)."""
    import torch
    plan = TbPlan([(k - 24, 3 * (k + 4), 0, 1, 2)] * n_blocks, max_iterations=1)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    pay = torch.randint(0, 2, (plan.bits_total,), dtype=torch.uint8, device="cuda", generator=gen)
    rm = torch.empty(plan.rx_floats, dtype=torch.uint8, device="cuda")
    cb = torch.empty(plan.post_floats, dtype=torch.uint8, device="cuda")
    plan.encode(pay.data_ptr(), rm.data_ptr(), cb.data_ptr())
    rate = k / (3.0 * (k + 4))
    n0 = 1.0 / (2.0 * rate * 10.0 ** (ebn0_db / 10.0))  # QPSK: Es = 1 carries 2 coded bits
    ch = ChannelPlan([(int(e), 2, 0, n0) for e in plan.es])
    rx = torch.empty(plan.rx_floats, dtype=torch.float32, device="cuda")
    ch.run(seed, rm.data_ptr(), rx.data_ptr())
    der = DeratePlan(plan.ks, plan.es)
    llr = torch.empty(der.llr_floats, dtype=torch.float32, device="cuda")
    der.run(rx.data_ptr(), llr.data_ptr())
    torch.cuda.synchronize()
    return llr, cb


def _decode(plan: Plan, llr, post, reps: int):
    import torch
    s = torch.cuda.current_stream()
    plan.decode(llr.data_ptr(), post.data_ptr(), 0, 0, s.cuda_stream)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        plan.decode(llr.data_ptr(), post.data_ptr(), 0, 0, s.cuda_stream)
    b.record(s)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune(k: int, decoder: int = DECODER_WINDOW_LOG, iterations: int = 8, n_blocks: int = 2048,
         ebn0_db=(0.25, 0.5, 0.75), min_segments: Sequence[int] = (32, 64, 96, 128, 192),
         warmups: Sequence[int] = (8, 16, 24, 32), waves: Sequence[int] = (0, 4, 2),
         extrinsic_bits: Sequence[int] = (16,), block_waves: Sequence[int] = (1, 3), reps: int = 3,
         seed: int = 1) -> Dict:
    """Returns {"best": {...}, "table": [...], "points": [per Eb/N0: exact errors and intervals], ...}."""
    import torch
    pts = [float(ebn0_db)] if np.isscalar(ebn0_db) else [float(e) for e in ebn0_db]
    batches = [calibration_batch(k, n_blocks, e, seed + i) for i, e in enumerate(pts)]
    timed_at = len(pts) // 2
    post = torch.empty(n_blocks * k, dtype=torch.float32, device="cuda")
    nbits = n_blocks * k

    def errors(plan):
        out = []
        for llr, cb in batches:
            plan.decode(llr.data_ptr(), post.data_ptr(), 0, 0, torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            wrong = (post >= 0) != cb.to(torch.bool)
            out.append((int(wrong.sum()), int(wrong.view(n_blocks, k).any(dim=1).sum())))
        return out

    exact = errors(Plan([k] * n_blocks, DECODER_EXACT, iterations))
    points = []
    for e, (be, ke) in zip(pts, exact):
        lo_b, hi_b = clopper_pearson(be, nbits)
        lo_k, hi_k = clopper_pearson(ke, n_blocks)
        points.append({"ebn0_db": e, "exact_bit_errors": be, "exact_block_errors": ke,
                       "bit_interval": [math.floor(lo_b * nbits), math.ceil(hi_b * nbits)],
                       "block_interval": [math.floor(lo_k * n_blocks), math.ceil(hi_k * n_blocks)]})
    seen = set()
    table: List[Candidate] = []
    for ms, g, w, xb, bw in itertools.product(min_segments, warmups, waves, extrinsic_bits, block_waves):
        try:
            geo = segment_geometry(k, decoder, ms, g, xb, bw)
        except ValueError:
            continue  # segment length outside what the kernel supports
        key = (geo["segment_length"], geo["warmup"], w, xb)
        if key in seen:
            continue
        seen.add(key)
        c = Candidate(ms, g, w, xb, bw, geo["waves_per_block"], geo["segment_length"], geo["segments"],
                      geo["blocks_per_wave"])
        plan = Plan([k] * n_blocks, decoder, iterations, min_segment=ms, warmup=g, waves_per_cu=w, extrinsic_bits=xb,
                    max_waves_per_block=bw)
        c.per_point = errors(plan)
        c.bit_errors = sum(x[0] for x in c.per_point)
        c.block_errors = sum(x[1] for x in c.per_point)
        c.qualifies = all(p["bit_interval"][0] <= be <= p["bit_interval"][1] and
                          p["block_interval"][0] <= ke <= p["block_interval"][1]
                          for p, (be, ke) in zip(points, c.per_point))
        c.ms = _decode(plan, batches[timed_at][0], post, reps)
        table.append(c)
        del plan
    ok = [c for c in table if c.qualifies]
    best = min(ok or table, key=lambda c: c.ms)
    return {"k": k, "decoder": decoder, "iterations": iterations, "n_blocks": n_blocks, "ebn0_db": pts,
            "timed_at_db": pts[timed_at], "bits": nbits, "points": points,
            "exact_bit_errors": [p["exact_bit_errors"] for p in points],
            "interval": [p["bit_interval"] for p in points],
            "best": asdict(best), "table": [asdict(c) for c in table]}


def options_for(result: Dict) -> Dict[str, int]:
    """Keyword arguments for Plan / options() from a tune() result."""
    b = result["best"]
    return {"min_segment": b["min_segment"], "warmup": b["warmup"], "waves_per_cu": b["waves_per_cu"],
            "extrinsic_bits": b["extrinsic_bits"], "max_waves_per_block": b.get("max_waves_per_block", 1)}
