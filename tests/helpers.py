"""Shared test helpers: seeded synthetic channel and tolerance checks."""
from __future__ import annotations

import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


def golden():
    return np.load(GOLDEN, allow_pickle=False)


def bpsk_awgn_llr(streams, k, ebn0_db, rng):
    """BPSK c -> 2c-1 (LLR > 0 means bit 1) over AWGN at Eb/N0 for the
    rate K/(3K+12) mother code; LLR = 2y/sigma^2 in float32."""
    rate = k / (3 * k + 12)
    sigma2 = 1.0 / (2.0 * rate * 10.0 ** (ebn0_db / 10.0))
    out = []
    for s in streams:
        y = (2.0 * np.asarray(s, np.float64) - 1.0) + math.sqrt(sigma2) * rng.standard_normal(len(s))
        out.append((2.0 * y / sigma2).astype(np.float32))
    return out


def random_blocks(n, k, ebn0_db, seed, encode):
    """n code blocks of K info bits: (bits [n,K], llr [n, 3(K+4)])."""
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2, (n, k)).astype(np.uint8)
    llr = np.zeros((n, 3 * (k + 4)), np.float32)
    for i in range(n):
        d = encode(bits[i])
        llr[i] = np.concatenate(bpsk_awgn_llr(d, k, ebn0_db, rng))
    return bits, llr


def clopper_pearson(x, n, alpha=0.05):
    from scipy.stats import beta

    lo = 0.0 if x == 0 else beta.ppf(alpha / 2, x, n - x + 1)
    hi = 1.0 if x == n else beta.ppf(1 - alpha / 2, x + 1, n - x)
    return lo, hi


def assert_llr_parity(got, ref, atol=1e-4, rtol=0.0, hard_tol=1e-3):
    """APP LLRs within atol (+ rtol*|ref|) and identical hard decisions except
    where |LLR| < hard_tol."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape
    assert np.all(np.isfinite(got))
    err = np.abs(got - ref)
    lim = atol + rtol * np.abs(ref)
    bad = np.flatnonzero(err > lim)
    assert bad.size == 0, f"{bad.size} LLRs off; worst {err.max():.3g} at {int(np.argmax(err))} (ref {ref[np.argmax(err)]:.4g})"
    hd = (got >= 0) != (ref >= 0)
    assert not np.any(hd & (np.abs(ref) >= hard_tol)), "hard decisions differ"
