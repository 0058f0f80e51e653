"""CPU restatement of the channel front end (QAM mapping, max-log demapping).

TEST INFRASTRUCTURE ONLY: the checker of the HIP channel kernels (tests/ only).
Not a reference function: the reference has no channel code.  The mappings
are 3GPP TS 36.211 section 7.1 (QPSK, 16QAM, 64QAM, unit average energy);
the demapper is max-log by brute force over all 2^Qm points:
    LLR_b = (min_{s: b=0} |y - h s|^2 - min_{s: b=1} |y - h s|^2) / N0,
positive = bit 1 (the decoder's convention).
"""
from __future__ import annotations

import numpy as np

SCALE = {2: 1.0 / np.sqrt(2.0), 4: 1.0 / np.sqrt(10.0), 6: 1.0 / np.sqrt(42.0)}


def _pam(bits):
    """Level of one dimension from its bits (first bit = sign), 36.211 Gray order."""
    s = 1 - 2 * bits[0]
    if len(bits) == 1:
        return s * 1
    if len(bits) == 2:
        return s * (1 + 2 * bits[1])
    return s * (4 - (1 - 2 * bits[1]) * (2 - (1 - 2 * bits[2])))


def constellation(qm):
    """All 2^qm points; index = bits b0..b(qm-1) read as a binary number, b0 first."""
    pts = np.zeros(1 << qm, np.complex128)
    for v in range(1 << qm):
        b = [(v >> (qm - 1 - j)) & 1 for j in range(qm)]
        pts[v] = SCALE[qm] * complex(_pam(b[0::2]), _pam(b[1::2]))
    return pts


def qam_map(bits, qm):
    bits = np.asarray(bits, np.int64).reshape(-1, qm)
    idx = np.zeros(len(bits), np.int64)
    for j in range(qm):
        idx = (idx << 1) | bits[:, j]
    return constellation(qm)[idx]


def maxlog_llr(y, h, n0, qm):
    """y, h: complex arrays (symbols); n0: scalar or per symbol.  Returns the
    LLRs of every bit, symbol by symbol (qm per symbol)."""
    pts = constellation(qm)
    d = np.abs(np.asarray(y)[:, None] - np.asarray(h)[:, None] * pts[None, :]) ** 2  # [n, 2^qm]
    out = np.zeros((len(y), qm))
    for j in range(qm):
        one = ((np.arange(1 << qm) >> (qm - 1 - j)) & 1).astype(bool)
        out[:, j] = (d[:, ~one].min(axis=1) - d[:, one].min(axis=1)) / n0
    return out.reshape(-1)
