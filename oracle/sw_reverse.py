"""FP64 CPU reference of the differentiable decoder, with reverse mode by torch.autograd.

TEST INFRASTRUCTURE ONLY: the checker of the HIP forward and adjoint kernels
of the differentiable decoder (tests/ and bench.py's cpu_baseline leg only).

The decoder (not a reference function; the reference has no differentiable
path) is a sliding-window Log-MAP turbo decoder with learnable extrinsic
scales.  Per SISO half (lte_turbo.cpp:410-486 arithmetic in exact form):
  gamma_t(s,u) = 0.5 (x_u (sys_t + la_t) + z_p par_t), x, z = +-1 for bits 1/0,
  alpha/beta recursions by the Jacobian logarithm (logaddexp), renormalised
  to a maximum of 0 after every step, ext_t = L1_t - L0_t - sys_t - la_t.
The block of K steps is cut into M = K / S segments.  Segment q owns steps
[qS, qS + S); its alpha starts g steps early at qS - g from all-zero metrics
(or at step 0 from state 0 when qS - g <= 0), its beta starts g steps late at
qS + S + g from all-zero metrics (or at K from the tail beta of
lte_turbo.cpp:357-378 when qS + S + g >= K).  M = 1 is the reference's
full-length SISO.
Iterations (lte_turbo.cpp:776-790) with scales w[0 .. 2I-1]: the extrinsic of
half h is multiplied by w[h] wherever it is used, i.e.
  la2 = w[2i] * pi(e1),  la1 = w[2i+1] * e2 (natural order),
  post = d0 + w[2I-2] * e1 + w[2I-1] * e2   (last iteration's e1, e2).
With all w = 1 this is the reference decoder's iteration.
"""
from __future__ import annotations

import numpy as np
import torch

NEG = -1.0e30


def _trellis():
    """(pred_state[n, j], pred_u[n, j], pred_par[n, j]) for the two
    transitions into each state n, and (next[s, u], par[s, u])."""
    nxt = np.zeros((8, 2), np.int64)
    par = np.zeros((8, 2), np.int64)
    for s in range(8):
        s0, s1, s2 = s & 1, (s >> 1) & 1, (s >> 2) & 1
        for u in range(2):
            fb = u ^ s1 ^ s2
            par[s, u] = fb ^ s0 ^ s2
            nxt[s, u] = (fb << 2) | (s2 << 1) | s1
    pred = [[] for _ in range(8)]
    for s in range(8):
        for u in range(2):
            pred[nxt[s, u]].append((s, u, par[s, u]))
    ps = np.array([[p[0] for p in pred[n]] for n in range(8)])
    pu = np.array([[p[1] for p in pred[n]] for n in range(8)])
    pp = np.array([[p[2] for p in pred[n]] for n in range(8)])
    return nxt, par, ps, pu, pp


NXT, PAR, PS, PU, PP = _trellis()


def _term_inputs(s):
    out, cur = [], s
    for _ in range(3):
        u = ((cur >> 1) ^ (cur >> 2)) & 1
        out.append(u)
        cur = NXT[cur, u]
    return out


def tail_beta(ts, tp):
    """ts, tp: [N, 3] -> [N, 8] (lte_turbo.cpp:357-378, normalised)."""
    cols = []
    for s in range(8):
        cur, m = s, torch.zeros_like(ts[:, 0])
        for t, u in enumerate(_term_inputs(s)):
            p = PAR[cur, u]
            m = m + 0.5 * ((1.0 if u else -1.0) * ts[:, t]) + 0.5 * ((1.0 if p else -1.0) * tp[:, t])
            cur = NXT[cur, u]
        cols.append(m if cur == 0 else torch.full_like(m, NEG))
    b = torch.stack(cols, dim=1)
    return b - b.max(dim=1, keepdim=True).values


def _gamma(sa, par, u, p):
    return 0.5 * ((2.0 * u - 1.0) * sa + (2.0 * p - 1.0) * par)


def siso_sw(sys, par, la, ts, tp, seg, warm):
    """One SISO half on [N, K] inputs; returns ext [N, K]."""
    n, k = sys.shape
    m = k // seg
    assert m * seg == k and 0 <= warm
    sa = sys + la
    starts = np.arange(m) * seg
    # alpha over local steps j = 0 .. warm + seg - 1, t = qS - warm + j
    a = torch.zeros(n, m, 8, dtype=sys.dtype)
    first = torch.from_numpy(starts - warm <= 0)
    a0 = torch.full((8,), NEG, dtype=sys.dtype)
    a0[0] = 0.0
    a = torch.where(first[None, :, None], a0.expand(n, m, 8), a)
    alphas = []  # alpha_t for t = qS .. qS + S - 1
    for j in range(warm + seg):
        t = starts - warm + j
        if j >= warm:
            alphas.append(a)
        act = torch.from_numpy(t >= 0)
        tc = np.clip(t, 0, k - 1)
        sa_t, pr_t = sa[:, tc], par[:, tc]  # [N, M]
        cand = []
        for jj in range(2):
            g = _gamma(sa_t[..., None], pr_t[..., None], torch.from_numpy(PU[:, jj]).to(sys.dtype),
                       torch.from_numpy(PP[:, jj]).to(sys.dtype))
            cand.append(a[:, :, PS[:, jj]] + g)
        nxt = torch.logaddexp(cand[0], cand[1])
        nxt = nxt - nxt.max(dim=2, keepdim=True).values
        a = torch.where(act[None, :, None], nxt, a)
    # beta over local steps j = warm + seg - 1 .. 0, t = qS + j
    last = torch.from_numpy(starts + seg + warm >= k)
    tb = tail_beta(ts, tp)  # [N, 8]
    b = torch.where(last[None, :, None], tb[:, None, :].expand(n, m, 8), torch.zeros(n, m, 8, dtype=sys.dtype))
    betas = [None] * seg  # beta_{t+1} for t = qS .. qS + S - 1
    for j in range(warm + seg - 1, -1, -1):
        t = starts + j
        if j < seg:
            betas[j] = b
        act = torch.from_numpy(t < k)
        tc = np.clip(t, 0, k - 1)
        sa_t, pr_t = sa[:, tc], par[:, tc]
        cand = []
        for u in range(2):
            g = _gamma(sa_t[..., None], pr_t[..., None], torch.tensor(float(u), dtype=sys.dtype),
                       torch.from_numpy(PAR[:, u]).to(sys.dtype))
            cand.append(b[:, :, NXT[:, u]] + g)
        nb = torch.logaddexp(cand[0], cand[1])
        nb = nb - nb.max(dim=2, keepdim=True).values
        b = torch.where(act[None, :, None], nb, b)
    ext = []
    for j in range(seg):
        t = starts + j
        sa_t, pr_t = sa[:, t], par[:, t]
        lu = []
        for u in range(2):
            g = _gamma(sa_t[..., None], pr_t[..., None], torch.tensor(float(u), dtype=sys.dtype),
                       torch.from_numpy(PAR[:, u]).to(sys.dtype))
            lu.append(torch.logsumexp(alphas[j] + g + betas[j][:, :, NXT[:, u]], dim=2))
        ext.append(lu[1] - lu[0] - sa_t)  # [N, M]
    return torch.stack(ext, dim=2).reshape(n, k)


def tail_llrs(llr, k):
    """(us, up, ls, lp) [N, 3] each, the decoder's read-out (lte_turbo.cpp:758-773)."""
    d = k + 4
    d0, d1, d2 = llr[:, :d], llr[:, d:2 * d], llr[:, 2 * d:]
    us = torch.stack([d0[:, k], d0[:, k + 2], torch.zeros_like(d0[:, k])], 1)
    up = torch.stack([d1[:, k], d0[:, k + 1], d1[:, k + 2]], 1)
    ls = torch.stack([d1[:, k + 3], d2[:, k], d2[:, k + 2]], 1)
    lp = torch.stack([d2[:, k + 1], d0[:, k + 3], d2[:, k + 3]], 1)
    return us, up, ls, lp


def decode(llr, scales, k, perm, seg, warm, iterations):
    """Weighted sliding-window decoder.  llr [N, 3(K+4)] float64, scales
    [2 * iterations]; perm: QPP permutation (numpy).  Returns post [N, K]."""
    d = k + 4
    d0, d1, d2 = llr[:, :d], llr[:, d:2 * d], llr[:, 2 * d:]
    us, up, ls, lp = tail_llrs(llr, k)
    inv = np.argsort(perm)
    sys1, sys2 = d0[:, :k], d0[:, perm]
    la1 = torch.zeros_like(sys1)
    e1 = e2 = None
    for it in range(iterations):
        e1 = siso_sw(sys1, d1[:, :k], la1, us, up, seg, warm)
        la2 = scales[2 * it] * e1[:, perm]
        e2p = siso_sw(sys2, d2[:, :k], la2, ls, lp, seg, warm)
        e2 = e2p[:, inv]
        la1 = scales[2 * it + 1] * e2
    i = iterations
    return sys1 + scales[2 * i - 2] * e1 + scales[2 * i - 1] * e2


def bce_loss(post, bits):
    """Mean binary cross-entropy of APP LLRs (bit 1 > 0) against the true bits."""
    return torch.nn.functional.binary_cross_entropy_with_logits(post, bits.to(post.dtype))
