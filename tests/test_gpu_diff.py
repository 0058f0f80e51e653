"""Differentiable decoder on the GPU: forward against the FP64 CPU reference,
hand-written adjoint against central finite differences (FP64 build) and
against the FP64 CPU reverse-mode gradient (float32 build).

Tolerances are stated per test."""
import numpy as np
import pytest
import torch

import oracle as O
from oracle import sw_reverse as R
import turbo_cpp_amd as T
from turbo_cpp_amd.autograd import DiffPlan, WeightedTurboDecoder, weighted_decode
from tests.helpers import bpsk_awgn_llr

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _blocks(k, n, ebn0, seed):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2, (n, k)).astype(np.uint8)
    llr = np.stack([np.concatenate(bpsk_awgn_llr(T.encode_code_block_streams(b), k, ebn0, rng)) for b in bits])
    return bits, llr.astype(np.float32)


def _geom(k, min_segment, warm):
    s = max(min_segment or 32, (k + 63) // 64)
    while k % s:
        s += 1
    m = k // s
    return s, (min(warm or 32, s) if m > 1 else 0)


def _ref_post(llr, scales, k, it, min_segment=0, warm=0, grad=False):
    s, g = _geom(k, min_segment, warm)
    x = torch.tensor(llr, dtype=torch.float64, requires_grad=grad)
    w = torch.tensor(scales, dtype=torch.float64, requires_grad=grad)
    post = R.decode(x, w, k, np.array(O.qpp_perm(k), np.int64), s, g, it)
    return x, w, post


@pytest.mark.parametrize("k,min_segment", [(40, 0), (40, 8), (1024, 0)])
def test_forward_fp64_matches_reference(k, min_segment):
    """FP64 build against the FP64 CPU reference: 1e-9 absolute."""
    _, llr = _blocks(k, 3, 0.5, k)
    sc = np.linspace(0.7, 1.2, 12)
    _, _, ref = _ref_post(llr, sc, k, 6, min_segment)
    plan = DiffPlan([k] * 3, 6, min_segment, dtype=torch.float64)
    _say("fwd fp64", k, min_segment)
    post, _ = plan.forward(torch.tensor(llr, dtype=torch.float64, device="cuda").reshape(-1),
                           torch.tensor(sc, dtype=torch.float64, device="cuda"))
    err = np.max(np.abs(post.cpu().numpy().reshape(3, k) - ref.detach().numpy()))
    assert err < 1e-9, err


def test_forward_plain_scales_equal_exact_decoder():
    """With w = 1 and one segment per block the forward is the reference
    decoder (lte_turbo.cpp:736-813): float32 build within 1e-3 of the exact
    HIP decoder (fp32 rounding of different exp/log forms over 12 halves)."""
    k = 544
    _, llr = _blocks(k, 4, 0.5, 3)
    dec = WeightedTurboDecoder([k] * 4, iterations=6, min_segment=k)
    _say("fwd plain == exact")
    post = dec(torch.tensor(llr, device="cuda")).detach().cpu().numpy()
    ref, _, _ = T.decode_batch_host([k] * 4, llr.reshape(-1), T.DECODER_EXACT, 6)
    assert np.max(np.abs(post.reshape(-1) - ref) / (1.0 + np.abs(ref))) < 1e-3


@pytest.mark.parametrize("min_segment", [0, 8])
def test_adjoint_fp64_gradcheck_central_differences(min_segment):
    """torch.autograd.gradcheck (central finite differences, FP64) of the HIP
    adjoint at K=40: one segment (exact BCJR) and five segments of 8 with
    8-step warm-ups.  eps 1e-6, atol 1e-6, rtol 1e-4."""
    k = 40
    _, llr = _blocks(k, 2, 0.0, 7)
    plan = DiffPlan([k] * 2, 6, min_segment, dtype=torch.float64)
    x = torch.tensor(llr * 0.3, dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.tensor(np.linspace(0.6, 1.1, 12), dtype=torch.float64, device="cuda", requires_grad=True)
    _say("gradcheck", min_segment)
    assert torch.autograd.gradcheck(lambda a, b: weighted_decode(a, b, plan), (x, w), eps=1e-6, atol=1e-6,
                                    rtol=1e-4, nondet_tol=0.0)


@pytest.mark.parametrize("k,min_segment,warm,ebn0",
                         [(40, 8, 8, -1.0), (1024, 0, 0, 0.5), (1024, 0, 16, 0.5), (4096, 0, 0, 0.5)])
def test_adjoint_fp32_matches_fp64_reverse_mode(k, min_segment, warm, ebn0):
    """float32 adjoint against torch.autograd in FP64 on the CPU reference:
    relative L2 error of both gradients below 2e-3.  The SNRs leave
    decision errors, so the gradients are not vanishingly small (where every
    bit is right the BCE gradients underflow towards 0 and any float32
    evaluation, the CPU reference's included, is off by ~1e-2 relative).
    K <= 2048 runs the LDS-staged kernels; at K=4096 a job's block copies
    (68 KB) exceed the 40 KB of LDS the plan gives a job, so the float32
    kernels that stage inputs per lane in global memory run."""
    bits, llr = _blocks(k, 2, ebn0, 11)
    sc = np.linspace(0.8, 1.1, 12)
    x, w, post = _ref_post(llr, sc, k, 6, min_segment, warm, grad=True)
    R.bce_loss(post, torch.tensor(bits)).backward()
    plan = DiffPlan([k] * 2, 6, min_segment, warm)
    xg = torch.tensor(llr, device="cuda", requires_grad=True)
    wg = torch.tensor(sc, dtype=torch.float32, device="cuda", requires_grad=True)
    _say("fp32 adjoint", k, min_segment, warm)
    p = weighted_decode(xg, wg, plan)
    torch.nn.functional.binary_cross_entropy_with_logits(p, torch.tensor(bits, device="cuda").float()).backward()
    for got, ref in ((xg.grad, x.grad), (wg.grad, w.grad)):
        got = got.cpu().double().numpy().reshape(-1)
        ref = ref.numpy().reshape(-1)
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        _say("relative L2 error", rel)
        assert rel < 2e-3, rel


def test_adjoint_fp64_global_staging_matches_reverse_mode():
    """FP64 adjoint at K=1024 against torch.autograd in FP64 on the CPU
    reference: relative L2 error of both gradients below 1e-9.  Two FP64
    blocks of K=1024 need 70 KB of block copies in LDS, more than the plan
    gives a job, so this runs the path that stages inputs per lane in global
    memory in FP64 (the float32 test above covers both paths)."""
    k = 1024
    bits, llr = _blocks(k, 2, 0.5, 13)
    sc = np.linspace(0.8, 1.1, 12)
    x, w, post = _ref_post(llr, sc, k, 6, grad=True)
    R.bce_loss(post, torch.tensor(bits)).backward()
    plan = DiffPlan([k] * 2, 6, dtype=torch.float64)
    xg = torch.tensor(llr, dtype=torch.float64, device="cuda", requires_grad=True)
    wg = torch.tensor(sc, dtype=torch.float64, device="cuda", requires_grad=True)
    _say("fp64 adjoint, global staging", k)
    p = weighted_decode(xg, wg, plan)
    torch.nn.functional.binary_cross_entropy_with_logits(
        p, torch.tensor(bits, dtype=torch.float64, device="cuda")).backward()
    for got, ref in ((xg.grad, x.grad), (wg.grad, w.grad)):
        got = got.cpu().numpy().reshape(-1)
        ref = ref.numpy().reshape(-1)
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        _say("relative L2 error", rel)
        assert rel < 1e-9, rel


@pytest.mark.parametrize("k,n", [(1024, 64), (4096, 6)])
def test_adjoint_fp32_is_deterministic(k, n):
    """The float32 gradients are a fixed sum order throughout (combine terms
    in a fixed order per position, the tail adjoints by segment, the scale
    partials per job and then in job order): two forward + backward runs on
    the same inputs give the same bits.  K=1024: 32 jobs of the LDS-staged
    kernels; K=4096: 6 jobs of the global-staging ones."""
    bits, llr = _blocks(k, n, 0.5, 17)
    plan = DiffPlan([k] * n, 6)
    x = torch.tensor(llr, device="cuda").reshape(-1).contiguous()
    w = torch.linspace(0.8, 1.1, 12, device="cuda")
    gp = torch.tensor(np.random.default_rng(3).standard_normal(plan.post_elems), dtype=torch.float32,
                      device="cuda")
    _say("fp32 determinism", k, n)
    outs = []
    for _ in range(2):
        post, saved = plan.forward(x, w)
        g_llr, g_sc = plan.backward(x, w, saved, gp)
        torch.cuda.synchronize()
        outs.append((post.cpu().numpy(), g_llr.cpu().numpy(), g_sc.cpu().numpy()))
    for a, b in zip(*outs):
        assert np.isfinite(a).all()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_training_lowers_the_loss():
    """A few Adam steps on the scales lower the BCE loss (float32)."""
    k, n = 1024, 64
    bits, llr = _blocks(k, n, 0.5, 5)
    dec = WeightedTurboDecoder([k] * n, iterations=6)
    x = torch.tensor(llr, device="cuda")
    y = torch.tensor(bits, device="cuda").float()
    opt = torch.optim.Adam(dec.parameters(), lr=0.05)
    losses = []
    _say("training")
    for _ in range(15):
        opt.zero_grad()
        loss = torch.nn.functional.binary_cross_entropy_with_logits(dec(x), y)
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    assert losses[-1] < losses[0]
    assert np.all(np.isfinite(dec.scales.detach().cpu().numpy()))


def test_diff_plan_errors():
    with pytest.raises(ValueError):
        DiffPlan([40], iterations=0)
    with pytest.raises(IndexError):
        DiffPlan([41])
    plan = DiffPlan([40])
    with pytest.raises(ValueError):
        plan.forward(torch.zeros(10, device="cuda"), torch.ones(12, device="cuda"))
