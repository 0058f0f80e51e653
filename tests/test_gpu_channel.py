"""Channel front end on the GPU against the CPU restatement (oracle/channel.py):
36.211 Gray QAM mapping, max-log demapping by brute force over the
constellation, noise and fading statistics.

Tolerances: demapped LLRs within 1e-4 (1 + |LLR|) of the float64 brute force
(float32 arithmetic); the fused kernel's LLRs equal the standalone
demapper's on its own y and h bit for bit; noise and fading moments within
3% of their targets on 2^17 symbols."""
import numpy as np
import pytest

import turbo_cpp_amd as T
from oracle import channel as CH

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _cplx(t):
    a = t.cpu().numpy().reshape(-1, 2).astype(np.float64)
    return a[:, 0] + 1j * a[:, 1]


@pytest.mark.parametrize("qm", [2, 4, 6])
@pytest.mark.parametrize("rayleigh", [0, 1])
def test_demapper_matches_brute_force(qm, rayleigh):
    import torch
    rng = np.random.default_rng(qm * 10 + rayleigh)
    nsym = 3000
    y = rng.normal(0, 1.0, (nsym, 2)).astype(np.float32)
    h = (rng.normal(0, 0.7, (nsym, 2)) if rayleigh else np.tile([1.0, 0.0], (nsym, 1))).astype(np.float32)
    n0 = 0.3
    plan = T.ChannelPlan([(1200 * qm, qm, rayleigh, n0), (1800 * qm, qm, rayleigh, n0)])
    llr = torch.empty(plan.n_bits, dtype=torch.float32, device="cuda")
    yd, hd = torch.from_numpy(y.reshape(-1)).cuda(), torch.from_numpy(h.reshape(-1)).cuda()
    _say("demap", qm, rayleigh)
    plan.demap(yd.data_ptr(), hd.data_ptr() if rayleigh else 0, llr.data_ptr())
    ref = CH.maxlog_llr(y[:, 0] + 1j * y[:, 1].astype(np.float64), h[:, 0] + 1j * h[:, 1].astype(np.float64), n0, qm)
    got = llr.cpu().numpy()
    assert np.all(np.abs(got - ref) <= 1e-4 * (1.0 + np.abs(ref))), np.max(np.abs(got - ref))


@pytest.mark.parametrize("qm,rayleigh", [(2, 0), (4, 1), (6, 0), (6, 1)])
def test_fused_channel_matches_mapping_demapper_and_statistics(qm, rayleigh):
    import torch
    rng = np.random.default_rng(7 + qm)
    nsym = 1 << 17
    bits = rng.integers(0, 2, nsym * qm).astype(np.uint8)
    n0 = 0.2
    plan = T.ChannelPlan([(nsym * qm // 2, qm, rayleigh, n0), (nsym * qm // 2, qm, rayleigh, n0)])
    b = torch.from_numpy(bits).cuda()
    llr = torch.empty(plan.n_bits, dtype=torch.float32, device="cuda")
    y = torch.empty(2 * nsym, dtype=torch.float32, device="cuda")
    h = torch.empty(2 * nsym, dtype=torch.float32, device="cuda")
    _say("channel", qm, rayleigh)
    plan.run(1234, b.data_ptr(), llr.data_ptr(), y.data_ptr(), h.data_ptr())
    llr2 = torch.empty_like(llr)
    plan.demap(y.data_ptr(), h.data_ptr(), llr2.data_ptr())
    assert torch.equal(llr, llr2)
    yc, hc = _cplx(y), _cplx(h)
    x = CH.qam_map(bits, qm)
    noise = yc - hc * x
    assert abs(np.mean(np.abs(noise) ** 2) / n0 - 1.0) < 0.03
    assert abs(np.mean(noise)) < 0.01
    if rayleigh:
        assert abs(np.mean(np.abs(hc) ** 2) - 1.0) < 0.03
    else:
        assert np.all(hc == 1.0)
    # same seed, same draw; another seed, another draw
    llr3 = torch.empty_like(llr)
    plan.run(1234, b.data_ptr(), llr3.data_ptr())
    assert torch.equal(llr, llr3)
    plan.run(99, b.data_ptr(), llr3.data_ptr())
    assert not torch.equal(llr, llr3)
    # the fused LLRs are the brute-force max-log LLRs of the kernel's own y, h
    got = llr.cpu().numpy()
    ref = CH.maxlog_llr(yc, hc, n0, qm)
    assert np.all(np.abs(got - ref) <= 1e-4 * (1.0 + np.abs(ref)))
    # sign convention bit 1 > 0: raw bit error rate far below 1/2 at Es/N0 = 7 dB
    assert np.mean((got > 0) != bits.astype(bool)) < 0.3


def test_encode_channel_decode_end_to_end():
    """GPU encoder -> 16QAM over AWGN -> soft de-rate matching, decode and TB
    assembly on the GPU: every transport block decodes at 12 dB Es/N0 (rate
    2/3, Shannon limit 7.3 dB)."""
    import torch
    rng = np.random.default_rng(3)
    tbs = []
    for b in (1000, 6120, 20000, 40000):
        c = T._segment_sizes(b + 24)
        nv = sum(3 * (k + 4) for k in [c[4]] * c[2] + [c[3]] * c[1])
        tbs.append((b, (nv // 2) // 4 * 4, 0, 1, 4))  # rate ~2/3
    pays = [rng.integers(0, 2, t[0]).astype(np.uint8) for t in tbs]
    plan = T.TbPlan(tbs, max_iterations=8, early_stop=True)
    pay = torch.from_numpy(np.concatenate(pays)).cuda()
    rm = torch.empty(plan.rx_floats, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    plan.encode(pay.data_ptr(), rm.data_ptr(), stream=s)
    segs, off = [], 0
    for e in plan.es:
        segs.append((int(e), 4, 0, 10 ** (-12 / 10)))
    ch = T.ChannelPlan(segs)
    llr = torch.empty(plan.rx_floats, dtype=torch.float32, device="cuda")
    ch.run(5, rm.data_ptr(), llr.data_ptr(), stream=s)
    bits = torch.empty(plan.bits_total, dtype=torch.uint8, device="cuda")
    ok = torch.zeros(plan.n_tb, dtype=torch.int32, device="cuda")
    _say("encode -> channel -> decode")
    plan.decode(llr.data_ptr(), bits.data_ptr(), ok.data_ptr(), stream=s)
    torch.cuda.synchronize()
    assert bool(ok.all()) and torch.equal(bits, pay)
