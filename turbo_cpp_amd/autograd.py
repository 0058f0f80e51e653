"""torch.autograd binding of the differentiable decoder (tc_diff_* of the C ABI).

The forward and the adjoint are the library's HIP kernels; this module only
passes torch device pointers and the current stream through ctypes.

    dec = WeightedTurboDecoder([1024] * 2048, iterations=6)   # 12 scales, init 1.0
    post = dec(llr)            # llr: [N, 3(K+4)] float32 on cuda, [d0|d1|d2] per block
    loss = bce(post, bits); loss.backward()                   # grads to dec.scales and llr
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np
import torch

from . import _check, lib


def _bind():
    L = lib()
    if not getattr(L, "_diff_bound", False):
        L.tc_diff_plan_create.argtypes = [np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS"), C.c_size_t,
                                          C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(C.c_void_p)]
        L.tc_diff_plan_destroy.argtypes = [C.c_void_p]
        L.tc_diff_plan_destroy.restype = None
        L.tc_diff_plan_sizes.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")]
        L.tc_diff_forward.argtypes = [C.c_void_p] * 6
        L.tc_diff_backward.argtypes = [C.c_void_p] * 8
        L._diff_bound = True
    return L


class DiffPlan:
    """Work space and geometry of the differentiable decoder for a fixed set
    of code block sizes (float32 or float64)."""

    def __init__(self, ks: Sequence[int], iterations: int = 6, min_segment: int = 0, warmup: int = 0,
                 dtype: torch.dtype = torch.float32):
        if dtype not in (torch.float32, torch.float64):
            raise ValueError("dtype must be float32 or float64")
        L = _bind()
        self.ks = np.ascontiguousarray(ks, dtype=np.uint32)
        self.dtype = dtype
        self.iterations = iterations
        h = C.c_void_p()
        _check(L.tc_diff_plan_create(self.ks, len(self.ks), iterations, min_segment, warmup,
                                     int(dtype == torch.float64), C.byref(h)))
        self._h = h
        sz = np.zeros(4, np.uint64)
        _check(L.tc_diff_plan_sizes(h, sz))
        self.llr_elems, self.post_elems, self.halves, self.saved_elems = (int(x) for x in sz)

    def _check_tensor(self, t: torch.Tensor, n: int, what: str):
        if t.dtype != self.dtype or not t.is_cuda or not t.is_contiguous() or t.numel() != n:
            raise ValueError(f"{what}: need a contiguous {self.dtype} cuda tensor of {n} elements")

    def forward(self, llr: torch.Tensor, scales: torch.Tensor):
        """Returns (post [post_elems], saved extrinsics for backward)."""
        self._check_tensor(llr, self.llr_elems, "llr")
        self._check_tensor(scales, self.halves, "scales")
        post = torch.empty(self.post_elems, dtype=self.dtype, device=llr.device)
        saved = torch.empty(self.saved_elems, dtype=self.dtype, device=llr.device)
        s = torch.cuda.current_stream(llr.device).cuda_stream
        _check(_bind().tc_diff_forward(self._h, llr.data_ptr(), scales.data_ptr(), post.data_ptr(),
                                       saved.data_ptr(), s))
        return post, saved

    def backward(self, llr, scales, saved, grad_post):
        self._check_tensor(grad_post, self.post_elems, "grad_post")
        g_llr = torch.empty(self.llr_elems, dtype=self.dtype, device=llr.device)
        g_sc = torch.empty(self.halves, dtype=self.dtype, device=llr.device)
        s = torch.cuda.current_stream(llr.device).cuda_stream
        _check(_bind().tc_diff_backward(self._h, llr.data_ptr(), scales.data_ptr(), saved.data_ptr(),
                                        grad_post.data_ptr(), g_llr.data_ptr(), g_sc.data_ptr(), s))
        return g_llr, g_sc

    def close(self):
        if getattr(self, "_h", None):
            _bind().tc_diff_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _WeightedDecode(torch.autograd.Function):
    @staticmethod
    def forward(ctx, llr, scales, plan):
        flat = llr.contiguous().reshape(-1)
        sc = scales.contiguous()
        post, saved = plan.forward(flat, sc)
        ctx.plan = plan
        ctx.shape = llr.shape
        ctx.save_for_backward(flat, sc, saved)
        return post

    @staticmethod
    def backward(ctx, grad_post):
        flat, sc, saved = ctx.saved_tensors
        g_llr, g_sc = ctx.plan.backward(flat, sc, saved, grad_post.contiguous().reshape(-1))
        return g_llr.reshape(ctx.shape), g_sc, None


def weighted_decode(llr: torch.Tensor, scales: torch.Tensor, plan: DiffPlan) -> torch.Tensor:
    """Posterior LLRs (all blocks' K values back to back; [N, K] when llr is
    [N, 3(K+4)] with one K) of the scaled sliding-window decoder."""
    post = _WeightedDecode.apply(llr, scales, plan)
    if llr.dim() == 2 and len(set(plan.ks.tolist())) == 1:
        post = post.reshape(llr.shape[0], int(plan.ks[0]))
    return post


class WeightedTurboDecoder(torch.nn.Module):
    """Turbo decoder with one learnable extrinsic scale per half-iteration
    (2 x iterations parameters, initialised to 1.0 = the plain decoder)."""

    def __init__(self, ks: Sequence[int], iterations: int = 6, min_segment: int = 0, warmup: int = 0,
                 dtype: torch.dtype = torch.float32):
        super().__init__()
        self.plan = DiffPlan(ks, iterations, min_segment, warmup, dtype)
        self.scales = torch.nn.Parameter(torch.ones(2 * iterations, dtype=dtype, device="cuda"))

    def forward(self, llr: torch.Tensor) -> torch.Tensor:
        return weighted_decode(llr, self.scales, self.plan)
