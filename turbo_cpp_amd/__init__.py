"""turbo_cpp_amd: MI355X-native LTE turbo decoder (HIP, gfx950).

Python mirror of the Wiseung/Turbo_CPP C++ API (include/turbo_cpp/lte_turbo.h)
over the C ABI of ``libturbo_cpp_amd.so`` (include/turbo_cpp/c_api.h), plus
the batched device-side decoder (:class:`Plan`).  Names, argument meaning and
error behaviour follow the reference; its exception types map to
``ValueError`` (std::invalid_argument), ``IndexError`` (std::out_of_range)
and ``RuntimeError`` (std::runtime_error).

The library must be built (``python -m turbo_cpp_amd.build``); importing
fails loudly when it is missing -- there is no CPU fallback for the GPU
entry points.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import enum
import os
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libturbo_cpp_amd.so")

TC_OK, TC_ERR_INVALID_ARGUMENT, TC_ERR_OUT_OF_RANGE, TC_ERR_RUNTIME = 0, 1, 2, 3
DECODER_EXACT, DECODER_WINDOW_LOG, DECODER_WINDOW_LIN, DECODER_EXACT_SERIAL = 0, 1, 2, 3
STAGING_AUTO, STAGING_FUSED, STAGING_SEPARATE = 0, 1, 2  # tc_tb_plan_set_staging
CRC24A, CRC24B, CRC16, CRC8 = 0, 1, 2, 3


class TcCbDesc(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("block_index", "k_r", "d_r", "e_r", "filler_count", "qpp_index")]


class TcOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_uint32), ("early_stop", C.c_int32), ("decoder", C.c_int32),
                ("min_segment", C.c_uint32), ("warmup", C.c_uint32), ("waves_per_cu", C.c_uint32),
                ("extrinsic_bits", C.c_uint32), ("max_waves_per_block", C.c_uint32)]


class TcTbDesc(C.Structure):
    _fields_ = [("transport_block_bits", C.c_uint64), ("g_total", C.c_uint64), ("rv", C.c_uint32),
                ("num_layers", C.c_uint32), ("mod_order", C.c_uint32), ("code_block_crc", C.c_uint32)]


class TcChannelSeg(C.Structure):
    _fields_ = [("n_bits", C.c_uint64), ("qm", C.c_uint32), ("rayleigh", C.c_uint32), ("n0", C.c_float),
                ("pad", C.c_uint32)]


class TcSegmentInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("k", "segment_length", "segments", "blocks_per_wave", "warmup",
                                          "extrinsic_bits", "waves_per_block")]


_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")

_lib = None


def lib() -> C.CDLL:
    """The loaded C ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m turbo_cpp_amd.build`")
        # One HIP runtime per process.  PyTorch wheels carry their own
        # libamdhip64 whose dependents ask for "libamdhip64.so"; the library
        # here asks for "libamdhip64.so.7".  Loading torch first lets this
        # library bind to torch's runtime; loading this library first would
        # make torch map a second runtime that then finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.tc_last_error.restype = C.c_char_p
        L.tc_device_count.argtypes = [C.POINTER(C.c_int32)]
        L.tc_segment_geometry.argtypes = [C.c_uint32, C.POINTER(TcOptions), C.POINTER(TcSegmentInfo)]
        L.tc_decode_code_block.argtypes = [C.POINTER(TcCbDesc), _f32p, C.c_size_t, _f32p, C.c_size_t, _f32p,
                                           C.c_size_t, C.POINTER(TcOptions), _f32p, C.POINTER(C.c_int32),
                                           C.POINTER(C.c_uint32)]
        L.tc_decode_code_block_noalloc.argtypes = L.tc_decode_code_block.argtypes
        L.tc_plan_create.argtypes = [_u32p, C.c_void_p, C.c_size_t, C.POINTER(TcOptions), C.POINTER(C.c_void_p)]
        L.tc_plan_destroy.argtypes = [C.c_void_p]
        L.tc_plan_destroy.restype = None
        L.tc_plan_llr_floats.argtypes = [C.c_void_p]
        L.tc_plan_llr_floats.restype = C.c_size_t
        L.tc_plan_post_floats.argtypes = [C.c_void_p]
        L.tc_plan_post_floats.restype = C.c_size_t
        L.tc_plan_decode.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.tc_decode_batch_host.argtypes = [_u32p, C.c_void_p, C.c_size_t, C.POINTER(TcOptions), _f32p, _f32p,
                                           C.c_void_p, C.c_void_p]
        L.tc_crc_append.argtypes = [C.c_int32, _u8p, C.c_size_t, _u8p]
        L.tc_crc_check.argtypes = [C.c_int32, _u8p, C.c_size_t, C.POINTER(C.c_int32)]
        L.tc_segment.argtypes = [_u8p, C.c_size_t, _u64p, C.c_void_p]
        L.tc_qpp_entry.argtypes = [C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.tc_encode_code_block.argtypes = [_u8p, C.c_uint32, C.c_uint32, _u8p, _u8p, _u8p]
        L.tc_compute_e.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                   C.POINTER(C.c_uint64)]
        L.tc_rate_match.argtypes = [_u8p, _u8p, _u8p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, _u8p]
        L.tc_derate_match.argtypes = [_u8p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, _f32p]
        L.tc_decode_code_block_cpu.argtypes = [C.POINTER(TcCbDesc), _f32p, _f32p, _f32p, C.c_uint32, C.c_int32,
                                               _f32p, C.POINTER(C.c_int32), C.POINTER(C.c_uint32)]
        L.tc_tb_plan_create.argtypes = [C.POINTER(TcTbDesc), C.c_size_t, C.POINTER(TcOptions),
                                        C.POINTER(C.c_void_p)]
        L.tc_tb_plan_destroy.argtypes = [C.c_void_p]
        L.tc_tb_plan_destroy.restype = None
        L.tc_tb_plan_sizes.argtypes = [C.c_void_p, _u64p]
        L.tc_tb_plan_code_blocks.argtypes = [C.c_void_p, _u32p, _u32p, _u64p]
        L.tc_tb_plan_decode.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32] + [C.c_void_p] * 6
        L.tc_plan_decode_rx.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int32,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.tc_tb_plan_decode_multi.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p] + [C.c_void_p] * 6
        L.tc_tb_plan_set_hold.argtypes = [C.c_void_p, C.c_int32]
        L.tc_tb_plan_set_staging.argtypes = [C.c_void_p, C.c_int32]
        L.tc_tb_plan_soft_combine.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
        L.tc_tb_plan_encode.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.tc_encode_transport_blocks_host.argtypes = [C.POINTER(TcTbDesc), C.c_size_t, _u8p, _u8p, C.c_void_p]
        L.tc_decode_transport_blocks_host.argtypes = [C.POINTER(TcTbDesc), C.c_size_t, C.POINTER(TcOptions), _f32p,
                                                      _u8p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.tc_channel_plan_create.argtypes = [C.POINTER(TcChannelSeg), C.c_size_t, C.POINTER(C.c_void_p)]
        L.tc_channel_plan_destroy.argtypes = [C.c_void_p]
        L.tc_channel_plan_destroy.restype = None
        L.tc_channel_plan_sizes.argtypes = [C.c_void_p, _u64p]
        L.tc_channel_llr.argtypes = [C.c_void_p, C.c_uint64] + [C.c_void_p] * 5
        L.tc_qam_demap.argtypes = [C.c_void_p] * 5
        L.tc_derate_plan_create.argtypes = [_u32p, C.c_void_p, _u64p, C.c_void_p, C.c_size_t,
                                            C.POINTER(C.c_void_p)]
        L.tc_derate_plan_destroy.argtypes = [C.c_void_p]
        L.tc_derate_plan_destroy.restype = None
        L.tc_derate_plan_rx_floats.argtypes = [C.c_void_p]
        L.tc_derate_plan_rx_floats.restype = C.c_size_t
        L.tc_derate_plan_llr_floats.argtypes = [C.c_void_p]
        L.tc_derate_plan_llr_floats.restype = C.c_size_t
        L.tc_derate_plan_run.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
        L.tc_derate_match_soft.argtypes = [_f32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32,
                                           _f32p]
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status == TC_OK:
        return
    msg = lib().tc_last_error().decode(errors="replace")
    if status == TC_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == TC_ERR_OUT_OF_RANGE:
        raise IndexError(msg)
    raise RuntimeError(msg)


def device_count() -> int:
    n = C.c_int32(0)
    _check(lib().tc_device_count(C.byref(n)))
    return n.value


def _u8(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint8)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


# ------------------------------------------------------- API data types -----
class LteChannelType(enum.IntEnum):
    UL_SCH = 0
    DL_SCH = 1
    PCH = 2
    MCH = 3


class DecoderBackend(enum.IntEnum):
    CpuReference = 0
    GpuStage1Exact = 1
    GpuStage2Windowed = 2


@dataclasses.dataclass
class LteTurboCodecParams:
    channel_type: LteChannelType = LteChannelType.DL_SCH
    transport_block_bits: int = 0
    g_total: int = 0
    rv: int = 0
    num_layers: int = 1
    mod_order: int = 2
    n_soft: int = 0
    max_iterations: int = 8
    enable_early_stop: bool = False


@dataclasses.dataclass
class CodeBlockDescriptor:
    block_index: int = 0
    k_r: int = 0
    d_r: int = 0
    e_r: int = 0
    filler_count: int = 0
    qpp_index: int = 0

    def c(self) -> TcCbDesc:
        return TcCbDesc(self.block_index, self.k_r, self.d_r, self.e_r, self.filler_count, self.qpp_index)


@dataclasses.dataclass
class EncodedCodeBlock:
    descriptor: CodeBlockDescriptor
    cb_bits: np.ndarray
    systematic: np.ndarray
    parity0: np.ndarray
    parity1: np.ndarray
    rate_matched_bits: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.uint8))


@dataclasses.dataclass
class EncodedTransportBlock:
    params: LteTurboCodecParams
    tb_bits_with_crc: np.ndarray
    code_blocks: List[EncodedCodeBlock]
    transport_bits: np.ndarray


@dataclasses.dataclass
class DecodedTransportBlock:
    tb_bits: np.ndarray
    code_block_crc_ok: List[bool]
    transport_block_crc_ok: bool
    iterations_used: int
    posterior_llr: Optional[np.ndarray]


@dataclasses.dataclass
class SegmentationResult:
    c: int
    c_plus: int
    c_minus: int
    k_plus: int
    k_minus: int
    filler_bits: int
    l: int
    code_blocks: List[np.ndarray]


def to_string(channel_type: LteChannelType) -> str:
    return LteChannelType(channel_type).name


# ---------------------------------------------------------- host codec ------
def qpp_entry_for_k(k: int):
    f1, f2, idx = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _check(lib().tc_qpp_entry(k, C.byref(f1), C.byref(f2), C.byref(idx)))
    return (k, f1.value, f2.value)


def qpp_index(k: int) -> int:
    f1, f2, idx = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _check(lib().tc_qpp_entry(k, C.byref(f1), C.byref(f2), C.byref(idx)))
    return idx.value


def qpp_interleaver_table():
    """All 188 (K, f1, f2) entries (36.212 Table 5.1.3-3)."""
    ks = list(range(40, 513, 8)) + list(range(528, 1025, 16)) + list(range(1056, 2049, 32)) + \
        list(range(2112, 6145, 64))
    return [qpp_entry_for_k(k) for k in ks]


def qpp_permutation(k: int) -> np.ndarray:
    _, f1, f2 = qpp_entry_for_k(k)
    i = np.arange(k, dtype=np.int64)
    return (f1 * i + f2 * i * i) % k


def qpp_inverse_permutation(k: int) -> np.ndarray:
    p = qpp_permutation(k)
    inv = np.empty_like(p)
    inv[p] = np.arange(k)
    return inv


def _crc_append(kind: int, bits) -> np.ndarray:
    bits = _u8(bits)
    out = np.zeros(len(bits) + (24 if kind < 2 else 16 if kind == 2 else 8), np.uint8)
    _check(lib().tc_crc_append(kind, bits, len(bits), out))
    return out


def _crc_check(kind: int, bits) -> bool:
    bits = _u8(bits)
    ok = C.c_int32(0)
    _check(lib().tc_crc_check(kind, bits, len(bits), C.byref(ok)))
    return bool(ok.value)


def append_crc24a(bits): return _crc_append(CRC24A, bits)
def append_crc24b(bits): return _crc_append(CRC24B, bits)
def append_crc16(bits): return _crc_append(CRC16, bits)
def append_crc8(bits): return _crc_append(CRC8, bits)
def check_crc24a(bits): return _crc_check(CRC24A, bits)
def check_crc24b(bits): return _crc_check(CRC24B, bits)
def check_crc16(bits): return _crc_check(CRC16, bits)
def check_crc8(bits): return _crc_check(CRC8, bits)


def segment_transport_block(bits_with_crc) -> SegmentationResult:
    bits = _u8(bits_with_crc)
    seg = np.zeros(7, np.uint64)
    _check(lib().tc_segment(bits, len(bits), seg, None))
    c, cp, cm, kp, km, f, l = (int(v) for v in seg)
    blocks = np.zeros(cp * kp + cm * km, np.uint8)
    _check(lib().tc_segment(bits, len(bits), seg, blocks.ctypes.data))
    out, off = [], 0
    for r in range(c):
        k = km if r < cm else kp
        out.append(blocks[off:off + k].copy())
        off += k
    return SegmentationResult(c, cp, cm, kp, km, f, l, out)


def encode_code_block_streams(cb_bits, filler: int = 0):
    """(d0, d1, d2) uint8 streams of K+4 bits each."""
    cb_bits = _u8(cb_bits)
    k = len(cb_bits)
    d = [np.zeros(k + 4, np.uint8) for _ in range(3)]
    _check(lib().tc_encode_code_block(cb_bits, k, filler, d[0], d[1], d[2]))
    return d


def turbo_encode_code_block(descriptor: CodeBlockDescriptor, cb_bits) -> EncodedCodeBlock:
    cb_bits = _u8(cb_bits)
    if len(cb_bits) != descriptor.k_r:
        raise ValueError("code block length mismatch")
    d0, d1, d2 = encode_code_block_streams(cb_bits, descriptor.filler_count)
    desc = dataclasses.replace(descriptor, d_r=descriptor.k_r + 4)
    return EncodedCodeBlock(desc, cb_bits.copy(), d0, d1, d2)


def compute_e(r: int, g_total: int, c: int, num_layers: int, mod_order: int) -> int:
    e = C.c_uint64()
    _check(lib().tc_compute_e(r, g_total, c, num_layers, mod_order, C.byref(e)))
    return e.value


def rate_match(enc: EncodedCodeBlock, e: int, rv: int) -> np.ndarray:
    out = np.zeros(e, np.uint8)
    _check(lib().tc_rate_match(enc.systematic, enc.parity0, enc.parity1, len(enc.systematic),
                               enc.descriptor.filler_count, e, rv, out))
    return out


def encode_transport_block(params: LteTurboCodecParams, transport_bits) -> EncodedTransportBlock:
    """lte_turbo.cpp:690-720."""
    bits = _u8(transport_bits)
    if params.transport_block_bits != len(bits):
        raise ValueError("transport block size does not match params")
    with_crc = append_crc24a(bits)
    seg = segment_transport_block(with_crc)
    blocks = []
    for r, cb in enumerate(seg.code_blocks):
        d = CodeBlockDescriptor(r, len(cb), len(cb) + 4,
                                compute_e(r, params.g_total, seg.c, params.num_layers, params.mod_order),
                                seg.filler_bits if r == 0 else 0, qpp_index(len(cb)))
        enc = turbo_encode_code_block(d, cb)
        enc.rate_matched_bits = rate_match(enc, d.e_r, params.rv)
        blocks.append(enc)
    return EncodedTransportBlock(params, with_crc, blocks, bits.copy())


def derate_match_code_block(params: LteTurboCodecParams, descriptor: CodeBlockDescriptor,
                            rate_matched_bits) -> np.ndarray:
    bits = _u8(rate_matched_bits)
    out = np.zeros(3 * descriptor.d_r, np.float32)
    _check(lib().tc_derate_match(bits, len(bits), descriptor.d_r, descriptor.filler_count, params.rv, out))
    return out


# ------------------------------------------------------------- decoders -----
def _spans(d0, d1, d2):
    return _f32(d0), _f32(d1), _f32(d2)


def turbo_decode_code_block_cpu(descriptor: CodeBlockDescriptor, d0, d1, d2, max_iterations: int,
                                enable_early_stop: bool):
    """Returns (posterior, crc_ok, iterations_used)."""
    d0, d1, d2 = _spans(d0, d1, d2)
    if not (len(d0) == len(d1) == len(d2) == descriptor.d_r):
        raise ValueError("invalid LLR sizes for turbo decode")
    post = np.zeros(descriptor.k_r, np.float32)
    ok, it = C.c_int32(0), C.c_uint32(0)
    cd = descriptor.c()
    _check(lib().tc_decode_code_block_cpu(C.byref(cd), d0, d1, d2, max_iterations, int(enable_early_stop), post,
                                          C.byref(ok), C.byref(it)))
    return post, bool(ok.value), it.value


def options(decoder: int = DECODER_WINDOW_LOG, max_iterations: int = 8, early_stop: bool = False,
            min_segment: int = 0, warmup: int = 0, waves_per_cu: int = 0, extrinsic_bits: int = 0,
            max_waves_per_block: int = 0) -> TcOptions:
    """tc_options; extrinsic_bits: width of the extrinsics held between
    half-iterations (16 or 32; 0 = the decoder's default: 16 for the window
    decoders, 32 for the exact ones); max_waves_per_block: waves one code
    block may span (1-3, 0 = 1; window log/lin decoders), which lets the
    segment length go down to ceil(K / (64 n))."""
    return TcOptions(max_iterations, int(early_stop), decoder, min_segment, warmup, waves_per_cu, extrinsic_bits,
                     max_waves_per_block)


def _gpu_decode(decoder, descriptor, d0, d1, d2, max_iterations, enable_early_stop, **kw):
    d0, d1, d2 = _spans(d0, d1, d2)
    post = np.zeros(descriptor.k_r, np.float32)
    ok, it = C.c_int32(0), C.c_uint32(0)
    o = options(decoder, max_iterations, enable_early_stop, **kw)
    cd = descriptor.c()
    _check(lib().tc_decode_code_block(C.byref(cd), d0, len(d0), d1, len(d1), d2, len(d2), C.byref(o), post,
                                      C.byref(ok), C.byref(it)))
    return post, bool(ok.value), it.value


def decode_code_block_noalloc(decoder, descriptor, d0, d1, d2, max_iterations, enable_early_stop, **kw):
    """tc_decode_code_block_noalloc: the per-block GPU decode without device
    allocation (inputs, outputs and work space in the library's static device
    array).  Calls that overlap on one GPU race."""
    d0, d1, d2 = _spans(d0, d1, d2)
    post = np.zeros(descriptor.k_r, np.float32)
    ok, it = C.c_int32(0), C.c_uint32(0)
    o = options(decoder, max_iterations, enable_early_stop, **kw)
    cd = descriptor.c()
    _check(lib().tc_decode_code_block_noalloc(C.byref(cd), d0, len(d0), d1, len(d1), d2, len(d2), C.byref(o), post,
                                              C.byref(ok), C.byref(it)))
    return post, bool(ok.value), it.value


def turbo_decode_code_block_gpu_stage1(descriptor, d0, d1, d2, max_iterations, enable_early_stop):
    """Exact Log-MAP on the GPU (GpuStage1Exact)."""
    return _gpu_decode(DECODER_EXACT, descriptor, d0, d1, d2, max_iterations, enable_early_stop)


def turbo_decode_code_block_gpu_stage2(descriptor, d0, d1, d2, max_iterations, enable_early_stop, **kw):
    """Segment-parallel decoder on the GPU (GpuStage2Windowed)."""
    return _gpu_decode(DECODER_WINDOW_LOG, descriptor, d0, d1, d2, max_iterations, enable_early_stop, **kw)


def turbo_decode_code_block_gpu_stage1_tile8_experimental(descriptor, d0, d1, d2, max_iterations,
                                                          enable_early_stop):
    """Exact Log-MAP on the GPU, one full-length recursion per block (the
    serial form; the same posteriors bit for bit as the segment-parallel one)."""
    return _gpu_decode(DECODER_EXACT_SERIAL, descriptor, d0, d1, d2, max_iterations, enable_early_stop)


def segment_geometry(k: int, decoder: int = DECODER_WINDOW_LOG, min_segment: int = 0, warmup: int = 0,
                     extrinsic_bits: int = 0, max_waves_per_block: int = 0):
    o = options(decoder, 8, False, min_segment, warmup, 0, extrinsic_bits, max_waves_per_block)
    g = TcSegmentInfo()
    _check(lib().tc_segment_geometry(k, C.byref(o), C.byref(g)))
    return dict(k=g.k, segment_length=g.segment_length, segments=g.segments,
                blocks_per_wave=g.blocks_per_wave, warmup=g.warmup, extrinsic_bits=g.extrinsic_bits,
                waves_per_block=g.waves_per_block)


def decode_batch_host(ks: Sequence[int], llr: np.ndarray, decoder: int = DECODER_WINDOW_LOG,
                      max_iterations: int = 8, early_stop: bool = False, fillers=None,
                      min_segment: int = 0, warmup: int = 0, extrinsic_bits: int = 0,
                      max_waves_per_block: int = 0):
    """Decode packed host LLRs (block r: [d0|d1|d2], 3(K_r+4) floats).

    Returns (posterior (sum K_r floats), crc_ok[n], iterations_used[n])."""
    ks = np.ascontiguousarray(ks, dtype=np.uint32)
    n = len(ks)
    llr = _f32(llr).reshape(-1)
    if llr.size != int(np.sum(3 * (ks.astype(np.int64) + 4))):
        raise ValueError("packed LLR length does not match the code block sizes")
    post = np.zeros(int(ks.astype(np.int64).sum()), np.float32)
    crc = np.zeros(n, np.int32)
    its = np.zeros(n, np.uint32)
    fil = None if fillers is None else np.ascontiguousarray(fillers, dtype=np.uint32)
    o = options(decoder, max_iterations, early_stop, min_segment, warmup, 0, extrinsic_bits, max_waves_per_block)
    _check(lib().tc_decode_batch_host(ks, None if fil is None else fil.ctypes.data, n, C.byref(o), llr, post,
                                      crc.ctypes.data, its.ctypes.data))
    return post, crc.astype(bool), its


class Plan:
    """Batched device decoder for a fixed set of code blocks (tc_plan_*).

    ``decode`` takes device pointers (e.g. ``tensor.data_ptr()``) in the packed
    layout and an optional hipStream_t handle (e.g.
    ``torch.cuda.current_stream().cuda_stream``)."""

    def __init__(self, ks: Sequence[int], decoder: int = DECODER_WINDOW_LOG, max_iterations: int = 8,
                 early_stop: bool = False, fillers=None, min_segment: int = 0, warmup: int = 0,
                 waves_per_cu: int = 0, extrinsic_bits: int = 0, max_waves_per_block: int = 0):
        self.ks = np.ascontiguousarray(ks, dtype=np.uint32)
        self._fil = None if fillers is None else np.ascontiguousarray(fillers, dtype=np.uint32)
        o = options(decoder, max_iterations, early_stop, min_segment, warmup, waves_per_cu, extrinsic_bits,
                    max_waves_per_block)
        h = C.c_void_p()
        _check(lib().tc_plan_create(self.ks, None if self._fil is None else self._fil.ctypes.data, len(self.ks),
                                    C.byref(o), C.byref(h)))
        self._h = h
        self.llr_floats = lib().tc_plan_llr_floats(h)
        self.post_floats = lib().tc_plan_post_floats(h)

    def decode(self, llr_ptr: int, post_ptr: int, crc_ptr: int = 0, iters_ptr: int = 0, stream: int = 0):
        _check(lib().tc_plan_decode(self._h, llr_ptr, post_ptr, crc_ptr or None, iters_ptr or None,
                                    stream or None))

    def decode_rx(self, derate: "DeratePlan", rx_ptrs, post_ptr: int, rvs=None, accumulate: bool = False,
                  keep_ptr: int = 0, crc_ptr: int = 0, iters_ptr: int = 0, stream: int = 0):
        """Decode from received LLRs of 1..4 transmissions (device pointers),
        de-rate matched inside the decoder's input staging (tc_plan_decode_rx)."""
        n = len(rx_ptrs)
        ptrs = (C.c_void_p * max(n, 1))(*rx_ptrs)
        rv = None if rvs is None else (C.c_int32 * n)(*rvs)
        _check(lib().tc_plan_decode_rx(self._h, derate._h, n, ptrs, rv, int(accumulate), keep_ptr or None,
                                       post_ptr or None, crc_ptr or None, iters_ptr or None, stream or None))

    def close(self):
        if getattr(self, "_h", None):
            lib().tc_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ChannelPlan:
    """Channel front end on the GPU (tc_channel_*): segments of (n_bits, qm,
    rayleigh, n0) back to back.  ``run`` maps coded bits to QAM, fades, adds
    noise and demaps (max-log) in one pass; ``demap`` demaps given y (and h)."""

    def __init__(self, segs):
        arr = (TcChannelSeg * max(len(segs), 1))()
        for i, (nb, qm, ray, n0) in enumerate(segs):
            arr[i] = TcChannelSeg(int(nb), int(qm), int(bool(ray)), float(n0), 0)
        self._segs = arr
        h = C.c_void_p()
        _check(lib().tc_channel_plan_create(arr, len(segs), C.byref(h)))
        self._h = h
        sz = np.zeros(2, np.uint64)
        _check(lib().tc_channel_plan_sizes(h, sz))
        self.n_bits, self.n_sym = int(sz[0]), int(sz[1])

    def run(self, seed: int, bits_ptr: int, llr_ptr: int, y_ptr: int = 0, h_ptr: int = 0, stream: int = 0):
        _check(lib().tc_channel_llr(self._h, seed, bits_ptr or None, llr_ptr or None, y_ptr or None, h_ptr or None,
                                    stream or None))

    def demap(self, y_ptr: int, h_ptr: int, llr_ptr: int, stream: int = 0):
        _check(lib().tc_qam_demap(self._h, y_ptr or None, h_ptr or None, llr_ptr or None, stream or None))

    def close(self):
        if getattr(self, "_h", None):
            lib().tc_channel_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def derate_match_soft(rx, k: int, filler: int = 0, rv: int = 0, out=None) -> np.ndarray:
    """Soft de-rate matching of one code block on the GPU (tc_derate_match_soft).

    rx: the E received LLRs (bit 1 > 0).  Returns [d0|d1|d2] (3(K+4) floats).
    With ``out`` given, the result is added into it (HARQ combining)."""
    rx = _f32(rx).reshape(-1)
    acc = out is not None
    if out is None:
        out = np.zeros(3 * (k + 4), np.float32)
    elif out.dtype != np.float32 or not out.flags.c_contiguous or out.size != 3 * (k + 4):
        raise ValueError("out must be a contiguous float32 array of 3(K+4)")
    _check(lib().tc_derate_match_soft(rx, rx.size, k, filler, rv, int(acc), out))
    return out


class DeratePlan:
    """Batched soft de-rate matching on the GPU (tc_derate_plan_*).

    ``run`` takes device pointers: ``rx`` holds the E_r received LLRs of every
    block back to back, ``llr`` receives the decoder input layout that
    :class:`Plan` reads.  ``rv=-1`` uses the per-block rv given here."""

    def __init__(self, ks: Sequence[int], es: Sequence[int], fillers=None, rvs=None):
        self.ks = np.ascontiguousarray(ks, dtype=np.uint32)
        self.es = np.ascontiguousarray(es, dtype=np.uint64)
        if self.es.size != self.ks.size:
            raise ValueError("one E per code block")
        self._fil = None if fillers is None else np.ascontiguousarray(fillers, dtype=np.uint32)
        self._rv = None if rvs is None else np.ascontiguousarray(rvs, dtype=np.uint32)
        h = C.c_void_p()
        _check(lib().tc_derate_plan_create(self.ks, None if self._fil is None else self._fil.ctypes.data, self.es,
                                           None if self._rv is None else self._rv.ctypes.data, len(self.ks),
                                           C.byref(h)))
        self._h = h
        self.rx_floats = lib().tc_derate_plan_rx_floats(h)
        self.llr_floats = lib().tc_derate_plan_llr_floats(h)

    def run(self, rx_ptr: int, llr_ptr: int, rv: int = -1, accumulate: bool = False, stream: int = 0):
        _check(lib().tc_derate_plan_run(self._h, rx_ptr or None, rv, int(accumulate), llr_ptr or None,
                                        stream or None))

    def close(self):
        if getattr(self, "_h", None):
            lib().tc_derate_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _tb_descs(tbs):
    arr = (TcTbDesc * max(len(tbs), 1))()
    for i, t in enumerate(tbs):
        if isinstance(t, LteTurboCodecParams):
            t = (t.transport_block_bits, t.g_total, t.rv, t.num_layers, t.mod_order)
        arr[i] = TcTbDesc(*[int(x) for x in t])
    return arr


class TbPlan:
    """Batched transport-block decoding on the GPU (tc_tb_plan_*).

    ``tbs``: LteTurboCodecParams or tuples (B, g_total, rv, num_layers,
    mod_order[, code_block_crc]).  ``decode`` takes device pointers; outputs
    may be 0."""

    def __init__(self, tbs, decoder: int = DECODER_WINDOW_LOG, max_iterations: int = 8,
                 early_stop: bool = False, min_segment: int = 0, warmup: int = 0, extrinsic_bits: int = 0,
                 max_waves_per_block: int = 0):
        self.n_tb = len(tbs)
        self._descs = _tb_descs(tbs)
        o = options(decoder, max_iterations, early_stop, min_segment, warmup, 0, extrinsic_bits, max_waves_per_block)
        h = C.c_void_p()
        _check(lib().tc_tb_plan_create(self._descs, self.n_tb, C.byref(o), C.byref(h)))
        self._h = h
        sz = np.zeros(4, np.uint64)
        _check(lib().tc_tb_plan_sizes(h, sz))
        self.n_cb, self.rx_floats, self.bits_total, self.post_floats = (int(x) for x in sz)
        self.ks = np.zeros(max(self.n_cb, 1), np.uint32)
        self.fillers = np.zeros(max(self.n_cb, 1), np.uint32)
        self.es = np.zeros(max(self.n_cb, 1), np.uint64)
        _check(lib().tc_tb_plan_code_blocks(h, self.ks, self.fillers, self.es))
        self.ks, self.fillers, self.es = self.ks[:self.n_cb], self.fillers[:self.n_cb], self.es[:self.n_cb]

    def decode(self, rx_ptr: int, bits_ptr: int = 0, tb_crc_ptr: int = 0, cb_crc_ptr: int = 0,
               iters_ptr: int = 0, post_ptr: int = 0, rv: int = -1, accumulate: bool = False, stream: int = 0):
        _check(lib().tc_tb_plan_decode(self._h, rx_ptr or None, rv, int(accumulate), bits_ptr or None,
                                       tb_crc_ptr or None, cb_crc_ptr or None, iters_ptr or None,
                                       post_ptr or None, stream or None))

    def decode_multi(self, rx_ptrs, rvs=None, bits_ptr: int = 0, tb_crc_ptr: int = 0, cb_crc_ptr: int = 0,
                     iters_ptr: int = 0, post_ptr: int = 0, stream: int = 0):
        """HARQ in one call: combine 1..4 transmissions (rv per transmission) and decode."""
        n = len(rx_ptrs)
        ptrs = (C.c_void_p * max(n, 1))(*rx_ptrs)
        rv = None if rvs is None else (C.c_int32 * n)(*rvs)
        _check(lib().tc_tb_plan_decode_multi(self._h, n, ptrs, rv, bits_ptr or None, tb_crc_ptr or None,
                                             cb_crc_ptr or None, iters_ptr or None, post_ptr or None,
                                             stream or None))

    def set_hold(self, hold: bool):
        """Whether decodes from received LLRs keep the decoder input for HARQ (default True)."""
        _check(lib().tc_tb_plan_set_hold(self._h, int(hold)))

    def set_staging(self, mode: int):
        """Where the window decoders de-rate match received transmissions:
        STAGING_FUSED (in the decoder's input staging), STAGING_SEPARATE (a
        de-rate kernel per transmission first) or STAGING_AUTO (the default,
        the separate form).  Same results."""
        _check(lib().tc_tb_plan_set_staging(self._h, int(mode)))

    def soft_combine(self, rx_ptr: int, rv: int = -1, accumulate: bool = False, stream: int = 0):
        """De-rate match one transmission into the held decoder input (no decode)."""
        _check(lib().tc_tb_plan_soft_combine(self._h, rx_ptr or None, rv, int(accumulate), stream or None))

    def encode(self, payload_ptr: int, rm_ptr: int, cb_ptr: int = 0, rv: int = -1, stream: int = 0):
        """Transmit side: payload bits (bits_total bytes) -> rate-matched bits
        (rx_floats bytes) and optionally code-block bits (post_floats bytes)."""
        _check(lib().tc_tb_plan_encode(self._h, payload_ptr or None, rv, rm_ptr or None, cb_ptr or None,
                                       stream or None))

    def close(self):
        if getattr(self, "_h", None):
            lib().tc_tb_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def decode_transport_blocks(tbs, rx, decoder: int = DECODER_WINDOW_LOG, max_iterations: int = 8,
                            early_stop: bool = False):
    """Decode many transport blocks from host LLRs on the GPU.

    Returns (payload bits per TB (list), tb_crc_ok bool[n_tb], cb_crc_ok
    bool[n_cb], iterations uint32[n_tb], posterior float32[sum K])."""
    descs = _tb_descs(tbs)
    n = len(tbs)
    rx = _f32(rx).reshape(-1)
    b = [int(descs[i].transport_block_bits) for i in range(n)]
    ks, es = _tb_sizes(descs, n)
    if rx.size != sum(es):  # the C entry point reads sum(E_r) floats
        raise ValueError(f"rx holds {rx.size} LLRs, the transport blocks' code blocks take {sum(es)}")
    bits = np.zeros(max(sum(b), 1), np.uint8)
    tb = np.zeros(max(n, 1), np.int32)
    o = options(decoder, max_iterations, early_stop)
    cb = np.zeros(max(len(ks), 1), np.int32)
    its = np.zeros(max(n, 1), np.uint32)
    post = np.zeros(max(sum(ks), 1), np.float32)
    _check(lib().tc_decode_transport_blocks_host(descs, n, C.byref(o), rx, bits, tb.ctypes.data, cb.ctypes.data,
                                                 its.ctypes.data, post.ctypes.data))
    out, off = [], 0
    for x in b:
        out.append(bits[off:off + x].copy())
        off += x
    return out, tb[:n].astype(bool), cb[:len(ks)].astype(bool), its[:n], post[:sum(ks)]


def _standard_geometry(params: LteTurboCodecParams, encoded: EncodedTransportBlock) -> bool:
    """True when the blocks are those encode_transport_block makes for params."""
    try:
        if params.transport_block_bits == 0 or params.rv > 3:  # device de-rate tables cover rv 0..3
            return False
        c, cp, cm, kp, km, f, _ = _segment_sizes(params.transport_block_bits + 24)
        if c != len(encoded.code_blocks):
            return False
        for r, b in enumerate(encoded.code_blocks):
            d = b.descriptor
            k = km if r < cm else kp
            e = compute_e(r, params.g_total, c, params.num_layers, params.mod_order)
            if (d.k_r, d.d_r, d.filler_count, d.e_r, len(b.rate_matched_bits)) != (k, k + 4, f if r == 0 else 0, e, e):
                return False
        return True
    except (ValueError, IndexError, RuntimeError):
        return False


def encode_transport_blocks(tbs, payloads):
    """GPU encode of many transport blocks (tc_encode_transport_blocks_host).

    payloads: one 0/1 array of B bits per TB.  Returns (rate-matched bits of
    every code block back to back, code-block bits back to back)."""
    descs = _tb_descs(tbs)
    n = len(tbs)
    if len(payloads) != n:
        raise ValueError("one payload per transport block")
    for i, p in enumerate(payloads):  # the C entry point reads sum(B) bytes
        if np.asarray(p).size != int(descs[i].transport_block_bits):
            raise ValueError(f"payload {i} has {np.asarray(p).size} bits, its descriptor says "
                             f"{int(descs[i].transport_block_bits)}")
    pay = np.concatenate([_u8(p) for p in payloads]) if n else np.zeros(1, np.uint8)
    ks, es = _tb_sizes(descs, n)
    nk, ne = sum(ks), sum(es)
    rm = np.zeros(max(ne, 1), np.uint8)
    cb = np.zeros(max(nk, 1), np.uint8)
    _check(lib().tc_encode_transport_blocks_host(descs, n, pay, rm, cb.ctypes.data))
    return rm[:ne], cb[:nk]


def _tb_sizes(descs, n):
    """K_r and E_r of every code block of n TB descriptors, as tc_tb_plan_create
    derives them (a code_block_crc descriptor is one block of B + 24 bits)."""
    ks, es = [], []
    for i in range(n):
        d = descs[i]
        b = int(d.transport_block_bits)
        if d.code_block_crc:
            c, blocks = 1, [b + 24]
        else:
            c, cp, cm, kp, km, _, _ = _segment_sizes(b + 24)
            blocks = [km] * cm + [kp] * cp
        ks += blocks
        es += [compute_e(r, int(d.g_total), c, int(d.num_layers), int(d.mod_order)) for r in range(c)]
    return ks, es


def _segment_sizes(b: int):
    """(c, c_plus, c_minus, k_plus, k_minus, filler, l) of a b-bit TB (CRC included)."""
    seg = np.zeros(7, np.uint64)
    _check(lib().tc_segment(np.zeros(b, np.uint8), b, seg, None))
    return [int(x) for x in seg]


def decode_transport_block(params: LteTurboCodecParams, encoded: EncodedTransportBlock,
                           backend: DecoderBackend) -> DecodedTransportBlock:
    """lte_turbo.cpp:891-980.  GPU backends batch all code blocks in one launch;
    for blocks laid out as encode_transport_block makes them, de-rate
    matching and TB assembly run on the device too."""
    blocks = encoded.code_blocks
    if backend != DecoderBackend.CpuReference and _standard_geometry(params, encoded):
        rx = np.concatenate([np.where(_u8(b.rate_matched_bits) > 0, 8.0, -8.0).astype(np.float32) for b in blocks])
        dec = DECODER_EXACT if backend == DecoderBackend.GpuStage1Exact else DECODER_WINDOW_LOG
        bits, tb_ok, cb_ok, its, post = decode_transport_blocks([params], rx, dec, params.max_iterations,
                                                                params.enable_early_stop)
        return DecodedTransportBlock(bits[0], [bool(x) for x in cb_ok], bool(tb_ok[0]), int(its[0]), post)
    llrs = [derate_match_code_block(params, b.descriptor, b.rate_matched_bits) for b in blocks]
    posts, oks, iters = [], [], []
    if backend == DecoderBackend.CpuReference:
        for b, l in zip(blocks, llrs):
            d = b.descriptor.d_r
            p, ok, it = turbo_decode_code_block_cpu(b.descriptor, l[:d], l[d:2 * d], l[2 * d:],
                                                    params.max_iterations, params.enable_early_stop)
            posts.append(p)
            oks.append(ok)
            iters.append(it)
    else:
        dec = DECODER_EXACT if backend == DecoderBackend.GpuStage1Exact else DECODER_WINDOW_LOG
        ks = [b.descriptor.k_r for b in blocks]
        post, crc, its = decode_batch_host(ks, np.concatenate(llrs), dec, params.max_iterations,
                                           params.enable_early_stop,
                                           fillers=[b.descriptor.filler_count for b in blocks])
        off = 0
        for i, k in enumerate(ks):
            posts.append(post[off:off + k])
            off += k
            oks.append(bool(crc[i]))
            iters.append(int(its[i]) if params.max_iterations else 0)
    tb = []
    cb_ok = []
    for b, p in zip(blocks, posts):
        bits = (p >= 0).astype(np.uint8)[b.descriptor.filler_count:]
        ok = True
        if len(blocks) > 1:
            ok = check_crc24b(bits)
            bits = bits[:-24]
        tb.append(bits)
        cb_ok.append(ok)
    tb_with_crc = np.concatenate(tb) if tb else np.zeros(0, np.uint8)
    return DecodedTransportBlock(tb_with_crc[:-24].copy(), cb_ok, check_crc24a(tb_with_crc),
                                 max(iters) if iters else 0, np.concatenate(posts) if posts else np.zeros(0))
