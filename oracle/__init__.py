"""CPU checker for turbo_cpp_amd -- TEST INFRASTRUCTURE ONLY.

Python access (ctypes) to two libraries built by ``oracle/Makefile``:

* ``liboracle.so``: ``turbo_oracle.c``, a C restatement of the reference's
  CPU codec chain (Wiseung/Turbo_CPP ``src/lte_turbo.cpp``) plus the windowed
  decoder that the batched HIP kernel runs.
* ``_ref/libturbo_ref.so``: the reference's own ``src/lte_turbo.cpp`` compiled
  CPU-only with ``ref_shim.cpp``; present only where it was built.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  The product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libturbo_ref.so")

CRC24A, CRC24B, CRC16, CRC8 = 0, 1, 2, 3

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_szp = np.ctypeslib.ndpointer(np.uintp, flags="C_CONTIGUOUS")


def build(with_reference: bool | None = None) -> None:
    """Compile the restatement, and the reference build when its tree exists."""
    targets = ["all"]
    if with_reference is None:
        with_reference = os.path.isdir("/root/reference/src")
    if with_reference:
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", _HERE] + targets, check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build(with_reference=False)
        L = C.CDLL(_LIB_PATH)
        L.or_qpp_count.restype = C.c_size_t
        L.or_qpp_entry.argtypes = [C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.or_qpp_index.argtypes = [C.c_uint32]
        L.or_qpp_index.restype = C.c_int
        L.or_qpp_perm.argtypes = [C.c_uint32, _u32p]
        L.or_crc_len.argtypes = [C.c_int]
        L.or_crc_len.restype = C.c_uint32
        L.or_crc_append.argtypes = [C.c_int, _u8p, C.c_size_t, _u8p]
        L.or_crc_check.argtypes = [C.c_int, _u8p, C.c_size_t]
        L.or_crc_check.restype = C.c_int
        L.or_segment.argtypes = [_u8p, C.c_size_t, _szp, C.c_void_p]
        L.or_segment.restype = C.c_int
        L.or_encode_cb.argtypes = [_u8p, C.c_uint32, C.c_uint32, _u8p, _u8p, _u8p]
        L.or_encode_cb.restype = C.c_int
        L.or_compute_e.argtypes = [C.c_size_t] * 5 + [C.POINTER(C.c_size_t)]
        L.or_compute_e.restype = C.c_int
        L.or_rate_match.argtypes = [_u8p, _u8p, _u8p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_uint32, _u8p]
        L.or_derate_match.argtypes = [_u8p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_uint32, _f32p]
        L.or_derate_match_soft.argtypes = [_f32p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_uint32, _f32p, C.c_int]
        L.or_assemble_tb.argtypes = [_f32p, _u32p, _u32p, C.c_size_t, C.c_size_t, _u8p, C.c_void_p,
                                     C.POINTER(C.c_int)]
        L.or_siso.argtypes = [_f32p, _f32p, _f32p, C.c_uint32, _f32p, _f32p, _f32p]
        L.or_decode_cb.argtypes = [_f32p, _f32p, _f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                   _f32p, C.POINTER(C.c_int), C.POINTER(C.c_uint32)]
        L.or_decode_cb.restype = C.c_int
        L.or_decode_window.argtypes = [_f32p, _f32p, _f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                       C.c_uint32, C.c_uint32, C.c_int, C.c_uint32, _f32p,
                                       C.POINTER(C.c_int), C.POINTER(C.c_uint32)]
        L.or_decode_window.restype = C.c_int
        L.or_decode_exact_seg.argtypes = [_f32p, _f32p, _f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                          C.c_uint32, _f32p, C.POINTER(C.c_int), C.POINTER(C.c_uint32), C.c_void_p]
        L.or_decode_exact_seg.restype = C.c_int
        L.or_round_half.argtypes = [C.c_float]
        L.or_round_half.restype = C.c_float
        L.or_brute_app.argtypes = [_f32p, _f32p, _f32p, C.c_uint32, _f32p, _f32p, _f64p]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def ref():
    """The compiled reference (CPU-only); raises if it was not built."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(_REF_PATH)
        R = C.CDLL(_REF_PATH)
        R.ref_qpp_table.argtypes = [_u32p]
        R.ref_qpp_table.restype = C.c_size_t
        R.ref_qpp_perm.argtypes = [C.c_uint32, _u32p]
        R.ref_crc_append.argtypes = [C.c_int, _u8p, C.c_size_t, _u8p]
        R.ref_crc_check.argtypes = [C.c_int, _u8p, C.c_size_t]
        R.ref_segment.argtypes = [_u8p, C.c_size_t, _szp, C.c_void_p]
        R.ref_encode_cb.argtypes = [_u8p, C.c_size_t, C.c_size_t, _u8p, _u8p, _u8p]
        R.ref_encode_tb.argtypes = [C.c_size_t, C.c_size_t, C.c_uint32, C.c_uint32, C.c_uint32, _u8p,
                                    C.POINTER(C.c_size_t), _szp, C.c_void_p, C.c_void_p]
        R.ref_derate.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, C.c_uint32, _u8p, _f32p]
        R.ref_decode_cb.argtypes = [C.c_size_t, C.c_size_t, _f32p, _f32p, _f32p, C.c_uint32, C.c_int, _f32p,
                                    C.POINTER(C.c_int), C.POINTER(C.c_uint32)]
        R.ref_decode_many.argtypes = [C.c_size_t, C.c_size_t, _f32p, C.c_uint32, _f32p, C.c_int]
        R.ref_tb_roundtrip.argtypes = [C.c_size_t, C.c_size_t, C.c_uint32, C.c_uint32, C.c_int, _u8p, _u8p,
                                       C.POINTER(C.c_int), C.POINTER(C.c_uint32), C.c_void_p,
                                       C.POINTER(C.c_size_t)]
        R.ref_decode_tb.argtypes = [C.c_size_t, C.c_size_t, C.c_uint32, C.c_uint32, C.c_int, _u8p, _u8p, _u8p,
                                    C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_uint32), _f32p]
        _ref = R
    return _ref


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ----------------------------------------------------------------- restatement
def qpp_table():
    L = lib()
    out = []
    for i in range(L.or_qpp_count()):
        k, f1, f2 = C.c_uint32(), C.c_uint32(), C.c_uint32()
        L.or_qpp_entry(i, C.byref(k), C.byref(f1), C.byref(f2))
        out.append((k.value, f1.value, f2.value))
    return out


def qpp_perm(k):
    p = np.zeros(k, np.uint32)
    lib().or_qpp_perm(k, p)
    return p


def crc_append(kind, bits):
    bits = _u8(bits)
    out = np.zeros(len(bits) + lib().or_crc_len(kind), np.uint8)
    lib().or_crc_append(kind, bits, len(bits), out)
    return out


def crc_check(kind, bits):
    bits = _u8(bits)
    return bool(lib().or_crc_check(kind, bits, len(bits)))


def segment(bits):
    """Returns (dict of segmentation fields, list of code-block arrays)."""
    bits = _u8(bits)
    seg = np.zeros(7, np.uintp)
    st = lib().or_segment(bits, len(bits), seg, None)
    if st == -1:
        raise ValueError("transport block cannot be empty")
    if st != 0:
        raise RuntimeError("failed to select K+/K-")
    c, cp, cm, kp, km, f, l = (int(v) for v in seg)
    total = cp * kp + cm * km
    blocks = np.zeros(total, np.uint8)
    lib().or_segment(bits, len(bits), seg, blocks.ctypes.data)
    ks = [km] * cm + [kp] * cp
    out, off = [], 0
    for k in ks:
        out.append(blocks[off:off + k].copy())
        off += k
    return dict(c=c, c_plus=cp, c_minus=cm, k_plus=kp, k_minus=km, filler_bits=f, l=l), out


def encode_cb(cb_bits, filler=0):
    cb_bits = _u8(cb_bits)
    k = len(cb_bits)
    d = [np.zeros(k + 4, np.uint8) for _ in range(3)]
    if lib().or_encode_cb(cb_bits, k, filler, d[0], d[1], d[2]) != 0:
        raise ValueError("unsupported K")
    return d


def compute_e(r, g_total, c, nl, qm):
    e = C.c_size_t()
    if lib().or_compute_e(r, g_total, c, nl, qm, C.byref(e)) != 0:
        raise ValueError("invalid rate-matching parameters")
    return e.value


def rate_match(d0, d1, d2, filler, e, rv):
    out = np.zeros(e, np.uint8)
    lib().or_rate_match(_u8(d0), _u8(d1), _u8(d2), len(d0), filler, e, rv, out)
    return out


def derate_match(bits, d, filler, rv):
    bits = _u8(bits)
    out = np.zeros(3 * d, np.float32)
    lib().or_derate_match(bits, len(bits), d, filler, rv, out)
    return out


def derate_match_soft(llr, d, filler, rv, out=None):
    """Soft de-rate matching; with out given, adds into it (HARQ combining)."""
    llr = _f32(llr)
    acc = out is not None
    if out is None:
        out = np.zeros(3 * d, np.float32)
    lib().or_derate_match_soft(llr, len(llr), d, filler, rv, out, int(acc))
    return out


def assemble_tb(post, ks, fillers, b):
    """Hard decisions, filler strip, CRC24B per block and CRC24A per TB.

    Returns (payload bits (b), per-block CRC ok (bool array), TB CRC ok)."""
    ks = np.ascontiguousarray(ks, dtype=np.uint32)
    fs = np.ascontiguousarray(fillers, dtype=np.uint32)
    bits = np.zeros(b, np.uint8)
    cb = np.zeros(len(ks), np.int32)
    tb = C.c_int()
    if lib().or_assemble_tb(_f32(post), ks, fs, len(ks), b, bits, cb.ctypes.data, C.byref(tb)):
        raise ValueError("code blocks do not add up to b + 24 bits")
    return bits, cb.astype(bool), bool(tb.value)


def siso(sys, parity, apriori, tail_sys, tail_par):
    k = len(sys)
    ext = np.zeros(k, np.float32)
    lib().or_siso(_f32(sys), _f32(parity), _f32(apriori), k, _f32(tail_sys), _f32(tail_par), ext)
    return ext


def decode_cb(d0, d1, d2, k, filler=0, max_iterations=8, early_stop=False):
    post = np.zeros(k, np.float32)
    ok, it = C.c_int(0), C.c_uint32(0)
    if lib().or_decode_cb(_f32(d0), _f32(d1), _f32(d2), k, filler, max_iterations, int(early_stop),
                          post, C.byref(ok), C.byref(it)) != 0:
        raise ValueError("unsupported K")
    return post, bool(ok.value), it.value


def decode_window(d0, d1, d2, k, seg_len, warmup=0, linear=False, filler=0, max_iterations=8,
                  early_stop=False, ex_bits=32):
    """The windowed decoder (or_decode_window); ex_bits=16 holds the
    extrinsics between half-iterations as fp16, as the GPU does for long
    segments (segment_geometry()["extrinsic_bits"])."""
    post = np.zeros(k, np.float32)
    ok, it = C.c_int(0), C.c_uint32(0)
    if lib().or_decode_window(_f32(d0), _f32(d1), _f32(d2), k, filler, max_iterations, int(early_stop),
                              seg_len, warmup, int(linear), ex_bits, post, C.byref(ok), C.byref(it)) != 0:
        raise ValueError("bad K / segment length / extrinsic width")
    return post, bool(ok.value), it.value


def decode_exact_seg(d0, d1, d2, k, seg_len, filler=0, max_iterations=8, early_stop=False, work=None):
    """Segment-parallel decode with exact boundaries (or_decode_exact_seg).
    work: optional uint64 array of 4, accumulated (halves, fix-up rounds,
    chunks recomputed by the slowest segment per round, chunks per pass)."""
    post = np.zeros(k, np.float32)
    ok, it = C.c_int(0), C.c_uint32(0)
    w = None if work is None else work.ctypes.data
    if lib().or_decode_exact_seg(_f32(d0), _f32(d1), _f32(d2), k, filler, max_iterations, int(early_stop),
                                 seg_len, post, C.byref(ok), C.byref(it), w) != 0:
        raise ValueError("bad K / segment length")
    return post, bool(ok.value), it.value


def round_half(x):
    """or_round_half over an array (float32 in, float32 out)."""
    f = lib().or_round_half
    return np.array([f(float(v)) for v in np.asarray(x, np.float32).ravel()], np.float32)


def brute_app(sys, parity, apriori, tail_sys, tail_par):
    k = len(sys)
    out = np.zeros(k, np.float64)
    lib().or_brute_app(_f32(sys), _f32(parity), _f32(apriori), k, _f32(tail_sys), _f32(tail_par), out)
    return out


# ------------------------------------------------------- reference (compiled)
def ref_qpp_table():
    out = np.zeros(3 * 188, np.uint32)
    n = ref().ref_qpp_table(out)
    return [tuple(int(x) for x in out[3 * i:3 * i + 3]) for i in range(n)]


def ref_qpp_perm(k):
    p = np.zeros(k, np.uint32)
    if ref().ref_qpp_perm(k, p) != 0:
        raise IndexError("unsupported QPP block size")
    return p


def ref_crc_append(kind, bits):
    bits = _u8(bits)
    out = np.zeros(len(bits) + 24, np.uint8)
    n = ref().ref_crc_append(kind, bits, len(bits), out)
    return out[:len(bits) + n]


def ref_crc_check(kind, bits):
    bits = _u8(bits)
    return bool(ref().ref_crc_check(kind, bits, len(bits)))


def ref_segment(bits):
    bits = _u8(bits)
    seg = np.zeros(7, np.uintp)
    st = ref().ref_segment(bits, len(bits), seg, None)
    if st:
        raise ValueError(st)
    c, cp, cm, kp, km, f, l = (int(v) for v in seg)
    blocks = np.zeros(cp * kp + cm * km, np.uint8)
    ref().ref_segment(bits, len(bits), seg, blocks.ctypes.data)
    return dict(c=c, c_plus=cp, c_minus=cm, k_plus=kp, k_minus=km, filler_bits=f, l=l), blocks


def ref_encode_cb(cb_bits, filler=0):
    cb_bits = _u8(cb_bits)
    k = len(cb_bits)
    d = [np.zeros(k + 4, np.uint8) for _ in range(3)]
    if ref().ref_encode_cb(cb_bits, k, filler, d[0], d[1], d[2]) != 0:
        raise ValueError("encode failed")
    return d


def ref_decode_tb(payload, g_total, rm, rv=0, max_iterations=8, early_stop=False):
    """The reference's decode_transport_block (CpuReference) on given
    rate-matched bits.  Returns (bits, cb_ok, tb_ok, iterations, posterior)."""
    payload = _u8(payload)
    rm = _u8(rm)
    bits = np.zeros(len(payload), np.uint8)
    cb = np.zeros(64, np.int32)
    tb = C.c_int()
    it = C.c_uint32()
    post = np.zeros(len(payload) + 24 * 64 + 6144 * 2, np.float32)
    st = ref().ref_decode_tb(len(payload), g_total, rv, max_iterations, int(early_stop), payload, rm, bits,
                             cb.ctypes.data, C.byref(tb), C.byref(it), post)
    if st:
        raise ValueError(st)
    return bits, cb, bool(tb.value), int(it.value), post


def ref_encode_tb(bits, g_total, rv=0, nl=1, qm=2):
    bits = _u8(bits)
    n = C.c_size_t()
    desc = np.zeros(6 * 16, np.uintp)
    rm = np.zeros(max(g_total, 1) + 64, np.uint8)
    cbb = np.zeros(len(bits) + 24 * 16 + 6144, np.uint8)
    st = ref().ref_encode_tb(len(bits), g_total, rv, nl, qm, bits, C.byref(n), desc, rm.ctypes.data,
                             cbb.ctypes.data)
    if st:
        raise ValueError(st)
    descs = [tuple(int(x) for x in desc[6 * r:6 * r + 6]) for r in range(n.value)]
    out, off, coff = [], 0, 0
    for (_, k, d, e, f, q) in descs:
        out.append(dict(k=k, d=d, e=e, filler=f, qpp_index=q, rm=rm[off:off + e].copy(),
                        cb_bits=cbb[coff:coff + k].copy()))
        off += e
        coff += k
    return out


def ref_derate(bits, k, filler, rv):
    bits = _u8(bits)
    out = np.zeros(3 * (k + 4), np.float32)
    if ref().ref_derate(k, filler, len(bits), rv, bits, out) != 0:
        raise ValueError("derate failed")
    return out


def ref_decode_cb(d0, d1, d2, k, filler=0, max_iterations=8, early_stop=False):
    post = np.zeros(k, np.float32)
    ok, it = C.c_int(0), C.c_uint32(0)
    st = ref().ref_decode_cb(k, filler, _f32(d0), _f32(d1), _f32(d2), max_iterations, int(early_stop), post,
                             C.byref(ok), C.byref(it))
    if st:
        raise ValueError(st)
    return post, bool(ok.value), it.value


def ref_decode_many(llr, k, max_iterations, threads):
    """llr: (n, 3(k+4)) float32.  Returns (n, k) posteriors."""
    llr = _f32(llr)
    n = llr.shape[0]
    post = np.zeros((n, k), np.float32)
    st = ref().ref_decode_many(n, k, llr.reshape(-1), max_iterations, post.reshape(-1), threads)
    if st:
        raise ValueError(st)
    return post
