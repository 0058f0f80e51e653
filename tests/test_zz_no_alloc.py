"""tc_decode_code_block_noalloc (no device allocation) against
tc_decode_code_block: same kernels, so identical posteriors, CRC flags and
iteration counts, bit for bit.  Two calls in a row per decoder, the second
with a smaller block, so what the first call left in the shared static array
cannot leak into the second.  Marked gpu."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import turbo_cpp_amd as T
from tests.helpers import bpsk_awgn_llr

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _block(k, ebn0, seed):
    rng = np.random.default_rng(seed)
    cb = rng.integers(0, 2, k).astype(np.uint8)
    return bpsk_awgn_llr(O.encode_cb(cb), k, ebn0, rng)


def _call(fn, k, llr, opts):
    """One decode through the C ABI entry point fn; returns (post, crc, iters)."""
    d = T.TcCbDesc(0, k, k + 4, 0, 0, 0)
    post = np.full(k, np.nan, np.float32)
    ok, it = C.c_int32(-1), C.c_uint32(12345)
    st = fn(C.byref(d), llr[0], len(llr[0]), llr[1], len(llr[1]), llr[2], len(llr[2]), C.byref(opts), post,
            C.byref(ok), C.byref(it))
    assert st == T.TC_OK, T.lib().tc_last_error()
    return post, ok.value, it.value


CASES = [
    # decoder, early stop, (K, Eb/N0) of the first and the second (smaller) call
    (T.DECODER_EXACT, False, (6144, 0.5), (1024, 0.5)),
    (T.DECODER_EXACT, True, (1024, 1.5), (40, 3.0)),
    (T.DECODER_EXACT_SERIAL, False, (6144, 0.5), (544, 0.5)),
    (T.DECODER_WINDOW_LOG, False, (6144, 0.5), (544, 0.5)),
    (T.DECODER_WINDOW_LOG, True, (6144, 1.5), (104, 3.0)),
    (T.DECODER_WINDOW_LIN, True, (2048, 1.0), (352, 2.0)),
]


@pytest.mark.parametrize("decoder,es,first,second", CASES)
def test_noalloc_equals_allocating_entry(decoder, es, first, second):
    L = T.lib()
    for i, (k, eb) in enumerate((first, second)):
        _say(f"noalloc decoder={decoder} es={es} call {i} K={k}")
        llr = _block(k, eb, 100 * k + i)
        opts = T.options(decoder, 8, es)
        want = _call(L.tc_decode_code_block, k, llr, opts)
        got = _call(L.tc_decode_code_block_noalloc, k, llr, opts)
        assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32))
        assert got[1:] == want[1:]


def test_noalloc_rejects_bad_sizes():
    L = T.lib()
    llr = _block(40, 1.0, 1)
    d = T.TcCbDesc(0, 40, 44, 0, 0, 0)
    post = np.zeros(40, np.float32)
    ok, it = C.c_int32(0), C.c_uint32(0)
    opts = T.options(T.DECODER_EXACT, 8, False)
    st = L.tc_decode_code_block_noalloc(C.byref(d), llr[0], 43, llr[1], 44, llr[2], 44, C.byref(opts), post,
                                        C.byref(ok), C.byref(it))
    assert st == T.TC_ERR_INVALID_ARGUMENT
    d.k_r = 41  # not a QPP size
    st = L.tc_decode_code_block_noalloc(C.byref(d), llr[0], 44, llr[1], 44, llr[2], 44, C.byref(opts), post,
                                        C.byref(ok), C.byref(it))
    assert st == T.TC_ERR_OUT_OF_RANGE
    # a block on several waves needs that many waves' work space: more than the static array holds
    big = _block(6144, 1.0, 2)
    d = T.TcCbDesc(0, 6144, 6148, 0, 0, 0)
    post = np.zeros(6144, np.float32)
    opts = T.options(T.DECODER_WINDOW_LOG, 8, False, 32, 16, 0, 0, 3)
    st = L.tc_decode_code_block_noalloc(C.byref(d), big[0], 6148, big[1], 6148, big[2], 6148, C.byref(opts), post,
                                        C.byref(ok), C.byref(it))
    assert st == T.TC_ERR_INVALID_ARGUMENT and b"waves" in L.tc_last_error()
