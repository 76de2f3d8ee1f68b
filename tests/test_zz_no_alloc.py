"""The identity key switch that allocates nothing
(cufhe_identity_key_switch_lvl10_noalloc) against the allocating entry point
cufhe_identity_key_switch_lvl10, through the C ABI: two calls in a row, the
second with fewer TLWEs (the shared scratch keeps the first call's
contents), bit-identical outputs; the size limit is rejected before any
launch."""
import ctypes

import numpy as np
import pytest

import conftest
import cufhe_amd as C
from conftest import load_lvl1_keys

pytestmark = pytest.mark.gpu

W0, W1 = C.TLWE0_WORDS, C.TLWE1_WORDS
MAX_JOBS = 4096


def _bind():
    L = C.lib()
    f = L.cufhe_identity_key_switch_lvl10_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    f.restype = ctypes.c_int
    g = L.cufhe_identity_key_switch_lvl10
    g.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    g.restype = ctypes.c_int
    return f, g


def test_noalloc_key_switch_equals_allocating_entry_point(keys_lvl1):
    C.check(C.lib().cufhe_initialize(), "initialize")
    conftest._LOADED_LVL1["tag"] = None  # load the keys whatever ran before
    load_lvl1_keys(("fft", True), keys_lvl1["bk_kb"], keys_lvl1["ksk"],
                   backend=C.BACKEND_FFT, variant=C.BR_KEY_BUNDLE)
    noalloc, alloc = _bind()
    rng = np.random.default_rng(4242)
    st = C.Stream()
    for n in (1500, 333):
        bits = rng.integers(0, 2, n).astype(np.uint8)
        cts = C.encrypt_bits(1, int(rng.integers(1 << 30)), keys_lvl1["k1"], bits)
        din = C.DeviceBuffer.from_numpy(cts)
        dref = C.DeviceBuffer(n * W0 * 4)
        dout = C.DeviceBuffer(n * W0 * 4)
        C.check(alloc(dref.ptr, din.ptr, n, st.handle), "identity_key_switch_lvl10")
        C.check(noalloc(dout.ptr, din.ptr, n, st.handle), "identity_key_switch_lvl10_noalloc")
        st.synchronize()
        ref = dref.download(np.uint32, (n, W0))
        out = dout.download(np.uint32, (n, W0))
        assert np.array_equal(out, ref), n
        assert list(C.decrypt_bits(0, keys_lvl1["k0"], out)) == list(bits)
    # beyond the maximum: an argument error, nothing launched
    rc = noalloc(dout.ptr, din.ptr, MAX_JOBS + 1, st.handle)
    assert rc == -1  # CUFHE_ERR_ARG
    st.destroy()
