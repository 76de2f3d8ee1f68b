"""CMUX / TRGSW API (reference CMUXNTT and TRGSW2NTT: src/cufhe_gates_gpu.cu:
116-133, src/bootstrap_gpu.cu:165-186, kernels :700-813 FFT and :925-1041
NTT): res = c0 + ExtProd(c1 - c0, C) for batches of CMUXes, on both lvl1
backends.

* NTT backend: bit-exact against the oracle's NTT-mode external product
  (the reference NTT path's arithmetic) on random TRLWEs and real TRGSW
  encryptions.
* FFT backend: against the oracle's exact-integer external product, at the
  stated tolerance (>= 99.9% of coefficients equal, |delta| <= 2 Torus32
  units; FP64 rounding of near-half values).
* Semantics on valid ciphertexts: the result decrypts to c1's message when
  the TRGSW encrypts 1 and to c0's when it encrypts 0; descriptor batches
  with repeated selectors and res aliasing c0 give the same results."""
import ctypes

import numpy as np
import pytest

import conftest
import cufhe_amd as C
import oracle as O
from conftest import load_lvl1_keys

pytestmark = pytest.mark.gpu
N = C.N1


def _keys(name, keys_lvl1, keys_bb):
    conftest._LOADED_LVL1["tag"] = None
    if name == "fft":
        load_lvl1_keys(("fft", True), keys_lvl1["bk_kb"], keys_lvl1["ksk"],
                       backend=C.BACKEND_FFT, variant=C.BR_KEY_BUNDLE)
        return keys_lvl1["k1"], O.EXACT
    load_lvl1_keys(("ntt", "bb"), keys_bb["bk"], keys_bb["ksk_subset"],
                   backend=C.BACKEND_NTT, variant=C.BR_BLOCK_BINARY, subset=True)
    return keys_bb["k1"], O.NTT


def _transform(trgsws):
    n = trgsws.shape[0]
    nb = C.lib().cufhe_trgsw_transform_bytes_lvl1()
    assert nb in (65536, 32768)
    d_in = C.DeviceBuffer.from_numpy(trgsws)
    d_out = C.DeviceBuffer(n * nb)
    C.check(C.lib().cufhe_trgsw_to_transform_lvl1(d_in.ptr, d_out.ptr, n, None), "transform")
    C.check(C.lib().cufhe_synchronize(), "sync")
    return d_out, nb


def _extract0(trlwes):
    """SampleExtract index 0 (reference bootstrap_gpu.cu:1201-1221) of rows of TRLWEs."""
    a, b = trlwes[:, :N], trlwes[:, N:]
    t = np.zeros((trlwes.shape[0], N + 1), np.uint32)
    t[:, 0] = a[:, 0]
    t[:, 1:N] = (0 - a[:, :0:-1].astype(np.uint64)).astype(np.uint32)
    t[:, N] = b[:, 0]
    return t


@pytest.mark.parametrize("backend", ["ntt", "fft"])
def test_cmux_batch_matches_oracle(keys_lvl1, keys_bb, backend):
    k1, mode = _keys(backend, keys_lvl1, keys_bb)
    rng = np.random.default_rng(70 + (backend == "fft"))
    n = 24
    sel = rng.integers(0, 2, n)
    trgsw = C.trgsw_encrypt(71, k1, sel, n)
    c1 = rng.integers(0, 1 << 32, (n, 2 * N), dtype=np.uint64).astype(np.uint32)
    c0 = rng.integers(0, 1 << 32, (n, 2 * N), dtype=np.uint64).astype(np.uint32)
    d_cs, nb = _transform(trgsw)
    d1, d0 = C.DeviceBuffer.from_numpy(c1), C.DeviceBuffer.from_numpy(c0)
    dres = C.DeviceBuffer(c1.nbytes)
    C.check(C.lib().cufhe_cmux_lvl1(dres.ptr, d_cs.ptr, d1.ptr, d0.ptr, n, None), "cmux")
    C.check(C.lib().cufhe_synchronize(), "sync")
    got = dres.download(np.uint32, c1.shape)
    want = np.stack([c0[i] + O.external_product(1, mode, c1[i] - c0[i], trgsw[i]) for i in range(n)])
    if mode == O.NTT:
        assert np.array_equal(got, want)
    else:
        diff = np.abs((got - want).view(np.int32).astype(np.int64))
        assert np.mean(diff == 0) >= 0.999 and diff.max() <= 2


@pytest.mark.parametrize("backend", ["ntt", "fft"])
def test_cmux_selects_and_descriptors(keys_lvl1, keys_bb, backend):
    k1, _ = _keys(backend, keys_lvl1, keys_bb)
    rng = np.random.default_rng(80 + (backend == "fft"))
    mu = 1 << 29
    n = 16
    m1 = rng.integers(0, 2, (n, N))
    m0 = rng.integers(0, 2, (n, N))
    p1 = np.where(m1 == 1, mu, (1 << 32) - mu).astype(np.uint32)
    p0 = np.where(m0 == 1, mu, (1 << 32) - mu).astype(np.uint32)
    c1 = C.trlwe_encrypt(81, k1, p1)
    c0 = C.trlwe_encrypt(82, k1, p0)
    sel = np.array([0, 1] * 4, np.int64)  # 8 TRGSWs, each used by two CMUXes
    d_cs, nb = _transform(C.trgsw_encrypt(83, k1, sel, len(sel)))
    d1, d0 = C.DeviceBuffer.from_numpy(c1), C.DeviceBuffer.from_numpy(c0)
    # contiguous form: CMUX i uses TRGSW i % 8 through a descriptor batch, and
    # writes in place over c0
    B = 2 * N * 4
    descs = np.zeros((n, 4), np.uint64)
    for i in range(n):
        descs[i] = [d0.ptr.value + i * B, d_cs.ptr.value + (i % 8) * nb, d1.ptr.value + i * B,
                    d0.ptr.value + i * B]
    dd = C.DeviceBuffer.from_numpy(descs)
    C.check(C.lib().cufhe_cmux_batch_lvl1(dd.ptr, n, None), "cmux batch")
    C.check(C.lib().cufhe_synchronize(), "sync")
    got = d0.download(np.uint32, c0.shape)
    want_bits = np.where(sel[np.arange(n) % 8][:, None] == 1, m1, m0)[:, 0]
    assert list(C.decrypt_bits(1, k1, _extract0(got))) == list(want_bits)
    # the same CMUXes through the contiguous entry point, from fresh inputs
    cs_rep = d_cs.download(np.uint8, (8, nb))[np.arange(n) % 8]
    d_rep = C.DeviceBuffer.from_numpy(cs_rep)
    d0b = C.DeviceBuffer.from_numpy(c0)
    dres = C.DeviceBuffer(c0.nbytes)
    C.check(C.lib().cufhe_cmux_lvl1(dres.ptr, d_rep.ptr, d1.ptr, d0b.ptr, n, None), "cmux")
    C.check(C.lib().cufhe_synchronize(), "sync")
    assert np.array_equal(dres.download(np.uint32, c0.shape), got)
