"""Shared fixtures. GPU tests are marked ``gpu``; everything else runs on CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def keys_lvl1():
    """Seed-0 keys: secret keys, lvl01 BK (plain and key bundle), lvl10 KSK."""
    import cufhe_amd as C
    k0, k1, k2 = C.secret_key_gen(0)
    return {
        "k0": k0, "k1": k1, "k2": k2,
        "bk_kb": C.bk_gen(1, True, 0, k0, k1),
        "bk": C.bk_gen(1, False, 0, k0, k1),
        "ksk": C.ksk_gen(1, 0, k1, k0),
    }


@pytest.fixture(scope="session")
def keys_lvl2(keys_lvl1):
    import cufhe_amd as C
    k0, k2 = keys_lvl1["k0"], keys_lvl1["k2"]
    return {
        "k0": k0, "k2": k2,
        "bk_kb": C.bk_gen(2, True, 0, k0, k2),
        "ksk": C.ksk_gen(2, 0, k2, k0),
    }


@pytest.fixture(scope="session")
def keys_bb():
    """Block-binary lvl0 key (blocks of 4, at most one bit each) whose lvl1 key
    starts with it; plain lvl01 BK (one TRGSW per bit) and the subset KSK."""
    import cufhe_amd as C
    k0, k1, k2 = C.secret_key_gen(3, block_binary=True)
    return {
        "k0": k0, "k1": k1,
        "bk": C.bk_gen(1, False, 3, k0, k1),
        "ksk_subset": C.ksk_gen(1, 3, k1, k0, subset=True),
    }


_LOADED_LVL1 = {"tag": None}


def load_lvl1_keys(tag, bk, ksk, backend=0, variant=1, subset=False):
    """Load lvl1 gate keys on the GPU unless the same set (tag) is loaded."""
    import cufhe_amd as C
    if _LOADED_LVL1["tag"] == tag:
        return
    _LOADED_LVL1["tag"] = None
    C.initialize_lvl01(bk, ksk, backend=backend, variant=variant, subset=subset)
    _LOADED_LVL1["tag"] = tag


def signed32(x):
    return np.asarray(x, np.uint32).view(np.int32).astype(np.int64)


def signed64(x):
    return np.asarray(x, np.uint64).view(np.int64)
