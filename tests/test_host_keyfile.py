"""Key files of the host layer (tfhe_keyfile_*; C++ SecretKey/EvalKey
save/load and the Python helpers): round trips and malformed input. CPU."""
import os
import subprocess

import numpy as np
import pytest

import cufhe_amd as C
from conftest import ROOT


def test_cpp_keyfile_round_trip(tmp_path):
    exe = os.path.join(ROOT, "tests", "cpp", "test_keyfile.bin")
    if not os.path.exists(exe):
        pytest.skip("build with `python cufhe_amd/build.py` first")
    p = subprocess.run([exe, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=240)
    print(p.stdout)
    assert p.returncode == 0 and "ALL PASS" in p.stdout


def test_python_keyfile_round_trip(tmp_path):
    k0, k1, k2 = C.secret_key_gen(5)
    ksk = C.ksk_gen(1, 5, k1, k0)
    path = str(tmp_path / "keys.bin")
    C.save_keyfile(path, {"sklvl0": k0, "sklvl1": k1, "sklvl2": k2, "iksklvl10": ksk},
                   key_bundle=True)
    flags, arrays = C.load_keyfile(path)
    assert flags["key_bundle"] and not flags["ntt"]
    assert set(arrays) == {"sklvl0", "sklvl1", "sklvl2", "iksklvl10"}
    for name, a in (("sklvl0", k0), ("sklvl1", k1), ("sklvl2", k2), ("iksklvl10", ksk)):
        assert arrays[name].dtype == a.dtype and np.array_equal(arrays[name], a), name
    with pytest.raises(C.CufheError):
        C.save_keyfile(path, {"iksklvl10": ksk[:-1]})
